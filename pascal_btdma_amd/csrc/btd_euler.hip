// Launchers of the euler3d sweep assembly and of the field <-> sweep moves.
#include "btd_euler.h"

namespace btd {

namespace {
constexpr int kRowsPerBlock = 32;

template <int LAYOUT, int AXIS>
hipError_t assemble_la(const double* U, const double* F, const Grid& g, double cc, double cd,
                       double* A, double* B, double* C, double* D, int64_t nsys, int64_t nrow,
                       hipStream_t s) {
    const dim3 grid((unsigned)((nsys + 63) / 64), (unsigned)((nrow + kRowsPerBlock - 1) / kRowsPerBlock));
    if constexpr (AXIS == 0)
        hipLaunchKernelGGL((k_euler_assemble_x<LAYOUT>), grid, dim3(64), 0, s, U, F, g, cc, cd, A, B, C, D,
                           nsys, nrow, kRowsPerBlock);
    else
        hipLaunchKernelGGL((k_euler_assemble<LAYOUT, AXIS>), grid, dim3(64), 0, s, U, F, g, cc, cd, A, B,
                           C, D, nsys, nrow, kRowsPerBlock);
    return hipGetLastError();
}

template <int LAYOUT, int AXIS>
hipError_t move_la(double* F, double* D, const Grid& g, int m, int64_t nsys, int64_t nrow,
                   bool scatter, hipStream_t s) {
    const dim3 grid((unsigned)((nsys + 63) / 64), (unsigned)((nrow + kRowsPerBlock - 1) / kRowsPerBlock));
    if (scatter)
        hipLaunchKernelGGL((k_sweep_move<LAYOUT, AXIS, true>), grid, dim3(64), 0, s, F, D, g, m, nsys,
                           nrow, kRowsPerBlock);
    else
        hipLaunchKernelGGL((k_sweep_move<LAYOUT, AXIS, false>), grid, dim3(64), 0, s, F, D, g, m, nsys,
                           nrow, kRowsPerBlock);
    return hipGetLastError();
}

template <int LAYOUT>
hipError_t assemble_l(int axis, const double* U, const double* F, const Grid& g, double cc,
                      double cd, double* A, double* B, double* C, double* D, int64_t nsys,
                      int64_t nrow, hipStream_t s) {
    if (axis == 0) return assemble_la<LAYOUT, 0>(U, F, g, cc, cd, A, B, C, D, nsys, nrow, s);
    if (axis == 1) return assemble_la<LAYOUT, 1>(U, F, g, cc, cd, A, B, C, D, nsys, nrow, s);
    return assemble_la<LAYOUT, 2>(U, F, g, cc, cd, A, B, C, D, nsys, nrow, s);
}

template <int LAYOUT>
hipError_t move_l(int axis, double* F, double* D, const Grid& g, int m, int64_t nsys, int64_t nrow,
                  bool scatter, hipStream_t s) {
    if (axis == 0) return move_la<LAYOUT, 0>(F, D, g, m, nsys, nrow, scatter, s);
    if (axis == 1) return move_la<LAYOUT, 1>(F, D, g, m, nsys, nrow, scatter, s);
    return move_la<LAYOUT, 2>(F, D, g, m, nsys, nrow, scatter, s);
}
}  // namespace

hipError_t launch_euler_assemble(int layout, int axis, const double* U, const double* F, int n1,
                                 int n2, int n3, double cc, double cd, double* A, double* B,
                                 double* C, double* D, int nsys, int nrow, hipStream_t s) {
    const Grid g{n1, n2, n3};
    if (layout == LAYOUT_TILED)
        return assemble_l<LAYOUT_TILED>(axis, U, F, g, cc, cd, A, B, C, D, nsys, nrow, s);
    return assemble_l<LAYOUT_REF_GPU>(axis, U, F, g, cc, cd, A, B, C, D, nsys, nrow, s);
}

namespace {
// field geometry of the haloed U and of a field without halo, for the x-line transposes
XField xfield_halo(const Grid& g) {
    return XField{g.n2, g.n1 + 2, (g.n1 + 2) * (g.n2 + 2), (g.n1 + 2) * (g.n2 + 2) * (g.n3 + 2), 1, 0};
}
XField xfield_plain(const Grid& g) { return XField{g.n2, g.n1, g.n1 * g.n2, g.n1 * g.n2 * g.n3, 0, 0}; }

template <int LAYOUT>
hipError_t xlines(bool to_lines, double* F, double* Ln, const XField& f, int nv, int64_t nsys,
                  int64_t L, hipStream_t s) {
    const dim3 grid((unsigned)((nsys + 63) / 64), (unsigned)((L + 63) / 64));
    if (to_lines)
        hipLaunchKernelGGL((k_xlines<LAYOUT, true>), grid, dim3(256), 0, s, F, Ln, f, nv, nsys, L);
    else
        hipLaunchKernelGGL((k_xlines<LAYOUT, false>), grid, dim3(256), 0, s, F, Ln, f, nv, nsys, L);
    return hipGetLastError();
}

template <int LAYOUT, int AXIS>
hipError_t sweep_cyclic_la(const double* U, const double* rhs, const Grid& g, double cc, double cd,
                           double* Wc, double* We, double* Wd, double* Ut, double* F, int64_t nsys,
                           int64_t nrow, int* flag, hipStream_t s) {
    const dim3 grid((unsigned)((nsys + 63) / 64));
    if constexpr (AXIS == 0) {
        // x lines are the field's contiguous direction: move U (with its halo cells along x) and
        // rhs into line order, sweep there (x over the rhs lines), and move x back
        hipError_t e = xlines<LAYOUT>(true, const_cast<double*>(U), Ut, xfield_halo(g), 5, nsys, nrow + 2, s);
        if (e == hipSuccess) e = xlines<LAYOUT>(true, const_cast<double*>(rhs), Wd, xfield_plain(g), 5, nsys, nrow, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_euler_sweep_cyclic<LAYOUT, AXIS, true>), grid, dim3(64), 0, s, Ut, Wd, g, cc,
                           cd, Wc, We, Wd, Wd, nsys, nrow, flag);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        return xlines<LAYOUT>(false, F, Wd, xfield_plain(g), 5, nsys, nrow, s);
    } else {
        hipLaunchKernelGGL((k_euler_sweep_cyclic<LAYOUT, AXIS, false>), grid, dim3(64), 0, s, U, rhs, g,
                           cc, cd, Wc, We, Wd, F, nsys, nrow, flag);
        return hipGetLastError();
    }
}

template <int LAYOUT>
hipError_t sweep_cyclic_l(int axis, const double* U, const double* rhs, const Grid& g, double cc,
                          double cd, double* Wc, double* We, double* Wd, double* Ut, double* F,
                          int64_t nsys, int64_t nrow, int* flag, hipStream_t s) {
    if (axis == 0) return sweep_cyclic_la<LAYOUT, 0>(U, rhs, g, cc, cd, Wc, We, Wd, Ut, F, nsys, nrow, flag, s);
    if (axis == 1) return sweep_cyclic_la<LAYOUT, 1>(U, rhs, g, cc, cd, Wc, We, Wd, Ut, F, nsys, nrow, flag, s);
    return sweep_cyclic_la<LAYOUT, 2>(U, rhs, g, cc, cd, Wc, We, Wd, Ut, F, nsys, nrow, flag, s);
}
}  // namespace

hipError_t launch_euler_sweep_cyclic(int layout, int axis, const double* U, const double* rhs,
                                     int n1, int n2, int n3, double cc, double cd, double* work,
                                     double* F, int nsys, int nrow, int* flag, hipStream_t s) {
    const Grid g{n1, n2, n3};
    const int64_t nm = layout_elems(layout, nsys, nrow, 25), nv = layout_elems(layout, nsys, nrow, 5);
    double* Wc = work;
    double* We = work + nm;
    double* Wd = work + 2 * nm;
    double* Ut = work + 2 * nm + nv;  // x sweeps: U's lines, nrow + 2 rows
    if (layout == LAYOUT_TILED)
        return sweep_cyclic_l<LAYOUT_TILED>(axis, U, rhs, g, cc, cd, Wc, We, Wd, Ut, F, nsys, nrow, flag, s);
    return sweep_cyclic_l<LAYOUT_REF_GPU>(axis, U, rhs, g, cc, cd, Wc, We, Wd, Ut, F, nsys, nrow, flag, s);
}

hipError_t launch_sweep_move(int layout, int axis, double* F, double* D, int n1, int n2, int n3,
                             int m, int nsys, int nrow, bool scatter, hipStream_t s) {
    const Grid g{n1, n2, n3};
    if (axis == 0) {  // through LDS tiles: the field's contiguous direction is the sweep's
        if (layout == LAYOUT_TILED) return xlines<LAYOUT_TILED>(!scatter, F, D, xfield_plain(g), m, nsys, nrow, s);
        return xlines<LAYOUT_REF_GPU>(!scatter, F, D, xfield_plain(g), m, nsys, nrow, s);
    }
    if (layout == LAYOUT_TILED) return move_l<LAYOUT_TILED>(axis, F, D, g, m, nsys, nrow, scatter, s);
    return move_l<LAYOUT_REF_GPU>(axis, F, D, g, m, nsys, nrow, scatter, s);
}

}  // namespace btd
