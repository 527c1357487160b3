// Sweep assembly for the weno2d implicit step (Fortran_Original/examples/weno2d/
// weno_mpi_gpu.f90): the m = 3 block form of (I - gamma dt L) for the x or y sweep of a
// scalar field under frozen WENO5 upwinding, built straight into the solve's layout. The
// reference computes the five stencil coefficients of every interface into coef_x / coef_y
// (compute_weno_coeffs_{x,y}_kernel, :430-485), then reads them back to fill the blocks
// (build_matrix_{x,y}_kernel_weno5, :504-719); here one lane owns one block row of one line
// and computes the coefficients of its four interfaces from the nine cells they touch, in
// registers, so the coefficients never go to memory.
//
// The field is Fortran-ordered with a three-cell halo, U(1-3:n1+3, 1-3:n2+3), filled by the
// caller (periodic for the reference). The x sweep's systems are the rows j (sys = j - 1),
// its block rows the cell triples (3I-2, 3I-1, 3I) along i; the y sweep's systems are the
// columns i, its block rows the triples along j: the numbering of trA(n2sub, n1sub/m, m, m)
// and trA(n1sub, n2sub/m, m, m). vel holds each line's advection speed (c_x of row j, c_y of
// column i); vel >= 0 takes the left-biased stencil.
#include "btd_internal.h"
#include "btd_kernels.h"

namespace btd {

namespace {

// WENO5 linear weights and epsilon (weno_mpi_gpu.f90:77-80).
constexpr double kW0 = 1.0 / 10.0, kW1 = 6.0 / 10.0, kW2 = 3.0 / 10.0, kWenoEps = 1.0e-6;
constexpr int kWenoHalo = 3;
constexpr int kWenoRows = 16;  // block rows per lane
constexpr int kWenoTile = 3 * kWenoRows + 7;  // cells of a staged line (+1: fewer LDS bank conflicts)

// Left-biased WENO5 reconstruction at the interface to the right of g2 from the cells
// g0..g4: the coefficient of each cell (weno5_coeffs_left, :381-398). The right-biased form
// (weno5_coeffs_right, :400-418) is this applied to the mirrored stencil.
__device__ __forceinline__ void weno5_left(double g0, double g1, double g2, double g3, double g4,
                                           double (&c)[5]) {
#pragma clang fp contract(off)
    const double t0 = g0 - 2.0 * g1 + g2, u0 = g0 - 4.0 * g1 + 3.0 * g2;
    const double t1 = g1 - 2.0 * g2 + g3, u1 = g1 - g3;
    const double t2 = g2 - 2.0 * g3 + g4, u2 = 3.0 * g2 - 4.0 * g3 + g4;
    const double b0 = (13.0 / 12.0) * (t0 * t0) + 0.25 * (u0 * u0);
    const double b1 = (13.0 / 12.0) * (t1 * t1) + 0.25 * (u1 * u1);
    const double b2 = (13.0 / 12.0) * (t2 * t2) + 0.25 * (u2 * u2);
    const double e0 = kWenoEps + b0, e1 = kWenoEps + b1, e2 = kWenoEps + b2;
    const double a0 = kW0 / (e0 * e0), a1 = kW1 / (e1 * e1), a2 = kW2 / (e2 * e2);
    const double s = a0 + a1 + a2;
    const double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
    c[0] = w0 / 3.0;
    c[1] = -7.0 * w0 / 6.0 - w1 / 6.0;
    c[2] = 11.0 * w0 / 6.0 + 5.0 * w1 / 6.0 + w2 / 3.0;
    c[3] = w1 / 3.0 + 5.0 * w2 / 6.0;
    c[4] = -w2 / 6.0;
}

// The three blocks of block row r of a line from its cells 3r-2 .. 3r+6 (f[0..8]): the WENO5
// coefficients of the row's four interfaces, then each cell's 7-point stencil of I - gdt L
// (scale = gdt c / h) spread over the block columns (build_matrix_{x,y}_kernel_weno5).
__device__ __forceinline__ void weno_row_blocks(const double (&f)[9], bool left, double scale,
                                                double (&Ab)[3][3], double (&Bb)[3][3], double (&Cb)[3][3]) {
#pragma clang fp contract(off)
    double cf[4][5];  // interfaces 3r .. 3r+3 (interface k to the right of cell k)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (left) {
            weno5_left(f[u], f[u + 1], f[u + 2], f[u + 3], f[u + 4], cf[u]);
        } else {
            double c[5];
            weno5_left(f[u + 5], f[u + 4], f[u + 3], f[u + 2], f[u + 1], c);
#pragma unroll
            for (int k = 0; k < 5; ++k) cf[u][k] = c[4 - k];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Ab[i][j] = 0.0;
            Bb[i][j] = i == j ? 1.0 : 0.0;
            Cb[i][j] = 0.0;
        }
#pragma unroll
    for (int irow = 0; irow < 3; ++irow) {
        // cell 3r + irow + 1: f_hat to its right (interface irow + 1) minus f_hat to its left
        double L[7];
#pragma unroll
        for (int o = 0; o < 7; ++o) L[o] = 0.0;
        // offsets -3..3 of L[0..6]: the left-biased stencils start two cells back, the
        // right-biased one (indices kept constant so that L stays in registers)
        if (left) {
#pragma unroll
            for (int k = 0; k < 5; ++k) L[1 + k] += cf[irow + 1][k];
#pragma unroll
            for (int k = 0; k < 5; ++k) L[k] -= cf[irow][k];
        } else {
#pragma unroll
            for (int k = 0; k < 5; ++k) L[2 + k] += cf[irow + 1][k];
#pragma unroll
            for (int k = 0; k < 5; ++k) L[1 + k] -= cf[irow][k];
        }
#pragma unroll
        for (int o = 0; o < 7; ++o) {
            const double val = scale * L[o];
            const int col = irow + o - 3;  // column relative to this block row's first cell
            if (col < 0)
                Ab[irow][col + 3] += val;
            else if (col < 3)
                Bb[irow][col] += val;
            else
                Cb[irow][col - 3] += val;
        }
    }
}

template <int LAYOUT, int AXIS>
__global__ __launch_bounds__(64) void k_weno_assemble(const double* __restrict__ U, const double* __restrict__ vel,
                                                      int n1, int n2, double h, double gdt,
                                                      double* __restrict__ A, double* __restrict__ B,
                                                      double* __restrict__ C, int64_t nsys, int64_t nrow) {
    // the reference's operation order, unfused, so that the blocks are the oracle's bit for bit
#pragma clang fp contract(off)
    const int64_t s0 = (int64_t)blockIdx.x * 64, sys = s0 + threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * kWenoRows;
    const int64_t r1 = r0 + kWenoRows < nrow ? r0 + kWenoRows : nrow;
    const int64_t ld = (int64_t)n1 + 2 * kWenoHalo;  // leading dimension of U
    // x sweep: the lines are rows of U, so the wavefront's 64 lines are staged through LDS,
    // each row's cells 3 r0 - 2 .. 3 r1 + 3 loaded by consecutive lanes (whole lines of U)
    __shared__ double tile[AXIS == 0 ? 64 : 1][kWenoTile];
    if constexpr (AXIS == 0) {
        const int ncell = 3 * (int)(r1 - r0) + 6;  // <= 54 < 64: one cell per lane and row
        const int t = threadIdx.x;
        const double* row = U + 3 * r0 + t + ld * (s0 + kWenoHalo);
        const int rows = nsys - s0 < 64 ? (int)(nsys - s0) : 64;
        if (t < ncell)
            for (int jj = 0; jj < rows; ++jj) tile[jj][t] = row[ld * jj];
        __syncthreads();
    }
    if (sys >= nsys) return;
    // cell c (1-based, -2..n+3) of this line
    auto cell = [&](int64_t c) {
        if constexpr (AXIS == 0) return tile[threadIdx.x][c - (3 * r0 - 2)];
        return U[(sys + kWenoHalo) + ld * (c + 2)];
    };
    const double v = vel[sys];
    const bool left = v >= 0.0;
    const double scale = (gdt * v) / h;
    const Addr<LAYOUT> am(sys, nsys, nrow, 9);
    double f[9];  // cells 3r-2 .. 3r+6 (1-based) of block row r: a window sliding by three cells
#pragma unroll
    for (int t = 0; t < 6; ++t) f[t + 3] = cell(3 * r0 - 2 + t);
    for (int64_t r = r0; r < r1; ++r) {
#pragma unroll
        for (int t = 0; t < 6; ++t) f[t] = f[t + 3];
#pragma unroll
        for (int t = 6; t < 9; ++t) f[t] = cell(3 * r - 2 + t);
        double Ab[3][3], Bb[3][3], Cb[3][3];
        weno_row_blocks(f, left, scale, Ab, Bb, Cb);
        store_mat<3>(A, am, r, Ab);
        store_mat<3>(B, am, r, Bb);
        store_mat<3>(C, am, r, Cb);
    }
}

// Block source of stage 1 for the matrix-free periodic sweep: a lane's blocks are computed from
// U as its rows advance (a sliding nine-cell window), never stored.
template <int AXIS>
struct WenoBlocks {
    const double* U;
    const double* vel;
    int n1;
    double h, gdt;
    struct Lane {
        const double* U;
        int64_t ld, sys, cur;
        bool left;
        double scale, f[9], A[3][3], B[3][3], C[3][3];
        __device__ double cell(int64_t c) const {
            return AXIS == 0 ? U[(c + 2) + ld * (sys + kWenoHalo)] : U[(sys + kWenoHalo) + ld * (c + 2)];
        }
        __device__ void row(int64_t g) {
            if (g == cur + 1) {
#pragma unroll
                for (int t = 0; t < 6; ++t) f[t] = f[t + 3];
#pragma unroll
                for (int t = 6; t < 9; ++t) f[t] = cell(3 * g - 2 + t);
            } else {
#pragma unroll
                for (int t = 0; t < 9; ++t) f[t] = cell(3 * g - 2 + t);
            }
            cur = g;
            weno_row_blocks(f, left, scale, A, B, C);
        }
        __device__ void b(double (&R)[3][3]) const {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) R[i][j] = B[i][j];
        }
        __device__ void a_mat(double (&X)[3][3]) const {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) X[i][j] = A[i][j];
        }
        __device__ void a_col(int k, double (&a)[3]) const {
#pragma unroll
            for (int i = 0; i < 3; ++i) a[i] = A[i][k];
        }
        __device__ void c_col(int j, double (&c)[3]) const {
#pragma unroll
            for (int i = 0; i < 3; ++i) c[i] = C[i][j];
        }
    };
    __device__ Lane lane(int64_t sys) const {
        Lane l;
        l.U = U;
        l.ld = (int64_t)n1 + 2 * kWenoHalo;
        l.sys = sys;
        l.cur = -2;
        const double v = vel[sys];
        l.left = v >= 0.0;
        l.scale = (gdt * v) / h;
        return l;
    }
};

template <int LAYOUT, int AXIS>
hipError_t weno_la(const double* U, const double* vel, int n1, int n2, double h, double gdt, double* A,
                   double* B, double* C, int64_t nsys, int64_t nrow, hipStream_t s) {
    const dim3 grid((unsigned)((nsys + 63) / 64), (unsigned)((nrow + kWenoRows - 1) / kWenoRows));
    hipLaunchKernelGGL((k_weno_assemble<LAYOUT, AXIS>), grid, dim3(64), 0, s, U, vel, n1, n2, h, gdt, A, B, C,
                       nsys, nrow);
    return hipGetLastError();
}

template <int LAYOUT>
hipError_t weno_l(int axis, const double* U, const double* vel, int n1, int n2, double h, double gdt,
                  double* A, double* B, double* C, int64_t nsys, int64_t nrow, hipStream_t s) {
    if (axis == 0) return weno_la<LAYOUT, 0>(U, vel, n1, n2, h, gdt, A, B, C, nsys, nrow, s);
    return weno_la<LAYOUT, 1>(U, vel, n1, n2, h, gdt, A, B, C, nsys, nrow, s);
}

// Field F(n1, n2) (Fortran order, no halo) <-> the sweep's vector d in the solve's layout:
// element p of block row r of line sys is cell 3r + p + 1 of that line (pack_scalar_to_xblock
// / unpack_xblock_to_scalar, weno_mpi_gpu.f90:279-301, and for the y sweep the transposes
// around it, :332-356). One lane per line and kWenoRows block rows; TO_D gathers, else scatters.
template <int LAYOUT, int AXIS, bool TO_D>
__global__ __launch_bounds__(64) void k_weno_move(double* __restrict__ F, double* __restrict__ D, int n1,
                                                  int64_t nsys, int64_t nrow) {
    const int64_t s0 = (int64_t)blockIdx.x * 64, sys = s0 + threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * kWenoRows;
    const int64_t r1 = r0 + kWenoRows < nrow ? r0 + kWenoRows : nrow;
    const Addr<LAYOUT> av(sys, nsys, nrow, 3);
    if constexpr (AXIS == 0) {
        // the lines are rows of F: staged through LDS so that F is read or written in whole
        // runs of cells by consecutive lanes
        __shared__ double tile[64][3 * kWenoRows + 1];
        const int nc = 3 * (int)(r1 - r0), t = threadIdx.x;
        const int rows = nsys - s0 < 64 ? (int)(nsys - s0) : 64;
        double* row = F + 3 * r0 + t + (int64_t)n1 * s0;
        if constexpr (TO_D) {
            if (t < nc)
                for (int jj = 0; jj < rows; ++jj) tile[jj][t] = row[(int64_t)n1 * jj];
            __syncthreads();
            if (sys < nsys)
                for (int64_t r = r0; r < r1; ++r)
#pragma unroll
                    for (int p = 0; p < 3; ++p) D[av.at(r, p)] = tile[t][3 * (r - r0) + p];
        } else {
            if (sys < nsys)
                for (int64_t r = r0; r < r1; ++r)
#pragma unroll
                    for (int p = 0; p < 3; ++p) tile[t][3 * (r - r0) + p] = D[av.at(r, p)];
            __syncthreads();
            if (t < nc)
                for (int jj = 0; jj < rows; ++jj) row[(int64_t)n1 * jj] = tile[jj][t];
        }
        return;
    }
    if (sys >= nsys) return;
    for (int64_t r = r0; r < r1; ++r) {
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const int64_t c = 3 * r + p;  // 0-based cell along the line
            const int64_t f = AXIS == 0 ? c + (int64_t)n1 * sys : sys + (int64_t)n1 * c;
            if constexpr (TO_D)
                D[av.at(r, p)] = F[f];
            else
                F[f] = D[av.at(r, p)];
        }
    }
}

template <int LAYOUT, int AXIS>
hipError_t move_la(bool to_d, double* F, double* D, int n1, int64_t nsys, int64_t nrow, hipStream_t s) {
    const dim3 grid((unsigned)((nsys + 63) / 64), (unsigned)((nrow + kWenoRows - 1) / kWenoRows));
    if (to_d)
        hipLaunchKernelGGL((k_weno_move<LAYOUT, AXIS, true>), grid, dim3(64), 0, s, F, D, n1, nsys, nrow);
    else
        hipLaunchKernelGGL((k_weno_move<LAYOUT, AXIS, false>), grid, dim3(64), 0, s, F, D, n1, nsys, nrow);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_weno_move(int layout, int axis, bool to_d, double* F, double* D, int n1, int nsys, int nrow,
                            hipStream_t s) {
    if (layout == LAYOUT_TILED)
        return axis == 0 ? move_la<LAYOUT_TILED, 0>(to_d, F, D, n1, nsys, nrow, s)
                         : move_la<LAYOUT_TILED, 1>(to_d, F, D, n1, nsys, nrow, s);
    return axis == 0 ? move_la<LAYOUT_REF_GPU, 0>(to_d, F, D, n1, nsys, nrow, s)
                     : move_la<LAYOUT_REF_GPU, 1>(to_d, F, D, n1, nsys, nrow, s);
}

hipError_t launch_weno_sweep_cyclic(int layout, int axis, const double* U, const double* vel, int n1, int n2,
                                    double h, double gdt, double* Aw, double* Cw, double* D, int nsys, int nrow,
                                    const int* pf, const int* pl, int npart, double* redA, double* redC,
                                    double* redD, int* flag, hipStream_t s) {
    (void)n2;
    if (axis == 0)
        return launch_two_stage_m<3, WenoBlocks<0>>(layout, Aw, nullptr, Cw, D, nsys, nrow, pf, pl, npart, redA,
                                                    redC, redD, flag, s, nullptr, true,
                                                    WenoBlocks<0>{U, vel, n1, h, gdt});
    return launch_two_stage_m<3, WenoBlocks<1>>(layout, Aw, nullptr, Cw, D, nsys, nrow, pf, pl, npart, redA, redC,
                                                redD, flag, s, nullptr, true, WenoBlocks<1>{U, vel, n1, h, gdt});
}

hipError_t launch_weno_assemble(int layout, int axis, const double* U, const double* vel, int n1, int n2,
                                double h, double gdt, double* A, double* B, double* C, int nsys, int nrow,
                                hipStream_t s) {
    if (layout == LAYOUT_TILED) return weno_l<LAYOUT_TILED>(axis, U, vel, n1, n2, h, gdt, A, B, C, nsys, nrow, s);
    return weno_l<LAYOUT_REF_GPU>(axis, U, vel, n1, n2, h, gdt, A, B, C, nsys, nrow, s);
}

}  // namespace btd
