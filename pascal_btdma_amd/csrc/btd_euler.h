// Sweep assembly for the euler3d ADI step (Fortran_Original/examples/euler3d/
// euler3d_mpi_gpu.f90): the block system of one sweep is built from the conserved field
// straight into the solve's layout, with the system numbering the reference produces by
// transposing (x sweep: sys = j + n2 k, rows i; y sweep: sys = i + n1 k, rows j; z sweep:
// sys = i + n1 j, rows k). This replaces build_matrix_{x,y,z}_kernel (:803-929) plus the
// transposes of dU, trA, trB, trC before the solve (:931-1002) and of dU after it
// (:959-974, :1004-1017).
//
// Fields are Fortran-ordered: U(0:n1+1, 0:n2+1, 0:n3+1, 5) with a one-cell halo,
// F(n1, n2, n3, m) without. One lane owns one system and walks a chunk of its rows, keeping
// the primitive variables of the previous, current and next cell in registers, so each row
// reads one new cell of U.
#pragma once
#include "btd_kernels.h"

namespace btd {

constexpr double kGamma = 1.4;  // euler3d_mpi_gpu.f90:15

struct Prim {
    double rho, u, v, w, E;
};

// The strides of the three grid axes (and of the variable index) in the haloed field.
struct Grid {
    int64_t n1, n2, n3;
    __device__ __forceinline__ int64_t sj() const { return n1 + 2; }
    __device__ __forceinline__ int64_t sk() const { return (n1 + 2) * (n2 + 2); }
    __device__ __forceinline__ int64_t sv() const { return (n1 + 2) * (n2 + 2) * (n3 + 2); }
};

// (i, j, k) (0-based interior indices) of system `sys`, row `row` of the sweep along AXIS.
template <int AXIS>
__device__ __forceinline__ void sweep_cell(const Grid& g, int64_t sys, int64_t row, int64_t& i,
                                           int64_t& j, int64_t& k) {
    if constexpr (AXIS == 0) {
        i = row;
        j = sys % g.n2;
        k = sys / g.n2;
    } else if constexpr (AXIS == 1) {
        j = row;
        i = sys % g.n1;
        k = sys / g.n1;
    } else {
        k = row;
        i = sys % g.n1;
        j = sys / g.n1;
    }
}

// primitive variables of one cell from its five conserved ones
__device__ __forceinline__ Prim prim_of(double rho, double ru, double rv, double rw, double rE) {
#pragma clang fp contract(off)
    Prim p;
    p.rho = rho;
    p.u = ru / rho;
    p.v = rv / rho;
    p.w = rw / rho;
    p.E = rE / rho;
    return p;
}

__device__ __forceinline__ Prim load_prim(const double* __restrict__ U, int64_t off, int64_t sv) {
    return prim_of(U[off], U[off + sv], U[off + 2 * sv], U[off + 3 * sv], U[off + 4 * sv]);
}

// Flux Jacobian along AXIS at one cell: compute_Amat (:749-765), compute_Bmat (:767-783),
// compute_Cmat (:785-801), J[iv][jv], the reference's operation order, no FMA contraction.
template <int AXIS>
__device__ __forceinline__ void euler_jacobian(const Prim& p, double (&J)[5][5]) {
#pragma clang fp contract(off)
    const double uu = p.u, vv = p.v, ww = p.w, E = p.E;
    const double q2 = uu * uu + vv * vv + ww * ww;
    const double H = E + (kGamma - 1.0) * (E - 0.5 * q2);
    const double c1 = kGamma - 1.0, c2 = 0.5 * c1;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) J[i][j] = 0.0;
    if constexpr (AXIS == 0) {
        J[0][1] = 1.0;
        J[1][0] = c2 * q2 - uu * uu; J[1][1] = (3.0 - kGamma) * uu;
        J[1][2] = -c1 * vv; J[1][3] = -c1 * ww; J[1][4] = c1;
        J[2][0] = -uu * vv; J[2][1] = vv; J[2][2] = uu;
        J[3][0] = -uu * ww; J[3][1] = ww; J[3][3] = uu;
        J[4][0] = (c2 * q2 - H) * uu; J[4][1] = H - c1 * (uu * uu);
        J[4][2] = -c1 * uu * vv; J[4][3] = -c1 * uu * ww; J[4][4] = kGamma * uu;
    } else if constexpr (AXIS == 1) {
        J[0][2] = 1.0;
        J[1][0] = -uu * vv; J[1][1] = vv; J[1][2] = uu;
        J[2][0] = c2 * q2 - vv * vv; J[2][1] = -c1 * uu;
        J[2][2] = (3.0 - kGamma) * vv; J[2][3] = -c1 * ww; J[2][4] = c1;
        J[3][0] = -vv * ww; J[3][2] = ww; J[3][3] = vv;
        J[4][0] = (c2 * q2 - H) * vv; J[4][1] = -c1 * uu * vv;
        J[4][2] = H - c1 * (vv * vv); J[4][3] = -c1 * vv * ww; J[4][4] = kGamma * vv;
    } else {
        J[0][3] = 1.0;
        J[1][0] = -uu * ww; J[1][1] = ww; J[1][3] = uu;
        J[2][0] = -vv * ww; J[2][2] = ww; J[2][3] = vv;
        J[3][0] = c2 * q2 - ww * ww; J[3][1] = -c1 * uu; J[3][2] = -c1 * vv;
        J[3][3] = (3.0 - kGamma) * ww; J[3][4] = c1;
        J[4][0] = (c2 * q2 - H) * ww; J[4][1] = -c1 * uu * ww;
        J[4][2] = -c1 * vv * ww; J[4][3] = H - c1 * (ww * ww); J[4][4] = kGamma * ww;
    }
}

// Row blocks of the sweep from the Jacobians of the previous and next cell:
//   A = -cc J_prev - cd I,  B = (1 + 2 cd) I,  C = cc J_next - cd I  (build_matrix_*, :836-848).
__device__ __forceinline__ void euler_blocks(const double (&Jl)[5][5], const double (&Jr)[5][5],
                                             double cc, double cd, double (&A)[5][5],
                                             double (&B)[5][5], double (&C)[5][5]) {
#pragma clang fp contract(off)
    const double bd = 1.0 + 2.0 * cd;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const double a = -cc * Jl[i][j], c = cc * Jr[i][j];
            A[i][j] = i == j ? a - cd : a;
            B[i][j] = i == j ? bd : 0.0;
            C[i][j] = i == j ? c - cd : c;
        }
}

// Assembly: rows [r0, r0 + R) of 64 systems per block (blockIdx.y = row chunk). Writes A, B,
// C (and, when F is given, d gathered from the field F) in the solve's LAYOUT.
template <int LAYOUT, int AXIS>
__global__ __launch_bounds__(64) void k_euler_assemble(const double* __restrict__ U,
                                                       const double* __restrict__ F, Grid g,
                                                       double cc, double cd, double* __restrict__ A,
                                                       double* __restrict__ B, double* __restrict__ C,
                                                       double* __restrict__ D, int64_t nsys,
                                                       int64_t nrow, int R) {
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) return;
    const int64_t r0 = (int64_t)blockIdx.y * R;
    const int64_t r1 = r0 + R < nrow ? r0 + R : nrow;
    const Addr<LAYOUT> am(sys, nsys, nrow, 25), av(sys, nsys, nrow, 5);
    const int64_t sv = g.sv();
    const int64_t step = AXIS == 0 ? 1 : (AXIS == 1 ? g.sj() : g.sk());
    int64_t i, j, k;
    sweep_cell<AXIS>(g, sys, r0, i, j, k);
    int64_t off = (i + 1) + (j + 1) * g.sj() + (k + 1) * g.sk();  // cell of row r0
    const int64_t fstep = AXIS == 0 ? 1 : (AXIS == 1 ? g.n1 : g.n1 * g.n2);
    const int64_t fsv = g.n1 * g.n2 * g.n3;
    int64_t foff = i + g.n1 * (j + g.n2 * k);
    Prim prev = load_prim(U, off - step, sv), cur = load_prim(U, off, sv);
    for (int64_t r = r0; r < r1; ++r) {
        const Prim next = load_prim(U, off + step, sv);
        double Jl[5][5], Jr[5][5], Ab[5][5], Bb[5][5], Cb[5][5];
        euler_jacobian<AXIS>(prev, Jl);
        euler_jacobian<AXIS>(next, Jr);
        euler_blocks(Jl, Jr, cc, cd, Ab, Bb, Cb);
        store_mat<5>(A, am, r, Ab);
        store_mat<5>(B, am, r, Bb);
        store_mat<5>(C, am, r, Cb);
        if (F) {
            double v[5];
#pragma unroll
            for (int e = 0; e < 5; ++e) v[e] = F[foff + e * fsv];
            store_vec<5>(D, av, r, v);
        }
        prev = cur;
        cur = next;
        off += step;
        foff += fstep;
    }
}

// The x sweep's assembly: its lines are the field's contiguous direction, so a lane walking its
// own line would touch a separate cache line per lane and load. The workgroup's 64 lines are
// staged through LDS instead, kXSub rows at a time (U cells r-1 .. r+kXSub, F cells r ..
// r+kXSub-1 of each line), consecutive lanes loading consecutive cells of a line; each lane then
// builds its rows from the tiles exactly as k_euler_assemble does (same operations, same result).
constexpr int kXSub = 8;
template <int LAYOUT>
__global__ __launch_bounds__(64) void k_euler_assemble_x(const double* __restrict__ U,
                                                         const double* __restrict__ F, Grid g,
                                                         double cc, double cd, double* __restrict__ A,
                                                         double* __restrict__ B, double* __restrict__ C,
                                                         double* __restrict__ D, int64_t nsys,
                                                         int64_t nrow, int R) {
    constexpr int kUc = kXSub + 2;  // U cells per line and sub-chunk
    __shared__ double tu[5][64][kUc + 1];
    __shared__ double tf[5][64][kXSub + 1];
    const int t = threadIdx.x;
    const int64_t s0 = (int64_t)blockIdx.x * 64, sys = s0 + t;
    const int lines = nsys - s0 < 64 ? (int)(nsys - s0) : 64;
    const int64_t r0 = (int64_t)blockIdx.y * R;
    const int64_t r1 = r0 + R < nrow ? r0 + R : nrow;
    const Addr<LAYOUT> am(sys, nsys, nrow, 25), av(sys, nsys, nrow, 5);
    const int64_t sv = g.sv(), fsv = g.n1 * g.n2 * g.n3;
    const unsigned n2 = (unsigned)g.n2;
    for (int64_t rs = r0; rs < r1; rs += kXSub) {
        const int cnt = r1 - rs < kXSub ? (int)(r1 - rs) : kXSub;
        __syncthreads();  // the previous sub-chunk's tiles have been read
        for (int idx = t; idx < lines * kUc; idx += 64) {
            const int l = idx / kUc, c = idx - l * kUc;
            if (c > cnt + 1) continue;
            const unsigned ls = (unsigned)(s0 + l);
            const int64_t j = ls % n2, k = ls / n2;
            const int64_t off = (rs + c) + (j + 1) * g.sj() + (k + 1) * g.sk();  // interior cell rs-1+c
#pragma unroll
            for (int e = 0; e < 5; ++e) tu[e][l][c] = U[off + e * sv];
        }
        if (F) {
            for (int idx = t; idx < lines * kXSub; idx += 64) {
                const int l = idx / kXSub, c = idx - l * kXSub;
                if (c >= cnt) continue;
                const unsigned ls = (unsigned)(s0 + l);
                const int64_t foff = (rs + c) + g.n1 * (int64_t)ls;  // i + n1 (j + n2 k), ls = j + n2 k
#pragma unroll
                for (int e = 0; e < 5; ++e) tf[e][l][c] = F[foff + e * fsv];
            }
        }
        __syncthreads();
        if (sys < nsys) {
            auto cell = [&](int c) { return prim_of(tu[0][t][c], tu[1][t][c], tu[2][t][c], tu[3][t][c], tu[4][t][c]); };
            Prim prev = cell(0), cur = cell(1);
            for (int q = 0; q < cnt; ++q) {
                const Prim next = cell(q + 2);
                double Jl[5][5], Jr[5][5], Ab[5][5], Bb[5][5], Cb[5][5];
                euler_jacobian<0>(prev, Jl);
                euler_jacobian<0>(next, Jr);
                euler_blocks(Jl, Jr, cc, cd, Ab, Bb, Cb);
                store_mat<5>(A, am, rs + q, Ab);
                store_mat<5>(B, am, rs + q, Bb);
                store_mat<5>(C, am, rs + q, Cb);
                if (F) {
                    double v[5];
#pragma unroll
                    for (int e = 0; e < 5; ++e) v[e] = tf[e][t][q];
                    store_vec<5>(D, av, rs + q, v);
                }
                prev = cur;
                cur = next;
            }
        }
    }
}

// Matrix-free periodic sweep: the solve of btdma_many_cycl (mod_btdma_cpu.f90:80-183;
// mod_btdma_gpu_v2.f90:352-616) on the sweep's system, with every block computed from U where
// the solve needs it, so A, B, C are never stored. Row 0 couples to row n-1 (periodic halo);
// x_0 is kept as a separate unknown and rows 1..n-1 are eliminated with the fill-in column E
// (E_1 = -A_1, E_{n-1} = -C_{n-1}):
//   forward, q = 1..n-1:  R = B_q - A_q C'_{q-1},
//                         [d'_q | E'_q | C'_q] = R^{-1} [rhs_q - A_q d'_{q-1} | X_q - A_q E'_{q-1} | C_q]
//   so that x_q = d'_q + E'_q x_0 - C'_q x_{q+1} (C'_{n-1} = 0). Row 1 expressed through x_0
//   alone, x_1 = alpha + beta x_0, is composed during the same sweep (alpha, beta, gamma start as
//   d'_1, E'_1, -C'_1 and take alpha += gamma d'_q, beta += gamma E'_q, gamma = -gamma C'_q), with
//   the three kept in LDS. Then row 0:
//   (B_0 + A_0 E'_{n-1} + C_0 beta) x_0 = rhs_0 - A_0 d'_{n-1} - C_0 alpha,
//   and one backward sweep x_q = d'_q + E'_q x_0 - C'_q x_{q+1} writes x into the field F.
// C', E', d' of rows 1..n-2 go to the workspace (Wc, We, Wd, the plan's layout); each lane reads
// its own field cells before writing them, so F may be the rhs field itself. Needs n >= 3.
// LINES = true reads U and rhs from the sweep's lines (U with its two halo rows, rows 0..n+1, and
// rhs rows 0..n-1, both in LAYOUT, E = 5; filled by k_xlines) and writes x to the rhs lines: a
// lane reads each row of rhs / d' before it writes that row, so rhs, Wd and F may be one buffer.
template <int AXIS>
__device__ __forceinline__ void euler_row_blocks(const Prim& prev, const Prim& next, double cc,
                                                 double cd, double (&A)[5][5], double (&C)[5][5]) {
    double Jl[5][5], Jr[5][5], Bb[5][5];
    euler_jacobian<AXIS>(prev, Jl);
    euler_jacobian<AXIS>(next, Jr);
    euler_blocks(Jl, Jr, cc, cd, A, Bb, C);
}

template <int LAYOUT, int AXIS, bool LINES>
__global__ __launch_bounds__(64) void k_euler_sweep_cyclic(const double* __restrict__ U,
                                                           const double* rhs, Grid g, double cc,
                                                           double cd, double* __restrict__ Wc,
                                                           double* __restrict__ We, double* Wd,
                                                           double* F, int64_t nsys, int64_t nrow,
                                                           int* __restrict__ flag) {
    constexpr int M = 5;
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) return;
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
    const Addr<LAYOUT> au(sys, nsys, nrow + 2, M);  // LINES: U's line, rows 0..nrow+1
    __shared__ double sp[(M + 2 * M * M) * 64];  // alpha, beta, gamma: [element][lane]
    const int lane = threadIdx.x;
    auto SA = [&](int i) -> double& { return sp[i * 64 + lane]; };
    auto SB = [&](int i, int j) -> double& { return sp[(M + i * M + j) * 64 + lane]; };
    auto SG = [&](int i, int j) -> double& { return sp[(M + M * M + i * M + j) * 64 + lane]; };
    const double bd = 1.0 + 2.0 * cd;
    const int64_t sv = g.sv();
    const int64_t step = AXIS == 0 ? 1 : (AXIS == 1 ? g.sj() : g.sk());
    const int64_t fstep = AXIS == 0 ? 1 : (AXIS == 1 ? g.n1 : g.n1 * g.n2);
    const int64_t fsv = g.n1 * g.n2 * g.n3;
    int64_t i0, j0, k0;
    sweep_cell<AXIS>(g, sys, 0, i0, j0, k0);
    const int64_t off0 = (i0 + 1) + (j0 + 1) * g.sj() + (k0 + 1) * g.sk();  // cell of row 0
    const int64_t foff0 = i0 + g.n1 * (j0 + g.n2 * k0);
    // the cell of row q (q = -1..nrow) and element e of row q of rhs / x
    auto prim = [&](int64_t q) -> Prim {
        if constexpr (LINES) return load_prim(U, au.at(q + 1, 0), au.estride());
        return load_prim(U, off0 + q * step, sv);
    };
    auto fo = [&](int64_t q, int e) -> int64_t {
        if constexpr (LINES) return av.at(q, e);
        return foff0 + q * fstep + e * fsv;
    };
    bool ok = true;
    double Cp[M][M], Ep[M][M], dp[M];
    Prim prev = prim(0), cur = prim(1);
    for (int64_t q = 1; q < nrow; ++q) {
        const bool last = q == nrow - 1;
        const Prim next = prim(q + 1);
        double A[M][M], C[M][M], R[M][M], X[M][M], v[M];
        euler_row_blocks<AXIS>(prev, next, cc, cd, A, C);
#pragma unroll
        for (int e = 0; e < M; ++e) v[e] = rhs[fo(q, e)];
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) {
                R[i][j] = i == j ? bd : 0.0;
                X[i][j] = q == 1 ? -A[i][j] : (last ? -C[i][j] : 0.0);
            }
        if (q > 1) {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    const double a = A[i][k];
                    v[i] = fma(-a, dp[k], v[i]);
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        R[i][j] = fma(-a, Cp[k][j], R[i][j]);
                        X[i][j] = fma(-a, Ep[k][j], X[i][j]);
                    }
                }
        }
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, v);
        solve_rows<M>(R, piv, X);
        if (!last) {
            solve_rows<M>(R, piv, C);
        } else {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) C[i][j] = 0.0;
        }
        // x_1 = alpha + beta x_0 + gamma x_{q+1}
        if (q == 1) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                SA(i) = v[i];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    SB(i, j) = X[i][j];
                    SG(i, j) = -C[i][j];
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                double gi[M];
#pragma unroll
                for (int k = 0; k < M; ++k) gi[k] = SG(i, k);
                double a = SA(i);
#pragma unroll
                for (int k = 0; k < M; ++k) a = fma(gi[k], v[k], a);
                SA(i) = a;
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    double b = SB(i, j), c = 0.0;
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        b = fma(gi[k], X[k][j], b);
                        c = fma(-gi[k], C[k][j], c);
                    }
                    SB(i, j) = b;
                    SG(i, j) = c;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            dp[i] = v[i];
#pragma unroll
            for (int j = 0; j < M; ++j) {
                Cp[i][j] = C[i][j];
                Ep[i][j] = X[i][j];
            }
        }
        if (!last) {
            store_mat<M>(Wc, am, q, Cp);
            store_mat<M>(We, am, q, Ep);
            store_vec<M>(Wd, av, q, dp);
        }
        prev = cur;
        cur = next;
    }
    // row 0: (Ep, dp) hold row n-1
    double x0[M];
    {
        double A0[M][M], C0[M][M], R[M][M];
        const Prim pl = prim(-1), pr = prim(1);
        euler_row_blocks<AXIS>(pl, pr, cc, cd, A0, C0);
#pragma unroll
        for (int e = 0; e < M; ++e) x0[e] = rhs[fo(0, e)];
#pragma unroll
        for (int i = 0; i < M; ++i) {
#pragma unroll
            for (int k = 0; k < M; ++k)
                x0[i] = fma(-A0[i][k], dp[k], fma(-C0[i][k], SA(k), x0[i]));
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double s = i == j ? bd : 0.0;
#pragma unroll
                for (int k = 0; k < M; ++k) s = fma(A0[i][k], Ep[k][j], fma(C0[i][k], SB(k, j), s));
                R[i][j] = s;
            }
        }
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, x0);
    }
    // backward: x_{n-1} = d'_{n-1} + E'_{n-1} x_0, then x_q = d'_q + E'_q x_0 - C'_q x_{q+1}
    double xn[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double s = dp[i];
#pragma unroll
        for (int k = 0; k < M; ++k) s = fma(Ep[i][k], x0[k], s);
        xn[i] = s;
    }
#pragma unroll
    for (int e = 0; e < M; ++e) F[fo(nrow - 1, e)] = xn[e];
    for (int64_t q = nrow - 2; q >= 1; --q) {
        double c[M][M], e_[M][M], d[M];
        load_mat_c<M>(Wc, am, q, c);
        load_mat_c<M>(We, am, q, e_);
        load_vec<M>(Wd, av, q, d);
        double x[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            double s = d[i];
#pragma unroll
            for (int k = 0; k < M; ++k) s = fma(e_[i][k], x0[k], fma(-c[i][k], xn[k], s));
            x[i] = s;
        }
#pragma unroll
        for (int e = 0; e < M; ++e) {
            F[fo(q, e)] = x[e];
            xn[e] = x[e];
        }
    }
#pragma unroll
    for (int e = 0; e < M; ++e) F[fo(0, e)] = x0[e];
    if (!ok) report_singular(flag, sys);
}

// x-line transposes. Along x the lines are the field's contiguous direction, so lanes that own
// consecutive x lines touch a different cache line each and refetch it after eviction. These
// kernels move a field to (TO_LINES) or from the x sweep's order in LAYOUT through 64 x 64 LDS
// tiles: a 256-thread block owns 64 lines (one tile of systems) and 64 consecutive x positions,
// reads each line's 64 doubles as one contiguous run and writes each row's 64 lanes as one
// contiguous run. Field geometry: element (x, line, v) of the field is at
//   x + hx + sj * (j + h) + sk * (k + h) + sv * v,   line = j + n2 * k,
// with halo offsets hx (x direction) and h (j, k); the lines hold L positions per line.
struct XField {
    int64_t n2, sj, sk, sv;
    int h, hx;
};

template <int LAYOUT, bool TO_LINES>
__global__ __launch_bounds__(256) void k_xlines(double* __restrict__ F, double* __restrict__ Ln,
                                                XField f, int nv, int64_t nsys, int64_t L) {
    __shared__ double tile[64][65];
    const int t = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t sys0 = (int64_t)blockIdx.x * 64;
    const int64_t x0 = (int64_t)blockIdx.y * 64;
    for (int v = 0; v < nv; ++v) {
        if constexpr (TO_LINES) {
            // read: lane t = x position, 16 lines per wave
            for (int l = w; l < 64; l += 4) {
                const int64_t sys = sys0 + l, x = x0 + t;
                double val = 0.0;
                if (sys < nsys && x < L) {
                    const int64_t j = sys % f.n2, k = sys / f.n2;
                    val = F[x + f.hx + f.sj * (j + f.h) + f.sk * (k + f.h) + f.sv * v];
                }
                tile[t][l] = val;
            }
            __syncthreads();
            // write: lane t = line, 16 x positions per wave
            for (int xx = w; xx < 64; xx += 4) {
                const int64_t x = x0 + xx, sys = sys0 + t;
                if (x < L && sys < nsys) {
                    const Addr<LAYOUT> a(sys, nsys, L, nv);
                    Ln[a.at(x, v)] = tile[xx][t];
                }
            }
        } else {
            for (int xx = w; xx < 64; xx += 4) {
                const int64_t x = x0 + xx, sys = sys0 + t;
                double val = 0.0;
                if (x < L && sys < nsys) {
                    const Addr<LAYOUT> a(sys, nsys, L, nv);
                    val = Ln[a.at(x, v)];
                }
                tile[xx][t] = val;
            }
            __syncthreads();
            for (int l = w; l < 64; l += 4) {
                const int64_t sys = sys0 + l, x = x0 + t;
                if (sys < nsys && x < L) {
                    const int64_t j = sys % f.n2, k = sys / f.n2;
                    F[x + f.hx + f.sj * (j + f.h) + f.sk * (k + f.h) + f.sv * v] = tile[t][l];
                }
            }
        }
        __syncthreads();
    }
}

// Field F(n1, n2, n3, m) <-> sweep vector in LAYOUT (SCATTER = false: gather into D,
// true: scatter D into F). Rows [r0, r0 + R) of 64 systems per block.
template <int LAYOUT, int AXIS, bool SCATTER>
__global__ __launch_bounds__(64) void k_sweep_move(double* __restrict__ F, double* __restrict__ D,
                                                   Grid g, int m, int64_t nsys, int64_t nrow,
                                                   int R) {
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) return;
    const int64_t r0 = (int64_t)blockIdx.y * R;
    const int64_t r1 = r0 + R < nrow ? r0 + R : nrow;
    const Addr<LAYOUT> av(sys, nsys, nrow, m);
    int64_t i, j, k;
    sweep_cell<AXIS>(g, sys, r0, i, j, k);
    const int64_t fstep = AXIS == 0 ? 1 : (AXIS == 1 ? g.n1 : g.n1 * g.n2);
    const int64_t fsv = g.n1 * g.n2 * g.n3;
    int64_t foff = i + g.n1 * (j + g.n2 * k);
    for (int64_t r = r0; r < r1; ++r) {
        for (int e = 0; e < m; ++e) {
            if constexpr (SCATTER)
                F[foff + e * fsv] = D[av.at(r, e)];
            else
                D[av.at(r, e)] = F[foff + e * fsv];
        }
        foff += fstep;
    }
}

}  // namespace btd
