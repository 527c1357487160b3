// gfx950 kernel templates of the batched block-Thomas solve, FP64 (instantiated per m
// in btd_inst.hip).
//
// One lane owns one system (or one row partition of one system) and walks its rows in
// order; the m x m blocks of the current row live in VGPRs. Consecutive lanes own
// consecutive systems, so with either the reference GPU layout or the tiled layout a
// wavefront's load of one block element is 64 consecutive doubles (512 B).
//
// Kernels:
//   k_plain      single-partition solve, btdma_many (mod_btdma_cpu.f90:15-78;
//                GPU: mod_btdma_gpu_v2.f90:241-350, pascal_btdma_cuda.cu:274-334).
//   k_modi       stage 1 of the two-stage solve over one partition per lane,
//                btdma_many_modi (mod_btdma_cpu.f90:305-443; mod_btdma_gpu_v2.f90:618-866).
//   k_reduced    plain solve of the 2P-row reduced system (B = I), on-device;
//   k_reduced_twisted  the same system solved twisted by two groups of m lanes per system
//                (btd_coop.h), for few systems with many partitions.
//   k_plain_twisted  the single-partition solve of a tiled plan, twisted, m lanes per half
//                system (btd_coop.h), for plans with too few systems to fill the device.
//   k_update     stage 3, btdma_many_update (mod_btdma_cpu.f90:511-535).
//   k_generate   seeded input generator (bench and tests).
#pragma once
#include <type_traits>

#include "btd_coop.h"
#include "btd_internal.h"

namespace btd {

// ------------------------------------------------------------------------ addressing --
// Offset of element e of (sys,row) is base + row*rs + e*es. For the tiled layout the
// element stride is the compile-time constant 64.
template <int LAYOUT>
struct Addr {
    int64_t base, rs, es;
    __device__ __forceinline__ Addr(int64_t sys, int64_t nsys, int64_t nrow, int64_t E) {
        if constexpr (LAYOUT == LAYOUT_TILED) {
            base = (sys / kTile) * nrow * E * kTile + (sys % kTile);
            rs = E * kTile;
            es = kTile;
        } else {
            base = sys;
            rs = nsys;
            es = nsys * nrow;
        }
    }
    __device__ __forceinline__ int64_t estride() const {
        if constexpr (LAYOUT == LAYOUT_TILED) return kTile;
        return es;
    }
    __device__ __forceinline__ int64_t at(int64_t row, int e) const {
        return base + row * rs + (int64_t)e * estride();
    }
};

// ----------------------------------------------------------------------- row helpers --
// Lanes nsys <= sys < 64*ceil(nsys/64) of the tiled layout own no system. Kernels whose
// outputs callers reduce over the whole flat array (gradients, solutions of the
// differentiable solves) write zeros to those slots, so the padding never holds stale
// memory. No-op for the reference GPU layout, which has no padding.
template <int E, int LAYOUT, typename T>
__device__ __forceinline__ void zero_pad_lane(T* p, int64_t sys, int64_t nsys, int64_t nrow) {
    if constexpr (LAYOUT == LAYOUT_TILED) {
        const Addr<LAYOUT> a(sys, nsys, nrow, E);
        for (int64_t q = 0; q < nrow; ++q)
#pragma unroll
            for (int e = 0; e < E; ++e) p[a.at(q, e)] = T(0);
    }
}

// Block loads take FP64 or FP32 storage (T) and return FP64 values.
template <int M, int LAYOUT, typename T>
__device__ __forceinline__ void load_mat(const T* __restrict__ p, const Addr<LAYOUT>& a,
                                         int64_t row, double (&X)[M][M]) {
    const T* r = p + a.at(row, 0);
    const int64_t es = a.estride();
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int i = 0; i < M; ++i) X[i][j] = __builtin_nontemporal_load(r + (i + M * j) * es);
}

template <int M, int LAYOUT, typename T>
__device__ __forceinline__ void load_col(const T* p, const Addr<LAYOUT>& a, int64_t row,
                                         int j, double (&x)[M]) {
    const T* r = p + a.at(row, 0);
    const int64_t es = a.estride();
#pragma unroll
    for (int i = 0; i < M; ++i) x[i] = r[(i + M * j) * es];
}

template <int M, int LAYOUT>
__device__ __forceinline__ void store_col(double* p, const Addr<LAYOUT>& a, int64_t row, int j,
                                          const double (&x)[M]) {
    double* r = p + a.at(row, 0);
    const int64_t es = a.estride();
#pragma unroll
    for (int i = 0; i < M; ++i) r[(i + M * j) * es] = x[i];
}

template <int M, int LAYOUT>
__device__ __forceinline__ void load_vec(const double* p, const Addr<LAYOUT>& a, int64_t row,
                                         double (&x)[M]) {
    const double* r = p + a.at(row, 0);
    const int64_t es = a.estride();
#pragma unroll
    for (int i = 0; i < M; ++i) x[i] = r[i * es];
}

template <int M, int LAYOUT>
__device__ __forceinline__ void store_vec(double* p, const Addr<LAYOUT>& a, int64_t row,
                                          const double (&x)[M]) {
    double* r = p + a.at(row, 0);
    const int64_t es = a.estride();
#pragma unroll
    for (int i = 0; i < M; ++i) r[i * es] = x[i];
}

// R <- R - A*Y and v <- v - A*y, A streamed by columns.
template <int M, int LAYOUT, typename T>
__device__ __forceinline__ void gemm_sub(const T* __restrict__ Ap, const Addr<LAYOUT>& am,
                                         int64_t row, const double (&Y)[M][M], const double (&y)[M],
                                         double (&R)[M][M], double (&v)[M]) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
        double ak[M];
        load_col<M>(Ap, am, row, k, ak);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            v[i] = fma(-ak[i], y[k], v[i]);
#pragma unroll
            for (int j = 0; j < M; ++j) R[i][j] = fma(-ak[i], Y[k][j], R[i][j]);
        }
    }
}

// The same update in the oracle's order (mod_btdma_cpu.f90:49-53 through dgemm/dgemv): the
// products A*Y and A*y are summed over k from the first term, unfused, and subtracted at the
// end: R <- R - (A*Y), v <- v - (A*y).
template <int M, int LAYOUT, typename T>
__device__ __forceinline__ void gemm_sub_ref(const T* __restrict__ Ap, const Addr<LAYOUT>& am,
                                             int64_t row, const double (&Y)[M][M], const double (&y)[M],
                                             double (&R)[M][M], double (&v)[M]) {
#pragma clang fp contract(off)
    double S[M][M], t[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        double ak[M];
        load_col<M>(Ap, am, row, k, ak);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            t[i] = k == 0 ? ak[i] * y[0] : t[i] + ak[i] * y[k];
#pragma unroll
            for (int j = 0; j < M; ++j) S[i][j] = k == 0 ? ak[i] * Y[0][j] : S[i][j] + ak[i] * Y[k][j];
        }
    }
#pragma unroll
    for (int i = 0; i < M; ++i) {
        v[i] = v[i] - t[i];
#pragma unroll
        for (int j = 0; j < M; ++j) R[i][j] = R[i][j] - S[i][j];
    }
}

template <int M>
__device__ __forceinline__ void solve_rows(const double (&R)[M][M], const int (&piv)[M],
                                           double (&X)[M][M]) {
#pragma unroll
    for (int j = 0; j < M; ++j) {
        double c[M];
#pragma unroll
        for (int i = 0; i < M; ++i) c[i] = X[i][j];
        lu_solve<M>(R, piv, c);
#pragma unroll
        for (int i = 0; i < M; ++i) X[i][j] = c[i];
    }
}

template <int M, int LAYOUT>
__device__ __forceinline__ void store_mat(double* p, const Addr<LAYOUT>& a, int64_t row,
                                          const double (&X)[M][M]) {
    double* r = p + a.at(row, 0);
    const int64_t es = a.estride();
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int i = 0; i < M; ++i) r[(i + M * j) * es] = X[i][j];
}

template <int M, int LAYOUT>
__device__ __forceinline__ void load_mat_c(const double* p, const Addr<LAYOUT>& a, int64_t row,
                                           double (&X)[M][M]) {
    const double* r = p + a.at(row, 0);
    const int64_t es = a.estride();
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int i = 0; i < M; ++i) X[i][j] = r[(i + M * j) * es];
}

template <int M>
__device__ __forceinline__ void mm(const double (&X)[M][M], const double (&Y)[M][M],
                                   double (&Z)[M][M]) {
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double s = X[i][0] * Y[0][j];
#pragma unroll
            for (int k = 1; k < M; ++k) s = fma(X[i][k], Y[k][j], s);
            Z[i][j] = s;
        }
}

template <int M>
__device__ __forceinline__ void mv(const double (&X)[M][M], const double (&y)[M], double (&z)[M]) {
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double s = X[i][0] * y[0];
#pragma unroll
        for (int k = 1; k < M; ++k) s = fma(X[i][k], y[k], s);
        z[i] = s;
    }
}

// ----------------------------------------------------------------------------- plain --
// Forward: [d'_q | C'_q] = (B_q - A_q C'_{q-1})^{-1} [d_q - A_q d'_{q-1} | C_q], written in
// place over C_q and d_q; backward: x_q = d'_q - C'_q x_{q+1}. A and B are only read.
// SAVE additionally keeps, per row, the LU factors of R_q = B_q - A_q C'_{q-1} (matrix
// layout, reciprocal pivots on the diagonal) and the pivot rows packed 4 bits per step in
// one 64-bit word (vector layout with one element per row), for the adjoint solve.
template <int M>
__device__ __forceinline__ uint64_t pack_piv(const int (&piv)[M]) {
    uint64_t w = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) w |= (uint64_t)piv[k] << (4 * k);
    return w;
}

template <int M>
__device__ __forceinline__ void unpack_piv(uint64_t w, int (&piv)[M]) {
#pragma unroll
    for (int k = 0; k < M; ++k) piv[k] = (int)((w >> (4 * k)) & 15u);
}

// x <- x - (C x_next) with the product summed over j from its first term, unfused: the
// oracle's back substitution (mod_btdma_cpu.f90:70-75 through dgemv).
template <int M>
__device__ __forceinline__ void back_sub_ref(const double (&c)[M][M], const double (&xn)[M], double (&x)[M]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double s = c[i][0] * xn[0];
#pragma unroll
        for (int j = 1; j < M; ++j) s = s + c[i][j] * xn[j];
        x[i] = x[i] - s;
    }
}

// A zero pivot in system `sys`: raise the plan's flag (word 0) and the system's own word
// (1 + sys). Words only go from 0 to 1, so concurrent writers of one system agree.
__device__ __forceinline__ void report_singular(int* flag, int64_t sys) {
    atomicOr(flag, 1);
    flag[1 + sys] = 1;
}

// Blocks stored as TB (double, or float for the FP32-storage solve, which computes in FP64);
// C'_q goes to Cw (C itself for the in-place FP64 solve); d is read from Din and d', then x,
// written to D (Din == D in place). ORD: every block operation in the
// oracle's order and rounding (gemm_sub_ref, lu_factor_ref, lu_solve_ref, and the back
// substitution as x_q - (C'_q x_{q+1}) with the product summed from its first term, unfused),
// so x equals the oracle's btdma_many bit for bit.
template <int M, int LAYOUT, bool SAVE, typename TB, bool ORD = false>
__device__ __forceinline__ void plain_solve(const TB* __restrict__ A, const TB* __restrict__ B,
                                            const TB* C, double* Cw, const double* Din, double* D,
                                            double* __restrict__ F, uint64_t* __restrict__ PV,
                                            int nsys, int nrow, int* __restrict__ flag) {
    static_assert(!(SAVE && ORD), "the adjoint's saved factors use the default order");
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) {
        // out of place (the forward of the differentiable solve), the outputs' padding lanes of
        // the tiled layout get zeros instead of whatever their allocation held
        if constexpr (SAVE) {
            if ((const void*)D != (const void*)Din) zero_pad_lane<M, LAYOUT>(D, sys, nsys, nrow);
            if ((const void*)Cw != (const void*)C) zero_pad_lane<M * M, LAYOUT>(Cw, sys, nsys, nrow);
        }
        return;
    }
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M), ap(sys, nsys, nrow, 1);
    double Cp[M][M], dp[M];
    bool ok = true;
    for (int64_t q = 0; q < nrow; ++q) {
        double R[M][M], v[M];
        load_mat<M>(B, am, q, R);
        load_vec<M>(Din, av, q, v);
        if (q > 0) {
            if constexpr (ORD)
                gemm_sub_ref<M>(A, am, q, Cp, dp, R, v);
            else
                gemm_sub<M>(A, am, q, Cp, dp, R, v);
        }
        int piv[M];
        if constexpr (ORD)
            ok = lu_factor_ref<M>(R, piv) && ok;
        else
            ok = lu_factor<M>(R, piv) && ok;
        if constexpr (SAVE) {
            double* r = F + am.at(q, 0);
#pragma unroll
            for (int j = 0; j < M; ++j)
#pragma unroll
                for (int i = 0; i < M; ++i) r[(i + M * j) * am.estride()] = R[i][j];
            PV[ap.at(q, 0)] = pack_piv<M>(piv);
        }
        if constexpr (ORD)
            lu_solve_ref<M>(R, piv, v);
        else
            lu_solve<M>(R, piv, v);
#pragma unroll
        for (int i = 0; i < M; ++i) dp[i] = v[i];
        store_vec<M>(D, av, q, dp);
        if (q + 1 < nrow) {
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double c[M];
                load_col<M>(C, am, q, j, c);
                if constexpr (ORD)
                    lu_solve_ref<M>(R, piv, c);
                else
                    lu_solve<M>(R, piv, c);
#pragma unroll
                for (int i = 0; i < M; ++i) Cp[i][j] = c[i];
                store_col<M>(Cw, am, q, j, c);
            }
        } else if constexpr (SAVE) {
            // the last row has no C' (its C is ignored); out of place its block gets zeros
            if ((const void*)Cw != (const void*)C) {
                const double z[M][M] = {};
                store_mat<M>(Cw, am, q, z);
            }
        }
    }
    // back substitution; dp holds x_{n-1}
    for (int64_t q = nrow - 2; q >= 0; --q) {
        double x[M], c[M][M];
        load_vec<M>(D, av, q, x);
        load_mat<M>(static_cast<const double*>(Cw), am, q, c);
        if constexpr (ORD) {
            back_sub_ref<M>(c, dp, x);
        } else {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) x[i] = fma(-c[i][j], dp[j], x[i]);
        }
#pragma unroll
        for (int i = 0; i < M; ++i) dp[i] = x[i];
        store_vec<M>(D, av, q, x);
    }
    if (!ok) report_singular(flag, sys);
}

// The one-lane solve computes in the oracle's operation order (ORD): measured at the same time
// as the fused order on the m = 5, n = 512, 131072-system batch (14.90 / 14.90 ms Euler-like,
// 14.70 / 14.71 ms on conditioned blocks), it returns the oracle's x bit for bit.
// C'_q goes to Cw: C itself for the in-place solve, a plan's workspace for BTD_FORM_WORKSPACE
// (C then only read).
template <int M, int LAYOUT>
__global__ __launch_bounds__(64) void k_plain(const double* __restrict__ A,
                                              const double* __restrict__ B, const double* C, double* Cw,
                                              double* D, int nsys, int nrow, int* __restrict__ flag) {
    plain_solve<M, LAYOUT, false, double, true>(A, B, C, Cw, D, D, nullptr, nullptr, nsys, nrow, flag);
}

// R <- R - X*Y and v <- v - X*y with X in registers: gemm_sub's operations in its order.
template <int M>
__device__ __forceinline__ void gemm_sub_reg(const double (&X)[M][M], const double (&Y)[M][M],
                                             const double (&y)[M], double (&R)[M][M], double (&v)[M]) {
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
        for (int i = 0; i < M; ++i) {
            v[i] = fma(-X[i][k], y[k], v[i]);
#pragma unroll
            for (int j = 0; j < M; ++j) R[i][j] = fma(-X[i][k], Y[k][j], R[i][j]);
        }
}

// Single-partition solve with two lanes per system, one from each end, blocks in registers:
// the two wavefronts of a 128-thread workgroup take the same 64 systems (one tile). With
// h = (n-1)/2, the top wavefront eliminates rows 0..h-1 downwards as k_plain does (C'_q =
// R_q^{-1} C_q over C_q, d'_q over d_q), the bottom one rows n-1..h+1 upwards (S_q = B_q -
// C_q A''_{q+1}, A''_q = S_q^{-1} A_q over C_q, which row q has just consumed, d''_q =
// S_q^{-1}(d_q - C_q d''_{q+1}) over d_q). They exchange their carried block and vector
// through LDS, both solve the meeting row
//   (B_h - A_h C'_{h-1} - C_h A''_{h+1}) x_h = d_h - A_h d'_{h-1} - C_h d''_{h+1}
// by the same operations in the same order, and substitute outwards: x_q = d'_q - C'_q x_{q+1}
// (q < h), x_q = d''_q - A''_q x_{q-1} (q > h). Half the dependent rows of k_plain per lane for
// the same bytes (C' / A'' and d' / d'' written and read back once, A not written): for batches
// too small to fill the device one lane per system. A of row 0 and C of row n-1 are not used.
template <int M, int LAYOUT, bool AHEAD, bool LIGHT = false>
__global__ __launch_bounds__(2 * kTile, (AHEAD || LIGHT) ? 1 : 2) void k_plain_halves(const double* __restrict__ A,
                                                                          const double* __restrict__ B,
                                                                          double* C, double* D, int nsys,
                                                                          int nrow, int* __restrict__ flag) {
    constexpr int X = M * M + M;  // LDS slots of a wavefront's carry
    __shared__ double xs[2][X * kTile];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile);  // 0 top, 1 bottom
    const int lane = threadIdx.x % kTile;
    const bool top = wv == 0;
    const int64_t sys = (int64_t)blockIdx.x * kTile + lane;
    if (sys >= nsys) return;  // a system's two lanes leave together: both wavefronts map lanes alike
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
    const int h = (nrow - 1) / 2;
    const int L = top ? h : nrow - 1 - h;  // rows this lane eliminates before the meeting
    const int64_t g0 = top ? 0 : nrow - 1, dg = top ? 1 : -1;
    const double* Ar = top ? A : C;  // couples row g to the row eliminated before it
    const double* Cr = top ? C : A;  // to the row eliminated after it
    double Cp[M][M], dp[M];          // the carry: C'_{g} / A''_{g} and d'_{g} / d''_{g}
#pragma unroll
    for (int i = 0; i < M; ++i) {
        dp[i] = 0.0;
#pragma unroll
        for (int j = 0; j < M; ++j) Cp[i][j] = 0.0;
    }
    // AHEAD: the operands of the next row (after the last one, of the meeting row h) are loaded
    // while the current row is eliminated, and the next row's d' and carried block while the
    // current x is formed. LIGHT: only the next row's A (its couplings to the carried row) and
    // d. Row h is written by neither wavefront before the meeting, and A of row 0 / C of row n-1
    // are loaded but not used.
    constexpr int MA = AHEAD ? M : 1;
    constexpr int ML = (AHEAD || LIGHT) ? M : 1;
    double nA[ML][ML], nB[MA][MA], nC[MA][MA], nd[ML];
    auto fetch = [&](int64_t g) {
        if constexpr (AHEAD) {
            load_mat_c<M>(Ar, am, g, nA);
            load_mat_c<M>(B, am, g, nB);
            load_mat_c<M>(Cr, am, g, nC);
            load_vec<M>(D, av, g, nd);
        } else if constexpr (LIGHT) {
            load_mat_c<M>(Ar, am, g, nA);
            load_vec<M>(D, av, g, nd);
        }
    };
    fetch(g0);
    bool ok = true;
    for (int t = 0; t < L; ++t) {
        const int64_t g = g0 + t * dg;
        double R[M][M], v[M];
        int piv[M];
        if constexpr (AHEAD) {
            double cA[M][M], cC[M][M];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                v[i] = nd[i];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    R[i][j] = nB[i][j];
                    cA[i][j] = nA[i][j];
                    cC[i][j] = nC[i][j];
                }
            }
            fetch(g + dg);
            if (t > 0) gemm_sub_reg<M>(cA, Cp, dp, R, v);
            ok = lu_factor<M>(R, piv) && ok;
            lu_solve<M>(R, piv, v);
#pragma unroll
            for (int i = 0; i < M; ++i) dp[i] = v[i];
            store_vec<M>(D, av, g, dp);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double c[M];
#pragma unroll
                for (int i = 0; i < M; ++i) c[i] = cC[i][j];
                lu_solve<M>(R, piv, c);
#pragma unroll
                for (int i = 0; i < M; ++i) Cp[i][j] = c[i];
                store_col<M>(C, am, g, j, c);
            }
        } else if constexpr (LIGHT) {
            // A and d of this row were loaded during the previous row: A C' is formed while B
            // arrives, then the next row's A and d are requested
            load_mat<M>(B, am, g, R);
            double S[M][M], sv[M];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                v[i] = nd[i];
                sv[i] = 0.0;
#pragma unroll
                for (int j = 0; j < M; ++j) S[i][j] = 0.0;
            }
            if (t > 0) {
#pragma unroll
                for (int k = 0; k < M; ++k)
#pragma unroll
                    for (int i = 0; i < M; ++i) {
                        sv[i] = fma(nA[i][k], dp[k], sv[i]);
#pragma unroll
                        for (int j = 0; j < M; ++j) S[i][j] = fma(nA[i][k], Cp[k][j], S[i][j]);
                    }
            }
            fetch(g + dg);
#pragma unroll
            for (int i = 0; i < M; ++i) {
                v[i] -= sv[i];
#pragma unroll
                for (int j = 0; j < M; ++j) R[i][j] -= S[i][j];
            }
            ok = lu_factor<M>(R, piv) && ok;
            lu_solve<M>(R, piv, v);
#pragma unroll
            for (int i = 0; i < M; ++i) dp[i] = v[i];
            store_vec<M>(D, av, g, dp);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double c[M];
                load_col<M>(Cr, am, g, j, c);
                lu_solve<M>(R, piv, c);
#pragma unroll
                for (int i = 0; i < M; ++i) Cp[i][j] = c[i];
                store_col<M>(C, am, g, j, c);
            }
        } else {
            load_mat<M>(B, am, g, R);
            load_vec<M>(D, av, g, v);
            if (t > 0) gemm_sub<M>(Ar, am, g, Cp, dp, R, v);
            ok = lu_factor<M>(R, piv) && ok;
            lu_solve<M>(R, piv, v);
#pragma unroll
            for (int i = 0; i < M; ++i) dp[i] = v[i];
            store_vec<M>(D, av, g, dp);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double c[M];
                load_col<M>(Cr, am, g, j, c);
                lu_solve<M>(R, piv, c);
#pragma unroll
                for (int i = 0; i < M; ++i) Cp[i][j] = c[i];
                store_col<M>(C, am, g, j, c);
            }
        }
    }
    // the meeting row: each wavefront posts its carry and reads its partner's
    double* mine = xs[wv];
    const double* other = xs[1 - wv];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        mine[(M * M + i) * kTile + lane] = dp[i];
#pragma unroll
        for (int j = 0; j < M; ++j) mine[(i * M + j) * kTile + lane] = Cp[i][j];
    }
    __syncthreads();
    double Yt[M][M], yt[M], Yb[M][M], yb[M];  // top's C'_{h-1}, d'_{h-1}; bottom's A''_{h+1}, d''_{h+1}
#pragma unroll
    for (int i = 0; i < M; ++i) {
        yt[i] = top ? dp[i] : other[(M * M + i) * kTile + lane];
        yb[i] = top ? other[(M * M + i) * kTile + lane] : dp[i];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            Yt[i][j] = top ? Cp[i][j] : other[(i * M + j) * kTile + lane];
            Yb[i][j] = top ? other[(i * M + j) * kTile + lane] : Cp[i][j];
        }
    }
    double xh[M];
    {
        double R[M][M];
        if constexpr (AHEAD) {
            double Ah[M][M], Ch[M][M];  // the top wavefront fetched A_h as its Ar, the bottom one as Cr
#pragma unroll
            for (int i = 0; i < M; ++i) {
                xh[i] = nd[i];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    R[i][j] = nB[i][j];
                    Ah[i][j] = top ? nA[i][j] : nC[i][j];
                    Ch[i][j] = top ? nC[i][j] : nA[i][j];
                }
            }
            if (h > 0) gemm_sub_reg<M>(Ah, Yt, yt, R, xh);
            if (h < nrow - 1) gemm_sub_reg<M>(Ch, Yb, yb, R, xh);
        } else {
            load_mat<M>(B, am, h, R);
            load_vec<M>(D, av, h, xh);
            if (h > 0) gemm_sub<M>(A, am, h, Yt, yt, R, xh);
            if (h < nrow - 1) gemm_sub<M>(C, am, h, Yb, yb, R, xh);
        }
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, xh);
    }
    __syncthreads();  // both wavefronts have read D_h and C_h before the top one writes x_h
    if (top) store_vec<M>(D, av, h, xh);
    // outward substitution from the meeting row
    double xn[M];
#pragma unroll
    for (int i = 0; i < M; ++i) xn[i] = xh[i];
    double px[MA], pc[MA][MA];
    auto fetch_back = [&](int64_t g) {
        if constexpr (AHEAD) {
            load_vec<M>(D, av, g, px);
            load_mat_c<M>(static_cast<const double*>(C), am, g, pc);
        }
    };
    if (L > 0) fetch_back(g0 + (L - 1) * dg);
    for (int t = L - 1; t >= 0; --t) {
        const int64_t g = g0 + t * dg;
        double x[M], c[M][M];
        if constexpr (AHEAD) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                x[i] = px[i];
#pragma unroll
                for (int j = 0; j < M; ++j) c[i][j] = pc[i][j];
            }
            if (t > 0) fetch_back(g - dg);
        } else {
            load_vec<M>(D, av, g, x);
            load_mat<M>(static_cast<const double*>(C), am, g, c);
        }
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) x[i] = fma(-c[i][j], xn[j], x[i]);
#pragma unroll
        for (int i = 0; i < M; ++i) xn[i] = x[i];
        store_vec<M>(D, av, g, x);
    }
    if (!ok) report_singular(flag, sys);
}

// FP32-stored blocks, FP64 arithmetic: the FP64 solve of the blocks' values (bit for bit what
// k_plain computes on their FP64 copies, and so the oracle's solve of those values), reading 4
// bytes per block element; C' (FP64) goes to the workspace Cw, x replaces D.
template <int M, int LAYOUT>
__global__ __launch_bounds__(64) void k_plain_f32s(const float* __restrict__ A,
                                                   const float* __restrict__ B, const float* C,
                                                   double* Cw, double* D, int nsys, int nrow,
                                                   int* __restrict__ flag) {
    plain_solve<M, LAYOUT, false, float, true>(A, B, C, Cw, D, D, nullptr, nullptr, nsys, nrow, flag);
}

// The forward solve that keeps the block factors for the adjoint. C' goes to Cw and d', then x,
// to D, with d read from Din: in place with Cw == C and Din == D, or with C and Din only read.
template <int M, int LAYOUT>
__global__ __launch_bounds__(64) void k_plain_save(const double* __restrict__ A,
                                                   const double* __restrict__ B, const double* C,
                                                   double* Cw, const double* Din, double* D,
                                                   double* __restrict__ F, uint64_t* __restrict__ PV,
                                                   int nsys, int nrow, int* __restrict__ flag) {
    plain_solve<M, LAYOUT, true, double>(A, B, C, Cw, Din, D, F, PV, nsys, nrow, flag);
}

// --------------------------------------------------------------------------- adjoint --
// T^T lambda = w with the block factorization T = L U of the forward solve (L: diagonal
// R_q, sub-diagonal A_q; U: identity diagonal, super-diagonal C'_q):
//   U^T y = w:       y_0 = w_0,  y_q = w_q - C'_{q-1}^T y_{q-1}
//   L^T lambda = y:  R_q^T lambda_q = y_q - A_{q+1}^T lambda_{q+1}   (q = n-1 .. 0)
// then the gradients of L(x) with dL/dx = w:
//   dd = lambda, dA_q = -lambda_q x_{q-1}^T, dB_q = -lambda_q x_q^T, dC_q = -lambda_q x_{q+1}^T
// (dA_0 = 0 and dC_{n-1} = 0: those blocks are ignored by the solve).
template <int M, int LAYOUT>
__global__ __launch_bounds__(64) void k_adjoint(const double* __restrict__ A,
                                                const double* __restrict__ Cp,
                                                const double* __restrict__ F,
                                                const uint64_t* __restrict__ PV,
                                                const double* __restrict__ X,
                                                const double* __restrict__ W, double* Lam,
                                                double* __restrict__ dA, double* __restrict__ dB,
                                                double* __restrict__ dC, int nsys, int nrow) {
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) {
        zero_pad_lane<M, LAYOUT>(Lam, sys, nsys, nrow);
        zero_pad_lane<M * M, LAYOUT>(dA, sys, nsys, nrow);
        zero_pad_lane<M * M, LAYOUT>(dB, sys, nsys, nrow);
        zero_pad_lane<M * M, LAYOUT>(dC, sys, nsys, nrow);
        return;
    }
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M), ap(sys, nsys, nrow, 1);
    double y[M];
    load_vec<M>(W, av, 0, y);
    store_vec<M>(Lam, av, 0, y);
    for (int64_t q = 1; q < nrow; ++q) {
        double c[M][M], w[M];
        load_mat_c<M>(Cp, am, q - 1, c);
        load_vec<M>(W, av, q, w);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) w[i] = fma(-c[j][i], y[j], w[i]);
#pragma unroll
        for (int i = 0; i < M; ++i) y[i] = w[i];
        store_vec<M>(Lam, av, q, y);
    }
    double t[M], xq[M], xq1[M];
#pragma unroll
    for (int i = 0; i < M; ++i) t[i] = xq1[i] = 0.0;
    load_vec<M>(X, av, nrow - 1, xq);
    for (int64_t q = nrow - 1; q >= 0; --q) {
        double z[M], f[M][M], xm[M];
        int piv[M];
        load_vec<M>(Lam, av, q, z);
        load_mat_c<M>(F, am, q, f);
        unpack_piv<M>(PV[ap.at(q, 0)], piv);
#pragma unroll
        for (int i = 0; i < M; ++i) z[i] -= t[i];
        lu_solve_t<M>(f, piv, z);
        store_vec<M>(Lam, av, q, z);
        if (q > 0) {
            load_vec<M>(X, av, q - 1, xm);
        } else {
#pragma unroll
            for (int i = 0; i < M; ++i) xm[i] = 0.0;
        }
        double* ra = dA + am.at(q, 0);
        double* rb = dB + am.at(q, 0);
        double* rc = dC + am.at(q, 0);
        const int64_t es = am.estride();
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int i = 0; i < M; ++i) {
                ra[(i + M * j) * es] = -z[i] * xm[j];
                rb[(i + M * j) * es] = -z[i] * xq[j];
                rc[(i + M * j) * es] = -z[i] * xq1[j];
            }
        if (q > 0) {
            double a[M][M];
            load_mat_c<M>(A, am, q, a);
#pragma unroll
            for (int i = 0; i < M; ++i) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < M; ++j) s = fma(a[j][i], z[j], s);
                t[i] = s;
            }
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            xq1[i] = xq[i];
            xq[i] = xm[i];
        }
    }
}

// ---------------------------------------------------------------------- FP32 + refine --
// Blocks stored in FP32; d and x in FP64. Sweep 0 is an FP32 block-Thomas solve (its
// C'_q kept in the FP32 workspace Wc, d'_q in Wv) with FP64 back substitution into x.
// Each refinement sweep forms the FP64 residual r_q = d_q - A_q x_{q-1} - B_q x_q -
// C_q x_{q+1} from the FP32-stored blocks, solves T delta = r with the same FP32 block
// factorization (R_q recomputed, C'_q recomputed in registers for the next row and read
// back from Wc for the back substitution) and updates x += delta in FP64.
template <int M>
__device__ __forceinline__ bool lu_factor_f(float (&a)[M][M], int (&piv)[M]) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        float best = fabsf(a[k][k]);
        int p = k;
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const float v = fabsf(a[i][k]);
            const bool c = v > best;
            best = c ? v : best;
            p = c ? i : p;
        }
        piv[k] = p;
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const bool c = (p == i);
#pragma unroll
            for (int j = k; j < M; ++j) {
                const float t = a[k][j];
                a[k][j] = c ? a[i][j] : t;
                a[i][j] = c ? t : a[i][j];
            }
        }
        ok = ok && (a[k][k] != 0.0f);
        const float r = 1.0f / a[k][k];
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const float l = a[i][k] * r;
            a[i][k] = l;
#pragma unroll
            for (int j = k + 1; j < M; ++j) a[i][j] = fmaf(-l, a[k][j], a[i][j]);
        }
        a[k][k] = r;
    }
    return ok;
}

template <int M>
__device__ __forceinline__ void lu_solve_f(const float (&a)[M][M], const int (&piv)[M], float (&x)[M]) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const bool c = (piv[k] == i);
            const float t = x[k];
            x[k] = c ? x[i] : t;
            x[i] = c ? t : x[i];
        }
#pragma unroll
        for (int i = k + 1; i < M; ++i) x[i] = fmaf(-a[i][k], x[k], x[i]);
    }
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
        float s = x[k];
#pragma unroll
        for (int j = k + 1; j < M; ++j) s = fmaf(-a[k][j], x[j], s);
        x[k] = s * a[k][k];
    }
}

// TR = true solves the transposed system T^T x = d from the same stored blocks: its row q
// has sub-diagonal C_{q-1}^T, diagonal B_q^T and super-diagonal A_{q+1}^T (the adjoint of the
// FP32-stored solve).
template <int M, int LAYOUT, bool TR>
__global__ __launch_bounds__(64) void k_f32r(const float* __restrict__ A, const float* __restrict__ B,
                                             const float* __restrict__ C,
                                             const double* __restrict__ D, double* X,
                                             float* Wc, float* Wv, int nsys, int nrow, int sweeps,
                                             int* __restrict__ flag) {
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) {
        zero_pad_lane<M, LAYOUT>(X, sys, nsys, nrow);
        return;
    }
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
    const int64_t es = am.estride(), vs = av.estride();
    bool ok = true;
    for (int sweep = 0; sweep <= sweeps; ++sweep) {
        // forward: FP32 factorization, FP32 right-hand side (d, or the FP64 residual rounded)
        float Cp[M][M], dp[M];
        double xm[M], xc[M];  // x_{q-1}, x_q of the previous iterate (sweeps >= 1)
        if (sweep > 0) {
#pragma unroll
            for (int i = 0; i < M; ++i) xm[i] = 0.0;
            load_vec<M>(X, av, 0, xc);
        }
        for (int64_t q = 0; q < nrow; ++q) {
            float R[M][M], a[M][M], v[M];
            const bool has_next = q + 1 < nrow;
            // element (i, j) of this row's sub-diagonal, diagonal and super-diagonal block
            const float* rb = B + am.at(q, 0);
            const float* ra = TR ? C + am.at(q > 0 ? q - 1 : 0, 0) : A + am.at(q, 0);
            const float* rc = TR ? A + am.at(has_next ? q + 1 : q, 0) : C + am.at(q, 0);
#pragma unroll
            for (int j = 0; j < M; ++j)
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    const int e = TR ? (j + M * i) : (i + M * j);
                    R[i][j] = rb[e * es];
                    a[i][j] = q > 0 ? ra[e * es] : 0.0f;
                }
            double r[M];
            load_vec<M>(D, av, q, r);
            float cq[M][M];
#pragma unroll
            for (int j = 0; j < M; ++j)
#pragma unroll
                for (int i = 0; i < M; ++i) cq[i][j] = has_next ? rc[(TR ? (j + M * i) : (i + M * j)) * es] : 0.0f;
            if (sweep > 0) {
                double xp[M];
                if (has_next) {
                    load_vec<M>(X, av, q + 1, xp);
                } else {
#pragma unroll
                    for (int i = 0; i < M; ++i) xp[i] = 0.0;
                }
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    double s = r[i];
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        s = fma(-(double)a[i][j], xm[j], s);
                        s = fma(-(double)R[i][j], xc[j], s);
                        s = fma(-(double)cq[i][j], xp[j], s);
                    }
                    r[i] = s;
                }
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    xm[i] = xc[i];
                    xc[i] = xp[i];
                }
            }
#pragma unroll
            for (int i = 0; i < M; ++i) v[i] = (float)r[i];
            if (q > 0) {
#pragma unroll
                for (int k = 0; k < M; ++k)
#pragma unroll
                    for (int i = 0; i < M; ++i) {
                        v[i] = fmaf(-a[i][k], dp[k], v[i]);
#pragma unroll
                        for (int j = 0; j < M; ++j) R[i][j] = fmaf(-a[i][k], Cp[k][j], R[i][j]);
                    }
            }
            int piv[M];
            const bool okq = lu_factor_f<M>(R, piv);
            if (sweep == 0) ok = ok && okq;
            lu_solve_f<M>(R, piv, v);
#pragma unroll
            for (int i = 0; i < M; ++i) dp[i] = v[i];
            float* wv = Wv + av.at(q, 0);
#pragma unroll
            for (int i = 0; i < M; ++i) wv[i * vs] = v[i];
            if (has_next) {
                float* wc = Wc + am.at(q, 0);
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    float c[M];
#pragma unroll
                    for (int i = 0; i < M; ++i) c[i] = cq[i][j];
                    lu_solve_f<M>(R, piv, c);
#pragma unroll
                    for (int i = 0; i < M; ++i) {
                        Cp[i][j] = c[i];
                        if (sweep == 0) wc[(i + M * j) * es] = c[i];
                    }
                }
            }
        }
        // backward in FP64: delta_q = d'_q - C'_q delta_{q+1}; x = delta (sweep 0) or x += delta
        double dn[M];
#pragma unroll
        for (int i = 0; i < M; ++i) dn[i] = (double)dp[i];
        {
            double x[M];
            if (sweep > 0) load_vec<M>(X, av, nrow - 1, x);
#pragma unroll
            for (int i = 0; i < M; ++i) x[i] = sweep > 0 ? x[i] + dn[i] : dn[i];
            store_vec<M>(X, av, nrow - 1, x);
        }
        for (int64_t q = nrow - 2; q >= 0; --q) {
            const float* wc = Wc + am.at(q, 0);
            const float* wv = Wv + av.at(q, 0);
            double dl[M];
#pragma unroll
            for (int i = 0; i < M; ++i) dl[i] = (double)wv[i * vs];
#pragma unroll
            for (int j = 0; j < M; ++j)
#pragma unroll
                for (int i = 0; i < M; ++i) dl[i] = fma(-(double)wc[(i + M * j) * es], dn[j], dl[i]);
            double x[M];
            if (sweep > 0) load_vec<M>(X, av, q, x);
#pragma unroll
            for (int i = 0; i < M; ++i) {
                dn[i] = dl[i];
                x[i] = sweep > 0 ? x[i] + dl[i] : dl[i];
            }
            store_vec<M>(X, av, q, x);
        }
    }
    if (!ok) report_singular(flag, sys);
}

// ------------------------------------------------------------------------------ modi --
// Stage 1 over partition `part` (rows [r0, r0+N)) of each system, one lane per partition
// (m <= 6; k_modi_coop below for larger m), mod_btdma_cpu.f90:305-443.
// Forward sweep (in place: Â over A, Ĉ over C, d̂ over D for rows 0..N-2):
//   [d̂|Â|Ĉ]_q = R^{-1}[d_q - A_q d̂_{q-1} | -A_q Â_{q-1} | C_q],  R = B_q - A_q Ĉ_{q-1}
// (rows 0 and 1: R = B_q and the right-hand side [d_q | A_q | C_q]), so that
//   x_q + Â_q x_0' + Ĉ_q x_{q+1} = d̂_q   (x_0' = first row of the partition, q >= 1).
// The reference then rewrites every row in a backward sweep. Here the closing step's only
// need from that sweep, row 1 expressed through x_0' and x_{N-1}
//   x_1 = alpha - beta x_0' - gamma x_{N-1},
// is composed during the forward sweep (alpha, beta, gamma start as row 1's d̂, Â, Ĉ and take
// alpha -= gamma d̂_k, beta -= gamma Â_k, gamma = -gamma Ĉ_k for k = 2..N-2), with the three
// kept in LDS ([element][lane], sp); stage 3 uses the forward values directly, so the
// partition is read once and nothing is re-read. Reduced rows go to the plan's buffers
// (tiled, 2P rows per system):
//   row 2*part   : (Â_1, I, Ĉ_1, d̂_1) after the close, couples x_1 to x_0 and x_N
//   row 2*part+1 : (Â_N, I, Ĉ_N, d̂_N) forward values, couples x_N to x_1 and x_{N+1}
// Where stage 1 takes the blocks of its rows: the arrays A, B, C (MemBlocks, the default), or a
// block source that computes them for the lane's line as the rows advance (the matrix-free
// weno2d sweep, btd_weno.hip): BS::lane(sys) gives an object with row(g), which makes row g's
// blocks current, and b(R), a_col(k, a), a_mat(X), c_col(j, c), which read them.
struct MemBlocks {};

// AHEAD (blocks in memory only): row q+1's A, B, C, d are loaded while row q is eliminated, for
// batches with too few partitions to hide HBM latency behind other wavefronts; the extra 80
// doubles (m = 5) a lane holds cost occupancy, so such launches ask for one wavefront per SIMD.
template <int M, int LAYOUT, class BS = MemBlocks, bool AHEAD = false>
__global__ __launch_bounds__(64, AHEAD ? 1 : 2) void k_modi(double* A, const double* __restrict__ B, double* C,
                                             double* D, int nsys, int nrow,
                                             const int* __restrict__ part_first,
                                             const int* __restrict__ part_len, int npart,
                                             double* __restrict__ redA, double* __restrict__ redC,
                                             double* __restrict__ redD, int* __restrict__ flag, BS bs) {
    constexpr bool kMem = std::is_same<BS, MemBlocks>::value;
    static_assert(kMem || !AHEAD, "rows are loaded ahead only from memory");
    const int ntiles = (nsys + 63) / 64;
    const int tile = blockIdx.x % ntiles, part = blockIdx.x / ntiles;
    const int64_t sys = (int64_t)tile * 64 + threadIdx.x;
    if (sys >= nsys) return;
    {
        const int64_t r0 = part_first[part];
        const int N = part_len[part];
        const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
        extern __shared__ double sp[];  // alpha (M), beta (M*M), gamma (M*M), [element][lane]
        const int lane = threadIdx.x;
        auto SA = [&](int i) -> double& { return sp[i * 64 + lane]; };
        auto SB = [&](int i, int j) -> double& { return sp[(M + i * M + j) * 64 + lane]; };
        auto SG = [&](int i, int j) -> double& { return sp[(M + M * M + i * M + j) * 64 + lane]; };
        bool ok = true;
        double Ah[M][M], Ch[M][M], dh[M];
        auto src = [&] {
            if constexpr (kMem) return 0;
            else return bs.lane(sys);
        }();
        // AHEAD: the next row's blocks (natural form) and right-hand side
        double nA[AHEAD ? M : 1][AHEAD ? M : 1], nB[AHEAD ? M : 1][AHEAD ? M : 1];
        double nC[AHEAD ? M : 1][AHEAD ? M : 1], nd[AHEAD ? M : 1];
        auto fetch_next = [&](int64_t g) {
            if constexpr (AHEAD) {
                load_mat_c<M>(A, am, g, nA);
                load_mat_c<M>(B, am, g, nB);
                load_mat_c<M>(C, am, g, nC);
                load_vec<M>(D, av, g, nd);
            }
        };
        fetch_next(r0);
        for (int q = 0; q < N; ++q) {
            const int64_t g = r0 + q;
            double R[M][M], v[M], X[M][M];
            double cA[AHEAD ? M : 1][AHEAD ? M : 1], cC[AHEAD ? M : 1][AHEAD ? M : 1];
            if constexpr (AHEAD) {
    #pragma unroll
                for (int i = 0; i < M; ++i) {
                    v[i] = nd[i];
    #pragma unroll
                    for (int j = 0; j < M; ++j) {
                        R[i][j] = nB[i][j];
                        cA[i][j] = nA[i][j];
                        cC[i][j] = nC[i][j];
                    }
                }
                fetch_next(q + 1 < N ? g + 1 : g);
            } else {
                if constexpr (kMem) {
                    load_mat<M>(B, am, g, R);
                } else {
                    src.row(g);
                    src.b(R);
                }
                load_vec<M>(D, av, g, v);
            }
            if (q >= 2) {
    #pragma unroll
                for (int i = 0; i < M; ++i)
    #pragma unroll
                    for (int j = 0; j < M; ++j) X[i][j] = 0.0;
    #pragma unroll
                for (int k = 0; k < M; ++k) {
                    double ak[M];
                    if constexpr (AHEAD) {
    #pragma unroll
                        for (int i = 0; i < M; ++i) ak[i] = cA[i][k];
                    } else if constexpr (kMem) {
                        load_col<M>(A, am, g, k, ak);
                    } else {
                        src.a_col(k, ak);
                    }
    #pragma unroll
                    for (int i = 0; i < M; ++i) {
                        v[i] = fma(-ak[i], dh[k], v[i]);
    #pragma unroll
                        for (int j = 0; j < M; ++j) {
                            R[i][j] = fma(-ak[i], Ch[k][j], R[i][j]);
                            X[i][j] = fma(-ak[i], Ah[k][j], X[i][j]);
                        }
                    }
                }
            } else {
                if constexpr (AHEAD) {
    #pragma unroll
                    for (int i = 0; i < M; ++i)
    #pragma unroll
                        for (int j = 0; j < M; ++j) X[i][j] = cA[i][j];
                } else if constexpr (kMem) {
                    load_mat_c<M>(A, am, g, X);
                } else {
                    src.a_mat(X);
                }
            }
            int piv[M];
            ok = lu_factor<M>(R, piv) && ok;
            lu_solve<M>(R, piv, v);
            solve_rows<M>(R, piv, X);
            const bool keep = q + 1 < N;  // the last row only feeds the reduced system
    #pragma unroll
            for (int j = 0; j < M; ++j) {
                double c[M];
                if constexpr (AHEAD) {
    #pragma unroll
                    for (int i = 0; i < M; ++i) c[i] = cC[i][j];
                } else if constexpr (kMem) {
                    load_col<M>(C, am, g, j, c);
                } else {
                    src.c_col(j, c);
                }
                lu_solve<M>(R, piv, c);
    #pragma unroll
                for (int i = 0; i < M; ++i) Ch[i][j] = c[i];
                if (keep) store_col<M>(C, am, g, j, c);
            }
    #pragma unroll
            for (int i = 0; i < M; ++i) {
                dh[i] = v[i];
    #pragma unroll
                for (int j = 0; j < M; ++j) Ah[i][j] = X[i][j];
            }
            if (keep) {
                store_mat<M>(A, am, g, Ah);
                store_vec<M>(D, av, g, dh);
            }
            if (q == 1) {
    #pragma unroll
                for (int i = 0; i < M; ++i) {
                    SA(i) = dh[i];
    #pragma unroll
                    for (int j = 0; j < M; ++j) {
                        SB(i, j) = Ah[i][j];
                        SG(i, j) = Ch[i][j];
                    }
                }
            } else if (q >= 2 && keep) {
    #pragma unroll
                for (int i = 0; i < M; ++i) {
                    double gi[M];
    #pragma unroll
                    for (int k = 0; k < M; ++k) gi[k] = SG(i, k);
                    double a = SA(i);
    #pragma unroll
                    for (int k = 0; k < M; ++k) a = fma(-gi[k], dh[k], a);
                    SA(i) = a;
    #pragma unroll
                    for (int j = 0; j < M; ++j) {
                        double b = SB(i, j), c = 0.0;
    #pragma unroll
                        for (int k = 0; k < M; ++k) {
                            b = fma(-gi[k], Ah[k][j], b);
                            c = fma(-gi[k], Ch[k][j], c);
                        }
                        SB(i, j) = b;
                        SG(i, j) = c;
                    }
                }
            }
        }
        const Addr<LAYOUT_TILED> rm(sys, 0, 2 * npart, M * M), rv(sys, 0, 2 * npart, M);
        store_mat<M>(redA, rm, 2 * part + 1, Ah);
        store_mat<M>(redC, rm, 2 * part + 1, Ch);
        store_vec<M>(redD, rv, 2 * part + 1, dh);
    #pragma unroll
        for (int i = 0; i < M; ++i) {
            dh[i] = SA(i);
    #pragma unroll
            for (int j = 0; j < M; ++j) {
                Ah[i][j] = SB(i, j);
                Ch[i][j] = SG(i, j);
            }
        }
        // close, q = 0: R = I - Ĉ_0 beta; [d̂|Â|Ĉ]_0 = R^{-1}[d̂_0 - Ĉ_0 alpha | Â_0 | -Ĉ_0 gamma]
        // with (dh, Ah, Ch) = (alpha, beta, gamma), row 1 composed over the whole partition.
        {
            double A0[M][M], C0[M][M], d0[M], R[M][M], X[M][M], T[M][M], t[M];
            load_mat_c<M>(A, am, r0, A0);
            load_mat_c<M>(C, am, r0, C0);
            load_vec<M>(D, av, r0, d0);
            mv<M>(C0, dh, t);
    #pragma unroll
            for (int i = 0; i < M; ++i) d0[i] -= t[i];
            mm<M>(C0, Ah, T);
    #pragma unroll
            for (int i = 0; i < M; ++i)
    #pragma unroll
                for (int j = 0; j < M; ++j) R[i][j] = (i == j ? 1.0 : 0.0) - T[i][j];
            mm<M>(C0, Ch, X);
    #pragma unroll
            for (int i = 0; i < M; ++i)
    #pragma unroll
                for (int j = 0; j < M; ++j) X[i][j] = -X[i][j];
            int piv[M];
            ok = lu_factor<M>(R, piv) && ok;
            lu_solve<M>(R, piv, d0);
            solve_rows<M>(R, piv, A0);
            solve_rows<M>(R, piv, X);
            store_mat<M>(redA, rm, 2 * part, A0);
            store_mat<M>(redC, rm, 2 * part, X);
            store_vec<M>(redD, rv, 2 * part, d0);
        }
        if (!ok) report_singular(flag, sys);
    }
}

// ------------------------------------------------------------------------ modi, twisted --
// Stage 1 with each partition swept from both ends, by the two wavefronts of a workgroup (the top
// and the bottom half of partition blockIdx.x / ntiles of the 64 systems of tile
// blockIdx.x mod ntiles, lane l on system 64 * tile + l), blocks in memory, partitions of at
// least 4 rows: each lane's chain of dependent rows is half the partition's, for batches with
// too few partitions to fill the device. The direction is uniform over a wavefront, so the
// arrays it reads are too. The top wavefront runs k_modi's forward sweep (rows loaded ahead)
// over rows 0..k, k = ceil(N/2) - 1; the bottom one runs the same sweep mirrored, over rows N-1
// down to k+1, with A and C exchanging roles. A lane's rows then satisfy, with its anchor x_a
// (x_0' top, x_{N-1} bottom) and the next row of its sweep x_s (x_{q+1} top, x_{q-1} bottom),
//   x_q + Â_q x_a + Ĉ_q x_s = d̂_q   (in place: d̂ over D, Â over the array of A's role, Ĉ over
//                                    the array of C's role: A and C top, C and A bottom),
// and row 1 (top) / row N-2 (bottom) is composed on the way, as in k_modi, as
// x = alpha - beta x_a - gamma x_s(last row). The two last rows p (k and k+1) of a system meet:
// with the partner's last row x_n + H x_f + K x_p = h (x_f the partner's anchor), each lane solves
//   x_p = a - b x_a - c x_f,  S = I - Ĉ_p K,  a = S^{-1}(d̂_p - Ĉ_p h),  b = S^{-1} Â_p,
//   c = -S^{-1} Ĉ_p H,
// and stores it as row p (its Ĉ now the coefficient of x_f); the partner's solution
// x_n = a' - b' x_f - c' x_a finishes the composition:
//   alpha -= gamma a',  beta -= gamma c',  gamma <- -gamma b'.
// The close of row 0 (top) / row N-1 (bottom) is k_modi's; the top lane writes reduced row
// 2*part, the bottom lane row 2*part+1 with its Â (the coefficient of x_N) as the reduced C and
// its Ĉ (of x_0') as the reduced A: the reduced system of k_modi, to rounding. Stage 3 is
// k_update_twisted. Each wavefront's composition and its side of the meeting's two exchanges
// share its LDS slots ([wavefront][element][lane], sp). AHEAD: each row's blocks are loaded
// while the previous one is eliminated (one wavefront per SIMD); without, at the start of their
// own row step.
template <int M, int LAYOUT, bool AHEAD = true>
__global__ __launch_bounds__(2 * kTile, AHEAD ? 1 : 2) void k_modi_twisted(
    double* A, const double* __restrict__ B, double* C, double* D, int nsys, int nrow,
    const int* __restrict__ part_first, const int* __restrict__ part_len, int npart, double* __restrict__ redA,
    double* __restrict__ redC, double* __restrict__ redD, int* __restrict__ flag) {
    constexpr int E2 = 2 * M * M + M;  // LDS slots of one wavefront
    const int ntiles = (nsys + kTile - 1) / kTile;
    const int tile = blockIdx.x % ntiles, part = blockIdx.x / ntiles;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile);  // 0 top, 1 bottom
    const int lane = threadIdx.x % kTile;
    const bool top = wv == 0;
    const int64_t sys = (int64_t)tile * kTile + lane;
    if (sys >= nsys) return;  // a system's two lanes leave together: both wavefronts map lanes alike
    const int64_t r0 = part_first[part];
    const int N = part_len[part];  // >= 4
    const int k = (N + 1) / 2 - 1;
    const int L = top ? k + 1 : N - 1 - k;  // rows of this lane's sweep, >= 2
    const int64_t g0 = top ? r0 : r0 + N - 1, dg = top ? 1 : -1;
    double* Ar = top ? A : C;  // couples row q to the previous row of the sweep
    double* Cr = top ? C : A;  // to the next one
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
    extern __shared__ double sp_all[];
    double* sp = sp_all + wv * E2 * kTile;         // this wavefront's slots
    double* pp = sp_all + (1 - wv) * E2 * kTile;   // the partner's
    auto SA = [&](int i) -> double& { return sp[i * kTile + lane]; };
    auto SB = [&](int i, int j) -> double& { return sp[(M + i * M + j) * kTile + lane]; };
    auto SG = [&](int i, int j) -> double& { return sp[(M + M * M + i * M + j) * kTile + lane]; };
    auto PA = [&](int i) { return pp[i * kTile + lane]; };
    auto PB = [&](int i, int j) { return pp[(M + i * M + j) * kTile + lane]; };
    auto PG = [&](int i, int j) { return pp[(M + M * M + i * M + j) * kTile + lane]; };
    bool ok = true;
    double Ah[M][M], Ch[M][M], dh[M];
    constexpr int MA = AHEAD ? M : 1;
    double nA[MA][MA], nB[MA][MA], nC[MA][MA], nd[MA];
    auto fetch_next = [&](int64_t g) {
        if constexpr (AHEAD) {
            load_mat_c<M>(Ar, am, g, nA);
            load_mat_c<M>(B, am, g, nB);
            load_mat_c<M>(Cr, am, g, nC);
            load_vec<M>(D, av, g, nd);
        }
    };
    fetch_next(g0);
    for (int q = 0; q < L; ++q) {
        const int64_t g = g0 + q * dg;
        double R[M][M], v[M], X[M][M], cA[M][M], cC[M][M];
        if constexpr (AHEAD) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                v[i] = nd[i];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    R[i][j] = nB[i][j];
                    cA[i][j] = nA[i][j];
                    cC[i][j] = nC[i][j];
                }
            }
            fetch_next(q + 1 < L ? g + dg : g);
        } else {
            load_mat_c<M>(Ar, am, g, cA);
            load_mat<M>(B, am, g, R);
            load_mat_c<M>(Cr, am, g, cC);
            load_vec<M>(D, av, g, v);
        }
        if (q >= 2) {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) X[i][j] = 0.0;
#pragma unroll
            for (int kk = 0; kk < M; ++kk)
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    v[i] = fma(-cA[i][kk], dh[kk], v[i]);
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        R[i][j] = fma(-cA[i][kk], Ch[kk][j], R[i][j]);
                        X[i][j] = fma(-cA[i][kk], Ah[kk][j], X[i][j]);
                    }
                }
        } else {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) X[i][j] = cA[i][j];
        }
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, v);
        solve_rows<M>(R, piv, X);
        solve_rows<M>(R, piv, cC);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            dh[i] = v[i];
#pragma unroll
            for (int j = 0; j < M; ++j) {
                Ah[i][j] = X[i][j];
                Ch[i][j] = cC[i][j];
            }
        }
        if (q + 1 < L) {  // the last row is stored after the meeting
            store_mat<M>(Ar, am, g, Ah);
            store_mat<M>(Cr, am, g, Ch);
            store_vec<M>(D, av, g, dh);
        }
        if (q == 1) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                SA(i) = dh[i];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    SB(i, j) = Ah[i][j];
                    SG(i, j) = Ch[i][j];
                }
            }
        } else if (q >= 2) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                double gi[M];
#pragma unroll
                for (int kk = 0; kk < M; ++kk) gi[kk] = SG(i, kk);
                double a = SA(i);
#pragma unroll
                for (int kk = 0; kk < M; ++kk) a = fma(-gi[kk], dh[kk], a);
                SA(i) = a;
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    double b = SB(i, j), c = 0.0;
#pragma unroll
                    for (int kk = 0; kk < M; ++kk) {
                        b = fma(-gi[kk], Ah[kk][j], b);
                        c = fma(-gi[kk], Ch[kk][j], c);
                    }
                    SB(i, j) = b;
                    SG(i, j) = c;
                }
            }
        }
    }
    // the composition to registers: its slots carry the meeting's exchanges
    double al[M], be[M][M], ga[M][M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        al[i] = SA(i);
#pragma unroll
        for (int j = 0; j < M; ++j) {
            be[i][j] = SB(i, j);
            ga[i][j] = SG(i, j);
        }
    }
    // exchange 1: the last rows (d̂, Â, Ĉ)_p; S = I - Ĉ_p K, [a | b | c] = S^{-1}[d̂_p - Ĉ_p h | Â_p | -Ĉ_p H]
#pragma unroll
    for (int i = 0; i < M; ++i) {
        SA(i) = dh[i];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            SB(i, j) = Ah[i][j];
            SG(i, j) = Ch[i][j];
        }
    }
    __syncthreads();
    double S[M][M], ma[M], mc[M][M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double t = 0.0;
#pragma unroll
        for (int kk = 0; kk < M; ++kk) t = fma(Ch[i][kk], PA(kk), t);
        ma[i] = dh[i] - t;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double s = 0.0, h = 0.0;
#pragma unroll
            for (int kk = 0; kk < M; ++kk) {
                s = fma(Ch[i][kk], PG(kk, j), s);
                h = fma(Ch[i][kk], PB(kk, j), h);
            }
            S[i][j] = (i == j ? 1.0 : 0.0) - s;
            mc[i][j] = -h;
        }
    }
    __syncthreads();  // the partner has read this lane's row before it is overwritten
    {
        int piv[M];
        ok = lu_factor<M>(S, piv) && ok;
        lu_solve<M>(S, piv, ma);
        solve_rows<M>(S, piv, Ah);
        solve_rows<M>(S, piv, mc);
    }
    const int64_t gp = g0 + (L - 1) * dg;
    store_vec<M>(D, av, gp, ma);
    store_mat<M>(Ar, am, gp, Ah);
    store_mat<M>(Cr, am, gp, mc);
    // exchange 2: (a, b, c); the partner's finishes the composition
#pragma unroll
    for (int i = 0; i < M; ++i) {
        SA(i) = ma[i];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            SB(i, j) = Ah[i][j];
            SG(i, j) = mc[i][j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double gi[M];
#pragma unroll
        for (int kk = 0; kk < M; ++kk) gi[kk] = ga[i][kk];
        double a = al[i];
#pragma unroll
        for (int kk = 0; kk < M; ++kk) a = fma(-gi[kk], PA(kk), a);
        dh[i] = a;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double b = be[i][j], c = 0.0;
#pragma unroll
            for (int kk = 0; kk < M; ++kk) {
                b = fma(-gi[kk], PG(kk, j), b);
                c = fma(-gi[kk], PB(kk, j), c);
            }
            Ah[i][j] = b;
            Ch[i][j] = c;
        }
    }
    // close, as in k_modi, with (dh, Ah, Ch) = (alpha, beta, gamma) through the whole partition
    {
        double A0[M][M], C0[M][M], d0[M], R[M][M], X[M][M], T[M][M], t[M];
        load_mat_c<M>(Ar, am, g0, A0);
        load_mat_c<M>(Cr, am, g0, C0);
        load_vec<M>(D, av, g0, d0);
        mv<M>(C0, dh, t);
#pragma unroll
        for (int i = 0; i < M; ++i) d0[i] -= t[i];
        mm<M>(C0, Ah, T);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) R[i][j] = (i == j ? 1.0 : 0.0) - T[i][j];
        mm<M>(C0, Ch, X);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) X[i][j] = -X[i][j];
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, d0);
        solve_rows<M>(R, piv, A0);
        solve_rows<M>(R, piv, X);
        const Addr<LAYOUT_TILED> rm(sys, 0, 2 * npart, M * M), rv(sys, 0, 2 * npart, M);
        const int rr = 2 * part + (top ? 0 : 1);
        if (top) {
            store_mat<M>(redA, rm, rr, A0);
            store_mat<M>(redC, rm, rr, X);
        } else {
            store_mat<M>(redA, rm, rr, X);
            store_mat<M>(redC, rm, rr, A0);
        }
        store_vec<M>(redD, rv, rr, d0);
    }
    if (!ok) report_singular(flag, sys);
}

// XCDs of the device: workgroups are dispatched to them round robin (workgroup b runs on XCD
// b mod kXcds), each XCD with its own L2.
constexpr int kXcds = 8;

// Stage 1, cooperative: the arithmetic of k_modi (forward modified sweep, row 1's relation
// composed on the way, the close, the two reduced rows) with a partition's block rows spread
// over m lanes (lane i owns row i of every block). Each row is eliminated by coop_eliminate
// with the right-hand sides [d | X | C], X = A_q for q < 2 and -A_q Â_{q-1} after; Ĉ and Â
// are carried to the next row in LDS, d̂ whole in every lane; alpha, beta, gamma are kept
// row-wise (lane i: alpha_i, beta(i,:), gamma(i,:)). floor(64/m) consecutive systems of one
// partition per wavefront, WV wavefronts per workgroup on consecutive systems; the items of
// one partition go to the workgroups of one XCD in order, so neighbouring systems' shared
// lines meet in one L2. Used for m >= 7, where the one-lane form spills its blocks to scratch
// memory.
// Rows k_modi_coop loads ahead: 4 for m <= 5 (a deeper ring costs the occupancy of a second
// wavefront per SIMD; 2 rows measured no faster at m = 5), 1 from m = 6.
template <int M>
constexpr int kModiAhead = M <= 5 ? 4 : 1;

// LOCK: the WV wavefronts of a workgroup (consecutive systems of one partition) meet at a
// barrier (s_barrier, no memory wait) before each row's loads, so the cache lines their
// systems share (floor(64/m) systems a wavefront do not fill whole 128-byte lines) are
// requested together and come from HBM once; every wavefront then runs every row step, those
// past the last system on a real one, storing nothing.
template <int M, int LAYOUT, int WV, bool LOCK = false>
__global__ __launch_bounds__(64 * WV) void k_modi_coop(double* A, const double* __restrict__ B, double* C,
                                                       double* D, int nsys, int nrow,
                                                       const int* __restrict__ part_first,
                                                       const int* __restrict__ part_len, int npart,
                                                       double* __restrict__ redA, double* __restrict__ redC,
                                                       double* __restrict__ redD, int* __restrict__ flag) {
    constexpr int G = kTile / M;  // systems per wavefront
    constexpr int E = M * M;
    __shared__ double shw[WV][(5 * M + 3) * kTile];
    const int nchunks = (nsys + G - 1) / G;
    const int ngroups = (nchunks + WV - 1) / WV;
    const int nitems = npart * ngroups;
    const int per = (nitems + kXcds - 1) / kXcds;
    const int u = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (u >= nitems) return;
    const int w = threadIdx.x / kTile, lane = threadIdx.x % kTile;
    const int part = u / ngroups, chunk0 = (u % ngroups) * WV + w;
    if (!LOCK && chunk0 >= nchunks) return;  // the wavefronts share no barrier
    const bool live = chunk0 < nchunks;
    const int chunk = live ? chunk0 : nchunks - 1;
    double* sh = shw[w];
    const int g = lane / M, i = lane % M;
    const bool grouped = g < G;
    const int64_t s0 = (int64_t)chunk * G;
    const int64_t sys = s0 + (grouped ? g : 0);
    const bool mine = live && grouped && sys < nsys;
    const int64_t se = grouped && sys < nsys ? sys : s0;  // a real system either way
    // lanes past the last whole group read group 0's posts and write only their own slots
    const int base = grouped ? g * M : 0;
    const Addr<LAYOUT> am(se, nsys, nrow, E), av(se, nsys, nrow, M);
    const Addr<LAYOUT_TILED> rm(se, 0, 2 * npart, E), rv(se, 0, 2 * npart, M);
    const int64_t r0 = part_first[part];
    const int N = part_len[part];
    double* carry = sh;                     // slots [0, M) Ĉ, [M, 2M) Â (slot k, lane base+j: (k, j))
    double* xb = sh + (2 * M + 1) * kTile;  // the elimination exchange, 3M + 2 slots
    bool ok = true;
    // rows loaded F ahead (one wavefront per item has few others to hide HBM latency behind);
    // a row past the partition's end repeats its last row (loaded, never used)
    constexpr int F = kModiAhead<M>;
    double pa[F][M], pb[F][M], pc[F][M], pd[F];
    auto fetch = [&](double (&fa)[M], double (&fb)[M], double (&fc)[M], double& fd, int q) {
        const int64_t gr = r0 + (q < N ? q : N - 1);
#pragma unroll
        for (int k = 0; k < M; ++k) {
            fa[k] = A[am.at(gr, i + M * k)];
            fb[k] = __builtin_nontemporal_load(B + am.at(gr, i + M * k));
            fc[k] = C[am.at(gr, i + M * k)];
        }
        fd = D[av.at(gr, i)];
    };
#pragma unroll
    for (int t = 0; t < F; ++t) fetch(pa[t], pb[t], pc[t], pd[t], t);
    double dp[M], beta[M], gamma[M], alpha = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) dp[k] = beta[k] = gamma[k] = 0.0;
    for (int q0 = 0; q0 < N; q0 += F)
#pragma unroll
    for (int t = 0; t < F; ++t) {
        const int q = q0 + t;
        if (q >= N) break;  // uniform: every lane of the wavefront is in this partition
        const int64_t gr = r0 + q;
        double r[M], rhs[2 * M + 1];
        rhs[0] = pd[t];
        if (q >= 2) {
#pragma unroll
            for (int j = 0; j < M; ++j) {
                r[j] = pb[t][j];
                rhs[1 + j] = 0.0;
                rhs[1 + M + j] = pc[t][j];
            }
#pragma unroll
            for (int k = 0; k < M; ++k) {
                const double a = pa[t][k];
                rhs[0] = fma(-a, dp[k], rhs[0]);
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    r[j] = fma(-a, carry[k * kTile + base + j], r[j]);
                    rhs[1 + j] = fma(-a, carry[(M + k) * kTile + base + j], rhs[1 + j]);
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < M; ++j) {
                r[j] = pb[t][j];
                rhs[1 + j] = pa[t][j];
                rhs[1 + M + j] = pc[t][j];
            }
        }
        if constexpr (LOCK) __builtin_amdgcn_s_barrier();
        fetch(pa[t], pb[t], pc[t], pd[t], q + F);
        double xc[2 * M], xd[M];
        ok = !coop_eliminate<M, 2>(r, rhs, xb, i, base, xc, xd) && ok;
        const bool keep = q + 1 < N;  // the last row only feeds the reduced system
        if (mine) {
            if (keep) {
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    A[am.at(gr, k + M * i)] = xc[k];
                    C[am.at(gr, k + M * i)] = xc[M + k];
                }
                D[av.at(gr, i)] = lane_elem<M>(xd, i);
            } else {
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    redA[rm.at(2 * part + 1, k + M * i)] = xc[k];
                    redC[rm.at(2 * part + 1, k + M * i)] = xc[M + k];
                }
                redD[rv.at(2 * part + 1, i)] = lane_elem<M>(xd, i);
            }
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
            carry[k * kTile + lane] = xc[M + k];  // Ĉ(k, i)
            carry[(M + k) * kTile + lane] = xc[k];  // Â(k, i)
            dp[k] = xd[k];
        }
        wave_lds_sync();
        if (q == 1) {
            alpha = lane_elem<M>(xd, i);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                beta[j] = carry[(M + i) * kTile + base + j];
                gamma[j] = carry[i * kTile + base + j];
            }
        } else if (q >= 2 && keep) {
            // alpha -= gamma d̂_q, beta -= gamma Â_q, gamma = -gamma Ĉ_q (row i)
#pragma unroll
            for (int k = 0; k < M; ++k) alpha = fma(-gamma[k], xd[k], alpha);
            double nb[M], ng[M];
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double b = beta[j], c = 0.0;
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    b = fma(-gamma[k], carry[(M + k) * kTile + base + j], b);
                    c = fma(-gamma[k], carry[k * kTile + base + j], c);
                }
                nb[j] = b;
                ng[j] = c;
            }
#pragma unroll
            for (int j = 0; j < M; ++j) {
                beta[j] = nb[j];
                gamma[j] = ng[j];
            }
        }
    }
    // close, row 0: R = I - Ĉ_0 beta; [d̂|Â|Ĉ]_0 = R^{-1}[d̂_0 - Ĉ_0 alpha | Â_0 | -Ĉ_0 gamma], with
    // Â_0, Ĉ_0, d̂_0 read back where the group stored them and alpha, beta, gamma through LDS
    {
        double a0[M], c0[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            a0[j] = A[am.at(r0, i + M * j)];
            c0[j] = C[am.at(r0, i + M * j)];
        }
        const double d0 = D[av.at(r0, i)];
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < M; ++j) {
            carry[j * kTile + lane] = gamma[j];        // gamma(i, j): slot j, lane base+i
            carry[(M + j) * kTile + lane] = beta[j];   // beta(i, j)
        }
        carry[2 * M * kTile + lane] = alpha;
        wave_lds_sync();
        double r[M], rhs[2 * M + 1];
        rhs[0] = d0;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            r[j] = i == j ? 1.0 : 0.0;
            rhs[1 + j] = a0[j];
            rhs[1 + M + j] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double ck = c0[k];
            rhs[0] = fma(-ck, carry[2 * M * kTile + base + k], rhs[0]);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                r[j] = fma(-ck, carry[(M + j) * kTile + base + k], r[j]);
                rhs[1 + M + j] = fma(-ck, carry[j * kTile + base + k], rhs[1 + M + j]);
            }
        }
        double xc[2 * M], xd[M];
        ok = !coop_eliminate<M, 2>(r, rhs, xb, i, base, xc, xd) && ok;
        if (mine) {
#pragma unroll
            for (int k = 0; k < M; ++k) {
                redA[rm.at(2 * part, k + M * i)] = xc[k];
                redC[rm.at(2 * part, k + M * i)] = xc[M + k];
            }
            redD[rv.at(2 * part, i)] = lane_elem<M>(xd, i);
        }
    }
    if (mine && !ok) report_singular(flag, sys);
}

// ---------------------------------------------------------------------------- reduced --
// Plain solve of the reduced system, 2P rows with B = I, in the tiled reduced buffers.
// The solution replaces redD.
template <int M>
__global__ __launch_bounds__(64) void k_reduced(const double* __restrict__ redA, double* redC,
                                                double* redD, int nsys, int nrr,
                                                int* __restrict__ flag) {
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) return;
    const Addr<LAYOUT_TILED> am(sys, 0, nrr, M * M), av(sys, 0, nrr, M);
    double Cp[M][M], dp[M];
    bool ok = true;
    for (int q = 0; q < nrr; ++q) {
        double R[M][M], v[M];
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) R[i][j] = (i == j) ? 1.0 : 0.0;
        load_vec<M>(redD, av, q, v);
        if (q > 0) gemm_sub<M>(redA, am, q, Cp, dp, R, v);
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, v);
#pragma unroll
        for (int i = 0; i < M; ++i) dp[i] = v[i];
        store_vec<M>(redD, av, q, dp);
        if (q + 1 < nrr) {
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double c[M];
                load_col<M>(redC, am, q, j, c);
                lu_solve<M>(R, piv, c);
#pragma unroll
                for (int i = 0; i < M; ++i) Cp[i][j] = c[i];
                store_col<M>(redC, am, q, j, c);
            }
        }
    }
    for (int q = nrr - 2; q >= 0; --q) {
        double x[M], c[M][M];
        load_vec<M>(redD, av, q, x);
        load_mat_c<M>(redC, am, q, c);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) x[i] = fma(-c[i][j], dp[j], x[i]);
#pragma unroll
        for (int i = 0; i < M; ++i) dp[i] = x[i];
        store_vec<M>(redD, av, q, x);
    }
    if (!ok) report_singular(flag, sys);
}

// Wavefronts per workgroup of k_modi_coop (each on its own systems): 2 at m >= 9, where the
// kernel holds one wavefront per SIMD, measured 2-11% faster on m = 10 (256-4096 systems,
// 1024-2048 rows, P = 16-128); at m = 7, 8 one and two were within 3%, and workgroups whose
// wavefronts meet at a barrier before each row's loads were 3-20% slower in every case.
template <int M>
constexpr int modi_coop_waves() { return M >= 9 ? 2 : 1; }

// Smallest m whose stage 1 runs cooperatively (k_modi_coop) instead of one lane per partition.
#ifndef BTD_MODI_COOP_FROM
#define BTD_MODI_COOP_FROM 7
#endif
constexpr int kModiCoopFrom = BTD_MODI_COOP_FROM;

// Below kModiCoopFrom, stage 1 runs one lane per partition from this many partitions in the
// batch (nsys * npart), with the next row loaded ahead from kModiCoopBelow, and m lanes per
// partition below that (launch_two_stage_m).
#ifndef BTD_MODI_LANE_FROM
#define BTD_MODI_LANE_FROM 65536
#endif
constexpr int64_t kModiLaneFrom = BTD_MODI_LANE_FROM;
#ifndef BTD_MODI_COOP_BELOW
#define BTD_MODI_COOP_BELOW 16384
#endif
constexpr int64_t kModiCoopBelow = BTD_MODI_COOP_BELOW;
// From kModiTwistedFrom partitions in the batch (and below kModiTwistedBelow in the tiled layout),
// non-periodic solves whose partitions have at least 4 rows run both stages twisted
// (k_modi_twisted, k_update_twisted) instead; below kModiTwistedFrom the m lanes per partition of
// k_modi_coop give more lanes than the twisted form's two. Measured (m = 5, whole solve, twisted
// against the other forms): reference layout 16384 x 2048 P = 4 10.2 vs 20.2 ms, 512 x 65536
// P = 128 11.3 vs 20.1, 32768 x 512 P = 2 5.4 vs 9.7 (one lane per partition in that layout reads
// half as fast); tiled, from 65536 partitions, within -8% / +9% of one lane per partition
// (16384 x 2048 P = 4 10.4 vs 9.6, 32768 x 512 P = 2 4.8 vs 5.3), so one lane stays there; at
// 4096 partitions (512 x 65536 P = 8) 31.8 vs 30.2 ms cooperative, at 8192 (P = 16) 17.0 vs 20.5.
#ifndef BTD_MODI_TWISTED_BELOW
#define BTD_MODI_TWISTED_BELOW 65536
#endif
constexpr int64_t kModiTwistedBelow = BTD_MODI_TWISTED_BELOW;
#ifndef BTD_MODI_TWISTED_FROM
#define BTD_MODI_TWISTED_FROM 8192
#endif
constexpr int64_t kModiTwistedFrom = BTD_MODI_TWISTED_FROM;
#ifndef BTD_TWISTED_REF_AHEAD_MAX_M
#define BTD_TWISTED_REF_AHEAD_MAX_M 5
#endif
template <int M>
constexpr bool kTwistedRefAhead = M <= BTD_TWISTED_REF_AHEAD_MAX_M;

// Largest m for which the few-partition forms run (two lanes per partition from both ends, m
// lanes per partition in stage 1 and 3, the next row loaded ahead in stage 1): they were
// verified at m = 2-5 in both layouts. At m = 6 a
// solve with the reference layout, 8195 systems x 4 partitions (stage 1 loading ahead, whose m = 6
// build spills 1.2-1.3 KB per lane to scratch; stage 3 m lanes per partition, run for the first
// time at m = 6) ended in an illegal memory access whose cause reading the code has not shown, so
// from m = 6 both stages keep one lane per partition without loading ahead, as before these
// forms existed.
constexpr int kFewFormsMaxM = 5;

// Largest number of wavefronts per CU for which the reduced system is solved twisted.
#ifndef BTD_TWISTED_REDUCED_WAVES_PER_CU
#define BTD_TWISTED_REDUCED_WAVES_PER_CU 4
#endif
constexpr int kTwistedReducedWavesPerCU = BTD_TWISTED_REDUCED_WAVES_PER_CU;

// The reduced system (nrr = 2P rows, B = I, tiled reduced buffers) solved twisted: per system
// two groups of M lanes, one eliminating from the first row down, the other from the last row
// up, meeting in the middle (coop_twisted), batch_width(M) systems per wavefront. k_reduced
// gives each system one lane and takes nrr row latencies; this takes about nrr/2 shorter
// ones, which is what matters when there are few systems with many partitions. The result
// differs from k_reduced's only by rounding; redC is left overwritten.
template <int M>
__global__ __launch_bounds__(64) void k_reduced_twisted(const double* __restrict__ redA, double* redC,
                                                        double* redD, int nsys, int nrr,
                                                        int* __restrict__ flag) {
    __shared__ double sh[kCoopSlotsFor<M> * kTile];
    constexpr int W = batch_width(M);
    const int lane = threadIdx.x;
    const int g = lane / M, i = lane % M, slot = g / 2;
    const bool grouped = slot < W;
    const bool top = grouped ? (g % 2 == 0) : true;
    const int64_t first = (int64_t)blockIdx.x * W;  // < nsys: the grid is ceil(nsys / W)
    const int64_t sys = first + (grouped ? slot : 0);
    const bool mine = grouped && sys < nsys;
    const int64_t s_eff = mine ? sys : first;
    const int64_t tile = s_eff / kTile, l = s_eff % kTile;
    const StridedAddr<M * M * kTile, kTile> am{tile * nrr * M * M * kTile + l};
    const StridedAddr<M * kTile, kTile> av{tile * nrr * M * kTile + l};
    const int steps = (nrr - 1) - (nrr - 1) / 2;
    const int base = grouped ? g * M : 0;
    const int pbase = grouped ? (top ? base + M : base - M) : M;
    const bool ok = coop_twisted<M, true>(redA, nullptr, redC, redD, am, av, i, base, pbase, top,
                                          mine ? nrr : 0, steps, nrr, sh);
    if (!ok && mine) report_singular(flag, sys);
}

// Smallest m whose single-partition solve runs cooperatively (k_plain_coop).
#ifndef BTD_PLAIN_COOP_FROM
#define BTD_PLAIN_COOP_FROM 8
#endif
constexpr int kPlainCoopFrom = BTD_PLAIN_COOP_FROM;

// Single-partition solve, cooperative: coop_rows with m lanes per system (lane i owns row i of
// every block), floor(64/m) consecutive systems per wavefront, WV wavefronts per workgroup on
// consecutive systems, either layout; the workgroups are cut into one contiguous
// range per XCD, taken in order. For m >= 8 batches large enough to fill the device this way
// (one lane per system would hold too many registers). Same results as k_plain to rounding; C
// is left holding C'. With LOCK the wavefronts of a workgroup meet at a barrier before each
// row's loads, so the parts of a cache line that they read are requested together and the
// line is fetched once; unlocked, they drift apart by more than the line stays in L2 (one
// wavefront reads 8 systems, half a line, at m = 8: twice the compulsory bytes).
template <int M, int LAYOUT, int WV, bool LOCK>
__global__ __launch_bounds__(64 * WV) void k_plain_coop(const double* A, const double* __restrict__ B,
                                                        double* C, double* D, int nsys, int nrow,
                                                        int* __restrict__ flag) {
    __shared__ double sh[WV][kCoopSlotsFor<M> * kTile];
    constexpr int G = kTile / M;
    const int nitems = (nsys + G * WV - 1) / (G * WV);
    const int per = (nitems + kXcds - 1) / kXcds;
    const int u = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (u >= nitems) return;
    const int w = threadIdx.x / kTile, lane = threadIdx.x % kTile;
    const int g = lane / M, i = lane % M;
    const bool grouped = g < G;
    const int64_t s0 = ((int64_t)u * WV + w) * G;
    const int64_t sys = s0 + (grouped ? g : 0);
    const bool mine = grouped && sys < nsys;
    const int64_t se = mine ? sys : (s0 < nsys ? s0 : (int64_t)u * WV * G);  // a real system either way
    const Addr<LAYOUT> am(se, nsys, nrow, M * M), av(se, nsys, nrow, M);
    // lanes past the last whole group read group 0's posts and write only their own slots
    const bool ok = coop_rows<M, LOCK>(A, B, C, D, am, av, i, grouped ? g * M : 0, mine ? nrow : 0, nrow, nrow, sh[w]);
    if (!ok && mine) report_singular(flag, sys);
}

// Cooperative plain solve (npart = 1), twisted: per system two groups of M lanes
// (coop_twisted), batch_width(M) consecutive systems per wavefront and WV wavefronts per
// workgroup on consecutive systems, either layout, so a plan of few systems fills the device
// with 2m lanes per system instead of m. The workgroups are cut into one contiguous range per
// XCD, so neighbouring systems' shared lines meet in one L2; with LOCK the wavefronts of a
// workgroup meet at a barrier before each row's loads (coop_rows). C is left overwritten.
template <int M, int LAYOUT, int WV, bool LOCK>
__global__ __launch_bounds__(64 * WV) void k_plain_twisted(const double* A, const double* __restrict__ B,
                                                           double* C, double* D, int nsys, int nrow,
                                                           int* __restrict__ flag) {
    __shared__ double sh[WV][kCoopSlotsFor<M> * kTile];
    constexpr int W = batch_width(M);
    const int nitems = (nsys + W * WV - 1) / (W * WV);
    const int per = (nitems + kXcds - 1) / kXcds;
    const int u = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (u >= nitems) return;
    const int w = threadIdx.x / kTile, lane = threadIdx.x % kTile;
    const int g = lane / M, i = lane % M, slot = g / 2;
    const bool grouped = slot < W;
    const bool top = grouped ? (g % 2 == 0) : true;
    const int64_t s0 = ((int64_t)u * WV + w) * W;
    const int64_t sys = s0 + (grouped ? slot : 0);
    const bool mine = grouped && sys < nsys;
    const int64_t se = mine ? sys : (s0 < nsys ? s0 : (int64_t)u * WV * W);  // a real system either way
    const Addr<LAYOUT> am(se, nsys, nrow, M * M), av(se, nsys, nrow, M);
    const int steps = (nrow - 1) - (nrow - 1) / 2;
    // lanes past the last pair act as a top group over groups 0 and 1, writing only their
    // own lane's slots
    const int base = grouped ? g * M : 0;
    const int pbase = grouped ? (top ? base + M : base - M) : M;
    const bool ok = coop_twisted<M, false, LOCK>(A, B, C, D, am, av, i, base, pbase, top, mine ? nrow : 0,
                                                 steps, nrow, sh[w]);
    if (!ok && mine) report_singular(flag, sys);
}

template <int M, int LAYOUT, int WV, bool LOCK>
void launch_plain_twisted_as(const double* A, const double* B, double* C, double* D, int nsys, int nrow,
                             int* flag, hipStream_t s) {
    constexpr int per_item = batch_width(M) * WV;
    const int64_t items = ((int64_t)nsys + per_item - 1) / per_item;
    const dim3 grid((unsigned)(kXcds * ((items + kXcds - 1) / kXcds))), block(64 * WV);
    hipLaunchKernelGGL((k_plain_twisted<M, LAYOUT, WV, LOCK>), grid, block, 0, s, A, B, C, D, nsys, nrow, flag);
}

// Workgroup shape of the twisted solve, as measured (m = 2-10; 1024-16384 systems of 512-4096
// rows, both layouts): m >= 9 two locked wavefronts (1024 x 2048: 16.3 -> 13.5 ms); m <= 8
// four locked wavefronts (as in k_plain_coop) from kTwistedLockFrom twisted wavefronts of work
// per CU (m = 8 4096 x 4096 39.8 -> 25.2 ms, m = 7 16384 x 2048 52.5 -> 31.7 ms, m = 3
// 16384 x 2048 8.6 -> 4.8 ms; reference layout), one wavefront below. Between 2 and 4 per CU
// the locked shape won at m = 3, 5 (8192 x 1024 m = 3: 2.20 -> 1.32 ms) and lost at m = 2, 7, 8
// (2048 x 1024 m = 8: 4.44 -> 6.12 ms).
template <int M>
constexpr int kTwistedLockFrom = (M >= 3 && M <= 6) ? 2 : 4;

template <int M, int LAYOUT>
void launch_plain_twisted_auto(const double* A, const double* B, double* C, double* D, int nsys, int nrow,
                               int* flag, hipStream_t s) {
    if constexpr (M >= 9) {
        launch_plain_twisted_as<M, LAYOUT, 2, true>(A, B, C, D, nsys, nrow, flag, s);
    } else {
        const int64_t waves = ((int64_t)nsys + batch_width(M) - 1) / batch_width(M);
        if (waves >= kTwistedLockFrom<M> * (int64_t)device_cus())
            launch_plain_twisted_as<M, LAYOUT, 4, true>(A, B, C, D, nsys, nrow, flag, s);
        else
            launch_plain_twisted_as<M, LAYOUT, 1, false>(A, B, C, D, nsys, nrow, flag, s);
    }
}

template <int M>
hipError_t launch_plain_twisted_m(int layout, const double* A, const double* B, double* C, double* D,
                                  int nsys, int nrow, int* flag, hipStream_t s) {
    if constexpr (M < 2) {
        return hipErrorInvalidValue;
    } else {
        if (layout == LAYOUT_TILED)
            launch_plain_twisted_auto<M, LAYOUT_TILED>(A, B, C, D, nsys, nrow, flag, s);
        else
            launch_plain_twisted_auto<M, LAYOUT_REF_GPU>(A, B, C, D, nsys, nrow, flag, s);
        return hipGetLastError();
    }
}

// Banded outer-product gradients of the FP32-stored solve, computed in FP64 and stored in
// FP32: dA_q = -lambda_q x_{q-1}^T, dB_q = -lambda_q x_q^T, dC_q = -lambda_q x_{q+1}^T.
template <int M, int LAYOUT>
__global__ __launch_bounds__(64) void k_outer_f32(const double* __restrict__ Lam,
                                                  const double* __restrict__ X,
                                                  float* __restrict__ dA, float* __restrict__ dB,
                                                  float* __restrict__ dC, int nsys, int nrow) {
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) {
        zero_pad_lane<M * M, LAYOUT>(dA, sys, nsys, nrow);
        zero_pad_lane<M * M, LAYOUT>(dB, sys, nsys, nrow);
        zero_pad_lane<M * M, LAYOUT>(dC, sys, nsys, nrow);
        return;
    }
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
    const int64_t es = am.estride();
    double xm[M], xc[M], xp[M];
#pragma unroll
    for (int i = 0; i < M; ++i) xm[i] = 0.0;
    load_vec<M>(X, av, 0, xc);
    for (int64_t q = 0; q < nrow; ++q) {
        double l[M];
        load_vec<M>(Lam, av, q, l);
        if (q + 1 < nrow) {
            load_vec<M>(X, av, q + 1, xp);
        } else {
#pragma unroll
            for (int i = 0; i < M; ++i) xp[i] = 0.0;
        }
        float* ra = dA + am.at(q, 0);
        float* rb = dB + am.at(q, 0);
        float* rc = dC + am.at(q, 0);
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int i = 0; i < M; ++i) {
                ra[(i + M * j) * es] = (float)(-l[i] * xm[j]);
                rb[(i + M * j) * es] = (float)(-l[i] * xc[j]);
                rc[(i + M * j) * es] = (float)(-l[i] * xp[j]);
            }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            xm[i] = xc[i];
            xc[i] = xp[i];
        }
    }
}

// ----------------------------------------------------------------------------- cyclic --
// Periodic block Thomas, btdma_many_cycl (mod_btdma_cpu.f90:80-183; GPU
// mod_btdma_gpu_v2.f90:352-616): row 0 couples to x_{n-1} through A_0 and row n-1 to x_0
// through C_{n-1}. x_0 is kept as a separate unknown: rows 1..n-1 are eliminated with a fill-in
// block column E (E_1 = -A_1, E_{n-1} = -C_{n-1}, otherwise 0), so that after the forward and
// backward sweeps x_q = d'_q + E'_q x_0; then row 0 gives
//   (B_0 + A_0 E'_{n-1} + C_0 E'_1) x_0 = d_0 - A_0 d'_{n-1} - C_0 d'_1.
// In place: E'_q over A_q, C'_q over C_q, d'_q over D (rows >= 1); needs n >= 3. BI = true
// solves a system whose B blocks are the identity (the reduced system of the two-stage
// solve), B is then not read.
template <int M, int LAYOUT, bool BI>
__device__ __forceinline__ bool cyclic_solve(double* A, const double* __restrict__ B, double* C,
                                             double* D, const Addr<LAYOUT>& am,
                                             const Addr<LAYOUT>& av, int nrow) {
    bool ok = true;
    double Cp[M][M], Ep[M][M], dp[M];
    for (int q = 1; q < nrow; ++q) {
        const bool last = q == nrow - 1;
        double R[M][M], v[M], X[M][M];
        if constexpr (BI) {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) R[i][j] = (i == j) ? 1.0 : 0.0;
        } else {
            load_mat<M>(B, am, q, R);
        }
        load_vec<M>(D, av, q, v);
        if (q == 1) {
            load_mat_c<M>(A, am, q, X);
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) X[i][j] = -X[i][j];
        } else {
            if (last) {
                load_mat_c<M>(C, am, q, X);
#pragma unroll
                for (int i = 0; i < M; ++i)
#pragma unroll
                    for (int j = 0; j < M; ++j) X[i][j] = -X[i][j];
            } else {
#pragma unroll
                for (int i = 0; i < M; ++i)
#pragma unroll
                    for (int j = 0; j < M; ++j) X[i][j] = 0.0;
            }
#pragma unroll
            for (int k = 0; k < M; ++k) {
                double ak[M];
                load_col<M>(A, am, q, k, ak);
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    v[i] = fma(-ak[i], dp[k], v[i]);
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        R[i][j] = fma(-ak[i], Cp[k][j], R[i][j]);
                        X[i][j] = fma(-ak[i], Ep[k][j], X[i][j]);
                    }
                }
            }
        }
        // C_q with the row's other loads (small m): C is written in place, so a column loaded
        // after the previous column's store would cost one more memory round trip per column
        constexpr bool kCEarly = M <= 5;
        double Cq[kCEarly ? M : 1][kCEarly ? M : 1];
        if constexpr (kCEarly) {
            if (!last) load_mat_c<M>(C, am, q, Cq);
        }
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, v);
        solve_rows<M>(R, piv, X);
        if (!last) {
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double c[M];
                if constexpr (kCEarly) {
#pragma unroll
                    for (int i = 0; i < M; ++i) c[i] = Cq[i][j];
                } else {
                    load_col<M>(C, am, q, j, c);
                }
                lu_solve<M>(R, piv, c);
#pragma unroll
                for (int i = 0; i < M; ++i) Cp[i][j] = c[i];
                store_col<M>(C, am, q, j, c);
            }
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            dp[i] = v[i];
#pragma unroll
            for (int j = 0; j < M; ++j) Ep[i][j] = X[i][j];
        }
        store_mat<M>(A, am, q, Ep);
        store_vec<M>(D, av, q, dp);
    }
    // backward, q = n-2..1: d'_q -= C'_q d'_{q+1}; E'_q -= C'_q E'_{q+1} ((Ep, dp) = row q+1)
    for (int q = nrow - 2; q >= 1; --q) {
        double c[M][M], e[M][M], d[M];
        load_mat_c<M>(C, am, q, c);
        load_mat_c<M>(A, am, q, e);
        load_vec<M>(D, av, q, d);
#pragma unroll
        for (int i = 0; i < M; ++i) {
#pragma unroll
            for (int k = 0; k < M; ++k) d[i] = fma(-c[i][k], dp[k], d[i]);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double s = e[i][j];
#pragma unroll
                for (int k = 0; k < M; ++k) s = fma(-c[i][k], Ep[k][j], s);
                e[i][j] = s;
            }
        }
        store_mat<M>(A, am, q, e);
        store_vec<M>(D, av, q, d);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            dp[i] = d[i];
#pragma unroll
            for (int j = 0; j < M; ++j) Ep[i][j] = e[i][j];
        }
    }
    // row 0; (Ep, dp) now hold row 1
    double x0[M];
    {
        double a0[M][M], c0[M][M], en[M][M], dn[M], R[M][M];
        load_mat_c<M>(A, am, 0, a0);
        load_mat_c<M>(C, am, 0, c0);
        load_mat_c<M>(A, am, nrow - 1, en);
        load_vec<M>(D, av, nrow - 1, dn);
        load_vec<M>(D, av, 0, x0);
        if constexpr (BI) {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) R[i][j] = (i == j) ? 1.0 : 0.0;
        } else {
            load_mat_c<M>(B, am, 0, R);
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
#pragma unroll
            for (int k = 0; k < M; ++k) x0[i] = fma(-a0[i][k], dn[k], fma(-c0[i][k], dp[k], x0[i]));
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double s = R[i][j];
#pragma unroll
                for (int k = 0; k < M; ++k) s = fma(a0[i][k], en[k][j], fma(c0[i][k], Ep[k][j], s));
                R[i][j] = s;
            }
        }
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, x0);
        store_vec<M>(D, av, 0, x0);
    }
    for (int q = 1; q < nrow; ++q) {
        double e[M][M], d[M];
        load_mat_c<M>(A, am, q, e);
        load_vec<M>(D, av, q, d);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int k = 0; k < M; ++k) d[i] = fma(e[i][k], x0[k], d[i]);
        store_vec<M>(D, av, q, d);
    }
    return ok;
}

template <int M, int LAYOUT>
__global__ __launch_bounds__(64) void k_cyclic(double* A, const double* __restrict__ B, double* C,
                                               double* D, int nsys, int nrow,
                                               int* __restrict__ flag) {
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) return;
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
    if (!cyclic_solve<M, LAYOUT, false>(A, B, C, D, am, av, nrow)) report_singular(flag, sys);
}

// Periodic reduced system (2P rows, B = I) of the two-stage cyclic solve
// (btdma_many_cyclic_mpi, mod_btdma_cpu.f90:273-302).
template <int M>
__global__ __launch_bounds__(64) void k_reduced_cyclic(double* redA, double* redC, double* redD,
                                                       int nsys, int nrr, int* __restrict__ flag) {
    const int64_t sys = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (sys >= nsys) return;
    const Addr<LAYOUT_TILED> am(sys, 0, nrr, M * M), av(sys, 0, nrr, M);
    if (!cyclic_solve<M, LAYOUT_TILED, true>(redA, nullptr, redC, redD, am, av, nrr))
        report_singular(flag, sys);
}

// ----------------------------------------------------------------------------- update --
// Stage 3 (btdma_many_update, mod_btdma_cpu.f90:511-535): x_1 and x_N of each partition come
// from the reduced solution; the interior rows follow from the forward values of stage 1 by
// x_q = d̂_q - Â_q x_1 - Ĉ_q x_{q+1}, q = N-2 .. 1 (the reference uses the backward-rewritten
// rows and x_N instead; the two are equal in exact arithmetic).
template <int M, int LAYOUT>
__global__ __launch_bounds__(64) void k_update(const double* __restrict__ A,
                                               const double* __restrict__ C, double* D, int nsys,
                                               int nrow, const int* __restrict__ part_first,
                                               const int* __restrict__ part_len, int npart,
                                               const double* __restrict__ redD) {
    const int ntiles = (nsys + 63) / 64;
    const int tile = blockIdx.x % ntiles, part = blockIdx.x / ntiles;
    const int64_t sys = (int64_t)tile * 64 + threadIdx.x;
    if (sys >= nsys) return;
    const int64_t r0 = part_first[part];
    const int N = part_len[part];
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
    const Addr<LAYOUT_TILED> rv(sys, 0, 2 * npart, M);
    double x1[M], xn[M];
    load_vec<M>(redD, rv, 2 * part, x1);
    load_vec<M>(redD, rv, 2 * part + 1, xn);
    store_vec<M>(D, av, r0, x1);
    store_vec<M>(D, av, r0 + N - 1, xn);
    for (int q = N - 2; q >= 1; --q) {
        const int64_t g = r0 + q;
        double a[M][M], c[M][M], x[M];
        load_mat<M>(A, am, g, a);
        load_mat<M>(C, am, g, c);
        load_vec<M>(D, av, g, x);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) x[i] = fma(-a[i][j], x1[j], fma(-c[i][j], xn[j], x[i]));
#pragma unroll
        for (int i = 0; i < M; ++i) xn[i] = x[i];
        store_vec<M>(D, av, g, x);
    }
}

// Stage 3 after k_modi_twisted, the same two wavefronts per tile and partition: x_0' and x_{N-1}
// come from the reduced solution; each lane walks its half back toward its anchor x_a,
//   x_q = d̂_q - Â_q x_a - Ĉ_q x_s,   q = p .. 1 of its sweep,
// x_s starting at the partner's anchor (row p's Ĉ is its coefficient) and then the row just
// computed; the next row's blocks are loaded while the current one is formed.
template <int M, int LAYOUT>
__global__ __launch_bounds__(2 * kTile) void k_update_twisted(const double* __restrict__ A,
                                                              const double* __restrict__ C, double* D, int nsys,
                                                              int nrow, const int* __restrict__ part_first,
                                                              const int* __restrict__ part_len, int npart,
                                                              const double* __restrict__ redD) {
    const int ntiles = (nsys + kTile - 1) / kTile;
    const int tile = blockIdx.x % ntiles, part = blockIdx.x / ntiles;
    const bool top = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile) == 0;
    const int64_t sys = (int64_t)tile * kTile + threadIdx.x % kTile;
    if (sys >= nsys) return;
    const int64_t r0 = part_first[part];
    const int N = part_len[part];  // >= 4
    const int k = (N + 1) / 2 - 1;
    const int L = top ? k + 1 : N - 1 - k;
    const int64_t g0 = top ? r0 : r0 + N - 1, dg = top ? 1 : -1;
    const double* Ar = top ? A : C;
    const double* Cr = top ? C : A;
    const Addr<LAYOUT> am(sys, nsys, nrow, M * M), av(sys, nsys, nrow, M);
    const Addr<LAYOUT_TILED> rv(sys, 0, 2 * npart, M);
    double xa[M], xs[M];
    load_vec<M>(redD, rv, 2 * part + (top ? 0 : 1), xa);
    load_vec<M>(redD, rv, 2 * part + (top ? 1 : 0), xs);
    store_vec<M>(D, av, g0, xa);
    double na[M][M], nc[M][M], nx[M];
    auto fetch = [&](int q) {
        const int64_t g = g0 + q * dg;
        load_mat<M>(Ar, am, g, na);
        load_mat<M>(Cr, am, g, nc);
        load_vec<M>(D, av, g, nx);
    };
    fetch(L - 1);
    for (int q = L - 1; q >= 1; --q) {
        double a[M][M], c[M][M], x[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            x[i] = nx[i];
#pragma unroll
            for (int j = 0; j < M; ++j) {
                a[i][j] = na[i][j];
                c[i][j] = nc[i][j];
            }
        }
        fetch(q > 1 ? q - 1 : q);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) x[i] = fma(-a[i][j], xa[j], fma(-c[i][j], xs[j], x[i]));
#pragma unroll
        for (int i = 0; i < M; ++i) xs[i] = x[i];
        store_vec<M>(D, av, g0 + q * dg, x);
    }
}

// Stage 3 with a partition's block rows over m lanes (lane i owns row i), for the batches whose
// stage 1 runs k_modi_coop: the same x_q = d̂_q - Â_q x_1 - Ĉ_q x_{q+1} (the same operation
// order as k_update, so the same bits), with x_{q+1} exchanged through LDS and the rows loaded
// kAhead ahead; floor(64/m) consecutive systems of one partition per wavefront, WV wavefronts
// (LOCK: meeting before each row's loads, as in k_modi_coop) per workgroup, the items of a
// partition on one XCD.
template <int M>
struct UpdateDepth {
    static constexpr int kAhead = M <= 2 ? 12 : M <= 3 ? 8 : M <= 5 ? 5 : 3;
};

template <int M, int LAYOUT, int WV = 1, bool LOCK = false>
__global__ __launch_bounds__(64 * WV) void k_update_coop(const double* __restrict__ A,
                                                         const double* __restrict__ C, double* D, int nsys,
                                                         int nrow, const int* __restrict__ part_first,
                                                         const int* __restrict__ part_len, int npart,
                                                         const double* __restrict__ redD) {
    constexpr int G = kTile / M;  // systems per wavefront
    constexpr int K = UpdateDepth<M>::kAhead;
    __shared__ double shw[WV][kTile];
    const int nchunks = (nsys + G - 1) / G;
    const int ngroups = (nchunks + WV - 1) / WV;
    const int nitems = npart * ngroups;
    const int per = (nitems + kXcds - 1) / kXcds;
    const int u = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (u >= nitems) return;
    const int w = threadIdx.x / kTile, lane = threadIdx.x % kTile;
    const int part = u / ngroups, chunk0 = (u % ngroups) * WV + w;
    if (!LOCK && chunk0 >= nchunks) return;  // the wavefronts share no barrier
    const bool live = chunk0 < nchunks;
    const int chunk = live ? chunk0 : nchunks - 1;
    double* sh = shw[w];
    const int g = lane / M, i = lane % M;
    const bool grouped = g < G;
    const int64_t s0 = (int64_t)chunk * G;
    const int64_t sys = s0 + (grouped ? g : 0);
    const bool mine = live && grouped && sys < nsys;
    const int64_t se = grouped && sys < nsys ? sys : s0;  // a real system either way
    const int base = grouped ? g * M : 0;
    const Addr<LAYOUT> am(se, nsys, nrow, M * M), av(se, nsys, nrow, M);
    const Addr<LAYOUT_TILED> rv(se, 0, 2 * npart, M);
    const int64_t r0 = part_first[part];
    const int N = part_len[part];  // >= 3
    double x1[M];
#pragma unroll
    for (int j = 0; j < M; ++j) x1[j] = redD[rv.at(2 * part, j)];
    double xn = redD[rv.at(2 * part + 1, i)];  // element i of x_{q+1}
    if (mine) {
        D[av.at(r0, i)] = lane_elem<M>(x1, i);
        D[av.at(r0 + N - 1, i)] = xn;
    }
    double pa[K][M], pc[K][M], pd[K];
    auto fetch = [&](double (&fa)[M], double (&fc)[M], double& fd, int q) {
        const int64_t gr = r0 + (q < 1 ? 1 : q);  // rows 1 .. N-2
#pragma unroll
        for (int j = 0; j < M; ++j) {
            fa[j] = A[am.at(gr, i + M * j)];
            fc[j] = C[am.at(gr, i + M * j)];
        }
        fd = D[av.at(gr, i)];
    };
#pragma unroll
    for (int t = 0; t < K; ++t) fetch(pa[t], pc[t], pd[t], N - 2 - t);
    for (int q0 = N - 2; q0 >= 1; q0 -= K)
#pragma unroll
        for (int t = 0; t < K; ++t) {
            const int q = q0 - t;
            if (q < 1) break;  // uniform: every lane of the wavefront is in this partition
            sh[lane] = xn;
            wave_lds_sync();
            double x = pd[t];
#pragma unroll
            for (int j = 0; j < M; ++j) x = fma(-pa[t][j], x1[j], fma(-pc[t][j], sh[base + j], x));
            wave_lds_sync();
            if constexpr (LOCK) __builtin_amdgcn_s_barrier();
            fetch(pa[t], pc[t], pd[t], q - K);
            if (mine) D[av.at(r0 + q, i)] = x;
            xn = x;
        }
}

// --------------------------------------------------------------------------- generate --
template <int M>
__global__ __launch_bounds__(256) void k_generate(int kind, uint64_t seed, int layout, int nsys,
                                                  int nrow, double* A, double* B, double* C,
                                                  double* D, float* Af, float* Bf, float* Cf) {
    const int64_t npad = pad_nsys(nsys);
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= npad * nrow) return;
    const int64_t row = idx / npad, sys = idx % npad;
    if (sys >= nsys && layout != LAYOUT_TILED) return;  // only the tiled layout has padding lanes
    double a[M][M], b[M][M], c[M][M], d[M];
    if (sys < nsys) {
        gen_row<M>(kind, seed, sys, row, nrow, a, b, c, d);
    } else {  // the tiled layout's padding lanes: zeros, so every element of the arrays is defined
        for (int j = 0; j < M; ++j) {
            d[j] = 0.0;
            for (int i = 0; i < M; ++i) a[i][j] = b[i][j] = c[i][j] = 0.0;
        }
    }
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            const int64_t o = elem_offset(layout, sys, row, i + M * j, M * M, nsys, nrow);
            if (A) A[o] = a[i][j];
            if (B) B[o] = b[i][j];
            if (C) C[o] = c[i][j];
            if (Af) Af[o] = (float)a[i][j];
            if (Bf) Bf[o] = (float)b[i][j];
            if (Cf) Cf[o] = (float)c[i][j];
        }
    if (D)
        for (int i = 0; i < M; ++i) D[elem_offset(layout, sys, row, i, M, nsys, nrow)] = d[i];
}

// ---------------------------------------------------------------------- launch templates --
template <int M, int LAYOUT, int WV, bool LOCK>
void launch_plain_coop_as(const double* A, const double* B, double* C, double* D, int nsys, int nrow,
                          int* flag, hipStream_t s) {
    constexpr int per_item = (kTile / M) * WV;
    const int64_t items = ((int64_t)nsys + per_item - 1) / per_item;
    const dim3 grid((unsigned)(kXcds * ((items + kXcds - 1) / kXcds))), block(64 * WV);
    hipLaunchKernelGGL((k_plain_coop<M, LAYOUT, WV, LOCK>), grid, block, 0, s, A, B, C, D, nsys, nrow, flag);
}

// The single-partition solve for m >= 8, as measured (m = 8-10, 1024-32768 systems of
// 128-4096 rows, both layouts), by the wavefronts of cooperative work per CU (floor(64/m)
// systems a wavefront):
// * m = 8 (two cooperative wavefronts per SIMD), above kCoopLockFrom per CU (2 reference, 4
//   tiled: above 4096 / 8192 systems on 256 CUs): cooperative, locked workgroups of 2 (tiled)
//   or 4 (reference) wavefronts, 13-33% faster than one unlocked wavefront, reading the
//   compulsory bytes instead of twice them (16384 x 2048: 43.4 -> 34.2 ms reference, 36.4 ->
//   24.6 ms tiled); up to that the twisted form (latency-bound: two groups per system halve the
//   sequential rows): 4096 x 4096 35.8 -> 23.8 ms (reference), 1024 x 2048 10.3 -> 5.7 ms. The
//   twisted form's time jumps past those sizes (its wavefronts of four systems need another
//   round: 4096 -> 5120 x 2048 11.7 -> 20.6 ms reference, 8192 -> 8448 14.0 -> 20.1 ms tiled),
//   where the cooperative one is 0.82-0.88x its time (5120-7168 reference, 8448-12288 tiled).
// * m >= 9 (one wavefront per SIMD): twisted while its wavefronts (floor(32/m) systems each)
//   fit one per SIMD (up to 3072 systems at m = 9, 10 on 256 CUs: 1024 x 2048 26.4 -> 13.5 ms
//   reference, 18.5 -> 10.6 ms tiled against cooperative), cooperative in unlocked workgroups
//   of 4 wavefronts (5-29% faster than one) above (4096 x 1024 15.0 vs 17.6 ms twisted; m = 9
//   3456 x 2048 28.9 -> 23.7 ms reference, where the twisted wavefronts need a second round;
//   m = 10 3072 x 2048 stays twisted, 23.8 vs 27.6 ms cooperative).
#ifndef BTD_COOP_LOCK_FROM
#define BTD_COOP_LOCK_FROM 2
#endif
template <int LAYOUT>
constexpr int kCoopLockFrom = (LAYOUT == LAYOUT_TILED ? 2 : 1) * BTD_COOP_LOCK_FROM;

template <int M, int LAYOUT>
void launch_plain_coop(const double* A, const double* B, double* C, double* D, int nsys, int nrow,
                       int* flag, hipStream_t s) {
    const int64_t cus = device_cus();
    if constexpr (M >= 9) {
        constexpr int W = kTile / 2 / M;  // systems per twisted wavefront
        if (((int64_t)nsys + W - 1) / W <= 4 * cus)
            launch_plain_twisted_auto<M, LAYOUT>(A, B, C, D, nsys, nrow, flag, s);
        else
            launch_plain_coop_as<M, LAYOUT, 4, false>(A, B, C, D, nsys, nrow, flag, s);
    } else {
        const int64_t waves = ((int64_t)nsys + kTile / M - 1) / (kTile / M);
        if (waves <= (int64_t)kCoopLockFrom<LAYOUT> * cus)
            launch_plain_twisted_auto<M, LAYOUT>(A, B, C, D, nsys, nrow, flag, s);
        else if constexpr (LAYOUT == LAYOUT_TILED)
            launch_plain_coop_as<M, LAYOUT, 2, true>(A, B, C, D, nsys, nrow, flag, s);
        else
            launch_plain_coop_as<M, LAYOUT, 4, true>(A, B, C, D, nsys, nrow, flag, s);
    }
}

// Smallest m whose single-partition solve runs twisted below kPlainLaneFrom systems instead of
// one lane per system. Below 32768 systems one lane per system leaves CUs idle, and the
// twisted solve (2m lanes per system) measured faster at every m tried: m = 3 16384 x 2048
// 9.35 -> 4.93 ms, m = 5 22.6 -> 14.6 ms, m = 7 32.5 -> 31.7 ms (reference layout), m = 2, 3
// 4096 x 512 1.06 -> 0.21, 2.06 -> 0.47 ms.
#ifndef BTD_PLAIN_TWISTED_FROM
#define BTD_PLAIN_TWISTED_FROM 2
#endif
constexpr int kPlainTwistedFrom = BTD_PLAIN_TWISTED_FROM;
#ifndef BTD_PLAIN_LANE_FROM
#define BTD_PLAIN_LANE_FROM 32768
#endif
constexpr int kPlainLaneFrom = BTD_PLAIN_LANE_FROM;

template <int M>
void launch_plain_lane(int layout, const double* A, const double* B, const double* C, double* Cw, double* D,
                       int nsys, int nrow, int* flag, hipStream_t s) {
    const dim3 grid((nsys + 63) / 64), block(64);
    if (layout == LAYOUT_TILED)
        hipLaunchKernelGGL((k_plain<M, LAYOUT_TILED>), grid, block, 0, s, A, B, C, Cw, D, nsys, nrow, flag);
    else
        hipLaunchKernelGGL((k_plain<M, LAYOUT_REF_GPU>), grid, block, 0, s, A, B, C, Cw, D, nsys, nrow, flag);
}

// Batches on which the single-partition solve runs two lanes per system from both ends with
// the blocks in registers (k_plain_halves): from 24576 systems at m = 2, 12288 at m = 3-5, 6144
// at m = 6 and above 8192 at m = 7; below 131072 systems at m = 2-4 and 98304 at m = 5, at every
// size from m = 6. Below, the cooperative twisted solve (2m lanes per system) runs; above, one
// lane per system (k_plain). Where a lane's chain of dependent rows bounds the solve rather than
// HBM, each row's operands are loaded while the previous row is eliminated: up to 32768 systems
// at m <= 5 and below 20480 at m = 6 (AHEAD); at m = 7 whole rows ahead spill, so below 32768
// systems only A and d are loaded ahead and A C' is formed while B arrives (LIGHT). Measured on
// 2048-262144 systems of 256-2048 rows, both layouts:
//   - against the twisted solve and the two-stage solve at the library's partition count:
//     m = 5 16384 x 2048 11.0 -> 7.8 ms reference layout (P = 2 before), 12288 x 2048 7.3 ->
//     5.2 ms tiled; m = 3 16384 x 2048 4.2 -> 3.1 ms and m = 6 12288 x 2048 21.0 -> 13.4 ms,
//     6144 x 2048 15.3 -> 12.4 ms reference (5120: 1.4x slower); m = 7 8448-12288 x 2048
//     0.56-0.76x the time (above 8192 systems the twisted solve's wavefronts, four systems each,
//     exceed two per SIMD and need a second round: 8448 x 2048 21.4 -> 14.9 ms reference; at
//     8192 it is 1.02-1.12x faster); slower below those sizes, where the device is idle (m = 5
//     4096 x 2048 2.8 -> 6.1 ms), and at m = 2 below 24576, whose twisted wavefronts read whole
//     halves of lines (16384 x 2048 1.4 -> 1.6 ms);
//   - loads ahead against none: m = 5 16384 x 2048 14.9 -> 7.3 ms reference layout, 7.9 -> 6.8
//     tiled, 32768 x 2048 18.2 -> 14.3 / 13.2 -> 13.2 ms; m = 3 16384 x 2048 4.3 -> 3.1 ms; m = 6
//     16384 x 2048 19.0 -> 14.6 / 12.2 -> 11.9 ms, 24576 x 2048 1.07x slower tiled; from 65536
//     systems (two wavefronts per SIMD) 1.03-1.05x slower; m = 7 (LIGHT) 16384 x 2048 19.0 ->
//     16.0 ms reference, 17.4 -> 14.7 tiled, from 32768 systems within 4% (m = 6 LIGHT: 1.2-1.3x
//     slower than AHEAD);
//   - against one lane per system from 32768 systems: faster while one lane per system fills at
//     most one wavefront per SIMD (m = 5 32768 x 2048 19.9 -> 13.1 ms tiled, 30.0 -> 14.3
//     reference; m = 2 5.7 -> 3.1 ms tiled; m = 4 65536 x 1024 11.2 -> 8.9 ms); beyond, on the
//     same placement of the arrays in HBM (four placements each), m = 5 98304 x 512 1.00-1.13x
//     and 131072 x 512 1.00-1.04x, m = 3 131072 x 512 0.95-1.00x, m = 2, 4 131072 x 512
//     1.16-1.18x slower in the reference layout, while m = 6 stays 0.90-1.00x and m = 7
//     0.78-0.92x up to 262144 systems. k_plain performs the oracle's operations bit for bit, so
//     it keeps the sizes where the two are equal.
#ifndef BTD_PLAIN_HALVES_M_MASK
#define BTD_PLAIN_HALVES_M_MASK 0xFC  // bit m set: m = 2 ... 7
#endif
template <int M>
constexpr bool kPlainHalvesM = M <= 7 && ((BTD_PLAIN_HALVES_M_MASK >> M) & 1);
#ifdef BTD_PLAIN_HALVES_FROM
template <int M>
constexpr int kPlainHalvesFrom = BTD_PLAIN_HALVES_FROM;
#else
template <int M>
constexpr int kPlainHalvesFrom = M == 2 ? 24576 : M <= 5 ? 12288 : M == 6 ? 6144 : 8193;
#endif
#ifdef BTD_PLAIN_HALVES_BELOW
template <int M>
constexpr int kPlainHalvesBelow = BTD_PLAIN_HALVES_BELOW;
#else
template <int M>
constexpr int kPlainHalvesBelow = M >= 6 ? 0x7fffffff : M == 5 ? 98304 : 131072;
#endif
#ifdef BTD_PLAIN_HALVES_AHEAD_BELOW
template <int M>
constexpr int kPlainHalvesAheadBelow = M <= 6 ? BTD_PLAIN_HALVES_AHEAD_BELOW : 0;
#else
template <int M>
constexpr int kPlainHalvesAheadBelow = M <= 5 ? 32769 : M == 6 ? 20480 : 0;
#endif
// m = 7 below 32768 systems: only A and d one row ahead (LIGHT; A C' is formed while B arrives)
template <int M>
constexpr int kPlainHalvesLightBelow = M == 7 ? 32768 : 0;

template <int M, bool AHEAD, bool LIGHT = false>
void launch_plain_halves_as(int layout, const double* A, const double* B, double* C, double* D, int nsys,
                            int nrow, int* flag, hipStream_t s) {
    const dim3 grid((nsys + kTile - 1) / kTile), block(2 * kTile);
    if (layout == LAYOUT_TILED)
        hipLaunchKernelGGL((k_plain_halves<M, LAYOUT_TILED, AHEAD, LIGHT>), grid, block, 0, s, A, B, C, D, nsys,
                           nrow, flag);
    else
        hipLaunchKernelGGL((k_plain_halves<M, LAYOUT_REF_GPU, AHEAD, LIGHT>), grid, block, 0, s, A, B, C, D,
                           nsys, nrow, flag);
}

template <int M>
void launch_plain_halves(int layout, const double* A, const double* B, double* C, double* D, int nsys,
                         int nrow, int* flag, hipStream_t s) {
    if constexpr (kPlainHalvesLightBelow<M> > 0) {
        if (nsys < kPlainHalvesLightBelow<M>) {
            launch_plain_halves_as<M, false, true>(layout, A, B, C, D, nsys, nrow, flag, s);
            return;
        }
    }
    if constexpr (kPlainHalvesAheadBelow<M> > 0) {
        if (nsys < kPlainHalvesAheadBelow<M>) {
            launch_plain_halves_as<M, true>(layout, A, B, C, D, nsys, nrow, flag, s);
            return;
        }
    }
    launch_plain_halves_as<M, false>(layout, A, B, C, D, nsys, nrow, flag, s);
}

template <int M>
hipError_t launch_plain_m(int layout, const double* A, const double* B, double* C, double* D,
                                 int nsys, int nrow, int* flag, hipStream_t s) {
    if constexpr (M >= kPlainCoopFrom) {
        if (layout == LAYOUT_TILED)
            launch_plain_coop<M, LAYOUT_TILED>(A, B, C, D, nsys, nrow, flag, s);
        else
            launch_plain_coop<M, LAYOUT_REF_GPU>(A, B, C, D, nsys, nrow, flag, s);
    } else if constexpr (M >= 2 && M >= kPlainTwistedFrom) {
        if (kPlainHalvesM<M> && nsys >= kPlainHalvesFrom<M> && nsys < kPlainHalvesBelow<M>)
            launch_plain_halves<M>(layout, A, B, C, D, nsys, nrow, flag, s);
        else if (nsys >= kPlainLaneFrom)
            launch_plain_lane<M>(layout, A, B, C, C, D, nsys, nrow, flag, s);
        else if (layout == LAYOUT_TILED)
            launch_plain_twisted_auto<M, LAYOUT_TILED>(A, B, C, D, nsys, nrow, flag, s);
        else
            launch_plain_twisted_auto<M, LAYOUT_REF_GPU>(A, B, C, D, nsys, nrow, flag, s);
    } else {
        launch_plain_lane<M>(layout, A, B, C, C, D, nsys, nrow, flag, s);
    }
    return hipGetLastError();
}

template <int M>
hipError_t launch_plain_save_m(int layout, const double* A, const double* B, const double* C, double* Cw,
                               const double* Din, double* D, double* F, uint64_t* PV, int nsys, int nrow,
                               int* flag, hipStream_t s) {
    const dim3 grid((nsys + 63) / 64), block(64);
    if (layout == LAYOUT_TILED)
        hipLaunchKernelGGL((k_plain_save<M, LAYOUT_TILED>), grid, block, 0, s, A, B, C, Cw, Din, D, F, PV,
                           nsys, nrow, flag);
    else
        hipLaunchKernelGGL((k_plain_save<M, LAYOUT_REF_GPU>), grid, block, 0, s, A, B, C, Cw, Din, D, F, PV,
                           nsys, nrow, flag);
    return hipGetLastError();
}

template <int M>
hipError_t launch_adjoint_m(int layout, const double* A, const double* Cp, const double* F,
                            const uint64_t* PV, const double* X, const double* W, double* Lam,
                            double* dA, double* dB, double* dC, int nsys, int nrow, hipStream_t s) {
    const dim3 grid((nsys + 63) / 64), block(64);
    if (layout == LAYOUT_TILED)
        hipLaunchKernelGGL((k_adjoint<M, LAYOUT_TILED>), grid, block, 0, s, A, Cp, F, PV, X, W, Lam, dA,
                           dB, dC, nsys, nrow);
    else
        hipLaunchKernelGGL((k_adjoint<M, LAYOUT_REF_GPU>), grid, block, 0, s, A, Cp, F, PV, X, W, Lam,
                           dA, dB, dC, nsys, nrow);
    return hipGetLastError();
}

template <int M>
hipError_t launch_plain_f32s_m(int layout, const float* A, const float* B, const float* C, double* Cw,
                               double* D, int nsys, int nrow, int* flag, hipStream_t s) {
    const dim3 grid((nsys + 63) / 64), block(64);
    if (layout == LAYOUT_TILED)
        hipLaunchKernelGGL((k_plain_f32s<M, LAYOUT_TILED>), grid, block, 0, s, A, B, C, Cw, D, nsys, nrow, flag);
    else
        hipLaunchKernelGGL((k_plain_f32s<M, LAYOUT_REF_GPU>), grid, block, 0, s, A, B, C, Cw, D, nsys, nrow,
                           flag);
    return hipGetLastError();
}

// BTD_FORM_LAPACK_ORDER: the one-lane solve (the oracle's operation order) at any batch size and
// m; BTD_FORM_WORKSPACE: the same with C' written to the workspace Cw (Cw = C: in place).
template <int M>
hipError_t launch_plain_ref_m(int layout, const double* A, const double* B, const double* C, double* Cw,
                              double* D, int nsys, int nrow, int* flag, hipStream_t s) {
    launch_plain_lane<M>(layout, A, B, C, Cw, D, nsys, nrow, flag, s);
    return hipGetLastError();
}

template <int M>
hipError_t launch_f32r_m(int layout, const float* A, const float* B, const float* C,
                         const double* D, double* X, float* Wc, float* Wv, int nsys, int nrow,
                         int sweeps, int* flag, hipStream_t s, bool transposed) {
    const dim3 grid((nsys + 63) / 64), block(64);
    if (layout == LAYOUT_TILED) {
        if (transposed)
            hipLaunchKernelGGL((k_f32r<M, LAYOUT_TILED, true>), grid, block, 0, s, A, B, C, D, X, Wc, Wv,
                               nsys, nrow, sweeps, flag);
        else
            hipLaunchKernelGGL((k_f32r<M, LAYOUT_TILED, false>), grid, block, 0, s, A, B, C, D, X, Wc, Wv,
                               nsys, nrow, sweeps, flag);
    } else {
        if (transposed)
            hipLaunchKernelGGL((k_f32r<M, LAYOUT_REF_GPU, true>), grid, block, 0, s, A, B, C, D, X, Wc,
                               Wv, nsys, nrow, sweeps, flag);
        else
            hipLaunchKernelGGL((k_f32r<M, LAYOUT_REF_GPU, false>), grid, block, 0, s, A, B, C, D, X, Wc,
                               Wv, nsys, nrow, sweeps, flag);
    }
    return hipGetLastError();
}

template <int M>
hipError_t launch_outer_f32_m(int layout, const double* Lam, const double* X, float* dA, float* dB,
                              float* dC, int nsys, int nrow, hipStream_t s) {
    const dim3 grid((nsys + 63) / 64), block(64);
    if (layout == LAYOUT_TILED)
        hipLaunchKernelGGL((k_outer_f32<M, LAYOUT_TILED>), grid, block, 0, s, Lam, X, dA, dB, dC, nsys, nrow);
    else
        hipLaunchKernelGGL((k_outer_f32<M, LAYOUT_REF_GPU>), grid, block, 0, s, Lam, X, dA, dB, dC, nsys,
                           nrow);
    return hipGetLastError();
}

template <int M>
hipError_t launch_cyclic_m(int layout, double* A, const double* B, double* C, double* D, int nsys,
                           int nrow, int* flag, hipStream_t s) {
    const dim3 grid((nsys + 63) / 64), block(64);
    if (layout == LAYOUT_TILED)
        hipLaunchKernelGGL((k_cyclic<M, LAYOUT_TILED>), grid, block, 0, s, A, B, C, D, nsys, nrow, flag);
    else
        hipLaunchKernelGGL((k_cyclic<M, LAYOUT_REF_GPU>), grid, block, 0, s, A, B, C, D, nsys, nrow, flag);
    return hipGetLastError();
}

template <int M, class BS = MemBlocks>
hipError_t launch_two_stage_m(int layout, double* A, const double* B, double* C, double* D,
                              int nsys, int nrow, const int* pf, const int* pl, int npart,
                              double* redA, double* redC, double* redD, int* flag,
                              hipStream_t s, hipEvent_t* ev, bool cyclic, BS bs = BS{}) {
    const int ntiles = (nsys + 63) / 64;
    const dim3 block(64), gridp((unsigned)(ntiles * npart)), gridr((unsigned)ntiles);
    if (ev) (void)hipEventRecord(ev[0], s);
    const size_t lds = sizeof(double) * (2 * M * M + M) * 64;
    if constexpr (M < kModiCoopFrom) {
        if (lds > 65536) {  // above the default dynamic-LDS limit (m >= 8)
            (void)hipFuncSetAttribute((const void*)k_modi<M, LAYOUT_TILED, BS>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            (void)hipFuncSetAttribute((const void*)k_modi<M, LAYOUT_REF_GPU, BS>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        }
    }
    // Stage 1: m lanes per partition (k_modi_coop) for m >= kModiCoopFrom, and for smaller m
    // (non-periodic, blocks in memory) when the partitions are too few to fill the device one
    // lane each (fewer than kModiLaneFrom of them); otherwise one lane per partition (k_modi).
    // Periodic solves keep k_modi below kModiCoopFrom, so that a matrix-free sweep (k_modi with
    // its own block source) and the assembled periodic solve run the same arithmetic.
    // Below kModiCoopFrom, by the partitions in the batch (nsys * npart; m = 5, 512 x 65536,
    // stage 1 at P = 8 / 16 / 32 / 64): from kModiLaneFrom one lane per partition (k_modi); below
    // kModiTwistedBelow (tiled layout; any count in the reference layout) down to kModiTwistedFrom,
    // non-periodic with partitions of at least 4 rows, two lanes per partition from both ends
    // (k_modi_twisted: 21.1 / 10.8 / 7.0-7.4 / 6.2-7.4 ms, then k_update_twisted); otherwise from
    // kModiCoopBelow one lane per partition with the next row loaded during the current one (47.6 /
    // 24.4 / 12.7 / 7.5 ms) and below that m lanes per partition (k_modi_coop: 25.0 / 15.6 / 16.9 /
    // 13.0 ms), non-periodic solves only. Loading ahead changes no
    // arithmetic, so periodic solves with few partitions use it too and still equal the
    // matrix-free sweeps bit for bit. These few-partition forms are enabled for m <= kFewFormsMaxM
    // only (see kFewFormsMaxM).
    constexpr bool kCanCoop = M >= 2 && std::is_same<BS, MemBlocks>::value;
    bool coop = M >= kModiCoopFrom, ahead = false, twisted12 = false;
    if constexpr (kCanCoop && M <= kFewFormsMaxM) {
        const int64_t lanes = (int64_t)nsys * npart;
        twisted12 = !cyclic && lanes >= kModiTwistedFrom && (layout != LAYOUT_TILED || lanes < kModiTwistedBelow) &&
                    nrow >= 4 * npart;
        coop = !twisted12 && !cyclic && lanes < kModiCoopBelow;
        ahead = !twisted12 && !coop && lanes < kModiLaneFrom;
    }
    if (twisted12) {
        if constexpr (kCanCoop && M <= kFewFormsMaxM) {
            const dim3 block2(2 * kTile);
            const size_t lds2 = 2 * lds;  // each wavefront's slots
            if (layout == LAYOUT_TILED)
                hipLaunchKernelGGL((k_modi_twisted<M, LAYOUT_TILED>), gridp, block2, lds2, s, A, B, C, D, nsys,
                                   nrow, pf, pl, npart, redA, redC, redD, flag);
            else
                hipLaunchKernelGGL((k_modi_twisted<M, LAYOUT_REF_GPU, kTwistedRefAhead<M>>), gridp, block2, lds2,
                                   s, A, B, C, D, nsys, nrow, pf, pl, npart, redA, redC, redD, flag);
        }
    } else if (coop) {
        if constexpr (kCanCoop) {
            // the items of a partition on one XCD. floor(64/m) systems a wavefront leave partial
            // lines, which neighbouring wavefronts fetch again (m = 5, 512 x 65536, P = 32: 42 GB
            // read vs 21.5 GB one lane per partition); four wavefronts locked per workgroup read
            // them once but were slower at m = 5 (P = 8 / 16 / 32 / 64: 25.0 / 15.1 / 16.0 / 12.7 ->
            // 38.2 / 19.3 / 16.6 / 12.8 ms): with so few partitions the chain of rows, not HBM,
            // bounds stage 1, and locked workgroups occupy fewer CUs
            constexpr bool kLock = false;
            constexpr int WV = modi_coop_waves<M>();
            const int64_t chunks = (nsys + kTile / M - 1) / (kTile / M);
            const int64_t items = (int64_t)npart * ((chunks + WV - 1) / WV);
            const dim3 gridc((unsigned)(kXcds * ((items + kXcds - 1) / kXcds))), blockc(64 * WV);
            if (layout == LAYOUT_TILED)
                hipLaunchKernelGGL((k_modi_coop<M, LAYOUT_TILED, WV, kLock>), gridc, blockc, 0, s, A, B, C, D,
                                   nsys, nrow, pf, pl, npart, redA, redC, redD, flag);
            else
                hipLaunchKernelGGL((k_modi_coop<M, LAYOUT_REF_GPU, WV, kLock>), gridc, blockc, 0, s, A, B, C, D,
                                   nsys, nrow, pf, pl, npart, redA, redC, redD, flag);
        }
    } else {
        if constexpr (M < kModiCoopFrom) {
            bool launched = false;
            if constexpr (std::is_same<BS, MemBlocks>::value) {
                if (ahead) {
                    if (layout == LAYOUT_TILED)
                        hipLaunchKernelGGL((k_modi<M, LAYOUT_TILED, BS, true>), gridp, block, lds, s, A, B, C, D,
                                           nsys, nrow, pf, pl, npart, redA, redC, redD, flag, bs);
                    else
                        hipLaunchKernelGGL((k_modi<M, LAYOUT_REF_GPU, BS, true>), gridp, block, lds, s, A, B, C, D,
                                           nsys, nrow, pf, pl, npart, redA, redC, redD, flag, bs);
                    launched = true;
                }
            }
            if (!launched) {
                if (layout == LAYOUT_TILED)
                    hipLaunchKernelGGL((k_modi<M, LAYOUT_TILED, BS>), gridp, block, lds, s, A, B, C, D, nsys, nrow,
                                       pf, pl, npart, redA, redC, redD, flag, bs);
                else
                    hipLaunchKernelGGL((k_modi<M, LAYOUT_REF_GPU, BS>), gridp, block, lds, s, A, B, C, D, nsys,
                                       nrow, pf, pl, npart, redA, redC, redD, flag, bs);
            }
        }
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    // the twisted reduced solve when its wavefronts all fit on the device at once (one per
    // SIMD); otherwise one lane per system does more systems per instruction
    bool twisted = false;
    if constexpr (M >= 2) {
        const int64_t waves = ((int64_t)nsys + batch_width(M) - 1) / batch_width(M);
        if (!cyclic && waves <= kTwistedReducedWavesPerCU * (int64_t)device_cus()) {
            hipLaunchKernelGGL((k_reduced_twisted<M>), dim3((unsigned)waves), block, 0, s, redA, redC, redD,
                               nsys, 2 * npart, flag);
            twisted = true;
        }
    }
    if (cyclic)
        hipLaunchKernelGGL((k_reduced_cyclic<M>), gridr, block, 0, s, redA, redC, redD, nsys, 2 * npart, flag);
    else if (!twisted)
        hipLaunchKernelGGL((k_reduced<M>), gridr, block, 0, s, redA, redC, redD, nsys, 2 * npart, flag);
    if (ev) (void)hipEventRecord(ev[2], s);
    // stage 3: twisted after the twisted stage 1; otherwise m lanes per partition below
    // kModiLaneFrom partitions (M >= 2)
    bool coop3 = false;
    if constexpr (M >= 2 && M <= kFewFormsMaxM) coop3 = !cyclic && (int64_t)nsys * npart < kModiLaneFrom;
    if (twisted12) {
        if constexpr (kCanCoop && M <= kFewFormsMaxM) {
            const dim3 block2(2 * kTile);
            if (layout == LAYOUT_TILED)
                hipLaunchKernelGGL((k_update_twisted<M, LAYOUT_TILED>), gridp, block2, 0, s, A, C, D, nsys, nrow,
                                   pf, pl, npart, redD);
            else
                hipLaunchKernelGGL((k_update_twisted<M, LAYOUT_REF_GPU>), gridp, block2, 0, s, A, C, D, nsys, nrow,
                                   pf, pl, npart, redD);
        }
    } else if (coop3) {
        if constexpr (M >= 2) {
            // four wavefronts locked per workgroup (their shared lines read once: m = 5, 512 x 65536,
            // P = 16 / 32 / 64: 6.5 / 5.6 / 5.4 -> 4.6 / 4.6 / 3.6 ms) once there are two
            // wavefronts of work per CU; below that one wavefront per workgroup keeps more CUs busy
            // (P = 8: 7.4 ms unlocked, 9.1 locked)
            const int64_t chunks = (nsys + kTile / M - 1) / (kTile / M);
            const bool lock = (int64_t)npart * chunks >= 2 * (int64_t)device_cus();
            const int WV = lock ? 4 : 1;
            const int64_t items = (int64_t)npart * ((chunks + WV - 1) / WV);
            const dim3 gridu((unsigned)(kXcds * ((items + kXcds - 1) / kXcds))), blocku(64 * WV);
            if (lock) {
                if (layout == LAYOUT_TILED)
                    hipLaunchKernelGGL((k_update_coop<M, LAYOUT_TILED, 4, true>), gridu, blocku, 0, s, A, C, D,
                                       nsys, nrow, pf, pl, npart, redD);
                else
                    hipLaunchKernelGGL((k_update_coop<M, LAYOUT_REF_GPU, 4, true>), gridu, blocku, 0, s, A, C, D,
                                       nsys, nrow, pf, pl, npart, redD);
            } else {
                if (layout == LAYOUT_TILED)
                    hipLaunchKernelGGL((k_update_coop<M, LAYOUT_TILED, 1, false>), gridu, blocku, 0, s, A, C, D,
                                       nsys, nrow, pf, pl, npart, redD);
                else
                    hipLaunchKernelGGL((k_update_coop<M, LAYOUT_REF_GPU, 1, false>), gridu, blocku, 0, s, A, C, D,
                                       nsys, nrow, pf, pl, npart, redD);
            }
        }
    } else if (layout == LAYOUT_TILED) {
        hipLaunchKernelGGL((k_update<M, LAYOUT_TILED>), gridp, block, 0, s, A, C, D, nsys, nrow, pf, pl,
                           npart, redD);
    } else {
        hipLaunchKernelGGL((k_update<M, LAYOUT_REF_GPU>), gridp, block, 0, s, A, C, D, nsys, nrow, pf,
                           pl, npart, redD);
    }
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

template <int M>
hipError_t launch_generate_m(int kind, uint64_t seed, int layout, int nsys, int nrow,
                                    double* A, double* B, double* C, double* D, float* Af,
                                    float* Bf, float* Cf, hipStream_t s) {
    const int64_t total = pad_nsys(nsys) * (int64_t)nrow;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    hipLaunchKernelGGL((k_generate<M>), grid, block, 0, s, kind, seed, layout, nsys, nrow, A, B, C, D,
                       Af, Bf, Cf);
    return hipGetLastError();
}

}  // namespace btd
