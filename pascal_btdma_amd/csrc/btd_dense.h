// Per-lane m x m block algebra held in registers (m is a compile-time constant, so every
// loop unrolls and every array index is static).
//
// LU with partial pivoting follows LAPACK dgetf2 (the routine behind the reference CPU
// oracle's dgesv, Fortran_Original/src/mod_btdma_cpu.f90:38): the pivot of column k is
// the FIRST row of maximal |a(i,k)|, i >= k; the multipliers are a(i,k) * (1/pivot).
// Rows k and p are exchanged only in columns k..m-1, and every right-hand side applies
// exchange k and elimination k in the same order, which gives the same solution as
// LAPACK's whole-row exchange. The reciprocal of each pivot is kept on the diagonal, so
// back substitution multiplies where dgetrs divides (a rounding difference only).
// Per-lane pivoting is data dependent, so exchanges are branch-free selects.
#pragma once

#include <hip/hip_runtime.h>

namespace btd {

template <int M>
__device__ __forceinline__ bool lu_factor(double (&a)[M][M], int (&piv)[M]) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        double best = fabs(a[k][k]);
        int p = k;
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const double v = fabs(a[i][k]);
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
                const double t = a[k][j];
                a[k][j] = c ? a[i][j] : t;
                a[i][j] = c ? t : a[i][j];
            }
        }
        ok = ok && (a[k][k] != 0.0);
        const double r = 1.0 / a[k][k];
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const double l = a[i][k] * r;
            a[i][k] = l;
#pragma unroll
            for (int j = k + 1; j < M; ++j) a[i][j] = fma(-l, a[k][j], a[i][j]);
        }
        a[k][k] = r;
    }
    return ok;
}

// x <- A^{-1} x with the factors from lu_factor.
template <int M>
__device__ __forceinline__ void lu_solve(const double (&a)[M][M], const int (&piv)[M], double (&x)[M]) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const bool c = (piv[k] == i);
            const double t = x[k];
            x[k] = c ? x[i] : t;
            x[i] = c ? t : x[i];
        }
#pragma unroll
        for (int i = k + 1; i < M; ++i) x[i] = fma(-a[i][k], x[k], x[i]);
    }
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
        double s = x[k];
#pragma unroll
        for (int j = k + 1; j < M; ++j) s = fma(-a[k][j], x[j], s);
        x[k] = s * a[k][k];
    }
}

// x <- A^{-T} x with the factors from lu_factor (A = P^T L U, so A^T y = x is solved as
// U^T z = x, L^T w = z, y = P^T w with the exchanges applied in reverse order).
template <int M>
__device__ __forceinline__ void lu_solve_t(const double (&a)[M][M], const int (&piv)[M], double (&x)[M]) {
#pragma unroll
    for (int k = 0; k < M; ++k) {  // U^T z = x (U^T lower, diagonal holds 1/u_kk)
        double s = x[k];
#pragma unroll
        for (int j = 0; j < k; ++j) s = fma(-a[j][k], x[j], s);
        x[k] = s * a[k][k];
    }
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {  // L^T, then undo exchange k
        double s = x[k];
#pragma unroll
        for (int i = k + 1; i < M; ++i) s = fma(-a[i][k], x[i], s);
        x[k] = s;
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const bool c = (piv[k] == i);
            const double t = x[k];
            x[k] = c ? x[i] : t;
            x[i] = c ? t : x[i];
        }
    }
}

// ---- the oracle's operation order --------------------------------------------------------
// lu_factor_ref / lu_solve_ref compute what lu_factor / lu_solve compute, in the order and with
// the roundings of LAPACK's unblocked dgetf2 + dgetrs as the reference CPU path calls them
// through dgesv (Fortran_Original/src/mod_btdma_cpu.f90:38,58): multipliers a(i,k) * (1/pivot)
// (a(i,k) / pivot when |pivot| < DBL_MIN, as dgetf2 does below sfmin), the trailing update
// a(i,j) - l * a(k,j) unfused, the pivot itself kept on the diagonal, unit-lower substitution
// x(i) - x(k) * l(i,k) for k ascending, and the upper solve column by column from the last:
// x(k) / u(k,k) (IEEE division), then x(i) - x(k) * u(i,k) for every i < k. No fused
// multiply-add anywhere, so each result is the oracle's bit for bit.
template <int M>
__device__ __forceinline__ bool lu_factor_ref(double (&a)[M][M], int (&piv)[M]) {
#pragma clang fp contract(off)
    bool ok = true;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        double best = fabs(a[k][k]);
        int p = k;
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const double v = fabs(a[i][k]);
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
                const double t = a[k][j];
                a[k][j] = c ? a[i][j] : t;
                a[i][j] = c ? t : a[i][j];
            }
        }
        const double pv = a[k][k];
        ok = ok && (pv != 0.0);
        if (fabs(pv) >= 2.2250738585072014e-308) {  // DBL_MIN, dgetf2's sfmin
            const double r = 1.0 / pv;
#pragma unroll
            for (int i = k + 1; i < M; ++i) a[i][k] = a[i][k] * r;
        } else {
#pragma unroll
            for (int i = k + 1; i < M; ++i) a[i][k] = a[i][k] / pv;
        }
#pragma unroll
        for (int i = k + 1; i < M; ++i)
#pragma unroll
            for (int j = k + 1; j < M; ++j) a[i][j] = a[i][j] - a[i][k] * a[k][j];
    }
    return ok;
}

template <int M>
__device__ __forceinline__ void lu_solve_ref(const double (&a)[M][M], const int (&piv)[M], double (&x)[M]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const bool c = (piv[k] == i);
            const double t = x[k];
            x[k] = c ? x[i] : t;
            x[i] = c ? t : x[i];
        }
#pragma unroll
        for (int i = k + 1; i < M; ++i) x[i] = x[i] - x[k] * a[i][k];
    }
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
        x[k] = x[k] / a[k][k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] = x[i] - x[k] * a[i][k];
    }
}

}  // namespace btd
