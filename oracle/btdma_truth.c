/*
 * Quad-precision truth for the batched block-Thomas solve. TEST INFRASTRUCTURE ONLY: the
 * product library never links or calls this file; tests and bench.py's accuracy report
 * load it (inside liboracle.so) as the checker.
 *
 * The FP64 oracle (btdma_oracle.c) restates the reference's CPU path, whose dgesv rounds
 * like any FP64 solve: on blocks of condition number up to 1e6 it differs from the exact
 * solution by about cond * 1e-16 per system, as much as an FP64 device solve does. To say
 * which of two FP64 results is nearer the exact solution, this file solves the same
 * FP64 inputs with the same algorithm (block Thomas, btdma_many of
 * Fortran_Original/src/mod_btdma_cpu.f90:15-78, each block solved by LU with partial
 * pivoting as dgesv does) carried out in IEEE binary128 (__float128, 113-bit significand,
 * unit roundoff 9.6e-35). Its own error is about cond * 1e-34, so rounded to FP64 the
 * result is the exact solution to within an ulp or two for every input used here.
 * The algorithm order does not matter at this precision; only the inputs do, and they
 * are the caller's FP64 values, read exactly.
 */
#include "btdma_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef __float128 qreal;

static qreal qabs(qreal v) { return v < 0 ? -v : v; }

/* a (m x m, column major) <- its LU with partial pivoting (first max |a|); b (m x k,
 * column major) <- a^{-1} b. Returns 0, or j+1 for a zero pivot in column j. */
static int q_gesv(int m, int k, qreal *a, int *piv, qreal *b) {
    for (int j = 0; j < m; ++j) {
        int p = j;
        qreal best = qabs(a[j + m * j]);
        for (int i = j + 1; i < m; ++i) {
            qreal v = qabs(a[i + m * j]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        piv[j] = p;
        if (a[p + m * j] == 0) return j + 1;
        if (p != j) {
            for (int c = 0; c < m; ++c) {
                qreal t = a[j + m * c];
                a[j + m * c] = a[p + m * c];
                a[p + m * c] = t;
            }
            for (int c = 0; c < k; ++c) {
                qreal t = b[j + m * c];
                b[j + m * c] = b[p + m * c];
                b[p + m * c] = t;
            }
        }
        for (int i = j + 1; i < m; ++i) {
            qreal l = a[i + m * j] / a[j + m * j];
            a[i + m * j] = l;
            for (int c = j + 1; c < m; ++c) a[i + m * c] -= l * a[j + m * c];
            for (int c = 0; c < k; ++c) b[i + m * c] -= l * b[j + m * c];
        }
    }
    for (int c = 0; c < k; ++c) {
        qreal *x = b + (size_t)m * c;
        for (int i = m - 1; i >= 0; --i) {
            qreal s = x[i];
            for (int j = i + 1; j < m; ++j) s -= a[i + m * j] * x[j];
            x[i] = s / a[i + m * i];
        }
    }
    return 0;
}

#define QMAT(arr, sys, row) (&(arr)[(size_t)m * m * ((size_t)(sys) + (size_t)nsys * (row))])
#define QVEC(arr, sys, row) (&(arr)[(size_t)m * ((size_t)(sys) + (size_t)nsys * (row))])

/* One system: C'_q = R_q^{-1} C_q, d'_q = R_q^{-1} (d_q - A_q d'_{q-1}) with
 * R_q = B_q - A_q C'_{q-1}, then x_q = d'_q - C'_q x_{q+1}, all in binary128. */
static int truth_one(int n, int nsys, int m, const double *a, const double *b, const double *c,
                     const double *d, double *x, int sys, qreal *cp, qreal *dp, qreal *r, int *piv) {
    const size_t mm2 = (size_t)m * m;
    const int k = m + 1; /* right-hand sides [d | C] */
    qreal *rhs = (qreal *)malloc(sizeof(qreal) * (size_t)m * k);
    if (!rhs) return -1;
    int info = 0;
    for (int q = 0; q < n; ++q) {
        const double *aq = QMAT(a, sys, q), *bq = QMAT(b, sys, q), *cq = QMAT(c, sys, q);
        const double *dq = QVEC(d, sys, q);
        for (size_t e = 0; e < mm2; ++e) r[e] = (qreal)bq[e];
        for (int i = 0; i < m; ++i) rhs[i] = (qreal)dq[i];
        for (size_t e = 0; e < mm2; ++e) rhs[m + e] = q + 1 < n ? (qreal)cq[e] : 0;
        if (q > 0) {
            const qreal *cpm = cp + mm2 * (q - 1), *dpm = dp + (size_t)m * (q - 1);
            for (int i = 0; i < m; ++i) {
                for (int l = 0; l < m; ++l) {
                    const qreal ail = (qreal)aq[i + m * l];
                    rhs[i] -= ail * dpm[l];
                    for (int j = 0; j < m; ++j) r[i + m * j] -= ail * cpm[l + m * j];
                }
            }
        }
        int st = q_gesv(m, k, r, piv, rhs);
        if (st && !info) info = st;
        memcpy(dp + (size_t)m * q, rhs, sizeof(qreal) * m);
        memcpy(cp + mm2 * q, rhs + m, sizeof(qreal) * mm2);
    }
    for (int q = n - 2; q >= 0; --q) {
        qreal *xq = dp + (size_t)m * q;
        const qreal *xn = dp + (size_t)m * (q + 1), *cq = cp + mm2 * q;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) xq[i] -= cq[i + m * j] * xn[j];
    }
    for (int q = 0; q < n; ++q)
        for (int i = 0; i < m; ++i) QVEC(x, sys, q)[i] = (double)dp[(size_t)m * q + i];
    free(rhs);
    return info;
}

int ora_btdma_many_q(int n, int nsys, int m, const double *a, const double *b, const double *c,
                     const double *d, double *x) {
    int info = 0;
#ifdef _OPENMP
#pragma omp parallel reduction(max : info)
#endif
    {
        const size_t mm2 = (size_t)m * m;
        qreal *cp = (qreal *)malloc(sizeof(qreal) * mm2 * (size_t)n);
        qreal *dp = (qreal *)malloc(sizeof(qreal) * (size_t)m * n);
        qreal *r = (qreal *)malloc(sizeof(qreal) * mm2);
        int *piv = (int *)malloc(sizeof(int) * (size_t)m);
        if (!cp || !dp || !r || !piv) {
            info = -1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
            for (int sys = 0; sys < nsys; ++sys) {
                int st = truth_one(n, nsys, m, a, b, c, d, x, sys, cp, dp, r, piv);
                if (st < 0)
                    info = 1 << 30; /* out of memory */
                else if (st > info)
                    info = st;
            }
        }
        free(cp);
        free(dp);
        free(r);
        free(piv);
    }
    return info;
}
