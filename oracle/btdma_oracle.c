/*
 * CPU oracle for the batched block-Thomas solve. TEST INFRASTRUCTURE ONLY: the product
 * library never links or calls this file; tests, the smoke check and bench.py's
 * CPU-baseline leg load it as the checker.
 *
 * Each function restates the reference routine cited above it, with the same loop
 * structure and the same in-place conventions. Dense m x m work that the reference
 * hands to BLAS/LAPACK (dgemm, dgemv, dgesv) is written out here from the published
 * LAPACK algorithm (dgetf2 right-looking LU with idamax partial pivoting, then dgetrs:
 * dlaswp, unit-lower forward substitution, upper back substitution with division by
 * the diagonal). Modern reference LAPACK reaches dgetrf2 (recursive) for small m; it
 * selects the same pivots and differs from dgetf2 only in rounding order.
 *
 * Systems are split across OpenMP threads in contiguous ranges; every thread runs the
 * reference's row-outer, system-inner loops over its own range, so the arithmetic per
 * system is unchanged by the thread count.
 */
#include "btdma_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MIDX(p, q, m) ((p) + (m) * (q))

static int g_threads = 0; /* 0 = OpenMP default */
static int g_last_threads = 1;

void ora_set_num_threads(int nthreads) { g_threads = nthreads; }
int ora_num_threads(void) { return g_last_threads; }

static int team_size(void) {
#ifdef _OPENMP
    int t = g_threads > 0 ? g_threads : omp_get_max_threads();
    return t < 1 ? 1 : t;
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- dense kernels -- */

/* idamax: first index of max |x| (LAPACK/BLAS convention). */
static int idamax_col(int n, const double *x) {
    int best = 0;
    double bv = fabs(x[0]);
    for (int i = 1; i < n; ++i) {
        double v = fabs(x[i]);
        if (v > bv) {
            bv = v;
            best = i;
        }
    }
    return best;
}

/* dgetf2: unblocked right-looking LU with partial pivoting, column major, lda. */
static int lu_dgetf2(int n, double *a, int lda, int *ipiv) {
    int info = 0;
    const double sfmin = DBL_MIN;
    for (int j = 0; j < n; ++j) {
        int jp = j + idamax_col(n - j, &a[j + (size_t)lda * j]);
        ipiv[j] = jp + 1;
        if (a[jp + (size_t)lda * j] != 0.0) {
            if (jp != j) {
                for (int k = 0; k < n; ++k) {
                    double t = a[j + (size_t)lda * k];
                    a[j + (size_t)lda * k] = a[jp + (size_t)lda * k];
                    a[jp + (size_t)lda * k] = t;
                }
            }
            double piv = a[j + (size_t)lda * j];
            if (fabs(piv) >= sfmin) {
                double r = 1.0 / piv;
                for (int i = j + 1; i < n; ++i) a[i + (size_t)lda * j] *= r;
            } else {
                for (int i = j + 1; i < n; ++i) a[i + (size_t)lda * j] /= piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        /* rank-1 update of the trailing matrix (dger) */
        for (int k = j + 1; k < n; ++k) {
            double t = a[j + (size_t)lda * k];
            if (t != 0.0) {
                for (int i = j + 1; i < n; ++i) a[i + (size_t)lda * k] -= a[i + (size_t)lda * j] * t;
            }
        }
    }
    return info;
}

/* dgetrs 'N' after dgetf2. */
static void lu_dgetrs(int n, int nrhs, const double *a, int lda, const int *ipiv, double *b, int ldb) {
    for (int i = 0; i < n; ++i) { /* dlaswp, forward */
        int ip = ipiv[i] - 1;
        if (ip != i) {
            for (int k = 0; k < nrhs; ++k) {
                double t = b[i + (size_t)ldb * k];
                b[i + (size_t)ldb * k] = b[ip + (size_t)ldb * k];
                b[ip + (size_t)ldb * k] = t;
            }
        }
    }
    for (int k = 0; k < nrhs; ++k) {
        double *x = &b[(size_t)ldb * k];
        for (int j = 0; j < n; ++j) { /* unit lower */
            double t = x[j];
            if (t != 0.0)
                for (int i = j + 1; i < n; ++i) x[i] -= t * a[i + (size_t)lda * j];
        }
        for (int j = n - 1; j >= 0; --j) { /* upper */
            if (x[j] != 0.0) {
                x[j] /= a[j + (size_t)lda * j];
                double t = x[j];
                for (int i = 0; i < j; ++i) x[i] -= t * a[i + (size_t)lda * j];
            }
        }
    }
}

int ora_dgesv(int n, int nrhs, double *a, int lda, int *ipiv, double *b, int ldb) {
    int info = lu_dgetf2(n, a, lda, ipiv);
    if (info == 0) lu_dgetrs(n, nrhs, a, lda, ipiv, b, ldb);
    return info;
}

/* out = X * Y (m x m, column major), like dgemm('N','N',...,alpha=1,beta=0). */
static void mm(int m, const double *X, const double *Y, double *out) {
    for (int j = 0; j < m; ++j) {
        for (int i = 0; i < m; ++i) out[MIDX(i, j, m)] = 0.0;
        for (int l = 0; l < m; ++l) {
            double t = Y[MIDX(l, j, m)];
            for (int i = 0; i < m; ++i) out[MIDX(i, j, m)] += t * X[MIDX(i, l, m)];
        }
    }
}

/* out = X * v, like dgemv('N',...,alpha=1,beta=0). */
static void mv(int m, const double *X, const double *v, double *out) {
    for (int i = 0; i < m; ++i) out[i] = 0.0;
    for (int l = 0; l < m; ++l) {
        double t = v[l];
        for (int i = 0; i < m; ++i) out[i] += t * X[MIDX(i, l, m)];
    }
}

int ora_para(int sta, int end, int rank, int nprocs, int *indx_a, int *indx_b) {
    int n = end - sta + 1;
    int tmp1 = n / nprocs, tmp2 = n % nprocs;
    int aa, bb;
    if (rank < tmp2) {
        aa = rank;
        bb = 1;
    } else {
        aa = tmp2;
        bb = 0;
    }
    *indx_a = rank * tmp1 + aa + sta;
    *indx_b = *indx_a + tmp1 + bb - 1;
    return tmp1 + bb;
}

/* ------------------------------------------------------------ per-range bodies -- */

#define MAT(arr, sys, row) (&(arr)[(size_t)m * m * ((size_t)(sys) + (size_t)nsys * (row))])
#define VEC(arr, sys, row) (&(arr)[(size_t)m * ((size_t)(sys) + (size_t)nsys * (row))])

typedef struct {
    double *sol, *rr, *t1, *t2, *t3, *v1, *v2, *eye;
    int *ipiv;
} work_t;

static int work_alloc(work_t *w, int m) {
    size_t mm2 = (size_t)m * m;
    w->sol = (double *)malloc(sizeof(double) * (size_t)m * (2 * m + 1));
    w->rr = (double *)malloc(sizeof(double) * mm2);
    w->t1 = (double *)malloc(sizeof(double) * mm2);
    w->t2 = (double *)malloc(sizeof(double) * mm2);
    w->t3 = (double *)malloc(sizeof(double) * mm2);
    w->v1 = (double *)malloc(sizeof(double) * m);
    w->v2 = (double *)malloc(sizeof(double) * m);
    w->eye = (double *)calloc(mm2, sizeof(double));
    w->ipiv = (int *)malloc(sizeof(int) * m);
    if (!w->sol || !w->rr || !w->t1 || !w->t2 || !w->t3 || !w->v1 || !w->v2 || !w->eye || !w->ipiv)
        return -1;
    for (int i = 0; i < m; ++i) w->eye[MIDX(i, i, m)] = 1.0;
    return 0;
}

static void work_free(work_t *w) {
    free(w->sol);
    free(w->rr);
    free(w->t1);
    free(w->t2);
    free(w->t3);
    free(w->v1);
    free(w->v2);
    free(w->eye);
    free(w->ipiv);
}

static inline int imax(int a, int b) { return a > b ? a : b; }

/* mod_btdma_cpu.f90:15-78 over systems [s0, s1). */
static int many_range(int n, int nsys, int m, double *a, double *b, double *c, double *d, int s0,
                      int s1, work_t *w) {
    size_t mm2 = (size_t)m * m;
    int info = 0;
    double *Sol = w->sol, *RR = w->rr, *AkCkm = w->t1, *AkDkm = w->v1, *CkDkp = w->v2;
    for (int sys = s0; sys < s1; ++sys) { /* q = 1 */
        memcpy(Sol, VEC(d, sys, 0), sizeof(double) * m);
        memcpy(Sol + m, MAT(c, sys, 0), sizeof(double) * mm2);
        memcpy(RR, MAT(b, sys, 0), sizeof(double) * mm2);
        info = imax(info, ora_dgesv(m, m + 1, RR, m, w->ipiv, Sol, m));
        memcpy(VEC(d, sys, 0), Sol, sizeof(double) * m);
        memcpy(MAT(c, sys, 0), Sol + m, sizeof(double) * mm2);
    }
    for (int q = 1; q < n; ++q) {
        for (int sys = s0; sys < s1; ++sys) {
            mm(m, MAT(a, sys, q), MAT(c, sys, q - 1), AkCkm);
            mv(m, MAT(a, sys, q), VEC(d, sys, q - 1), AkDkm);
            const double *dq = VEC(d, sys, q), *bq = MAT(b, sys, q);
            for (int i = 0; i < m; ++i) Sol[i] = dq[i] - AkDkm[i];
            memcpy(Sol + m, MAT(c, sys, q), sizeof(double) * mm2);
            for (size_t k = 0; k < mm2; ++k) RR[k] = bq[k] - AkCkm[k];
            info = imax(info, ora_dgesv(m, m + 1, RR, m, w->ipiv, Sol, m));
            memcpy(VEC(d, sys, q), Sol, sizeof(double) * m);
            memcpy(MAT(c, sys, q), Sol + m, sizeof(double) * mm2);
        }
    }
    for (int q = n - 2; q >= 0; --q) {
        for (int sys = s0; sys < s1; ++sys) {
            mv(m, MAT(c, sys, q), VEC(d, sys, q + 1), CkDkp);
            double *dq = VEC(d, sys, q);
            for (int i = 0; i < m; ++i) dq[i] -= CkDkp[i];
        }
    }
    return info;
}

/* mod_btdma_cpu.f90:305-443 over systems [s0, s1). */
static int modi_range(int m, int nsys, int nrow, double *A, double *B, double *C, double *D,
                      double *rdA, double *rdB, double *rdC, double *rdD, int s0, int s1,
                      work_t *w) {
    size_t mm2 = (size_t)m * m;
    int info = 0;
    double *Sol = w->sol, *RR = w->rr;
    double *AqAqm = w->t1, *AqCqm = w->t2, *Tmp = w->t3; /* Tmp: CqAqp, then CqCqp */
    double *AqDqm = w->v1, *CqDqp = w->v2;
    for (int q = 0; q < 2; ++q) { /* q = 1, 2 */
        for (int sys = s0; sys < s1; ++sys) {
            memcpy(Sol, VEC(D, sys, q), sizeof(double) * m);
            memcpy(Sol + m, MAT(A, sys, q), sizeof(double) * mm2);
            memcpy(Sol + m + mm2, MAT(C, sys, q), sizeof(double) * mm2);
            memcpy(RR, MAT(B, sys, q), sizeof(double) * mm2);
            info = imax(info, ora_dgesv(m, 2 * m + 1, RR, m, w->ipiv, Sol, m));
            memcpy(VEC(D, sys, q), Sol, sizeof(double) * m);
            memcpy(MAT(A, sys, q), Sol + m, sizeof(double) * mm2);
            memcpy(MAT(C, sys, q), Sol + m + mm2, sizeof(double) * mm2);
        }
    }
    for (int q = 2; q < nrow; ++q) { /* q = 3..nrow_sub */
        for (int sys = s0; sys < s1; ++sys) {
            mv(m, MAT(A, sys, q), VEC(D, sys, q - 1), AqDqm);
            mm(m, MAT(A, sys, q), MAT(A, sys, q - 1), AqAqm);
            mm(m, MAT(A, sys, q), MAT(C, sys, q - 1), AqCqm);
            const double *dq = VEC(D, sys, q), *bq = MAT(B, sys, q);
            for (int i = 0; i < m; ++i) Sol[i] = dq[i] - AqDqm[i];
            for (size_t k = 0; k < mm2; ++k) Sol[m + k] = -AqAqm[k];
            memcpy(Sol + m + mm2, MAT(C, sys, q), sizeof(double) * mm2);
            for (size_t k = 0; k < mm2; ++k) RR[k] = bq[k] - AqCqm[k];
            info = imax(info, ora_dgesv(m, 2 * m + 1, RR, m, w->ipiv, Sol, m));
            memcpy(VEC(D, sys, q), Sol, sizeof(double) * m);
            memcpy(MAT(A, sys, q), Sol + m, sizeof(double) * mm2);
            memcpy(MAT(C, sys, q), Sol + m + mm2, sizeof(double) * mm2);
        }
    }
    for (int q = nrow - 3; q >= 1; --q) { /* q = nrow_sub-2 .. 2 */
        for (int sys = s0; sys < s1; ++sys) {
            mv(m, MAT(C, sys, q), VEC(D, sys, q + 1), CqDqp);
            mm(m, MAT(C, sys, q), MAT(A, sys, q + 1), Tmp);
            double *dq = VEC(D, sys, q), *aq = MAT(A, sys, q), *cq = MAT(C, sys, q);
            for (int i = 0; i < m; ++i) dq[i] -= CqDqp[i];
            for (size_t k = 0; k < mm2; ++k) aq[k] -= Tmp[k];
            mm(m, cq, MAT(C, sys, q + 1), Tmp);
            for (size_t k = 0; k < mm2; ++k) cq[k] = -Tmp[k];
        }
    }
    for (int sys = s0; sys < s1; ++sys) { /* q = 1 close */
        mv(m, MAT(C, sys, 0), VEC(D, sys, 1), CqDqp);
        mm(m, MAT(C, sys, 0), MAT(C, sys, 1), AqCqm);  /* CqCqp */
        mm(m, MAT(C, sys, 0), MAT(A, sys, 1), AqAqm);  /* CqAqp */
        const double *d0 = VEC(D, sys, 0);
        for (int i = 0; i < m; ++i) Sol[i] = d0[i] - CqDqp[i];
        memcpy(Sol + m, MAT(A, sys, 0), sizeof(double) * mm2);
        for (size_t k = 0; k < mm2; ++k) Sol[m + mm2 + k] = -AqCqm[k];
        for (size_t k = 0; k < mm2; ++k) RR[k] = w->eye[k] - AqAqm[k];
        info = imax(info, ora_dgesv(m, 2 * m + 1, RR, m, w->ipiv, Sol, m));
        memcpy(VEC(D, sys, 0), Sol, sizeof(double) * m);
        memcpy(MAT(A, sys, 0), Sol + m, sizeof(double) * mm2);
        memcpy(MAT(C, sys, 0), Sol + m + mm2, sizeof(double) * mm2);
    }
    /* reduced rows (:432-441) */
    for (int sys = s0; sys < s1; ++sys) {
        for (int r = 0; r < 2; ++r) {
            int q = r == 0 ? 0 : nrow - 1;
            memcpy(MAT(rdA, sys, r), MAT(A, sys, q), sizeof(double) * mm2);
            memcpy(MAT(rdC, sys, r), MAT(C, sys, q), sizeof(double) * mm2);
            memcpy(VEC(rdD, sys, r), VEC(D, sys, q), sizeof(double) * m);
            memcpy(MAT(rdB, sys, r), w->eye, sizeof(double) * mm2);
        }
    }
    return info;
}

/* mod_btdma_cpu.f90:511-535 over systems [s0, s1). */
static void update_range(int m, int nsys, int nrow, const double *A, const double *C, double *D,
                         const double *rdD, int s0, int s1, work_t *w) {
    double *AqD0 = w->v1, *CqDn = w->v2;
    for (int sys = s0; sys < s1; ++sys) {
        memcpy(VEC(D, sys, 0), VEC(rdD, sys, 0), sizeof(double) * m);
        memcpy(VEC(D, sys, nrow - 1), VEC(rdD, sys, 1), sizeof(double) * m);
    }
    for (int q = 1; q < nrow - 1; ++q) {
        for (int sys = s0; sys < s1; ++sys) {
            mv(m, MAT(A, sys, q), VEC(D, sys, 0), AqD0);
            mv(m, MAT(C, sys, q), VEC(D, sys, nrow - 1), CqDn);
            double *dq = VEC(D, sys, q);
            for (int i = 0; i < m; ++i) dq[i] = dq[i] - AqD0[i] - CqDn[i];
        }
    }
}

/* mod_btdma_cpu.f90:80-183 over systems [s0, s1). e is scratch (m,m,nsys,n). */
static int cycl_range(int n, int nsys, int m, double *a, double *b, double *c, double *d, double *e,
                      int s0, int s1, work_t *w) {
    size_t mm2 = (size_t)m * m;
    int info = 0;
    double *Sol = w->sol, *RR = w->rr, *AkCkm = w->t1, *AkEkm = w->t2, *CkEkp = w->t3;
    double *AkDkm = w->v1, *CkDkp = w->v2;
    for (int q = 0; q < n; ++q)
        for (int sys = s0; sys < s1; ++sys) memset(MAT(e, sys, q), 0, sizeof(double) * mm2);
    for (int sys = s0; sys < s1; ++sys) {
        const double *a2 = MAT(a, sys, 1), *cn = MAT(c, sys, n - 1);
        double *e2 = MAT(e, sys, 1), *en = MAT(e, sys, n - 1);
        for (size_t k = 0; k < mm2; ++k) e2[k] = -a2[k];
        for (size_t k = 0; k < mm2; ++k) en[k] = -cn[k];
    }
    for (int sys = s0; sys < s1; ++sys) { /* q = 2 */
        memcpy(Sol, VEC(d, sys, 1), sizeof(double) * m);
        memcpy(Sol + m, MAT(e, sys, 1), sizeof(double) * mm2);
        memcpy(Sol + m + mm2, MAT(c, sys, 1), sizeof(double) * mm2);
        memcpy(RR, MAT(b, sys, 1), sizeof(double) * mm2);
        info = imax(info, ora_dgesv(m, 2 * m + 1, RR, m, w->ipiv, Sol, m));
        memcpy(VEC(d, sys, 1), Sol, sizeof(double) * m);
        memcpy(MAT(e, sys, 1), Sol + m, sizeof(double) * mm2);
        memcpy(MAT(c, sys, 1), Sol + m + mm2, sizeof(double) * mm2);
    }
    for (int q = 2; q < n; ++q) { /* q = 3..n */
        for (int sys = s0; sys < s1; ++sys) {
            mm(m, MAT(a, sys, q), MAT(c, sys, q - 1), AkCkm);
            mv(m, MAT(a, sys, q), VEC(d, sys, q - 1), AkDkm);
            mm(m, MAT(a, sys, q), MAT(e, sys, q - 1), AkEkm);
            const double *dq = VEC(d, sys, q), *eq = MAT(e, sys, q), *bq = MAT(b, sys, q);
            for (int i = 0; i < m; ++i) Sol[i] = dq[i] - AkDkm[i];
            for (size_t k = 0; k < mm2; ++k) Sol[m + k] = eq[k] - AkEkm[k];
            memcpy(Sol + m + mm2, MAT(c, sys, q), sizeof(double) * mm2);
            for (size_t k = 0; k < mm2; ++k) RR[k] = bq[k] - AkCkm[k];
            info = imax(info, ora_dgesv(m, 2 * m + 1, RR, m, w->ipiv, Sol, m));
            memcpy(VEC(d, sys, q), Sol, sizeof(double) * m);
            memcpy(MAT(e, sys, q), Sol + m, sizeof(double) * mm2);
            memcpy(MAT(c, sys, q), Sol + m + mm2, sizeof(double) * mm2);
        }
    }
    for (int q = n - 2; q >= 1; --q) { /* q = n-1..2 */
        for (int sys = s0; sys < s1; ++sys) {
            mv(m, MAT(c, sys, q), VEC(d, sys, q + 1), CkDkp);
            mm(m, MAT(c, sys, q), MAT(e, sys, q + 1), CkEkp);
            double *dq = VEC(d, sys, q), *eq = MAT(e, sys, q);
            for (int i = 0; i < m; ++i) dq[i] -= CkDkp[i];
            for (size_t k = 0; k < mm2; ++k) eq[k] -= CkEkp[k];
        }
    }
    for (int sys = s0; sys < s1; ++sys) {
        mm(m, MAT(a, sys, 0), MAT(e, sys, n - 1), AkEkm);
        mm(m, MAT(c, sys, 0), MAT(e, sys, 1), CkEkp);
        mv(m, MAT(a, sys, 0), VEC(d, sys, n - 1), AkDkm);
        mv(m, MAT(c, sys, 0), VEC(d, sys, 1), CkDkp);
        const double *b0 = MAT(b, sys, 0);
        for (size_t k = 0; k < mm2; ++k) RR[k] = b0[k] + AkEkm[k] + CkEkp[k];
        double *d0 = VEC(d, sys, 0);
        for (int i = 0; i < m; ++i) d0[i] = d0[i] - AkDkm[i] - CkDkp[i];
        info = imax(info, ora_dgesv(m, 1, RR, m, w->ipiv, d0, m));
    }
    for (int q = 1; q < n; ++q) {
        for (int sys = s0; sys < s1; ++sys) {
            mv(m, MAT(e, sys, q), VEC(d, sys, 0), AkDkm); /* EkDk */
            double *dq = VEC(d, sys, q);
            for (int i = 0; i < m; ++i) dq[i] += AkDkm[i];
        }
    }
    return info;
}

/* ------------------------------------------------------------- public wrappers -- */

#define PARALLEL_OVER_SYSTEMS(BODY)                                                     \
    int info_all = 0;                                                                   \
    int nt = team_size();                                                               \
    if (nt > nsys) nt = nsys;                                                           \
    g_last_threads = nt;                                                                \
    int alloc_fail = 0;                                                                 \
    _Pragma("omp parallel num_threads(nt) reduction(max : info_all) reduction(| : alloc_fail)") \
    {                                                                                   \
        int tid = 0, nth = 1;                                                           \
        ORA_TID(tid, nth);                                                              \
        int s0 = (int)(((long long)nsys * tid) / nth);                                  \
        int s1 = (int)(((long long)nsys * (tid + 1)) / nth);                            \
        work_t w;                                                                       \
        if (work_alloc(&w, m) != 0) {                                                   \
            alloc_fail = 1;                                                             \
        } else {                                                                        \
            BODY;                                                                       \
        }                                                                               \
        work_free(&w);                                                                  \
    }                                                                                   \
    if (alloc_fail) return -2;

#ifdef _OPENMP
#define ORA_TID(tid, nth)                   \
    do {                                    \
        tid = omp_get_thread_num();         \
        nth = omp_get_num_threads();        \
    } while (0)
#else
#define ORA_TID(tid, nth) \
    do {                  \
    } while (0)
#endif

int ora_btdma_many(int n, int nsys, int m, double *a, double *b, double *c, double *d) {
    if (n < 1 || nsys < 1 || m < 1) return -1;
    PARALLEL_OVER_SYSTEMS(info_all = imax(info_all, many_range(n, nsys, m, a, b, c, d, s0, s1, &w)));
    return info_all;
}

int ora_btdma_many_modi(int m, int nsys, int nrow, double *A, double *B, double *C, double *D,
                        double *rdA, double *rdB, double *rdC, double *rdD) {
    if (nrow < 3 || nsys < 1 || m < 1) return -1;
    PARALLEL_OVER_SYSTEMS(info_all = imax(info_all, modi_range(m, nsys, nrow, A, B, C, D, rdA, rdB,
                                                               rdC, rdD, s0, s1, &w)));
    return info_all;
}

void ora_btdma_many_update(int m, int nsys, int nrow, const double *A, const double *C, double *D,
                           const double *rdD) {
    if (nrow < 2 || nsys < 1 || m < 1) return;
    int nt = team_size();
    if (nt > nsys) nt = nsys;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0, nth = 1;
        ORA_TID(tid, nth);
        int s0 = (int)(((long long)nsys * tid) / nth);
        int s1 = (int)(((long long)nsys * (tid + 1)) / nth);
        work_t w;
        if (work_alloc(&w, m) == 0) update_range(m, nsys, nrow, A, C, D, rdD, s0, s1, &w);
        work_free(&w);
    }
}

int ora_btdma_many_mpi(int m, int nsys, int n, int npart, double *A, double *B, double *C,
                       double *D) {
    if (m < 1 || nsys < 1 || npart < 1 || n < 3 * npart) return -1;
    size_t mm2 = (size_t)m * m;
    size_t rmat = mm2 * (size_t)nsys * 2, rvec = (size_t)m * nsys * 2;
    size_t tmat = mm2 * (size_t)nsys * 2 * npart, tvec = (size_t)m * nsys * 2 * npart;
    double *rdA = malloc(sizeof(double) * rmat), *rdB = malloc(sizeof(double) * rmat);
    double *rdC = malloc(sizeof(double) * rmat), *rdD = malloc(sizeof(double) * rvec);
    double *trA = malloc(sizeof(double) * tmat), *trB = malloc(sizeof(double) * tmat);
    double *trC = malloc(sizeof(double) * tmat), *trD = malloc(sizeof(double) * tvec);
    int info = 0;
    if (!rdA || !rdB || !rdC || !rdD || !trA || !trB || !trC || !trD) {
        info = -2;
        goto done;
    }
    /* stage 1 on every partition, then place its two reduced rows at rows 2r, 2r+1 */
    for (int r = 0; r < npart; ++r) {
        int ia, ib;
        int nrow = ora_para(0, n - 1, r, npart, &ia, &ib);
        size_t om = mm2 * (size_t)nsys * ia, ov = (size_t)m * nsys * ia;
        int i1 = ora_btdma_many_modi(m, nsys, nrow, A + om, B + om, C + om, D + ov, rdA, rdB, rdC, rdD);
        info = imax(info, i1);
        memcpy(trA + mm2 * (size_t)nsys * 2 * r, rdA, sizeof(double) * rmat);
        memcpy(trB + mm2 * (size_t)nsys * 2 * r, rdB, sizeof(double) * rmat);
        memcpy(trC + mm2 * (size_t)nsys * 2 * r, rdC, sizeof(double) * rmat);
        memcpy(trD + (size_t)m * nsys * 2 * r, rdD, sizeof(double) * rvec);
    }
    info = imax(info, ora_btdma_many(2 * npart, nsys, m, trA, trB, trC, trD));
    for (int r = 0; r < npart; ++r) {
        int ia, ib;
        int nrow = ora_para(0, n - 1, r, npart, &ia, &ib);
        size_t om = mm2 * (size_t)nsys * ia, ov = (size_t)m * nsys * ia;
        ora_btdma_many_update(m, nsys, nrow, A + om, C + om, D + ov, trD + (size_t)m * nsys * 2 * r);
    }
done:
    free(rdA);
    free(rdB);
    free(rdC);
    free(rdD);
    free(trA);
    free(trB);
    free(trC);
    free(trD);
    return info;
}

/* btdma_many_cyclic_mpi (mod_btdma_cpu.f90:273-302) with the npart row partitions held in
 * one process: stage 1 and stage 3 as in ora_btdma_many_mpi, the reduced system of 2*npart
 * rows solved periodically (btdma_many_cycl). Requires n/npart >= 3 and 2*npart >= 3. */
int ora_btdma_many_cyclic_mpi(int m, int nsys, int n, int npart, double *A, double *B, double *C,
                              double *D) {
    if (m < 1 || nsys < 1 || npart < 2 || n < 3 * npart) return -1;
    size_t mm2 = (size_t)m * m;
    size_t rmat = mm2 * (size_t)nsys * 2, rvec = (size_t)m * nsys * 2;
    size_t tmat = mm2 * (size_t)nsys * 2 * npart, tvec = (size_t)m * nsys * 2 * npart;
    double *rdA = malloc(sizeof(double) * rmat), *rdB = malloc(sizeof(double) * rmat);
    double *rdC = malloc(sizeof(double) * rmat), *rdD = malloc(sizeof(double) * rvec);
    double *trA = malloc(sizeof(double) * tmat), *trB = malloc(sizeof(double) * tmat);
    double *trC = malloc(sizeof(double) * tmat), *trD = malloc(sizeof(double) * tvec);
    int info = 0;
    if (!rdA || !rdB || !rdC || !rdD || !trA || !trB || !trC || !trD) {
        info = -2;
        goto done;
    }
    for (int r = 0; r < npart; ++r) {
        int ia, ib;
        int nrow = ora_para(0, n - 1, r, npart, &ia, &ib);
        size_t om = mm2 * (size_t)nsys * ia, ov = (size_t)m * nsys * ia;
        info = imax(info, ora_btdma_many_modi(m, nsys, nrow, A + om, B + om, C + om, D + ov, rdA, rdB,
                                              rdC, rdD));
        memcpy(trA + mm2 * (size_t)nsys * 2 * r, rdA, sizeof(double) * rmat);
        memcpy(trB + mm2 * (size_t)nsys * 2 * r, rdB, sizeof(double) * rmat);
        memcpy(trC + mm2 * (size_t)nsys * 2 * r, rdC, sizeof(double) * rmat);
        memcpy(trD + (size_t)m * nsys * 2 * r, rdD, sizeof(double) * rvec);
    }
    info = imax(info, ora_btdma_many_cycl(2 * npart, nsys, m, trA, trB, trC, trD));
    for (int r = 0; r < npart; ++r) {
        int ia, ib;
        int nrow = ora_para(0, n - 1, r, npart, &ia, &ib);
        size_t om = mm2 * (size_t)nsys * ia, ov = (size_t)m * nsys * ia;
        ora_btdma_many_update(m, nsys, nrow, A + om, C + om, D + ov, trD + (size_t)m * nsys * 2 * r);
    }
done:
    free(rdA);
    free(rdB);
    free(rdC);
    free(rdD);
    free(trA);
    free(trB);
    free(trC);
    free(trD);
    return info;
}

int ora_btdma_many_cycl(int n, int nsys, int m, double *a, double *b, double *c, double *d) {
    if (n < 3 || nsys < 1 || m < 1) return -1;
    double *e = (double *)malloc(sizeof(double) * (size_t)m * m * nsys * n);
    if (!e) return -2;
    int info_all = 0;
    int nt = team_size();
    if (nt > nsys) nt = nsys;
    g_last_threads = nt;
#pragma omp parallel num_threads(nt) reduction(max : info_all)
    {
        int tid = 0, nth = 1;
        ORA_TID(tid, nth);
        int s0 = (int)(((long long)nsys * tid) / nth);
        int s1 = (int)(((long long)nsys * (tid + 1)) / nth);
        work_t w;
        if (work_alloc(&w, m) == 0)
            info_all = imax(info_all, cycl_range(n, nsys, m, a, b, c, d, e, s0, s1, &w));
        else
            info_all = -2;
        work_free(&w);
    }
    free(e);
    return info_all;
}

int ora_btdma_many_t(int n, int nsys, int m, const double *a, const double *b, const double *c,
                     double *w) {
    if (n < 1 || nsys < 1 || m < 1) return -1;
    size_t mm2 = (size_t)m * m, tot = mm2 * (size_t)nsys * n;
    double *at = calloc(tot, sizeof(double)), *bt = malloc(sizeof(double) * tot);
    double *ct = calloc(tot, sizeof(double));
    if (!at || !bt || !ct) {
        free(at);
        free(bt);
        free(ct);
        return -2;
    }
    for (int q = 0; q < n; ++q) {
        for (int sys = 0; sys < nsys; ++sys) {
            for (int i = 0; i < m; ++i) {
                for (int j = 0; j < m; ++j) {
                    MAT(bt, sys, q)[MIDX(i, j, m)] = MAT(b, sys, q)[MIDX(j, i, m)];
                    if (q > 0) MAT(at, sys, q)[MIDX(i, j, m)] = MAT(c, sys, q - 1)[MIDX(j, i, m)];
                    if (q < n - 1) MAT(ct, sys, q)[MIDX(i, j, m)] = MAT(a, sys, q + 1)[MIDX(j, i, m)];
                }
            }
        }
    }
    int info = ora_btdma_many(n, nsys, m, at, bt, ct, w);
    free(at);
    free(bt);
    free(ct);
    return info;
}
