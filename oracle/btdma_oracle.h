/*
 * CPU oracle: an FP64 restatement, in plain C, of the reference's CPU block-Thomas
 * solver (Fortran_Original/src/mod_btdma_cpu.f90) and of the LAPACK dgesv contract it
 * relies on.
 *
 * TEST INFRASTRUCTURE ONLY. Only the test suite, the smoke check and the CPU-baseline
 * leg of bench.py load this library, and only as the checker. The product library
 * (pascal_btdma_amd) never links or calls it.
 *
 * Layout: every array uses the reference CPU layout, column-major inside a block:
 *   a(m,m,nsys,n):  a[p + m*q + m*m*sys + m*m*nsys*row]
 *   d(m,nsys,n):    d[p + m*sys + m*nsys*row]
 * (mod_btdma_cpu.f90:18). All solves are in place like the reference: A, C and D are
 * overwritten, the solution is returned in D, B is read only except where the
 * reference itself overwrites it (it never does).
 *
 * Pinning status of each function (tests/test_oracle.py):
 *   ora_dgesv            pinned: scipy.linalg.lu_factor pivots/factors, scipy solve
 *   ora_btdma_many       pinned: Study signature CSV (closed form), dense solve,
 *                        whole-block-row permutation invariance
 *   ora_btdma_many_mpi   pinned: dense solve, P invariance, Study signatures at np=2,4,8
 *   ora_btdma_many_cycl  pinned: dense periodic solve (n >= 3)
 *   ora_btdma_many_cyclic_mpi  pinned: dense periodic solve, P invariance
 *   ora_btdma_many_t     pinned: dense transposed solve
 *   ora_btdma_many_q     pinned: exact rational solves (Python fractions) of small systems,
 *                        incl. blocks of condition number up to 1e6; equality with
 *                        ora_btdma_many on well-conditioned systems
 */
#ifndef BTDMA_ORACLE_H
#define BTDMA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* LAPACK dgesv semantics: LU with partial pivoting (first max |a| in the column,
 * like idamax), row interchanges recorded 1-based in ipiv, then forward/back
 * substitution for nrhs right-hand sides. Returns info (0, or k>0 if U(k,k) == 0). */
int ora_dgesv(int n, int nrhs, double *a, int lda, int *ipiv, double *b, int ldb);

/* mpiutil_para (mpiutil.f90:27-48): near-equal split of [sta,end] into nprocs parts;
 * the first mod(n,nprocs) parts get one extra element. Returns the part size. */
int ora_para(int sta, int end, int rank, int nprocs, int *indx_a, int *indx_b);

/* btdma_many (mod_btdma_cpu.f90:15-78): plain block Thomas over n rows. */
int ora_btdma_many(int n, int nsys, int m, double *a, double *b, double *c, double *d);

/* btdma_many_modi (mod_btdma_cpu.f90:305-443): local modified sweep over one partition
 * of nrow rows; fills the two reduced rows per system (rd*(m,m,nsys,2), rdD(m,nsys,2)). */
int ora_btdma_many_modi(int m, int nsys, int nrow, double *A, double *B, double *C, double *D,
                        double *rdA, double *rdB, double *rdC, double *rdD);

/* btdma_many_update (mod_btdma_cpu.f90:511-535). */
void ora_btdma_many_update(int m, int nsys, int nrow, const double *A, const double *C, double *D,
                           const double *rdD);

/* btdma_many_mpi (mod_btdma_cpu.f90:242-271) with the nprocs row partitions held in one
 * process: partition r owns rows ora_para(0,n-1,r,npart); the all-to-all becomes a
 * gather of the reduced rows in the order [r0 first, r0 last, r1 first, ...]
 * (mpiutil.f90:50-91). Requires n/npart >= 3. Returns max dgesv info. */
int ora_btdma_many_mpi(int m, int nsys, int n, int npart, double *A, double *B, double *C,
                       double *D);

/* btdma_many_cycl (mod_btdma_cpu.f90:80-183): periodic block Thomas. A(row 1) couples to
 * x_n, C(row n) couples to x_1. Requires n >= 3. */
int ora_btdma_many_cycl(int n, int nsys, int m, double *a, double *b, double *c, double *d);

/* btdma_many_cyclic_mpi (mod_btdma_cpu.f90:273-302) with npart in-process row partitions:
 * stage 1, a periodic reduced system of 2*npart rows, stage 3. npart >= 2, n/npart >= 3. */
int ora_btdma_many_cyclic_mpi(int m, int nsys, int n, int npart, double *A, double *B, double *C,
                              double *D);

/* Transposed solve T^T lambda = w for the (non-periodic) system given by a, b, c,
 * assembled as the block-tridiagonal system with blocks a'_q = c_{q-1}^T,
 * b'_q = b_q^T, c'_q = a_{q+1}^T and solved by ora_btdma_many. a, b, c are not
 * modified; w is overwritten by lambda. */
int ora_btdma_many_t(int n, int nsys, int m, const double *a, const double *b, const double *c,
                     double *w);

/* Truth solve (btdma_truth.c): the block-Thomas solve of the same FP64 inputs carried out in
 * IEEE binary128 and rounded to FP64 at the end, i.e. the exact solution to about an ulp for
 * block condition numbers far beyond 1e6. a, b, c, d are read only (CPU layout; C of the last
 * row and A of the first are ignored); x (m,nsys,n) receives the solution. Returns the first
 * zero-pivot column + 1 (0 if none). */
int ora_btdma_many_q(int n, int nsys, int m, const double *a, const double *b, const double *c,
                     const double *d, double *x);

/* Threads used by the last call (OpenMP over systems). */
int ora_num_threads(void);
void ora_set_num_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
