/*
 * pascal_btdma_amd: C ABI of the MI355X (gfx950) batched block-tridiagonal solver.
 *
 * Drop-in for the many-system plan-create / solve entry points of the reference:
 *   CUDA C++ port   CUDA_CXX_Port/include/pascal_btdma_cuda.hpp:88-166
 *   CUDA Fortran    Fortran_Original/src/mod_btdma_gpu_v2.f90:70-239
 * One GPU, no MPI: the reference's communicator of nprocs row partitions becomes
 * `npart` row partitions of every system held in the same HBM, and the reduced system
 * of 2*npart rows is solved on the device (no all-to-all).
 *
 * Conventions (same as the reference):
 *  - Row q of system s: A_q x_{q-1} + B_q x_q + C_q x_{q+1} = d_q; A of the first row and
 *    C of the last row are ignored.
 *  - Device arrays, solved in place: D returns x; A, C and D are overwritten (A only
 *    when npart > 1); B is only read.
 *  - Blocks are column-major: element (p,q) of a block is e = p + m*q.
 *  - Layouts (element e of block row `row` of system `sys`):
 *      BTD_LAYOUT_REF_GPU  sys + nsys*(row + nrow*e)   (pascal_btdma_cuda.hpp:46-65)
 *      BTD_LAYOUT_REF_CPU  e + E*(sys + nsys*row)      (mod_btdma_cpu.f90:18; conversion only)
 *      BTD_LAYOUT_TILED    ((tile*nrow + row)*E + e)*64 + lane, sys = 64*tile + lane,
 *                          nsys padded up to a multiple of 64 (this library's native form)
 *    with E = m*m for blocks and E = m for vectors.
 *  - Every call is asynchronous on the given HIP stream (NULL = default stream) unless
 *    stated otherwise. A plan holds device scratch and must not be used by two calls at
 *    once.
 * Errors are returned as btd_status; btd_last_error() gives a thread-local message.
 */
#ifndef PASCAL_BTDMA_AMD_H
#define PASCAL_BTDMA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BTD_MAX_M 10

typedef struct btd_plan_s *btd_plan_t;

/* Status codes. The reference throws C++ exceptions; they map as noted. */
typedef enum btd_status {
    BTD_OK = 0,
    BTD_ERR_INVALID_ARG = 1,     /* std::invalid_argument from create (pascal_btdma_cuda.cu:962-967) */
    BTD_ERR_NOT_CREATED = 2,     /* std::runtime_error, uncreated plan (pascal_btdma_cuda.cu:865-867) */
    BTD_ERR_DIM_MISMATCH = 3,    /* std::invalid_argument, dims != plan (pascal_btdma_cuda.cu:868-870) */
    BTD_ERR_HIP = 4,             /* CudaError (pascal_btdma_cuda.hpp:20-22,43-44) */
    BTD_ERR_SINGULAR = 5,        /* new: a zero pivot in some block LU (the reference ignores info) */
    BTD_ERR_NOT_IMPLEMENTED = 6  /* std::logic_error from solve_cyclic (pascal_btdma_cuda.cu:1063-1075) */
} btd_status;

typedef enum btd_layout {
    BTD_LAYOUT_REF_GPU = 0,
    BTD_LAYOUT_REF_CPU = 1,
    BTD_LAYOUT_TILED = 2
} btd_layout;

/* How a plan's single-partition solve runs (btd_plan_set_form). */
typedef enum btd_form {
    BTD_FORM_PLAIN = 0,   /* the library's choice: one lane per system (m = 1, or m <= 7 in
                           * large batches; blocks in VGPRs; every block operation in the
                           * order and rounding of the reference CPU path's BLAS/LAPACK calls,
                           * so x is that path's result bit for bit), m lanes per system
                           * (m >= 8; one block row each, pivoting through LDS) or, for smaller
                           * batches, twisted (these two agree with it to rounding) */
    BTD_FORM_TWISTED = 1, /* two groups of m lanes per system, one eliminating from each end */
    BTD_FORM_LAPACK_ORDER = 2, /* the one-lane solve at every batch size and m: each block
                                * operation in the order and rounding of the reference CPU path's
                                * BLAS/LAPACK calls (dgemm/dgemv sums from the first term, dgetf2
                                * + dgetrs, IEEE division, no fused multiply-add) */
    BTD_FORM_WORKSPACE = 3     /* the solve of BTD_FORM_LAPACK_ORDER with C' written to a
                                * workspace the plan owns instead of over C: C is only read */
} btd_form;

/* Most workspaces btd_plan_place_workspace tries in one call. */
#define BTD_MAX_WORKSPACE_CANDIDATES 16

/* Phase times in seconds; the fields of BtdmaSolveTimings (pascal_btdma_cuda.hpp:24-39)
 * and BTDMA_TIMING_gpu_v2 (mod_btdma_gpu_v2.f90:38-45). On one GPU the two exchange
 * phases do not exist and read 0. Measured with HIP events on the solve stream. */
typedef struct btd_timings {
    double total;
    double local_compute;
    double forward_exchange;
    double reduced_compute;
    double backward_exchange;
    double update_compute;
} btd_timings;

/* Replaces BtdmaGpuPlan::create(m, nsys, nrow, MPI_Comm, mode) (pascal_btdma_cuda.hpp:128-132,
 * pascal_btdma_cuda.cu:956-1005) and btdma_makeplan_gpu_v2(plan, m, nsys, nrow_sub, comm)
 * (mod_btdma_gpu_v2.f90:70-109). `npart` takes the place of the communicator size:
 * npart = 1 is the single-partition solve; npart > 1 splits every system's rows with
 * btd_partition_1d and needs nrow >= 3*npart; npart = 0 lets the library choose
 * (btd_auto_npart; read the choice back with btd_plan_npart). layout is BTD_LAYOUT_REF_GPU
 * or BTD_LAYOUT_TILED. 1 <= m <= BTD_MAX_M, nsys >= 1, nrow >= 2. Allocates device scratch
 * on the current device. */
btd_status btd_plan_create(btd_plan_t *plan, int m, int nsys, int nrow, int npart, int layout);

/* New: the partition count a plan created with npart = 0 gets on one GPU, where partitions are
 * an occupancy choice rather than a rank count: 1 (the single-kernel solve) for nsys >= 32768
 * at m = 1, 12288 at m = 3, 5 (two lanes per system there), 8192 at m = 2, 4, 4096 at m = 6
 * and 3072 at m >= 7 (two, m or 2m lanes per system). Otherwise, for m <= 5 (both stages two lanes per partition): 1 below 128
 * rows when nsys < 8192, else about 16384 / nsys partitions (at least 2 from 8192 systems) of at
 * least 16 rows, rounded down to a power of two; for m = 6: 1 up to 128 rows, else about
 * 65536 / nsys of at least 8 rows, a power of two; for m >= 7: 1 below 128 rows and for
 * 1024 systems or more of at most 192 rows, else about 65536 / nsys (one wave per SIMD) of at
 * least 8 rows. 1 whenever that is < 2. Calibrated on
 * MI355X over m = 2-8, 256-16384 systems, 64-2048 rows. Pure function (no device use). */
int btd_auto_npart(int m, int nsys, int nrow);

/* The plan's partition count (the one chosen when it was created with npart = 0). */
btd_status btd_plan_npart(btd_plan_t plan, int *npart);

/* New (no counterpart in the reference): choose how an npart = 1 plan's non-cyclic solve runs.
 * BTD_FORM_PLAIN (the default) is the library's choice: one lane per system for m = 1, for
 * m = 2-4 from 131072 systems and m = 5 from 98304; two lanes per system, one eliminating
 * downwards from row 0 and one upwards from row nrow-1 with the blocks in registers, meeting in
 * the middle, from 24576 systems at m = 2, 12288 at m = 3-5, 6144 at m = 6, above 8192 at m = 7
 * (up to those one-lane sizes; at every size from m = 6); for m >= 8 m lanes per system (one
 * block row each); and the twisted form below for batches too small to fill the device those
 * ways: 2 <= m <= 7 below those sizes, m = 8 up to 16 systems per CU (reference layout) or 32
 * (tiled), i.e. 4096 / 8192 systems on 256 CUs, m = 9, 10 up to 12 systems per CU (3072). BTD_FORM_TWISTED (m >= 2, either layout) always gives each system two
 * groups of m lanes, one eliminating downwards from row 0 and one upwards from row nrow-1,
 * meeting in the middle: half the row latency per system (on large batches its wavefronts
 * share cache lines that leave L2 between their reads, so it moves more bytes). Same
 * zero-pivot reporting; results agree with the one-lane solve to rounding; C is left
 * overwritten. BTD_FORM_LAPACK_ORDER (npart = 1, any m) runs the one-lane solve at every batch
 * size: every block operation of mod_btdma_cpu.f90:15-78 in the order of the reference's dgemm,
 * dgemv and dgesv as reference (unblocked dgetf2/dgetrs) BLAS/LAPACK performs them, so x is the
 * FP64 CPU path's result bit for bit, not merely within rounding of it (slower than the
 * library's choice on batches that do not fill the device one lane per system).
 * BTD_FORM_WORKSPACE (npart = 1, any m) is that solve with C'_q written to a workspace of the
 * layout's size of an m*m block array, allocated by this call and owned by the plan, instead of
 * over C: C is only read, so a caller that solves the same blocks again need not restore C, and
 * the library, not the caller, decides where the written array lies (btd_plan_place_workspace).
 * Same x, bit for bit, as BTD_FORM_LAPACK_ORDER. Setting another form frees the workspace.
 * Cyclic and the other solve variants are unaffected. */
btd_status btd_plan_set_form(btd_plan_t plan, int form);
btd_status btd_plan_form(btd_plan_t plan, int *form);

/* New (no counterpart in the reference), for plans in BTD_FORM_WORKSPACE: picks where in HBM the
 * workspace lies. The time of the one-lane solve depends on the physical pages of the two arrays
 * it writes (the workspace and D), which no allocation call chooses: on MI355X, m = 5, n = 512,
 * 131072 systems, the same solve of the same arrays took 13.2 ms or 14.9 ms depending on the
 * allocation, each allocation keeping its time. This call allocates up to `candidates` - 1 more
 * workspaces (as many as fit), times two solves of the caller's arrays with each (D restored
 * from a copy before each solve and at the end, so D returns unchanged), keeps the fastest
 * workspace and frees the others. ms (NULL or `candidates` entries) receives each candidate's
 * time in milliseconds, -1 for those not allocated; candidate 0 is the plan's workspace before
 * the call. The solves raise the plan's zero-pivot words as a solve would. Synchronizes `stream`
 * and allocates: not capturable. The choice holds for these arrays; another D may prefer
 * another workspace. 1 <= candidates <= BTD_MAX_WORKSPACE_CANDIDATES. */
btd_status btd_plan_place_workspace(btd_plan_t plan, const double *a, const double *b, const double *c,
                                    double *d, int m, int nsys, int nrow, int candidates, double *ms,
                                    void *stream);

/* New (no counterpart in the reference): btd_plan_place_workspace for the FP32-storage solve
 * (btd_solve_noncyclic_f32s with work == NULL, which then writes C' to the plan's workspace):
 * the candidates are timed with that solve of the FP32 blocks a, b, c and the FP64 d, which is
 * restored and returned unchanged. Same rules and limits as btd_plan_place_workspace. */
btd_status btd_plan_place_workspace_f32s(btd_plan_t plan, const float *a, const float *b, const float *c,
                                         double *d, int candidates, double *ms, void *stream);

/* Replaces BtdmaGpuPlan::destroy (pascal_btdma_cuda.hpp:134, pascal_btdma_cuda.cu:1007-1033)
 * and btdma_cleanplan_gpu_v2 (mod_btdma_gpu_v2.f90:111-129). NULL is accepted. */
btd_status btd_plan_destroy(btd_plan_t plan);

/* Replaces solve_noncyclic(plan, a, b, c, d, m, nsys, nrow, stream)
 * (pascal_btdma_cuda.hpp:137-145, pascal_btdma_cuda.cu:1035-1045) and
 * btdma_many_mpi_gpu_v2(A,B,C,D,m,nsys,nrow_sub,plan) (mod_btdma_gpu_v2.f90:131-167).
 * In place, asynchronous. Zero pivots are recorded in the plan; see btd_plan_check. */
btd_status btd_solve_noncyclic(btd_plan_t plan, double *a, const double *b, double *c, double *d,
                               int m, int nsys, int nrow, void *stream);

/* Replaces solve_noncyclic_profiled (pascal_btdma_cuda.hpp:147-156, pascal_btdma_cuda.cu:1047-1061)
 * and btdma_many_mpi_gpu_v2_profiled (mod_btdma_gpu_v2.f90:169-198). Synchronises the
 * stream, fills *timings (may be NULL) and returns BTD_ERR_SINGULAR if this solve met a
 * zero pivot (earlier unchecked zero pivots of the plan are dropped first); the record then
 * stays for btd_plan_info. */
btd_status btd_solve_noncyclic_profiled(btd_plan_t plan, double *a, const double *b, double *c,
                                        double *d, int m, int nsys, int nrow, btd_timings *timings,
                                        void *stream);

/* Periodic solve: replaces btdma_many_cycl_mpi_gpu_v2 (mod_btdma_gpu_v2.f90:200-239) and
 * solve_cyclic (pascal_btdma_cuda.hpp:158-166), which throws logic_error in the reference
 * C++ port. Row 0 couples to x_{n-1} through A of row 0 and row n-1 to x_0 through C of
 * row n-1. npart > 1: stage 1 as in btd_solve_noncyclic, a periodic reduced system of
 * 2*npart rows, stage 3 (btdma_many_cyclic_mpi, mod_btdma_cpu.f90:273-302). npart = 1:
 * the same over partitions the library chooses (about 65536 / nsys of at least 8 rows, 1 from
 * 32768 systems (8192 for m = 2, 4, 4096 for m >= 7) and below 128 rows at m >= 5; where that is
 * < 2, 4, or 2 below 32 rows; faster than one periodic sweep per system at every size
 * measured), with their scratch allocated by the plan's first periodic solve: that first call
 * allocates and copies synchronously, and is refused (BTD_ERR_INVALID_ARG) when its stream is
 * being captured, so make one periodic call before capturing one in a graph; below 6 rows one
 * periodic block-Thomas sweep per system (needs nrow >= 3). A, C, D are overwritten. */
btd_status btd_solve_cyclic(btd_plan_t plan, double *a, const double *b, double *c, double *d,
                            int m, int nsys, int nrow, void *stream);

/* Forward solve that also keeps the block factorization for btd_solve_adjoint (no
 * counterpart in the reference; the differentiable-solve extension). Same in-place
 * contract as btd_solve_noncyclic (npart must be 1): c returns C'_q, d returns x.
 * factors: btd_layout_elems(layout, nsys, nrow, m*m) doubles, the LU factors of
 * R_q = B_q - A_q C'_{q-1} per row. pivots: btd_layout_elems(layout, nsys, nrow, 1)
 * 64-bit words, the pivot rows of R_q packed 4 bits per elimination step. */
btd_status btd_solve_noncyclic_save(btd_plan_t plan, const double *a, const double *b, double *c,
                                    double *d, double *factors, uint64_t *pivots, void *stream);

/* btd_solve_noncyclic_save with c and d only read (no counterpart in the reference): C'_q goes to
 * c_out and x to x (each the size of c and d), so a caller that keeps its inputs (the
 * differentiable solve) copies nothing first. c_out may be c and x may be d (then it is
 * btd_solve_noncyclic_save); other overlaps are not allowed. Same x and C'_q (q < nrow - 1) bit
 * for bit; out of place, c_out's last row (no C' there: the last row's C is ignored) and the
 * tiled layout's padding lanes of c_out and x are written with zeros. */
btd_status btd_solve_noncyclic_save_out(btd_plan_t plan, const double *a, const double *b, const double *c,
                                        const double *d, double *c_out, double *x, double *factors,
                                        uint64_t *pivots, void *stream);

/* Adjoint of the solve for L(x) with dL/dx = w (no counterpart in the reference):
 * lambda = T^{-T} w from the factors of btd_solve_noncyclic_save (a: the original A,
 * c_factored: C' as returned in c, x: the solution), then dd = lambda and the banded
 * outer products dA_q = -lambda_q x_{q-1}^T, dB_q = -lambda_q x_q^T,
 * dC_q = -lambda_q x_{q+1}^T (dA of the first row and dC of the last row are 0). */
btd_status btd_solve_adjoint(btd_plan_t plan, const double *a, const double *c_factored,
                             const double *factors, const uint64_t *pivots, const double *x,
                             const double *w, double *lambda, double *da, double *db, double *dc,
                             void *stream);

/* Mixed precision (no counterpart in the reference): blocks a, b, c stored in FP32,
 * d and x in FP64. An FP32 block-Thomas solve followed by `sweeps` (0..8) iterative-
 * refinement sweeps, each with the residual r = d - T x formed in FP64 from the
 * FP32-stored blocks and the correction x += T^{-1} r from the FP32 factorization.
 * a, b, c, d are only read. work: btd_layout_elems(layout, nsys, nrow, m*m) +
 * btd_layout_elems(layout, nsys, nrow, m) floats of device scratch. npart must be 1. */
btd_status btd_solve_noncyclic_f32r(btd_plan_t plan, const float *a, const float *b, const float *c,
                                    const double *d, double *x, float *work, int sweeps,
                                    void *stream);

/* FP32-stored blocks, FP64 arithmetic (no counterpart in the reference): the FP64 solve of the
 * blocks' values, one lane per system in the reference CPU path's BLAS/LAPACK operation order,
 * so bit for bit what that path (and btd_solve_noncyclic's one-lane solve) computes on FP64
 * copies of a, b, c, reading 4 bytes per block element. a, b, c are only read; d is in place (x returned in d);
 * work: btd_layout_elems(layout, nsys, nrow, m*m) doubles of device scratch (holds C'), or NULL
 * for the plan's workspace (a plan in BTD_FORM_WORKSPACE; its placement for this solve is picked
 * by btd_plan_place_workspace_f32s). npart must be 1. */
btd_status btd_solve_noncyclic_f32s(btd_plan_t plan, const float *a, const float *b, const float *c,
                                    double *d, double *work, void *stream);

/* Adjoint of btd_solve_noncyclic_f32r (no counterpart in the reference): lambda solves the
 * transposed system T^T lambda = w from the same FP32-stored blocks a, b, c (as given for
 * T), by an FP32 solve plus `sweeps` refinement sweeps with FP64 residuals. Same workspace
 * as btd_solve_noncyclic_f32r; npart must be 1. */
btd_status btd_solve_transposed_f32r(btd_plan_t plan, const float *a, const float *b, const float *c,
                                     const double *w, double *lambda, float *work, int sweeps,
                                     void *stream);

/* dA_q = -lambda_q x_{q-1}^T, dB_q = -lambda_q x_q^T, dC_q = -lambda_q x_{q+1}^T, computed in
 * FP64 and stored as FP32 block arrays (the gradients of FP32-stored blocks). */
btd_status btd_outer_grads_f32(btd_plan_t plan, const double *lambda, const double *x, float *da,
                               float *db, float *dc, void *stream);

/* ---- euler3d ADI sweeps: assembly straight into the solve's layout ----------------------
 * Fields are Fortran-ordered device arrays: U(0:n1+1, 0:n2+1, 0:n3+1, 5), the conserved
 * variables with their one-cell halo already filled; fields without halo are
 * F(n1, n2, n3, m). The sweep along `axis` (0 = x, 1 = y, 2 = z) is a batch of nsys systems
 * of nrow block rows numbered as the reference numbers them after its transposes:
 *   x: nrow = n1, nsys = n2*n3, sys = j + n2*k;   y: nrow = n2, nsys = n1*n3, sys = i + n1*k;
 *   z: nrow = n3, nsys = n1*n2, sys = i + n1*j     (0-based interior i, j, k).
 * The plan must have that nrow and nsys (and m = 5 for the assembly). */

/* Replaces build_matrix_{x,y,z}_kernel (euler3d_mpi_gpu.f90:803-929) followed by
 * transpose_{jk,ik}_kernel / transpose_{jk,ik}_5d_kernel (:931-957, :976-1002): writes the
 * sweep's blocks a, b, c (A = -cc J(U_prev) - cd I, B = (1 + 2cd) I, C = cc J(U_next) - cd I,
 * cc = dt/(4h), cd = eps*dt/(2h^2), J the flux Jacobian along the axis, compute_Amat/Bmat/Cmat
 * :749-801) and, if rhs is given, d = rhs reordered, in the plan's layout. rhs and d are
 * both NULL or both given. Asynchronous. */
btd_status btd_euler_assemble(btd_plan_t plan, int axis, const double *U, const double *rhs, int n1,
                              int n2, int n3, double h, double dt, double eps, double *a, double *b,
                              double *c, double *d, void *stream);

/* Matrix-free periodic sweep (no reference counterpart; it does what build_matrix_*,
 * the transposes, btdma_many_cycl_mpi_gpu_v2 and the transpose back do at
 * euler3d_mpi_gpu.f90:1260-1325 for one sweep, single partition): solves the sweep's periodic
 * system with the blocks of btd_euler_assemble computed inside the solve, and writes the
 * solution into the field out(n1, n2, n3, 5); rhs(n1, n2, n3, 5) is the right-hand side
 * field, and out may be rhs itself (the reference overwrites dU). npart must be 1 and
 * nrow >= 3. work: device scratch of 2*btd_layout_elems(layout, nsys, nrow, 25) +
 * btd_layout_elems(layout, nsys, nrow, 5) + btd_layout_elems(layout, nsys, nrow + 2, 5)
 * doubles. Zero pivots are recorded in the plan. */
btd_status btd_euler_sweep_cyclic(btd_plan_t plan, int axis, const double *U, const double *rhs,
                                  int n1, int n2, int n3, double h, double dt, double eps,
                                  double *work, double *out, void *stream);

/* ---- weno2d implicit sweeps: assembly straight into the solve's layout ------------------
 * Replaces compute_weno_coeffs_{x,y}_kernel + build_matrix_{x,y}_kernel_weno5
 * (Fortran_Original/examples/weno2d/weno_mpi_gpu.f90:430-485, :504-719): writes the blocks a,
 * b, c of the sweep's periodic system (I - gdt L) with m = 3 (three consecutive cells per block
 * row), L the frozen-WENO5 upwind advection operator of a scalar field; the five stencil
 * coefficients of each interface are computed in the kernel and never stored. U is the
 * Fortran-ordered field with its three-cell halo filled, U(1-3:n1+3, 1-3:n2+3); vel[s] the
 * advection speed of line s (x sweep, axis 0: c_x of row j = s+1, nsys = n2, nrow = n1/3;
 * y sweep, axis 1: c_y of column i = s+1, nsys = n1, nrow = n2/3); vel >= 0 upwinds from the
 * left. h is the cell width, gdt = gamma * dt. The plan must have m = 3 and that nsys and
 * nrow. Asynchronous. No counterpart to the reference's API (its weno2d example's kernels). */
btd_status btd_weno_assemble(btd_plan_t plan, int axis, const double *U, const double *vel, int n1, int n2,
                             double h, double gdt, double *a, double *b, double *c, void *stream);

/* Matrix-free periodic weno2d sweep (no reference counterpart; it does what
 * build_matrix_{x,y}_kernel_weno5 and btdma_many_cycl_mpi_gpu_v2 do at weno_mpi_gpu.f90:1020,
 * :1042/:1058, without storing A, B, C): solves (I - gdt L) x = d for the sweep along `axis`,
 * the two-stage periodic solve with stage 1 computing each block row from U and vel as
 * btd_weno_assemble would write it (so the result equals btd_weno_assemble followed by
 * btd_solve_cyclic bit for bit). d: the right-hand side in the plan's layout
 * (btd_weno_gather), replaced by x. work: device scratch of 2 * btd_layout_elems(layout,
 * nsys, nrow, 9) doubles. npart > 1 plans use their partitions, npart = 1 plans those of
 * btd_solve_cyclic (allocated by the plan's first periodic call, which is synchronous and
 * refused while its stream is being captured; needs nrow >= 6). Zero pivots are recorded in the
 * plan. Asynchronous once the partitions exist. */
btd_status btd_weno_sweep_cyclic(btd_plan_t plan, int axis, const double *U, const double *vel, int n1, int n2,
                                 double h, double gdt, double *work, double *d, void *stream);

/* Scalar field F(n1, n2) (Fortran order, no halo) -> the weno2d sweep's vector d (m = 3, cell
 * 3r + p + 1 of line s at element p of block row r of system s), in the plan's layout; and
 * back. Replace pack_scalar_to_xblock_kernel / unpack_xblock_to_scalar_kernel
 * (weno_mpi_gpu.f90:279-301) for the x sweep and, for the y sweep, the transposes of dU around
 * it (transpose_xy_d_kernel / transpose_yx_d_kernel, :332-356). Asynchronous. */
btd_status btd_weno_gather(btd_plan_t plan, int axis, const double *field, int n1, int n2, double *d,
                           void *stream);
btd_status btd_weno_scatter(btd_plan_t plan, int axis, const double *d, int n1, int n2, double *field,
                            void *stream);

/* Field F(n1, n2, n3, m) -> the sweep's vector d in the plan's layout (transpose_jk_kernel /
 * transpose_ik_kernel, :931-943, :976-988, for callers that assemble their own blocks). */
btd_status btd_sweep_gather(btd_plan_t plan, int axis, const double *field, int n1, int n2, int n3,
                            double *d, void *stream);

/* The sweep's vector d -> field F(n1, n2, n3, m) (transpose_kj_kernel :959-974,
 * transpose_ki_kernel :1004-1017). */
btd_status btd_sweep_scatter(btd_plan_t plan, int axis, const double *d, int n1, int n2, int n3,
                             double *field, void *stream);

/* New (the reference ignores LAPACK/elimination info): synchronises the plan's last
 * stream use, stores 1 in *singular if any block LU met a zero pivot since the last
 * check (0 otherwise), and clears the flag and the per-system record. */
btd_status btd_plan_check(btd_plan_t plan, int *singular);

/* New, the per-system form of btd_plan_check (the reference's dgesv `info`, which it
 * ignores, mod_btdma_cpu.f90:38): synchronises like btd_plan_check and fills the host array
 * info[nsys] with 1 for every system whose solve met a zero pivot since the last check
 * (system index in the caller's numbering), 0 for the others; then clears the record. */
btd_status btd_plan_info(btd_plan_t plan, int *info);

/* Thread-local message for the last error on this thread ("" if none). */
const char *btd_last_error(void);

/* matrix_index / vector_index (pascal_btdma_cuda.hpp:46-65): offsets in the reference GPU
 * layout. p is the block row, q the block column. */
size_t btd_matrix_index(int sys, int row, int p, int q, int nsys, int nrow, int m);
size_t btd_vector_index(int sys, int row, int p, int nsys, int nrow);

/* partition_1d (pascal_btdma_cuda.cu:921-930) / mpiutil_para (mpiutil.f90:27-48). */
void btd_partition_1d(int start, int end, int nprocs, int rank, int *first, int *last);

/* Number of doubles an array occupies in a layout (E = m*m for blocks, m for vectors). */
size_t btd_layout_elems(int layout, int nsys, int nrow, int elems_per_block);

/* Layout conversion on the device (src and dst must not overlap). */
btd_status btd_convert_layout(int src_layout, int dst_layout, int elems_per_block, int nsys,
                              int nrow, const double *src, double *dst, void *stream);

/* Seeded synthetic inputs written on the device in `layout`. kind: 0 = Study diagonal
 * workload (A=0.25I, B=2I, C=0.25I, d=1), 1 = diagonally dominant rows with each whole
 * block row permuted, 2 = Euler-like (J based, sigma 0.4), 3 = conditioned B blocks
 * (cond log-uniform in [1e1,1e6]) with A, B, C rounded to FP32, 4 = A, C uniform in [-1,1],
 * B uniform in [-1,1] plus (row sum |A| + row sum |C| + 1) on the diagonal with only B's rows
 * then permuted. Any pointer may be NULL. */
btd_status btd_generate(int kind, uint64_t seed, int m, int nsys, int nrow, int layout, double *a,
                        double *b, double *c, double *d, void *stream);

/* Same values as btd_generate, rounded to FP32 for the A, B, C arrays (d stays FP64). */
btd_status btd_generate_f32(int kind, uint64_t seed, int m, int nsys, int nrow, int layout,
                            float *a, float *b, float *c, double *d, void *stream);

/* Host copy of the generator for systems [sys0, sys0+count): natural row-major blocks,
 * a[((row*count + s)*m + i)*m + j] = A_row(i,j) of system sys0+s; d[(row*count+s)*m + i].
 * Bit-identical to btd_generate. Any pointer may be NULL. No device use. */
btd_status btd_generate_host(int kind, uint64_t seed, int m, int nsys, int nrow, int sys0,
                             int count, double *a, double *b, double *c, double *d);

/* ---- Mixed-size batches (no counterpart in the reference, whose plan has one m and one
 * nrow): systems of different block size m (1..10) and row count n (>= 2) solved by one
 * persistent kernel that takes work items (packs of same-m systems) from a queue in
 * longest-first order. Packed layout: systems of one m, longest first, are interleaved in
 * packs of W systems (W = 64 for m = 1, floor(32/m) otherwise): element e of row `row` of
 * system s is at mat_off(s) + (row*m*m + e)*W for blocks and vec_off(s) + (row*m + e)*W for
 * vectors (btd_batch_locate gives both offsets and W). In place like btd_solve_noncyclic. */
typedef struct btd_batch_s *btd_batch_t;

/* m[s], n[s] for s < nsys (host arrays, copied). */
btd_status btd_batch_create(btd_batch_t *batch, int nsys, const int *m, const int *n);
btd_status btd_batch_destroy(btd_batch_t batch);
/* Doubles needed by the packed block arrays (A, B, C) and vector arrays (D). */
btd_status btd_batch_sizes(btd_batch_t batch, size_t *mat_elems, size_t *vec_elems, int *npacks);
/* Where system sys starts in the packed arrays and the stride W between its elements. */
btd_status btd_batch_locate(btd_batch_t batch, int sys, size_t *mat_off, size_t *vec_off, int *stride);
/* Launch parameters (what the autotuner picks): placement 0 = one lane per system, blocks
 * in VGPRs; 1 = as 0 but m >= 6 with blocks in LDS; 2, 3, 4 = m lanes per system (one block
 * row per lane, pivot search and row exchange through LDS) for m >= 6, m >= 4, m >= 2;
 * 5 = as 4, with the items whose cost exceeds total / grid solved twisted (two groups of m
 * lanes per system eliminating from both ends, half the rows each; results agree with the
 * other placements to rounding, not bit for bit); 6 = as 4 with every cooperative item
 * twisted; grid = persistent workgroups of 64 lanes. Synchronous: waits for all work queued on
 * the device (hipDeviceSynchronize) before it rewrites the batch's work-item table, so a solve or
 * graph replay of the batch in flight is never handed a half-written table; not to be called
 * while a stream is being captured. */
btd_status btd_batch_set_params(btd_batch_t batch, int placement, int grid);
/* The seeded generators of btd_generate, per system (system index s, its own n). */
btd_status btd_batch_generate(btd_batch_t batch, int kind, uint64_t seed, double *a, double *b,
                              double *c, double *d, void *stream);
/* Asynchronous, capturable in a graph (the work counter is reset by a kernel). */
btd_status btd_batch_solve(btd_batch_t batch, double *a, const double *b, double *c, double *d,
                           void *stream);
btd_status btd_batch_check(btd_batch_t batch, int *singular);

/* ---- Graph capture of any sequence of the asynchronous calls above on one stream
 * (hipStreamBeginCapture / hipGraphInstantiate / hipGraphLaunch); other streams may join
 * the capture through events. The stream must not be the default stream. */
typedef struct btd_graph_s *btd_graph_t;
btd_status btd_graph_begin(void *stream);
btd_status btd_graph_end(void *stream, btd_graph_t *graph);
btd_status btd_graph_launch(btd_graph_t graph, void *stream);
btd_status btd_graph_destroy(btd_graph_t graph);

/* ---- Allocation-free solve ------------------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The non-periodic
 * solve of btd_solve_noncyclic (same kernels, same in-place contract and layouts; npart = 0
 * lets the library choose as btd_plan_create does, npart = 1 the single-partition solve) with
 * no plan: the scratch a plan would allocate (the two-stage solve's reduced system of
 * 2*npart rows per system, its partition table, the zero-pivot words) lives in one static
 * device array of BTD_NOALLOC_SCRATCH_BYTES per device, held until the process ends. A solve
 * needing more (about pad64(nsys) * 2*npart * (2m^2 + m) * 8 bytes plus 4*(nsys + 1)) is
 * refused with BTD_ERR_INVALID_ARG before anything runs. The call allocates nothing (no
 * hipMalloc, hipFree or allocating library routine); its first call on a device looks up the
 * array's address there. Every call on a device shares that device's array, so calls that
 * overlap on one device (other streams, other threads) race; calls in one stream are ordered.
 * Zero pivots are not reported (the words are cleared by every call). Asynchronous. */
#define BTD_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
btd_status btd_solve_noncyclic_noalloc(double *a, const double *b, double *c, double *d, int m, int nsys,
                                       int nrow, int npart, int layout, void *stream);

/* Library version string. */
const char *btd_version(void);

#ifdef __cplusplus
}
#endif
#endif
