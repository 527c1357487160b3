// Host dispatch from a runtime block size m to the per-m instantiations (btd_inst.hip),
// plus the layout-conversion kernel.
#include "btd_internal.h"

namespace btd {

#define BTD_DECL_PLAIN_TWISTED(M) \
    hipError_t launch_plain_twisted_##M(int, const double*, const double*, double*, double*, int, int, int*, hipStream_t);
BTD_DECL_PLAIN_TWISTED(1)
BTD_DECL_PLAIN_TWISTED(2)
BTD_DECL_PLAIN_TWISTED(3)
BTD_DECL_PLAIN_TWISTED(4)
BTD_DECL_PLAIN_TWISTED(5)
BTD_DECL_PLAIN_TWISTED(6)
BTD_DECL_PLAIN_TWISTED(7)
BTD_DECL_PLAIN_TWISTED(8)
BTD_DECL_PLAIN_TWISTED(9)
BTD_DECL_PLAIN_TWISTED(10)
#undef BTD_DECL_PLAIN_TWISTED

hipError_t launch_plain_1(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_1(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_1(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_1(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_2(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_2(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_2(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_2(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_3(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_3(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_3(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_3(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_4(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_4(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_4(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_4(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_5(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_5(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_5(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_5(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_6(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_6(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_6(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_6(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_7(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_7(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_7(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_7(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_8(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_8(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_8(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_8(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_9(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_9(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_9(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_9(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_10(int, const double*, const double*, double*, double*, int, int, int*,
                          hipStream_t);
hipError_t launch_two_stage_10(int, double*, const double*, double*, double*, int, int, const int*,
                              const int*, int, double*, double*, double*, int*, hipStream_t,
                              hipEvent_t*, bool);
hipError_t launch_cyclic_10(int, double*, const double*, double*, double*, int, int, int*, hipStream_t);
hipError_t launch_generate_10(int, uint64_t, int, int, int, double*, double*, double*, double*,
                             float*, float*, float*, hipStream_t);

hipError_t launch_plain_save_1(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_1(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_1(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_1(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_1(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_1(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_2(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_2(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_2(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_2(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_2(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_2(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_3(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_3(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_3(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_3(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_3(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_3(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_4(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_4(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_4(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_4(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_4(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_4(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_5(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_5(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_5(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_5(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_5(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_5(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_6(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_6(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_6(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_6(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_6(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_6(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_7(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_7(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_7(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_7(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_7(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_7(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_8(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_8(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_8(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_8(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_8(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_8(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_9(int, const double*, const double*, const double*, double*, const double*,
                               double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_9(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_9(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_9(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_9(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_9(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

hipError_t launch_plain_save_10(int, const double*, const double*, const double*, double*, const double*,
                                double*, double*,
                               uint64_t*, int, int, int*, hipStream_t);
hipError_t launch_adjoint_10(int, const double*, const double*, const double*, const uint64_t*,
                            const double*, const double*, double*, double*, double*, double*, int,
                            int, hipStream_t);
hipError_t launch_f32r_10(int, const float*, const float*, const float*, const double*, double*,
                         float*, float*, int, int, int, int*, hipStream_t, bool);
hipError_t launch_outer_f32_10(int, const double*, const double*, float*, float*, float*, int, int,
                              hipStream_t);
hipError_t launch_plain_f32s_10(int, const float*, const float*, const float*, double*, double*, int,
                               int, int*, hipStream_t);
hipError_t launch_plain_ref_10(int, const double*, const double*, const double*, double*, double*, int, int, int*,
                              hipStream_t);

// ---------------------------------------------------------------------------- convert --
__global__ __launch_bounds__(256) void k_convert(int src_layout, int dst_layout, int E, int nsys,
                                                 int nrow, const double* __restrict__ src,
                                                 double* __restrict__ dst) {
    const int64_t total = (int64_t)nsys * nrow * E;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total) return;
    // idx enumerates (e, row, sys) with sys fastest
    const int64_t sys = idx % nsys, rest = idx / nsys;
    const int64_t row = rest % nrow, e = rest / nrow;
    dst[elem_offset(dst_layout, sys, row, e, E, nsys, nrow)] =
        src[elem_offset(src_layout, sys, row, e, E, nsys, nrow)];
}

#define BTD_DISPATCH_M(m, CALL)                  \
    switch (m) {                                 \
        case 1: return CALL(1);                  \
        case 2: return CALL(2);                  \
        case 3: return CALL(3);                  \
        case 4: return CALL(4);                  \
        case 5: return CALL(5);                  \
        case 6: return CALL(6);                  \
        case 7: return CALL(7);                  \
        case 8: return CALL(8);                  \
        case 9: return CALL(9);                  \
        case 10: return CALL(10);                \
        default: return hipErrorInvalidValue;    \
    }

hipError_t launch_plain(int m, int layout, const double* A, const double* B, double* C, double* D,
                        int nsys, int nrow, int* flag, hipStream_t s) {
#define CALL(M) launch_plain_##M(layout, A, B, C, D, nsys, nrow, flag, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_plain_twisted(int m, int layout, const double* A, const double* B, double* C, double* D,
                                int nsys, int nrow, int* flag, hipStream_t s) {
#define CALL(M) launch_plain_twisted_##M(layout, A, B, C, D, nsys, nrow, flag, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_two_stage(int m, int layout, double* A, const double* B, double* C, double* D,
                            int nsys, int nrow, const int* pf, const int* pl, int npart,
                            double* redA, double* redC, double* redD, int* flag, hipStream_t s,
                            hipEvent_t* ev, bool cyclic) {
#define CALL(M) \
    launch_two_stage_##M(layout, A, B, C, D, nsys, nrow, pf, pl, npart, redA, redC, redD, flag, s, ev, cyclic)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_generate(int m, int kind, uint64_t seed, int layout, int nsys, int nrow, double* A,
                           double* B, double* C, double* D, float* Af, float* Bf, float* Cf,
                           hipStream_t s) {
#define CALL(M) launch_generate_##M(kind, seed, layout, nsys, nrow, A, B, C, D, Af, Bf, Cf, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_plain_save(int m, int layout, const double* A, const double* B, const double* C,
                             double* Cw, const double* Din, double* D, double* F, uint64_t* PV, int nsys,
                             int nrow, int* flag, hipStream_t s) {
#define CALL(M) launch_plain_save_##M(layout, A, B, C, Cw, Din, D, F, PV, nsys, nrow, flag, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_adjoint(int m, int layout, const double* A, const double* Cp, const double* F,
                          const uint64_t* PV, const double* X, const double* W, double* Lam,
                          double* dA, double* dB, double* dC, int nsys, int nrow, hipStream_t s) {
#define CALL(M) launch_adjoint_##M(layout, A, Cp, F, PV, X, W, Lam, dA, dB, dC, nsys, nrow, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_f32r(int m, int layout, const float* A, const float* B, const float* C,
                       const double* D, double* X, float* Wc, float* Wv, int nsys, int nrow,
                       int sweeps, int* flag, hipStream_t s, bool transposed) {
#define CALL(M) launch_f32r_##M(layout, A, B, C, D, X, Wc, Wv, nsys, nrow, sweeps, flag, s, transposed)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_cyclic(int m, int layout, double* A, const double* B, double* C, double* D,
                         int nsys, int nrow, int* flag, hipStream_t s) {
#define CALL(M) launch_cyclic_##M(layout, A, B, C, D, nsys, nrow, flag, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_plain_f32s(int m, int layout, const float* A, const float* B, const float* C,
                             double* Cw, double* D, int nsys, int nrow, int* flag, hipStream_t s) {
#define CALL(M) launch_plain_f32s_##M(layout, A, B, C, Cw, D, nsys, nrow, flag, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_plain_ref(int m, int layout, const double* A, const double* B, const double* C, double* Cw,
                            double* D, int nsys, int nrow, int* flag, hipStream_t s) {
#define CALL(M) launch_plain_ref_##M(layout, A, B, C, Cw, D, nsys, nrow, flag, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_outer_f32(int m, int layout, const double* Lam, const double* X, float* dA,
                            float* dB, float* dC, int nsys, int nrow, hipStream_t s) {
#define CALL(M) launch_outer_f32_##M(layout, Lam, X, dA, dB, dC, nsys, nrow, s)
    BTD_DISPATCH_M(m, CALL)
#undef CALL
}

hipError_t launch_convert(int src_layout, int dst_layout, int E, int nsys, int nrow,
                          const double* src, double* dst, hipStream_t s) {
    const int64_t total = (int64_t)nsys * nrow * E;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_convert, grid, block, 0, s, src_layout, dst_layout, E, nsys, nrow, src, dst);
    return hipGetLastError();
}

}  // namespace btd
