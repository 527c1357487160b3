// Instantiates the kernel templates for one block size m = BTD_M (the build compiles this
// file once per m, so the ten instantiations compile in parallel).
#include "btd_kernels.h"

#ifndef BTD_M
#error "compile with -DBTD_M=<m>"
#endif

#define BTD_CAT2(a, b) a##b
#define BTD_CAT(a, b) BTD_CAT2(a, b)
#define BTD_NAME(f) BTD_CAT(f, BTD_M)

namespace btd {

hipError_t BTD_NAME(launch_plain_)(int layout, const double* A, const double* B, double* C,
                                   double* D, int nsys, int nrow, int* flag, hipStream_t s) {
    return launch_plain_m<BTD_M>(layout, A, B, C, D, nsys, nrow, flag, s);
}

hipError_t BTD_NAME(launch_plain_twisted_)(int layout, const double* A, const double* B, double* C,
                                           double* D, int nsys, int nrow, int* flag, hipStream_t s) {
    return launch_plain_twisted_m<BTD_M>(layout, A, B, C, D, nsys, nrow, flag, s);
}

hipError_t BTD_NAME(launch_two_stage_)(int layout, double* A, const double* B, double* C,
                                       double* D, int nsys, int nrow, const int* pf,
                                       const int* pl, int npart, double* redA, double* redC,
                                       double* redD, int* flag, hipStream_t s, hipEvent_t* ev,
                                       bool cyclic) {
    return launch_two_stage_m<BTD_M>(layout, A, B, C, D, nsys, nrow, pf, pl, npart, redA, redC,
                                     redD, flag, s, ev, cyclic);
}

hipError_t BTD_NAME(launch_cyclic_)(int layout, double* A, const double* B, double* C, double* D,
                                    int nsys, int nrow, int* flag, hipStream_t s) {
    return launch_cyclic_m<BTD_M>(layout, A, B, C, D, nsys, nrow, flag, s);
}

hipError_t BTD_NAME(launch_generate_)(int kind, uint64_t seed, int layout, int nsys, int nrow,
                                      double* A, double* B, double* C, double* D, float* Af,
                                      float* Bf, float* Cf, hipStream_t s) {
    return launch_generate_m<BTD_M>(kind, seed, layout, nsys, nrow, A, B, C, D, Af, Bf, Cf, s);
}

hipError_t BTD_NAME(launch_plain_save_)(int layout, const double* A, const double* B, const double* C,
                                        double* Cw, const double* Din, double* D, double* F, uint64_t* PV,
                                        int nsys, int nrow, int* flag, hipStream_t s) {
    return launch_plain_save_m<BTD_M>(layout, A, B, C, Cw, Din, D, F, PV, nsys, nrow, flag, s);
}

hipError_t BTD_NAME(launch_adjoint_)(int layout, const double* A, const double* Cp,
                                     const double* F, const uint64_t* PV, const double* X,
                                     const double* W, double* Lam, double* dA, double* dB,
                                     double* dC, int nsys, int nrow, hipStream_t s) {
    return launch_adjoint_m<BTD_M>(layout, A, Cp, F, PV, X, W, Lam, dA, dB, dC, nsys, nrow, s);
}

hipError_t BTD_NAME(launch_f32r_)(int layout, const float* A, const float* B, const float* C,
                                  const double* D, double* X, float* Wc, float* Wv, int nsys,
                                  int nrow, int sweeps, int* flag, hipStream_t s, bool transposed) {
    return launch_f32r_m<BTD_M>(layout, A, B, C, D, X, Wc, Wv, nsys, nrow, sweeps, flag, s, transposed);
}

hipError_t BTD_NAME(launch_plain_f32s_)(int layout, const float* A, const float* B, const float* C,
                                        double* Cw, double* D, int nsys, int nrow, int* flag,
                                        hipStream_t s) {
    return launch_plain_f32s_m<BTD_M>(layout, A, B, C, Cw, D, nsys, nrow, flag, s);
}

hipError_t BTD_NAME(launch_plain_ref_)(int layout, const double* A, const double* B, const double* C,
                                       double* Cw, double* D, int nsys, int nrow, int* flag, hipStream_t s) {
    return launch_plain_ref_m<BTD_M>(layout, A, B, C, Cw, D, nsys, nrow, flag, s);
}

hipError_t BTD_NAME(launch_outer_f32_)(int layout, const double* Lam, const double* X, float* dA,
                                       float* dB, float* dC, int nsys, int nrow, hipStream_t s) {
    return launch_outer_f32_m<BTD_M>(layout, Lam, X, dA, dB, dC, nsys, nrow, s);
}

}  // namespace btd
