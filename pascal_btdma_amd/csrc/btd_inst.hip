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

hipError_t BTD_NAME(launch_two_stage_)(int layout, double* A, const double* B, double* C,
                                       double* D, int nsys, int nrow, const int* pf,
                                       const int* pl, int npart, double* redA, double* redC,
                                       double* redD, int* flag, hipStream_t s, hipEvent_t* ev) {
    return launch_two_stage_m<BTD_M>(layout, A, B, C, D, nsys, nrow, pf, pl, npart, redA, redC,
                                     redD, flag, s, ev);
}

hipError_t BTD_NAME(launch_generate_)(int kind, uint64_t seed, int layout, int nsys, int nrow,
                                      double* A, double* B, double* C, double* D, float* Af,
                                      float* Bf, float* Cf, hipStream_t s) {
    return launch_generate_m<BTD_M>(kind, seed, layout, nsys, nrow, A, B, C, D, Af, Bf, Cf, s);
}

}  // namespace btd
