// Launchers for the mixed-size batch kernels (btd_batch.h).
#include <algorithm>

#include "btd_batch.h"

namespace btd {

hipError_t launch_batch(const BatchDev& b, int kind, int grid, const double* A, const double* B,
                        double* C, double* D, int* counter, int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, counter);
    if (kind == 1)
        hipLaunchKernelGGL((k_batch<1>), dim3(grid), dim3(64), 0, s, b, A, B, C, D, counter, flag);
    else if (kind == 2)
        hipLaunchKernelGGL((k_batch<2>), dim3(grid), dim3(64), 0, s, b, A, B, C, D, counter, flag);
    else
        hipLaunchKernelGGL((k_batch<0>), dim3(grid), dim3(64), 0, s, b, A, B, C, D, counter, flag);
    return hipGetLastError();
}

hipError_t launch_batch_generate(const BatchDev& b, const int* slot_sys, int max_rows, int kind,
                                 uint64_t seed, double* A, double* B, double* C, double* D,
                                 hipStream_t s) {
    const dim3 grid((unsigned)((max_rows + 3) / 4), (unsigned)std::min(b.npacks, 65535)), block(256);
    hipLaunchKernelGGL(k_batch_generate, grid, block, 0, s, b, slot_sys, kind, seed, A, B, C, D);
    return hipGetLastError();
}

}  // namespace btd
