// btd_solve_noncyclic_noalloc: the non-periodic solve without a plan and without any device
// allocation, for callers that cannot allocate (code under graph capture, tight serving loops,
// harnesses that forbid hipMalloc). Its kernels are those of btd_solve_noncyclic; the scratch
// a plan would allocate (the two-stage solve's reduced system, the partition table and the
// zero-pivot words) lives in one file-scope device array instead.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/pascal_btdma_amd.h"
#include "btd_internal.h"

// The static scratch: BTD_NOALLOC_SCRATCH_BYTES (64 MiB) on every device the library runs on,
// held until the process ends. In host code `scratch` names a host-side stand-in, never the
// device memory; kernels get the address hipGetSymbolAddress returns on each device
// (scratch_on_device), and host code uses only that address.
static __device__ __align__(256) unsigned char scratch[BTD_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kScratchBytes = BTD_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// The device address of `scratch` on the current device, looked up once per device.
hipError_t scratch_on_device(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

// Clears the zero-pivot words and writes the partition table [first..., len...] of
// btd_partition_1d(0, nrow - 1, npart, r) (a kernel, not a memset or a host copy, so the call
// can be captured in a graph and replayed).
__global__ void k_noalloc_init(int* flag, int nflag, int* part, int nrow, int npart) {
    const int total = nflag + (part ? npart : 0);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        if (i < nflag) {
            flag[i] = 0;
        } else {
            const int r = i - nflag, base = nrow / npart, rem = nrow % npart;
            part[r] = r * base + (r < rem ? r : rem);
            part[npart + r] = base + (r < rem ? 1 : 0);
        }
    }
}

btd_status fail(btd_status s, const std::string& msg) {
    btd::set_last_error(msg.c_str());
    return s;
}

}  // namespace

extern "C" {

btd_status btd_solve_noncyclic_noalloc(double* a, const double* b, double* c, double* d, int m, int nsys,
                                       int nrow, int npart, int layout, void* stream) {
    btd::set_last_error("");
    const char* what = "btd_solve_noncyclic_noalloc";
    if (m < 1 || m > BTD_MAX_M) return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": requires 1 <= m <= 10");
    if (nsys < 1 || nrow < 2)
        return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": requires nsys > 0 and nrow >= 2");
    if (npart < 0) return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": requires npart >= 0");
    if (npart == 0) npart = btd_auto_npart(m, nsys, nrow);
    if (npart > 1 && nrow < 3 * npart)
        return fail(BTD_ERR_INVALID_ARG,
                    std::string(what) + ": the two-stage solve needs at least 3 rows per partition (nrow >= 3*npart)");
    if (layout != BTD_LAYOUT_REF_GPU && layout != BTD_LAYOUT_TILED)
        return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": layout must be BTD_LAYOUT_REF_GPU or BTD_LAYOUT_TILED");
    if (!a || !b || !c || !d) return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": NULL array");

    // scratch layout: zero-pivot words (1 + nsys ints) | partition table (2 npart ints) |
    // reduced A, C (tiles * 2 npart * m^2 doubles each) | reduced D (tiles * 2 npart * m)
    const size_t flag_bytes = align256(sizeof(int) * (1 + (size_t)nsys));
    size_t part_bytes = 0, redm = 0, redv = 0;
    if (npart > 1) {
        const size_t rows = (size_t)btd::pad_nsys(nsys) * 2 * (size_t)npart;
        part_bytes = align256(sizeof(int) * 2 * (size_t)npart);
        redm = align256(sizeof(double) * rows * m * m);
        redv = align256(sizeof(double) * rows * m);
    }
    const size_t need = flag_bytes + part_bytes + 2 * redm + redv;
    if (need > kScratchBytes)
        return fail(BTD_ERR_INVALID_ARG,
                    std::string(what) + ": this solve needs " + std::to_string(need) +
                        " bytes of scratch, more than the static scratch array's " + std::to_string(kScratchBytes) +
                        " (fewer partitions or systems per call, or a plan: btd_plan_create + btd_solve_noncyclic)");

    void* base = nullptr;
    hipError_t e = scratch_on_device(&base);
    if (e != hipSuccess) return fail(BTD_ERR_HIP, std::string(what) + ": scratch address: " + hipGetErrorString(e));
    unsigned char* p = static_cast<unsigned char*>(base);
    int* flag = reinterpret_cast<int*>(p);
    int* part = npart > 1 ? reinterpret_cast<int*>(p + flag_bytes) : nullptr;
    double* redA = npart > 1 ? reinterpret_cast<double*>(p + flag_bytes + part_bytes) : nullptr;
    double* redC = npart > 1 ? reinterpret_cast<double*>(p + flag_bytes + part_bytes + redm) : nullptr;
    double* redD = npart > 1 ? reinterpret_cast<double*>(p + flag_bytes + part_bytes + 2 * redm) : nullptr;

    hipStream_t s = (hipStream_t)stream;
    const int nflag = 1 + nsys, items = nflag + (npart > 1 ? npart : 0);
    const int grid = (items + 255) / 256 < 1024 ? (items + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_noalloc_init, dim3(grid), dim3(256), 0, s, flag, nflag, part, nrow, npart);
    e = hipGetLastError();
    if (e == hipSuccess) {
        if (npart == 1)
            e = btd::launch_plain(m, layout, a, b, c, d, nsys, nrow, flag, s);
        else
            e = btd::launch_two_stage(m, layout, a, b, c, d, nsys, nrow, part, part + npart, npart, redA, redC,
                                      redD, flag, s, nullptr, false);
    }
    if (e != hipSuccess) return fail(BTD_ERR_HIP, std::string(what) + ": launch: " + hipGetErrorString(e));
    return BTD_OK;
}

}  // extern "C"
