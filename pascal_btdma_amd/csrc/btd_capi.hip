// Host side of the C ABI: plan lifetime, argument checks, dispatch and timing.
// The reference's equivalents are BtdmaGpuPlan::create/destroy and solve_noncyclic_impl
// (CUDA_CXX_Port/src/pascal_btdma_cuda.cu:853-1075).
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pascal_btdma_amd.h"
#include "btd_internal.h"

namespace {

thread_local std::string g_err;

btd_status fail(btd_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

btd_status hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return BTD_ERR_HIP;
}

#define BTD_HIP(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

}  // namespace

void btd::set_last_error(const char* msg) { g_err = msg; }

int btd::device_cus() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 256;
    if (dev < 64) {
        const int c = cached[dev].load(std::memory_order_relaxed);
        if (c > 0) return c;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (dev < 64) cached[dev].store(cus, std::memory_order_relaxed);
    return cus;
}

struct btd_plan_s {
    int m = 0, nsys = 0, nrow = 0, npart = 1, layout = BTD_LAYOUT_TILED;
    int form = BTD_FORM_PLAIN;  // how an npart = 1 solve runs (btd_plan_set_form)
    int device = 0;
    bool created = false;
    std::vector<int> first, len;  // partition rows (host)
    int* d_part = nullptr;        // [first..., len...] on the device
    double* redA = nullptr;       // reduced system (tiled, 2*npart rows); B = I is implicit
    double* redC = nullptr;
    double* redD = nullptr;
    int* d_flag = nullptr;  // zero-pivot flag (word 0), then one word per system
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    hipStream_t last_stream = nullptr;
    // The periodic solve of an npart = 1 plan runs as a two-stage solve over cyc_npart
    // partitions (0 until its first call sets them up; 1: too few rows, one periodic sweep).
    int cyc_npart = 0;
    int* cyc_part = nullptr;
    double *cyc_redA = nullptr, *cyc_redC = nullptr, *cyc_redD = nullptr;
    // BTD_FORM_WORKSPACE: C' of the single-partition solve goes here (the layout's elements of an
    // m*m block array), allocated when the form is set, placed by btd_plan_place_workspace
    double* ws = nullptr;
};

static void plan_free(btd_plan_s* p) {
    if (!p) return;
    for (void* q : {(void*)p->d_part, (void*)p->redA, (void*)p->redC, (void*)p->redD, (void*)p->cyc_part,
                    (void*)p->cyc_redA, (void*)p->cyc_redC, (void*)p->cyc_redD, (void*)p->ws})
        if (q) (void)hipFree(q);
    if (p->d_flag) (void)hipFree(p->d_flag);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    delete p;
}

extern "C" {

const char* btd_version(void) { return "pascal_btdma_amd 0.1 (gfx950)"; }

const char* btd_last_error(void) { return g_err.c_str(); }

size_t btd_matrix_index(int sys, int row, int p, int q, int nsys, int nrow, int m) {
    return (size_t)sys + (size_t)nsys * ((size_t)row + (size_t)nrow * ((size_t)p + (size_t)m * q));
}

size_t btd_vector_index(int sys, int row, int p, int nsys, int nrow) {
    return (size_t)sys + (size_t)nsys * ((size_t)row + (size_t)nrow * p);
}

void btd_partition_1d(int start, int end, int nprocs, int rank, int* first, int* last) {
    const int n = end - start + 1;
    const int base = n / nprocs, rem = n % nprocs;
    *first = rank * base + start + std::min(rank, rem);
    *last = *first + base - 1 + (rem > rank ? 1 : 0);
}

size_t btd_layout_elems(int layout, int nsys, int nrow, int elems_per_block) {
    return (size_t)btd::layout_elems(layout, nsys, nrow, elems_per_block);
}

// The partition count for stage-1 kernels with one lane per partition (the periodic solve's,
// and m >= 7): about one wave per SIMD (65536 lanes), >= 8 rows per partition; 1 from 32768
// systems (4096 for m >= 7, 8192 for m = 2, 4), and for systems of fewer than 128 rows at m >= 5,
// where the three launches and the reduced solve of the two-stage form cost more than the
// twisted solve's n/2 row steps (64 rows, 256-2048 systems: m = 5 1.2-1.65x, m = 8 0.96-1.43x
// faster single-partition).
static int auto_npart_one_lane(int m, int nsys, int nrow) {
    constexpr int64_t kLanes = 65536;
    if (nsys < 1 || nrow < 2) return 1;
    const int64_t one = m >= 7 ? 4096 : (m == 2 || m == 4) ? 8192 : kLanes / 2;
    if ((int64_t)nsys >= one || (m >= 5 && nrow < 128)) return 1;
    const int64_t want = (kLanes + nsys - 1) / nsys;
    const int64_t most = nrow / 8;
    const int64_t p = want < most ? want : most;
    return p < 2 ? 1 : (int)p;
}

static int pow2_floor(int64_t p) {
    int q = 1;
    while ((int64_t)q * 2 <= p) q *= 2;
    return q;
}

int btd_auto_npart(int m, int nsys, int nrow) {
    // One GPU: partitions are an occupancy choice. The single-kernel solve takes a batch that
    // fills the device on its own: 32768 systems at m = 1 (one lane per system); 12288 for m = 3
    // and 5, whose single-partition solve gives every system two lanes there, one from each end
    // (12288-24576 systems of 512 and 2048 rows, both layouts: 0.70-0.87x the time of the best
    // P in 2-8; at 8192 0.97-1.16x, so P = 2 stays there); 8192 for m = 2 and 4, whose twisted
    // solve's wavefronts read whole halves or quarters of lines; 4096 for m = 6 and 3072 for
    // m >= 7, whose single-partition solve is twisted or gives every system two, m or 2m lanes
    // (m = 7-10 at 3072 systems of 128-256 rows, both layouts: the one-lane rule's P = 22 took
    // 1.08-1.50x the time of P = 1).
    // Smaller batches, measured over m = 2-8, 256-16384 systems, 64-2048 rows (P = 1..128, the
    // best against this rule: geometric mean 1.02x, against the one-lane rule 1.10x):
    //   m <= 5 (stage 1 and 3 twisted, two lanes per partition): about 16384 partitions,
    //     >= 16 rows each, a power of two; at least 2 from 8192 systems; 1 below 128 rows
    //     for fewer than 8192 systems;
    //   m = 6: 1 up to 128 rows, else the one-lane rule's count rounded to a power of two;
    //   m >= 7: the one-lane rule (auto_npart_one_lane) below 3072 systems, except 1 for
    //     1024 systems or more of at most 192 rows (m = 7-10, 1024-2048 systems of 128-192
    //     rows, both layouts: P = 1 within 6% of the best P in 37 of 48 cases, at worst 1.12x;
    //     the one-lane rule's 16-24 partitions up to 1.46x the time of P = 1).
    if (nsys < 1 || nrow < 2) return 1;
    if (m >= 7) return (nsys >= 3072 || (nsys >= 1024 && nrow <= 192)) ? 1 : auto_npart_one_lane(m, nsys, nrow);
    const int64_t one = m == 6 ? 4096 : (m == 2 || m == 4) ? 8192 : (m == 3 || m == 5) ? 12288 : 32768;
    if ((int64_t)nsys >= one) return 1;
    int64_t want, most;
    if (m == 6) {
        if (nrow <= 128) return 1;
        want = (65536 + (int64_t)nsys - 1) / nsys;
        most = nrow / 8;
    } else {
        if (nrow < 128 && nsys < 8192) return 1;
        want = (16384 + (int64_t)nsys - 1) / nsys;
        if (nsys >= 8192 && want < 2) want = 2;
        most = nrow / 16;
    }
    const int64_t p = want < most ? want : most;
    return p < 2 ? 1 : pow2_floor(p);
}

// Row partitions of every system (btd_partition_1d) on the host and as a device table
// [first..., len...], and the tiled reduced system of 2*npart rows per system.
static hipError_t alloc_partitions(int m, int nsys, int nrow, int npart, std::vector<int>* first,
                                   std::vector<int>* len, int** d_part, double** redA, double** redC,
                                   double** redD) {
    std::vector<int> f(npart), n(npart);
    for (int r = 0; r < npart; ++r) {
        int a, b;
        btd_partition_1d(0, nrow - 1, npart, r, &a, &b);
        f[r] = a;
        n[r] = b - a + 1;
    }
    std::vector<int> tab(f);
    tab.insert(tab.end(), n.begin(), n.end());
    hipError_t e = hipMalloc(d_part, sizeof(int) * tab.size());
    if (e == hipSuccess) e = hipMemcpy(*d_part, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice);
    const size_t tiles = (size_t)btd::pad_nsys(nsys);
    const size_t rm = tiles * 2 * npart * m * m, rv = tiles * 2 * npart * m;
    if (e == hipSuccess) e = hipMalloc(redA, rm * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(redC, rm * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(redD, rv * sizeof(double));
    if (first) *first = f;
    if (len) *len = n;
    return e;
}

// Partitions of the periodic solve of an npart = 1 plan. The two-stage periodic solve
// (k_modi / k_modi_coop, a periodic reduced system, k_update) measured faster than one
// periodic sweep per lane (k_cyclic) at every size tried (m = 2-10, 256-131072 systems of
// 12-1024 rows: 1.3-1.9x at m <= 5 on large batches, 3-13x at m >= 7), with the partition
// count of btd_auto_npart where that is >= 2 and otherwise 4 (2 for fewer than 32 rows; P = 2,
// 3, 4, 8 were within 5% of each other on large batches). Fewer than 6 rows keep k_cyclic.
static int cyclic_npart(int m, int nsys, int nrow) {
    int p = auto_npart_one_lane(m, nsys, nrow);
    if (p < 2) p = nrow >= 32 ? 4 : 2;
    return nrow >= 3 * p ? p : 1;
}

btd_status btd_plan_create(btd_plan_t* plan, int m, int nsys, int nrow, int npart, int layout) {
    g_err.clear();
    if (!plan) return fail(BTD_ERR_INVALID_ARG, "btd_plan_create: plan pointer is NULL");
    *plan = nullptr;
    if (m < 1 || m > BTD_MAX_M)
        return fail(BTD_ERR_INVALID_ARG, "btd_plan_create: requires 1 <= m <= 10");
    if (nsys < 1 || nrow < 2)
        return fail(BTD_ERR_INVALID_ARG, "btd_plan_create: requires nsys > 0 and nrow >= 2");
    if (npart < 0)
        return fail(BTD_ERR_INVALID_ARG, "btd_plan_create: requires npart >= 1 (or 0: chosen by the library)");
    if (npart == 0) npart = btd_auto_npart(m, nsys, nrow);
    if (npart > 1 && nrow < 3 * npart)
        return fail(BTD_ERR_INVALID_ARG,
                    "btd_plan_create: the two-stage solve needs at least 3 rows per partition "
                    "(nrow >= 3*npart)");
    if (layout != BTD_LAYOUT_REF_GPU && layout != BTD_LAYOUT_TILED)
        return fail(BTD_ERR_INVALID_ARG,
                    "btd_plan_create: layout must be BTD_LAYOUT_REF_GPU or BTD_LAYOUT_TILED "
                    "(convert CPU-layout arrays with btd_convert_layout)");
    auto* p = new btd_plan_s;
    p->m = m;
    p->nsys = nsys;
    p->nrow = nrow;
    p->npart = npart;
    p->layout = layout;
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipMalloc(&p->d_flag, sizeof(int) * (1 + (size_t)nsys));
    if (e == hipSuccess) e = hipMemset(p->d_flag, 0, sizeof(int) * (1 + (size_t)nsys));
    for (int i = 0; i < 5 && e == hipSuccess; ++i) e = hipEventCreate(&p->ev[i]);
    if (e == hipSuccess && npart > 1)
        e = alloc_partitions(m, nsys, nrow, npart, &p->first, &p->len, &p->d_part, &p->redA, &p->redC,
                             &p->redD);
    if (e != hipSuccess) {
        plan_free(p);
        return hip_fail(e, "btd_plan_create");
    }
    p->created = true;
    *plan = p;
    return BTD_OK;
}

btd_status btd_plan_npart(btd_plan_t plan, int* npart) {
    g_err.clear();
    if (!plan || !plan->created) return fail(BTD_ERR_NOT_CREATED, "btd_plan_npart: uncreated plan");
    if (!npart) return fail(BTD_ERR_INVALID_ARG, "btd_plan_npart: NULL npart");
    *npart = plan->npart;
    return BTD_OK;
}

static size_t ws_bytes(const btd_plan_s* p) {
    return sizeof(double) * btd_layout_elems(p->layout, p->nsys, p->nrow, p->m * p->m);
}

// Leaving BTD_FORM_WORKSPACE frees the workspace (after the plan's last solve stream is done).
static void drop_workspace(btd_plan_s* p) {
    if (!p->ws) return;
    if (p->last_stream) (void)hipStreamSynchronize(p->last_stream);
    (void)hipDeviceSynchronize();
    (void)hipFree(p->ws);
    p->ws = nullptr;
}

btd_status btd_plan_set_form(btd_plan_t plan, int form) {
    g_err.clear();
    if (!plan || !plan->created) return fail(BTD_ERR_NOT_CREATED, "btd_plan_set_form: uncreated plan");
    if (form == BTD_FORM_PLAIN) {
        drop_workspace(plan);
        plan->form = form;
        return BTD_OK;
    }
    if (form == BTD_FORM_LAPACK_ORDER || form == BTD_FORM_WORKSPACE) {
        if (plan->npart != 1)
            return fail(BTD_ERR_INVALID_ARG, form == BTD_FORM_WORKSPACE
                                                 ? "btd_plan_set_form: the workspace form needs npart == 1"
                                                 : "btd_plan_set_form: the LAPACK-order form needs npart == 1");
        if (form == BTD_FORM_WORKSPACE && !plan->ws) {
            void* p = nullptr;
            hipError_t e = hipMalloc(&p, ws_bytes(plan));
            if (e != hipSuccess) return hip_fail(e, "btd_plan_set_form: workspace");
            plan->ws = static_cast<double*>(p);
        }
        if (form == BTD_FORM_LAPACK_ORDER) drop_workspace(plan);
        plan->form = form;
        return BTD_OK;
    }
    if (form != BTD_FORM_TWISTED) return fail(BTD_ERR_INVALID_ARG, "btd_plan_set_form: unknown form");
    if (plan->npart != 1 || plan->m < 2)
        return fail(BTD_ERR_INVALID_ARG, "btd_plan_set_form: the twisted form needs npart == 1 and m >= 2");
    drop_workspace(plan);
    plan->form = form;
    return BTD_OK;
}

btd_status btd_plan_form(btd_plan_t plan, int* form) {
    g_err.clear();
    if (!plan || !plan->created) return fail(BTD_ERR_NOT_CREATED, "btd_plan_form: uncreated plan");
    if (!form) return fail(BTD_ERR_INVALID_ARG, "btd_plan_form: NULL form");
    *form = plan->form;
    return BTD_OK;
}

btd_status btd_plan_destroy(btd_plan_t plan) {
    g_err.clear();
    plan_free(plan);
    return BTD_OK;
}

static btd_status check_solve(btd_plan_t plan, const void* a, const void* b, const void* c,
                              const void* d, int m, int nsys, int nrow) {
    if (!plan || !plan->created)
        return fail(BTD_ERR_NOT_CREATED, "solve called with an uncreated plan");
    if (m != plan->m || nsys != plan->nsys || nrow != plan->nrow)
        return fail(BTD_ERR_DIM_MISMATCH, "solve dimensions do not match the plan");
    if (!a || !b || !c || !d) return fail(BTD_ERR_INVALID_ARG, "solve: NULL array");
    return BTD_OK;
}

// The periodic partitions of an npart = 1 plan, allocated by its first periodic solve
// (hipMalloc and a synchronous table copy). That first call cannot be captured into a graph:
// it is refused with a clear error when `s` is capturing.
static btd_status ensure_cyclic_partitions(btd_plan_t plan, const char* what, hipStream_t s) {
    if (plan->cyc_npart != 0) return BTD_OK;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    BTD_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return fail(BTD_ERR_INVALID_ARG,
                    std::string(what) + ": the plan's first periodic solve allocates its partition scratch and "
                                        "cannot be captured; make one periodic call before capturing");
    const int pc = cyclic_npart(plan->m, plan->nsys, plan->nrow);
    if (pc > 1) {
        hipError_t e = alloc_partitions(plan->m, plan->nsys, plan->nrow, pc, nullptr, nullptr, &plan->cyc_part,
                                        &plan->cyc_redA, &plan->cyc_redC, &plan->cyc_redD);
        if (e != hipSuccess) {  // free what was allocated; the next call tries again
            for (void** q : {(void**)&plan->cyc_part, (void**)&plan->cyc_redA, (void**)&plan->cyc_redC,
                             (void**)&plan->cyc_redD}) {
                if (*q) (void)hipFree(*q);
                *q = nullptr;
            }
            return hip_fail(e, (std::string(what) + ": partition scratch").c_str());
        }
    }
    plan->cyc_npart = pc;
    return BTD_OK;
}

static btd_status solve_impl(btd_plan_t plan, double* a, const double* b, double* c, double* d,
                             hipStream_t s, bool timed, bool cyclic = false) {
    plan->last_stream = s;
    if (plan->npart == 1 && cyclic) {
        btd_status st = ensure_cyclic_partitions(plan, "btd_solve_cyclic", s);
        if (st != BTD_OK) return st;
        if (plan->cyc_npart > 1)
            BTD_HIP(btd::launch_two_stage(plan->m, plan->layout, a, b, c, d, plan->nsys, plan->nrow,
                                          plan->cyc_part, plan->cyc_part + plan->cyc_npart, plan->cyc_npart,
                                          plan->cyc_redA, plan->cyc_redC, plan->cyc_redD, plan->d_flag, s,
                                          nullptr, true));
        else
            BTD_HIP(btd::launch_cyclic(plan->m, plan->layout, a, b, c, d, plan->nsys, plan->nrow,
                                       plan->d_flag, s));
    } else if (plan->npart == 1) {
        if (timed) BTD_HIP(hipEventRecord(plan->ev[0], s));
        if (plan->form == BTD_FORM_TWISTED)
            BTD_HIP(btd::launch_plain_twisted(plan->m, plan->layout, a, b, c, d, plan->nsys, plan->nrow,
                                              plan->d_flag, s));
        else if (plan->form == BTD_FORM_LAPACK_ORDER || plan->form == BTD_FORM_WORKSPACE)
            BTD_HIP(btd::launch_plain_ref(plan->m, plan->layout, a, b, c,
                                          plan->form == BTD_FORM_WORKSPACE ? plan->ws : c, d, plan->nsys,
                                          plan->nrow, plan->d_flag, s));
        else
            BTD_HIP(btd::launch_plain(plan->m, plan->layout, a, b, c, d, plan->nsys, plan->nrow,
                                      plan->d_flag, s));
        if (timed) BTD_HIP(hipEventRecord(plan->ev[1], s));
    } else {
        BTD_HIP(btd::launch_two_stage(plan->m, plan->layout, a, b, c, d, plan->nsys, plan->nrow,
                                      plan->d_part, plan->d_part + plan->npart, plan->npart,
                                      plan->redA, plan->redC, plan->redD, plan->d_flag, s,
                                      timed ? plan->ev : nullptr, cyclic));
    }
    return BTD_OK;
}

btd_status btd_solve_noncyclic(btd_plan_t plan, double* a, const double* b, double* c, double* d,
                               int m, int nsys, int nrow, void* stream) {
    g_err.clear();
    btd_status st = check_solve(plan, a, b, c, d, m, nsys, nrow);
    if (st != BTD_OK) return st;
    return solve_impl(plan, a, b, c, d, (hipStream_t)stream, false);
}

static double ev_seconds(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms * 1e-3;
}

// Reads the zero-pivot flag after the plan's last stream has drained; when it is set, copies
// the per-system words to `info` (if given) and clears flag and words.
static btd_status read_status(btd_plan_t plan, int* singular, int* info) {
    BTD_HIP(hipStreamSynchronize(plan->last_stream));
    int h = 0;
    BTD_HIP(hipMemcpy(&h, plan->d_flag, sizeof(int), hipMemcpyDeviceToHost));
    if (h) {
        if (info)
            BTD_HIP(hipMemcpy(info, plan->d_flag + 1, sizeof(int) * (size_t)plan->nsys,
                              hipMemcpyDeviceToHost));
        BTD_HIP(hipMemset(plan->d_flag, 0, sizeof(int) * (1 + (size_t)plan->nsys)));
    } else if (info) {
        std::memset(info, 0, sizeof(int) * (size_t)plan->nsys);
    }
    if (singular) *singular = h ? 1 : 0;
    return BTD_OK;
}

// ------------------------------------------------------------------ euler3d sweeps --
static btd_status check_sweep(btd_plan_t plan, int axis, int n1, int n2, int n3, const char* what) {
    if (!plan || !plan->created) return fail(BTD_ERR_NOT_CREATED, std::string(what) + ": uncreated plan");
    if (axis < 0 || axis > 2) return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": axis must be 0, 1 or 2");
    if (n1 < 1 || n2 < 1 || n3 < 1) return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": grid sizes must be >= 1");
    const int64_t nrow = axis == 0 ? n1 : (axis == 1 ? n2 : n3);
    const int64_t nsys = axis == 0 ? (int64_t)n2 * n3 : (axis == 1 ? (int64_t)n1 * n3 : (int64_t)n1 * n2);
    if (nrow != plan->nrow || nsys != plan->nsys)
        return fail(BTD_ERR_DIM_MISMATCH,
                    std::string(what) + ": the plan's nrow/nsys do not match the sweep's grid line "
                                        "length/count");
    return BTD_OK;
}

btd_status btd_euler_assemble(btd_plan_t plan, int axis, const double* U, const double* rhs, int n1,
                              int n2, int n3, double h, double dt, double eps, double* a, double* b,
                              double* c, double* d, void* stream) {
    g_err.clear();
    btd_status st = check_sweep(plan, axis, n1, n2, n3, "btd_euler_assemble");
    if (st != BTD_OK) return st;
    if (plan->m != 5) return fail(BTD_ERR_INVALID_ARG, "btd_euler_assemble: the Euler system has m = 5");
    if (!U || !a || !b || !c) return fail(BTD_ERR_INVALID_ARG, "btd_euler_assemble: NULL array");
    if ((rhs == nullptr) != (d == nullptr))
        return fail(BTD_ERR_INVALID_ARG, "btd_euler_assemble: rhs and d must both be given or both be NULL");
    if (!(h > 0.0)) return fail(BTD_ERR_INVALID_ARG, "btd_euler_assemble: h must be > 0");
    const double cc = dt / (4.0 * h);
    const double cd = eps * dt / (2.0 * (h * h));
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_euler_assemble(plan->layout, axis, U, rhs, n1, n2, n3, cc, cd, a, b, c, d,
                                       plan->nsys, plan->nrow, (hipStream_t)stream));
    return BTD_OK;
}

// The weno2d sweep along `axis` of an n1 x n2 grid against the plan (m = 3, cell triples).
static btd_status check_weno(btd_plan_t plan, int axis, int n1, int n2, const char* what) {
    const std::string w(what);
    if (!plan || !plan->created) return fail(BTD_ERR_NOT_CREATED, w + ": uncreated plan");
    if (axis != 0 && axis != 1) return fail(BTD_ERR_INVALID_ARG, w + ": axis must be 0 or 1");
    if (n1 < 1 || n2 < 1) return fail(BTD_ERR_INVALID_ARG, w + ": grid sizes must be >= 1");
    if (plan->m != 3) return fail(BTD_ERR_INVALID_ARG, w + ": the block rows are cell triples (m = 3)");
    const int n = axis == 0 ? n1 : n2, nsys = axis == 0 ? n2 : n1;
    if (n % 3 != 0) return fail(BTD_ERR_INVALID_ARG, w + ": the sweep's cell count must be a multiple of 3");
    if (plan->nrow != n / 3 || plan->nsys != nsys)
        return fail(BTD_ERR_DIM_MISMATCH, w + ": the plan's nrow/nsys do not match the sweep");
    return BTD_OK;
}

btd_status btd_weno_assemble(btd_plan_t plan, int axis, const double* U, const double* vel, int n1, int n2,
                             double h, double gdt, double* a, double* b, double* c, void* stream) {
    g_err.clear();
    btd_status st = check_weno(plan, axis, n1, n2, "btd_weno_assemble");
    if (st != BTD_OK) return st;
    if (!U || !vel || !a || !b || !c) return fail(BTD_ERR_INVALID_ARG, "btd_weno_assemble: NULL array");
    if (!(h > 0.0)) return fail(BTD_ERR_INVALID_ARG, "btd_weno_assemble: h must be > 0");
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_weno_assemble(plan->layout, axis, U, vel, n1, n2, h, gdt, a, b, c, plan->nsys, plan->nrow,
                                      (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_weno_sweep_cyclic(btd_plan_t plan, int axis, const double* U, const double* vel, int n1, int n2,
                                 double h, double gdt, double* work, double* d, void* stream) {
    g_err.clear();
    btd_status st = check_weno(plan, axis, n1, n2, "btd_weno_sweep_cyclic");
    if (st != BTD_OK) return st;
    if (!U || !vel || !work || !d) return fail(BTD_ERR_INVALID_ARG, "btd_weno_sweep_cyclic: NULL array");
    if (!(h > 0.0)) return fail(BTD_ERR_INVALID_ARG, "btd_weno_sweep_cyclic: h must be > 0");
    const int* part = plan->d_part;
    double *ra = plan->redA, *rc = plan->redC, *rd = plan->redD;
    int np = plan->npart;
    if (np == 1) {
        st = ensure_cyclic_partitions(plan, "btd_weno_sweep_cyclic", (hipStream_t)stream);
        if (st != BTD_OK) return st;
        if (plan->cyc_npart < 2)
            return fail(BTD_ERR_INVALID_ARG, "btd_weno_sweep_cyclic: needs at least 6 block rows (18 cells)");
        np = plan->cyc_npart;
        part = plan->cyc_part;
        ra = plan->cyc_redA;
        rc = plan->cyc_redC;
        rd = plan->cyc_redD;
    }
    const size_t nm = btd_layout_elems(plan->layout, plan->nsys, plan->nrow, 9);
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_weno_sweep_cyclic(plan->layout, axis, U, vel, n1, n2, h, gdt, work, work + nm, d, plan->nsys,
                                          plan->nrow, part, part + np, np, ra, rc, rd, plan->d_flag,
                                          (hipStream_t)stream));
    return BTD_OK;
}

static btd_status weno_move(btd_plan_t plan, int axis, bool to_d, double* field, double* d, int n1, int n2,
                            void* stream, const char* what) {
    g_err.clear();
    btd_status st = check_weno(plan, axis, n1, n2, what);
    if (st != BTD_OK) return st;
    if (!field || !d) return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": NULL array");
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_weno_move(plan->layout, axis, to_d, field, d, n1, plan->nsys, plan->nrow,
                                  (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_weno_gather(btd_plan_t plan, int axis, const double* field, int n1, int n2, double* d,
                           void* stream) {
    return weno_move(plan, axis, true, const_cast<double*>(field), d, n1, n2, stream, "btd_weno_gather");
}

btd_status btd_weno_scatter(btd_plan_t plan, int axis, const double* d, int n1, int n2, double* field,
                            void* stream) {
    return weno_move(plan, axis, false, field, const_cast<double*>(d), n1, n2, stream, "btd_weno_scatter");
}

btd_status btd_euler_sweep_cyclic(btd_plan_t plan, int axis, const double* U, const double* rhs,
                                  int n1, int n2, int n3, double h, double dt, double eps,
                                  double* work, double* out, void* stream) {
    g_err.clear();
    btd_status st = check_sweep(plan, axis, n1, n2, n3, "btd_euler_sweep_cyclic");
    if (st != BTD_OK) return st;
    if (plan->m != 5) return fail(BTD_ERR_INVALID_ARG, "btd_euler_sweep_cyclic: the Euler system has m = 5");
    if (plan->npart != 1)
        return fail(BTD_ERR_INVALID_ARG, "btd_euler_sweep_cyclic: needs a plan with npart == 1");
    if (plan->nrow < 3) return fail(BTD_ERR_INVALID_ARG, "btd_euler_sweep_cyclic: needs nrow >= 3");
    if (!U || !rhs || !work || !out) return fail(BTD_ERR_INVALID_ARG, "btd_euler_sweep_cyclic: NULL array");
    if (!(h > 0.0)) return fail(BTD_ERR_INVALID_ARG, "btd_euler_sweep_cyclic: h must be > 0");
    const double cc = dt / (4.0 * h);
    const double cd = eps * dt / (2.0 * (h * h));
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_euler_sweep_cyclic(plan->layout, axis, U, rhs, n1, n2, n3, cc, cd, work, out,
                                           plan->nsys, plan->nrow, plan->d_flag, (hipStream_t)stream));
    return BTD_OK;
}

static btd_status sweep_move(btd_plan_t plan, int axis, double* field, double* d, int n1, int n2,
                             int n3, bool scatter, void* stream, const char* what) {
    g_err.clear();
    btd_status st = check_sweep(plan, axis, n1, n2, n3, what);
    if (st != BTD_OK) return st;
    if (!field || !d) return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": NULL array");
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_sweep_move(plan->layout, axis, field, d, n1, n2, n3, plan->m, plan->nsys,
                                   plan->nrow, scatter, (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_sweep_gather(btd_plan_t plan, int axis, const double* field, int n1, int n2, int n3,
                            double* d, void* stream) {
    return sweep_move(plan, axis, const_cast<double*>(field), d, n1, n2, n3, false, stream,
                      "btd_sweep_gather");
}

btd_status btd_sweep_scatter(btd_plan_t plan, int axis, const double* d, int n1, int n2, int n3,
                             double* field, void* stream) {
    return sweep_move(plan, axis, field, const_cast<double*>(d), n1, n2, n3, true, stream,
                      "btd_sweep_scatter");
}

btd_status btd_plan_check(btd_plan_t plan, int* singular) {
    g_err.clear();
    if (!plan || !plan->created) return fail(BTD_ERR_NOT_CREATED, "btd_plan_check: uncreated plan");
    return read_status(plan, singular, nullptr);
}

btd_status btd_plan_info(btd_plan_t plan, int* info) {
    g_err.clear();
    if (!plan || !plan->created) return fail(BTD_ERR_NOT_CREATED, "btd_plan_info: uncreated plan");
    if (!info) return fail(BTD_ERR_INVALID_ARG, "btd_plan_info: NULL info");
    return read_status(plan, nullptr, info);
}

static btd_status check_single_partition(btd_plan_t plan, const char* what);

// Placement of a plan's workspace by timing: `solve(ws)` runs the workspace form's solve of the
// caller's arrays with candidate workspace ws on stream s, d (db bytes) restored from a copy
// before each solve and at the end. Shared by the FP64 and the FP32-storage solves.
extern "C++" {
template <class Solve>
static btd_status place_workspace_impl(btd_plan_t plan, double* d, int candidates, double* ms, hipStream_t s,
                                       const char* what, Solve solve) {
    if (plan->form != BTD_FORM_WORKSPACE || !plan->ws)
        return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": the plan's form is not BTD_FORM_WORKSPACE");
    if (candidates < 1 || candidates > BTD_MAX_WORKSPACE_CANDIDATES)
        return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": candidates must be in 1.." +
                                             std::to_string(BTD_MAX_WORKSPACE_CANDIDATES));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    BTD_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": allocates and synchronizes; cannot be captured");
    if (ms)
        for (int k = 0; k < candidates; ++k) ms[k] = -1.0;
    const size_t wb = ws_bytes(plan);
    const size_t db = sizeof(double) * btd_layout_elems(plan->layout, plan->nsys, plan->nrow, plan->m);
    BTD_HIP(hipStreamSynchronize(s));
    void* dsave = nullptr;
    BTD_HIP(hipMalloc(&dsave, db));
    std::vector<double*> cand{plan->ws};
    for (int k = 1; k < candidates; ++k) {  // as many as fit; the plan's own is the first
        void* p = nullptr;
        if (hipMalloc(&p, wb) != hipSuccess) {
            (void)hipGetLastError();
            break;
        }
        cand.push_back(static_cast<double*>(p));
    }
    std::vector<double> t(cand.size(), 0.0);
    hipError_t e = hipMemcpyAsync(dsave, d, db, hipMemcpyDeviceToDevice, s);
    // two timed solves on each candidate (the first of a fresh allocation maps its pages): the
    // faster counts
    for (size_t k = 0; k < cand.size() && e == hipSuccess; ++k) {
        double best = 0.0;
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
            e = hipMemcpyAsync(d, dsave, db, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = hipEventRecord(plan->ev[0], s);
            if (e == hipSuccess) e = solve(cand[k]);
            if (e == hipSuccess) e = hipEventRecord(plan->ev[1], s);
            if (e == hipSuccess) e = hipEventSynchronize(plan->ev[1]);
            if (e == hipSuccess) {
                const double v = ev_seconds(plan->ev[0], plan->ev[1]) * 1e3;
                best = rep == 0 || v < best ? v : best;
            }
        }
        t[k] = best;
    }
    size_t pick = 0;
    for (size_t k = 1; k < cand.size(); ++k)
        if (t[k] < t[pick]) pick = k;
    if (e == hipSuccess) e = hipMemcpyAsync(d, dsave, db, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) plan->ws = cand[pick];  // on failure the plan keeps its own
    for (size_t k = 0; k < cand.size(); ++k)
        if (cand[k] != plan->ws) (void)hipFree(cand[k]);
    (void)hipFree(dsave);
    if (e != hipSuccess) return hip_fail(e, what);
    if (ms)
        for (size_t k = 0; k < cand.size(); ++k) ms[k] = t[k];
    plan->last_stream = s;
    return BTD_OK;
}
}  // extern "C++"

btd_status btd_plan_place_workspace(btd_plan_t plan, const double* a, const double* b, const double* c,
                                    double* d, int m, int nsys, int nrow, int candidates, double* ms,
                                    void* stream) {
    g_err.clear();
    btd_status st = check_solve(plan, a, b, c, d, m, nsys, nrow);
    if (st != BTD_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    return place_workspace_impl(plan, d, candidates, ms, s, "btd_plan_place_workspace", [&](double* ws) {
        return btd::launch_plain_ref(plan->m, plan->layout, a, b, c, ws, d, plan->nsys, plan->nrow, plan->d_flag, s);
    });
}

btd_status btd_plan_place_workspace_f32s(btd_plan_t plan, const float* a, const float* b, const float* c,
                                         double* d, int candidates, double* ms, void* stream) {
    g_err.clear();
    btd_status st = check_single_partition(plan, "btd_plan_place_workspace_f32s");
    if (st != BTD_OK) return st;
    if (!a || !b || !c || !d) return fail(BTD_ERR_INVALID_ARG, "btd_plan_place_workspace_f32s: NULL array");
    hipStream_t s = (hipStream_t)stream;
    return place_workspace_impl(plan, d, candidates, ms, s, "btd_plan_place_workspace_f32s", [&](double* ws) {
        return btd::launch_plain_f32s(plan->m, plan->layout, a, b, c, ws, d, plan->nsys, plan->nrow, plan->d_flag, s);
    });
}

btd_status btd_solve_noncyclic_profiled(btd_plan_t plan, double* a, const double* b, double* c,
                                        double* d, int m, int nsys, int nrow, btd_timings* t,
                                        void* stream) {
    g_err.clear();
    btd_status st = check_solve(plan, a, b, c, d, m, nsys, nrow);
    if (st != BTD_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    // this call reports its own solve only: earlier unchecked zero pivots are dropped
    BTD_HIP(hipMemsetAsync(plan->d_flag, 0, sizeof(int) * (1 + (size_t)plan->nsys), s));
    st = solve_impl(plan, a, b, c, d, s, true);
    if (st != BTD_OK) return st;
    const int last = plan->npart == 1 ? 1 : 3;
    BTD_HIP(hipEventSynchronize(plan->ev[last]));
    if (t) {
        *t = btd_timings{};
        if (plan->npart == 1) {
            t->local_compute = ev_seconds(plan->ev[0], plan->ev[1]);
        } else {
            t->local_compute = ev_seconds(plan->ev[0], plan->ev[1]);
            t->reduced_compute = ev_seconds(plan->ev[1], plan->ev[2]);
            t->update_compute = ev_seconds(plan->ev[2], plan->ev[3]);
        }
        t->total = ev_seconds(plan->ev[0], plan->ev[last]);
    }
    int sing = 0;
    BTD_HIP(hipMemcpy(&sing, plan->d_flag, sizeof(int), hipMemcpyDeviceToHost));
    if (sing)  // the flag and per-system words stay set for btd_plan_info / btd_plan_check
        return fail(BTD_ERR_SINGULAR, "a block LU met a zero pivot (btd_plan_info names the systems)");
    return BTD_OK;
}

static btd_status check_single_partition(btd_plan_t plan, const char* what) {
    if (!plan || !plan->created) return fail(BTD_ERR_NOT_CREATED, std::string(what) + ": uncreated plan");
    if (plan->npart != 1)
        return fail(BTD_ERR_INVALID_ARG, std::string(what) + ": needs a plan with npart == 1");
    return BTD_OK;
}

btd_status btd_solve_noncyclic_save(btd_plan_t plan, const double* a, const double* b, double* c,
                                    double* d, double* factors, uint64_t* pivots, void* stream) {
    g_err.clear();
    btd_status st = check_single_partition(plan, "btd_solve_noncyclic_save");
    if (st != BTD_OK) return st;
    if (!a || !b || !c || !d || !factors || !pivots)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_noncyclic_save: NULL array");
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_plain_save(plan->m, plan->layout, a, b, c, c, d, d, factors, pivots, plan->nsys,
                                   plan->nrow, plan->d_flag, (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_solve_noncyclic_save_out(btd_plan_t plan, const double* a, const double* b, const double* c,
                                        const double* d, double* c_out, double* x, double* factors,
                                        uint64_t* pivots, void* stream) {
    g_err.clear();
    btd_status st = check_single_partition(plan, "btd_solve_noncyclic_save_out");
    if (st != BTD_OK) return st;
    if (!a || !b || !c || !d || !c_out || !x || !factors || !pivots)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_noncyclic_save_out: NULL array");
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_plain_save(plan->m, plan->layout, a, b, c, c_out, d, x, factors, pivots, plan->nsys,
                                   plan->nrow, plan->d_flag, (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_solve_adjoint(btd_plan_t plan, const double* a, const double* c_factored,
                             const double* factors, const uint64_t* pivots, const double* x,
                             const double* w, double* lambda, double* da, double* db, double* dc,
                             void* stream) {
    g_err.clear();
    btd_status st = check_single_partition(plan, "btd_solve_adjoint");
    if (st != BTD_OK) return st;
    if (!a || !c_factored || !factors || !pivots || !x || !w || !lambda || !da || !db || !dc)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_adjoint: NULL array");
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_adjoint(plan->m, plan->layout, a, c_factored, factors, pivots, x, w, lambda,
                                da, db, dc, plan->nsys, plan->nrow, (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_solve_noncyclic_f32r(btd_plan_t plan, const float* a, const float* b, const float* c,
                                    const double* d, double* x, float* work, int sweeps,
                                    void* stream) {
    g_err.clear();
    btd_status st = check_single_partition(plan, "btd_solve_noncyclic_f32r");
    if (st != BTD_OK) return st;
    if (!a || !b || !c || !d || !x || !work)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_noncyclic_f32r: NULL array");
    if (sweeps < 0 || sweeps > 8)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_noncyclic_f32r: sweeps must be in [0, 8]");
    plan->last_stream = (hipStream_t)stream;
    float* wc = work;
    float* wv = work + btd::layout_elems(plan->layout, plan->nsys, plan->nrow, plan->m * plan->m);
    BTD_HIP(btd::launch_f32r(plan->m, plan->layout, a, b, c, d, x, wc, wv, plan->nsys, plan->nrow,
                             sweeps, plan->d_flag, (hipStream_t)stream, false));
    return BTD_OK;
}

btd_status btd_solve_noncyclic_f32s(btd_plan_t plan, const float* a, const float* b, const float* c,
                                    double* d, double* work, void* stream) {
    g_err.clear();
    btd_status st = check_single_partition(plan, "btd_solve_noncyclic_f32s");
    if (st != BTD_OK) return st;
    if (!a || !b || !c || !d)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_noncyclic_f32s: NULL array");
    if (!work) {  // the plan's workspace (BTD_FORM_WORKSPACE), placed by btd_plan_place_workspace_f32s
        if (plan->form != BTD_FORM_WORKSPACE || !plan->ws)
            return fail(BTD_ERR_INVALID_ARG,
                        "btd_solve_noncyclic_f32s: NULL work needs a plan in BTD_FORM_WORKSPACE");
        work = plan->ws;
    }
    plan->last_stream = (hipStream_t)stream;
    BTD_HIP(btd::launch_plain_f32s(plan->m, plan->layout, a, b, c, work, d, plan->nsys, plan->nrow,
                                   plan->d_flag, (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_solve_transposed_f32r(btd_plan_t plan, const float* a, const float* b, const float* c,
                                     const double* w, double* lambda, float* work, int sweeps,
                                     void* stream) {
    g_err.clear();
    btd_status st = check_single_partition(plan, "btd_solve_transposed_f32r");
    if (st != BTD_OK) return st;
    if (!a || !b || !c || !w || !lambda || !work)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_transposed_f32r: NULL array");
    if (sweeps < 0 || sweeps > 8)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_transposed_f32r: sweeps must be in [0, 8]");
    plan->last_stream = (hipStream_t)stream;
    float* wc = work;
    float* wv = work + btd::layout_elems(plan->layout, plan->nsys, plan->nrow, plan->m * plan->m);
    BTD_HIP(btd::launch_f32r(plan->m, plan->layout, a, b, c, w, lambda, wc, wv, plan->nsys, plan->nrow,
                             sweeps, plan->d_flag, (hipStream_t)stream, true));
    return BTD_OK;
}

btd_status btd_outer_grads_f32(btd_plan_t plan, const double* lambda, const double* x, float* da,
                               float* db, float* dc, void* stream) {
    g_err.clear();
    btd_status st = check_single_partition(plan, "btd_outer_grads_f32");
    if (st != BTD_OK) return st;
    if (!lambda || !x || !da || !db || !dc)
        return fail(BTD_ERR_INVALID_ARG, "btd_outer_grads_f32: NULL array");
    BTD_HIP(btd::launch_outer_f32(plan->m, plan->layout, lambda, x, da, db, dc, plan->nsys, plan->nrow,
                                  (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_solve_cyclic(btd_plan_t plan, double* a, const double* b, double* c, double* d,
                            int m, int nsys, int nrow, void* stream) {
    g_err.clear();
    btd_status st = check_solve(plan, a, b, c, d, m, nsys, nrow);
    if (st != BTD_OK) return st;
    if (plan->npart == 1 && nrow < 3)
        return fail(BTD_ERR_INVALID_ARG, "btd_solve_cyclic: the periodic solve needs nrow >= 3");
    return solve_impl(plan, a, b, c, d, (hipStream_t)stream, false, true);
}

btd_status btd_convert_layout(int src_layout, int dst_layout, int E, int nsys, int nrow,
                              const double* src, double* dst, void* stream) {
    g_err.clear();
    if (E < 1 || nsys < 1 || nrow < 1 || !src || !dst || src_layout < 0 || src_layout > 2 ||
        dst_layout < 0 || dst_layout > 2)
        return fail(BTD_ERR_INVALID_ARG, "btd_convert_layout: bad argument");
    BTD_HIP(btd::launch_convert(src_layout, dst_layout, E, nsys, nrow, src, dst, (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_generate(int kind, uint64_t seed, int m, int nsys, int nrow, int layout, double* a,
                        double* b, double* c, double* d, void* stream) {
    g_err.clear();
    if (kind < 0 || kind >= btd::kGenKinds || m < 1 || m > BTD_MAX_M || nsys < 1 || nrow < 1 || layout < 0 ||
        layout > 2)
        return fail(BTD_ERR_INVALID_ARG, "btd_generate: bad argument");
    BTD_HIP(btd::launch_generate(m, kind, seed, layout, nsys, nrow, a, b, c, d, nullptr, nullptr,
                                 nullptr, (hipStream_t)stream));
    return BTD_OK;
}

btd_status btd_generate_f32(int kind, uint64_t seed, int m, int nsys, int nrow, int layout,
                            float* a, float* b, float* c, double* d, void* stream) {
    g_err.clear();
    if (kind < 0 || kind >= btd::kGenKinds || m < 1 || m > BTD_MAX_M || nsys < 1 || nrow < 1 || layout < 0 ||
        layout > 2)
        return fail(BTD_ERR_INVALID_ARG, "btd_generate_f32: bad argument");
    BTD_HIP(btd::launch_generate(m, kind, seed, layout, nsys, nrow, nullptr, nullptr, nullptr, d, a,
                                 b, c, (hipStream_t)stream));
    return BTD_OK;
}

}  // extern "C"

template <int M>
static void gen_host_m(int kind, uint64_t seed, int nrow, int sys0, int count, double* a,
                       double* b, double* c, double* d) {
    for (int row = 0; row < nrow; ++row)
        for (int s = 0; s < count; ++s) {
            double A[M][M], B[M][M], C[M][M], dd[M];
            btd::gen_row<M>(kind, seed, (int64_t)sys0 + s, row, nrow, A, B, C, dd);
            const size_t o = ((size_t)row * count + s) * M * M;
            for (int i = 0; i < M; ++i)
                for (int j = 0; j < M; ++j) {
                    if (a) a[o + i * M + j] = A[i][j];
                    if (b) b[o + i * M + j] = B[i][j];
                    if (c) c[o + i * M + j] = C[i][j];
                }
            if (d)
                for (int i = 0; i < M; ++i) d[((size_t)row * count + s) * M + i] = dd[i];
        }
}

extern "C" btd_status btd_generate_host(int kind, uint64_t seed, int m, int nsys, int nrow,
                                        int sys0, int count, double* a, double* b, double* c,
                                        double* d) {
    g_err.clear();
    if (kind < 0 || kind >= btd::kGenKinds || m < 1 || m > BTD_MAX_M || nsys < 1 || nrow < 1 || sys0 < 0 ||
        count < 0 || sys0 + count > nsys)
        return fail(BTD_ERR_INVALID_ARG, "btd_generate_host: bad argument");
    switch (m) {
        case 1: gen_host_m<1>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 2: gen_host_m<2>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 3: gen_host_m<3>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 4: gen_host_m<4>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 5: gen_host_m<5>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 6: gen_host_m<6>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 7: gen_host_m<7>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 8: gen_host_m<8>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 9: gen_host_m<9>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        case 10: gen_host_m<10>(kind, seed, nrow, sys0, count, a, b, c, d); break;
        default: break;
    }
    return BTD_OK;
}
