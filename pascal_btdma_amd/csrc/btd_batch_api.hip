// Host side of the mixed-size batch (packing, autotuned launch parameters) and of the
// graph-capture helpers. No counterpart in the reference (its plan has one m and one
// nrow); the batch extends the plan/solve contract to mixed m and uneven n.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/pascal_btdma_amd.h"
#include "btd_internal.h"


namespace {
btd_status bfail(btd_status s, const std::string& m) {
    btd::set_last_error(m.c_str());
    return s;
}
}  // namespace

struct btd_batch_s {
    int nsys = 0, npacks = 0, max_rows = 0;
    std::vector<int> m, n;
    // packs: systems of one m, sorted longest first, batch_width(m) per pack
    std::vector<int> pack_m, pack_rows, pack_count, pack_slot0, slot_rows, slot_sys;
    std::vector<int64_t> mat_off, vec_off;
    std::vector<int> sys_pack, sys_slot;
    int64_t mat_elems = 0, vec_elems = 0;
    void* dev = nullptr;  // one allocation holding every device table
    btd::BatchDev bd{};
    const int* d_slot_sys = nullptr;
    int* d_counter = nullptr;
    int* d_flag = nullptr;
    int* d_items = nullptr;  // [count] [pack0 | npack | mode] x max_items
    int max_items = 0;
    int placement = 0;
    int grid = 0;
    hipStream_t last_stream = nullptr;
};

static void batch_free(btd_batch_s* b) {
    if (!b) return;
    if (b->dev) (void)hipFree(b->dev);
    delete b;
}

// Work items for a placement policy: 0 every pack one lane per system in VGPRs; 1 m >= 6 in
// LDS (one lane per system); 2 m >= 6 cooperative (m lanes per system, two packs per item);
// 3 m >= 4 cooperative; 4 every m >= 2 cooperative; 5 as 4, with every item whose cost
// exceeds a wave's mean share of the batch (total cost / grid) split into two twisted items
// (one pack each, two groups of m lanes per system, half the rows per group), so that the
// longest systems stop setting the end of the launch; 6 as 4 with every cooperative item
// twisted. One-lane items take floor(64/W) packs. Longest estimated work first.
static btd_status build_items(btd_batch_s* b, int placement, int grid) {
    struct Item {
        int pack, npack, mode;
        double cost;
    };
    std::vector<Item> items;
    const int coop_from = placement == 2 ? 6 : placement == 3 ? 4 : placement >= 4 ? 2 : 1000;
    auto coop_cost = [](double rows, double mm) { return rows * mm * (mm + 4) * 4.0; };
    // runs of consecutive packs with the same m
    std::vector<std::pair<int, int>> runs;
    for (int p = 0; p < b->npacks;) {
        int q = p;
        while (q < b->npacks && b->pack_m[q] == b->pack_m[p]) ++q;
        runs.push_back({p, q});
        p = q;
    }
    double share = 0.0;  // a wave's mean share of the cooperative cost
    if (placement == 5) {
        for (auto [p, q] : runs) {
            const int m = b->pack_m[p];
            if (m < coop_from) continue;
            for (int k = p; k < q; k += 2) share += coop_cost(b->pack_rows[k], m);
        }
        share /= grid;
    }
    for (auto [p, q] : runs) {
        const int m = b->pack_m[p];
        const double mm = m;
        if (m >= coop_from) {
            for (int k = p; k < q; k += 2) {
                const int np = std::min(2, q - k);
                const int rows = b->pack_rows[k];
                if (placement == 6 || (placement == 5 && coop_cost(rows, mm) > share)) {
                    for (int j = 0; j < np; ++j)
                        items.push_back({k + j, 1, 3, coop_cost((b->pack_rows[k + j] + 1) / 2, mm)});
                } else {
                    items.push_back({k, np, 2, coop_cost(rows, mm)});
                }
            }
        } else {
            const int per = btd::kTile / btd::batch_width(m);
            const int mode = (placement == 1 && m >= 6) ? 1 : 0;
            for (int k = p; k < q; k += per)
                items.push_back({k, std::min(per, q - k), mode, (double)b->pack_rows[k] * mm * mm * (mm + 4)});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
    if ((int)items.size() > b->max_items) return bfail(BTD_ERR_INVALID_ARG, "batch: too many work items");
    // the count goes with the table (the int before it): a graph captured earlier replays
    // whatever item list is current, never a count that does not match it
    std::vector<int> host(1 + 3 * (size_t)b->max_items, 0);
    host[0] = (int)items.size();
    for (size_t k = 0; k < items.size(); ++k) {
        host[1 + k] = items[k].pack;
        host[1 + b->max_items + k] = items[k].npack;
        host[1 + 2 * b->max_items + k] = items[k].mode;
    }
    // a solve or graph replay of this batch still in flight (on any stream) reads the table:
    // let every queued kernel finish before the synchronous copy rewrites it
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(b->d_items - 1, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bfail(BTD_ERR_HIP, std::string("batch items: ") + hipGetErrorString(e));
    b->placement = placement;
    return BTD_OK;
}

extern "C" {

btd_status btd_batch_create(btd_batch_t* batch, int nsys, const int* m, const int* n) {
    if (!batch || !m || !n || nsys < 1) return bfail(BTD_ERR_INVALID_ARG, "btd_batch_create: bad argument");
    *batch = nullptr;
    for (int s = 0; s < nsys; ++s) {
        if (m[s] < 1 || m[s] > BTD_MAX_M)
            return bfail(BTD_ERR_INVALID_ARG, "btd_batch_create: requires 1 <= m <= 10 for every system");
        if (n[s] < 2) return bfail(BTD_ERR_INVALID_ARG, "btd_batch_create: requires n >= 2 for every system");
    }
    auto* b = new btd_batch_s;
    b->nsys = nsys;
    b->m.assign(m, m + nsys);
    b->n.assign(n, n + nsys);
    std::vector<int> idx(nsys);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) {
        if (m[a] != m[c]) return m[a] < m[c];
        return n[a] > n[c];
    });
    b->sys_pack.assign(nsys, 0);
    b->sys_slot.assign(nsys, 0);
    int64_t mo = 0, vo = 0;
    for (size_t i = 0; i < idx.size();) {
        const int mm = m[idx[i]];
        const int W = btd::batch_width(mm);
        size_t j = i;
        while (j < idx.size() && j - i < (size_t)W && m[idx[j]] == mm) ++j;
        const int p = (int)b->pack_m.size();
        const int rows = n[idx[i]];
        b->pack_m.push_back(mm);
        b->pack_rows.push_back(rows);
        b->pack_count.push_back((int)(j - i));
        b->pack_slot0.push_back((int)b->slot_rows.size());
        b->mat_off.push_back(mo);
        b->vec_off.push_back(vo);
        for (size_t l = i; l < j; ++l) {
            b->slot_rows.push_back(n[idx[l]]);
            b->slot_sys.push_back(idx[l]);
            b->sys_pack[idx[l]] = p;
            b->sys_slot[idx[l]] = (int)(l - i);
        }
        mo += (int64_t)rows * mm * mm * W;
        vo += (int64_t)rows * mm * W;
        b->max_rows = std::max(b->max_rows, rows);
        i = j;
    }
    b->npacks = (int)b->pack_m.size();
    b->mat_elems = mo;
    b->vec_elems = vo;
    // device tables in one allocation
    const size_t P = b->npacks, S = b->slot_rows.size();
    b->max_items = (int)P;  // at most one item per pack (twisted); fewer otherwise
    const size_t bytes_i = sizeof(int) * (P * 4 + S * 2 + 3 + 3 * (size_t)b->max_items);
    const size_t bytes = ((bytes_i + 15) / 16) * 16 + sizeof(int64_t) * 2 * P;
    std::vector<char> host(bytes, 0);
    int* hi = (int*)host.data();
    size_t o = 0;
    auto put = [&](const std::vector<int>& v) {
        std::memcpy(hi + o, v.data(), v.size() * sizeof(int));
        size_t at = o;
        o += v.size();
        return at;
    };
    const size_t o_m = put(b->pack_m), o_rows = put(b->pack_rows), o_cnt = put(b->pack_count),
                 o_s0 = put(b->pack_slot0), o_sr = put(b->slot_rows), o_ss = put(b->slot_sys);
    const size_t o_counter = o, o_flag = o + 1, o_items = o + 3;  // o + 2: the item count
    int64_t* h64 = (int64_t*)(host.data() + ((bytes_i + 15) / 16) * 16);
    std::memcpy(h64, b->mat_off.data(), sizeof(int64_t) * P);
    std::memcpy(h64 + P, b->vec_off.data(), sizeof(int64_t) * P);
    hipError_t e = hipMalloc(&b->dev, bytes);
    if (e == hipSuccess) e = hipMemcpy(b->dev, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        batch_free(b);
        return bfail(BTD_ERR_HIP, std::string("btd_batch_create: ") + hipGetErrorString(e));
    }
    int* di = (int*)b->dev;
    int64_t* d64 = (int64_t*)((char*)b->dev + ((bytes_i + 15) / 16) * 16);
    b->d_items = di + o_items;
    b->bd = btd::BatchDev{b->d_items, b->d_items + b->max_items, b->d_items + 2 * b->max_items,
                          di + o_m, di + o_rows, di + o_cnt, di + o_s0, di + o_sr, d64, d64 + P,
                          b->d_items - 1, b->npacks};
    b->d_slot_sys = di + o_ss;
    b->d_counter = di + o_counter;
    b->d_flag = di + o_flag;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    b->grid = 4 * cus;  // one 64-lane wave per SIMD
    btd_status st = build_items(b, 0, b->grid);
    if (st != BTD_OK) {
        batch_free(b);
        return st;
    }
    *batch = b;
    return BTD_OK;
}

btd_status btd_batch_destroy(btd_batch_t batch) {
    batch_free(batch);
    return BTD_OK;
}

btd_status btd_batch_sizes(btd_batch_t b, size_t* mat_elems, size_t* vec_elems, int* npacks) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_sizes: NULL batch");
    if (mat_elems) *mat_elems = (size_t)b->mat_elems;
    if (vec_elems) *vec_elems = (size_t)b->vec_elems;
    if (npacks) *npacks = b->npacks;
    return BTD_OK;
}

btd_status btd_batch_locate(btd_batch_t b, int sys, size_t* mat_off, size_t* vec_off, int* stride) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_locate: NULL batch");
    if (sys < 0 || sys >= b->nsys) return bfail(BTD_ERR_INVALID_ARG, "btd_batch_locate: sys out of range");
    const int p = b->sys_pack[sys], l = b->sys_slot[sys];
    if (mat_off) *mat_off = (size_t)(b->mat_off[p] + l);
    if (vec_off) *vec_off = (size_t)(b->vec_off[p] + l);
    if (stride) *stride = btd::batch_width(b->pack_m[p]);
    return BTD_OK;
}

btd_status btd_batch_set_params(btd_batch_t b, int placement, int grid) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_set_params: NULL batch");
    if (placement < 0 || placement > 6 || grid < 1)
        return bfail(BTD_ERR_INVALID_ARG, "btd_batch_set_params: placement in 0..6, grid >= 1");
    btd_status st = build_items(b, placement, grid);
    if (st != BTD_OK) return st;
    b->grid = grid;
    return BTD_OK;
}

btd_status btd_batch_generate(btd_batch_t b, int kind, uint64_t seed, double* a, double* bb,
                              double* c, double* d, void* stream) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_generate: NULL batch");
    if (!a || !bb || !c || !d || kind < 0 || kind >= btd::kGenKinds)
        return bfail(BTD_ERR_INVALID_ARG, "btd_batch_generate: bad argument");
    hipError_t e = btd::launch_batch_generate(b->bd, b->d_slot_sys, b->max_rows, kind, seed, a, bb, c,
                                              d, (hipStream_t)stream);
    if (e != hipSuccess) return bfail(BTD_ERR_HIP, std::string("btd_batch_generate: ") + hipGetErrorString(e));
    return BTD_OK;
}

btd_status btd_batch_solve(btd_batch_t b, double* a, const double* bb, double* c, double* d,
                           void* stream) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_solve: NULL batch");
    if (!a || !bb || !c || !d) return bfail(BTD_ERR_INVALID_ARG, "btd_batch_solve: NULL array");
    b->last_stream = (hipStream_t)stream;
    const int kind = b->placement == 0 ? 0 : b->placement == 1 ? 1 : 2;
    hipError_t e = btd::launch_batch(b->bd, kind, b->grid, a, bb, c, d, b->d_counter, b->d_flag,
                                     (hipStream_t)stream);
    if (e != hipSuccess) return bfail(BTD_ERR_HIP, std::string("btd_batch_solve: ") + hipGetErrorString(e));
    return BTD_OK;
}

btd_status btd_batch_check(btd_batch_t b, int* singular) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_check: NULL batch");
    hipError_t e = hipStreamSynchronize(b->last_stream);
    int h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, b->d_flag, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && h) e = hipMemset(b->d_flag, 0, sizeof(int));
    if (e != hipSuccess) return bfail(BTD_ERR_HIP, std::string("btd_batch_check: ") + hipGetErrorString(e));
    if (singular) *singular = h ? 1 : 0;
    return BTD_OK;
}

// ---------------------------------------------------------------------------- graphs --
struct btd_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

btd_status btd_graph_begin(void* stream) {
    if (!stream) return bfail(BTD_ERR_INVALID_ARG, "btd_graph_begin: capture needs a non-default stream");
    hipError_t e = hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail(BTD_ERR_HIP, std::string("btd_graph_begin: ") + hipGetErrorString(e));
    return BTD_OK;
}

btd_status btd_graph_end(void* stream, btd_graph_t* graph) {
    if (!stream || !graph) return bfail(BTD_ERR_INVALID_ARG, "btd_graph_end: bad argument");
    *graph = nullptr;
    auto* g = new btd_graph_s;
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return bfail(BTD_ERR_HIP, std::string("btd_graph_end: ") + hipGetErrorString(e));
    }
    *graph = g;
    return BTD_OK;
}

btd_status btd_graph_launch(btd_graph_t graph, void* stream) {
    if (!graph) return bfail(BTD_ERR_NOT_CREATED, "btd_graph_launch: NULL graph");
    hipError_t e = hipGraphLaunch(graph->exec, (hipStream_t)stream);
    if (e != hipSuccess) return bfail(BTD_ERR_HIP, std::string("btd_graph_launch: ") + hipGetErrorString(e));
    return BTD_OK;
}

btd_status btd_graph_destroy(btd_graph_t graph) {
    if (!graph) return BTD_OK;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return BTD_OK;
}

}  // extern "C"
