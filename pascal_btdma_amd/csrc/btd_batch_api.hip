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
    int nsys = 0, ntiles = 0, max_rows = 0;
    std::vector<int> m, n;
    std::vector<int> tile_m, tile_rows, tile_count, lane_rows, lane_sys;
    std::vector<int64_t> mat_off, vec_off;
    std::vector<int> sys_tile, sys_lane;
    int64_t mat_elems = 0, vec_elems = 0;
    void* dev = nullptr;  // one allocation holding every device table
    btd::BatchDev bd{};
    const int* d_lane_sys = nullptr;
    int* d_counter = nullptr;
    int* d_flag = nullptr;
    int* d_items = nullptr;  // [tile | lane0 | mode] x max_items
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

// Work items for a placement policy: 0 every tile one lane per system in VGPRs; 1 m >= 6 in
// LDS (one lane per system); 2 m >= 6 cooperative (m lanes per system, floor(64/m) systems
// per item); 3 m >= 4 cooperative; 4 every m >= 2 cooperative; 5 as 4, with every item whose
// cost exceeds a wave's mean share of the batch (total cost / grid) made twisted (two groups
// of m lanes per system, floor(64/(2m)) systems per item, half the rows per group), so that
// the longest systems stop setting the end of the launch. Longest estimated work first.
static btd_status build_items(btd_batch_s* b, int placement, int grid) {
    struct Item {
        int tile, lane0, mode;
        double cost;
    };
    std::vector<Item> items;
    const int coop_from = placement == 2 ? 6 : placement == 3 ? 4 : placement >= 4 ? 2 : 1000;
    auto coop_cost = [](double rows, double mm) { return rows * mm * (mm + 4) * 4.0; };
    double share = 0.0;  // a wave's mean share of the cooperative cost
    if (placement == 5) {
        for (int t = 0; t < b->ntiles; ++t) {
            const int m = b->tile_m[t];
            if (m < coop_from) continue;
            const int G = btd::kTile / m;
            for (int l0 = 0; l0 < b->tile_count[t]; l0 += G)
                share += coop_cost(b->lane_rows[t * btd::kTile + l0], m);
        }
        share /= grid;
    }
    for (int t = 0; t < b->ntiles; ++t) {
        const int m = b->tile_m[t];
        const double mm = m;
        if (m >= coop_from) {
            const int G = btd::kTile / m, G2 = btd::kTile / (2 * m);
            for (int l0 = 0; l0 < b->tile_count[t];) {
                const int rows = b->lane_rows[t * btd::kTile + l0];
                if (placement == 5 && G2 >= 1 && coop_cost(rows, mm) > share) {
                    items.push_back({t, l0, 3, coop_cost((rows - 1) - (rows - 1) / 2 + 1, mm)});
                    l0 += G2;
                } else {
                    items.push_back({t, l0, 2, coop_cost(rows, mm)});
                    l0 += G;
                }
            }
        } else {
            const int mode = (placement == 1 && m >= 6) ? 1 : 0;
            items.push_back({t, 0, mode, (double)b->tile_rows[t] * mm * mm * (mm + 4)});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
    if ((int)items.size() > b->max_items) return bfail(BTD_ERR_INVALID_ARG, "batch: too many work items");
    // the count goes with the table (the int before it): a graph captured earlier replays
    // whatever item list is current, never a count that does not match it
    std::vector<int> host(1 + 3 * (size_t)b->max_items, 0);
    host[0] = (int)items.size();
    for (size_t k = 0; k < items.size(); ++k) {
        host[1 + k] = items[k].tile;
        host[1 + b->max_items + k] = items[k].lane0;
        host[1 + 2 * b->max_items + k] = items[k].mode;
    }
    hipError_t e = hipMemcpy(b->d_items - 1, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice);
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
    b->sys_tile.assign(nsys, 0);
    b->sys_lane.assign(nsys, 0);
    int64_t mo = 0, vo = 0;
    for (size_t i = 0; i < idx.size();) {
        const int mm = m[idx[i]];
        size_t j = i;
        while (j < idx.size() && j - i < (size_t)btd::kTile && m[idx[j]] == mm) ++j;
        const int t = (int)b->tile_m.size();
        const int rows = n[idx[i]];
        b->tile_m.push_back(mm);
        b->tile_rows.push_back(rows);
        b->tile_count.push_back((int)(j - i));
        b->mat_off.push_back(mo);
        b->vec_off.push_back(vo);
        for (int l = 0; l < btd::kTile; ++l) {
            const bool used = i + l < j;
            b->lane_rows.push_back(used ? n[idx[i + l]] : 0);
            b->lane_sys.push_back(used ? idx[i + l] : -1);
            if (used) {
                b->sys_tile[idx[i + l]] = t;
                b->sys_lane[idx[i + l]] = l;
            }
        }
        mo += (int64_t)rows * mm * mm * btd::kTile;
        vo += (int64_t)rows * mm * btd::kTile;
        b->max_rows = std::max(b->max_rows, rows);
        i = j;
    }
    b->ntiles = (int)b->tile_m.size();
    b->mat_elems = mo;
    b->vec_elems = vo;
    // device tables in one allocation
    const size_t T = b->ntiles;
    b->max_items = (int)T * ((btd::kTile + 1) / 2);  // cooperative m = 2 cuts a tile into 2 items
    const size_t bytes_i = sizeof(int) * (T * 3 + T * btd::kTile * 2 + 3 + 3 * (size_t)b->max_items);
    const size_t bytes = ((bytes_i + 15) / 16) * 16 + sizeof(int64_t) * 2 * T;
    std::vector<char> host(bytes, 0);
    int* hi = (int*)host.data();
    size_t o = 0;
    auto put = [&](const std::vector<int>& v) {
        std::memcpy(hi + o, v.data(), v.size() * sizeof(int));
        size_t at = o;
        o += v.size();
        return at;
    };
    const size_t o_m = put(b->tile_m), o_rows = put(b->tile_rows), o_cnt = put(b->tile_count),
                 o_lr = put(b->lane_rows), o_ls = put(b->lane_sys);
    const size_t o_counter = o, o_flag = o + 1, o_items = o + 3;  // o + 2: the item count
    int64_t* h64 = (int64_t*)(host.data() + ((bytes_i + 15) / 16) * 16);
    std::memcpy(h64, b->mat_off.data(), sizeof(int64_t) * T);
    std::memcpy(h64 + T, b->vec_off.data(), sizeof(int64_t) * T);
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
                          di + o_m, di + o_rows, di + o_cnt, di + o_lr, d64, d64 + T, b->d_items - 1,
                          b->ntiles};
    b->d_lane_sys = di + o_ls;
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

btd_status btd_batch_sizes(btd_batch_t b, size_t* mat_elems, size_t* vec_elems, int* ntiles) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_sizes: NULL batch");
    if (mat_elems) *mat_elems = (size_t)b->mat_elems;
    if (vec_elems) *vec_elems = (size_t)b->vec_elems;
    if (ntiles) *ntiles = b->ntiles;
    return BTD_OK;
}

btd_status btd_batch_locate(btd_batch_t b, int sys, size_t* mat_off, size_t* vec_off) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_locate: NULL batch");
    if (sys < 0 || sys >= b->nsys) return bfail(BTD_ERR_INVALID_ARG, "btd_batch_locate: sys out of range");
    const int t = b->sys_tile[sys], l = b->sys_lane[sys];
    if (mat_off) *mat_off = (size_t)(b->mat_off[t] + l);
    if (vec_off) *vec_off = (size_t)(b->vec_off[t] + l);
    return BTD_OK;
}

btd_status btd_batch_set_params(btd_batch_t b, int placement, int grid) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_set_params: NULL batch");
    if (placement < 0 || placement > 5 || grid < 1)
        return bfail(BTD_ERR_INVALID_ARG, "btd_batch_set_params: placement in 0..5, grid >= 1");
    btd_status st = build_items(b, placement, grid);
    if (st != BTD_OK) return st;
    b->grid = grid;
    return BTD_OK;
}

btd_status btd_batch_generate(btd_batch_t b, int kind, uint64_t seed, double* a, double* bb,
                              double* c, double* d, void* stream) {
    if (!b) return bfail(BTD_ERR_NOT_CREATED, "btd_batch_generate: NULL batch");
    if (!a || !bb || !c || !d || kind < 0 || kind > 3)
        return bfail(BTD_ERR_INVALID_ARG, "btd_batch_generate: bad argument");
    hipError_t e = btd::launch_batch_generate(b->bd, b->d_lane_sys, b->max_rows, kind, seed, a, bb, c,
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
