// Mixed-size batches: systems of different block size m and row count n solved by one
// persistent kernel.
//
// Packing: systems are grouped by m, sorted by n (longest first) and cut into packs of
// W = batch_width(m) systems (64 for m = 1, else floor(32/m): the systems of one twisted work
// item). A pack is stored interleaved with its own row count (its longest system's):
// element e of row `row` of the system in slot s of pack p is at mat_off[p] + (row*E + e)*W + s
// (E = m*m), vectors likewise with E = m. So the W systems of a pack that a work item solves
// together occupy whole, contiguous cache lines: a cooperative load instruction (one block
// element per lane, m lanes per system) reads M*W consecutive doubles.
// Work items (consecutive packs of one m): one lane per system over floor(64/W) packs;
// cooperative, m lanes per system over 2 packs; or twisted, two groups of m lanes per system
// over 1 pack. Items are ordered by decreasing estimated cost (longest-processing-time
// first) and consumed through an atomic counter by a grid sized to the device; every wave
// exits when the counter passes the number of items.
#pragma once

#include "btd_kernels.h"

namespace btd {


// Plain block Thomas for one lane of a tile (rows 0..n-1), blocks in VGPRs.
template <int M>
__device__ __forceinline__ bool plain_lane(const double* __restrict__ A, const double* __restrict__ B,
                                           double* C, double* D, int64_t mo, int64_t vo, int n) {
    constexpr int E = M * M;
    constexpr int64_t es = batch_width(M);
    double Cp[M][M], dp[M];
    bool ok = true;
    for (int q = 0; q < n; ++q) {
        const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
        double R[M][M], v[M];
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int i = 0; i < M; ++i) R[i][j] = __builtin_nontemporal_load(B + rm + (i + M * j) * es);
#pragma unroll
        for (int i = 0; i < M; ++i) v[i] = D[rv + i * es];
        if (q > 0) {
#pragma unroll
            for (int k = 0; k < M; ++k) {
                double ak[M];
#pragma unroll
                for (int i = 0; i < M; ++i) ak[i] = __builtin_nontemporal_load(A + rm + (i + M * k) * es);
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    v[i] = fma(-ak[i], dp[k], v[i]);
#pragma unroll
                    for (int j = 0; j < M; ++j) R[i][j] = fma(-ak[i], Cp[k][j], R[i][j]);
                }
            }
        }
        int piv[M];
        ok = lu_factor<M>(R, piv) && ok;
        lu_solve<M>(R, piv, v);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            dp[i] = v[i];
            D[rv + i * es] = v[i];
        }
        if (q + 1 < n) {
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double c[M];
#pragma unroll
                for (int i = 0; i < M; ++i) c[i] = C[rm + (i + M * j) * es];
                lu_solve<M>(R, piv, c);
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    Cp[i][j] = c[i];
                    C[rm + (i + M * j) * es] = c[i];
                }
            }
        }
    }
    for (int q = n - 2; q >= 0; --q) {
        const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
        double x[M];
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = D[rv + i * es];
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int i = 0; i < M; ++i) x[i] = fma(-C[rm + (i + M * j) * es], dp[j], x[i]);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            dp[i] = x[i];
            D[rv + i * es] = x[i];
        }
    }
    return ok;
}

// The same solve with the row's block R, its pivot order and the carried C' in LDS
// (lane-interleaved, [element][lane]); the pivot exchange is then a swap of two row indices
// instead of register selects over whole rows. Used for large m when the autotuner finds
// it faster. sR: M*M*64 doubles, sC: M*M*64 doubles, sP: M*64 ints.
template <int M>
__device__ bool plain_lane_lds(const double* __restrict__ A, const double* __restrict__ B,
                               double* C, double* D, int64_t mo, int64_t vo, int n,
                               double* __restrict__ sR, double* __restrict__ sC,
                               int* __restrict__ sP) {
    constexpr int E = M * M;
    constexpr int64_t es = batch_width(M);
    const int lane = threadIdx.x;
#define BTD_R(i, j) sR[((i) + M * (j)) * kTile + lane]
#define BTD_CP(i, j) sC[((i) + M * (j)) * kTile + lane]
#define BTD_P(i) sP[(i) * kTile + lane]
    double dp[M];
    bool ok = true;
    for (int q = 0; q < n; ++q) {
        const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
        double v[M];
        for (int e = 0; e < E; ++e) sR[e * kTile + lane] = __builtin_nontemporal_load(B + rm + e * es);
#pragma unroll
        for (int i = 0; i < M; ++i) v[i] = D[rv + i * es];
        if (q > 0) {
#pragma unroll
            for (int k = 0; k < M; ++k) {
                double ak[M];
#pragma unroll
                for (int i = 0; i < M; ++i) ak[i] = __builtin_nontemporal_load(A + rm + (i + M * k) * es);
#pragma unroll
                for (int i = 0; i < M; ++i) v[i] = fma(-ak[i], dp[k], v[i]);
                for (int j = 0; j < M; ++j) {
                    const double ckj = BTD_CP(k, j);
#pragma unroll
                    for (int i = 0; i < M; ++i) BTD_R(i, j) = fma(-ak[i], ckj, BTD_R(i, j));
                }
            }
        }
        for (int i = 0; i < M; ++i) BTD_P(i) = i;
        for (int k = 0; k < M; ++k) {
            int p = k;
            double best = fabs(BTD_R(BTD_P(k), k));
            for (int i = k + 1; i < M; ++i) {
                const double t = fabs(BTD_R(BTD_P(i), k));
                if (t > best) {
                    best = t;
                    p = i;
                }
            }
            const int pk = BTD_P(p);
            BTD_P(p) = BTD_P(k);
            BTD_P(k) = pk;
            ok = ok && (BTD_R(pk, k) != 0.0);
            const double r = 1.0 / BTD_R(pk, k);
            BTD_R(pk, k) = r;
            for (int i = k + 1; i < M; ++i) {
                const int pi = BTD_P(i);
                const double l = BTD_R(pi, k) * r;
                BTD_R(pi, k) = l;
                for (int j = k + 1; j < M; ++j) BTD_R(pi, j) = fma(-l, BTD_R(pk, j), BTD_R(pi, j));
            }
        }
        int perm[M];
#pragma unroll
        for (int i = 0; i < M; ++i) perm[i] = BTD_P(i);
        // P R = L U with L, U stored in the physical rows perm[i]
        auto solve = [&](double (&x)[M]) {
            double y[M];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                double s = 0.0;
#pragma unroll
                for (int r = 0; r < M; ++r) s = (perm[i] == r) ? x[r] : s;
#pragma unroll
                for (int k = 0; k < i; ++k) s = fma(-BTD_R(perm[i], k), y[k], s);
                y[i] = s;
            }
#pragma unroll
            for (int i = M - 1; i >= 0; --i) {
                double s = y[i];
#pragma unroll
                for (int j = i + 1; j < M; ++j) s = fma(-BTD_R(perm[i], j), x[j], s);
                x[i] = s * BTD_R(perm[i], i);
            }
        };
        solve(v);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            dp[i] = v[i];
            D[rv + i * es] = v[i];
        }
        if (q + 1 < n) {
            for (int j = 0; j < M; ++j) {
                double c[M];
#pragma unroll
                for (int i = 0; i < M; ++i) c[i] = C[rm + (i + M * j) * es];
                solve(c);
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    BTD_CP(i, j) = c[i];
                    C[rm + (i + M * j) * es] = c[i];
                }
            }
        }
    }
#undef BTD_R
#undef BTD_CP
#undef BTD_P
    for (int q = n - 2; q >= 0; --q) {
        const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
        double x[M];
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = D[rv + i * es];
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int i = 0; i < M; ++i) x[i] = fma(-C[rm + (i + M * j) * es], dp[j], x[i]);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            dp[i] = x[i];
            D[rv + i * es] = x[i];
        }
    }
    return ok;
}

// A cooperative item: the systems of packs p0 (and p0 + 1 when npack = 2), m lanes each.
template <int M>
__device__ __forceinline__ bool coop_item(const BatchDev& b, const double* A, const double* B,
                                          double* C, double* D, int p0, int npack, double* sh) {
    constexpr int W = batch_width(M);
    const int lane = threadIdx.x;
    const int g = lane / M, i = lane % M;
    const bool grouped = g < 2 * W && g / W < npack;
    const int pk = grouped ? p0 + g / W : p0, slot = g % W;
    const bool mine = grouped && slot < b.pack_count[pk];
    const int n = mine ? b.slot_rows[b.pack_slot0[pk] + slot] : 0;
    const int nmax = b.pack_rows[p0];  // systems are sorted longest first
    const int64_t mo = b.mat_off[pk] + (mine ? slot : 0), vo = b.vec_off[pk] + (mine ? slot : 0);
    // lanes past the last whole group read group 0's posts and write only their own lane's
    // slots, which no group reads
    const StridedAddr<M * M * W, W> am{mo};
    const StridedAddr<M * W, W> av{vo};
    // the row form at every m: the batch kernel also holds the one-lane blocks, and the
    // column form's registers on top of them spill (the batch ran 2x slower with it)
    return coop_rows<M, false, false>(A, B, C, D, am, av, i, grouped ? g * M : 0, n, nmax, b.pack_rows[pk], sh);
}

// A twisted item: the systems of pack p, two groups of m lanes each.
template <int M>
__device__ __forceinline__ bool twisted_item(const BatchDev& b, const double* A, const double* B,
                                             double* C, double* D, int p, double* sh) {
    constexpr int W = batch_width(M);
    const int lane = threadIdx.x;
    const int g = lane / M, i = lane % M;
    const int slot = g / 2;
    const bool grouped = slot < W;
    const bool top = grouped ? (g % 2 == 0) : true;
    const bool mine = grouped && slot < b.pack_count[p];
    const int n = mine ? b.slot_rows[b.pack_slot0[p] + slot] : 0;
    const int nmax = b.pack_rows[p];
    const int steps = (nmax - 1) - (nmax - 1) / 2;
    const int64_t mo = b.mat_off[p] + (mine ? slot : 0), vo = b.vec_off[p] + (mine ? slot : 0);
    // lanes past the last pair act as a top group over groups 0 and 1, writing only their
    // own lane's slots
    const int base = grouped ? g * M : 0;
    const int pbase = grouped ? (top ? base + M : base - M) : M;
    const StridedAddr<M * M * W, W> am{mo};
    const StridedAddr<M * W, W> av{vo};
    return coop_twisted<M, false, false, false>(A, B, C, D, am, av, i, base, pbase, top, n, steps, nmax, sh);
}

// Persistent solve of a packed batch. KIND selects the code a placement needs, so that the
// register allocation of the kernel is that of the paths it can run: 0 one lane per system
// for every m (VGPRs); 1 one lane per system, m >= 6 with blocks in LDS (105 KB of LDS);
// 2 one lane per system for m <= 5, m lanes per system (cooperative items, mode 2) and two
// groups of m lanes per system (twisted items, mode 3).
template <int KIND>
__global__ __launch_bounds__(64) void k_batch(BatchDev b, const double* __restrict__ A,
                                              const double* __restrict__ B, double* C, double* D,
                                              int* __restrict__ counter, int* __restrict__ flag) {
    bool ok = true;
    const int lane = threadIdx.x;
    const int nitems = b.nitems[0];
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __shfl(t, 0);
        if (t >= nitems) break;
        const int p0 = b.item_pack[t], npack = b.item_npack[t];
        const int mode = b.item_mode[t];
        const int m = b.pack_m[p0];
        if constexpr (KIND == 2) {
            if (mode >= 2) {
                __shared__ double sh[kCoopSlots * kTile];
                if (mode == 2) {
                    switch (m) {
                        case 2: ok = coop_item<2>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        case 3: ok = coop_item<3>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        case 4: ok = coop_item<4>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        case 5: ok = coop_item<5>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        case 6: ok = coop_item<6>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        case 7: ok = coop_item<7>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        case 8: ok = coop_item<8>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        case 9: ok = coop_item<9>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        case 10: ok = coop_item<10>(b, A, B, C, D, p0, npack, sh) && ok; break;
                        default: break;
                    }
                } else {
                    switch (m) {
                        case 2: ok = twisted_item<2>(b, A, B, C, D, p0, sh) && ok; break;
                        case 3: ok = twisted_item<3>(b, A, B, C, D, p0, sh) && ok; break;
                        case 4: ok = twisted_item<4>(b, A, B, C, D, p0, sh) && ok; break;
                        case 5: ok = twisted_item<5>(b, A, B, C, D, p0, sh) && ok; break;
                        case 6: ok = twisted_item<6>(b, A, B, C, D, p0, sh) && ok; break;
                        case 7: ok = twisted_item<7>(b, A, B, C, D, p0, sh) && ok; break;
                        case 8: ok = twisted_item<8>(b, A, B, C, D, p0, sh) && ok; break;
                        case 9: ok = twisted_item<9>(b, A, B, C, D, p0, sh) && ok; break;
                        case 10: ok = twisted_item<10>(b, A, B, C, D, p0, sh) && ok; break;
                        default: break;
                    }
                }
                // the next item's first posts must not overtake this item's last reads
                wave_lds_sync();
                continue;
            }
        }
        // one lane per system: lane l holds slot l % W of pack p0 + l / W
        const int W = batch_width(m);
        const int pk = p0 + lane / W, slot = lane % W;
        if (lane / W < npack && slot < b.pack_count[pk]) {
            const int n = b.slot_rows[b.pack_slot0[pk] + slot];
            const int64_t mo = b.mat_off[pk] + slot, vo = b.vec_off[pk] + slot;
            if constexpr (KIND == 1) {
                if (mode == 1) {
                    __shared__ double sR[kMaxM * kMaxM * kTile];
                    __shared__ double sC[kMaxM * kMaxM * kTile];
                    __shared__ int sP[kMaxM * kTile];
                    switch (m) {
                        case 6: ok = plain_lane_lds<6>(A, B, C, D, mo, vo, n, sR, sC, sP) && ok; break;
                        case 7: ok = plain_lane_lds<7>(A, B, C, D, mo, vo, n, sR, sC, sP) && ok; break;
                        case 8: ok = plain_lane_lds<8>(A, B, C, D, mo, vo, n, sR, sC, sP) && ok; break;
                        case 9: ok = plain_lane_lds<9>(A, B, C, D, mo, vo, n, sR, sC, sP) && ok; break;
                        case 10: ok = plain_lane_lds<10>(A, B, C, D, mo, vo, n, sR, sC, sP) && ok; break;
                        default: break;
                    }
                }
            }
            if (mode == 0) {
                switch (m) {
                    case 1: ok = plain_lane<1>(A, B, C, D, mo, vo, n) && ok; break;
                    case 2: ok = plain_lane<2>(A, B, C, D, mo, vo, n) && ok; break;
                    case 3: ok = plain_lane<3>(A, B, C, D, mo, vo, n) && ok; break;
                    case 4: ok = plain_lane<4>(A, B, C, D, mo, vo, n) && ok; break;
                    case 5: ok = plain_lane<5>(A, B, C, D, mo, vo, n) && ok; break;
                    default: break;
                }
                if constexpr (KIND == 0) {
                    switch (m) {
                        case 6: ok = plain_lane<6>(A, B, C, D, mo, vo, n) && ok; break;
                        case 7: ok = plain_lane<7>(A, B, C, D, mo, vo, n) && ok; break;
                        case 8: ok = plain_lane<8>(A, B, C, D, mo, vo, n) && ok; break;
                        case 9: ok = plain_lane<9>(A, B, C, D, mo, vo, n) && ok; break;
                        case 10: ok = plain_lane<10>(A, B, C, D, mo, vo, n) && ok; break;
                        default: break;
                    }
                }
            }
        }
    }
    if (!ok) atomicOr(flag, 1);
}

// Work-counter reset as a kernel (a memset node would not replay reliably in a graph).
__global__ void k_batch_reset(int* counter) { counter[0] = 0; }

// Generator for a packed batch: one thread per (pack, row, slot). Storage of a pack that holds
// no system row (slots past the pack's systems, rows past a shorter system's end) is written
// with zeros, so every element of a generated batch is defined.
template <int M>
__device__ __forceinline__ void gen_batch_elem(int kind, uint64_t seed, int64_t sys, int row, int n,
                                               int64_t mo, int64_t vo, double* A, double* B,
                                               double* C, double* D, bool real) {
    constexpr int64_t es = batch_width(M);
    double a[M][M], bb[M][M], c[M][M], d[M];
    if (real) {
        gen_row<M>(kind, seed, sys, row, n, a, bb, c, d);
    } else {
        for (int j = 0; j < M; ++j) {
            d[j] = 0.0;
            for (int i = 0; i < M; ++i) a[i][j] = bb[i][j] = c[i][j] = 0.0;
        }
    }
    const int64_t rm = mo + (int64_t)row * M * M * es, rv = vo + (int64_t)row * M * es;
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            A[rm + (i + M * j) * es] = a[i][j];
            B[rm + (i + M * j) * es] = bb[i][j];
            C[rm + (i + M * j) * es] = c[i][j];
        }
    for (int i = 0; i < M; ++i) D[rv + i * es] = d[i];
}

__global__ __launch_bounds__(256) void k_batch_generate(BatchDev b, const int* __restrict__ slot_sys,
                                                        int kind, uint64_t seed, double* A, double* B,
                                                        double* C, double* D) {
    const int row = blockIdx.x * 4 + threadIdx.x / kTile, slot = threadIdx.x % kTile;
    for (int p = blockIdx.y; p < b.npacks; p += gridDim.y) {
        if (slot >= batch_width(b.pack_m[p]) || row >= b.pack_rows[p]) continue;  // not storage of pack p
        const bool used = slot < b.pack_count[p];
        const int n = used ? b.slot_rows[b.pack_slot0[p] + slot] : 0;
        const bool real = used && row < n;
        const int64_t sys = used ? slot_sys[b.pack_slot0[p] + slot] : 0;
        const int64_t mo = b.mat_off[p] + slot, vo = b.vec_off[p] + slot;
        switch (b.pack_m[p]) {
            case 1: gen_batch_elem<1>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 2: gen_batch_elem<2>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 3: gen_batch_elem<3>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 4: gen_batch_elem<4>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 5: gen_batch_elem<5>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 6: gen_batch_elem<6>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 7: gen_batch_elem<7>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 8: gen_batch_elem<8>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 9: gen_batch_elem<9>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            case 10: gen_batch_elem<10>(kind, seed, sys, row, n, mo, vo, A, B, C, D, real); break;
            default: break;
        }
    }
}

}  // namespace btd
