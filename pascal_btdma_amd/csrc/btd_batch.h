// Mixed-size batches: systems of different block size m and row count n solved by one
// persistent kernel.
//
// Packing: systems are grouped by m, sorted by n (longest first) and cut into tiles of up
// to 64 systems of the same m. Each tile is stored like the tiled layout with its own row
// count (the longest system of the tile): element e of row `row` of the system in lane l of
// tile t is at tile_mat_off[t] + (row*E + e)*64 + l (E = m*m), vectors likewise with E = m.
// Work items: a whole tile (one lane per system) or, for the cooperative placement, a
// chunk of floor(64/m) systems of a tile with m lanes per system. Items are ordered by
// decreasing estimated cost (longest-processing-time first) and consumed through an
// atomic counter by a grid sized to the device; every wave exits when the counter passes
// the number of items.
#pragma once

#include "btd_kernels.h"

namespace btd {


// Plain block Thomas for one lane of a tile (rows 0..n-1), blocks in VGPRs.
template <int M>
__device__ __forceinline__ bool plain_lane(const double* __restrict__ A, const double* __restrict__ B,
                                           double* C, double* D, int64_t mo, int64_t vo, int n) {
    constexpr int E = M * M;
    const int64_t es = kTile;
    double Cp[M][M], dp[M];
    bool ok = true;
    for (int q = 0; q < n; ++q) {
        const int64_t rm = mo + (int64_t)q * E * kTile, rv = vo + (int64_t)q * M * kTile;
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
        const int64_t rm = mo + (int64_t)q * E * kTile, rv = vo + (int64_t)q * M * kTile;
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
    const int64_t es = kTile;
    const int lane = threadIdx.x;
#define BTD_R(i, j) sR[((i) + M * (j)) * kTile + lane]
#define BTD_CP(i, j) sC[((i) + M * (j)) * kTile + lane]
#define BTD_P(i) sP[(i) * kTile + lane]
    double dp[M];
    bool ok = true;
    for (int q = 0; q < n; ++q) {
        const int64_t rm = mo + (int64_t)q * E * kTile, rv = vo + (int64_t)q * M * kTile;
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
        const int64_t rm = mo + (int64_t)q * E * kTile, rv = vo + (int64_t)q * M * kTile;
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

// Orders the LDS accesses of the lanes of one wavefront. The batch kernel runs one-wave
// (64-lane) workgroups, so LDS-only acquire/release fences around a wave barrier suffice and
// outstanding global loads (the prefetched next row) are not waited for, unlike with a full
// workgroup barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// 1/x from the hardware reciprocal refined by two Newton steps (error ~1 ulp); shorter on
// the pivot's critical path than an IEEE division.
__device__ __forceinline__ double fast_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Cooperative solve: m lanes per system, lane i owns block row i of every block. LDS
// (sh, [slot][lane], 5m+5 slots) holds the carried C'_{q-1} (column j in lane base+j) and
// d'_{q-1}, and two alternating exchange buffers, so every pivot step needs one barrier:
// each lane posts its candidate |r(k)| with its whole augmented row, every lane picks the
// pivot (largest |r(k)| among rows not yet chosen, lowest lane on a tie) and eliminates
// with it; the buffer written at step k is rewritten only at step k+2, after the barrier
// of step k+1 that every lane passes once done reading. After the elimination each lane
// posts its row once more and back-substitutes one right-hand-side column on its own
// (lane i: column i of C'_q; every lane: d'_q). The next row's blocks are loaded while the
// current row is eliminated. All lanes of the wave run the same number of row steps (nmax),
// so every barrier is reached by the whole workgroup; lanes past their system's n, and
// lanes of no system (n = 0), compute on zeros and store nothing.
template <int M>
__device__ bool coop_rows(const double* __restrict__ A, const double* __restrict__ B, double* C,
                          double* D, int64_t mo, int64_t vo, int i, int base, int n, int nmax,
                          double* __restrict__ sh) {
    constexpr int E = M * M;
    constexpr int W = 2 * M + 2;          // exchange slots: row (M), rhs (M+1), candidate (1)
    double* carry = sh;                   // slots [0, M]: C'(k, lane-base) and d'(lane-base)
    double* xbuf0 = sh + (M + 1) * kTile; // two exchange buffers of W slots
    double* xbuf1 = xbuf0 + W * kTile;
    const int lane = threadIdx.x;
    const int64_t es = kTile;
    bool ok = true;
    double pa[M], pb[M], pc[M], pd;
    auto fetch = [&](int q) {
        const bool act = q < n, has_next = q + 1 < n;
        const int64_t rm = mo + (int64_t)q * E * kTile, rv = vo + (int64_t)q * M * kTile;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            pa[k] = (act && q > 0) ? __builtin_nontemporal_load(A + rm + (i + M * k) * es) : 0.0;
            pb[k] = act ? __builtin_nontemporal_load(B + rm + (i + M * k) * es) : 0.0;
            pc[k] = has_next ? C[rm + (i + M * k) * es] : 0.0;
        }
        pd = act ? D[rv + i * es] : 0.0;
    };
    fetch(0);
    double xlast = 0.0;
    for (int q = 0; q < nmax; ++q) {
        const bool act = q < n;
        const bool has_next = q + 1 < n;
        const int64_t rm = mo + (int64_t)q * E * kTile, rv = vo + (int64_t)q * M * kTile;
        double r[M], rhs[M + 1];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            r[j] = pb[j];
            rhs[1 + j] = pc[j];
        }
        rhs[0] = pd;
        if (q > 0) {
#pragma unroll
            for (int k = 0; k < M; ++k) {
                rhs[0] = fma(-pa[k], carry[M * kTile + base + k], rhs[0]);
#pragma unroll
                for (int j = 0; j < M; ++j) r[j] = fma(-pa[k], carry[k * kTile + base + j], r[j]);
            }
        }
        if (q + 1 < nmax) fetch(q + 1);
        int rank = -1;
        int plane[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            double* xb = (k & 1) ? xbuf1 : xbuf0;
#pragma unroll
            for (int j = 0; j < M; ++j) xb[j * kTile + lane] = r[j];
#pragma unroll
            for (int j = 0; j <= M; ++j) xb[(M + j) * kTile + lane] = rhs[j];
            xb[(2 * M + 1) * kTile + lane] = rank < 0 ? fabs(r[k]) : -1.0;
            wave_lds_sync();
            int p = 0;
            double best = -2.0;
#pragma unroll
            for (int t = 0; t < M; ++t) {
                const double v = xb[(2 * M + 1) * kTile + base + t];
                const bool c = v > best;
                best = c ? v : best;
                p = c ? t : p;
            }
            plane[k] = p;
            const int pl = base + p;
            const double pk = xb[k * kTile + pl];
            ok = ok && (!act || pk != 0.0);
            if (i == p) {
                rank = k;
            } else if (rank < 0) {
                const double l = r[k] * fast_rcp(pk);
#pragma unroll
                for (int j = k + 1; j < M; ++j) r[j] = fma(-l, xb[j * kTile + pl], r[j]);
#pragma unroll
                for (int j = 0; j <= M; ++j) rhs[j] = fma(-l, xb[(M + j) * kTile + pl], rhs[j]);
            }
        }
        // post the eliminated rows; lane i solves column i+1 (C'(:, i)), every lane column 0
        double* xb = (M & 1) ? xbuf1 : xbuf0;
#pragma unroll
        for (int j = 0; j < M; ++j) xb[j * kTile + lane] = r[j];
#pragma unroll
        for (int j = 0; j <= M; ++j) xb[(M + j) * kTile + lane] = rhs[j];
        wave_lds_sync();
        double xc[M], xd[M];
#pragma unroll
        for (int k = M - 1; k >= 0; --k) {
            const int pl = base + plane[k];
            double sc = 0.0;
#pragma unroll
            for (int t = 0; t < M; ++t) sc = (i == t) ? xb[(M + 1 + t) * kTile + pl] : sc;
            double sd = xb[M * kTile + pl];
#pragma unroll
            for (int j = k + 1; j < M; ++j) {
                const double u = xb[j * kTile + pl];
                sc = fma(-u, xc[j], sc);
                sd = fma(-u, xd[j], sd);
            }
            const double rk = fast_rcp(xb[k * kTile + pl]);
            xc[k] = sc * rk;
            xd[k] = sd * rk;
        }
        double di = xd[0];
#pragma unroll
        for (int t = 1; t < M; ++t) di = (i == t) ? xd[t] : di;
        if (act) {
            D[rv + i * es] = di;
            xlast = di;
        }
        if (has_next) {
#pragma unroll
            for (int k = 0; k < M; ++k) C[rm + (k + M * i) * es] = xc[k];
        }
#pragma unroll
        for (int k = 0; k < M; ++k) carry[k * kTile + lane] = xc[k];  // C'(k, i) in lane base+i
        carry[M * kTile + lane] = di;                                   // d'(i) in lane base+i
        wave_lds_sync();
    }
    // back substitution over the rows: x_q(i) = d'_q(i) - sum_j C'_q(i,j) x_{q+1}(j); the
    // next row's C'(i,:) and d'(i) are loaded while the current row is finished.
    double xnext = xlast;
    double cn[M], dn = 0.0;
    auto bfetch = [&](int q) {
        const bool act = q >= 0 && q < n - 1;
        const int64_t rm = mo + (int64_t)(act ? q : 0) * E * kTile, rv = vo + (int64_t)(act ? q : 0) * M * kTile;
#pragma unroll
        for (int j = 0; j < M; ++j) cn[j] = act ? C[rm + (i + M * j) * es] : 0.0;
        dn = act ? D[rv + i * es] : 0.0;
    };
    bfetch(nmax - 2);
    for (int q = nmax - 2; q >= 0; --q) {
        const bool act = q < n - 1;
        double cq[M], x = dn;
#pragma unroll
        for (int j = 0; j < M; ++j) cq[j] = cn[j];
        bfetch(q - 1);
        sh[lane] = xnext;
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < M; ++j) x = fma(-cq[j], sh[base + j], x);
        if (act) {
            D[vo + (int64_t)q * M * kTile + i * es] = x;
            xnext = x;
        }
        wave_lds_sync();
    }
    return ok;
}

template <int M>
__device__ __forceinline__ bool coop_item(const BatchDev& b, const double* A, const double* B,
                                          double* C, double* D, int tile, int lane0, double* sh) {
    constexpr int G = kTile / M;  // systems per wave
    const int lane = threadIdx.x;
    const int g = lane / M, i = lane % M;
    const int slot = lane0 + g;
    const int count = b.tile_count[tile];
    const bool mine = g < G && slot < count;
    const int n = mine ? b.lane_rows[tile * kTile + slot] : 0;
    const int nmax = b.lane_rows[tile * kTile + lane0];  // systems are sorted longest first
    const int64_t mo = b.mat_off[tile] + (mine ? slot : 0), vo = b.vec_off[tile] + (mine ? slot : 0);
    return coop_rows<M>(A, B, C, D, mo, vo, mine ? i : 0, mine ? g * M : lane, n, nmax, sh);
}

// Persistent solve of a packed batch. KIND selects the code a placement needs, so that the
// register allocation of the kernel is that of the paths it can run: 0 one lane per system
// for every m (VGPRs); 1 one lane per system, m >= 6 with blocks in LDS (105 KB of LDS);
// 2 one lane per system for m <= 5 and m lanes per system (cooperative items).
template <int KIND>
__global__ __launch_bounds__(64) void k_batch(BatchDev b, const double* __restrict__ A,
                                              const double* __restrict__ B, double* C, double* D,
                                              int* __restrict__ counter, int* __restrict__ flag) {
    bool ok = true;
    const int lane = threadIdx.x;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __shfl(t, 0);
        if (t >= b.nitems) break;
        const int tile = b.item_tile[t];
        const int mode = b.item_mode[t];
        const int m = b.tile_m[tile];
        if constexpr (KIND == 2) {
            if (mode == 2) {
                __shared__ double sh[(5 * kMaxM + 5) * kTile];
                const int l0 = b.item_lane0[t];
                switch (m) {
                    case 2: ok = coop_item<2>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    case 3: ok = coop_item<3>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    case 4: ok = coop_item<4>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    case 5: ok = coop_item<5>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    case 6: ok = coop_item<6>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    case 7: ok = coop_item<7>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    case 8: ok = coop_item<8>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    case 9: ok = coop_item<9>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    case 10: ok = coop_item<10>(b, A, B, C, D, tile, l0, sh) && ok; break;
                    default: break;
                }
                continue;
            }
        }
        if (lane < b.tile_count[tile]) {
            const int n = b.lane_rows[tile * kTile + lane];
            const int64_t mo = b.mat_off[tile] + lane, vo = b.vec_off[tile] + lane;
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

// Generator for a packed batch: one thread per (tile, row, lane).
template <int M>
__device__ __forceinline__ void gen_batch_elem(int kind, uint64_t seed, int64_t sys, int row, int n,
                                               int64_t mo, int64_t vo, double* A, double* B,
                                               double* C, double* D) {
    double a[M][M], bb[M][M], c[M][M], d[M];
    gen_row<M>(kind, seed, sys, row, n, a, bb, c, d);
    const int64_t rm = mo + (int64_t)row * M * M * kTile, rv = vo + (int64_t)row * M * kTile;
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            A[rm + (i + M * j) * kTile] = a[i][j];
            B[rm + (i + M * j) * kTile] = bb[i][j];
            C[rm + (i + M * j) * kTile] = c[i][j];
        }
    for (int i = 0; i < M; ++i) D[rv + i * kTile] = d[i];
}

__global__ __launch_bounds__(256) void k_batch_generate(BatchDev b, const int* __restrict__ lane_sys,
                                                        int kind, uint64_t seed, double* A, double* B,
                                                        double* C, double* D) {
    const int tile = blockIdx.y;
    const int row = blockIdx.x * 4 + threadIdx.x / kTile, lane = threadIdx.x % kTile;
    if (lane >= b.tile_count[tile]) return;
    const int n = b.lane_rows[tile * kTile + lane];
    if (row >= n) return;
    const int64_t sys = lane_sys[tile * kTile + lane];
    const int64_t mo = b.mat_off[tile] + lane, vo = b.vec_off[tile] + lane;
    switch (b.tile_m[tile]) {
        case 1: gen_batch_elem<1>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 2: gen_batch_elem<2>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 3: gen_batch_elem<3>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 4: gen_batch_elem<4>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 5: gen_batch_elem<5>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 6: gen_batch_elem<6>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 7: gen_batch_elem<7>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 8: gen_batch_elem<8>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 9: gen_batch_elem<9>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        case 10: gen_batch_elem<10>(kind, seed, sys, row, n, mo, vo, A, B, C, D); break;
        default: break;
    }
}

}  // namespace btd
