// Cooperative block-Thomas rows: a system's block rows spread over m lanes of a wavefront
// (lane i owns row i of every block; from kColFromM, lane j owns column j), pivot search and
// row exchange through LDS. Used by the
// mixed-size batch (btd_batch.h), the single-partition cooperative and twisted solves, stage 1
// of the two-stage solve for m >= 7 and the twisted solve of its reduced system (btd_kernels.h). Storage: element e (column-major within the block) of row q of a lane's
// system is at am.at(q, e) for blocks and av.at(q, e) for vectors; StridedAddr gives the
// pack and tiled forms with compile-time strides, the plans' Addr the reference GPU layout.
#pragma once

#include "btd_common.h"

namespace btd {

// Element e of row q at base + q*RS + e*ES, strides fixed at compile time.
template <int64_t RS, int64_t ES>
struct StridedAddr {
    int64_t base;
    __device__ __forceinline__ int64_t at(int64_t q, int e) const { return base + q * RS + (int64_t)e * ES; }
};

// Orders the LDS accesses of the lanes of one wavefront. In the cooperative kernels every
// wavefront exchanges data only among its own lanes, through its own LDS region, and the LDS
// performs the operations of a wavefront in the order they are issued, so wavefront-scope
// fences (which keep the compiler from reordering the accesses and emit no wait) suffice: an
// exchange costs one LDS round trip, not a drain of every outstanding load.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

// Rows of the forward (kFwdAhead) and backward (kBwdAhead) sweeps of the cooperative solve
// kept in flight: one wavefront per work item has no other waves to hide HBM latency
// behind, so the next rows' blocks are loaded while the current row is eliminated.
template <int M>
struct CoopDepth {
    static constexpr int kFwdAhead = M <= 2 ? 8 : M <= 3 ? 6 : M <= 5 ? 4 : 1;
    static constexpr int kBwdAhead = M <= 3 ? 16 : M <= 5 ? 12 : 8;
};

// LDS of a cooperative work item, [slot][lane] (64 doubles a slot): slots [0, M) the carried
// block of each group (slot k, lane base+j: element (k, j)), slot M its carried right-hand
// side (lane base+i: element i), then the 2M+2 slots of the elimination exchange.
// Sized per m where the kernel is (a cooperative workgroup of m = 5 needs 18 slots, not 33).
template <int M>
constexpr int kCoopSlotsFor = 3 * M + 3;
constexpr int kCoopSlots = kCoopSlotsFor<kMaxM>;  // any m (the mixed-size batch kernel)

// Elimination of one block row by a group of M lanes: lane i holds row i of R and of the
// right-hand sides rhs = (d, S), S of NS blocks of M columns whose R^{-1} S is carried on.
// Pivot step k: every lane posts its row's columns k..M-1, its rhs and its candidate |r(k)|
// (-1 once chosen); every lane reads the M candidates, picks the pivot (largest, lowest lane
// on a tie), reads the pivot row and eliminates with it, a lane that is already chosen or is
// the pivot with multiplier 0. Then each lane posts its row once more and back-substitutes
// column i of every block of S on its own: xc[c*M + k] = (R^{-1} S)(k, c*M + i); and, in every
// lane, the whole xd = R^{-1} d. Returns 1 if a pivot was zero. The code is free of
// lane-dependent branches so that LDS reads issue in batches. LDS: xb holds (NS+1)*M + 2
// slots: the row [0, M), d at M, S at [M+1, M+1+NS*M), the candidate after them.
//
// Up to kGatherMaxM the same result comes from coop_eliminate_gather: one exchange instead of
// M + 1 (each lane reads the group's whole block once and repeats the elimination in its
// registers, bit for bit the operations of the exchanged form; outputs compared equal on every
// solve form). Measured against the exchanged form: the latency-bound small batches 7-25% faster
// (the reference Study's 1024 x 128 case at m = 2 / 3 / 5: 0.061 / 0.090 / 0.157 -> 0.051 /
// 0.068 / 0.146 ms; 1024 x 64 at m = 3: 0.050 -> 0.040 ms; twisted m = 5 16384 x 2048 13.0 ->
// 12.7 ms), the mixed-size batch unchanged; at m = 8, where a lane then holds 80 more doubles,
// 7-15% slower (cooperative 16384 x 1024 12.8 -> 14.5 ms, Study m = 8 0.31 -> 0.36 ms), so m >= 7
// keep the exchanged form.
#ifndef BTD_COOP_GATHER_MAX_M
#define BTD_COOP_GATHER_MAX_M 6
#endif
constexpr int kGatherMaxM = BTD_COOP_GATHER_MAX_M;

template <int M, int NS>
__device__ __forceinline__ int coop_eliminate_gather(const double (&r)[M], const double (&rhs)[NS * M + 1],
                                                     double* __restrict__ xb, int i, int base,
                                                     double (&xc)[NS * M], double (&xd)[M]);

template <int M, int NS = 1>
__device__ __forceinline__ int coop_eliminate(double (&r)[M], double (&rhs)[NS * M + 1],
                                              double* __restrict__ xb, int i, int base,
                                              double (&xc)[NS * M], double (&xd)[M]) {
    if constexpr (M <= kGatherMaxM) return coop_eliminate_gather<M, NS>(r, rhs, xb, i, base, xc, xd);
    constexpr int NR = NS * M + 1;     // right-hand-side entries of a row
    constexpr int CAND = M + NR;       // the candidate's slot
    const int lane = threadIdx.x & (kTile - 1);  // the lane in its wavefront
    int rank = -1;
    int plane[M];
    int zero_pivot = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
        for (int j = k; j < M; ++j) xb[j * kTile + lane] = r[j];
#pragma unroll
        for (int j = 0; j < NR; ++j) xb[(M + j) * kTile + lane] = rhs[j];
        xb[CAND * kTile + lane] = rank < 0 ? fabs(r[k]) : -1.0;
        wave_lds_sync();
        double cand[M];
#pragma unroll
        for (int t = 0; t < M; ++t) cand[t] = xb[CAND * kTile + base + t];
        int p = 0;
        double best = cand[0];
#pragma unroll
        for (int t = 1; t < M; ++t) {
            const bool c = cand[t] > best;
            best = c ? cand[t] : best;
            p = c ? t : p;
        }
        plane[k] = p;
        const int pl = base + p;
        double prow[M], prhs[NR];
#pragma unroll
        for (int j = k; j < M; ++j) prow[j] = xb[j * kTile + pl];
#pragma unroll
        for (int j = 0; j < NR; ++j) prhs[j] = xb[(M + j) * kTile + pl];
        zero_pivot |= prow[k] == 0.0 ? 1 : 0;
        const bool elim = rank < 0 && i != p;
        const double l = elim ? r[k] * fast_rcp(prow[k]) : 0.0;
        rank = (rank < 0 && i == p) ? k : rank;
#pragma unroll
        for (int j = k + 1; j < M; ++j) r[j] = fma(-l, prow[j], r[j]);
#pragma unroll
        for (int j = 0; j < NR; ++j) rhs[j] = fma(-l, prhs[j], rhs[j]);
    }
#pragma unroll
    for (int j = 0; j < M; ++j) xb[j * kTile + lane] = r[j];
#pragma unroll
    for (int j = 0; j < NR; ++j) xb[(M + j) * kTile + lane] = rhs[j];
    wave_lds_sync();
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
        const int pl = base + plane[k];
        double sc[NS];
#pragma unroll
        for (int c = 0; c < NS; ++c) sc[c] = xb[(M + 1 + c * M + i) * kTile + pl];
        double sd = xb[M * kTile + pl];
#pragma unroll
        for (int j = k + 1; j < M; ++j) {
            const double u = xb[j * kTile + pl];
#pragma unroll
            for (int c = 0; c < NS; ++c) sc[c] = fma(-u, xc[c * M + j], sc[c]);
            sd = fma(-u, xd[j], sd);
        }
        const double rk = fast_rcp(xb[k * kTile + pl]);
#pragma unroll
        for (int c = 0; c < NS; ++c) xc[c * M + k] = sc[c] * rk;
        xd[k] = sd * rk;
    }
    return zero_pivot;
}

// coop_eliminate with one exchange: every lane posts its row of R and of the right-hand sides
// once; every lane then reads its group's M rows (R, the d column and, of each S block, column
// i) and eliminates them in registers exactly as the exchanged form does across lanes: at step k
// the pivot is the row not yet chosen with the largest |R(t, k)|, the lowest row on a tie; each
// row not yet chosen is updated with l = R(t, k) * fast_rcp(pivot), R(t, j) = fma(-l, R(p, j),
// R(t, j)) and the same on its right-hand sides; back substitution follows the pivots' order.
// The rows are kept in pivot order by branch-free exchanges (the original row numbers ride
// along for the tie rule), so every value is computed by the same operations on the same
// operands as in the exchanged form: the same bits (zeros aside, whose sign may differ).
template <int M, int NS>
__device__ __forceinline__ int coop_eliminate_gather(const double (&r)[M], const double (&rhs)[NS * M + 1],
                                                     double* __restrict__ xb, int i, int base,
                                                     double (&xc)[NS * M], double (&xd)[M]) {
    constexpr int NR = NS * M + 1;
    const int lane = threadIdx.x & (kTile - 1);  // the lane in its wavefront
#pragma unroll
    for (int j = 0; j < M; ++j) xb[j * kTile + lane] = r[j];
#pragma unroll
    for (int j = 0; j < NR; ++j) xb[(M + j) * kTile + lane] = rhs[j];
    wave_lds_sync();
    double R[M][M], X[M][NS + 1];  // row t: R(t, :), d(t), S(t, c*M + i) for each block c
    int org[M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
#pragma unroll
        for (int j = 0; j < M; ++j) R[t][j] = xb[j * kTile + base + t];
        X[t][0] = xb[M * kTile + base + t];
#pragma unroll
        for (int c = 0; c < NS; ++c) X[t][1 + c] = xb[(M + 1 + c * M + i) * kTile + base + t];
        org[t] = t;
    }
    wave_lds_sync();  // every read of this exchange before the next one's writes
    int zero_pivot = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        double best = fabs(R[k][k]);
        int p = k, po = org[k];
#pragma unroll
        for (int t = k + 1; t < M; ++t) {
            const double v = fabs(R[t][k]);
            const bool c = v > best || (v == best && org[t] < po);
            best = c ? v : best;
            p = c ? t : p;
            po = c ? org[t] : po;
        }
#pragma unroll
        for (int t = k + 1; t < M; ++t) {
            const bool c = p == t;
#pragma unroll
            for (int j = k; j < M; ++j) {
                const double u = R[k][j];
                R[k][j] = c ? R[t][j] : u;
                R[t][j] = c ? u : R[t][j];
            }
#pragma unroll
            for (int j = 0; j < NS + 1; ++j) {
                const double u = X[k][j];
                X[k][j] = c ? X[t][j] : u;
                X[t][j] = c ? u : X[t][j];
            }
            const int o = org[k];
            org[k] = c ? org[t] : o;
            org[t] = c ? o : org[t];
        }
        zero_pivot |= R[k][k] == 0.0 ? 1 : 0;
        const double rk = fast_rcp(R[k][k]);
#pragma unroll
        for (int t = k + 1; t < M; ++t) {
            const double l = R[t][k] * rk;
#pragma unroll
            for (int j = k + 1; j < M; ++j) R[t][j] = fma(-l, R[k][j], R[t][j]);
#pragma unroll
            for (int j = 0; j < NS + 1; ++j) X[t][j] = fma(-l, X[k][j], X[t][j]);
        }
    }
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
        double sc[NS];
#pragma unroll
        for (int c = 0; c < NS; ++c) sc[c] = X[k][1 + c];
        double sd = X[k][0];
#pragma unroll
        for (int j = k + 1; j < M; ++j) {
            const double u = R[k][j];
#pragma unroll
            for (int c = 0; c < NS; ++c) sc[c] = fma(-u, xc[c * M + j], sc[c]);
            sd = fma(-u, xd[j], sd);
        }
        const double rk = fast_rcp(R[k][k]);
#pragma unroll
        for (int c = 0; c < NS; ++c) xc[c * M + k] = sc[c] * rk;
        xd[k] = sd * rk;
    }
    return zero_pivot;
}

// Element i of a vector held whole in every lane, picked with bit masks: written as a chain
// of selects, the compiler turns it into an indexed load and keeps the whole vector in
// scratch memory (stored on every row step).
template <int M>
__device__ __forceinline__ double lane_elem(const double (&v)[M], int i) {
    unsigned long long bits = 0;
#pragma unroll
    for (int t = 0; t < M; ++t)
        bits |= (unsigned long long)__double_as_longlong(v[t]) & (0ull - (unsigned long long)(i == t));
    return __longlong_as_double((long long)bits);
}

// ---------------------------------------------------------------------------------------
// Column form of the cooperative row step (m >= kColFromM). Lane j of a group owns column j
// of R_q, of the block R_q is solved against (C_q; A_q in the bottom half of a twisted solve)
// and of the result C'_q; every lane carries the whole right-hand side. The carried column of
// C'_{q-1} then never leaves its lane: the only block exchanged is A_q (each lane posts its
// column, every lane reads all of it for R_q(:, j) = B_q(:, j) - A_q C'_{q-1}(:, j) and
// d_q - A_q d'_{q-1}), and the pivot search is local to the lane that owns the column. A row
// takes m + 2 LDS round trips (the row form: 2m + 2) and about half the LDS operations. Every
// value is computed by the same operations on the same operands as in the row form (the same
// multipliers r(t, k) * fast_rcp(pivot), the same fused updates and back substitution), so the
// results are the row form's bit for bit, except where two candidates tie exactly for a pivot:
// the row form then takes the lowest original row, this form the first row in dgetf2's swapped
// order. Measured against the row form (which waits on its exchanges most of its cycles at
// m >= 7: m = 8, 64% of wave cycles waiting, VALU busy 10%): the Study's m = 8 16384 x 2048
// case 34.2 -> 31.0 ms, m = 8 tiled 16384 x 1024 12.9 -> 12.0 ms, twisted m = 8 1024 x 2048
// 5.6 -> 5.0 ms and 4096 x 4096 25.4 -> 23.8 ms, m = 7 16384 x 1024 13.0 -> 11.4 ms, m = 10
// 16384 x 512 16.6 -> 15.9 ms, m = 9 8192 x 512 unchanged (12.2 / 12.4 ms). The mixed-size batch
// keeps the row form: its kernel also holds the one-lane blocks, and with the column form on
// top it spilled 1.3 KB per lane and ran 2x slower.
#ifndef BTD_COOP_COL_FROM_M
#define BTD_COOP_COL_FROM_M 7
#endif
constexpr int kColFromM = BTD_COOP_COL_FROM_M;

// Rows of the column form's forward sweep loaded ahead: one. Two took 256 VGPRs at m = 8 (one
// wavefront per SIMD instead of two) and measured slower there (the Study's m = 8 16384 x 2048
// case 31.0 -> 35.1 ms, twisted 4096 x 4096 23.8 -> 23.1 ms, m = 9, 10 within 4%).
template <int M>
struct ColDepth {
#ifdef BTD_COL_FWD_AHEAD
    static constexpr int kFwdAhead = BTD_COL_FWD_AHEAD;
#else
    static constexpr int kFwdAhead = 1;
#endif
};

// One block row in the column form. r: column j of R (destroyed); c: column j of the block
// solved against, replaced by column j of R^{-1} C (WITH_C); dv: the whole right-hand side,
// replaced by R^{-1} d. Pivot step k: lane k of the group (the owner of column k) picks the
// first largest |r| among rows k..M-1 (dgetf2's choice) and posts the pivot row index, the
// pivot's reciprocal (hardware reciprocal and two Newton steps) and the multipliers of the
// rows below it after the exchange of rows k and p; every lane exchanges rows k and p of its
// columns and eliminates. Back substitution: every lane posts its column of U and solves its
// own column and the whole d. xb: the group's exchange, [slot][lane], M + 2 slots. Returns 1
// if this lane's column met a zero pivot (lane k of a group sees the pivot of step k).
template <int M, bool WITH_C>
__device__ __forceinline__ int coop_eliminate_col(double (&r)[M], double (&c)[M], double (&dv)[M],
                                                  double* __restrict__ xb, int j, int base) {
    const int lane = threadIdx.x & (kTile - 1);  // the lane in its wavefront
    int zero_pivot = 0;
    double rk[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        double best = fabs(r[k]), pv = r[k];
        int p = k;
#pragma unroll
        for (int t = k + 1; t < M; ++t) {
            const double v = fabs(r[t]);
            const bool b = v > best;
            best = b ? v : best;
            pv = b ? r[t] : pv;
            p = b ? t : p;
        }
        zero_pivot |= (j == k && pv == 0.0) ? 1 : 0;
        const double rp = fast_rcp(pv);
#pragma unroll
        for (int t = k + 1; t < M; ++t) xb[t * kTile + lane] = (p == t ? r[k] : r[t]) * rp;
        xb[M * kTile + lane] = (double)p;
        xb[(M + 1) * kTile + lane] = rp;
        wave_lds_sync();
        const int src = base + k;
        const int pk = (int)xb[M * kTile + src];
        rk[k] = xb[(M + 1) * kTile + src];
        double l[M];
#pragma unroll
        for (int t = k + 1; t < M; ++t) l[t] = xb[t * kTile + src];
#pragma unroll
        for (int t = k + 1; t < M; ++t) {
            const bool s = pk == t;
            const double u = r[k], w = dv[k];
            r[k] = s ? r[t] : u;
            r[t] = s ? u : r[t];
            dv[k] = s ? dv[t] : w;
            dv[t] = s ? w : dv[t];
            if constexpr (WITH_C) {
                const double z = c[k];
                c[k] = s ? c[t] : z;
                c[t] = s ? z : c[t];
            }
        }
#pragma unroll
        for (int t = k + 1; t < M; ++t) {
            r[t] = fma(-l[t], r[k], r[t]);
            dv[t] = fma(-l[t], dv[k], dv[t]);
            if constexpr (WITH_C) c[t] = fma(-l[t], c[k], c[t]);
        }
    }
    // U(k, j') = r[k] of lane j' (j' > k)
#pragma unroll
    for (int t = 0; t < M - 1; ++t) xb[t * kTile + lane] = r[t];
    wave_lds_sync();
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {  // in place: c[t], dv[t] (t > k) already solved
        double sc = c[k], sd = dv[k];
#pragma unroll
        for (int t = k + 1; t < M; ++t) {
            const double u = xb[k * kTile + base + t];
            if constexpr (WITH_C) sc = fma(-u, c[t], sc);
            sd = fma(-u, dv[t], sd);
        }
        if constexpr (WITH_C) c[k] = sc * rk[k];
        dv[k] = sd * rk[k];
    }
    return zero_pivot;
}

// Forward sweep of coop_rows in the column form (lane j owns column j); the arguments are
// coop_rows'. d'_q(j) goes over d_q and C'_q(:, j) over C_q(:, j) as in the row form, so the
// backward sweep is shared. xlast: this lane's element of its system's last d'.
template <int M, bool LOCK, class AM, class AV>
__device__ __forceinline__ bool coop_fwd_col(const double* __restrict__ A, const double* __restrict__ B,
                                             double* C, double* D, const AM& am, const AV& av, int j,
                                             int base, int n, int nmax, int rows, double* __restrict__ sh,
                                             double& xlast) {
    constexpr int F = ColDepth<M>::kFwdAhead;
    const int lane = threadIdx.x & (kTile - 1);
    bool ok = true;
    double pa[F][M], pb[F][M], pc[F][M], pd[F];
    auto fetch = [&](double (&fa)[M], double (&fb)[M], double (&fc)[M], double& fd, int q) {
        q = q < rows ? q : rows - 1;
#pragma unroll
        for (int t = 0; t < M; ++t) {
            fa[t] = __builtin_nontemporal_load(A + am.at(q, t + M * j));
            fb[t] = __builtin_nontemporal_load(B + am.at(q, t + M * j));
            fc[t] = C[am.at(q, t + M * j)];
        }
        fd = D[av.at(q, j)];
    };
#pragma unroll
    for (int s = 0; s < F; ++s) fetch(pa[s], pb[s], pc[s], pd[s], s);
    double cc[M], dp[M];  // column j of C'_{q-1}; d'_{q-1}, whole
#pragma unroll
    for (int t = 0; t < M; ++t) {
        cc[t] = 0.0;
        dp[t] = 0.0;
    }
    for (int q0 = 0; q0 < nmax; q0 += F) {
#pragma unroll
        for (int s = 0; s < F; ++s) {
            const int q = q0 + s;
            if (q < nmax) {  // uniform: the last chunk may be partial
                const bool act = q < n;
                const bool has_next = q + 1 < n;
#pragma unroll
                for (int t = 0; t < M; ++t) sh[t * kTile + lane] = pa[s][t];  // A_q(t, j)
                sh[M * kTile + lane] = pd[s];                                   // d_q(j)
                wave_lds_sync();
                double r[M], c[M], dv[M];
#pragma unroll
                for (int t = 0; t < M; ++t) {
                    r[t] = pb[s][t];
                    c[t] = pc[s][t];
                    dv[t] = sh[M * kTile + base + t];
                }
                if (q > 0) {
#pragma unroll
                    for (int k = 0; k < M; ++k)
#pragma unroll
                        for (int t = 0; t < M; ++t) {
                            const double a = sh[t * kTile + base + k];
                            r[t] = fma(-a, cc[k], r[t]);
                            dv[t] = fma(-a, dp[k], dv[t]);
                        }
                }
                wave_lds_sync();
                if constexpr (LOCK) __builtin_amdgcn_s_barrier();
                fetch(pa[s], pb[s], pc[s], pd[s], q + F);
                const int zero_pivot = coop_eliminate_col<M, true>(r, c, dv, sh, j, base);
                ok = ok && !(act && zero_pivot);
                if (act) {
                    const double dj = lane_elem<M>(dv, j);
                    D[av.at(q, j)] = dj;
                    xlast = dj;
                }
                if (has_next) {
#pragma unroll
                    for (int t = 0; t < M; ++t) C[am.at(q, t + M * j)] = c[t];
                }
#pragma unroll
                for (int t = 0; t < M; ++t) {
                    cc[t] = c[t];
                    dp[t] = dv[t];
                }
                wave_lds_sync();
            }
        }
    }
    return ok;
}

// Both halves of coop_twisted up to and including the meeting row, in the column form (lane j
// owns column j; sub/sup as in coop_twisted). Stores d'_q(j) and the solved block's column j
// for every row a group eliminates, and x_h(j) from the top group; xh: the whole x_h.
template <int M, bool UNIT_B, bool LOCK, class AM, class AV>
__device__ __forceinline__ bool coop_twisted_fwd_col(const double* A, const double* __restrict__ B, double* C,
                                                     double* D, const AM& am, const AV& av, int j, int base,
                                                     int pbase, bool top, int n, int h, int cnt, int steps,
                                                     int rows, double* __restrict__ sh, double (&xh)[M]) {
    constexpr int F = ColDepth<M>::kFwdAhead;
    const int lane = threadIdx.x & (kTile - 1);
    const double* sub = top ? A : C;
    const double* sup = top ? C : A;
    auto clamp_row = [&](int q) { return q < 0 ? 0 : (q >= rows ? rows - 1 : q); };
    bool ok = true;
    double ps[F][M], pg[F][M], pu[F][M], pd[F];
    auto fetch = [&](double (&fs)[M], double (&fg)[M], double (&fu)[M], double& fd, int t) {
        const int q = clamp_row(top ? t : n - 1 - t);
#pragma unroll
        for (int k = 0; k < M; ++k) {
            fs[k] = sub[am.at(q, k + M * j)];
            if constexpr (UNIT_B)
                fg[k] = j == k ? 1.0 : 0.0;
            else
                fg[k] = __builtin_nontemporal_load(B + am.at(q, k + M * j));
            fu[k] = sup[am.at(q, k + M * j)];
        }
        fd = D[av.at(q, j)];
    };
#pragma unroll
    for (int s = 0; s < F; ++s) fetch(ps[s], pg[s], pu[s], pd[s], s);
    double cc[M], dp[M];  // column j of the carried block; the carried d, whole
#pragma unroll
    for (int t = 0; t < M; ++t) {
        cc[t] = 0.0;
        dp[t] = 0.0;
    }
    for (int t0 = 0; t0 < steps; t0 += F) {
#pragma unroll
        for (int s = 0; s < F; ++s) {
            const int t = t0 + s;
            if (t < steps) {  // uniform: the last chunk may be partial
                const bool act = t < cnt;
                const int q = clamp_row(top ? t : n - 1 - t);
#pragma unroll
                for (int k = 0; k < M; ++k) sh[k * kTile + lane] = ps[s][k];
                sh[M * kTile + lane] = pd[s];
                wave_lds_sync();
                double r[M], c[M], dv[M];
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    r[k] = pg[s][k];
                    c[k] = pu[s][k];
                    dv[k] = sh[M * kTile + base + k];
                }
                if (t > 0) {
#pragma unroll
                    for (int k = 0; k < M; ++k)
#pragma unroll
                        for (int u = 0; u < M; ++u) {
                            const double a = sh[u * kTile + base + k];
                            r[u] = fma(-a, cc[k], r[u]);
                            dv[u] = fma(-a, dp[k], dv[u]);
                        }
                }
                wave_lds_sync();
                if constexpr (LOCK) __builtin_amdgcn_s_barrier();
                fetch(ps[s], pg[s], pu[s], pd[s], t + F);
                const int zero_pivot = coop_eliminate_col<M, true>(r, c, dv, sh, j, base);
                ok = ok && !(act && zero_pivot);
                if (act) {
                    D[av.at(q, j)] = lane_elem<M>(dv, j);
#pragma unroll
                    for (int k = 0; k < M; ++k) C[am.at(q, k + M * j)] = c[k];
                }
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    cc[k] = act ? c[k] : cc[k];
                    dp[k] = act ? dv[k] : dp[k];
                }
                wave_lds_sync();
            }
        }
    }
    // meeting row h: (B_h - A_h C'_{h-1} - C_h A''_{h+1}) x_h = d_h - A_h d'_{h-1} - C_h d''_{h+1},
    // solved by both groups of a system from the same operands in the same order
    const int q = h;
    double r[M], dv[M], dummy[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        sh[k * kTile + lane] = cc[k];                                                   // carry column j
        sh[(M + 1 + k) * kTile + lane] = h > 0 ? A[am.at(q, k + M * j)] : 0.0;          // A_h(k, j)
        sh[(2 * M + 1 + k) * kTile + lane] = C[am.at(q, k + M * j)];                     // C_h(k, j)
        if constexpr (UNIT_B)
            r[k] = j == k ? 1.0 : 0.0;
        else
            r[k] = B[am.at(q, k + M * j)];
        dummy[k] = 0.0;
    }
    sh[M * kTile + lane] = lane_elem<M>(dp, j);       // element j of the carried d
    sh[(3 * M + 1) * kTile + lane] = D[av.at(q, j)];  // d_h(j)
    wave_lds_sync();
    const int tb = top ? base : pbase, bb = top ? pbase : base;
    double ct[M], cb[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        ct[k] = sh[k * kTile + tb + j];  // C'_{h-1}(k, j)
        cb[k] = sh[k * kTile + bb + j];  // A''_{h+1}(k, j)
        dv[k] = sh[(3 * M + 1) * kTile + base + k];
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double dt = sh[M * kTile + tb + k], db = sh[M * kTile + bb + k];
#pragma unroll
        for (int u = 0; u < M; ++u) {
            const double a = sh[(M + 1 + u) * kTile + base + k];      // A_h(u, k)
            const double cz = sh[(2 * M + 1 + u) * kTile + base + k];  // C_h(u, k)
            dv[u] = fma(-a, dt, dv[u]);
            dv[u] = fma(-cz, db, dv[u]);
            r[u] = fma(-a, ct[k], r[u]);
            r[u] = fma(-cz, cb[k], r[u]);
        }
    }
    wave_lds_sync();
    const int zero_pivot = coop_eliminate_col<M, false>(r, dummy, dv, sh, j, base);
    ok = ok && !(n > 0 && zero_pivot);
    if (n > 0 && top) D[av.at(q, j)] = lane_elem<M>(dv, j);
#pragma unroll
    for (int k = 0; k < M; ++k) xh[k] = dv[k];
    return ok;
}

// Cooperative solve: M lanes per system, lane i owns block row i of every block; row q is
// R_q = B_q - A_q C'_{q-1} eliminated with rhs (d_q - A_q d'_{q-1}, C_q) (coop_eliminate),
// C'_q stored over C_q and carried in LDS, d'_q over d_q and carried whole in registers;
// then x_q = d'_q - C'_q x_{q+1} upwards, x_{q+1} exchanged through LDS. Exchanges are
// ordered by wave_lds_sync only. Rows are loaded kFwdAhead (forward) and kBwdAhead
// (backward) rows ahead; every row below `rows` (the lane's pack's) is allocated for every
// slot of the pack, so loads are clamped to that range and need no other guard. All lanes of
// the wave run the same row steps (nmax, the item's longest system); lanes past their
// system's n compute on what the rows hold and store nothing.
// With LOCK each row step's loads are issued after a workgroup barrier (s_barrier, no memory
// wait), so the wavefronts of a workgroup, which must all run nmax steps, load each row together.
template <int M, bool LOCK = false, bool COL = (M >= kColFromM), class AM, class AV>
__device__ bool coop_rows(const double* __restrict__ A, const double* __restrict__ B, double* C,
                          double* D, const AM& am, const AV& av, int i, int base, int n, int nmax, int rows,
                          double* __restrict__ sh) {
    constexpr int G = CoopDepth<M>::kBwdAhead;
    const int lane = threadIdx.x & (kTile - 1);  // the lane in its wavefront
    bool ok = true;
    double xlast = 0.0;
    if constexpr (COL) {
        ok = coop_fwd_col<M, LOCK>(A, B, C, D, am, av, i, base, n, nmax, rows, sh, xlast);
    } else {
    constexpr int F = CoopDepth<M>::kFwdAhead;
    double* carry = sh;
    double* xb = sh + (M + 1) * kTile;
    double pa[F][M], pb[F][M], pc[F][M], pd[F];
    auto fetch = [&](double (&fa)[M], double (&fb)[M], double (&fc)[M], double& fd, int q) {
        q = q < rows ? q : rows - 1;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            fa[k] = __builtin_nontemporal_load(A + am.at(q, i + M * k));
            fb[k] = __builtin_nontemporal_load(B + am.at(q, i + M * k));
            fc[k] = C[am.at(q, i + M * k)];
        }
        fd = D[av.at(q, i)];
    };
#pragma unroll
    for (int s = 0; s < F; ++s) fetch(pa[s], pb[s], pc[s], pd[s], s);
    double dp[M];  // d'_{q-1}, whole, in every lane of the system
#pragma unroll
    for (int k = 0; k < M; ++k) dp[k] = 0.0;
    for (int q0 = 0; q0 < nmax; q0 += F) {
#pragma unroll
        for (int s = 0; s < F; ++s) {
            const int q = q0 + s;
            if (q < nmax) {  // uniform: the last chunk may be partial
                const bool act = q < n;
                const bool has_next = q + 1 < n;
                double r[M], rhs[M + 1];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    r[j] = pb[s][j];
                    rhs[1 + j] = pc[s][j];
                }
                rhs[0] = pd[s];
                if (q > 0) {
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        const double a = pa[s][k];
                        rhs[0] = fma(-a, dp[k], rhs[0]);
#pragma unroll
                        for (int j = 0; j < M; ++j) r[j] = fma(-a, carry[k * kTile + base + j], r[j]);
                    }
                }
                if constexpr (LOCK) __builtin_amdgcn_s_barrier();
                fetch(pa[s], pb[s], pc[s], pd[s], q + F);
                double xc[M], xd[M];
                const int zero_pivot = coop_eliminate<M>(r, rhs, xb, i, base, xc, xd);
                const double di = lane_elem<M>(xd, i);
                ok = ok && !(act && zero_pivot);
                if (act) {
                    D[av.at(q, i)] = di;
                    xlast = di;
                }
                if (has_next) {
#pragma unroll
                    for (int k = 0; k < M; ++k) C[am.at(q, k + M * i)] = xc[k];
                }
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    carry[k * kTile + lane] = xc[k];  // C'(k, i) in lane base+i
                    dp[k] = xd[k];
                }
                wave_lds_sync();
            }
        }
    }
    }
    // back substitution over the rows: x_q(i) = d'_q(i) - sum_j C'_q(i,j) x_{q+1}(j), with
    // C'_q(i, :) and d'_q(i) loaded kBwdAhead rows ahead (rows are clamped to the lane's pack).
    double xnext = xlast;
    double cn[G][M], dn[G];
    auto bfetch = [&](double (&fc)[M], double& fd, int q) {
        q = q < 0 ? 0 : (q < rows ? q : rows - 1);  // this lane's pack may be shorter than nmax
#pragma unroll
        for (int j = 0; j < M; ++j) fc[j] = C[am.at(q, i + M * j)];
        fd = D[av.at(q, i)];
    };
#pragma unroll
    for (int s = 0; s < G; ++s) bfetch(cn[s], dn[s], nmax - 2 - s);
    for (int q0 = nmax - 2; q0 >= 0; q0 -= G) {
#pragma unroll
        for (int s = 0; s < G; ++s) {
            const int q = q0 - s;
            if (q >= 0) {  // uniform: the last chunk may be partial
                const bool act = q >= 0 && q < n - 1;
                double cq[M], x = dn[s];
#pragma unroll
                for (int j = 0; j < M; ++j) cq[j] = cn[s][j];
                if constexpr (LOCK) __builtin_amdgcn_s_barrier();
                bfetch(cn[s], dn[s], q - G);
                sh[lane] = xnext;
                wave_lds_sync();
#pragma unroll
                for (int j = 0; j < M; ++j) x = fma(-cq[j], sh[base + j], x);
                if (act) {
                    D[av.at(q, i)] = x;
                    xnext = x;
                }
                wave_lds_sync();
            }
        }
    }
    return ok;
}

// Twisted cooperative solve: two groups of M lanes per system (lane i of each owns block row
// i). With h = (n-1)/2 the top group (`top`) eliminates rows 0..h-1 downwards as coop_rows
// does (C'_q = R_q^{-1} C_q over C_q), while the bottom group eliminates rows n-1..h+1
// upwards: S_q = B_q - C_q A''_{q+1}, A''_q = S_q^{-1} A_q stored over C_q (dead once row q
// is eliminated), d''_q = S_q^{-1}(d_q - C_q d''_{q+1}) over d_q. Both groups then solve the
// meeting row, (B_h - A_h C'_{h-1} - C_h A''_{h+1}) x_h = d_h - A_h d'_{h-1} - C_h d''_{h+1},
// and substitute outwards: x_q = d'_q - C'_q x_{q+1} (q < h), x_q = d''_q - A''_q x_{q-1}
// (q > h). Each group handles half of the rows, so a long system takes half the time of
// coop_rows for the same arithmetic. The groups of a system are base and pbase; `steps`
// (uniform over the wave) is the largest row count of a group of the item (the bottom
// group of its longest system); every row loaded is clamped below `rows`, the pack's row
// count. A group's carried block and rhs stay as its last row left them while the other
// groups finish. ES is the element stride of the storage; with UNIT_B every B_q is the
// identity and B is not read (the reduced system of the two-stage solve).
// LOCK: as in coop_rows (every wavefront of the workgroup runs the same `steps`).
template <int M, bool UNIT_B = false, bool LOCK = false, bool COL = (M >= kColFromM), class AM, class AV>
__device__ bool coop_twisted(const double* A, const double* __restrict__ B, double* C, double* D,
                             const AM& am, const AV& av, int i, int base, int pbase, bool top, int n,
                             int steps, int rows, double* __restrict__ sh) {
    constexpr int G = CoopDepth<M>::kBwdAhead;
    const int lane = threadIdx.x & (kTile - 1);  // the lane in its wavefront
    const int h = n > 0 ? (n - 1) / 2 : 0;
    const int cnt = n > 0 ? (top ? h : n - 1 - h) : 0;
    auto clamp_row = [&](int q) { return q < 0 ? 0 : (q >= rows ? rows - 1 : q); };
    bool ok = true;
    double xh[M];
    if constexpr (COL) {
        ok = coop_twisted_fwd_col<M, UNIT_B, LOCK>(A, B, C, D, am, av, i, base, pbase, top, n, h, cnt, steps,
                                                   rows, sh, xh);
    } else {
    constexpr int F = CoopDepth<M>::kFwdAhead;
    double* carry = sh;
    double* xb = sh + (M + 1) * kTile;
    const double* sub = top ? A : C;  // the block coupling row q to the row eliminated before it
    const double* sup = top ? C : A;  // the block coupling row q to the row eliminated after it
    double ps[F][M], pg[F][M], pu[F][M], pd[F];
    auto fetch = [&](double (&fs)[M], double (&fg)[M], double (&fu)[M], double& fd, int t) {
        const int q = clamp_row(top ? t : n - 1 - t);
#pragma unroll
        for (int k = 0; k < M; ++k) {
            fs[k] = sub[am.at(q, i + M * k)];
            if constexpr (UNIT_B)
                fg[k] = i == k ? 1.0 : 0.0;
            else
                fg[k] = __builtin_nontemporal_load(B + am.at(q, i + M * k));
            fu[k] = sup[am.at(q, i + M * k)];
        }
        fd = D[av.at(q, i)];
    };
#pragma unroll
    for (int s = 0; s < F; ++s) fetch(ps[s], pg[s], pu[s], pd[s], s);
    double dp[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        dp[k] = 0.0;
        carry[k * kTile + lane] = 0.0;
    }
    for (int t0 = 0; t0 < steps; t0 += F) {
#pragma unroll
        for (int s = 0; s < F; ++s) {
            const int t = t0 + s;
            if (t < steps) {  // uniform: the last chunk may be partial
                const bool act = t < cnt;
                const int q = clamp_row(top ? t : n - 1 - t);
                double r[M], rhs[M + 1];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    r[j] = pg[s][j];
                    rhs[1 + j] = pu[s][j];
                }
                rhs[0] = pd[s];
                if (t > 0) {
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        const double a = ps[s][k];
                        rhs[0] = fma(-a, dp[k], rhs[0]);
#pragma unroll
                        for (int j = 0; j < M; ++j) r[j] = fma(-a, carry[k * kTile + base + j], r[j]);
                    }
                }
                if constexpr (LOCK) __builtin_amdgcn_s_barrier();
                fetch(ps[s], pg[s], pu[s], pd[s], t + F);
                double xc[M], xd[M];
                const int zero_pivot = coop_eliminate<M>(r, rhs, xb, i, base, xc, xd);
                ok = ok && !(act && zero_pivot);
                if (act) {
                    D[av.at(q, i)] = lane_elem<M>(xd, i);
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        C[am.at(q, k + M * i)] = xc[k];
                        carry[k * kTile + lane] = xc[k];
                    }
                }
#pragma unroll
                for (int k = 0; k < M; ++k) dp[k] = act ? xd[k] : dp[k];
                wave_lds_sync();
            }
        }
    }
    // meeting row h: both groups of a system solve it from both carries
    carry[M * kTile + lane] = lane_elem<M>(dp, i);
    wave_lds_sync();
    const int tb = top ? base : pbase, bb = top ? pbase : base;
    {
        const int q = h;  // the meeting row
        double r[M], rhs[M + 1];
        double ah[M], ch[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            ah[k] = h > 0 ? A[am.at(q, i + M * k)] : 0.0;  // row 0 has no sub-diagonal block
            ch[k] = C[am.at(q, i + M * k)];
            if constexpr (UNIT_B)
                r[k] = i == k ? 1.0 : 0.0;
            else
                r[k] = B[am.at(q, i + M * k)];
            rhs[1 + k] = 0.0;
        }
        rhs[0] = D[av.at(q, i)];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            rhs[0] = fma(-ah[k], carry[M * kTile + tb + k], rhs[0]);
            rhs[0] = fma(-ch[k], carry[M * kTile + bb + k], rhs[0]);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                r[j] = fma(-ah[k], carry[k * kTile + tb + j], r[j]);
                r[j] = fma(-ch[k], carry[k * kTile + bb + j], r[j]);
            }
        }
        double xc[M];
        const int zero_pivot = coop_eliminate<M>(r, rhs, xb, i, base, xc, xh);
        ok = ok && !(n > 0 && zero_pivot);
        if (n > 0 && top) D[av.at(q, i)] = lane_elem<M>(xh, i);
    }
    }
    // outward substitution, kBwdAhead rows ahead
    double xprev = lane_elem<M>(xh, i);
    double cn[G][M], dn[G];
    auto bfetch = [&](double (&fc)[M], double& fd, int t) {
        const int q = clamp_row(top ? h - 1 - t : h + 1 + t);
#pragma unroll
        for (int j = 0; j < M; ++j) fc[j] = C[am.at(q, i + M * j)];
        fd = D[av.at(q, i)];
    };
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < G; ++s) bfetch(cn[s], dn[s], s);
    for (int t0 = 0; t0 < steps; t0 += G) {
#pragma unroll
        for (int s = 0; s < G; ++s) {
            const int t = t0 + s;
            if (t < steps) {  // uniform: the last chunk may be partial
                const bool act = t < cnt;
                const int q = clamp_row(top ? h - 1 - t : h + 1 + t);
                double cq[M], x = dn[s];
#pragma unroll
                for (int j = 0; j < M; ++j) cq[j] = cn[s][j];
                if constexpr (LOCK) __builtin_amdgcn_s_barrier();
                bfetch(cn[s], dn[s], t + G);
                sh[lane] = xprev;
                wave_lds_sync();
#pragma unroll
                for (int j = 0; j < M; ++j) x = fma(-cq[j], sh[base + j], x);
                if (act) {
                    D[av.at(q, i)] = x;
                    xprev = x;
                }
                wave_lds_sync();
            }
        }
    }
    return ok;
}

}  // namespace btd
