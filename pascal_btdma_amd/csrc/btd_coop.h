// Cooperative block-Thomas rows: a system's block rows spread over m lanes of a wavefront
// (lane i owns row i of every block), pivot search and row exchange through LDS. Used by the
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
template <int M, bool LOCK = false, class AM, class AV>
__device__ bool coop_rows(const double* __restrict__ A, const double* __restrict__ B, double* C,
                          double* D, const AM& am, const AV& av, int i, int base, int n, int nmax, int rows,
                          double* __restrict__ sh) {
    constexpr int F = CoopDepth<M>::kFwdAhead;
    constexpr int G = CoopDepth<M>::kBwdAhead;
    double* carry = sh;
    double* xb = sh + (M + 1) * kTile;
    const int lane = threadIdx.x & (kTile - 1);  // the lane in its wavefront
    bool ok = true;
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
    double xlast = 0.0;
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
template <int M, bool UNIT_B = false, bool LOCK = false, class AM, class AV>
__device__ bool coop_twisted(const double* A, const double* __restrict__ B, double* C, double* D,
                             const AM& am, const AV& av, int i, int base, int pbase, bool top, int n,
                             int steps, int rows, double* __restrict__ sh) {
    constexpr int F = CoopDepth<M>::kFwdAhead;
    constexpr int G = CoopDepth<M>::kBwdAhead;
    double* carry = sh;
    double* xb = sh + (M + 1) * kTile;
    const int lane = threadIdx.x & (kTile - 1);  // the lane in its wavefront
    const int h = n > 0 ? (n - 1) / 2 : 0;
    const int cnt = n > 0 ? (top ? h : n - 1 - h) : 0;
    const double* sub = top ? A : C;  // the block coupling row q to the row eliminated before it
    const double* sup = top ? C : A;  // the block coupling row q to the row eliminated after it
    auto clamp_row = [&](int q) { return q < 0 ? 0 : (q >= rows ? rows - 1 : q); };
    bool ok = true;
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
    double xh[M];
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
