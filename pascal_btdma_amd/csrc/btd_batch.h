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

// Orders the LDS accesses of the lanes of one wavefront. The batch kernel runs one-wave
// (64-lane) workgroups, and the LDS performs the operations of a wavefront in the order they
// are issued, so wavefront-scope fences (which keep the compiler from reordering the accesses
// and emit no wait) suffice: an exchange costs one LDS round trip, not a drain of every
// outstanding load.
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
constexpr int kCoopSlots = 3 * kMaxM + 3;

// Elimination of one block row by a group of M lanes: lane i holds row i of R and of the
// right-hand sides rhs = (d, S) (S the block whose R^{-1} S is carried to the next row).
// Pivot step k: every lane posts its row's columns k..M-1, its rhs and its candidate |r(k)|
// (-1 once chosen); every lane reads the M candidates, picks the pivot (largest, lowest lane
// on a tie), reads the pivot row and eliminates with it, a lane that is already chosen or is
// the pivot with multiplier 0. Then each lane posts its row once more and back-substitutes
// one rhs column on its own: xc = column i of R^{-1} S (xc[k] = (R^{-1} S)(k, i)) and, in
// every lane, the whole xd = R^{-1} d. Returns 1 if a pivot was zero. The code is free of
// lane-dependent branches so that LDS reads issue in batches.
template <int M>
__device__ __forceinline__ int coop_eliminate(double (&r)[M], double (&rhs)[M + 1], double* __restrict__ xb,
                                              int i, int base, double (&xc)[M], double (&xd)[M]) {
    const int lane = threadIdx.x;
    int rank = -1;
    int plane[M];
    int zero_pivot = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
        for (int j = k; j < M; ++j) xb[j * kTile + lane] = r[j];
#pragma unroll
        for (int j = 0; j <= M; ++j) xb[(M + j) * kTile + lane] = rhs[j];
        xb[(2 * M + 1) * kTile + lane] = rank < 0 ? fabs(r[k]) : -1.0;
        wave_lds_sync();
        double cand[M];
#pragma unroll
        for (int t = 0; t < M; ++t) cand[t] = xb[(2 * M + 1) * kTile + base + t];
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
        double prow[M], prhs[M + 1];
#pragma unroll
        for (int j = k; j < M; ++j) prow[j] = xb[j * kTile + pl];
#pragma unroll
        for (int j = 0; j <= M; ++j) prhs[j] = xb[(M + j) * kTile + pl];
        zero_pivot |= prow[k] == 0.0 ? 1 : 0;
        const bool elim = rank < 0 && i != p;
        const double l = elim ? r[k] * fast_rcp(prow[k]) : 0.0;
        rank = (rank < 0 && i == p) ? k : rank;
#pragma unroll
        for (int j = k + 1; j < M; ++j) r[j] = fma(-l, prow[j], r[j]);
#pragma unroll
        for (int j = 0; j <= M; ++j) rhs[j] = fma(-l, prhs[j], rhs[j]);
    }
#pragma unroll
    for (int j = 0; j < M; ++j) xb[j * kTile + lane] = r[j];
#pragma unroll
    for (int j = 0; j <= M; ++j) xb[(M + j) * kTile + lane] = rhs[j];
    wave_lds_sync();
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
        const int pl = base + plane[k];
        double sc = xb[(M + 1 + i) * kTile + pl];
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
    return zero_pivot;
}

// Element i of a vector held whole in every lane.
template <int M>
__device__ __forceinline__ double lane_elem(const double (&v)[M], int i) {
    double x = v[0];
#pragma unroll
    for (int t = 1; t < M; ++t) x = (i == t) ? v[t] : x;
    return x;
}

// Cooperative solve: M lanes per system, lane i owns block row i of every block; row q is
// R_q = B_q - A_q C'_{q-1} eliminated with rhs (d_q - A_q d'_{q-1}, C_q) (coop_eliminate),
// C'_q stored over C_q and carried in LDS, d'_q over d_q and carried whole in registers;
// then x_q = d'_q - C'_q x_{q+1} upwards, x_{q+1} exchanged through LDS. Exchanges are
// ordered by wave_lds_sync only. Rows are loaded kFwdAhead (forward) and kBwdAhead
// (backward) rows ahead; every row below `rows` (the lane's pack's) is allocated for every
// slot of the pack, so loads are clamped to that range and need no other guard. All lanes of
// the wave run the same row steps (nmax, the item's longest system, rounded up to the
// prefetch depth); lanes past their system's n compute on what the rows hold and store
// nothing.
template <int M>
__device__ bool coop_rows(const double* __restrict__ A, const double* __restrict__ B, double* C,
                          double* D, int64_t mo, int64_t vo, int i, int base, int n, int nmax, int rows,
                          double* __restrict__ sh) {
    constexpr int E = M * M;
    constexpr int F = CoopDepth<M>::kFwdAhead;
    constexpr int G = CoopDepth<M>::kBwdAhead;
    double* carry = sh;
    double* xb = sh + (M + 1) * kTile;
    const int lane = threadIdx.x;
    constexpr int64_t es = batch_width(M);
    bool ok = true;
    double pa[F][M], pb[F][M], pc[F][M], pd[F];
    auto fetch = [&](double (&fa)[M], double (&fb)[M], double (&fc)[M], double& fd, int q) {
        q = q < rows ? q : rows - 1;
        const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            fa[k] = __builtin_nontemporal_load(A + rm + (i + M * k) * es);
            fb[k] = __builtin_nontemporal_load(B + rm + (i + M * k) * es);
            fc[k] = C[rm + (i + M * k) * es];
        }
        fd = D[rv + i * es];
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
            const bool act = q < n;
            const bool has_next = q + 1 < n;
            const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
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
            fetch(pa[s], pb[s], pc[s], pd[s], q + F);
            double xc[M], xd[M];
            const int zero_pivot = coop_eliminate<M>(r, rhs, xb, i, base, xc, xd);
            const double di = lane_elem<M>(xd, i);
            ok = ok && !(act && zero_pivot);
            if (act) {
                D[rv + i * es] = di;
                xlast = di;
            }
            if (has_next) {
#pragma unroll
                for (int k = 0; k < M; ++k) C[rm + (k + M * i) * es] = xc[k];
            }
#pragma unroll
            for (int k = 0; k < M; ++k) {
                carry[k * kTile + lane] = xc[k];  // C'(k, i) in lane base+i
                dp[k] = xd[k];
            }
            wave_lds_sync();
        }
    }
    // back substitution over the rows: x_q(i) = d'_q(i) - sum_j C'_q(i,j) x_{q+1}(j), with
    // C'_q(i, :) and d'_q(i) loaded kBwdAhead rows ahead (rows are clamped to the lane's pack).
    double xnext = xlast;
    double cn[G][M], dn[G];
    auto bfetch = [&](double (&fc)[M], double& fd, int q) {
        q = q < 0 ? 0 : (q < rows ? q : rows - 1);  // this lane's pack may be shorter than nmax
        const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
#pragma unroll
        for (int j = 0; j < M; ++j) fc[j] = C[rm + (i + M * j) * es];
        fd = D[rv + i * es];
    };
#pragma unroll
    for (int s = 0; s < G; ++s) bfetch(cn[s], dn[s], nmax - 2 - s);
    for (int q0 = nmax - 2; q0 >= 0; q0 -= G) {
#pragma unroll
        for (int s = 0; s < G; ++s) {
            const int q = q0 - s;
            const bool act = q >= 0 && q < n - 1;
            double cq[M], x = dn[s];
#pragma unroll
            for (int j = 0; j < M; ++j) cq[j] = cn[s][j];
            bfetch(cn[s], dn[s], q - G);
            sh[lane] = xnext;
            wave_lds_sync();
#pragma unroll
            for (int j = 0; j < M; ++j) x = fma(-cq[j], sh[base + j], x);
            if (act) {
                D[vo + (int64_t)q * M * es + i * es] = x;
                xnext = x;
            }
            wave_lds_sync();
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
// groups finish.
template <int M>
__device__ bool coop_twisted(const double* A, const double* __restrict__ B, double* C, double* D,
                             int64_t mo, int64_t vo, int i, int base, int pbase, bool top, int n,
                             int steps, int rows, double* __restrict__ sh) {
    constexpr int E = M * M;
    constexpr int F = CoopDepth<M>::kFwdAhead;
    constexpr int G = CoopDepth<M>::kBwdAhead;
    double* carry = sh;
    double* xb = sh + (M + 1) * kTile;
    const int lane = threadIdx.x;
    constexpr int64_t es = batch_width(M);
    const int h = n > 0 ? (n - 1) / 2 : 0;
    const int cnt = n > 0 ? (top ? h : n - 1 - h) : 0;
    const double* sub = top ? A : C;  // the block coupling row q to the row eliminated before it
    const double* sup = top ? C : A;  // the block coupling row q to the row eliminated after it
    auto clamp_row = [&](int q) { return q < 0 ? 0 : (q >= rows ? rows - 1 : q); };
    bool ok = true;
    double ps[F][M], pg[F][M], pu[F][M], pd[F];
    auto fetch = [&](double (&fs)[M], double (&fg)[M], double (&fu)[M], double& fd, int t) {
        const int q = clamp_row(top ? t : n - 1 - t);
        const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            fs[k] = sub[rm + (i + M * k) * es];
            fg[k] = __builtin_nontemporal_load(B + rm + (i + M * k) * es);
            fu[k] = sup[rm + (i + M * k) * es];
        }
        fd = D[rv + i * es];
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
            const bool act = t < cnt;
            const int q = clamp_row(top ? t : n - 1 - t);
            const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
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
            fetch(ps[s], pg[s], pu[s], pd[s], t + F);
            double xc[M], xd[M];
            const int zero_pivot = coop_eliminate<M>(r, rhs, xb, i, base, xc, xd);
            ok = ok && !(act && zero_pivot);
            if (act) {
                D[rv + i * es] = lane_elem<M>(xd, i);
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    C[rm + (k + M * i) * es] = xc[k];
                    carry[k * kTile + lane] = xc[k];
                }
            }
#pragma unroll
            for (int k = 0; k < M; ++k) dp[k] = act ? xd[k] : dp[k];
            wave_lds_sync();
        }
    }
    // meeting row h: both groups of a system solve it from both carries
    carry[M * kTile + lane] = lane_elem<M>(dp, i);
    wave_lds_sync();
    const int tb = top ? base : pbase, bb = top ? pbase : base;
    double xh[M];
    {
        const int64_t rm = mo + (int64_t)h * E * es, rv = vo + (int64_t)h * M * es;
        double r[M], rhs[M + 1];
        double ah[M], ch[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            ah[k] = h > 0 ? A[rm + (i + M * k) * es] : 0.0;  // row 0 has no sub-diagonal block
            ch[k] = C[rm + (i + M * k) * es];
            r[k] = B[rm + (i + M * k) * es];
            rhs[1 + k] = 0.0;
        }
        rhs[0] = D[rv + i * es];
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
        if (n > 0 && top) D[rv + i * es] = lane_elem<M>(xh, i);
    }
    // outward substitution, kBwdAhead rows ahead
    double xprev = lane_elem<M>(xh, i);
    double cn[G][M], dn[G];
    auto bfetch = [&](double (&fc)[M], double& fd, int t) {
        const int q = clamp_row(top ? h - 1 - t : h + 1 + t);
        const int64_t rm = mo + (int64_t)q * E * es, rv = vo + (int64_t)q * M * es;
#pragma unroll
        for (int j = 0; j < M; ++j) fc[j] = C[rm + (i + M * j) * es];
        fd = D[rv + i * es];
    };
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < G; ++s) bfetch(cn[s], dn[s], s);
    for (int t0 = 0; t0 < steps; t0 += G) {
#pragma unroll
        for (int s = 0; s < G; ++s) {
            const int t = t0 + s;
            const bool act = t < cnt;
            const int q = clamp_row(top ? h - 1 - t : h + 1 + t);
            double cq[M], x = dn[s];
#pragma unroll
            for (int j = 0; j < M; ++j) cq[j] = cn[s][j];
            bfetch(cn[s], dn[s], t + G);
            sh[lane] = xprev;
            wave_lds_sync();
#pragma unroll
            for (int j = 0; j < M; ++j) x = fma(-cq[j], sh[base + j], x);
            if (act) {
                D[vo + (int64_t)q * M * es + i * es] = x;
                xprev = x;
            }
            wave_lds_sync();
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
    return coop_rows<M>(A, B, C, D, mo, vo, i, grouped ? g * M : 0, n, nmax, b.pack_rows[pk], sh);
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
    return coop_twisted<M>(A, B, C, D, mo, vo, i, base, pbase, top, n, steps, nmax, sh);
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

// Generator for a packed batch: one thread per (pack, row, slot).
template <int M>
__device__ __forceinline__ void gen_batch_elem(int kind, uint64_t seed, int64_t sys, int row, int n,
                                               int64_t mo, int64_t vo, double* A, double* B,
                                               double* C, double* D) {
    constexpr int64_t es = batch_width(M);
    double a[M][M], bb[M][M], c[M][M], d[M];
    gen_row<M>(kind, seed, sys, row, n, a, bb, c, d);
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
        if (slot >= b.pack_count[p]) continue;
        const int n = b.slot_rows[b.pack_slot0[p] + slot];
        if (row >= n) continue;
        const int64_t sys = slot_sys[b.pack_slot0[p] + slot];
        const int64_t mo = b.mat_off[p] + slot, vo = b.vec_off[p] + slot;
        switch (b.pack_m[p]) {
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
}

}  // namespace btd
