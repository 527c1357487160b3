// Shared host/device definitions: block layouts, seeded input generators and the
// per-lane dense block algebra (LU with partial pivoting) used by every kernel.
//
// Block convention: X[i][j] is row i, column j of an m x m block. Row q of a system is
//   A_q x_{q-1} + B_q x_q + C_q x_{q+1} = d_q,
// A_0 and C_{n-1} are ignored (Fortran_Original/src/mod_btdma_cpu.f90:15-78).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BTD_HD __host__ __device__ __forceinline__

namespace btd {

constexpr int kTile = 64;     // systems per tile = lanes per wavefront
constexpr int kMaxM = 10;     // largest block size handled

// Systems interleaved in one pack of a mixed-size batch: 64 for m = 1 (one lane each), else
// the systems of a twisted work item (two groups of m lanes per system in a wavefront).
BTD_HD constexpr int batch_width(int m) { return m == 1 ? kTile : kTile / (2 * m); }

// ------------------------------------------------------------------------------------
// Layouts. Element e of a block is p + m*q (p = block row, q = block column), the
// column-major order of the reference. A vector element is e = p.
//   REF_GPU  (pascal_btdma_cuda.hpp:46-65): sys + nsys*(row + nrow*e)
//   REF_CPU  (mod_btdma_cpu.f90:18):        e + E*(sys + nsys*row)
//   TILED    (this library):                ((tile*nrow + row)*E + e)*64 + lane,
//            sys = 64*tile + lane, nsys padded up to a multiple of 64.
// ------------------------------------------------------------------------------------
enum Layout : int { LAYOUT_REF_GPU = 0, LAYOUT_REF_CPU = 1, LAYOUT_TILED = 2 };

BTD_HD int64_t pad_nsys(int64_t nsys) { return (nsys + kTile - 1) / kTile * kTile; }

BTD_HD int64_t elem_offset(int layout, int64_t sys, int64_t row, int64_t e, int64_t E,
                           int64_t nsys, int64_t nrow) {
    if (layout == LAYOUT_REF_GPU) return sys + nsys * (row + nrow * e);
    if (layout == LAYOUT_REF_CPU) return e + E * (sys + nsys * row);
    const int64_t tile = sys / kTile, lane = sys % kTile;
    return ((tile * nrow + row) * E + e) * kTile + lane;
}

BTD_HD int64_t layout_elems(int layout, int64_t nsys, int64_t nrow, int64_t E) {
    return (layout == LAYOUT_TILED ? pad_nsys(nsys) : nsys) * nrow * E;
}

// ------------------------------------------------------------------------------------
// Generators. A counter-based hash of (seed, stream, sys, row, entry) gives every
// entry independently, so any subset of systems is regenerated exactly on the host.
// Host and device builds of these functions perform the same IEEE operations in the
// same order (contraction into FMA is switched off below), so they agree bit for bit.
// ------------------------------------------------------------------------------------
enum GenKind : int {
    GEN_DIAG = 0,    // A = 0.25 I, B = 2 I, C = 0.25 I, d = 1 (Study/README.md:40-50)
    GEN_DDPERM = 1,  // strictly diagonally dominant rows, whole block row permuted
    GEN_EULER = 2,   // A = -0.5 s J_{q-1}, C = 0.5 s J_{q+1}, B = I + s diag|J_q| + E
    GEN_COND = 3,    // B = U diag(sv) V^T with cond log-uniform [1e1,1e6]; FP32-rounded
    GEN_BPERM = 4,   // the small CPU-runnable configuration word for word: B = U + diag(row sums
                     // |A| + |C| + 1), then only B's rows permuted
};
constexpr int kGenKinds = 5;

enum GenStream : uint32_t { S_A = 1, S_B = 2, S_C = 3, S_D = 4, S_PERM = 5, S_J = 6, S_E = 7,
                            S_U = 8, S_V = 9, S_K = 10 };

BTD_HD uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

BTD_HD uint64_t row_key(uint64_t seed, uint32_t stream, int64_t sys, int64_t row) {
    return mix64(mix64(mix64(seed ^ ((uint64_t)stream << 56)) + (uint64_t)sys) + (uint64_t)row);
}

// uniform in [-1, 1): a 53-bit integer scaled by 2^-52, minus one (exact).
BTD_HD double u11(uint64_t key, uint32_t idx) {
    const uint64_t h = mix64(key + (uint64_t)(idx + 1) * 0xD1B54A32D192ED03ull);
    return (double)(h >> 11) * 0x1.0p-52 - 1.0;
}

#pragma clang fp contract(off)

BTD_HD double det_abs(double x) { return x < 0.0 ? -x : x; }

// 2^x for x in [-40, 0], built from exactly the same operations on host and device:
// x = k + f with integer k <= 0 and f in [0,1); 2^f by a degree-12 Taylor series of
// exp(f ln2) in Horner form; 2^k by repeated halving (exact).
BTD_HD double det_exp2(double x) {
    int k = 0;
    while (x < 0.0) { x = x + 1.0; k = k - 1; }
    const double t = x * 0.6931471805599453;
    double s = 1.0 / 479001600.0;
    const double inv[12] = {1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0,
                            1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5,
                            1.0, 1.0};
    for (int i = 0; i < 12; ++i) s = s * t + inv[i];
    for (int i = 0; i > k; --i) s = s * 0.5;
    return s;
}

// Random orthogonal m x m matrix: product of m Householder reflectors H(v_k),
// v_k uniform in [-1,1]^m (entry idx = k*m + i of the given key).
template <int M>
BTD_HD void gen_orthogonal(uint64_t key, double (&Q)[M][M]) {
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) Q[i][j] = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k < M; ++k) {
        double v[M];
        double vv = 0.0;
        for (int i = 0; i < M; ++i) {
            v[i] = u11(key, (uint32_t)(k * M + i));
            vv = vv + v[i] * v[i];
        }
        if (vv < 1e-30) continue;
        const double sc = 2.0 / vv;
        // Q <- Q H = Q - sc (Q v) v^T
        for (int i = 0; i < M; ++i) {
            double qv = 0.0;
            for (int j = 0; j < M; ++j) qv = qv + Q[i][j] * v[j];
            const double f = sc * qv;
            for (int j = 0; j < M; ++j) Q[i][j] = Q[i][j] - f * v[j];
        }
    }
}

// One block row (A_q, B_q, C_q, d_q) of system sys, natural form.
template <int M>
BTD_HD void gen_row(int kind, uint64_t seed, int64_t sys, int64_t row, int64_t nrow,
                    double (&A)[M][M], double (&B)[M][M], double (&C)[M][M], double (&d)[M]) {
    const bool first = (row == 0), last = (row == nrow - 1);
    if (kind == GEN_DIAG) {
        for (int i = 0; i < M; ++i) {
            for (int j = 0; j < M; ++j) {
                A[i][j] = (i == j) ? 0.25 : 0.0;
                B[i][j] = (i == j) ? 2.0 : 0.0;
                C[i][j] = (i == j) ? 0.25 : 0.0;
            }
            d[i] = 1.0;
        }
        return;
    }
    if (kind == GEN_BPERM) {
        // A, C ~ U[-1,1] (A of the first row and C of the last are 0), B = U[-1,1] + (row sum of
        // |A| + row sum of |C| + 1) on the diagonal, then the rows of B alone permuted (A, C, d
        // keep theirs), d ~ U[-1,1]
        const uint64_t ka = row_key(seed, S_A, sys, row), kb = row_key(seed, S_B, sys, row);
        const uint64_t kc = row_key(seed, S_C, sys, row), kd = row_key(seed, S_D, sys, row);
        const uint64_t kp = row_key(seed, S_PERM, sys, row);
        double b0[M][M];
        for (int i = 0; i < M; ++i) {
            double s = 1.0;
            for (int j = 0; j < M; ++j) {
                A[i][j] = first ? 0.0 : u11(ka, (uint32_t)(i * M + j));
                C[i][j] = last ? 0.0 : u11(kc, (uint32_t)(i * M + j));
                b0[i][j] = u11(kb, (uint32_t)(i * M + j));
            }
            for (int j = 0; j < M; ++j) s = s + det_abs(A[i][j]);
            for (int j = 0; j < M; ++j) s = s + det_abs(C[i][j]);
            b0[i][i] = b0[i][i] + s;
            d[i] = u11(kd, (uint32_t)i);
        }
        int perm[M];
        for (int i = 0; i < M; ++i) perm[i] = i;
        for (int k = M - 1; k > 0; --k) {
            const int j = (int)(mix64(kp + (uint64_t)k) % (uint64_t)(k + 1));
            const int t = perm[k];
            perm[k] = perm[j];
            perm[j] = t;
        }
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j) B[i][j] = b0[perm[i]][j];
        return;
    }
    if (kind == GEN_DDPERM) {
        const uint64_t ka = row_key(seed, S_A, sys, row), kb = row_key(seed, S_B, sys, row);
        const uint64_t kc = row_key(seed, S_C, sys, row), kd = row_key(seed, S_D, sys, row);
        const uint64_t kp = row_key(seed, S_PERM, sys, row);
        double a0[M][M], b0[M][M], c0[M][M], d0[M];
        for (int i = 0; i < M; ++i) {
            double s = 1.0;
            for (int j = 0; j < M; ++j) {
                a0[i][j] = first ? 0.0 : u11(ka, (uint32_t)(i * M + j));
                c0[i][j] = last ? 0.0 : u11(kc, (uint32_t)(i * M + j));
                b0[i][j] = u11(kb, (uint32_t)(i * M + j));
            }
            for (int j = 0; j < M; ++j) s = s + det_abs(a0[i][j]);
            for (int j = 0; j < M; ++j) s = s + det_abs(c0[i][j]);
            for (int j = 0; j < M; ++j) s = s + det_abs(b0[i][j]);
            b0[i][i] = b0[i][i] + s;
            d0[i] = u11(kd, (uint32_t)i);
        }
        int perm[M];
        for (int i = 0; i < M; ++i) perm[i] = i;
        for (int k = M - 1; k > 0; --k) {
            const int j = (int)(mix64(kp + (uint64_t)k) % (uint64_t)(k + 1));
            const int t = perm[k];
            perm[k] = perm[j];
            perm[j] = t;
        }
        for (int i = 0; i < M; ++i) {
            const int src = perm[i];
            for (int j = 0; j < M; ++j) {
                A[i][j] = a0[src][j];
                B[i][j] = b0[src][j];
                C[i][j] = c0[src][j];
            }
            d[i] = d0[src];
        }
        return;
    }
    if (kind == GEN_EULER) {
        const double sigma = 0.4;
        const uint64_t kj = row_key(seed, S_J, sys, row), ke = row_key(seed, S_E, sys, row);
        const uint64_t kjm = row_key(seed, S_J, sys, row - 1), kjp = row_key(seed, S_J, sys, row + 1);
        const uint64_t kd = row_key(seed, S_D, sys, row);
        for (int i = 0; i < M; ++i) {
            double s = 0.0;
            for (int j = 0; j < M; ++j) {
                const uint32_t ij = (uint32_t)(i * M + j);
                A[i][j] = first ? 0.0 : (-0.5 * sigma) * u11(kjm, ij);
                C[i][j] = last ? 0.0 : (0.5 * sigma) * u11(kjp, ij);
                B[i][j] = 0.1 * u11(ke, ij);
                s = s + det_abs(u11(kj, ij));
            }
            const double shift = 1.0 + sigma * s;
            B[i][i] = B[i][i] + shift;
            d[i] = u11(kd, (uint32_t)i);
        }
        return;
    }
    // GEN_COND
    {
        const uint64_t kk = row_key(seed, S_K, sys, row);
        const double log10k = 1.0 + 5.0 * (0.5 * u11(kk, 0) + 0.5);  // [1, 6)
        const double log2k = log10k * 3.321928094887362;
        double sv[M];
        for (int k = 0; k < M; ++k)
            sv[k] = (M == 1) ? 1.0 : det_exp2(-log2k * (double)k / (double)(M - 1));
        const double smin = sv[M - 1];
        double U[M][M], V[M][M];
        gen_orthogonal<M>(row_key(seed, S_U, sys, row), U);
        gen_orthogonal<M>(row_key(seed, S_V, sys, row), V);
        const uint64_t ka = row_key(seed, S_A, sys, row), kc = row_key(seed, S_C, sys, row);
        const uint64_t kd = row_key(seed, S_D, sys, row);
        const double cpl = 0.25 * smin / (double)M;
        for (int i = 0; i < M; ++i) {
            for (int j = 0; j < M; ++j) {
                double b = 0.0;
                for (int k = 0; k < M; ++k) b = b + (U[i][k] * sv[k]) * V[j][k];
                B[i][j] = (double)(float)b;
                const uint32_t ij = (uint32_t)(i * M + j);
                A[i][j] = first ? 0.0 : (double)(float)(cpl * u11(ka, ij));
                C[i][j] = last ? 0.0 : (double)(float)(cpl * u11(kc, ij));
            }
            d[i] = u11(kd, (uint32_t)i);
        }
    }
}

#pragma clang fp contract(fast)

}  // namespace btd
