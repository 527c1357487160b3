"""Pins the CPU oracle (oracle/btdma_oracle.c) before anything is checked against it.

Each pin is chosen so that a plausible mistake (a dropped term, a wrong sign or index,
a transposed block, a wrong pivot) fails it:
  * dgesv   vs scipy.linalg.lu_factor (LAPACK getrf): same pivot rows, same factors
  * plain / two-stage / cyclic / transposed solves vs dense numpy solves of the
    assembled (n*m) x (n*m) matrix (brute force)
  * the reference Study's solution signatures (tests/golden/study_signatures.csv,
    extracted from Study/result/btdma_solution_signature_260702_193138.csv) for the
    plain (np = 1) and two-stage (np = 2, 4, 8) paths
  * invariance of x under a permutation of the rows of whole block rows
  * P invariance of the two-stage solve and the N >= 3 guard
  * mpiutil_para's closed form
"""
import csv
import os

import numpy as np
import pytest
import scipy.linalg as sla

from helpers import cpu_layout

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rand_system(rng, n, nsys, m, cyclic=False, dominance=2.0):
    A = rng.uniform(-1, 1, (n, nsys, m, m))
    C = rng.uniform(-1, 1, (n, nsys, m, m))
    if not cyclic:
        A[0] = 0.0
        C[-1] = 0.0
    B = rng.uniform(-1, 1, (n, nsys, m, m))
    s = np.abs(A).sum(-1) + np.abs(C).sum(-1) + np.abs(B).sum(-1) + dominance
    B += s[..., None] * np.eye(m)
    # permute the rows of every whole block row so that pivoting is needed
    for q in range(n):
        for k in range(nsys):
            p = rng.permutation(m)
            A[q, k], B[q, k], C[q, k] = A[q, k][p], B[q, k][p], C[q, k][p]
    d = rng.uniform(-1, 1, (n, nsys, m))
    return A, B, C, d


def dense_x(oracle, A, B, C, d, cyclic=False, transpose=False):
    n, nsys, m = d.shape
    out = np.empty_like(d)
    for k in range(nsys):
        T = oracle.assemble_dense(A, B, C, k, cyclic=cyclic)
        if transpose:
            T = T.T
        out[:, k] = np.linalg.solve(T, d[:, k].reshape(-1)).reshape(n, m)
    return out


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 10])
def test_dgesv_matches_lapack_getrf(oracle, m):
    rng = np.random.default_rng(100 + m)
    for _ in range(20):
        a = rng.uniform(-1, 1, (m, m))
        b = rng.uniform(-1, 1, (m, 3))
        lu, ipiv, x, info = oracle.dgesv(a, b)
        assert info == 0
        lu_ref, piv_ref = sla.lu_factor(a)
        np.testing.assert_array_equal(ipiv - 1, piv_ref)  # same pivot rows
        np.testing.assert_allclose(lu, lu_ref, rtol=0, atol=1e-13 * np.abs(lu_ref).max())
        xr = np.linalg.solve(a, b)
        kappa = np.linalg.cond(a)
        assert np.abs(x - xr).max() <= 1e-14 * kappa * np.abs(xr).max()


def test_dgesv_reports_singular(oracle):
    a = np.array([[1.0, 2.0], [2.0, 4.0]])
    _, _, _, info = oracle.dgesv(a, np.ones(2))
    assert info == 2


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 10])
def test_plain_matches_dense(oracle, m):
    rng = np.random.default_rng(m)
    A, B, C, d = rand_system(rng, 9, 3, m)
    x, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    np.testing.assert_allclose(x, dense_x(oracle, A, B, C, d), rtol=0, atol=1e-13)


def test_plain_ignores_first_a_and_last_c(oracle):
    rng = np.random.default_rng(5)
    A, B, C, d = rand_system(rng, 7, 2, 3)
    x0, _ = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    A[0] = 7.0
    C[-1] = -3.0
    x1, _ = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    np.testing.assert_array_equal(x0, x1)


@pytest.mark.parametrize("m", [2, 3, 5])
def test_block_row_permutation_invariance(oracle, m):
    rng = np.random.default_rng(7 * m)
    A, B, C, d = rand_system(rng, 16, 4, m)
    x0, _ = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    for q in range(16):
        for k in range(4):
            p = rng.permutation(m)
            A[q, k], B[q, k], C[q, k], d[q, k] = A[q, k][p], B[q, k][p], C[q, k][p], d[q, k][p]
    x1, _ = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    np.testing.assert_allclose(x1, x0, rtol=0, atol=1e-13)


@pytest.mark.parametrize("m,n,npart", [(1, 12, 4), (2, 24, 8), (3, 20, 3), (5, 30, 5),
                                        (3, 9, 3), (8, 16, 2), (10, 12, 3), (3, 100, 7)])
def test_two_stage_matches_dense(oracle, m, n, npart):
    rng = np.random.default_rng(n * 10 + m)
    A, B, C, d = rand_system(rng, n, 3, m)
    x, info = oracle.solve_two_stage(cpu_layout(A), cpu_layout(B), cpu_layout(C), d, npart)
    assert info == 0
    np.testing.assert_allclose(x, dense_x(oracle, A, B, C, d), rtol=0, atol=1e-12)


def test_two_stage_p_invariance_and_guard(oracle):
    rng = np.random.default_rng(11)
    A, B, C, d = rand_system(rng, 48, 5, 3)
    xs = [oracle.solve_two_stage(cpu_layout(A), cpu_layout(B), cpu_layout(C), d, p)[0]
          for p in (1, 2, 4, 8, 16)]
    for x in xs[1:]:
        np.testing.assert_allclose(x, xs[0], rtol=0, atol=1e-13)
    # fewer than 3 rows in a partition is rejected (the modified sweep needs N >= 3)
    _, info = oracle.solve_two_stage(cpu_layout(A), cpu_layout(B), cpu_layout(C), d, 24)
    assert info == -1


@pytest.mark.parametrize("m", [1, 2, 3, 5])
def test_cyclic_matches_dense_periodic(oracle, m):
    rng = np.random.default_rng(300 + m)
    A, B, C, d = rand_system(rng, 10, 3, m, cyclic=True)
    x, info = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    np.testing.assert_allclose(x, dense_x(oracle, A, B, C, d, cyclic=True), rtol=0, atol=1e-12)


@pytest.mark.parametrize("m", [1, 3, 5])
def test_transposed_matches_dense(oracle, m):
    rng = np.random.default_rng(400 + m)
    A, B, C, d = rand_system(rng, 11, 3, m)
    lam, info = oracle.solve_transposed(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    np.testing.assert_allclose(lam, dense_x(oracle, A, B, C, d, transpose=True), rtol=0, atol=1e-12)


def test_para_closed_form(oracle):
    for n in range(1, 40):
        for nprocs in range(1, min(n, 12) + 1):
            nxt = 0
            for r in range(nprocs):
                size, ia, ib = oracle.para(0, n - 1, r, nprocs)
                assert ia == nxt and ib - ia + 1 == size
                assert size == n // nprocs + (1 if r < n % nprocs else 0)
                nxt = ib + 1
            assert nxt == n


def study_rows():
    with open(os.path.join(GOLDEN, "study_signatures.csv")) as f:
        return list(csv.DictReader(f))


def diag_system(n, nsys, m):
    eye = np.eye(m)
    A = np.broadcast_to(0.25 * eye, (n, nsys, m, m)).copy()
    B = np.broadcast_to(2.0 * eye, (n, nsys, m, m)).copy()
    C = A.copy()
    d = np.ones((n, nsys, m))
    return A, B, C, d


def signature(x, nsys_total):
    """Study signature of D = x in the reference GPU layout (Study/example_cuda_cxx_btdma_profile.cu:76-160),
    for a batch of identical systems of which x holds one or more."""
    n, nsys, m = x.shape
    scale = nsys_total / nsys
    return dict(solution_sum=x.sum() * scale, solution_l2=np.sqrt((x * x).sum() * scale),
                solution_linf=np.abs(x).max(), sample_z0=x[0, 0, 0], sample_zmid=x[n // 2, 0, 0],
                sample_zlast=x[n - 1, 0, 0])


@pytest.mark.parametrize("row", study_rows(), ids=lambda r: f"np{r['nranks']}_{r['n3']}_m{r['m']}")
def test_study_signatures(oracle, row):
    """The diagonal workload's signatures as published by the reference (H200 runs).
    All systems are identical, so one system is solved and the sums are scaled."""
    n, m, nranks = int(row["n3"]), int(row["m"]), int(row["nranks"])
    nsys_total = int(row["nsys"])
    A, B, C, d = diag_system(n, 1, m)
    if nranks == 1:
        x, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    else:
        x, info = oracle.solve_two_stage(cpu_layout(A), cpu_layout(B), cpu_layout(C), d, nranks)
    assert info == 0
    sig = signature(x, nsys_total)
    for k in ("solution_sum", "solution_l2"):
        assert abs(sig[k] / float(row[k]) - 1) < 1e-9, (k, sig[k], row[k])
    for k in ("solution_linf", "sample_z0", "sample_zmid", "sample_zlast"):
        assert abs(sig[k] - float(row[k])) < 1e-14, (k, sig[k], row[k])


def test_study_closed_form_scalar_tridiagonal():
    """x of the diagonal workload is the scalar tridiagonal (0.25, 2, 0.25) x = 1; its
    end value 0.45080666151703325 and mid value 0.4 are what the reference published."""
    n = 128
    ab = np.zeros((3, n))
    ab[0, 1:] = 0.25
    ab[1] = 2.0
    ab[2, :-1] = 0.25
    x = sla.solve_banded((1, 1), ab, np.ones(n))
    assert abs(x[0] - 0.4508066615170332) < 1e-15
    assert abs(x[n // 2] - 0.4) < 1e-15


@pytest.mark.parametrize("m,n,npart", [(1, 12, 2), (2, 24, 4), (3, 20, 3), (5, 30, 5), (3, 100, 7)])
def test_cyclic_two_stage_matches_dense_periodic(oracle, m, n, npart):
    rng = np.random.default_rng(700 + m * n)
    A, B, C, d = rand_system(rng, n, 3, m, cyclic=True)
    x, info = oracle.solve_cyclic_two_stage(cpu_layout(A), cpu_layout(B), cpu_layout(C), d, npart)
    assert info == 0
    np.testing.assert_allclose(x, dense_x(oracle, A, B, C, d, cyclic=True), rtol=0, atol=1e-12)
    x1, _ = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    np.testing.assert_allclose(x, x1, rtol=0, atol=1e-13)


# ------------------------------------------------------------------ truth (binary128) --

def _exact_dense_solve(T, rhs):
    """Gaussian elimination with partial pivoting in exact rational arithmetic
    (fractions.Fraction reads each double exactly); returns the exact solution as Fractions."""
    from fractions import Fraction
    N = T.shape[0]
    M = [[Fraction(float(v)) for v in row] + [Fraction(float(r))] for row, r in zip(T, rhs)]
    for k in range(N):
        p = max(range(k, N), key=lambda i: abs(M[i][k]))
        M[k], M[p] = M[p], M[k]
        for i in range(k + 1, N):
            if M[i][k] != 0:
                f = M[i][k] / M[k][k]
                M[i] = [a - f * b for a, b in zip(M[i], M[k])]
    x = [Fraction(0)] * N
    for k in range(N - 1, -1, -1):
        x[k] = (M[k][N] - sum(M[k][j] * x[j] for j in range(k + 1, N))) / M[k][k]
    return x


@pytest.mark.parametrize("kind,m,n,nsys", [(1, 3, 4, 3),   # DDPERM: permuted dominant rows
                                           (3, 5, 3, 3),   # COND: cond(B) log-uniform 1e1..1e6
                                           (3, 2, 6, 3),
                                           (2, 5, 3, 2)])  # EULER-like
def test_truth_equals_exact_rational_solution(oracle, kind, m, n, nsys):
    """The binary128 solve rounded to FP64 equals the exact solution of the FP64 inputs
    (rational arithmetic on the assembled dense matrix) to 2 ulp in every component."""
    import pascal_btdma_amd as pb
    A, B, C, d = pb.generate_host(pb.GenKind(kind), 4242 + m, m, nsys, n)
    xq, info = oracle.solve_truth(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    for s in range(nsys):
        T = oracle.assemble_dense(A, B, C, s)
        exact = _exact_dense_solve(T, d[:, s, :].reshape(-1))
        xe = np.array([float(v) for v in exact])  # correctly rounded
        got = xq[:, s, :].reshape(-1)
        ulp = np.spacing(np.abs(xe))
        assert np.all(np.abs(got - xe) <= 2 * ulp), (s, np.max(np.abs(got - xe) / ulp))


def test_truth_matches_fp64_oracle_when_well_conditioned(oracle):
    """On diagonally dominant rows the FP64 oracle is accurate to a few ulps, so the two agree
    to 1e-15; on blocks of condition number up to 1e6 the FP64 oracle's distance from the
    truth grows to about cond * 1e-16 (here above 1e-13 for some systems), which a correct
    truth must show and a truth that merely repeated the FP64 arithmetic would not."""
    import pascal_btdma_amd as pb
    from helpers import rel_err
    A, B, C, d = pb.generate_host(pb.GenKind.DDPERM, 17, 5, 64, 96)
    x, _ = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    xq, info = oracle.solve_truth(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    assert rel_err(x, xq).max() <= 1e-15
    A, B, C, d = pb.generate_host(pb.GenKind.COND, 17, 5, 64, 96)
    x, _ = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    xq, _ = oracle.solve_truth(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    e = rel_err(x, xq)
    assert e.max() > 1e-13 and e.max() < 1e-9
