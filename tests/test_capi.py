"""CPU-side checks of the product library: it loads, exports every symbol the public
header declares, validates arguments before touching the device, and its host
generator has the properties the workload definitions promise. No device compute."""
import os
import re

import numpy as np
import pytest

import pascal_btdma_amd as pb
from pascal_btdma_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "pascal_btdma_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(btd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _capi.SIGNATURES, f"{s} has no ctypes signature"
    assert set(_capi.SIGNATURES) <= set(syms)
    assert b"gfx950" in lib.btd_version()


def test_index_helpers_match_reference_formula():
    nsys, nrow, m = 7, 5, 3
    for sys in range(nsys):
        for row in range(nrow):
            for p in range(m):
                assert pb.vector_index(sys, row, p, nsys, nrow) == sys + nsys * (row + nrow * p)
                for q in range(m):
                    assert pb.matrix_index(sys, row, p, q, nsys, nrow, m) == \
                        sys + nsys * (row + nrow * (p + m * q))


def test_partition_matches_oracle_para(oracle):
    for n in (1, 7, 64, 100, 65536):
        for P in (1, 2, 3, 8, 32):
            if P > n:
                continue
            for r in range(P):
                f, l = pb.partition_1d(0, n - 1, P, r)
                size, ia, ib = oracle.para(0, n - 1, r, P)
                assert (f, l) == (ia, ib)


def test_layout_elems():
    assert pb.layout_elems(pb.Layout.REF_GPU, 100, 7, 9) == 100 * 7 * 9
    assert pb.layout_elems(pb.Layout.TILED, 100, 7, 9) == 128 * 7 * 9


@pytest.mark.parametrize("args,exc", [
    ((0, 10, 10, 1, pb.Layout.TILED), ValueError),
    ((11, 10, 10, 1, pb.Layout.TILED), ValueError),
    ((3, 0, 10, 1, pb.Layout.TILED), ValueError),
    ((3, 10, 1, 1, pb.Layout.TILED), ValueError),
    ((3, 10, 10, 4, pb.Layout.TILED), ValueError),   # needs nrow >= 3*npart
    ((3, 10, 10, -1, pb.Layout.TILED), ValueError),
    ((3, 10, 10, 1, pb.Layout.REF_CPU), ValueError),
])
def test_plan_create_rejects_bad_arguments(args, exc):
    with pytest.raises(exc):
        pb.BtdmaGpuPlan().create(*args)


def test_auto_npart_rule():
    """npart = 0: the single-kernel solve for large batches (8192 systems suffice for m = 2, 4,
    12288 for m = 3, 5, 4096 for m = 6, 3072 for m >= 7); for smaller m <= 5 batches about 16384 partitions of >= 16 rows (a power of
    two; at least 2 from 8192 systems; none below 128 rows under 8192 systems); for m = 6 the
    65536-lane count as a power of two (none up to 128 rows); for m >= 7 about 65536 lanes of
    >= 8 rows."""
    auto = lambda m, nsys, nrow: pb._capi.lib().btd_auto_npart(m, nsys, nrow)
    assert auto(5, 131072, 512) == 1          # the headline batch
    assert auto(5, 32768, 512) == 1
    assert auto(5, 1024, 128) == 8            # the Study's single-GPU case: nrow / 16
    assert auto(5, 512, 65536) == 32          # long systems: 16384 partitions
    assert auto(5, 700, 200) == 8             # min(24, 12) -> power of two
    assert auto(8, 4096, 2048) == 1           # m >= 7: m or 2m lanes per system fill the device
    assert auto(8, 2048, 2048) == 32
    assert auto(7, 4096, 2048) == 1 and auto(7, 2048, 2048) == 32 and auto(7, 700, 200) == 25
    assert auto(7, 3072, 256) == 1 and auto(10, 3072, 2048) == 1 and auto(9, 3071, 256) == 22
    assert auto(8, 1024, 128) == 1 and auto(10, 1024, 192) == 1 and auto(8, 1024, 193) == 24
    assert auto(8, 1023, 128) == 16
    assert auto(6, 4096, 2048) == 1 and auto(6, 2048, 2048) == 32 and auto(6, 700, 300) == 32
    assert auto(6, 1024, 128) == 1
    assert auto(2, 8192, 512) == 1 and auto(4, 8192, 512) == 1
    assert auto(3, 8192, 512) == 2 and auto(3, 8192, 64) == 2 and auto(5, 12287, 128) == 2
    assert auto(3, 16384, 64) == 1 and auto(5, 16384, 128) == 1
    assert auto(4, 4096, 2048) == 4 and auto(2, 256, 512) == 32 and auto(3, 256, 2048) == 64
    assert auto(5, 1024, 64) == 1 and auto(8, 256, 127) == 1 and auto(4, 1024, 64) == 1  # short systems
    assert auto(3, 1024, 15) == 1
    assert auto(3, 12288, 2048) == 1 and auto(5, 16384, 256) == 1 and auto(3, 32767, 512) == 1
    assert auto(5, 12287, 2048) == 2 and auto(3, 8192, 2048) == 2 and auto(1, 20480, 512) == 2
    for m, nsys, nrow in [(2, 7, 40), (10, 3000, 4096), (5, 1, 2), (3, 9000, 20), (6, 100, 129)]:
        p = auto(m, nsys, nrow)
        assert p == 1 or (p >= 2 and nrow >= 8 * p)


def test_uncreated_plan_errors():
    plan = pb.BtdmaGpuPlan()
    with pytest.raises(RuntimeError):
        pb.solve_noncyclic(plan, 1, 1, 1, 1, 3, 10, 10, stream=0)
    with pytest.raises(RuntimeError):
        pb.solve_cyclic(plan, 1, 1, 1, 1, 3, 10, 10, stream=0)


def test_generator_subset_is_consistent():
    full = pb.generate_host(pb.GenKind.DDPERM, 3, 3, 20, 9)
    part = pb.generate_host(pb.GenKind.DDPERM, 3, 3, 20, 9, sys0=5, count=4)
    for f, p in zip(full, part):
        np.testing.assert_array_equal(f[:, 5:9], p)


def test_generator_ddperm_rows_are_dominant_after_unpermuting():
    for m in (2, 3, 5, 8, 10):
        A, B, C, d = pb.generate_host(pb.GenKind.DDPERM, 11, m, 6, 7)
        assert np.all(A[0] == 0) and np.all(C[-1] == 0)
        for q in range(7):
            for s in range(6):
                jmax = np.abs(B[q, s]).argmax(axis=1)
                assert sorted(jmax) == list(range(m))  # a permutation of the columns
                for i in range(m):
                    j = jmax[i]
                    off = np.abs(B[q, s, i]).sum() - abs(B[q, s, i, j])
                    margin = abs(B[q, s, i, j]) - off - np.abs(A[q, s, i]).sum() - np.abs(C[q, s, i]).sum()
                    assert margin >= 1.0 - 1e-12


def test_generator_euler_couples_neighbouring_rows():
    A, B, C, d = pb.generate_host(pb.GenKind.EULER, 5, 5, 4, 9)
    # A_{q+1} = -0.2 J_q and C_{q-1} = 0.2 J_q, so C_{q-1} = -A_{q+1} exactly
    np.testing.assert_array_equal(C[:-2], -A[2:])
    assert np.all(A[0] == 0) and np.all(C[-1] == 0)


def test_generator_cond_blocks():
    A, B, C, d = pb.generate_host(pb.GenKind.COND, 9, 5, 16, 8)
    for X in (A, B, C):
        np.testing.assert_array_equal(X.astype(np.float32).astype(np.float64), X)
    k = np.linalg.cond(B.reshape(-1, 5, 5))
    assert k.min() > 5 and k.max() < 2e6
    assert np.log10(k).std() > 0.8  # spread over the log-uniform range


def test_generator_diag_is_study_workload():
    A, B, C, d = pb.generate_host(pb.GenKind.DIAG, 0, 3, 2, 4)
    np.testing.assert_array_equal(B, np.broadcast_to(2 * np.eye(3), B.shape))
    np.testing.assert_array_equal(A, np.broadcast_to(0.25 * np.eye(3), A.shape))
    np.testing.assert_array_equal(d, 1.0)


def test_new_entry_points_reject_an_uncreated_plan():
    """Every entry point without a reference counterpart checks its plan before any device
    use: a NULL plan gives BTD_ERR_NOT_CREATED (no GPU needed)."""
    lib = _capi.lib()
    NOT_CREATED = 2
    null = None
    calls = {
        "btd_solve_noncyclic_f32s": (null, 1, 1, 1, 1, 1, null),
        "btd_plan_place_workspace_f32s": (null, 1, 1, 1, 1, 2, None, null),
        "btd_solve_noncyclic_save_out": (null, 1, 1, 1, 1, 1, 1, 1, 1, null),
        "btd_solve_transposed_f32r": (null, 1, 1, 1, 1, 1, 1, 2, null),
        "btd_outer_grads_f32": (null, 1, 1, 1, 1, 1, null),
        "btd_euler_assemble": (null, 0, 1, 1, 4, 4, 4, 0.1, 0.1, 0.1, 1, 1, 1, 1, null),
        "btd_euler_sweep_cyclic": (null, 0, 1, 1, 4, 4, 4, 0.1, 0.1, 0.1, 1, 1, null),
        "btd_sweep_gather": (null, 0, 1, 4, 4, 4, 1, null),
        "btd_sweep_scatter": (null, 0, 1, 4, 4, 4, 1, null),
        "btd_plan_info": (null, 1),
        "btd_plan_npart": (null, None),
        "btd_plan_check": (null, None),
    }
    for name, args in calls.items():
        assert getattr(lib, name)(*args) == NOT_CREATED, name
        assert lib.btd_last_error() != b"", name


def test_generator_bperm_is_the_literal_small_configuration():
    """BPERM: A, C uniform in [-1, 1] (A of the first row and C of the last zero), d uniform;
    B's rows are a permutation of U + diag(1 + row sums of |A| and |C|) with U in [-1, 1): some
    row permutation of B minus that diagonal lies in [-1, 1), and A, C, d are not permuted (their
    rows line up with the diagonal shift). Permutations other than the identity occur."""
    import itertools
    m = 3
    A, B, C, d = pb.generate_host(pb.GenKind.BPERM, 21, m, 40, 6)
    assert np.all(A[0] == 0) and np.all(C[-1] == 0)
    for X in (A, C, d):
        assert np.all(np.abs(X) <= 1.0)
    s = 1.0 + np.abs(A).sum(-1) + np.abs(C).sum(-1)  # (n, nsys, m)
    nonid = 0
    for q in range(B.shape[0]):
        for k in range(B.shape[1]):
            found = None
            for perm in itertools.permutations(range(m)):
                b0 = B[q, k][np.argsort(perm)]  # undo: B[i] = b0[perm[i]]
                u = b0 - np.diag(s[q, k])
                if np.all(u >= -1.0) and np.all(u < 1.0):
                    found = perm
                    break
            assert found is not None, (q, k)
            nonid += found != tuple(range(m))
    assert nonid > 0

