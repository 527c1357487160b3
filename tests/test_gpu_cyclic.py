"""Periodic (cyclic) solve on the device vs the oracle (btdma_many_cycl and
btdma_many_cyclic_mpi restatements, both pinned to dense periodic solves)."""
import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout, layout_to_vec, nat_to_layout, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def periodic_system(rng, n, nsys, m):
    A = rng.uniform(-1, 1, (n, nsys, m, m))
    C = rng.uniform(-1, 1, (n, nsys, m, m))
    B = rng.uniform(-1, 1, (n, nsys, m, m))
    s = np.abs(A).sum(-1) + np.abs(C).sum(-1) + np.abs(B).sum(-1) + 1.0
    B += s[..., None] * np.eye(m)
    for q in range(n):
        for k in range(nsys):
            p = rng.permutation(m)
            A[q, k], B[q, k], C[q, k] = A[q, k][p], B[q, k][p], C[q, k][p]
    d = rng.uniform(-1, 1, (n, nsys, m))
    return A, B, C, d


def gpu_cyclic(cuda, A, B, C, d, layout, npart):
    import torch
    n, nsys, m = d.shape
    lay = int(layout)
    t = [torch.from_numpy(nat_to_layout(X, lay)).to(cuda) for X in (A, B, C, d)]
    with pb.BtdmaGpuPlan().create(m, nsys, n, npart, layout) as plan:
        pb.solve_cyclic(plan, *t, m, nsys, n)
        torch.cuda.synchronize()
        assert not plan.check()
    return layout_to_vec(t[3].cpu().numpy(), lay, n, nsys, m)


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 10])
@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
def test_cyclic_single_partition(cuda, oracle, m, layout):
    rng = np.random.default_rng(40 + m)
    A, B, C, d = periodic_system(rng, 17, 70, m)
    x = gpu_cyclic(cuda, A, B, C, d, layout, 1)
    xo, info = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    assert rel_err(x, xo).max() <= TOL


@pytest.mark.parametrize("m,n,npart", [(2, 24, 2), (3, 96, 4), (5, 96, 8), (3, 100, 7), (5, 9, 3),
                                        (8, 40, 3), (10, 30, 4)])
@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
def test_cyclic_two_stage(cuda, oracle, m, n, npart, layout):
    rng = np.random.default_rng(90 + n + m)
    A, B, C, d = periodic_system(rng, n, 67, m)
    x = gpu_cyclic(cuda, A, B, C, d, layout, npart)
    xo, info = oracle.solve_cyclic_two_stage(cpu_layout(A), cpu_layout(B), cpu_layout(C), d, npart)
    assert info == 0
    assert rel_err(x, xo).max() <= TOL
    x1, _ = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert rel_err(x, x1).max() <= TOL


def test_cyclic_three_rows(cuda, oracle):
    rng = np.random.default_rng(3)
    A, B, C, d = periodic_system(rng, 3, 64, 4)
    x = gpu_cyclic(cuda, A, B, C, d, pb.Layout.TILED, 1)
    xo, _ = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert rel_err(x, xo).max() <= TOL


@pytest.mark.parametrize("m", [2, 5, 8])
@pytest.mark.parametrize("n", [3, 4, 5, 6, 7])
def test_cyclic_single_partition_row_counts(cuda, oracle, m, n):
    """npart = 1: fewer than 6 rows run one periodic sweep per system (k_cyclic); from 6 rows
    the periodic solve runs over partitions the library chooses (here 2)."""
    rng = np.random.default_rng(7 * n + m)
    A, B, C, d = periodic_system(rng, n, 66, m)
    x = gpu_cyclic(cuda, A, B, C, d, pb.Layout.TILED, 1)
    xo, info = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    assert rel_err(x, xo).max() <= TOL


@pytest.mark.parametrize("m,nsys,n", [(3, 3000, 64), (8, 4100, 40), (3, 200, 33)])
def test_cyclic_single_partition_partition_counts(cuda, oracle, m, nsys, n):
    """npart = 1 periodic solves whose library-chosen partition count is 8 (64 rows / 8), 4 (a
    batch the single-partition solve fills the device with) and 4 (33 rows)."""
    rng = np.random.default_rng(m + n)
    A, B, C, d = periodic_system(rng, n, nsys, m)
    x = gpu_cyclic(cuda, A, B, C, d, pb.Layout.REF_GPU, 1)
    xo, info = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    assert rel_err(x, xo).max() <= TOL


def test_cyclic_and_plain_on_one_plan(cuda, oracle):
    """One npart = 1 plan: a periodic solve (which sets up its partitions), a non-periodic
    solve, then a periodic solve of other data reusing them; each against the oracle."""
    import torch
    m, nsys, n = 4, 90, 20
    lay = int(pb.Layout.TILED)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, pb.Layout.TILED) as plan:
        for k, cyc in enumerate((True, False, True)):
            rng = np.random.default_rng(500 + k)
            A, B, C, d = periodic_system(rng, n, nsys, m)
            t = [torch.from_numpy(nat_to_layout(X, lay)).to(cuda) for X in (A, B, C, d)]
            (pb.solve_cyclic if cyc else pb.solve_noncyclic)(plan, *t, m, nsys, n)
            torch.cuda.synchronize()
            assert not plan.check()
            x = layout_to_vec(t[3].cpu().numpy(), lay, n, nsys, m)
            if cyc:
                xo, _ = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
            else:
                A[0] = 0.0  # row 0's sub-diagonal block is outside a non-periodic system
                C[-1] = 0.0
                xo, _ = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
            assert rel_err(x, xo).max() <= TOL


def test_cyclic_first_call_refused_under_capture_then_graph_replays(cuda, oracle):
    """The first periodic solve of an npart = 1 plan allocates its partition scratch: under
    graph capture it is refused with an error instead of breaking the capture. After one eager
    periodic call the solve captures, and every replay equals the eager result bit for bit."""
    import torch
    m, nsys, n = 3, 130, 40
    lay = int(pb.Layout.TILED)
    rng = np.random.default_rng(77)
    A, B, C, d = periodic_system(rng, n, nsys, m)
    t0 = [torch.from_numpy(nat_to_layout(X, lay)).to(cuda) for X in (A, B, C, d)]
    t = [x.clone() for x in t0]
    s1 = torch.cuda.Stream()
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, pb.Layout.TILED) as plan:
        torch.cuda.synchronize()
        g = pb.Graph()
        g.begin(s1)
        with pytest.raises(ValueError, match="cannot be captured"):
            pb.solve_cyclic(plan, *t, m, nsys, n, stream=s1)
        try:
            g.end(s1)  # the capture recorded nothing
        except pb.BtdmaError:
            pass
        g.destroy()
        pb.solve_cyclic(plan, *t, m, nsys, n)  # eager: sets up the partitions
        torch.cuda.synchronize()
        assert not plan.check()
        eager = t[3].clone()
        xo, info = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
        assert info == 0
        assert rel_err(layout_to_vec(eager.cpu().numpy(), lay, n, nsys, m), xo).max() <= TOL
        for x, x0 in zip(t, t0):
            x.copy_(x0)
        torch.cuda.synchronize()
        g = pb.Graph()
        g.begin(s1)
        pb.solve_cyclic(plan, *t, m, nsys, n, stream=s1)
        g.end(s1)
        for _ in range(3):
            for x, x0 in zip(t, t0):
                x.copy_(x0)
            torch.cuda.synchronize()
            g.launch(s1)
            s1.synchronize()
            assert torch.equal(t[3], eager)
        g.destroy()
        assert not plan.check()
