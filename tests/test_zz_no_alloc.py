"""btd_solve_noncyclic_noalloc (the allocation-free solve: scratch in the library's static
device array) against btd_solve_noncyclic with a plan of the same partition count, through
the C ABI. Same kernels, so the results must be equal bit for bit: two calls in a row (the
second smaller), a single-partition call, graph replays of a captured call, and the refusal
of a solve whose scratch exceeds the array."""
import ctypes

import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout, layout_to_vec, rel_err

pytestmark = pytest.mark.gpu


def make(cuda, kind, seed, m, nsys, n, layout):
    import torch
    nm = pb.layout_elems(layout, nsys, n, m * m)
    nv = pb.layout_elems(layout, nsys, n, m)
    t = [torch.empty(nm, dtype=torch.float64, device=cuda) for _ in range(3)]
    t.append(torch.empty(nv, dtype=torch.float64, device=cuda))
    pb.generate(kind, seed, m, nsys, n, layout, *t)
    return t


def with_plan(t, m, nsys, n, npart, layout):
    import torch
    a, b, c, d = (x.clone() for x in t)
    with pb.BtdmaGpuPlan().create(m, nsys, n, npart, layout) as plan:
        pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
        torch.cuda.synchronize()
        assert not plan.check()
    return d


def no_alloc(t, m, nsys, n, npart, layout, stream=None):
    a, b, c, d = (x.clone() for x in t)
    pb.solve_noncyclic_noalloc(a, b, c, d, m, nsys, n, npart, layout, stream)
    return a, b, c, d


@pytest.mark.parametrize("layout", [pb.Layout.TILED, pb.Layout.REF_GPU])
def test_two_calls_in_a_row_match_the_plan_solve(cuda, oracle, layout):
    import torch
    cases = [(pb.GenKind.DDPERM, 11, 5, 512, 4096, 32),   # 14.4 MB of reduced system
             (pb.GenKind.DDPERM, 12, 3, 100, 96, 4)]      # smaller: stale scratch must not matter
    for kind, seed, m, nsys, n, npart in cases:
        t = make(cuda, kind, seed, m, nsys, n, layout)
        ref = with_plan(t, m, nsys, n, npart, layout)
        got = no_alloc(t, m, nsys, n, npart, layout)[3]
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (m, nsys, n, npart)
    # the last (smaller) case against the oracle as well
    A, B, C, d = pb.generate_host(kind, seed, m, nsys, n)
    xo, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    assert rel_err(layout_to_vec(got.cpu().numpy(), int(layout), n, nsys, m), xo).max() <= 1e-12


def test_single_partition_and_library_choice(cuda):
    import torch
    lay = pb.Layout.TILED
    for m, nsys, n, npart in ((5, 300, 64, 1), (4, 200, 400, 0)):
        t = make(cuda, pb.GenKind.EULER, 5 + m, m, nsys, n, lay)
        ref = with_plan(t, m, nsys, n, npart, lay)
        got = no_alloc(t, m, nsys, n, npart, lay)[3]
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (m, nsys, n, npart)


def test_graph_replays_equal_the_eager_call(cuda):
    """The call captured into a graph (no allocation, no memset node) replays to the eager
    result, with other eager solves of the same array between replays."""
    import torch
    m, nsys, n, npart, lay = 5, 256, 1024, 16, pb.Layout.TILED
    t = make(cuda, pb.GenKind.DDPERM, 21, m, nsys, n, lay)
    eager = no_alloc(t, m, nsys, n, npart, lay)[3]
    other = make(cuda, pb.GenKind.DDPERM, 22, 3, 64, 60, lay)
    a, b, c, d = (x.clone() for x in t)
    s1 = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = pb.Graph()
    g.begin(s1)
    pb.solve_noncyclic_noalloc(a, b, c, d, m, nsys, n, npart, lay, s1)
    g.end(s1)
    for _ in range(3):
        a.copy_(t[0])
        c.copy_(t[2])
        d.copy_(t[3])
        no_alloc(other, 3, 64, 60, 4, lay)  # eager use of the shared scratch between replays
        torch.cuda.synchronize()
        g.launch(s1)
        s1.synchronize()
        assert torch.equal(d, eager)
    g.destroy()


def test_refuses_more_scratch_than_the_array_holds(cuda):
    """m = 10, 4096 systems, 64 partitions would need ~0.84 GB of reduced system: refused
    before anything runs (the arrays are never touched, so small tensors suffice)."""
    import torch
    lib = pb._capi.lib()
    t = torch.zeros(16, dtype=torch.float64, device=cuda)
    p = t.data_ptr()
    st = lib.btd_solve_noncyclic_noalloc(p, p, p, p, 10, 4096, 4096, 64, int(pb.Layout.TILED), None)
    assert st == 1  # BTD_ERR_INVALID_ARG
    assert "bytes of scratch" in pb._capi.last_error()
    st = lib.btd_solve_noncyclic_noalloc(p, p, p, p, 11, 64, 64, 1, int(pb.Layout.TILED), None)
    assert st == 1
    torch.cuda.synchronize()
    assert torch.count_nonzero(t) == 0
