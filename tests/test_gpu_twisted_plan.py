"""The twisted form of a plan's single-partition solve (btd_plan_set_form, BTD_FORM_TWISTED):
two groups of m lanes per system, one eliminating from each end, against the oracle, in both
layouts."""
import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout, gpu_solve, nat_to_layout, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12
TILED = int(pb.Layout.TILED)


def oracle_plain(oracle, A, B, C, d):
    x, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    return x


@pytest.mark.parametrize("m,nsys,n", [(2, 70, 2), (3, 64, 3), (5, 130, 17), (8, 33, 64), (10, 7, 40),
                                      (4, 200, 101), (6, 65, 5), (7, 100, 9), (9, 20, 31), (5, 1, 4),
                                      (8, 200, 24), (10, 150, 19)])
@pytest.mark.parametrize("layout", [pb.Layout.TILED, pb.Layout.REF_GPU])
def test_twisted_matches_oracle(cuda, oracle, m, nsys, n, layout):
    """Every system (nsys not a multiple of 64, the last tile partial; n = 2 and 3 put the
    meeting row at 0 and 1) against the oracle's plain solve."""
    A, B, C, d = pb.generate_host(pb.GenKind.DDPERM, 300 + m + n, m, nsys, n)
    x, sing = gpu_solve(A, B, C, d, int(layout), form=int(pb.Form.TWISTED))
    assert not sing
    assert rel_err(x, oracle_plain(oracle, A, B, C, d)).max() <= TOL


@pytest.mark.parametrize("m,per_cu", [(2, 4), (3, 2), (5, 2), (6, 2)])
@pytest.mark.parametrize("layout", [pb.Layout.TILED, pb.Layout.REF_GPU])
def test_twisted_locked_workgroups(cuda, oracle, m, per_cu, layout):
    """From per_cu twisted wavefronts of work per CU (floor(32/m) systems each) the twisted
    solve runs four wavefronts per workgroup that meet at a barrier before each row's loads;
    a few systems past that, the last workgroup is partial."""
    import torch
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    nsys, n = per_cu * cus * (32 // m) + 5, 6
    A, B, C, d = pb.generate_host(pb.GenKind.DDPERM, 500 + m, m, nsys, n)
    x, sing = gpu_solve(A, B, C, d, int(layout), form=int(pb.Form.TWISTED))
    assert not sing
    assert rel_err(x, oracle_plain(oracle, A, B, C, d)).max() <= TOL


@pytest.mark.parametrize("kind", [pb.GenKind.EULER, pb.GenKind.COND, pb.GenKind.DIAG])
def test_twisted_other_generators(cuda, oracle, kind):
    A, B, C, d = pb.generate_host(kind, 17, 5, 96, 48)
    x, sing = gpu_solve(A, B, C, d, TILED, form=int(pb.Form.TWISTED))
    xo = oracle_plain(oracle, A, B, C, d)
    assert not sing
    if kind != pb.GenKind.COND:
        assert rel_err(x, xo).max() <= TOL
    else:
        # cond(B) up to 1e6: the oracle is itself an FP64 solve about cond * 1e-16 from the
        # exact solution; the twisted solve (another elimination order) must be as accurate as
        # the oracle against the binary128 truth, within a factor 10
        xt, info = oracle.solve_truth(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
        assert info == 0
        e_tw, e_or = rel_err(x, xt), rel_err(xo, xt)
        print(f"COND twisted vs truth {e_tw.max():.2e}, oracle vs truth {e_or.max():.2e}")
        assert e_tw.max() <= 10 * e_or.max()


def test_twisted_profiled_and_info(cuda):
    """The profiled solve reports the whole solve as local compute; zero block rows in the top
    half, at the meeting row and in the bottom half are then named per system."""
    import torch
    m, nsys, n = 4, 150, 21  # meeting row (n-1)/2 = 10
    A, B, C, d = pb.generate_host(pb.GenKind.DDPERM, 8, m, nsys, n)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, pb.Layout.TILED) as plan:
        plan.set_form(pb.Form.TWISTED)
        assert plan.form == pb.Form.TWISTED
        ta, tb, tc, td = (torch.from_numpy(nat_to_layout(X, TILED)).to(cuda) for X in (A, B, C, d))
        t = pb.solve_noncyclic_profiled(plan, ta, tb, tc, td, m, nsys, n)
        assert t.total > 0 and t.local_compute > 0 and t.reduced_compute == 0
        for s_, q in ((3, 2), (77, 10), (149, 18)):
            A[q, s_] = B[q, s_] = C[q, s_] = 0.0
        ta, tb, tc, td = (torch.from_numpy(nat_to_layout(X, TILED)).to(cuda) for X in (A, B, C, d))
        pb.solve_noncyclic(plan, ta, tb, tc, td, m, nsys, n)
        assert plan.info().nonzero()[0].tolist() == [3, 77, 149]
        assert not plan.check()


def test_twisted_form_rules(cuda):
    """The twisted form needs npart == 1 and m >= 2 (either layout); PLAIN is always allowed."""
    with pb.BtdmaGpuPlan().create(3, 64, 12, 1, pb.Layout.REF_GPU) as plan:
        plan.set_form(pb.Form.TWISTED)
        assert plan.form == pb.Form.TWISTED
        plan.set_form(pb.Form.PLAIN)
    with pb.BtdmaGpuPlan().create(3, 64, 12, 2, pb.Layout.TILED) as plan:
        with pytest.raises(ValueError):
            plan.set_form(pb.Form.TWISTED)
    with pb.BtdmaGpuPlan().create(1, 64, 12, 1, pb.Layout.TILED) as plan:
        with pytest.raises(ValueError):
            plan.set_form(pb.Form.TWISTED)
    with pb.BtdmaGpuPlan().create(3, 64, 12, 1, pb.Layout.TILED) as plan:
        assert plan.form == pb.Form.PLAIN
        with pytest.raises(ValueError):
            plan.set_form(7)
