"""FP32 block storage with FP64 residual refinement, on the device.

The oracle solves the same FP32-rounded blocks in FP64. Per-system normwise relative
error is reported for every case; the FP64 plain device solve on the same inputs gives
the accuracy floor two FP64 solvers share at that conditioning.
"""
import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout, gpu_solve, layout_to_vec, nat_to_layout, rel_err

pytestmark = pytest.mark.gpu


def run_f32r(cuda, A, B, C, d, layout, sweeps):
    import torch
    n, nsys, m = d.shape
    lay = int(layout)
    f32 = lambda X: torch.from_numpy(nat_to_layout(X, lay).astype(np.float32)).to(cuda)
    ta, tb, tc = f32(A), f32(B), f32(C)
    td = torch.from_numpy(nat_to_layout(d, lay)).to(cuda)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, layout) as plan:
        x = torch.zeros_like(td)
        work = torch.empty(pb.f32r_work_elems(plan), dtype=torch.float32, device=cuda)
        pb.solve_noncyclic_f32r(plan, ta, tb, tc, td, x, work, sweeps)
        torch.cuda.synchronize()
        assert not plan.check()
    return layout_to_vec(x.cpu().numpy(), lay, n, nsys, m)


def fp32_round(*Xs):
    return [X.astype(np.float32).astype(np.float64) for X in Xs]


@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
@pytest.mark.parametrize("m", [2, 3, 5])
def test_refinement_reaches_fp64_on_well_conditioned(cuda, oracle, m, layout):
    A, B, C, d = pb.generate_host(pb.GenKind.EULER, 60 + m, m, 100, 64)
    A, B, C = fp32_round(A, B, C)
    xo, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    errs = [rel_err(run_f32r(cuda, A, B, C, d, layout, s), xo).max() for s in range(4)]
    print(f"m={m} {layout.name} max rel err by sweeps: {['%.2e' % e for e in errs]}")
    assert 1e-9 < errs[0] < 1e-4          # an FP32 solve
    assert errs[2] <= 1e-12 and errs[3] <= 1e-12


def test_refinement_on_conditioned_blocks(cuda, oracle):
    """cond(B_q) log-uniform in [1e1, 1e6]. Accuracy is measured against the binary128 truth
    (the exact solution of the FP32-rounded blocks to about an ulp), since the FP64 oracle is
    itself an FP64 solve about cond * 1e-16 away from it. The contraction factor of FP32-factor
    refinement is about cond * 2^-24, so the error falls by at least 10x per sweep until it
    reaches the FP64 floor, and with the 8 sweeps allowed f32r is within 10x of the FP64 device
    solve's own error against the truth (the bench reports 0-8 sweeps)."""
    m, nsys, n = 5, 128, 64
    A, B, C, d = pb.generate_host(pb.GenKind.COND, 9, m, nsys, n)
    xo, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    xt, info = oracle.solve_truth(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    x64, _ = gpu_solve(A, B, C, d, int(pb.Layout.TILED), form=int(pb.Form.LAPACK_ORDER))
    np.testing.assert_array_equal(x64, xo)  # the one-lane FP64 solve is the oracle bit for bit
    floor = rel_err(x64, xt).max()
    errs = [rel_err(run_f32r(cuda, A, B, C, d, pb.Layout.TILED, s), xt) for s in range(9)]
    mx = [e.max() for e in errs]
    print(f"cond blocks: FP64 solve vs truth {floor:.2e}; f32r vs truth by sweeps "
          f"{['%.2e' % e for e in mx]}")
    for k in range(8):
        assert mx[k + 1] <= 0.1 * mx[k] or mx[k + 1] <= 10 * floor
    assert mx[8] <= 10 * floor


def test_f32r_rejects_bad_arguments(cuda):
    import torch
    with pb.BtdmaGpuPlan().create(3, 64, 8, 2, pb.Layout.TILED) as plan:
        t = torch.zeros(10, dtype=torch.float32, device=cuda)
        with pytest.raises(ValueError):  # npart must be 1
            pb._check(pb._capi.lib().btd_solve_noncyclic_f32r(plan._h, t.data_ptr(), t.data_ptr(),
                                                              t.data_ptr(), t.data_ptr(), t.data_ptr(),
                                                              t.data_ptr(), 2, None))
    with pb.BtdmaGpuPlan().create(3, 64, 8, 1, pb.Layout.TILED) as plan:
        with pytest.raises(ValueError):  # sweeps out of range
            pb._check(pb._capi.lib().btd_solve_noncyclic_f32r(plan._h, t.data_ptr(), t.data_ptr(),
                                                              t.data_ptr(), t.data_ptr(), t.data_ptr(),
                                                              t.data_ptr(), 9, None))


@pytest.mark.parametrize("m", list(range(1, 11)))
@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
def test_fp32_storage_is_the_fp64_solve_of_the_stored_values(cuda, m, layout):
    """btd_solve_noncyclic_f32s (FP32 blocks, FP64 arithmetic, one lane per system) equals,
    bit for bit, the FP64 solve btd_solve_noncyclic of the same blocks converted to FP64 in the
    BTD_FORM_LAPACK_ORDER form (one lane per system in the oracle's operation order, at every
    size and m); the blocks are left as they were. The default form runs other elimination
    orders by m and batch size (m lanes per system, 2m twisted, two lanes per system from both
    ends), which agree to rounding."""
    n = 13
    sizes = (70, 4096 + 5) if m > 1 else (70,)
    for nsys in sizes:
        x32, x64 = _f32s_and_fp64(cuda, m, nsys, n, layout, pb.Form.LAPACK_ORDER)
        assert np.array_equal(x32, x64), f"nsys={nsys}"
    x32, x64 = _f32s_and_fp64(cuda, m, sizes[-1], n, layout)
    assert rel_err(x32, x64).max() <= 1e-12


def _f32s_and_fp64(cuda, m, nsys, n, layout, form=None):
    """x from the FP32-storage solve and from the FP64 solve of the same (FP32-rounded) blocks
    (in `form`, else the plan's default), as natural-form arrays; asserts the FP32 blocks are
    unchanged."""
    import torch
    A, B, C, d = pb.generate_host(pb.GenKind.DDPERM, 300 + m, m, nsys, n)
    A, B, C = fp32_round(A, B, C)
    lay = int(layout)
    f32 = lambda X: torch.from_numpy(nat_to_layout(X, lay).astype(np.float32)).to(cuda)
    ta, tb, tc = f32(A), f32(B), f32(C)
    td = torch.from_numpy(nat_to_layout(d, lay)).to(cuda)
    keep = [t.clone() for t in (ta, tb, tc)]
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, layout) as plan:
        work = torch.empty(plan.matrix_elems(), dtype=torch.float64, device=cuda)
        x32 = td.clone()
        pb.solve_noncyclic_f32s(plan, ta, tb, tc, x32, work)
        a64, b64, c64, x64 = ta.double(), tb.double(), tc.double(), td.clone()
        if form is not None:
            plan.set_form(form)
        pb.solve_noncyclic(plan, a64, b64, c64, x64, m, nsys, n)
        torch.cuda.synchronize()
        assert not plan.check()
    for t, k in zip((ta, tb, tc), keep):
        assert torch.equal(t, k)
    return (layout_to_vec(x32.cpu().numpy(), lay, n, nsys, m),
            layout_to_vec(x64.cpu().numpy(), lay, n, nsys, m))


def test_fp32_storage_on_conditioned_blocks(cuda, oracle):
    """cond(B_q) log-uniform in [1e1, 1e6]: the FP32-storage solve needs no refinement; it is the
    oracle's FP64 solve of the stored values bit for bit, and the twisted FP64 device solve of
    those values is as close to the binary128 truth as the oracle, within a factor 10."""
    import torch
    m, nsys, n = 5, 128, 64
    A, B, C, d = pb.generate_host(pb.GenKind.COND, 9, m, nsys, n)
    xo, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    x64, _ = gpu_solve(A, B, C, d, int(pb.Layout.TILED))
    lay = int(pb.Layout.TILED)
    f32 = lambda X: torch.from_numpy(nat_to_layout(X, lay).astype(np.float32)).to(cuda)
    ta, tb, tc = f32(A), f32(B), f32(C)
    td = torch.from_numpy(nat_to_layout(d, lay)).to(cuda)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, pb.Layout.TILED) as plan:
        work = torch.empty(plan.matrix_elems(), dtype=torch.float64, device=cuda)
        pb.solve_noncyclic_f32s(plan, ta, tb, tc, td, work)
        torch.cuda.synchronize()
    x = layout_to_vec(td.cpu().numpy(), lay, n, nsys, m)
    # the FP32-storage solve computes in the oracle's operation order: its solve of the stored
    # values bit for bit
    np.testing.assert_array_equal(x, xo)
    # 128 systems: the FP64 solve runs twisted (another elimination order); against the
    # binary128 truth it must be as accurate as the oracle, within a factor 10
    xt, _ = oracle.solve_truth(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    e_tw, e_or = rel_err(x64, xt).max(), rel_err(xo, xt).max()
    print(f"cond up to 1e6: oracle vs truth {e_or:.2e}, twisted FP64 vs truth {e_tw:.2e}")
    assert e_tw <= 10 * e_or


@pytest.mark.parametrize("layout", [pb.Layout.TILED, pb.Layout.REF_GPU])
def test_fp32_storage_in_the_placed_workspace(cuda, layout):
    """work=None solves into the plan's workspace (Form.WORKSPACE): after
    place_workspace_f32s, whose timed solves return d unchanged and leave the blocks as they
    were, x equals the solve with a caller's work array bit for bit; without the workspace form
    work=None is refused."""
    import torch
    m, nsys, n = 5, 700, 40
    A, B, C, d = pb.generate_host(pb.GenKind.COND, 31, m, nsys, n)
    A, B, C = fp32_round(A, B, C)
    lay = int(layout)
    f32 = lambda X: torch.from_numpy(nat_to_layout(X, lay).astype(np.float32)).to(cuda)
    ta, tb, tc = f32(A), f32(B), f32(C)
    td = torch.from_numpy(nat_to_layout(d, lay)).to(cuda)
    keep = [t.clone() for t in (ta, tb, tc, td)]
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, layout) as plan:
        with pytest.raises(ValueError, match="BTD_FORM_WORKSPACE"):
            pb.solve_noncyclic_f32s(plan, ta, tb, tc, td.clone())
        ref = td.clone()
        pb.solve_noncyclic_f32s(plan, ta, tb, tc, ref, torch.empty(plan.matrix_elems(), dtype=torch.float64,
                                                                   device=cuda))
        plan.set_form(pb.Form.WORKSPACE)
        ms = plan.place_workspace_f32s(ta, tb, tc, td, candidates=3)
        assert len(ms) == 3 and ms[0] > 0
        torch.cuda.synchronize()
        for t, k in zip((ta, tb, tc, td), keep):
            assert torch.equal(t, k)
        x = td.clone()
        pb.solve_noncyclic_f32s(plan, ta, tb, tc, x)
        torch.cuda.synchronize()
        assert not plan.check()
    assert torch.equal(x, ref)


def test_fp32_storage_rejects_partitions(cuda):
    import torch
    t = torch.zeros(10, dtype=torch.float32, device=cuda)
    with pb.BtdmaGpuPlan().create(3, 64, 8, 2, pb.Layout.TILED) as plan:
        with pytest.raises(ValueError):
            pb._check(pb._capi.lib().btd_solve_noncyclic_f32s(plan._h, t.data_ptr(), t.data_ptr(),
                                                              t.data_ptr(), t.data_ptr(), t.data_ptr(),
                                                              None))
