"""euler3d ADI sweep assembly on the device, through the C ABI.

Assembly is compared bit for bit with the restatement in oracle/euler_assembly.py (same
operations, no FMA contraction on either side); a whole periodic sweep (assembly, the
periodic solve, scatter back to the field) is compared with the oracle's assembly and
periodic solve at <= 1e-12 per-system relative error.
"""
import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout, layout_to_mat, layout_to_vec, rel_err
from oracle import euler_assembly as ea

pytestmark = pytest.mark.gpu
H, DT, EPS = 0.05, 0.04, 0.01
GRIDS = [(9, 11, 13), (70, 3, 4), (4, 66, 3)]  # partial tiles; rows over several chunks


def upload(cuda, X):
    import torch
    return torch.from_numpy(np.ascontiguousarray(X).reshape(-1)).to(cuda)


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
def test_assembly_is_bit_exact(cuda, grid, axis, layout):
    import torch
    n1, n2, n3 = grid
    U = ea.vortex_field(n1, n2, n3, seed=axis + 10 * n1)
    rhs = np.random.default_rng(n2).standard_normal((5, n3, n2, n1))
    nrow, nsys = pb.sweep_shape(axis, n1, n2, n3)
    with pb.BtdmaGpuPlan().create(5, nsys, nrow, 1, layout) as plan:
        a, b, c = (torch.full((plan.matrix_elems(),), np.nan, dtype=torch.float64, device=cuda)
                   for _ in range(3))
        d = torch.full((plan.vector_elems(),), np.nan, dtype=torch.float64, device=cuda)
        pb.euler_assemble(plan, axis, upload(cuda, U), upload(cuda, rhs), n1, n2, n3, H, DT, EPS,
                          a, b, c, d)
        # blocks alone (no right-hand side): the same blocks
        a2, b2, c2 = (torch.full_like(a, np.nan) for _ in range(3))
        pb.euler_assemble(plan, axis, upload(cuda, U), None, n1, n2, n3, H, DT, EPS, a2, b2, c2)
        torch.cuda.synchronize()
    A, B, C, dref = ea.assemble(U, rhs, axis, H, DT, EPS)
    lay = int(layout)
    for got, ref in ((a, A), (b, B), (c, C), (a2, A), (b2, B), (c2, C)):
        np.testing.assert_array_equal(layout_to_mat(got.cpu().numpy(), lay, nrow, nsys, 5), ref)
    np.testing.assert_array_equal(layout_to_vec(d.cpu().numpy(), lay, nrow, nsys, 5), dref)


@pytest.mark.parametrize("grid", [(9, 70, 5), (130, 9, 8)], ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("m", [3, 5])
def test_gather_scatter(cuda, grid, axis, m):
    """Along x the move goes through the LDS-tiled line transpose (several x tiles and system
    tiles in the second grid)."""
    import torch
    n1, n2, n3 = grid
    F = np.random.default_rng(axis).standard_normal((m, n3, n2, n1))
    nrow, nsys = pb.sweep_shape(axis, n1, n2, n3)
    for layout in (pb.Layout.TILED, pb.Layout.REF_GPU):
        with pb.BtdmaGpuPlan().create(m, nsys, nrow, 1, layout) as plan:
            d = torch.zeros(plan.vector_elems(), dtype=torch.float64, device=cuda)
            back = torch.zeros(F.size, dtype=torch.float64, device=cuda)
            pb.sweep_gather(plan, axis, upload(cuda, F), n1, n2, n3, d)
            pb.sweep_scatter(plan, axis, d, n1, n2, n3, back)
            torch.cuda.synchronize()
        np.testing.assert_array_equal(layout_to_vec(d.cpu().numpy(), int(layout), nrow, nsys, m),
                                      ea.to_sweep(np.moveaxis(F, 0, -1), axis))
        np.testing.assert_array_equal(back.cpu().numpy().reshape(F.shape), F)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("npart", [1, 3])
def test_periodic_sweep_matches_oracle(cuda, oracle, axis, npart):
    """assemble -> periodic solve -> scatter, as the reference's
    build_matrix -> transposes -> btdma_many_cycl_mpi_gpu_v2 -> transpose back."""
    import torch
    n1, n2, n3 = 12, 10, 9
    U = ea.vortex_field(n1, n2, n3, seed=5 + axis)
    rhs = np.random.default_rng(7 + axis).standard_normal((5, n3, n2, n1))
    nrow, nsys = pb.sweep_shape(axis, n1, n2, n3)
    with pb.BtdmaGpuPlan().create(5, nsys, nrow, npart, pb.Layout.TILED) as plan:
        a, b, c = (torch.empty(plan.matrix_elems(), dtype=torch.float64, device=cuda) for _ in range(3))
        d = torch.empty(plan.vector_elems(), dtype=torch.float64, device=cuda)
        out = torch.zeros(rhs.size, dtype=torch.float64, device=cuda)
        pb.euler_assemble(plan, axis, upload(cuda, U), upload(cuda, rhs), n1, n2, n3, H, DT, EPS,
                          a, b, c, d)
        pb.solve_cyclic(plan, a, b, c, d, 5, nsys, nrow)
        pb.sweep_scatter(plan, axis, d, n1, n2, n3, out)
        torch.cuda.synchronize()
        assert not plan.check()
    A, B, C, dref = ea.assemble(U, rhs, axis, H, DT, EPS)
    if npart == 1:
        x, info = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), dref)
    else:
        x, info = oracle.solve_cyclic_two_stage(cpu_layout(A), cpu_layout(B), cpu_layout(C), dref, npart)
    assert info == 0
    xf = out.cpu().numpy().reshape(5, n3, n2, n1)
    xs = ea.to_sweep(np.moveaxis(xf, 0, -1), axis)
    assert rel_err(xs, x).max() <= 1e-12
    # and the solution solves the periodic system (residual of the dense matrix, a few systems)
    for s_ in (0, nsys // 2, nsys - 1):
        T = oracle.assemble_dense(A, B, C, s_, cyclic=True)
        r = T @ x[:, s_].reshape(-1) - dref[:, s_].reshape(-1)
        assert np.abs(r).max() <= 1e-12 * np.abs(dref[:, s_]).max()


def test_sweep_argument_errors(cuda):
    import torch
    n1, n2, n3 = 4, 5, 6
    U = torch.zeros(5 * 6 * 7 * 8, dtype=torch.float64, device=cuda)
    with pb.BtdmaGpuPlan().create(3, n2 * n3, n1, 1, pb.Layout.TILED) as plan:  # m != 5
        t = torch.zeros(plan.matrix_elems(), dtype=torch.float64, device=cuda)
        with pytest.raises(ValueError):
            pb.euler_assemble(plan, 0, U, None, n1, n2, n3, H, DT, EPS, t, t, t)
    with pb.BtdmaGpuPlan().create(5, n2 * n3, n1, 1, pb.Layout.TILED) as plan:
        t = torch.zeros(plan.matrix_elems(), dtype=torch.float64, device=cuda)
        with pytest.raises(ValueError):  # the plan is the x sweep's, not the y sweep's
            pb.euler_assemble(plan, 1, U, None, n1, n2, n3, H, DT, EPS, t, t, t)
        with pytest.raises(ValueError):  # axis out of range
            pb.euler_assemble(plan, 3, U, None, n1, n2, n3, H, DT, EPS, t, t, t)
        pb.euler_assemble(plan, 0, U.fill_(1.0), None, n1, n2, n3, H, DT, EPS, t, t, t)
        torch.cuda.synchronize()


@pytest.mark.parametrize("grid", [(12, 10, 9), (3, 70, 4), (5, 3, 66), (130, 9, 8)],
                         ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
def test_matrix_free_sweep_matches_oracle(cuda, oracle, grid, axis, layout):
    """btd_euler_sweep_cyclic (blocks computed inside the periodic solve, solution written to
    the field) against the oracle's assembly and periodic solve, <= 1e-12; written in place
    over the rhs field, as the reference overwrites dU. Includes nrow = 3."""
    import torch
    n1, n2, n3 = grid
    U = ea.vortex_field(n1, n2, n3, seed=20 + axis)
    rhs = np.random.default_rng(30 + axis).standard_normal((5, n3, n2, n1))
    nrow, nsys = pb.sweep_shape(axis, n1, n2, n3)
    if nrow < 3:
        pytest.skip("periodic sweep needs nrow >= 3")
    with pb.BtdmaGpuPlan().create(5, nsys, nrow, 1, layout) as plan:
        work = torch.full((pb.euler_sweep_work_elems(plan),), np.nan, dtype=torch.float64, device=cuda)
        f = upload(cuda, rhs)
        pb.euler_sweep_cyclic(plan, axis, upload(cuda, U), f, n1, n2, n3, H, DT, EPS, work, f)
        torch.cuda.synchronize()
        assert not plan.check()
    A, B, C, dref = ea.assemble(U, rhs, axis, H, DT, EPS)
    x, info = oracle.solve_cyclic(cpu_layout(A), cpu_layout(B), cpu_layout(C), dref)
    assert info == 0
    xs = ea.to_sweep(np.moveaxis(f.cpu().numpy().reshape(5, n3, n2, n1), 0, -1), axis)
    assert rel_err(xs, x).max() <= 1e-12


def test_matrix_free_sweep_argument_errors(cuda):
    import torch
    n1, n2, n3 = 6, 5, 4
    U = torch.ones(5 * 8 * 7 * 6, dtype=torch.float64, device=cuda)
    f = torch.zeros(5 * n1 * n2 * n3, dtype=torch.float64, device=cuda)
    with pb.BtdmaGpuPlan().create(5, n2 * n3, n1, 2, pb.Layout.TILED) as plan:  # npart != 1
        w = torch.zeros(pb.euler_sweep_work_elems(plan), dtype=torch.float64, device=cuda)
        with pytest.raises(ValueError):
            pb.euler_sweep_cyclic(plan, 0, U, f, n1, n2, n3, H, DT, EPS, w, f)
    with pb.BtdmaGpuPlan().create(5, n2 * n3, n1, 1, pb.Layout.TILED) as plan:
        with pytest.raises(ValueError):  # workspace too small
            pb.euler_sweep_cyclic(plan, 0, U, f, n1, n2, n3, H, DT, EPS, f, f)
