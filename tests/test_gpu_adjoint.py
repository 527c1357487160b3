"""Adjoint (differentiable solve) on the device, through the C ABI and torch.autograd.

Pins: lambda against the oracle's FP64 solve of the transposed system (rel err <= 1e-12,
inside the <= 1e-10 bar), the banded outer-product gradients against -lambda x^T built
from oracle vectors, and fourth-order central finite differences of the oracle's solve on
64 random entries (tolerance 1e-10; the stencil's own error is about 2e-12 here).
"""
import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout, layout_to_mat, layout_to_vec, nat_to_layout, rel_err

pytestmark = pytest.mark.gpu


def oracle_grads(oracle, A, B, C, d, w):
    x, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    lam, info = oracle.solve_transposed(cpu_layout(A), cpu_layout(B), cpu_layout(C), w)
    assert info == 0
    n = d.shape[0]
    xm = np.concatenate([np.zeros_like(x[:1]), x[:-1]])
    xp = np.concatenate([x[1:], np.zeros_like(x[:1])])
    dA = -lam[..., :, None] * xm[..., None, :]
    dB = -lam[..., :, None] * x[..., None, :]
    dC = -lam[..., :, None] * xp[..., None, :]
    return x, lam, dA, dB, dC


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
def test_adjoint_matches_transposed_oracle(cuda, oracle, m, layout):
    import torch
    nsys, n = 70, 13
    A, B, C, d = pb.generate_host(pb.GenKind.DDPERM, 500 + m, m, nsys, n)
    w = np.random.default_rng(m).uniform(-1, 1, d.shape)
    lay = int(layout)
    up = lambda X: torch.from_numpy(nat_to_layout(X, lay)).to(cuda)
    ta, tb, tc, td, tw = up(A), up(B), up(C), up(d), up(w)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, layout) as plan:
        F = torch.empty_like(tc)
        P = torch.empty(pb.layout_elems(layout, nsys, n, 1), dtype=torch.int64, device=cuda)
        pb.solve_noncyclic_save(plan, ta, tb, tc, td, F, P)
        lam = torch.empty_like(td)
        da, db, dc = (torch.empty_like(ta) for _ in range(3))
        pb.solve_adjoint(plan, ta, tc, F, P, td, tw, lam, da, db, dc)
        torch.cuda.synchronize()
        assert not plan.check()
    x, lo, dAo, dBo, dCo = oracle_grads(oracle, A, B, C, d, w)
    assert rel_err(layout_to_vec(td.cpu().numpy(), lay, n, nsys, m), x).max() <= 1e-12
    assert rel_err(layout_to_vec(lam.cpu().numpy(), lay, n, nsys, m), lo).max() <= 1e-12
    for got, ref in ((da, dAo), (db, dBo), (dc, dCo)):
        g = layout_to_mat(got.cpu().numpy(), lay, n, nsys, m)
        scale = np.abs(ref).max(axis=(0, 2, 3))
        assert (np.abs(g - ref).max(axis=(0, 2, 3)) / scale).max() <= 1e-12
    # ignored blocks get zero gradient
    assert np.all(layout_to_mat(da.cpu().numpy(), lay, n, nsys, m)[0] == 0)
    assert np.all(layout_to_mat(dc.cpu().numpy(), lay, n, nsys, m)[-1] == 0)


@pytest.mark.parametrize("m", [2, 5])
@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
def test_save_out_equals_the_in_place_save(cuda, m, layout):
    """solve_noncyclic_save_out (c and d only read, C' and x to their own arrays) gives the
    in-place solve_noncyclic_save's C', x, factors and pivots bit for bit and leaves c, d as
    they were; c_out = c, x = d is the in-place call itself. The out-of-place outputs start
    NaN-filled: their tiled padding lanes must come back zero (the generated inputs' are), and
    so must c_out's last row (the generator's C of the last row is zero, and the in-place call
    leaves it there)."""
    import torch
    nsys, n = 100, 17
    t = [torch.empty(pb.layout_elems(layout, nsys, n, m * m if k < 3 else m), dtype=torch.float64, device=cuda)
         for k in range(4)]
    pb.generate(pb.GenKind.DDPERM, 41 + m, m, nsys, n, layout, *t)
    a, b, c, d = t
    c0, d0 = c.clone(), d.clone()
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, layout) as plan:
        mk = lambda: (torch.empty_like(c), torch.empty(pb.layout_elems(layout, nsys, n, 1), dtype=torch.int64,
                                                        device=cuda))
        F1, P1 = mk()
        ci, di = c.clone(), d.clone()
        pb.solve_noncyclic_save(plan, a, b, ci, di, F1, P1)
        F2, P2 = mk()
        co, xo = torch.full_like(c, float("nan")), torch.full_like(d, float("nan"))
        pb.solve_noncyclic_save_out(plan, a, b, c, d, co, xo, F2, P2)
        F3, P3 = mk()
        cs, ds = c.clone(), d.clone()
        pb.solve_noncyclic_save_out(plan, a, b, cs, ds, cs, ds, F3, P3)
        torch.cuda.synchronize()
        assert not plan.check()
    assert torch.equal(c, c0) and torch.equal(d, d0)
    for got in ((co, xo, F2, P2), (cs, ds, F3, P3)):
        for k, (g, r) in enumerate(zip(got, (ci, di, F1, P1))):
            if k < 2:
                assert torch.equal(g, r)
            else:  # factors and pivots: the real systems' (their padding lanes are not written)
                lay, E = int(layout), (m * m if k == 2 else 1)
                view = lambda X: layout_to_vec(X.cpu().numpy(), lay, n, nsys, E)
                np.testing.assert_array_equal(view(g), view(r))


def fd_check(oracle, base, w, grads, count, seed, names=("A", "B", "C", "d")):
    """Largest |FD - g| / max(1, |g|) over `count` random entries (A of rows >= 1, C of rows
    <= n-2, B, d), the FD taken of the oracle's FP64 solve of that entry's system, loss
    L = sum(w * x). Blocks: fourth-order central stencil
    (L(-2h) - 8L(-h) + 8L(h) - L(2h)) / 12h with h = 1e-3 (truncation ~h^4, rounding
    ~1e-16 |L| / h; about 2e-12 on these systems); d: L is linear in d, so h = 1 and the
    stencil is exact up to rounding."""
    n, nsys, m = base["d"].shape
    rng = np.random.default_rng(seed)
    worst = 0.0
    for k in range(count):
        name = names[k % len(names)]
        s = int(rng.integers(nsys))
        q = int(rng.integers(1 if name == "A" else 0, n - 1 if name == "C" else n))
        idx = (q, s, int(rng.integers(m))) + (() if name == "d" else (int(rng.integers(m)),))
        h = 1.0 if name == "d" else 1e-3
        vals = {}
        for t in (-2, -1, 1, 2):
            arrs = {kk: v[:, s:s + 1].copy() for kk, v in base.items()}
            arrs[name][(idx[0], 0) + idx[2:]] += t * h
            xs, info = oracle.solve_plain(cpu_layout(arrs["A"]), cpu_layout(arrs["B"]),
                                          cpu_layout(arrs["C"]), arrs["d"])
            assert info == 0
            vals[t] = float((xs[:, 0] * w[:, s]).sum())
        fd = (vals[-2] - 8 * vals[-1] + 8 * vals[1] - vals[2]) / (12 * h)
        g = float(grads[name][idx])
        worst = max(worst, abs(fd - g) / max(1.0, abs(g)))
    return worst


def test_autograd_function_and_finite_differences(cuda, oracle):
    """WENO-like config shape reduced in nsys: m = 3, n = 256, L = sum(w * x); device
    gradients against fourth-order central differences at 1e-10."""
    import torch
    from pascal_btdma_amd.autograd import block_tridiagonal_solve
    m, nsys, n = 3, 64, 256
    lay = pb.Layout.TILED
    A, B, C, d = pb.generate_host(pb.GenKind.EULER, 44, m, nsys, n)
    w = np.random.default_rng(3).uniform(-1, 1, d.shape)
    up = lambda X: torch.from_numpy(nat_to_layout(X, int(lay))).to(cuda)
    ta, tb, tc, td, tw = (up(X).requires_grad_(i < 4) for i, X in enumerate((A, B, C, d, w)))
    plan = pb.BtdmaGpuPlan().create(m, nsys, n, 1, lay)
    x = block_tridiagonal_solve(plan, ta, tb, tc, td)
    (x * tw).sum().backward()
    x_, lo, dAo, dBo, dCo = oracle_grads(oracle, A, B, C, d, w)
    assert rel_err(layout_to_vec(x.detach().cpu().numpy(), int(lay), n, nsys, m), x_).max() <= 1e-12
    assert rel_err(layout_to_vec(td.grad.cpu().numpy(), int(lay), n, nsys, m), lo).max() <= 1e-12
    grads = {"A": layout_to_mat(ta.grad.cpu().numpy(), int(lay), n, nsys, m),
             "B": layout_to_mat(tb.grad.cpu().numpy(), int(lay), n, nsys, m),
             "C": layout_to_mat(tc.grad.cpu().numpy(), int(lay), n, nsys, m),
             "d": layout_to_vec(td.grad.cpu().numpy(), int(lay), n, nsys, m)}

    worst = fd_check(oracle, {"A": A, "B": B, "C": C, "d": d}, w, grads, 64, seed=7)
    print(f"finite-difference max rel err over 64 entries: {worst:.3e}")
    assert worst <= 1e-10
    plan.destroy()


@pytest.mark.parametrize("m", [2, 3, 5])
@pytest.mark.parametrize("layout", [pb.Layout.REF_GPU, pb.Layout.TILED])
def test_f32r_transposed_solve_and_fp32_gradients(cuda, oracle, m, layout):
    """FP32-stored blocks: lambda = T^{-T} w by the transposed FP32 solve plus refinement
    sweeps against the oracle's FP64 transposed solve of the same FP32-rounded blocks
    (<= 1e-12 after 2 sweeps on these well-conditioned rows), and the FP32 outer-product
    gradients against -lambda x^T rounded to FP32 (one FP32 rounding: <= 2^-24 relative)."""
    import torch
    nsys, n = 100, 64
    A, B, C, d = pb.generate_host(pb.GenKind.EULER, 80 + m, m, nsys, n)
    A, B, C = (X.astype(np.float32).astype(np.float64) for X in (A, B, C))
    w = np.random.default_rng(10 + m).uniform(-1, 1, d.shape)
    lay = int(layout)
    f32 = lambda X: torch.from_numpy(nat_to_layout(X, lay).astype(np.float32)).to(cuda)
    f64 = lambda X: torch.from_numpy(nat_to_layout(X, lay)).to(cuda)
    ta, tb, tc, tw, td = f32(A), f32(B), f32(C), f64(w), f64(d)
    x, lo, dAo, dBo, dCo = oracle_grads(oracle, A, B, C, d, w)
    errs = []
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, layout) as plan:
        work = torch.empty(pb.f32r_work_elems(plan), dtype=torch.float32, device=cuda)
        for sweeps in range(4):
            lam = torch.zeros_like(tw)
            pb.solve_transposed_f32r(plan, ta, tb, tc, tw, lam, work, sweeps)
            torch.cuda.synchronize()
            assert not plan.check()
            errs.append(rel_err(layout_to_vec(lam.cpu().numpy(), lay, n, nsys, m), lo).max())
        tx = f64(x)
        da, db, dc = (torch.empty_like(ta) for _ in range(3))
        pb.outer_grads_f32(plan, lam, tx, da, db, dc)
        torch.cuda.synchronize()
    print(f"m={m} {layout.name} transposed f32r err by sweeps: {['%.2e' % e for e in errs]}")
    assert 1e-9 < errs[0] < 1e-4
    assert errs[2] <= 1e-12 and errs[3] <= 1e-12
    for got, ref in ((da, dAo), (db, dBo), (dc, dCo)):
        g = layout_to_mat(got.cpu().numpy().astype(np.float64), lay, n, nsys, m)
        np.testing.assert_allclose(g, ref, rtol=2.0 ** -23, atol=1e-30)
    assert np.all(layout_to_mat(da.cpu().numpy(), lay, n, nsys, m)[0] == 0)
    assert np.all(layout_to_mat(dc.cpu().numpy(), lay, n, nsys, m)[-1] == 0)


def test_f32r_autograd_function_and_finite_differences(cuda, oracle):
    """block_tridiagonal_solve_f32r: d.grad against the oracle's transposed solve of the
    FP32-rounded blocks (<= 1e-12 with 3 sweeps) and against finite differences (1e-10),
    block gradients against -lambda x^T in FP32 and against finite differences within
    FP32 rounding (1e-6)."""
    import torch
    from pascal_btdma_amd.autograd import block_tridiagonal_solve_f32r
    m, nsys, n = 3, 64, 256
    lay = int(pb.Layout.TILED)
    A, B, C, d = pb.generate_host(pb.GenKind.EULER, 45, m, nsys, n)
    A, B, C = (X.astype(np.float32).astype(np.float64) for X in (A, B, C))
    w = np.random.default_rng(4).uniform(-1, 1, d.shape)
    f32 = lambda X: torch.from_numpy(nat_to_layout(X, lay).astype(np.float32)).to(cuda).requires_grad_()
    ta, tb, tc = f32(A), f32(B), f32(C)
    td = torch.from_numpy(nat_to_layout(d, lay)).to(cuda).requires_grad_()
    tw = torch.from_numpy(nat_to_layout(w, lay)).to(cuda)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, pb.Layout.TILED) as plan:
        x = block_tridiagonal_solve_f32r(plan, ta, tb, tc, td, 3)
        (x * tw).sum().backward()
        torch.cuda.synchronize()
    x_, lo, dAo, dBo, dCo = oracle_grads(oracle, A, B, C, d, w)
    assert rel_err(layout_to_vec(x.detach().cpu().numpy(), lay, n, nsys, m), x_).max() <= 1e-12
    assert rel_err(layout_to_vec(td.grad.cpu().numpy(), lay, n, nsys, m), lo).max() <= 1e-12
    assert ta.grad.dtype == torch.float32 and tb.grad.dtype == torch.float32
    grads = {"A": layout_to_mat(ta.grad.cpu().numpy().astype(np.float64), lay, n, nsys, m),
             "B": layout_to_mat(tb.grad.cpu().numpy().astype(np.float64), lay, n, nsys, m),
             "C": layout_to_mat(tc.grad.cpu().numpy().astype(np.float64), lay, n, nsys, m),
             "d": layout_to_vec(td.grad.cpu().numpy(), lay, n, nsys, m)}
    for name, ref in (("A", dAo), ("B", dBo), ("C", dCo)):
        np.testing.assert_allclose(grads[name], ref, rtol=2.0 ** -23, atol=1e-30)
    base = {"A": A, "B": B, "C": C, "d": d}
    worst_d = fd_check(oracle, base, w, grads, 16, seed=8, names=("d",))
    worst = fd_check(oracle, base, w, grads, 36, seed=9, names=("A", "B", "C"))
    print(f"f32r finite differences: FP64 d.grad {worst_d:.3e}, FP32 block grads {worst:.3e}")
    assert worst_d <= 1e-10
    assert worst <= 1e-6  # FP32-stored block gradients: one FP32 rounding of each entry


@pytest.mark.parametrize("m", [3, 5])
@pytest.mark.parametrize("variant", ["fp64", "f32r"])
def test_autograd_tiled_padding_is_zero(cuda, oracle, m, variant):
    """nsys = 100 in the tiled layout leaves 28 padding lanes in the last tile. Every
    gradient (and the FP32 path's x) must be finite over the whole flat tensor, with zeros
    in the padding, even when the allocator hands back memory that held NaN."""
    import torch
    from pascal_btdma_amd.autograd import block_tridiagonal_solve, block_tridiagonal_solve_f32r
    nsys, n = 100, 16
    lay = int(pb.Layout.TILED)
    A, B, C, d = pb.generate_host(pb.GenKind.EULER, 90 + m, m, nsys, n)
    if variant == "f32r":
        A, B, C = (X.astype(np.float32).astype(np.float64) for X in (A, B, C))
    w = np.random.default_rng(m).uniform(-1, 1, d.shape)
    bdt = torch.float32 if variant == "f32r" else torch.float64
    up = lambda X, dt: torch.from_numpy(nat_to_layout(X, lay)).to(cuda).to(dt).requires_grad_()
    ta, tb, tc = up(A, bdt), up(B, bdt), up(C, bdt)
    td = up(d, torch.float64)
    tw = torch.from_numpy(nat_to_layout(w, lay)).to(cuda)
    # leave NaN in freed blocks of every size the backward allocates
    for numel, dt in ((ta.numel(), bdt), (td.numel(), torch.float64)):
        junk = [torch.full((numel,), float("nan"), dtype=dt, device=cuda) for _ in range(8)]
        del junk
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, pb.Layout.TILED) as plan:
        if variant == "f32r":
            x = block_tridiagonal_solve_f32r(plan, ta, tb, tc, td, 2)
        else:
            x = block_tridiagonal_solve(plan, ta, tb, tc, td)
        (x * tw).sum().backward()
        torch.cuda.synchronize()
    pad_mask_v = torch.from_numpy(nat_to_layout(np.ones((n, nsys, m)), lay)).to(cuda) == 0
    pad_mask_m = torch.from_numpy(nat_to_layout(np.ones((n, nsys, m, m)), lay)).to(cuda) == 0
    assert int(pad_mask_v.sum()) == n * 28 * m
    for name, g, mask in (("dA", ta.grad, pad_mask_m), ("dB", tb.grad, pad_mask_m),
                          ("dC", tc.grad, pad_mask_m), ("dd", td.grad, pad_mask_v)):
        assert torch.isfinite(g).all(), name
        assert (g[mask] == 0).all(), name
        assert torch.isfinite(g.double().norm()), name
    if variant == "f32r":
        assert torch.isfinite(x).all() and (x.detach()[pad_mask_v] == 0).all()
    x_, lo, _, _, _ = oracle_grads(oracle, A, B, C, d, w)
    assert rel_err(layout_to_vec(td.grad.cpu().numpy(), lay, n, nsys, m), lo).max() <= 1e-12
