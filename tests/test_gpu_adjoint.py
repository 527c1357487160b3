"""Adjoint (differentiable solve) on the device, through the C ABI and torch.autograd.

Pins: lambda against the oracle's FP64 solve of the transposed system (rel err <= 1e-12,
inside the <= 1e-10 bar), the banded outer-product gradients against -lambda x^T built
from oracle vectors, and central finite differences (h = 1e-6) of the device solve on
64 random entries, whose own error floor (rounding of L / h) sets a 1e-6 tolerance.
"""
import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout, layout_to_vec, nat_to_layout, rel_err

pytestmark = pytest.mark.gpu


def layout_to_mat(flat, layout, n, nsys, m):
    """Inverse of nat_to_layout for blocks: (n, nsys, m, m) natural form."""
    v = layout_to_vec(flat, layout, n, nsys, m * m)  # (n, nsys, E) with e = i + m*j
    return np.swapaxes(v.reshape(n, nsys, m, m), -1, -2).copy()


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


def test_autograd_function_and_finite_differences(cuda, oracle):
    """WENO-like config shape reduced in nsys: m = 3, n = 256, L = sum(w * x)."""
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

    # central differences of the device solve, per-system loss, 64 random entries
    rng = np.random.default_rng(7)
    h = 1e-6
    base = {"A": A, "B": B, "C": C, "d": d}
    worst = 0.0
    with torch.no_grad():
        def loss_sys(arrs, s):
            xs = pb_solve(arrs)
            return float((xs[:, s] * w[:, s]).sum())

        def pb_solve(arrs):
            t = [up(arrs[k]) for k in ("A", "B", "C", "d")]
            with pb.BtdmaGpuPlan().create(m, nsys, n, 1, lay) as p2:
                pb.solve_noncyclic(p2, *t, m, nsys, n)
                torch.cuda.synchronize()
            return layout_to_vec(t[3].cpu().numpy(), int(lay), n, nsys, m)

        for k in range(64):
            name = ["A", "B", "C", "d"][k % 4]
            s = int(rng.integers(nsys))
            q = int(rng.integers(1 if name == "A" else 0, n - 1 if name == "C" else n))
            idx = (q, s, int(rng.integers(m))) + (() if name == "d" else (int(rng.integers(m)),))
            plus = {kk: v.copy() for kk, v in base.items()}
            minus = {kk: v.copy() for kk, v in base.items()}
            plus[name][idx] += h
            minus[name][idx] -= h
            fd = (loss_sys(plus, s) - loss_sys(minus, s)) / (2 * h)
            g = grads[name][idx]
            err = abs(fd - g) / max(1.0, abs(g))
            worst = max(worst, err)
    print(f"finite-difference max rel err over 64 entries: {worst:.3e}")
    assert worst <= 1e-6
    plan.destroy()
