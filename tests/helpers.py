"""Shared test helpers: numpy layout conversions written independently of the library,
the error metric, and a device round trip through the C ABI."""
from __future__ import annotations

import numpy as np

TILE = 64


def pad(nsys: int) -> int:
    return (nsys + TILE - 1) // TILE * TILE


# natural form: X[row, sys, i, j] (block row i, column j), d[row, sys, i]

def nat_to_layout(X: np.ndarray, layout: int) -> np.ndarray:
    """Flatten natural-form blocks/vectors into a layout (0 REF_GPU, 1 REF_CPU, 2 TILED)."""
    if X.ndim == 4:
        n, nsys, m, _ = X.shape
        E = m * m
        Xe = np.swapaxes(X, -1, -2).reshape(n, nsys, E)  # e = i + m*j (column-major)
    else:
        n, nsys, E = X.shape
        Xe = X
    if layout == 0:
        return np.ascontiguousarray(np.transpose(Xe, (2, 0, 1))).reshape(-1)  # [e][row][sys]
    if layout == 1:
        return np.ascontiguousarray(Xe).reshape(-1)  # [row][sys][e]
    npd = pad(nsys)
    Y = np.zeros((n, npd, E))
    Y[:, :nsys] = Xe
    Y = Y.reshape(n, npd // TILE, TILE, E)
    return np.ascontiguousarray(np.transpose(Y, (1, 0, 3, 2))).reshape(-1)  # [tile][row][e][lane]


def layout_to_vec(flat: np.ndarray, layout: int, n: int, nsys: int, m: int) -> np.ndarray:
    """Inverse of nat_to_layout for vectors: returns (n, nsys, m)."""
    if layout == 0:
        return np.transpose(flat[: m * n * nsys].reshape(m, n, nsys), (1, 2, 0)).copy()
    if layout == 1:
        return flat[: n * nsys * m].reshape(n, nsys, m).copy()
    npd = pad(nsys)
    Y = flat[: npd * n * m].reshape(npd // TILE, n, m, TILE)
    Y = np.transpose(Y, (1, 0, 3, 2)).reshape(n, npd, m)
    return Y[:, :nsys].copy()


def layout_to_mat(flat: np.ndarray, layout: int, n: int, nsys: int, m: int) -> np.ndarray:
    """Inverse of nat_to_layout for blocks: returns (n, nsys, m, m) natural form."""
    v = layout_to_vec(flat, layout, n, nsys, m * m)  # (n, nsys, E) with e = i + m*j
    return np.swapaxes(v.reshape(n, nsys, m, m), -1, -2).copy()


def cpu_layout(X: np.ndarray) -> np.ndarray:
    """natural -> reference CPU layout arrays for the oracle."""
    return np.ascontiguousarray(np.swapaxes(X, -1, -2)) if X.ndim == 4 else np.ascontiguousarray(X)


def rel_err(x: np.ndarray, xref: np.ndarray) -> np.ndarray:
    """Per-system normwise relative error max_rows,i |x - xref| / max |xref| -> (nsys,)."""
    num = np.abs(x - xref).max(axis=(0, 2))
    den = np.abs(xref).max(axis=(0, 2))
    return num / den


def gpu_solve(A, B, C, d, layout: int, npart: int = 1, profiled: bool = False, form: int = 0):
    """Upload natural-form numpy inputs, solve in place through the C ABI, return x (n,nsys,m)
    (and the timings when profiled); `form` is the plan's btd_form."""
    import torch
    import pascal_btdma_amd as pb
    n, nsys, m = d.shape
    dev = torch.device("cuda:0")
    ta = torch.from_numpy(nat_to_layout(A, layout)).to(dev)
    tb = torch.from_numpy(nat_to_layout(B, layout)).to(dev)
    tc = torch.from_numpy(nat_to_layout(C, layout)).to(dev)
    td = torch.from_numpy(nat_to_layout(d, layout)).to(dev)
    with pb.BtdmaGpuPlan().create(m, nsys, n, npart, pb.Layout(layout)) as plan:
        if form:
            plan.set_form(pb.Form(form))
        if profiled:
            t = pb.solve_noncyclic_profiled(plan, ta, tb, tc, td, m, nsys, n)
        else:
            pb.solve_noncyclic(plan, ta, tb, tc, td, m, nsys, n)
            t = None
        torch.cuda.synchronize()
        sing = plan.check()
    x = layout_to_vec(td.cpu().numpy(), layout, n, nsys, m)
    return (x, t, sing) if profiled else (x, sing)
