"""CPU restatement of the weno2d implicit-sweep assembly (Fortran_Original/examples/weno2d/
weno_mpi_gpu.f90): the frozen WENO5 stencil coefficients of every cell interface and the
m = 3 block form of (I - gamma dt L) that the x and y sweeps solve.

TEST INFRASTRUCTURE ONLY: imported by the test suite as the checker. The product package
never imports it.

Array conventions (numpy, C order):
* U: the scalar field with its three-cell halo, shape (n2+6, n1+6) = [j, i] (byte-identical
  to Fortran U(1-3:n1+3, 1-3:n2+3)); numpy index = Fortran index + 2.
* vel: the advection speed of each line of the sweep (x sweep: c_x of row j, length n2;
  y sweep: c_y of column i, length n1), weno_mpi_gpu.f90:438, :467.
* A sweep's system in natural form: blocks (nrow, nsys, 3, 3) = [row, sys, iv, jv] with
  nrow = n/3 (three consecutive cells per block row) and sys = j (x sweep) or i (y sweep),
  the numbering of trA(n2sub, n1sub/m, m, m) / trA(n1sub, n2sub/m, m, m) (:504-511, :641-648).
"""
from __future__ import annotations

import numpy as np

M = 3             # cells per block row (weno_mpi_gpu.f90:49)
NHALO = 3         # :51
D_W = (0.1, 0.6, 0.3)   # linear weights d_w0..d_w2 (:77-79)
WENO_EPS = 1.0e-6       # :80


def _weights(b0, b1, b2):
    a0 = D_W[0] / (WENO_EPS + b0) ** 2
    a1 = D_W[1] / (WENO_EPS + b1) ** 2
    a2 = D_W[2] / (WENO_EPS + b2) ** 2
    s = a0 + a1 + a2
    return a0 / s, a1 / s, a2 / s


def weno5_left(fm2, fm1, f0, fp1, fp2):
    """Coefficients of (fm2, fm1, f0, fp1, fp2) in the left-biased WENO5 reconstruction at the
    interface to the right of f0 (weno5_coeffs_left, :381-398)."""
    b0 = (13.0 / 12.0) * (fm2 - 2.0 * fm1 + f0) ** 2 + 0.25 * (fm2 - 4.0 * fm1 + 3.0 * f0) ** 2
    b1 = (13.0 / 12.0) * (fm1 - 2.0 * f0 + fp1) ** 2 + 0.25 * (fm1 - fp1) ** 2
    b2 = (13.0 / 12.0) * (f0 - 2.0 * fp1 + fp2) ** 2 + 0.25 * (3.0 * f0 - 4.0 * fp1 + fp2) ** 2
    w0, w1, w2 = _weights(b0, b1, b2)
    return (w0 / 3.0, -7.0 * w0 / 6.0 - w1 / 6.0, 11.0 * w0 / 6.0 + 5.0 * w1 / 6.0 + w2 / 3.0,
            w1 / 3.0 + 5.0 * w2 / 6.0, -w2 / 6.0)


def weno5_right(fm1, f0, fp1, fp2, fp3):
    """Coefficients of (fm1, f0, fp1, fp2, fp3) in the right-biased (mirrored) WENO5
    reconstruction at the interface to the right of f0 (weno5_coeffs_right, :400-418)."""
    b0 = (13.0 / 12.0) * (fp3 - 2.0 * fp2 + fp1) ** 2 + 0.25 * (fp3 - 4.0 * fp2 + 3.0 * fp1) ** 2
    b1 = (13.0 / 12.0) * (fp2 - 2.0 * fp1 + f0) ** 2 + 0.25 * (fp2 - f0) ** 2
    b2 = (13.0 / 12.0) * (fp1 - 2.0 * f0 + fm1) ** 2 + 0.25 * (3.0 * fp1 - 4.0 * f0 + fm1) ** 2
    w0, w1, w2 = _weights(b0, b1, b2)
    return (-w2 / 6.0, w1 / 3.0 + 5.0 * w2 / 6.0, 11.0 * w0 / 6.0 + 5.0 * w1 / 6.0 + w2 / 3.0,
            -7.0 * w0 / 6.0 - w1 / 6.0, w0 / 3.0)


def lines(U, axis, n1, n2):
    """The sweep's lines with their halos: (nsys, n + 6) [sys, cell], cell index = Fortran
    cell + 2, for the x (axis 0: lines j) or y (axis 1: lines i) sweep."""
    assert U.shape == (n2 + 2 * NHALO, n1 + 2 * NHALO)
    if axis == 0:
        return U[NHALO:NHALO + n2, :]
    return U[:, NHALO:NHALO + n1].T


def interface_coeffs(U, vel, axis, n1, n2):
    """coef(sys, k, 0..4) of interface k + 1/2, k = 0..n (compute_weno_coeffs_{x,y}_kernel,
    :430-485): left-biased cells k-2..k+2 for vel >= 0, right-biased k-1..k+3 otherwise."""
    L = lines(U, axis, n1, n2)
    n = n1 if axis == 0 else n2
    k = np.arange(0, n + 1)
    cell = lambda off: L[:, k + off + 2]  # Fortran cell k + off
    left = np.array(weno5_left(cell(-2), cell(-1), cell(0), cell(1), cell(2)))
    right = np.array(weno5_right(cell(-1), cell(0), cell(1), cell(2), cell(3)))
    pos = (np.asarray(vel) >= 0.0)[None, :, None]
    return np.moveaxis(np.where(pos, left, right), 0, -1)  # (nsys, n+1, 5)


def row_stencil(coef, vel, h, gdt):
    """M[i, i+off], off = -3..3, for every cell i (1-based) of every line: (nsys, n, 7) with
    the diagonal's identity excluded (build_matrix_{x,y}_kernel_weno5, :551-600, :686-705:
    f_hat_{i+1/2} with +1, f_hat_{i-1/2} with -1, times gamma dt c / h)."""
    nsys, n1p, _ = coef.shape
    n = n1p - 1
    vel = np.asarray(vel, dtype=np.float64)
    base = np.where(vel >= 0.0, -2, -1)
    L = np.zeros((nsys, n, 7))
    for s in range(nsys):
        for kk in range(5):
            L[s, :, 3 + base[s] + kk] += coef[s, 1:, kk]           # interface i + 1/2 (index i)
        for kk in range(5):
            L[s, :, 3 - 1 + base[s] + kk] -= coef[s, :n, kk]       # interface i - 1/2 (index i-1)
    scale = (gdt * vel) / h
    return L * scale[:, None, None]


def assemble(U, vel, axis, n1, n2, h, gdt):
    """Blocks A, B, C in natural form (nrow, nsys, 3, 3) of the sweep's periodic system
    (I - gamma dt L) x = rhs, three consecutive cells per block row."""
    n = n1 if axis == 0 else n2
    assert n % M == 0
    coef = interface_coeffs(U, vel, axis, n1, n2)
    Lv = row_stencil(coef, vel, h, gdt)
    nsys, nrow = coef.shape[0], n // M
    A = np.zeros((nrow, nsys, M, M))
    B = np.zeros((nrow, nsys, M, M))
    C = np.zeros((nrow, nsys, M, M))
    B[:, :, np.arange(M), np.arange(M)] = 1.0
    for r in range(nrow):
        for irow in range(M):
            i = r * M + irow  # 0-based cell
            for off in range(-3, 4):
                v = Lv[:, i, off + 3]
                col = irow + off  # 0-based column relative to this block
                if 0 <= col < M:
                    B[r, :, irow, col] += v
                elif col < 0:
                    A[r, :, irow, col + M] += v
                else:
                    C[r, :, irow, col - M] += v
    return A, B, C


def dense_operator(U, vel, axis, n1, n2, h, gdt):
    """(nsys, n, n) dense periodic matrices I - gamma dt L of the scalar lines, built cell by
    cell from the interface coefficients (independent of the block distribution)."""
    n = n1 if axis == 0 else n2
    coef = interface_coeffs(U, vel, axis, n1, n2)
    Lv = row_stencil(coef, vel, h, gdt)
    nsys = coef.shape[0]
    T = np.zeros((nsys, n, n))
    for i in range(n):
        T[:, i, i] = 1.0
        for off in range(-3, 4):
            T[:, i, (i + off) % n] += Lv[:, i, off + 3]
    return T
