"""CPU restatement of the euler3d ADI sweep assembly (Fortran_Original/examples/euler3d/
euler3d_mpi_gpu.f90): the flux Jacobians, the three build_matrix kernels and the transposes
that reorder a sweep's arrays so that the sweep direction is the block-row index.

TEST INFRASTRUCTURE ONLY: imported by the test suite and by bench.py's checks, as the
checker. The product package never imports it.

Array conventions (numpy, C order):
* U: the conserved field with its one-cell halo, shape (5, n3+2, n2+2, n1+2) = [v, k, j, i]
  (byte-identical to Fortran U(0:n1+1, 0:n2+1, 0:n3+1, 5)).
* dU: a field without halo, shape (5, n3, n2, n1) (Fortran dU(n1, n2, n3, 5)).
* Blocks on the grid: (n3, n2, n1, 5, 5) = [k, j, i, iv, jv].
* A sweep's system in natural form: blocks (n, nsys, 5, 5) = [row, sys, iv, jv], vectors
  (n, nsys, 5), with the reference's system numbering (see to_sweep).
"""
from __future__ import annotations

import numpy as np

GAMMA = 1.4  # euler3d_mpi_gpu.f90:15


def primitive(Uc):
    """(rho, u, v, w, E) from the conserved (rho, rho u, rho v, rho w, rho E), as
    build_matrix_*_kernel does before each Jacobian (:826-828)."""
    rho = Uc[0]
    return rho, Uc[1] / rho, Uc[2] / rho, Uc[3] / rho, Uc[4] / rho


def flux(Uc, axis):
    """Euler flux along x (0), y (1) or z (2), compute_fluxes_kernel (:682-693)."""
    rho, rhou, rhov, rhow, rhoE = Uc
    uu, vv, ww = rhou / rho, rhov / rho, rhow / rho
    p = (GAMMA - 1.0) * (rhoE - 0.5 * rho * (uu * uu + vv * vv + ww * ww))
    vel = (uu, vv, ww)[axis]
    mom = [rhou * vel, rhov * vel, rhow * vel]
    mom[axis] = mom[axis] + p
    return np.array([(rhou, rhov, rhow)[axis], *mom, (rhoE + p) * vel])


def jacobian(rho, uu, vv, ww, E, axis):
    """Flux Jacobian along `axis`: compute_Amat (:749-765), compute_Bmat (:767-783),
    compute_Cmat (:785-801). Returns J[..., iv, jv] (0-based) over the input's shape."""
    q2 = uu * uu + vv * vv + ww * ww
    H = E + (GAMMA - 1.0) * (E - 0.5 * q2)
    c1 = GAMMA - 1.0
    c2 = 0.5 * c1
    J = np.zeros(np.shape(rho) + (5, 5))
    if axis == 0:
        J[..., 0, 1] = 1.0
        J[..., 1, 0] = c2 * q2 - uu * uu
        J[..., 1, 1] = (3.0 - GAMMA) * uu
        J[..., 1, 2] = -c1 * vv
        J[..., 1, 3] = -c1 * ww
        J[..., 1, 4] = c1
        J[..., 2, 0] = -uu * vv
        J[..., 2, 1] = vv
        J[..., 2, 2] = uu
        J[..., 3, 0] = -uu * ww
        J[..., 3, 1] = ww
        J[..., 3, 3] = uu
        J[..., 4, 0] = (c2 * q2 - H) * uu
        J[..., 4, 1] = H - c1 * (uu * uu)
        J[..., 4, 2] = -c1 * uu * vv
        J[..., 4, 3] = -c1 * uu * ww
        J[..., 4, 4] = GAMMA * uu
    elif axis == 1:
        J[..., 0, 2] = 1.0
        J[..., 1, 0] = -uu * vv
        J[..., 1, 1] = vv
        J[..., 1, 2] = uu
        J[..., 2, 0] = c2 * q2 - vv * vv
        J[..., 2, 1] = -c1 * uu
        J[..., 2, 2] = (3.0 - GAMMA) * vv
        J[..., 2, 3] = -c1 * ww
        J[..., 2, 4] = c1
        J[..., 3, 0] = -vv * ww
        J[..., 3, 2] = ww
        J[..., 3, 3] = vv
        J[..., 4, 0] = (c2 * q2 - H) * vv
        J[..., 4, 1] = -c1 * uu * vv
        J[..., 4, 2] = H - c1 * (vv * vv)
        J[..., 4, 3] = -c1 * vv * ww
        J[..., 4, 4] = GAMMA * vv
    else:
        J[..., 0, 3] = 1.0
        J[..., 1, 0] = -uu * ww
        J[..., 1, 1] = ww
        J[..., 1, 3] = uu
        J[..., 2, 0] = -vv * ww
        J[..., 2, 2] = ww
        J[..., 2, 3] = vv
        J[..., 3, 0] = c2 * q2 - ww * ww
        J[..., 3, 1] = -c1 * uu
        J[..., 3, 2] = -c1 * vv
        J[..., 3, 3] = (3.0 - GAMMA) * ww
        J[..., 3, 4] = c1
        J[..., 4, 0] = (c2 * q2 - H) * ww
        J[..., 4, 1] = -c1 * uu * ww
        J[..., 4, 2] = -c1 * vv * ww
        J[..., 4, 3] = H - c1 * (ww * ww)
        J[..., 4, 4] = GAMMA * ww
    return J


def build_grid(U, axis, h, dt, eps):
    """build_matrix_{z,y,x}_kernel (:803-929) on the whole grid: (trA, trB, trC) as
    (n3, n2, n1, 5, 5) arrays, with
      trA = -cc J(U at the previous cell along the axis) - cd I,
      trB = (1 + 2 cd) I,
      trC =  cc J(U at the next cell) - cd I,   cc = dt / (4h), cd = eps dt / (2h^2).
    Every square is written as a product (x * x), the order of operations is the
    reference's, and nothing is fused, so a device restatement without FMA contraction
    reproduces these values bit for bit."""
    n3, n2, n1 = (s - 2 for s in U.shape[1:])
    cc = dt / (4.0 * h)
    cd = eps * dt / (2.0 * (h * h))
    ax = 3 - axis  # numpy axis of the sweep direction in U[v, k, j, i]

    def neighbour(delta):
        sl = [slice(None), slice(1, n3 + 1), slice(1, n2 + 1), slice(1, n1 + 1)]
        sl[ax] = slice(sl[ax].start + delta, sl[ax].stop + delta)
        return U[tuple(sl)]

    Jl = jacobian(*primitive(neighbour(-1)), axis)
    Jr = jacobian(*primitive(neighbour(+1)), axis)
    eye = np.eye(5)
    trA = -cc * Jl - cd * eye
    trB = np.broadcast_to((1.0 + 2.0 * cd) * eye, trA.shape).copy()
    trC = cc * Jr - cd * eye
    return trA, trB, trC


def to_sweep(F, axis):
    """Grid array (n3, n2, n1, ...) = [k, j, i, ...] to the sweep's (n, nsys, ...) with the
    reference's numbering: z sweep rows k, sys = i + n1 j (no transpose); y sweep rows j,
    sys = i + n1 k (transpose_jk_kernel :931-943, transpose_jk_5d_kernel :945-957); x sweep
    rows i, sys = j + n2 k (transpose_ik_kernel :976-988, transpose_ik_5d_kernel :990-1002)."""
    n3, n2, n1 = F.shape[:3]
    rest = F.shape[3:]
    if axis == 2:
        G = F
    elif axis == 1:
        G = np.swapaxes(F, 0, 1)          # (j, k, i, ...)
    else:
        G = np.moveaxis(F, 2, 0)          # (i, k, j, ...)
    return np.ascontiguousarray(G.reshape((G.shape[0], G.shape[1] * G.shape[2]) + rest))


def from_sweep(X, axis, n1, n2, n3):
    """Inverse of to_sweep for vectors (n, nsys, 5) -> (n3, n2, n1, 5)
    (transpose_kj_kernel :959-974, transpose_ki_kernel :1004-1017)."""
    m = X.shape[-1]
    if axis == 2:
        G = X.reshape(n3, n2, n1, m)
    elif axis == 1:
        G = np.swapaxes(X.reshape(n2, n3, n1, m), 0, 1)
    else:
        G = np.moveaxis(X.reshape(n1, n3, n2, m), 0, 2)
    return np.ascontiguousarray(G)


def assemble(U, dU, axis, h, dt, eps):
    """The sweep's block system as the reference hands it to btdma_many_cycl_mpi_gpu_v2
    (:1276, :1296, :1320): natural-form A, B, C (n, nsys, 5, 5) and d (n, nsys, 5), d being
    the right-hand side field dU reordered like the blocks."""
    blocks = [to_sweep(T, axis) for T in build_grid(U, axis, h, dt, eps)]
    d = to_sweep(np.moveaxis(dU, 0, -1), axis)
    return (*blocks, d)


def sweep_shape(axis, n1, n2, n3):
    """(nrow, nsys) of the sweep along `axis`."""
    return ((n1, n2 * n3), (n2, n1 * n3), (n3, n1 * n2))[axis]


def vortex_field(n1, n2, n3, seed=0, amp=0.05):
    """A smooth periodic test state with its periodic halo: rho in [0.8, 1.2], velocities
    of order 0.3, pressure near 1 (positive), plus a seeded perturbation of size amp."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n1) + 0.5) / n1
    y = (np.arange(n2) + 0.5) / n2
    z = (np.arange(n3) + 0.5) / n3
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    two_pi = 2 * np.pi
    rho = 1.0 + 0.2 * np.sin(two_pi * X) * np.cos(two_pi * Y)
    u = 0.3 * np.sin(two_pi * Y) + amp * rng.uniform(-1, 1, X.shape)
    v = 0.3 * np.cos(two_pi * Z) + amp * rng.uniform(-1, 1, X.shape)
    w = 0.2 * np.sin(two_pi * X) + amp * rng.uniform(-1, 1, X.shape)
    p = 1.0 + 0.1 * np.cos(two_pi * (X + Y + Z))
    E = p / ((GAMMA - 1.0) * rho) + 0.5 * (u * u + v * v + w * w)
    inner = np.stack([rho, rho * u, rho * v, rho * w, rho * E])
    return periodic_halo(inner)


def periodic_halo(inner):
    """(5, n3, n2, n1) -> (5, n3+2, n2+2, n1+2) with periodic halo cells (the one-rank case
    of the reference's halo exchange)."""
    return np.ascontiguousarray(np.pad(inner, ((0, 0), (1, 1), (1, 1), (1, 1)), mode="wrap"))
