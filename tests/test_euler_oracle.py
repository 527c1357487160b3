"""CPU pins of the euler3d sweep-assembly restatement (oracle/euler_assembly.py).

* The three flux Jacobians (compute_Amat/Bmat/Cmat, euler3d_mpi_gpu.f90:749-801) equal the
  central finite-difference Jacobians of the Euler flux (compute_fluxes_kernel :682-693):
  a wrong sign, index or term in either restatement fails this.
* The sweep numbering equals the reference's transpose kernels (:931-1017), restated here as
  explicit loops over Fortran indices.
* The assembled blocks have the ADI structure: B = (1 + 2 cd) I, and A + C = -2 cd I where the
  state is uniform along the sweep.
"""
import numpy as np
import pytest

from oracle import euler_assembly as ea


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_jacobian_is_the_flux_derivative(axis):
    U = ea.vortex_field(6, 5, 4, seed=3, amp=0.2)
    worst = 0.0
    for (k, j, i) in [(1, 1, 1), (2, 3, 4), (4, 5, 6), (3, 2, 5)]:
        Uc = U[:, k, j, i].copy()
        J = ea.jacobian(*ea.primitive(Uc), axis)
        for jv in range(5):
            h = 1e-6 * max(1.0, abs(Uc[jv]))
            up, um = Uc.copy(), Uc.copy()
            up[jv] += h
            um[jv] -= h
            col = (ea.flux(up, axis) - ea.flux(um, axis)) / (2 * h)
            worst = max(worst, np.abs(col - J[:, jv]).max())
    assert worst < 1e-8


def test_sweep_numbering_is_the_reference_transposes():
    n1, n2, n3, m = 4, 3, 5, 2
    rng = np.random.default_rng(0)
    # Fortran F(n1, n2, n3, m) -> numpy [k, j, i, v] grid form
    F = rng.standard_normal((n3, n2, n1, m))
    for axis in range(3):
        X = ea.to_sweep(F, axis)  # (nrow, nsys, m)
        nrow, nsys = ea.sweep_shape(axis, n1, n2, n3)
        assert X.shape == (nrow, nsys, m)
        for i in range(n1):
            for j in range(n2):
                for k in range(n3):
                    # transposed Fortran arrays: x: (n2, n3, n1, m) at (j, k, i); y: (n1, n3, n2, m)
                    # at (i, k, j); z: untransposed (n1, n2, n3, m) at (i, j, k)
                    if axis == 0:
                        lin = j + n2 * (k + n3 * i)
                    elif axis == 1:
                        lin = i + n1 * (k + n3 * j)
                    else:
                        lin = i + n1 * (j + n2 * k)
                    row, sys_ = divmod(lin, nsys)
                    np.testing.assert_array_equal(X[row, sys_], F[k, j, i])
        np.testing.assert_array_equal(ea.from_sweep(X, axis, n1, n2, n3), F)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_assembled_blocks_have_the_adi_structure(axis):
    n1, n2, n3 = 5, 4, 3
    h, dt, eps = 0.1, 0.05, 0.02
    cd = eps * dt / (2 * h * h)
    U = ea.vortex_field(n1, n2, n3, seed=1)
    dU = np.random.default_rng(2).standard_normal((5, n3, n2, n1))
    A, B, C, d = ea.assemble(U, dU, axis, h, dt, eps)
    np.testing.assert_array_equal(B, np.broadcast_to((1 + 2 * cd) * np.eye(5), B.shape))
    # uniform state along the sweep: the two Jacobians coincide
    Uu = U.copy()
    sl = [slice(None)] * 4
    for t in range(Uu.shape[3 - axis]):
        sl[3 - axis] = t
        src = [slice(None)] * 4
        src[3 - axis] = 1
        Uu[tuple(sl)] = U[tuple(src)]
    Au, _, Cu, _ = ea.assemble(Uu, dU, axis, h, dt, eps)
    np.testing.assert_allclose(Au + Cu, np.broadcast_to(-2 * cd * np.eye(5), Au.shape), atol=1e-15)
    # d is dU in the sweep order
    np.testing.assert_array_equal(d, ea.to_sweep(np.moveaxis(dU, 0, -1), axis))
