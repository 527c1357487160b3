"""Pins of the weno2d assembly oracle (oracle/weno_assembly.py) by what the mathematics fixes:
linear data give WENO5's linear weights and the fifth-order upwind coefficients, a step
gives the smooth stencil nearly all the weight, the coefficients of a reconstruction sum to
one, (I - gamma dt L) preserves constants, and the m = 3 block form of a line is the dense
periodic 7-point operator of its cells."""
import numpy as np
import pytest

from oracle import weno_assembly as wa

LINEAR_LEFT = (1 / 30, -13 / 60, 47 / 60, 9 / 20, -1 / 20)


def test_linear_data_give_the_linear_scheme():
    f = 0.3 + 0.7 * np.arange(-2, 4)  # cells -2..3 of a linear function
    left = wa.weno5_left(*f[0:5])
    right = wa.weno5_right(*f[1:6])
    np.testing.assert_allclose(left, LINEAR_LEFT, rtol=0, atol=1e-14)
    np.testing.assert_allclose(right, LINEAR_LEFT[::-1], rtol=0, atol=1e-14)


def test_step_weights_the_smooth_stencil():
    # cells -2..2 = 0, 0, 0, 1, 1: only the left-most three-cell stencil is smooth
    c = wa.weno5_left(0.0, 0.0, 0.0, 1.0, 1.0)
    np.testing.assert_allclose(c, (1 / 3, -7 / 6, 11 / 6, 0.0, 0.0), atol=1e-9)
    # mirrored: cells -1..3 = 1, 1, 0, 0, 0 seen from the right
    c = wa.weno5_right(1.0, 1.0, 0.0, 0.0, 0.0)
    np.testing.assert_allclose(c, (0.0, 0.0, 11 / 6, -7 / 6, 1 / 3), atol=1e-9)


@pytest.mark.parametrize("fn", [wa.weno5_left, wa.weno5_right])
def test_coefficients_sum_to_one(fn):
    rng = np.random.default_rng(1)
    f = rng.uniform(-1, 1, (5, 1000))
    np.testing.assert_allclose(np.sum(fn(*f), axis=0), 1.0, rtol=0, atol=1e-14)


def field(rng, n1, n2):
    """A periodic field with its three-cell halo, (n2 + 6, n1 + 6)."""
    core = rng.uniform(-1, 1, (n2, n1))
    return np.pad(core, wa.NHALO, mode="wrap")


@pytest.mark.parametrize("axis", [0, 1])
def test_blocks_are_the_dense_line_operator(axis):
    rng = np.random.default_rng(2 + axis)
    n1, n2 = 12, 9
    U = field(rng, n1, n2)
    nsys = n2 if axis == 0 else n1
    vel = rng.uniform(-2, 2, nsys)
    vel[0], vel[1] = 0.0, -0.5  # c = 0 takes the left-biased stencil
    A, B, C = wa.assemble(U, vel, axis, n1, n2, 1.0 / 12, 0.01)
    T = wa.dense_operator(U, vel, axis, n1, n2, 1.0 / 12, 0.01)
    nrow = A.shape[0]
    for s in range(nsys):
        D = np.zeros((3 * nrow, 3 * nrow))
        for r in range(nrow):
            blk = slice(3 * r, 3 * r + 3)
            D[blk, blk] += B[r, s]
            D[blk, slice(3 * ((r - 1) % nrow), 3 * ((r - 1) % nrow) + 3)] += A[r, s]
            D[blk, slice(3 * ((r + 1) % nrow), 3 * ((r + 1) % nrow) + 3)] += C[r, s]
        np.testing.assert_allclose(D, T[s], rtol=0, atol=1e-15)
        # constants are preserved: every row of I - gamma dt L sums to one
        np.testing.assert_allclose(D.sum(axis=1), 1.0, rtol=0, atol=1e-13)
        # upwinding: vel >= 0 reaches three cells back and two forward, else two back, three forward
        i = np.arange(3 * nrow)
        far = (i + 3) % (3 * nrow) if vel[s] >= 0 else (i - 3) % (3 * nrow)
        assert np.all(T[s][i, far] == 0.0)


def test_lines_follow_the_reference_numbering():
    n1, n2 = 6, 3
    U = np.zeros((n2 + 6, n1 + 6))
    U[3 + 1, 3 + 4] = 7.0  # Fortran U(5, 2): cell i = 5 of row j = 2
    assert wa.lines(U, 0, n1, n2)[1, 5 + 2] == 7.0   # x sweep: system j - 1, cell i + 2
    assert wa.lines(U, 1, n1, n2)[4, 2 + 2] == 7.0   # y sweep: system i - 1, cell j + 2
