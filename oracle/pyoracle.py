"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and the
CPU-baseline leg of bench.py, as the checker. The product package never imports it.

Arrays use the reference CPU layout of mod_btdma_cpu.f90:18, held in numpy as
C-ordered arrays of shape (n, nsys, m_col, m_row) for blocks and (n, nsys, m) for
vectors, which is byte-identical to Fortran a(m,m,nsys,n) / d(m,nsys,n).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.ora_dgesv.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int]
        L.ora_para.argtypes = [ctypes.c_int] * 4 + [_ip, _ip]
        L.ora_btdma_many.argtypes = [ctypes.c_int] * 3 + [_dp] * 4
        L.ora_btdma_many_mpi.argtypes = [ctypes.c_int] * 4 + [_dp] * 4
        L.ora_btdma_many_cycl.argtypes = [ctypes.c_int] * 3 + [_dp] * 4
        L.ora_btdma_many_cyclic_mpi.argtypes = [ctypes.c_int] * 4 + [_dp] * 4
        L.ora_btdma_many_t.argtypes = [ctypes.c_int] * 3 + [_dp] * 4
        L.ora_btdma_many_q.argtypes = [ctypes.c_int] * 3 + [_dp] * 5
        L.ora_set_num_threads.argtypes = [ctypes.c_int]
        L.ora_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def set_threads(n: int) -> None:
    lib().ora_set_num_threads(int(n))


def last_threads() -> int:
    return int(lib().ora_num_threads())


def dgesv(a: np.ndarray, b: np.ndarray):
    """LAPACK dgesv on column-major copies. a: (n,n), b: (n,) or (n,k), both row-major numpy.
    Returns (lu (row-major view of the factored matrix), ipiv (1-based), x, info)."""
    n = a.shape[0]
    af = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    bb = b.reshape(n, -1)
    bf = np.asfortranarray(bb, dtype=np.float64).copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().ora_dgesv(n, bf.shape[1], af.ctypes.data_as(_dp), n,
                           ipiv.ctypes.data_as(_ip), bf.ctypes.data_as(_dp), n)
    return af, ipiv, bf.reshape(b.shape), info


def para(sta: int, end: int, rank: int, nprocs: int):
    ia, ib = ctypes.c_int(), ctypes.c_int()
    n = lib().ora_para(sta, end, rank, nprocs, ctypes.byref(ia), ctypes.byref(ib))
    return n, ia.value, ib.value


def _prep(A, B, C, D):
    return [np.ascontiguousarray(x, dtype=np.float64).copy() for x in (A, B, C, D)]


def solve_plain(A, B, C, D):
    """btdma_many on copies; returns x (n, nsys, m) and info."""
    A, B, C, D = _prep(A, B, C, D)
    n, nsys, m = D.shape
    info = lib().ora_btdma_many(n, nsys, m, _p(A), _p(B), _p(C), _p(D))
    return D, info


def solve_two_stage(A, B, C, D, npart: int):
    """btdma_many_mpi with npart in-process row partitions; returns x and info."""
    A, B, C, D = _prep(A, B, C, D)
    n, nsys, m = D.shape
    info = lib().ora_btdma_many_mpi(m, nsys, n, npart, _p(A), _p(B), _p(C), _p(D))
    return D, info


def solve_cyclic(A, B, C, D):
    A, B, C, D = _prep(A, B, C, D)
    n, nsys, m = D.shape
    info = lib().ora_btdma_many_cycl(n, nsys, m, _p(A), _p(B), _p(C), _p(D))
    return D, info


def solve_cyclic_two_stage(A, B, C, D, npart: int):
    """btdma_many_cyclic_mpi with npart in-process row partitions (npart >= 2)."""
    A, B, C, D = _prep(A, B, C, D)
    n, nsys, m = D.shape
    info = lib().ora_btdma_many_cyclic_mpi(m, nsys, n, npart, _p(A), _p(B), _p(C), _p(D))
    return D, info


def solve_truth(A, B, C, D):
    """The same block-Thomas solve carried out in binary128 (btdma_truth.c) and rounded to
    FP64: the exact solution of the given FP64 inputs to about an ulp. Inputs untouched;
    returns x (n, nsys, m) and info."""
    A, B, C, D = _prep(A, B, C, D)
    n, nsys, m = D.shape
    X = np.empty_like(D)
    info = lib().ora_btdma_many_q(n, nsys, m, _p(A), _p(B), _p(C), _p(D), _p(X))
    return X, info


def solve_transposed(A, B, C, W):
    A, B, C, W = _prep(A, B, C, W)
    n, nsys, m = W.shape
    info = lib().ora_btdma_many_t(n, nsys, m, _p(A), _p(B), _p(C), _p(W))
    return W, info


# ------------------------------------------------------------------ layout helpers --
# numpy "natural" form used by tests: blocks X[row, sys, i, j] = X_q(i,j) (row i, col j).
# The CPU layout stores column-major blocks, i.e. numpy (n, nsys, j, i).

def to_cpu_layout(X: np.ndarray) -> np.ndarray:
    """natural (n, nsys, m, m)[i,j] -> reference CPU layout (n, nsys, m, m)[j,i]."""
    return np.ascontiguousarray(np.swapaxes(X, -1, -2))


def from_cpu_layout(X: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.swapaxes(X, -1, -2))


def assemble_dense(A, B, C, sys: int, cyclic: bool = False) -> np.ndarray:
    """Dense (n*m) x (n*m) matrix of one system from natural-form blocks."""
    n, _, m, _ = B.shape
    T = np.zeros((n * m, n * m))
    for q in range(n):
        T[q * m:(q + 1) * m, q * m:(q + 1) * m] = B[q, sys]
        if q > 0:
            T[q * m:(q + 1) * m, (q - 1) * m:q * m] = A[q, sys]
        elif cyclic:
            T[0:m, (n - 1) * m:n * m] += A[0, sys]
        if q < n - 1:
            T[q * m:(q + 1) * m, (q + 1) * m:(q + 2) * m] = C[q, sys]
        elif cyclic:
            T[(n - 1) * m:n * m, 0:m] += C[n - 1, sys]
    return T
