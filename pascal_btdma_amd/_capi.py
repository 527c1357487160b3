"""ctypes binding of libpascal_btdma_amd.so (declared in include/pascal_btdma_amd.h).

This is the binding a Python caller of the reference's C++ interface would add: the
library's entry points take plain device pointers and sizes, so torch tensors are passed
by data_ptr() and HIP streams by their handle. Loading fails loudly when the library has
not been built; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpascal_btdma_amd.so")

# Every symbol include/pascal_btdma_amd.h declares, with (restype, argtypes).
_vp = ctypes.c_void_p
_i = ctypes.c_int
_u64 = ctypes.c_uint64
_dp = ctypes.c_void_p  # device or host double* passed as an address
_d = ctypes.c_double


class BtdTimings(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "total", "local_compute", "forward_exchange", "reduced_compute",
        "backward_exchange", "update_compute")]


SIGNATURES = {
    "btd_plan_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _i, _i]),
    "btd_plan_destroy": (_i, [_vp]),
    "btd_solve_noncyclic": (_i, [_vp, _dp, _dp, _dp, _dp, _i, _i, _i, _vp]),
    "btd_solve_noncyclic_profiled": (_i, [_vp, _dp, _dp, _dp, _dp, _i, _i, _i,
                                          ctypes.POINTER(BtdTimings), _vp]),
    "btd_solve_cyclic": (_i, [_vp, _dp, _dp, _dp, _dp, _i, _i, _i, _vp]),
    "btd_solve_noncyclic_save": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "btd_solve_noncyclic_save_out": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "btd_solve_adjoint": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "btd_solve_noncyclic_f32r": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _vp]),
    "btd_solve_noncyclic_f32s": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "btd_solve_transposed_f32r": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _vp]),
    "btd_outer_grads_f32": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "btd_euler_assemble": (_i, [_vp, _i, _dp, _dp, _i, _i, _i, _d, _d, _d, _dp, _dp, _dp, _dp, _vp]),
    "btd_weno_assemble": (_i, [_vp, _i, _dp, _dp, _i, _i, _d, _d, _dp, _dp, _dp, _vp]),
    "btd_weno_sweep_cyclic": (_i, [_vp, _i, _dp, _dp, _i, _i, _d, _d, _dp, _dp, _vp]),
    "btd_weno_gather": (_i, [_vp, _i, _dp, _i, _i, _dp, _vp]),
    "btd_weno_scatter": (_i, [_vp, _i, _dp, _i, _i, _dp, _vp]),
    "btd_euler_sweep_cyclic": (_i, [_vp, _i, _dp, _dp, _i, _i, _i, _d, _d, _d, _dp, _dp, _vp]),
    "btd_sweep_gather": (_i, [_vp, _i, _dp, _i, _i, _i, _dp, _vp]),
    "btd_sweep_scatter": (_i, [_vp, _i, _dp, _i, _i, _i, _dp, _vp]),
    "btd_plan_check": (_i, [_vp, ctypes.POINTER(_i)]),
    "btd_auto_npart": (_i, [_i, _i, _i]),
    "btd_plan_npart": (_i, [_vp, ctypes.POINTER(_i)]),
    "btd_plan_set_form": (_i, [_vp, _i]),
    "btd_plan_form": (_i, [_vp, ctypes.POINTER(_i)]),
    "btd_plan_place_workspace": (_i, [_vp, _dp, _dp, _dp, _dp, _i, _i, _i, _i, ctypes.POINTER(_d), _vp]),
    "btd_plan_place_workspace_f32s": (_i, [_vp, _dp, _dp, _dp, _dp, _i, ctypes.POINTER(_d), _vp]),
    "btd_plan_info": (_i, [_vp, _vp]),
    "btd_last_error": (ctypes.c_char_p, []),
    "btd_matrix_index": (ctypes.c_size_t, [_i, _i, _i, _i, _i, _i, _i]),
    "btd_vector_index": (ctypes.c_size_t, [_i, _i, _i, _i, _i]),
    "btd_partition_1d": (None, [_i, _i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "btd_layout_elems": (ctypes.c_size_t, [_i, _i, _i, _i]),
    "btd_convert_layout": (_i, [_i, _i, _i, _i, _i, _dp, _dp, _vp]),
    "btd_generate": (_i, [_i, _u64, _i, _i, _i, _i, _dp, _dp, _dp, _dp, _vp]),
    "btd_generate_f32": (_i, [_i, _u64, _i, _i, _i, _i, _dp, _dp, _dp, _dp, _vp]),
    "btd_generate_host": (_i, [_i, _u64, _i, _i, _i, _i, _i, _dp, _dp, _dp, _dp]),
    "btd_version": (ctypes.c_char_p, []),
    "btd_batch_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "btd_batch_destroy": (_i, [_vp]),
    "btd_batch_sizes": (_i, [_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                             ctypes.POINTER(_i)]),
    "btd_batch_locate": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_size_t),
                              ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_i)]),
    "btd_batch_set_params": (_i, [_vp, _i, _i]),
    "btd_batch_generate": (_i, [_vp, _i, _u64, _dp, _dp, _dp, _dp, _vp]),
    "btd_batch_solve": (_i, [_vp, _dp, _dp, _dp, _dp, _vp]),
    "btd_batch_check": (_i, [_vp, ctypes.POINTER(_i)]),
    "btd_graph_begin": (_i, [_vp]),
    "btd_graph_end": (_i, [_vp, ctypes.POINTER(_vp)]),
    "btd_graph_launch": (_i, [_vp, _vp]),
    "btd_graph_destroy": (_i, [_vp]),
    "btd_solve_noncyclic_noalloc": (_i, [_dp, _dp, _dp, _dp, _i, _i, _i, _i, _i, _vp]),
}

_lib = None


def lib():
    """Load the library (once) and attach the declared signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m pascal_btdma_amd.build` "
                "(the HIP library is required; there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().btd_last_error()
    return msg.decode() if msg else ""
