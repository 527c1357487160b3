"""pascal_btdma_amd: batched block-tridiagonal (block Thomas) solver for one MI355X.

Python mirror of the reference's C++ interface (CUDA_CXX_Port/include/pascal_btdma_cuda.hpp):
``BtdmaGpuPlan.create / destroy``, ``solve_noncyclic``, ``solve_noncyclic_profiled``,
``solve_cyclic``, ``BtdmaSolveTimings``, ``matrix_index``, ``vector_index`` and
``partition_1d`` keep the reference's names, argument meaning and error behaviour
(``ValueError`` where the reference throws ``std::invalid_argument``, ``RuntimeError``
for ``std::runtime_error``, ``NotImplementedError`` for ``std::logic_error``).
Arrays are torch CUDA tensors (float64) in device memory; all work runs in the HIP
library through its C ABI.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass

from . import _capi

__all__ = [
    "Layout", "Form", "GenKind", "BtdmaGpuPlan", "BtdmaSolveTimings", "BtdmaError", "HipError",
    "SingularError", "solve_noncyclic", "solve_noncyclic_profiled", "solve_cyclic",
    "matrix_index", "vector_index", "partition_1d", "layout_elems", "convert_layout",
    "generate", "generate_f32", "generate_host", "kMaxBlockSize", "solve_noncyclic_save", "solve_noncyclic_save_out",
    "solve_adjoint", "solve_noncyclic_f32r", "f32r_work_elems", "solve_transposed_f32r", "solve_noncyclic_f32s",
    "outer_grads_f32", "euler_assemble", "euler_sweep_cyclic", "weno_assemble", "weno_sweep_cyclic", "weno_gather", "weno_scatter",
    "euler_sweep_work_elems", "sweep_gather", "sweep_scatter", "sweep_shape", "BtdmaBatch", "Graph",
]

kMaxBlockSize = 10


class Layout(enum.IntEnum):
    REF_GPU = 0   # sys + nsys*(row + nrow*e)   (pascal_btdma_cuda.hpp:46-65)
    REF_CPU = 1   # e + E*(sys + nsys*row)      (mod_btdma_cpu.f90:18)
    TILED = 2     # ((tile*nrow + row)*E + e)*64 + lane


class Form(enum.IntEnum):
    """How a plan's single-partition solve runs (btd_form)."""
    PLAIN = 0
    TWISTED = 1
    LAPACK_ORDER = 2
    WORKSPACE = 3


class GenKind(enum.IntEnum):
    DIAG = 0
    DDPERM = 1
    EULER = 2
    COND = 3
    BPERM = 4


class BtdmaError(RuntimeError):
    pass


class HipError(BtdmaError):
    """A HIP runtime call failed (the reference's CudaError)."""


class SingularError(BtdmaError, ArithmeticError):
    """A block LU met a zero pivot."""


_OK, _INVALID, _NOT_CREATED, _DIM, _HIP, _SINGULAR, _NOT_IMPL = range(7)


def _check(status: int) -> None:
    if status == _OK:
        return
    msg = _capi.last_error()
    if status in (_INVALID, _DIM):
        raise ValueError(msg)
    if status == _NOT_CREATED:
        raise RuntimeError(msg)
    if status == _HIP:
        raise HipError(msg)
    if status == _SINGULAR:
        raise SingularError(msg)
    if status == _NOT_IMPL:
        raise NotImplementedError(msg)
    raise BtdmaError(f"status {status}: {msg}")


@dataclass
class BtdmaSolveTimings:
    """Phase times in seconds (pascal_btdma_cuda.hpp:24-39)."""
    total: float = 0.0
    local_compute: float = 0.0
    forward_exchange: float = 0.0
    reduced_compute: float = 0.0
    backward_exchange: float = 0.0
    update_compute: float = 0.0

    def computation(self) -> float:
        return self.local_compute + self.reduced_compute + self.update_compute

    def communication(self) -> float:
        return self.forward_exchange + self.backward_exchange


def _ptr(t) -> int:
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _check_tensor(name, t, nelem, dtype):
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device (cuda) tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < nelem:
        raise ValueError(f"{name} holds {t.numel()} elements, the layout needs {nelem}")


class BtdmaGpuPlan:
    """Plan for a batch of nsys systems of nrow block rows (pascal_btdma_cuda.hpp:88-135).

    ``npart`` replaces the reference's MPI communicator: the rows of every system are
    split into npart partitions (partition_1d) solved by the two-stage method on the
    same GPU. Move-only semantics become: one plan object, explicit destroy, and
    ``with`` support.
    """

    def __init__(self):
        self._h = ctypes.c_void_p(None)
        self.m = self.nsys = self.nrow = 0
        self.npart = 1
        self.layout = Layout.REF_GPU
        self.created = False

    def create(self, m: int, nsys: int, nrow: int, npart: int = 1,
               layout: Layout = Layout.REF_GPU) -> "BtdmaGpuPlan":
        self.destroy()
        h = ctypes.c_void_p(None)
        _check(_capi.lib().btd_plan_create(ctypes.byref(h), int(m), int(nsys), int(nrow),
                                           int(npart), int(layout)))
        self._h = h
        if int(npart) == 0:  # chosen by the library
            p = ctypes.c_int(0)
            _check(_capi.lib().btd_plan_npart(h, ctypes.byref(p)))
            npart = p.value
        self.m, self.nsys, self.nrow, self.npart = int(m), int(nsys), int(nrow), int(npart)
        self.layout = Layout(layout)
        self.created = True
        return self

    def destroy(self) -> None:
        if self._h.value:
            _capi.lib().btd_plan_destroy(self._h)
        self._h = ctypes.c_void_p(None)
        self.created = False

    @property
    def form(self) -> int:
        """How the single-partition solve runs: Form.PLAIN, TWISTED, LAPACK_ORDER or WORKSPACE
        (btd_plan_form)."""
        f = ctypes.c_int(0)
        _check(_capi.lib().btd_plan_form(self._h, ctypes.byref(f)))
        return Form(f.value)

    def set_form(self, form: "Form") -> "BtdmaGpuPlan":
        """btd_plan_set_form: Form.TWISTED needs npart == 1 and m >= 2 (either layout);
        LAPACK_ORDER and WORKSPACE need npart == 1 (WORKSPACE allocates the plan's C' array)."""
        _check(_capi.lib().btd_plan_set_form(self._h, int(form)))
        return self

    def place_workspace(self, a, b, c, d, candidates: int = 8, stream=None) -> list[float]:
        """btd_plan_place_workspace: with Form.WORKSPACE set, time the solve of these arrays with
        up to `candidates` workspace allocations and keep the fastest; d returns unchanged.
        Returns each candidate's time in ms (-1: not allocated; candidate 0 was the plan's)."""
        _validate(self, a, b, c, d, self.m, self.nsys, self.nrow)
        ms = (ctypes.c_double * int(candidates))()
        _check(_capi.lib().btd_plan_place_workspace(self._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d), self.m,
                                                     self.nsys, self.nrow, int(candidates), ms,
                                                     _stream(stream)))
        return [float(x) for x in ms]

    def place_workspace_f32s(self, a, b, c, d, candidates: int = 8, stream=None) -> list[float]:
        """btd_plan_place_workspace_f32s: place_workspace for the FP32-storage solve
        (solve_noncyclic_f32s with work=None): a, b, c float32 blocks, d float64, unchanged."""
        import torch
        nm, nv = self.matrix_elems(), self.vector_elems()
        for name, t, n, dt in (("a", a, nm, torch.float32), ("b", b, nm, torch.float32),
                               ("c", c, nm, torch.float32), ("d", d, nv, torch.float64)):
            _check_tensor(name, t, n, dt)
        ms = (ctypes.c_double * int(candidates))()
        _check(_capi.lib().btd_plan_place_workspace_f32s(self._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                                          int(candidates), ms, _stream(stream)))
        return [float(x) for x in ms]

    def check(self) -> bool:
        """Synchronise; True if any block LU met a zero pivot since the last check."""
        s = ctypes.c_int(0)
        _check(_capi.lib().btd_plan_check(self._h, ctypes.byref(s)))
        return bool(s.value)

    def info(self):
        """Synchronise; int32 numpy array of nsys entries, 1 for each system whose solve met
        a zero pivot since the last check, else 0 (clears the record like check())."""
        import numpy as np
        out = np.zeros(self.nsys, dtype=np.int32)
        _check(_capi.lib().btd_plan_info(self._h, out.ctypes.data))
        return out

    def matrix_elems(self) -> int:
        return layout_elems(self.layout, self.nsys, self.nrow, self.m * self.m)

    def vector_elems(self) -> int:
        return layout_elems(self.layout, self.nsys, self.nrow, self.m)

    @property
    def handle(self) -> int:
        return self._h.value or 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _validate(plan, a, b, c, d, m, nsys, nrow):
    import torch
    if plan.created and (m, nsys, nrow) == (plan.m, plan.nsys, plan.nrow):
        nm, nv = plan.matrix_elems(), plan.vector_elems()
        for name, t, n in (("a", a, nm), ("b", b, nm), ("c", c, nm), ("d", d, nv)):
            _check_tensor(name, t, n, torch.float64)


def solve_noncyclic(plan: BtdmaGpuPlan, a, b, c, d, m: int, nsys: int, nrow: int,
                    stream=None) -> None:
    """In-place solve; x is returned in d (pascal_btdma_cuda.hpp:137-145)."""
    _validate(plan, a, b, c, d, m, nsys, nrow)
    _check(_capi.lib().btd_solve_noncyclic(plan._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                           int(m), int(nsys), int(nrow), _stream(stream)))


def solve_noncyclic_profiled(plan: BtdmaGpuPlan, a, b, c, d, m: int, nsys: int, nrow: int,
                             timings: BtdmaSolveTimings | None = None,
                             stream=None) -> BtdmaSolveTimings:
    """Solve and fill the six phase timings (pascal_btdma_cuda.hpp:147-156)."""
    _validate(plan, a, b, c, d, m, nsys, nrow)
    t = _capi.BtdTimings()
    _check(_capi.lib().btd_solve_noncyclic_profiled(plan._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                                    int(m), int(nsys), int(nrow),
                                                    ctypes.byref(t), _stream(stream)))
    out = timings if timings is not None else BtdmaSolveTimings()
    for f in ("total", "local_compute", "forward_exchange", "reduced_compute",
              "backward_exchange", "update_compute"):
        setattr(out, f, getattr(t, f))
    return out


def solve_cyclic(plan: BtdmaGpuPlan, a, b, c, d, m: int, nsys: int, nrow: int,
                 stream=None) -> None:
    """Periodic solve (pascal_btdma_cuda.hpp:158-166)."""
    _check(_capi.lib().btd_solve_cyclic(plan._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d), int(m),
                                        int(nsys), int(nrow), _stream(stream)))


def solve_noncyclic_noalloc(a, b, c, d, m: int, nsys: int, nrow: int, npart: int = 1,
                            layout: Layout = Layout.TILED, stream=None) -> None:
    """The non-periodic solve without a plan and without device allocation
    (btd_solve_noncyclic_noalloc): scratch in the library's static device array, at most
    BTD_NOALLOC_SCRATCH_BYTES (64 MiB); calls that overlap on one device race."""
    import torch
    nm = layout_elems(layout, nsys, nrow, m * m)
    nv = layout_elems(layout, nsys, nrow, m)
    for name, t, n in (("a", a, nm), ("b", b, nm), ("c", c, nm), ("d", d, nv)):
        _check_tensor(name, t, n, torch.float64)
    _check(_capi.lib().btd_solve_noncyclic_noalloc(_ptr(a), _ptr(b), _ptr(c), _ptr(d), int(m), int(nsys),
                                                   int(nrow), int(npart), int(layout), _stream(stream)))


def solve_noncyclic_save(plan: BtdmaGpuPlan, a, b, c, d, factors, pivots, stream=None) -> None:
    """Forward solve that keeps the block factorization (LU of R_q and pivots) for the
    adjoint. In place like solve_noncyclic: c returns C', d returns x."""
    import torch
    if plan.created:
        nm, nv1 = plan.matrix_elems(), layout_elems(plan.layout, plan.nsys, plan.nrow, 1)
        _check_tensor("factors", factors, nm, torch.float64)
        _check_tensor("pivots", pivots, nv1, torch.int64)
        _validate(plan, a, b, c, d, plan.m, plan.nsys, plan.nrow)
    _check(_capi.lib().btd_solve_noncyclic_save(plan._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                                _ptr(factors), _ptr(pivots), _stream(stream)))


def solve_noncyclic_save_out(plan: BtdmaGpuPlan, a, b, c, d, c_out, x, factors, pivots,
                             stream=None) -> None:
    """solve_noncyclic_save with c and d only read: C' goes to c_out and x to x (c_out may be
    c and x may be d; no other overlap)."""
    import torch
    if plan.created:
        nm, nv1 = plan.matrix_elems(), layout_elems(plan.layout, plan.nsys, plan.nrow, 1)
        _check_tensor("factors", factors, nm, torch.float64)
        _check_tensor("pivots", pivots, nv1, torch.int64)
        _check_tensor("c_out", c_out, nm, torch.float64)
        _check_tensor("x", x, plan.vector_elems(), torch.float64)
        _validate(plan, a, b, c, d, plan.m, plan.nsys, plan.nrow)
    _check(_capi.lib().btd_solve_noncyclic_save_out(plan._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                                    _ptr(c_out), _ptr(x), _ptr(factors), _ptr(pivots),
                                                    _stream(stream)))


def solve_adjoint(plan: BtdmaGpuPlan, a, c_factored, factors, pivots, x, w, lam, da, db, dc,
                  stream=None) -> None:
    """lambda = T^{-T} w and the gradients dA, dB, dC of L with dL/dx = w."""
    import torch
    if plan.created:
        nm, nv = plan.matrix_elems(), plan.vector_elems()
        for name, t, n in (("a", a, nm), ("c_factored", c_factored, nm), ("factors", factors, nm),
                           ("x", x, nv), ("w", w, nv), ("lam", lam, nv), ("da", da, nm),
                           ("db", db, nm), ("dc", dc, nm)):
            _check_tensor(name, t, n, torch.float64)
    _check(_capi.lib().btd_solve_adjoint(plan._h, _ptr(a), _ptr(c_factored), _ptr(factors),
                                         _ptr(pivots), _ptr(x), _ptr(w), _ptr(lam), _ptr(da),
                                         _ptr(db), _ptr(dc), _stream(stream)))


def f32r_work_elems(plan: BtdmaGpuPlan) -> int:
    """Floats of device scratch btd_solve_noncyclic_f32r needs for this plan."""
    return plan.matrix_elems() + plan.vector_elems()


def solve_noncyclic_f32r(plan: BtdmaGpuPlan, a, b, c, d, x, work, sweeps: int = 2,
                         stream=None) -> None:
    """FP32-stored blocks (a, b, c float32), FP64 d and x: FP32 solve plus `sweeps`
    refinement sweeps with FP64 residuals. a, b, c, d are only read."""
    import torch
    if plan.created:
        nm, nv = plan.matrix_elems(), plan.vector_elems()
        for name, t, n, dt in (("a", a, nm, torch.float32), ("b", b, nm, torch.float32),
                               ("c", c, nm, torch.float32), ("d", d, nv, torch.float64),
                               ("x", x, nv, torch.float64),
                               ("work", work, f32r_work_elems(plan), torch.float32)):
            _check_tensor(name, t, n, dt)
    _check(_capi.lib().btd_solve_noncyclic_f32r(plan._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                                _ptr(x), _ptr(work), int(sweeps), _stream(stream)))


def solve_noncyclic_f32s(plan: BtdmaGpuPlan, a, b, c, d, work=None, stream=None) -> None:
    """FP32-stored blocks (a, b, c float32, only read), FP64 arithmetic: d (float64) is
    replaced by x, the FP64 solve of the blocks' values; work holds C' (plan.matrix_elems()
    float64), or None for the workspace of a plan in Form.WORKSPACE (placed for this solve by
    plan.place_workspace_f32s)."""
    import torch
    if plan.created:
        nm, nv = plan.matrix_elems(), plan.vector_elems()
        checks = [("a", a, nm, torch.float32), ("b", b, nm, torch.float32),
                  ("c", c, nm, torch.float32), ("d", d, nv, torch.float64)]
        if work is not None:
            checks.append(("work", work, nm, torch.float64))
        for name, t, n, dt in checks:
            _check_tensor(name, t, n, dt)
    _check(_capi.lib().btd_solve_noncyclic_f32s(plan._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                                _ptr(work), _stream(stream)))


def solve_transposed_f32r(plan: BtdmaGpuPlan, a, b, c, w, lam, work, sweeps: int = 2,
                          stream=None) -> None:
    """lambda = T^{-T} w for the FP32-stored blocks a, b, c of T (as given for T): FP32
    solve of the transposed system plus `sweeps` FP64-residual refinement sweeps."""
    import torch
    if plan.created:
        nm, nv = plan.matrix_elems(), plan.vector_elems()
        for name, t, n, dt in (("a", a, nm, torch.float32), ("b", b, nm, torch.float32),
                               ("c", c, nm, torch.float32), ("w", w, nv, torch.float64),
                               ("lam", lam, nv, torch.float64),
                               ("work", work, f32r_work_elems(plan), torch.float32)):
            _check_tensor(name, t, n, dt)
    _check(_capi.lib().btd_solve_transposed_f32r(plan._h, _ptr(a), _ptr(b), _ptr(c), _ptr(w),
                                                 _ptr(lam), _ptr(work), int(sweeps),
                                                 _stream(stream)))


def outer_grads_f32(plan: BtdmaGpuPlan, lam, x, da, db, dc, stream=None) -> None:
    """FP32 gradients dA_q = -lambda_q x_{q-1}^T, dB_q = -lambda_q x_q^T,
    dC_q = -lambda_q x_{q+1}^T from FP64 lambda and x."""
    import torch
    if plan.created:
        nm, nv = plan.matrix_elems(), plan.vector_elems()
        for name, t, n, dt in (("lam", lam, nv, torch.float64), ("x", x, nv, torch.float64),
                               ("da", da, nm, torch.float32), ("db", db, nm, torch.float32),
                               ("dc", dc, nm, torch.float32)):
            _check_tensor(name, t, n, dt)
    _check(_capi.lib().btd_outer_grads_f32(plan._h, _ptr(lam), _ptr(x), _ptr(da), _ptr(db),
                                           _ptr(dc), _stream(stream)))


def sweep_shape(axis: int, n1: int, n2: int, n3: int):
    """(nrow, nsys) of the sweep along axis 0 (x), 1 (y) or 2 (z) of an n1 x n2 x n3 grid."""
    return ((n1, n2 * n3), (n2, n1 * n3), (n3, n1 * n2))[int(axis)]


def _check_fields(n1, n2, n3, m, **fields):
    import torch
    sizes = {"U": 5 * (n1 + 2) * (n2 + 2) * (n3 + 2)}
    for name, t in fields.items():
        if t is None:
            continue
        _check_tensor(name, t, sizes.get(name, m * n1 * n2 * n3), torch.float64)


def euler_assemble(plan: BtdmaGpuPlan, axis: int, U, rhs, n1: int, n2: int, n3: int, h: float,
                   dt: float, eps: float, a, b, c, d=None, stream=None) -> None:
    """Blocks (and, with rhs, the right-hand side) of the euler3d ADI sweep along `axis`
    from the haloed conserved field U(0:n1+1, 0:n2+1, 0:n3+1, 5) (Fortran order, flat float64
    device tensor), written in the plan's layout; replaces the reference's build_matrix
    kernel plus its transposes."""
    import torch
    if plan.created:
        nm, nv = plan.matrix_elems(), plan.vector_elems()
        _check_fields(n1, n2, n3, 5, U=U, rhs=rhs)
        for name, t, n in (("a", a, nm), ("b", b, nm), ("c", c, nm), ("d", d, nv)):
            if t is not None:
                _check_tensor(name, t, n, torch.float64)
    _check(_capi.lib().btd_euler_assemble(plan._h, int(axis), _ptr(U), _ptr(rhs), int(n1), int(n2),
                                          int(n3), float(h), float(dt), float(eps), _ptr(a), _ptr(b),
                                          _ptr(c), _ptr(d), _stream(stream)))


def weno_assemble(plan: BtdmaGpuPlan, axis: int, U, vel, n1: int, n2: int, h: float, gdt: float,
                  a, b, c, stream=None) -> None:
    """Blocks (m = 3) of the weno2d implicit sweep along `axis` (0 = x, 1 = y) from the scalar
    field U(1-3:n1+3, 1-3:n2+3) with its halo (Fortran order, flat float64 device tensor) and
    each line's advection speed `vel` (n2 values for x, n1 for y), in the plan's layout;
    replaces the reference's WENO5 coefficient and build_matrix kernels."""
    import torch
    if plan.created:
        nm = plan.matrix_elems()
        _check_tensor("U", U, (n1 + 6) * (n2 + 6), torch.float64)
        _check_tensor("vel", vel, n2 if axis == 0 else n1, torch.float64)
        for name, t in (("a", a), ("b", b), ("c", c)):
            _check_tensor(name, t, nm, torch.float64)
    _check(_capi.lib().btd_weno_assemble(plan._h, int(axis), _ptr(U), _ptr(vel), int(n1), int(n2), float(h),
                                         float(gdt), _ptr(a), _ptr(b), _ptr(c), _stream(stream)))


def weno_sweep_cyclic(plan: BtdmaGpuPlan, axis: int, U, vel, n1: int, n2: int, h: float, gdt: float,
                      work, d, stream=None) -> None:
    """Matrix-free periodic weno2d sweep: solves (I - gdt L) x = d along `axis` with the blocks
    computed inside stage 1 of the two-stage periodic solve (equal, bit for bit, to
    weno_assemble + solve_cyclic); `work` holds 2 * plan.matrix_elems() doubles."""
    import torch
    if plan.created:
        _check_tensor("U", U, (n1 + 6) * (n2 + 6), torch.float64)
        _check_tensor("vel", vel, n2 if axis == 0 else n1, torch.float64)
        _check_tensor("work", work, 2 * plan.matrix_elems(), torch.float64)
        _check_tensor("d", d, plan.vector_elems(), torch.float64)
    _check(_capi.lib().btd_weno_sweep_cyclic(plan._h, int(axis), _ptr(U), _ptr(vel), int(n1), int(n2), float(h),
                                             float(gdt), _ptr(work), _ptr(d), _stream(stream)))


def weno_gather(plan: BtdmaGpuPlan, axis: int, field, n1: int, n2: int, d, stream=None) -> None:
    """Scalar field F(n1, n2) (Fortran order) -> the weno2d sweep's vector d (m = 3) in the plan's
    layout."""
    import torch
    if plan.created:
        _check_tensor("field", field, n1 * n2, torch.float64)
        _check_tensor("d", d, plan.vector_elems(), torch.float64)
    _check(_capi.lib().btd_weno_gather(plan._h, int(axis), _ptr(field), int(n1), int(n2), _ptr(d), _stream(stream)))


def weno_scatter(plan: BtdmaGpuPlan, axis: int, d, n1: int, n2: int, field, stream=None) -> None:
    """The weno2d sweep's vector d -> the scalar field F(n1, n2)."""
    import torch
    if plan.created:
        _check_tensor("field", field, n1 * n2, torch.float64)
        _check_tensor("d", d, plan.vector_elems(), torch.float64)
    _check(_capi.lib().btd_weno_scatter(plan._h, int(axis), _ptr(d), int(n1), int(n2), _ptr(field), _stream(stream)))


def euler_sweep_work_elems(plan: BtdmaGpuPlan) -> int:
    """Doubles of device scratch btd_euler_sweep_cyclic needs for this plan."""
    return (2 * plan.matrix_elems() + plan.vector_elems()
            + layout_elems(plan.layout, plan.nsys, plan.nrow + 2, plan.m))


def euler_sweep_cyclic(plan: BtdmaGpuPlan, axis: int, U, rhs, n1: int, n2: int, n3: int, h: float,
                       dt: float, eps: float, work, out, stream=None) -> None:
    """Matrix-free periodic euler3d sweep along `axis`: the blocks btd_euler_assemble would
    write are computed inside the solve; the solution goes into the field `out`
    (n1, n2, n3, 5), which may be `rhs` itself."""
    import torch
    if plan.created:
        _check_fields(n1, n2, n3, 5, U=U, rhs=rhs, out=out)
        _check_tensor("work", work, euler_sweep_work_elems(plan), torch.float64)
    _check(_capi.lib().btd_euler_sweep_cyclic(plan._h, int(axis), _ptr(U), _ptr(rhs), int(n1),
                                              int(n2), int(n3), float(h), float(dt), float(eps),
                                              _ptr(work), _ptr(out), _stream(stream)))


def sweep_gather(plan: BtdmaGpuPlan, axis: int, field, n1: int, n2: int, n3: int, d,
                 stream=None) -> None:
    """Field F(n1, n2, n3, m) -> the sweep's vector d in the plan's layout."""
    import torch
    if plan.created:
        _check_fields(n1, n2, n3, plan.m, field=field)
        _check_tensor("d", d, plan.vector_elems(), torch.float64)
    _check(_capi.lib().btd_sweep_gather(plan._h, int(axis), _ptr(field), int(n1), int(n2), int(n3),
                                        _ptr(d), _stream(stream)))


def sweep_scatter(plan: BtdmaGpuPlan, axis: int, d, n1: int, n2: int, n3: int, field,
                  stream=None) -> None:
    """The sweep's vector d -> field F(n1, n2, n3, m)."""
    import torch
    if plan.created:
        _check_fields(n1, n2, n3, plan.m, field=field)
        _check_tensor("d", d, plan.vector_elems(), torch.float64)
    _check(_capi.lib().btd_sweep_scatter(plan._h, int(axis), _ptr(d), int(n1), int(n2), int(n3),
                                         _ptr(field), _stream(stream)))


class BtdmaBatch:
    """A batch of systems with their own block size m[s] (1..10) and row count n[s] (>= 2),
    packed into interleaved packs of equal m (64 systems for m = 1, floor(32/m) otherwise)
    and solved by one persistent kernel. With (mo, vo, w) = locate(s), element e of block row
    `row` of system s sits at mo + (row*m*m + e)*w in the block arrays and at
    vo + (row*m + e)*w in the vector array."""

    def __init__(self, m, n):
        import numpy as np
        self.m = np.ascontiguousarray(m, dtype=np.int32)
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        if self.m.shape != self.n.shape or self.m.ndim != 1:
            raise ValueError("m and n must be 1-D arrays of the same length")
        self._h = ctypes.c_void_p(None)
        ip = ctypes.POINTER(ctypes.c_int)
        _check(_capi.lib().btd_batch_create(ctypes.byref(self._h), len(self.m),
                                            self.m.ctypes.data_as(ip), self.n.ctypes.data_as(ip)))
        me, ve, nt = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
        _check(_capi.lib().btd_batch_sizes(self._h, ctypes.byref(me), ctypes.byref(ve), ctypes.byref(nt)))
        self.mat_elems, self.vec_elems, self.npacks = me.value, ve.value, nt.value
        self.placement, self.grid = 0, None

    @property
    def nsys(self) -> int:
        return len(self.m)

    def locate(self, sys: int) -> tuple[int, int, int]:
        """(first block element, first vector element, element stride) of system sys."""
        mo, vo, w = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
        _check(_capi.lib().btd_batch_locate(self._h, int(sys), ctypes.byref(mo), ctypes.byref(vo),
                                            ctypes.byref(w)))
        return mo.value, vo.value, w.value

    def set_params(self, placement: int, grid: int) -> None:
        _check(_capi.lib().btd_batch_set_params(self._h, int(placement), int(grid)))
        self.placement, self.grid = int(placement), int(grid)

    def generate(self, kind: GenKind, seed: int, a, b, c, d, stream=None) -> None:
        _check(_capi.lib().btd_batch_generate(self._h, int(kind), int(seed), _ptr(a), _ptr(b),
                                              _ptr(c), _ptr(d), _stream(stream)))

    def solve(self, a, b, c, d, stream=None) -> None:
        import torch
        for name, t, n in (("a", a, self.mat_elems), ("b", b, self.mat_elems),
                           ("c", c, self.mat_elems), ("d", d, self.vec_elems)):
            _check_tensor(name, t, n, torch.float64)
        _check(_capi.lib().btd_batch_solve(self._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                           _stream(stream)))

    def check(self) -> bool:
        s = ctypes.c_int(0)
        _check(_capi.lib().btd_batch_check(self._h, ctypes.byref(s)))
        return bool(s.value)

    def extract(self, flat, sys: int, vector: bool = True):
        """Host copy of one system's vector (n, m) or blocks (n, m, m) from a packed array."""
        import numpy as np
        m, n = int(self.m[sys]), int(self.n[sys])
        mo, vo, w = self.locate(sys)
        E = m if vector else m * m
        base = vo if vector else mo
        idx = base + (np.arange(n)[:, None] * E + np.arange(E)[None, :]) * w
        vals = flat[idx]
        return vals if vector else np.swapaxes(vals.reshape(n, m, m), -1, -2)

    def destroy(self) -> None:
        if self._h.value:
            _capi.lib().btd_batch_destroy(self._h)
        self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Graph:
    """A captured sequence of asynchronous solves replayed with one launch
    (btd_graph_begin / btd_graph_end / btd_graph_launch)."""

    def __init__(self):
        self._h = ctypes.c_void_p(None)

    def begin(self, stream) -> None:
        _check(_capi.lib().btd_graph_begin(_stream(stream)))

    def end(self, stream) -> None:
        _check(_capi.lib().btd_graph_end(_stream(stream), ctypes.byref(self._h)))

    def launch(self, stream=None) -> None:
        _check(_capi.lib().btd_graph_launch(self._h, _stream(stream)))

    def destroy(self) -> None:
        if self._h.value:
            _capi.lib().btd_graph_destroy(self._h)
        self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def matrix_index(sys: int, row: int, p: int, q: int, nsys: int, nrow: int, m: int) -> int:
    return int(_capi.lib().btd_matrix_index(sys, row, p, q, nsys, nrow, m))


def vector_index(sys: int, row: int, p: int, nsys: int, nrow: int) -> int:
    return int(_capi.lib().btd_vector_index(sys, row, p, nsys, nrow))


def partition_1d(start: int, end: int, nprocs: int, rank: int) -> tuple[int, int]:
    f, l = ctypes.c_int(), ctypes.c_int()
    _capi.lib().btd_partition_1d(start, end, nprocs, rank, ctypes.byref(f), ctypes.byref(l))
    return f.value, l.value


def layout_elems(layout: Layout, nsys: int, nrow: int, elems_per_block: int) -> int:
    return int(_capi.lib().btd_layout_elems(int(layout), nsys, nrow, elems_per_block))


def convert_layout(src_layout: Layout, dst_layout: Layout, elems_per_block: int, nsys: int,
                   nrow: int, src, dst, stream=None) -> None:
    _check(_capi.lib().btd_convert_layout(int(src_layout), int(dst_layout), elems_per_block,
                                          nsys, nrow, _ptr(src), _ptr(dst), _stream(stream)))


def generate(kind: GenKind, seed: int, m: int, nsys: int, nrow: int, layout: Layout,
             a=None, b=None, c=None, d=None, stream=None) -> None:
    _check(_capi.lib().btd_generate(int(kind), int(seed), m, nsys, nrow, int(layout), _ptr(a),
                                    _ptr(b), _ptr(c), _ptr(d), _stream(stream)))


def generate_f32(kind: GenKind, seed: int, m: int, nsys: int, nrow: int, layout: Layout,
                 a=None, b=None, c=None, d=None, stream=None) -> None:
    _check(_capi.lib().btd_generate_f32(int(kind), int(seed), m, nsys, nrow, int(layout),
                                        _ptr(a), _ptr(b), _ptr(c), _ptr(d), _stream(stream)))


def generate_host(kind: GenKind, seed: int, m: int, nsys: int, nrow: int, sys0: int = 0,
                  count: int | None = None):
    """Host copy of the generator: numpy arrays A, B, C (nrow, count, m, m) natural
    row-major blocks and d (nrow, count, m), bit-identical to the device generator."""
    import numpy as np
    count = nsys - sys0 if count is None else count
    A = np.empty((nrow, count, m, m))
    B = np.empty_like(A)
    C = np.empty_like(A)
    d = np.empty((nrow, count, m))
    _check(_capi.lib().btd_generate_host(int(kind), int(seed), m, nsys, nrow, sys0, count,
                                         A.ctypes.data, B.ctypes.data, C.ctypes.data,
                                         d.ctypes.data))
    return A, B, C, d
