"""Builds libpascal_btdma_amd.so in-tree with hipcc for gfx950.

Every HIP translation unit under csrc/ is compiled to an object in parallel, then the
objects are linked into one shared library next to this file. Usage:
    python -m pascal_btdma_amd.build [--verbose] [--resource-usage]
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpascal_btdma_amd.so")
ARCH = os.environ.get("BTD_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops: keep v_pk_{add,mul,fma}_f32 out of the generated code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]

# (source, extra defines, object name): the kernel templates are instantiated once per
# block size m so the instantiations compile in parallel.
UNITS = [("btd_inst.hip", [f"-DBTD_M={m}"], f"btd_inst_m{m}.o") for m in range(1, 11)] + [
    ("btd_dispatch.hip", [], "btd_dispatch.o"),
    ("btd_capi.hip", [], "btd_capi.o"),
    ("btd_batch.hip", [], "btd_batch.o"),
    ("btd_batch_api.hip", [], "btd_batch_api.o"),
    ("btd_euler.hip", [], "btd_euler.o"),
    ("btd_weno.hip", [], "btd_weno.o"),
    ("btd_noalloc.hip", [], "btd_noalloc.o"),
]


def _compile(unit, extra: list[str], verbose: bool) -> tuple[str, float, str]:
    src, defs, objname = unit
    obj = os.path.join(BUILD, objname)
    cmd = [HIPCC, *COMMON, *defs, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    out = (r.stdout + r.stderr) if verbose else ""
    return obj, time.time() - t0, out


def build(verbose: bool = False, resource_usage: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    extra = ["-Rpass-analysis=kernel-resource-usage"] if resource_usage else []
    # slowest first: the batch kernels, then the per-m solve kernels
    units = [u for u in UNITS if u[0] == "btd_batch.hip"] + list(reversed(UNITS[:10])) + \
        [u for u in UNITS[10:] if u[0] != "btd_batch.hip"]
    with cf.ThreadPoolExecutor(max_workers=int(os.environ.get("BTD_BUILD_JOBS", "8"))) as ex:
        results = list(ex.map(lambda u: _compile(u, extra, verbose or resource_usage), units))
    objs = [r[0] for r in results]
    for unit, (_, dt, out) in zip(units, results):
        if verbose:
            print(f"[build] {unit[2]}: {dt:.1f}s")
        if out:
            print(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_examples(verbose)
    build_fortran_examples(verbose)
    return LIB


def build_cpp_examples(verbose: bool = False) -> None:
    """C++ programs under tests/cpp that use the C++ mirror header (linked against the
    library with an $ORIGIN-relative rpath, so they run from any copy of the tree)."""
    root = os.path.dirname(HERE)
    src_dir = os.path.join(root, "tests", "cpp")
    if not os.path.isdir(src_dir):
        return
    out_dir = os.path.join(src_dir, "bin")
    os.makedirs(out_dir, exist_ok=True)
    for f in sorted(os.listdir(src_dir)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(out_dir, os.path.splitext(f)[0])
        cmd = [HIPCC, "-O2", "-std=c++17", os.path.join(src_dir, f), "-I", os.path.join(root, "include"),
               "-L", HERE, "-lpascal_btdma_amd", "-Wl,-rpath,$ORIGIN/../../../pascal_btdma_amd", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {f} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"[build] {exe}")


def build_fortran_examples(verbose: bool = False) -> None:
    """The Fortran binding modules (include/mod_btdma_amd.f90, include/mod_btdma_gpu_v2.f90) and the
    Fortran programs under tests/fortran, compiled with the image's amdflang (skipped if it is absent)."""
    root = os.path.dirname(HERE)
    src_dir = os.path.join(root, "tests", "fortran")
    fc = shutil.which("amdflang") or "/opt/rocm/lib/llvm/bin/amdflang"
    if not os.path.isdir(src_dir) or not os.path.exists(fc):
        return
    out_dir = os.path.join(src_dir, "bin")
    os.makedirs(out_dir, exist_ok=True)
    mod_objs = []
    for mod in ("mod_btdma_amd", "mod_btdma_gpu_v2"):
        mod_obj = os.path.join(out_dir, mod + ".o")
        cmd = [fc, "-O2", "-c", os.path.join(root, "include", mod + ".f90"), "-module-dir", out_dir,
               "-o", mod_obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building the Fortran module failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        mod_objs.append(mod_obj)
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "lib")
    for f in sorted(os.listdir(src_dir)):
        if not f.endswith(".f90"):
            continue
        exe = os.path.join(out_dir, os.path.splitext(f)[0])
        cmd = [fc, "-O2", os.path.join(src_dir, f), *mod_objs, "-I", out_dir, "-L", HERE, "-lpascal_btdma_amd",
               "-L", rocm_lib, "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../../pascal_btdma_amd", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {f} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"[build] {exe}")


if __name__ == "__main__":
    build(verbose="--verbose" in sys.argv, resource_usage="--resource-usage" in sys.argv)
    print(LIB)
