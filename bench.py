"""Benchmark of the batched block-Thomas solve on one MI355X.

Workload (default): the Euler-3D-like sweep, m = 5, n = 512 block rows, 131072 systems,
FP64, single partition, tiled layout; inputs generated on the device and resident in
HBM before the timed region. A step is one solve of the whole batch. The solve is in
place (C and D are overwritten), so C and D are restored from pristine copies before
every step, outside the event-timed region.

Prints one JSON line: block-rows/s, the HBM roofline of the solve kernel (algorithmic
bytes 8*(3m^2+2m) per block row over the HIP-event time of the kernel), the maximum
relative error vs the CPU oracle on a subset of systems, and the CPU baseline (the
oracle on a bounded sample of the same workload, on the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "block-rows/s + % HBM peak; FP32+refine vs FP64 plain at rel err ≤1e-12"

WORKLOADS = {
    "euler_m5": dict(kind=2, m=5, n=512, nsys=131072, seed=1,
                     name="Euler-3D-like sweep, m=5, n=512, 131072 systems, FP64 plain"),
    "ddperm_m3": dict(kind=1, m=3, n=64, nsys=1024, seed=0,
                      name="diagonally dominant permuted rows, m=3, n=64, 1024 systems, FP64 plain"),
}


def algorithmic_bytes(m: int, rows: int) -> float:
    """FP64 plain solve: read A, B, C, d once and write x once per block row."""
    return 8.0 * (3 * m * m + 2 * m) * rows


def tiled_vec_subset(flat: np.ndarray, n: int, m: int, tiles: list[int]) -> np.ndarray:
    out = []
    for t in tiles:
        blk = flat[t * n * m * 64:(t + 1) * n * m * 64].reshape(n, m, 64)
        out.append(np.transpose(blk, (0, 2, 1)))
    return np.concatenate(out, axis=1)  # (n, 64*len(tiles), m)


def cpu_baseline(w: dict, target_s: float = 12.0) -> dict:
    """The oracle (FP64, LAPACK dgesv semantics) on the host cores, on the first S systems
    of the same workload; S is sized so the timed solves take about target_s seconds."""
    import pascal_btdma_amd as pb
    from oracle import pyoracle
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    pyoracle.set_threads(threads)
    m, n = w["m"], w["n"]

    def gen(S):
        A, B, C, d = pb.generate_host(w["kind"], w["seed"], m, w["nsys"], n, 0, S)
        return [pyoracle.to_cpu_layout(X) for X in (A, B, C)] + [d]

    def run(arrs):
        A, B, C, d = (x.copy() for x in arrs)  # the solve is in place
        t0 = time.perf_counter()
        pyoracle.lib().ora_btdma_many(n, A.shape[1], m, pyoracle._p(A), pyoracle._p(B),
                                      pyoracle._p(C), pyoracle._p(d))
        return time.perf_counter() - t0

    S = max(threads, 64)
    dt = run(gen(S))
    S = int(min(w["nsys"], max(S, S * target_s / 3.0 / max(dt, 1e-3))))
    S = max(threads, S - S % threads)
    arrs = gen(S)
    total_t, total_rows, reps = 0.0, 0, 0
    while total_t < target_s and reps < 10:
        total_t += run(arrs)
        total_rows += S * n
        reps += 1
    return {"value": total_rows / total_t, "unit": "block-rows/s", "cores": pyoracle.last_threads(),
            "kind": "port",
            "sample": f"first {S} systems of the same workload, solved {reps}x "
                      f"({total_rows} block rows, {total_t:.1f} s) with the C oracle, "
                      f"OpenMP over systems"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="euler_m5", choices=sorted(WORKLOADS))
    ap.add_argument("--layout", default="tiled", choices=["tiled", "ref_gpu"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(independent systems shard across GPUs with no communication; see DESIGN.md)",
              file=sys.stderr)
        return 2

    import torch

    import pascal_btdma_amd as pb
    from oracle import pyoracle

    w = WORKLOADS[args.workload]
    m, n, nsys = w["m"], w["n"], w["nsys"]
    lay = pb.Layout.TILED if args.layout == "tiled" else pb.Layout.REF_GPU
    dev = torch.device("cuda:0")
    nm, nv = pb.layout_elems(lay, nsys, n, m * m), pb.layout_elems(lay, nsys, n, m)
    a, b, c = (torch.empty(nm, dtype=torch.float64, device=dev) for _ in range(3))
    d = torch.empty(nv, dtype=torch.float64, device=dev)
    pb.generate(w["kind"], w["seed"], m, nsys, n, lay, a, b, c, d)
    c0, d0 = c.clone(), d.clone()
    torch.cuda.synchronize()
    plan = pb.BtdmaGpuPlan().create(m, nsys, n, 1, lay)
    stream = torch.cuda.current_stream()

    def restore():
        c.copy_(c0)
        d.copy_(d0)

    for _ in range(args.warmup):
        restore()
        pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        restore()
        e0.record(stream)
        pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    times = np.array([e0.elapsed_time(e1) for e0, e1 in evs])  # ms
    singular = plan.check()
    ms = float(times.mean())
    rows = nsys * n
    value = rows / (ms * 1e-3)
    achieved = algorithmic_bytes(m, rows) / (ms * 1e-3) / 1e9

    # parity on a subset of systems: the first two and the last tiles
    ntiles = (nsys + 63) // 64
    tiles = sorted({0, 1, ntiles - 1})
    if lay == pb.Layout.TILED:
        x = tiled_vec_subset(d.cpu().numpy(), n, m, tiles)
        sys_ids = [t * 64 + j for t in tiles for j in range(64)]
    else:
        full = d.view(m, n, nsys)
        sys_ids = [s for t in tiles for s in range(t * 64, min(nsys, t * 64 + 64))]
        x = full[:, :, sys_ids].permute(1, 2, 0).cpu().numpy()
    errs = []
    for lo in range(0, len(sys_ids), 64):
        s0 = sys_ids[lo]
        A, B, C, dd = pb.generate_host(w["kind"], w["seed"], m, nsys, n, s0, 64)
        xo, info = pyoracle.solve_plain(pyoracle.to_cpu_layout(A), pyoracle.to_cpu_layout(B),
                                        pyoracle.to_cpu_layout(C), dd)
        xs = x[:, lo:lo + 64]
        errs.append((np.abs(xs - xo).max(axis=(0, 2)) / np.abs(xo).max(axis=(0, 2))).max())
    max_err = float(max(errs))

    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_plain_{args.workload}_{args.layout}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")

    cpu = None if args.no_cpu else cpu_baseline(w, args.cpu_seconds)
    line = {
        "metric": METRIC, "value": value, "unit": "block-rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded counter-based generator, generated on the device)",
        "config": {"workload": w["name"], "m": m, "n": n, "nsys": nsys, "npart": 1,
                   "layout": args.layout, "block_rows": rows},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "max_rel_err_vs_oracle": max_err, "parity_systems": len(sys_ids),
        "singular": singular, "ms_min": float(times.min()), "ms_max": float(times.max()),
        "wall_s_timed_region_incl_restore": wall,
        "kernel": f"k_plain<{m},{'TILED' if lay == pb.Layout.TILED else 'REF_GPU'}>",
    }
    print(json.dumps(line), flush=True)
    plan.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
