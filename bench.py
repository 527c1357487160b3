"""Benchmark of the batched block-Thomas solve on one MI355X.

Headline workload: the Euler-3D-like sweep, m = 5, n = 512 block rows, 131072 systems,
FP64, single partition, tiled layout; inputs generated on the device and resident in
HBM before the timed region. A step is one solve of the whole batch. Two forms of the same
kernel (k_plain, the same arithmetic and bits) are timed: in place (C' over C, the
reference's contract; C and D restored from pristine copies before every step) and the
plan's workspace form (C' to a plan-owned array placed by btd_plan_place_workspace, C only
read; D restored before every step); the faster one is the headline, both are reported.
Restores run outside the event-timed region (events are recorded on the solve's stream).

Prints one JSON line: block-rows/s, the HBM roofline of the solve kernel (algorithmic
bytes 8*(3m^2+2m) per block row over the HIP-event time of the kernel), the maximum
relative error vs the CPU oracle on a subset of systems, the CPU baseline (the oracle on
a bounded sample of the same workload, on the host cores), and under "extras" the other
configurations: the reference Study's single-GPU case, the small CPU-sized case, the
mixed-size batch with long partitioned systems in one graph, FP32 storage with FP64
refinement, and the differentiable solve (forward and adjoint).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
import traceback

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

# The reference's only published numbers for this path: Study, 1 x H200, CUDA C++ port,
# diagonal workload 32 x 32 x 128 (nsys 1024, n 128), solve time in ms
# (Study/result/tables/summary_by_case.csv:2-5).
STUDY_H200_MS = {2: 2.134, 3: 3.901, 5: 11.581, 8: 28.550}
# The Study's largest cases, 8 x H200 (one rank each), CUDA C++ port, device MPI mode, diagonal
# workload 128 x 128 x 2048 (nsys 16384, n 2048), solve time in ms
# (Study/result/tables/summary_by_case.csv:25,27,29).
STUDY_LARGE_H200X8_MS = {2: 6.579, 5: 29.902, 8: 84.719}


def algorithmic_bytes(m: int, rows: int) -> float:
    """FP64 plain solve: read A, B, C, d once and write x once per block row."""
    return 8.0 * (3 * m * m + 2 * m) * rows


def tiled_vec_subset(flat: np.ndarray, n: int, m: int, tiles: list[int]) -> np.ndarray:
    out = []
    for t in tiles:
        blk = flat[t * n * m * 64:(t + 1) * n * m * 64].reshape(n, m, 64)
        out.append(np.transpose(blk, (0, 2, 1)))
    return np.concatenate(out, axis=1)  # (n, 64*len(tiles), m)


def rel_err_sys(x: np.ndarray, xo: np.ndarray) -> np.ndarray:
    return np.abs(x - xo).max(axis=(0, 2)) / np.abs(xo).max(axis=(0, 2))


def event_time(fn, restore, steps: int, stream) -> np.ndarray:
    """Per-step HIP-event times (ms) of fn on `stream`, restore() outside the events. The
    device is drained before every restore: restore() works on torch's current stream, which
    need not be `stream`, and must neither overlap the previous step nor change its inputs
    while it runs."""
    import torch
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    for e0, e1 in evs:
        torch.cuda.synchronize()
        restore()
        e0.record(stream)
        fn()
        e1.record(stream)
    torch.cuda.synchronize()
    return np.array([e0.elapsed_time(e1) for e0, e1 in evs])


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
    S = int(min(w["nsys"], 16384, max(S, S * target_s / 3.0 / max(dt, 1e-3))))
    S = max(threads, S - S % threads)
    arrs = gen(S)
    total_t, total_rows, reps = 0.0, 0, 0
    while total_t < target_s and reps < 10:
        total_t += run(arrs)
        total_rows += S * n
        reps += 1
    cores = pyoracle.last_threads()
    # one thread, on a slice sized for about a quarter of target_s
    S1 = max(64, min(S, int(S / cores * 0.25 * target_s / max(total_t / reps, 1e-3))))
    pyoracle.set_threads(1)
    t1 = run([x[:, :S1].copy() for x in arrs])
    pyoracle.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": total_rows / total_t, "unit": "block-rows/s", "cores": cores,
            "kind": "port",
            "sample": f"first {S} systems of the same workload ({S * n} block rows), solved {reps}x "
                      f"in {total_t:.1f} s with the C oracle, OpenMP over systems",
            "single_thread_value": S1 * n / t1,
            "single_thread_sample": f"first {S1} systems, one thread, {t1:.1f} s",
            "cpu_model": model, "nproc": os.cpu_count()}


class ParityChecks:
    """Device results recorded during the GPU part, to be checked against the oracle in the
    CPU leg: each entry holds a function returning the natural-form inputs (from the
    product's host generator), the device solution x (n, S, m) for those S systems, a sink
    that receives the per-system relative errors against the oracle and, optionally, one that
    receives them against the binary128 truth solve of the same inputs."""

    def __init__(self):
        self.items = []

    def add(self, inputs, x, sink, truth_sink=None):
        self.items.append((inputs, np.asarray(x), sink, truth_sink))


def max_sink(out: dict, key: str, exact_key: str | None = None):
    """A sink that keeps the largest per-system error seen under out[key] and, with exact_key,
    whether every checked value equals the oracle's bit for bit."""
    def sink(e, x=None, xo=None):
        out[key] = max(out[key] or 0.0, float(np.max(e)))
        if exact_key is not None and x is not None:
            out[exact_key] = bool(out.get(exact_key, True) is not False and np.array_equal(x, xo))
    return sink


def cpu_leg(w: dict, target_s: float, checks: ParityChecks, baseline: bool) -> dict | None:
    """The CPU leg of the benchmark: times the oracle on a bounded sample of the headline
    workload (unless baseline is False) and checks every recorded device result against the
    oracle's solve of the same inputs."""
    from oracle import pyoracle
    cache = {}  # id(inputs) -> (inputs, oracle x, truth x or None)
    for inputs, x, sink, truth_sink in checks.items:
        if id(inputs) not in cache:
            A, B, C, dd = inputs()
            args = (pyoracle.to_cpu_layout(A), pyoracle.to_cpu_layout(B), pyoracle.to_cpu_layout(C), dd)
            xo, info = pyoracle.solve_plain(*args)
            cache[id(inputs)] = [inputs, xo, None, args]
        entry = cache[id(inputs)]
        sink(rel_err_sys(x, entry[1]), x, entry[1])
        if truth_sink is not None:
            if entry[2] is None:
                entry[2], _ = pyoracle.solve_truth(*entry[3])
            truth_sink(rel_err_sys(x, entry[2]))
    return cpu_baseline(w, target_s) if baseline else None


# ------------------------------------------------------------------------------ extras --

def extra_study(dev) -> dict:
    """The reference Study's single-GPU case (diagonal workload, 32 x 32 x 128) for m = 2, 3,
    5, 8: time of one solve at npart = 1 and at the best npart, vs the published H200 time."""
    import torch
    import pascal_btdma_amd as pb
    nsys, n = 1024, 128
    lay = pb.Layout.REF_GPU
    out = {}
    for m, h200 in STUDY_H200_MS.items():
        a, b, c, d = (torch.empty(nsys * n * (m * m if i < 3 else m), dtype=torch.float64, device=dev)
                      for i in range(4))
        pb.generate(pb.GenKind.DIAG, 0, m, nsys, n, lay, a, b, c, d)
        c0, d0 = c.clone(), d.clone()
        a0 = a.clone()
        res = {}
        for P in (1, 2, 4, 8, 16):
            with pb.BtdmaGpuPlan().create(m, nsys, n, P, lay) as plan:
                def restore():
                    a.copy_(a0)
                    c.copy_(c0)
                    d.copy_(d0)
                s = torch.cuda.current_stream()
                restore()
                pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
                t = event_time(lambda: pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n), restore, 5, s)
                res[P] = float(np.median(t))
        best = min(res, key=res.get)
        with pb.BtdmaGpuPlan().create(m, nsys, n, 0, lay) as plan:
            def restore():
                a.copy_(a0)
                c.copy_(c0)
                d.copy_(d0)
            t = event_time(lambda: pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n), restore, 5,
                           torch.cuda.current_stream())
            auto_np, ms_auto = plan.npart, float(np.median(t))
        out[f"m{m}"] = {"ms_npart1": res[1], "best_npart": best, "ms_best": res[best],
                        "auto_npart": auto_np, "ms_auto_npart": ms_auto,
                        "speedup_auto_npart": h200 / ms_auto,
                        "h200_ms_published": h200, "speedup_npart1": h200 / res[1],
                        "speedup_best": h200 / res[best],
                        "block_rows_per_s_best": nsys * n / (res[best] * 1e-3)}
    return out


def extra_study_large(dev) -> dict:
    """The Study's largest cases (diagonal workload, 128 x 128 x 2048, reference layout) for
    m = 2, 5, 8 on one GPU at the library's partition count, against the published time of the
    reference on 8 H200s."""
    import torch
    import pascal_btdma_amd as pb
    nsys, n = 16384, 2048
    lay = pb.Layout.REF_GPU
    out = {}
    for m, h200 in STUDY_LARGE_H200X8_MS.items():
        a, b, c, d = (torch.empty(nsys * n * (m * m if i < 3 else m), dtype=torch.float64, device=dev)
                      for i in range(4))
        pb.generate(pb.GenKind.DIAG, 0, m, nsys, n, lay, a, b, c, d)
        a0, c0, d0 = a.clone(), c.clone(), d.clone()

        def restore():
            a.copy_(a0)
            c.copy_(c0)
            d.copy_(d0)
        with pb.BtdmaGpuPlan().create(m, nsys, n, 0, lay) as plan:
            restore()
            pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
            t = event_time(lambda: pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n), restore, 5,
                           torch.cuda.current_stream())
            ms = float(np.median(t))
            out[f"m{m}"] = {"auto_npart": plan.npart, "ms": ms, "h200x8_ms_published": h200,
                            "speedup_vs_h200x8": h200 / ms, "block_rows_per_s": nsys * n / (ms * 1e-3)}
        del a, b, c, d, a0, c0, d0
        torch.cuda.empty_cache()
    return out


def extra_small(dev, checks: ParityChecks) -> dict:
    """The small case: 1024 systems, m = 3, n = 64 (the CPU-runnable size), FP64 plain; all
    systems are checked against the oracle in the CPU leg."""
    import torch
    import pascal_btdma_amd as pb
    w = WORKLOADS["ddperm_m3"]
    m, n, nsys = w["m"], w["n"], w["nsys"]
    lay = pb.Layout.TILED
    a, b, c = (torch.empty(pb.layout_elems(lay, nsys, n, m * m), dtype=torch.float64, device=dev)
               for _ in range(3))
    d = torch.empty(pb.layout_elems(lay, nsys, n, m), dtype=torch.float64, device=dev)
    pb.generate(w["kind"], w["seed"], m, nsys, n, lay, a, b, c, d)
    c0, d0 = c.clone(), d.clone()
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, lay) as plan:
        def restore():
            c.copy_(c0)
            d.copy_(d0)
        restore()
        pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
        t = event_time(lambda: pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n), restore, 10,
                       torch.cuda.current_stream())
    x = tiled_vec_subset(d.cpu().numpy(), n, m, list(range(nsys // 64)))
    ms = float(np.median(t))
    # the same batch with the partition count chosen by the library (npart = 0)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 0, lay) as plan:
        a0 = a.clone()

        def restore_a():
            a.copy_(a0)
            c.copy_(c0)
            d.copy_(d0)
        ta = event_time(lambda: pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n), restore_a, 10,
                        torch.cuda.current_stream())
        auto_np = plan.npart
    out = {"workload": w["name"], "ms": ms, "block_rows_per_s": nsys * n / (ms * 1e-3),
           "auto_npart": auto_np, "ms_auto_npart": float(np.median(ta)),
           "max_rel_err_vs_oracle_all_systems": None}
    checks.add(lambda: pb.generate_host(w["kind"], w["seed"], m, nsys, n), x,
               max_sink(out, "max_rel_err_vs_oracle_all_systems"))
    # the same shape with the configuration's generator taken word for word (only B's rows
    # permuted; pb.GenKind.BPERM), solved by the same plan form
    pb.generate(pb.GenKind.BPERM, w["seed"], m, nsys, n, lay, a, b, c, d)
    c0, d0 = c.clone(), d.clone()
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, lay) as plan:
        def restore_b():
            c.copy_(c0)
            d.copy_(d0)
        restore_b()
        pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
        tb = event_time(lambda: pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n), restore_b, 10,
                        torch.cuda.current_stream())
    xb = tiled_vec_subset(d.cpu().numpy(), n, m, list(range(nsys // 64)))
    r = out["literal_generator"] = {
        "workload": "the same shape, B = U[-1,1] + diag(row sums |A|+|C| + 1) with only B's rows permuted",
        "ms": float(np.median(tb)), "max_rel_err_vs_oracle_all_systems": None,
        "max_rel_err_vs_truth_all_systems": None}

    def truth_sink(e):
        r["max_rel_err_vs_truth_all_systems"] = float(np.max(e))
    checks.add(lambda: pb.generate_host(pb.GenKind.BPERM, w["seed"], m, nsys, n), xb,
               max_sink(r, "max_rel_err_vs_oracle_all_systems"), truth_sink)
    return out


def mixed_batch(dev):
    """The 20000 mixed-size systems (m uniform in {2,3,5,8,10}, n log-uniform in
    [32,1024], seeded), packed and generated: (batch, m, n, a, b, c0, d0)."""
    import torch
    import pascal_btdma_amd as pb
    rng = np.random.default_rng(2026)
    nb = 20000
    m = rng.choice(np.array([2, 3, 5, 8, 10]), nb)
    n = np.rint(np.exp(rng.uniform(np.log(32), np.log(1024), nb))).astype(np.int32)
    bt = pb.BtdmaBatch(m, n)
    a, b, c0 = (torch.empty(bt.mat_elems, dtype=torch.float64, device=dev) for _ in range(3))
    d0 = torch.empty(bt.vec_elems, dtype=torch.float64, device=dev)
    bt.generate(pb.GenKind.DDPERM, 7, a, b, c0, d0)
    return bt, m, n, a, b, c0, d0


def batch_probe(dev, placement: int, grid: int, reps: int = 3) -> dict:
    """Only the 20000-system mixed-size batch, `reps` solves at fixed launch parameters (for profilers)."""
    import torch
    bt, m, n, a, b, c0, d0 = mixed_batch(dev)
    bt.set_params(placement, grid)
    c, d = c0.clone(), d0.clone()
    t = event_time(lambda: bt.solve(a, b, c, d), lambda: (c.copy_(c0), d.copy_(d0)), reps,
                   torch.cuda.current_stream())
    return {"batch_probe": {"placement": placement, "grid": grid, "ms": [float(x) for x in t],
                            "singular": bt.check()}}


def extra_batch(dev, steps: int, checks: ParityChecks) -> dict:
    """Mixed sizes plus long systems: 20000 systems with m uniform in {2,3,5,8,10}, n log-uniform in [32,1024],
    plus 512 long systems (m = 5, n = 65536) split into P partitions, all replayed as one
    captured graph (the batch's persistent kernel and the long systems' three kernels, one
    after the other or on forked branches). The autotuner picks the batch placement/grid and
    times P over 8..256; the reported graph uses the best of the configuration's P = 8, 16, 32,
    with the graph at the best P overall beside it. The graph's topology and, when forked, the
    batch's grid are picked by timing them."""
    import torch
    import pascal_btdma_amd as pb
    from pascal_btdma_amd.tuner import tune_batch, tune_partitions
    bt, m, n, a, b, c0, d0 = mixed_batch(dev)
    nb = len(m)
    tb = tune_batch(bt, a, b, c0, d0, reps=2)
    lm, lnsys, ln = 5, 512, 65536
    lay = pb.Layout.TILED
    la0, lb, lc0 = (torch.empty(pb.layout_elems(lay, lnsys, ln, lm * lm), dtype=torch.float64, device=dev)
                    for _ in range(3))
    ld0 = torch.empty(pb.layout_elems(lay, lnsys, ln, lm), dtype=torch.float64, device=dev)
    pb.generate(pb.GenKind.DDPERM, 8, lm, lnsys, ln, lay, la0, lb, lc0, ld0)
    tp = tune_partitions(lm, lnsys, ln, la0, lb, lc0, ld0, lay, candidates=(8, 16, 32, 64, 128, 256),
                         reps=2)
    # the configuration's own partition counts are 8, 16 and 32: its number uses the best of
    # those; the best count overall (the library's choice) is reported beside it
    stated = {int(k): v for k, v in tp["table_ms"].items() if int(k) in (8, 16, 32)}
    p_stated = min(stated, key=stated.get)
    c, d, la, lc, ld = c0.clone(), d0.clone(), la0.clone(), lc0.clone(), ld0.clone()

    def restore():
        c.copy_(c0)
        d.copy_(d0)
        la.copy_(la0)
        lc.copy_(lc0)
        ld.copy_(ld0)

    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()

    def capture(plan, forked: bool, grid: int):
        """One graph holding the batch's persistent kernel (at `grid` workgroups) and the
        long systems' three kernels, either on two forked branches or one after the other."""
        bt.set_params(tb["placement"], grid)
        g = pb.Graph()
        fork, join = torch.cuda.Event(), torch.cuda.Event()
        torch.cuda.synchronize()
        g.begin(s1)
        if forked:
            fork.record(s1)
            s2.wait_event(fork)
        bt.solve(a, b, c, d, stream=s1)
        pb.solve_noncyclic(plan, la, lb, lc, ld, lm, lnsys, ln, stream=s2 if forked else s1)
        if forked:
            join.record(s2)
            s1.wait_event(join)
        g.end(s1)
        restore()
        torch.cuda.synchronize()
        g.launch(s1)
        s1.synchronize()
        t = event_time(lambda: g.launch(s1), lambda: (restore(), torch.cuda.synchronize()), steps, s1)
        return g, float(np.median(t))

    # sequential at the tuned grid; forked also with the batch on one or two workgroups per CU,
    # which leaves the other SIMDs to the long systems' kernels
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    forked_grids = sorted({tb["grid"], cus, 2 * cus})
    candidates = [("sequential", tb["grid"])] + [("forked", gr) for gr in forked_grids]

    def best_graph(plan):
        graphs = {}
        for topo, gr in candidates:
            graphs[(topo, gr)] = capture(plan, topo == "forked", gr)
        key, (g, ms_best) = min(graphs.items(), key=lambda kv: kv[1][1])
        for gg, _ in graphs.values():
            if gg is not g:
                gg.destroy()
        return key, g, ms_best, {f"{k[0]},grid={k[1]}": v[1] for k, v in graphs.items()}

    plan_lib = pb.BtdmaGpuPlan().create(lm, lnsys, ln, tp["npart"], lay)
    _, g_lib, ms_lib, cand_lib = best_graph(plan_lib)
    g_lib.destroy()
    tl_lib = event_time(lambda: pb.solve_noncyclic(plan_lib, la, lb, lc, ld, lm, lnsys, ln), restore, 3,
                        torch.cuda.current_stream())
    plan_lib.destroy()
    plan = pb.BtdmaGpuPlan().create(lm, lnsys, ln, p_stated, lay)
    (topology, grid_used), g, ms_best, cand = best_graph(plan)
    ms_seq = cand[f"sequential,grid={tb['grid']}"]
    ms_fork = min(v for k, v in cand.items() if k.startswith("forked"))
    t = np.array([ms_best])
    # the batch alone (at the tuned grid) and the long systems alone, for the record
    bt.set_params(tb["placement"], tb["grid"])
    tb_only = event_time(lambda: bt.solve(a, b, c, d), restore, 3, torch.cuda.current_stream())
    tl_only = event_time(lambda: pb.solve_noncyclic(plan, la, lb, lc, ld, lm, lnsys, ln), restore, 3,
                         torch.cuda.current_stream())
    # parity on a sample of batch systems and 2 long systems
    restore()
    torch.cuda.synchronize()
    g.launch(s1)
    s1.synchronize()
    hd = d.cpu().numpy()
    xl = tiled_vec_subset(ld[: ln * lm * 64].cpu().numpy(), ln, lm, [0])[:, :2]
    g.destroy()
    plan.destroy()
    rows_b = int(n.sum())
    rows_l = lnsys * ln
    bytes_b = float(sum(8.0 * (3 * mm * mm + 2 * mm) * nn for mm, nn in zip(m.tolist(), n.tolist())))
    bytes_l = algorithmic_bytes(lm, rows_l)
    ms = float(np.median(t))
    out = {"workload": "20000 systems m in {2,3,5,8,10}, n log-uniform [32,1024] + 512 systems m=5 "
                        "n=65536 in P partitions (P the best of 8, 16, 32), one graph launch",
            "ms": ms, "npart_long": p_stated, "block_rows": rows_b + rows_l,
            "block_rows_per_s": (rows_b + rows_l) / (ms * 1e-3),
            "algorithmic_GBs": (bytes_b + bytes_l) / (ms * 1e-3) / 1e9,
            "hbm_frac": (bytes_b + bytes_l) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "graph_topology": topology, "graph_batch_grid": grid_used, "graph_forked_ms": ms_fork,
            "graph_sequential_ms": ms_seq,
            "graph_candidates_ms": cand,
            "batch_only_ms": float(np.median(tb_only)), "long_only_ms": float(np.median(tl_only)),
            "long_only_GBs": bytes_l / (np.median(tl_only) * 1e-3) / 1e9,
            "library_npart_long": tp["npart"], "library_npart_graph_ms": ms_lib,
            "library_npart_graph_candidates_ms": cand_lib,
            "library_npart_long_only_ms": float(np.median(tl_lib)),
            "library_npart_block_rows_per_s": (rows_b + rows_l) / (ms_lib * 1e-3),
            "tuned_batch": tb, "tuned_partitions": tp, "batch_packs": bt.npacks,
            "max_rel_err_vs_oracle_sampled": None,
            "parity_sample": "every 997th batch system and the first 2 long systems"}
    sink = max_sink(out, "max_rel_err_vs_oracle_sampled")
    for s in range(0, nb, 997):
        checks.add(lambda s=s: pb.generate_host(pb.GenKind.DDPERM, 7, int(m[s]), nb, int(n[s]), s, 1),
                   bt.extract(hd, s)[:, None, :], sink)
    checks.add(lambda: pb.generate_host(pb.GenKind.DDPERM, 8, lm, lnsys, ln, 0, 2), xl, sink)
    return out


def extra_f32r(dev, steps: int, fp64_ms: float, checks: ParityChecks) -> dict:
    """FP32-stored conditioned blocks: m = 5, n = 512, 131072 systems, blocks with cond log-uniform [1e1, 1e6]
    stored in FP32. Modes: FP32 solve + s FP64-residual refinement sweeps (s = 0..8), FP32
    storage with FP64 arithmetic (no refinement), and the FP64 plain kernel on the same
    (FP32-rounded) values; times against the FP64 plain kernel's time on the headline workload. Every
    mode's error on a sample of systems is measured against the oracle on the same inputs and
    against the binary128 truth solve of those inputs (the oracle is itself an FP64 solve,
    about cond * 1e-16 from the exact solution)."""
    import torch
    import pascal_btdma_amd as pb
    m, n, nsys = 5, 512, 131072
    lay = pb.Layout.TILED
    a, b, c = (torch.empty(pb.layout_elems(lay, nsys, n, m * m), dtype=torch.float32, device=dev)
               for _ in range(3))
    d = torch.empty(pb.layout_elems(lay, nsys, n, m), dtype=torch.float64, device=dev)
    pb.generate_f32(pb.GenKind.COND, 3, m, nsys, n, lay, a, b, c, d)
    x = torch.empty_like(d)
    out = {"workload": "cond-controlled B (1e1..1e6), m=5, n=512, 131072 systems, FP32 blocks",
           "fp64_plain_ms_headline": fp64_ms}
    sample = lambda: pb.generate_host(pb.GenKind.COND, 3, m, nsys, n, 0, 128)

    def error_sinks(r):
        def sink(err, xd=None, xo=None):
            r["max_rel_err_vs_oracle"] = float(err.max())
            r["median_rel_err_vs_oracle"] = float(np.median(err))
            r["frac_systems_le_1e-12_vs_oracle"] = float((err <= 1e-12).mean())
            if xd is not None:
                r["bit_identical_to_oracle"] = bool(np.array_equal(xd, xo))

        def truth(err):
            r["max_rel_err_vs_truth"] = float(err.max())
            r["median_rel_err_vs_truth"] = float(np.median(err))
            r["frac_systems_le_1e-12_vs_truth"] = float((err <= 1e-12).mean())
        return sink, truth

    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, lay) as plan:
        work = torch.empty(pb.f32r_work_elems(plan), dtype=torch.float32, device=dev)
        for sweeps in range(9):
            pb.solve_noncyclic_f32r(plan, a, b, c, d, x, work, sweeps)
            t = event_time(lambda: pb.solve_noncyclic_f32r(plan, a, b, c, d, x, work, sweeps),
                           lambda: None, steps, torch.cuda.current_stream())
            xs = tiled_vec_subset(x[: 2 * n * m * 64].cpu().numpy(), n, m, [0, 1])
            ms = float(np.median(t))
            r = out[f"sweeps{sweeps}"] = {"ms": ms, "block_rows_per_s": nsys * n / (ms * 1e-3),
                                          "speedup_vs_fp64_plain": fp64_ms / ms}
            checks.add(sample, xs, *error_sinks(r))
        # FP32 storage with FP64 arithmetic: no refinement, the FP64 solve of the stored values
        work64 = torch.empty(plan.matrix_elems(), dtype=torch.float64, device=dev)
        xs_ = torch.empty_like(d)
        t = event_time(lambda: pb.solve_noncyclic_f32s(plan, a, b, c, xs_, work64),
                       lambda: xs_.copy_(d), steps, torch.cuda.current_stream())
        ms = float(np.median(t))
        r = out["fp32_storage_fp64_arithmetic"] = {
            "ms": ms, "block_rows_per_s": nsys * n / (ms * 1e-3), "speedup_vs_fp64_plain": fp64_ms / ms,
            "bytes_per_row_model": 28 * m * m + 32 * m}
        xs = tiled_vec_subset(xs_[: 2 * n * m * 64].cpu().numpy(), n, m, [0, 1])
        checks.add(sample, xs, *error_sinks(r))
        # the same solve with C' in the plan's workspace, placed for this solve by timing
        # candidates (btd_plan_place_workspace_f32s), as the headline's FP64 workspace is
        x_caller = xs_.clone()
        del work64
        plan.set_form(pb.Form.WORKSPACE)
        cand = plan.place_workspace_f32s(a, b, c, d, candidates=8)
        t = event_time(lambda: pb.solve_noncyclic_f32s(plan, a, b, c, xs_),
                       lambda: xs_.copy_(d), steps, torch.cuda.current_stream())
        torch.cuda.synchronize()
        ms = float(np.median(t))
        out["fp32_storage_fp64_arithmetic_placed"] = {
            "ms": ms, "block_rows_per_s": nsys * n / (ms * 1e-3), "speedup_vs_fp64_plain": fp64_ms / ms,
            "placement_candidates_ms": cand, "x_bit_identical_to_caller_workspace": bool(torch.equal(xs_, x_caller))}
        plan.set_form(pb.Form.PLAIN)
        del x_caller, xs_
        # the same mode on the headline's well-conditioned Euler-like blocks rounded to FP32
        pb.generate_f32(pb.GenKind.EULER, 1, m, nsys, n, lay, a, b, c, d)
        work64 = torch.empty(plan.matrix_elems(), dtype=torch.float64, device=dev)
        xs_ = torch.empty_like(d)
        t = event_time(lambda: pb.solve_noncyclic_f32s(plan, a, b, c, xs_, work64),
                       lambda: xs_.copy_(d), steps, torch.cuda.current_stream())
        ms = float(np.median(t))
        r = out["fp32_storage_fp64_arithmetic_euler_blocks"] = {
            "workload": "the headline's Euler-like m=5, n=512, 131072 systems with A, B, C rounded to FP32",
            "ms": ms, "speedup_vs_fp64_plain": fp64_ms / ms}

        def euler32():
            A_, B_, C_, d_ = pb.generate_host(pb.GenKind.EULER, 1, m, nsys, n, 0, 128)
            return (*(X.astype(np.float32).astype(np.float64) for X in (A_, B_, C_)), d_)
        xs = tiled_vec_subset(xs_[: 2 * n * m * 64].cpu().numpy(), n, m, [0, 1])
        checks.add(euler32, xs, *error_sinks(r))
        del work64, xs_
        pb.generate_f32(pb.GenKind.COND, 3, m, nsys, n, lay, a, b, c, d)
        # the FP64 plain kernel on the same (FP32-rounded) values, timed
        a64, b64, c64 = a.double(), b.double(), c.double()
        c640, d64 = c64.clone(), d.clone()
        t = event_time(lambda: pb.solve_noncyclic(plan, a64, b64, c64, d64, m, nsys, n),
                       lambda: (c64.copy_(c640), d64.copy_(d)), steps, torch.cuda.current_stream())
        ms = float(np.median(t))
        r = out["fp64_plain_same_inputs"] = {"ms": ms, "speedup_vs_fp64_plain": fp64_ms / ms}
        xs = tiled_vec_subset(d64[: 2 * n * m * 64].cpu().numpy(), n, m, [0, 1])
        checks.add(sample, xs, *error_sinks(r))
        del a64, b64, c64, c640, d64
    out["error_sample"] = ("first 128 systems, against the oracle on the same FP32-rounded blocks and "
                           "against the binary128 truth solve of those blocks")
    return out


def extra_adjoint(dev, steps: int) -> dict:
    """The differentiable solve: WENO-like m = 3, n = 256, 262144 systems through the autograd Functions:
    forward (with factors kept) and backward (adjoint kernel) times, raw kernel times, in
    FP64 and with FP32-stored blocks + FP64 refinement. The blocks are rounded to FP32
    values first, so that both variants solve the same systems."""
    import torch
    import pascal_btdma_amd as pb
    from pascal_btdma_amd.autograd import block_tridiagonal_solve, block_tridiagonal_solve_f32r
    m, n, nsys = 3, 256, 262144
    lay = pb.Layout.TILED
    a, b, c = (torch.empty(pb.layout_elems(lay, nsys, n, m * m), dtype=torch.float64, device=dev)
               for _ in range(3))
    d = torch.empty(pb.layout_elems(lay, nsys, n, m), dtype=torch.float64, device=dev)
    pb.generate(pb.GenKind.EULER, 4, m, nsys, n, lay, a, b, c, d)
    for t_ in (a, b, c):
        t_.copy_(t_.float().double())
    w = torch.rand_like(d) * 2 - 1
    plan = pb.BtdmaGpuPlan().create(m, nsys, n, 1, lay)
    for t_ in (a, b, c, d):
        t_.requires_grad_(True)
    s = torch.cuda.current_stream()
    fwd, bwd = [], []
    for it in range(steps + 1):
        for t_ in (a, b, c, d):
            t_.grad = None
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(s)
        x = block_tridiagonal_solve(plan, a, b, c, d)
        e1.record(s)
        x.backward(w)
        e2.record(s)
        torch.cuda.synchronize()
        if it > 0:
            fwd.append(e0.elapsed_time(e1))
            bwd.append(e1.elapsed_time(e2))
        del x
    with torch.no_grad():
        cw, xw = c.detach().clone(), d.detach().clone()
        F = torch.empty_like(cw)
        P = torch.empty(pb.layout_elems(lay, nsys, n, 1), dtype=torch.int64, device=dev)
        ad = a.detach()
        kf = event_time(lambda: pb.solve_noncyclic_save(plan, ad, b.detach(), cw, xw, F, P),
                        lambda: (cw.copy_(c.detach()), xw.copy_(d.detach())), steps, s)
        lam = torch.empty_like(xw)
        da, db, dc = (torch.empty_like(ad) for _ in range(3))
        kb = event_time(lambda: pb.solve_adjoint(plan, ad, cw, F, P, xw, w, lam, da, db, dc),
                        lambda: None, steps, s)
        x64, lam64 = xw.clone(), lam.clone()
        del cw, F, P, da, db, dc, ad
    # FP32-stored blocks, 2 refinement sweeps each way
    a32, b32, c32 = (t_.detach().float().requires_grad_(True) for t_ in (a, b, c))
    for t_ in (a, b, c):
        t_.grad = None
    fwd32, bwd32 = [], []
    for it in range(steps + 1):
        for t_ in (a32, b32, c32, d):
            t_.grad = None
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(s)
        x = block_tridiagonal_solve_f32r(plan, a32, b32, c32, d, 2)
        e1.record(s)
        x.backward(w)
        e2.record(s)
        torch.cuda.synchronize()
        if it > 0:
            fwd32.append(e0.elapsed_time(e1))
            bwd32.append(e1.elapsed_time(e2))
        if it < steps:
            del x
    err_x = float(((x.detach() - x64).abs().max() / x64.abs().max()).item())
    err_l = float(((d.grad - lam64).abs().max() / lam64.abs().max()).item())
    del x
    plan.destroy()
    rows = nsys * n
    f_ms, b_ms = float(np.median(fwd)), float(np.median(bwd))
    kf_ms, kb_ms = float(np.median(kf)), float(np.median(kb))
    f32_ms, b32_ms = float(np.median(fwd32)), float(np.median(bwd32))
    return {"workload": "WENO-like m=3, n=256, 262144 systems, FP64, autograd",
            "forward_ms": f_ms, "backward_ms": b_ms, "backward_over_forward": b_ms / f_ms,
            "kernel_forward_save_ms": kf_ms, "kernel_adjoint_ms": kb_ms,
            "kernel_backward_over_forward": kb_ms / kf_ms,
            "forward_block_rows_per_s": rows / (f_ms * 1e-3),
            "kernel_adjoint_alg_GBs": 8.0 * (6 * m * m + 5 * m) * rows / (kb_ms * 1e-3) / 1e9,
            "fp32r_forward_ms": f32_ms, "fp32r_backward_ms": b32_ms,
            "fp32r_backward_over_forward": b32_ms / f32_ms,
            "fp32r_speedup_vs_fp64": (f_ms + b_ms) / (f32_ms + b32_ms),
            "fp32r_max_rel_err_x_vs_fp64": err_x, "fp32r_max_rel_err_lambda_vs_fp64": err_l}


def euler_state(dev, n1: int, n2: int, n3: int):
    """A smooth periodic conserved state on an n1 x n2 x n3 grid with its one-cell periodic
    halo, as a flat float64 device tensor in Fortran order U(0:n1+1, 0:n2+1, 0:n3+1, 5)."""
    import torch
    two_pi, gamma = 2 * np.pi, 1.4
    ax = [(torch.arange(n, dtype=torch.float64, device=dev) + 0.5) / n for n in (n3, n2, n1)]
    Z, Y, X = torch.meshgrid(*ax, indexing="ij")
    rho = 1.0 + 0.2 * torch.sin(two_pi * X) * torch.cos(two_pi * Y)
    u, v, w = 0.3 * torch.sin(two_pi * Y), 0.3 * torch.cos(two_pi * Z), 0.2 * torch.sin(two_pi * X)
    p = 1.0 + 0.1 * torch.cos(two_pi * (X + Y + Z))
    E = p / ((gamma - 1.0) * rho) + 0.5 * (u * u + v * v + w * w)
    inner = torch.stack([rho, rho * u, rho * v, rho * w, rho * E])  # (5, n3, n2, n1)
    for dim in (1, 2, 3):  # periodic halo
        inner = torch.cat([inner.narrow(dim, inner.shape[dim] - 1, 1), inner, inner.narrow(dim, 0, 1)], dim)
    return inner.contiguous().reshape(-1)


def extra_euler(dev, steps: int, n: int = 256) -> dict:
    """The euler3d ADI step's three periodic sweeps (z, y, x) on an n^3 grid, m = 5: the
    assembled path (btd_euler_assemble -> btd_solve_cyclic -> btd_sweep_scatter) and the
    matrix-free sweep (btd_euler_sweep_cyclic), per sweep, with the agreement of the two; plus
    the device time of the transposes the reference runs around the y and x sweeps
    (euler3d_mpi_gpu.f90:1287-1325: dU, trA, trB, trC in and dU out), timed as torch copies of
    the same arrays, which the assembly makes unnecessary."""
    import torch
    import pascal_btdma_amd as pb
    n1 = n2 = n3 = n
    h, dt, eps = 1.0 / n, 0.5 / n, 0.01
    lay = pb.Layout.TILED
    s = torch.cuda.current_stream()
    U = euler_state(dev, n1, n2, n3)
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    rhs = torch.randn(5 * n1 * n2 * n3, dtype=torch.float64, device=dev, generator=g)
    out = {"workload": f"euler3d ADI step, {n}^3 grid, m=5, periodic sweeps, FP64",
           "cells": n1 * n2 * n3}
    tot = {"assembled": 0.0, "matrix_free": 0.0, "reference_transposes": 0.0}
    for axis, name in ((2, "z"), (1, "y"), (0, "x")):
        nrow, nsys = pb.sweep_shape(axis, n1, n2, n3)
        with pb.BtdmaGpuPlan().create(5, nsys, nrow, 1, lay) as plan:
            a, b, c = (torch.empty(plan.matrix_elems(), dtype=torch.float64, device=dev) for _ in range(3))
            d = torch.empty(plan.vector_elems(), dtype=torch.float64, device=dev)
            x1 = torch.empty_like(rhs)
            t_as = event_time(lambda: pb.euler_assemble(plan, axis, U, rhs, n1, n2, n3, h, dt, eps, a, b, c, d),
                              lambda: None, steps, s)
            t_sv = event_time(lambda: pb.solve_cyclic(plan, a, b, c, d, 5, nsys, nrow),
                              lambda: pb.euler_assemble(plan, axis, U, rhs, n1, n2, n3, h, dt, eps,
                                                        a, b, c, d), steps, s)
            t_sc = event_time(lambda: pb.sweep_scatter(plan, axis, d, n1, n2, n3, x1), lambda: None, steps, s)
            del a, b, c
            work = torch.empty(pb.euler_sweep_work_elems(plan), dtype=torch.float64, device=dev)
            x2 = torch.empty_like(rhs)
            t_mf = event_time(lambda: pb.euler_sweep_cyclic(plan, axis, U, rhs, n1, n2, n3, h, dt, eps,
                                                            work, x2), lambda: None, steps, s)
            torch.cuda.synchronize()
            assert not plan.check()
            diff = float(((x1 - x2).abs().max() / x1.abs().max()).item())
            del work, d
        # the reference's transposes for this sweep (none for z)
        t_tr = [0.0]
        if axis != 2:
            blk = torch.empty((25, n3, n2, n1), dtype=torch.float64, device=dev)
            vec = rhs.view(5, n3, n2, n1)
            perm = (0, 2, 1, 3) if axis == 1 else (0, 3, 1, 2)
            back = (0, 2, 1, 3) if axis == 1 else (0, 2, 3, 1)

            def transposes():
                for _ in range(3):
                    blk.permute(perm).contiguous()
                vv = vec.permute(perm).contiguous()
                vv.permute(back).contiguous()
            t_tr = event_time(transposes, lambda: None, steps, s)
            del blk
        torch.cuda.empty_cache()
        r = {"assemble_ms": float(np.median(t_as)), "solve_cyclic_ms": float(np.median(t_sv)),
             "scatter_ms": float(np.median(t_sc)), "matrix_free_ms": float(np.median(t_mf)),
             "reference_transposes_ms": float(np.median(t_tr)),
             "matrix_free_vs_assembled_max_rel_diff": diff}
        r["assembled_ms"] = r["assemble_ms"] + r["solve_cyclic_ms"] + r["scatter_ms"]
        r["matrix_free_speedup_vs_assembled"] = r["assembled_ms"] / r["matrix_free_ms"]
        r["matrix_free_GBs_traffic_model"] = 8.0 * 125 * nsys * nrow / (r["matrix_free_ms"] * 1e-3) / 1e9
        out[f"sweep_{name}"] = r
        tot["assembled"] += r["assembled_ms"]
        tot["matrix_free"] += r["matrix_free_ms"]
        tot["reference_transposes"] += r["reference_transposes_ms"]
    out["adi_step_assembled_ms"] = tot["assembled"]
    out["adi_step_matrix_free_ms"] = tot["matrix_free"]
    out["adi_step_reference_style_estimate_ms"] = tot["assembled"] + tot["reference_transposes"]
    out["matrix_free_speedup_vs_reference_style_estimate"] = \
        out["adi_step_reference_style_estimate_ms"] / tot["matrix_free"]
    out["cells_per_s_matrix_free_adi_step"] = out["cells"] / (tot["matrix_free"] * 1e-3)
    out["traffic_model"] = ("matrix-free sweep: 125 doubles per block row (U 5, rhs 5, C'/E'/d' "
                            "55 written and 55 read, x 5); assembled: assembly 90 + periodic solve "
                            "195 (two-stage over the plan's periodic partitions: A, B, C, d 80 read, "
                            "Â/Ĉ/d̂ 55 written and 55 read, x 5) + scatter 10")
    return out


def extra_weno(dev, steps: int, n: int = 3072) -> dict:
    """The weno2d example's implicit sweeps on an n x n grid, tiled layout, per-line speeds of
    both signs: the assembly of one x and one y sweep's m = 3 blocks from the haloed scalar
    field (btd_weno_assemble: device time and the rate of the block bytes written), and the
    whole sweep (gather, assemble, periodic solve over the plan's partitions, scatter), assembled
    and matrix-free (btd_weno_sweep_cyclic: the blocks computed inside the solve's stage 1)."""
    import torch
    import pascal_btdma_amd as pb
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    U = torch.rand((n + 6) * (n + 6), dtype=torch.float64, device=dev, generator=g)
    vel = torch.linspace(-1.0, 1.0, n, dtype=torch.float64, device=dev)
    out = {"workload": f"weno2d sweep assembly, {n}^2 grid, m=3, frozen WENO5, FP64"}
    s = torch.cuda.current_stream()
    for axis, name in ((0, "x"), (1, "y")):
        with pb.BtdmaGpuPlan().create(3, n, n // 3, 1, pb.Layout.TILED) as plan:
            a, b, c = (torch.empty(plan.matrix_elems(), dtype=torch.float64, device=dev) for _ in range(3))
            run = lambda: pb.weno_assemble(plan, axis, U, vel, n, n, 1.0 / n, 0.5 / n, a, b, c)
            run()
            t = event_time(run, lambda: None, steps, s)
            ms = float(np.median(t))
            written = 3 * 9 * 8 * n * (n // 3)
            out[f"sweep_{name}_ms"] = ms
            out[f"sweep_{name}_blocks_written_GBs"] = written / (ms * 1e-3) / 1e9
            # the whole sweep: field -> vector, blocks, periodic solve, vector -> field
            d = torch.empty(plan.vector_elems(), dtype=torch.float64, device=dev)
            F = torch.rand(n * n, dtype=torch.float64, device=dev, generator=g)

            def sweep():
                pb.weno_gather(plan, axis, F, n, n, d)
                pb.weno_assemble(plan, axis, U, vel, n, n, 1.0 / n, 0.5 / n, a, b, c)
                pb.solve_cyclic(plan, a, b, c, d, 3, n, n // 3)
                pb.weno_scatter(plan, axis, d, n, n, F)
            sweep()
            t = event_time(sweep, lambda: None, steps, s)
            out[f"sweep_{name}_pipeline_ms"] = float(np.median(t))
            # the same sweep matrix-free: the blocks computed inside stage 1, never stored
            work = torch.empty(2 * plan.matrix_elems(), dtype=torch.float64, device=dev)

            def sweep_mf():
                pb.weno_gather(plan, axis, F, n, n, d)
                pb.weno_sweep_cyclic(plan, axis, U, vel, n, n, 1.0 / n, 0.5 / n, work, d)
                pb.weno_scatter(plan, axis, d, n, n, F)
            sweep_mf()
            t = event_time(sweep_mf, lambda: None, steps, s)
            out[f"sweep_{name}_matrix_free_ms"] = float(np.median(t))
            if plan.check():
                out[f"sweep_{name}_singular"] = True
            del a, b, c, d, F, work
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="euler_m5", choices=sorted(WORKLOADS))
    ap.add_argument("--layout", default="tiled", choices=["tiled", "ref_gpu"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="headline workload only")
    ap.add_argument("--only", default="",
                    help="comma list of extras to run (study,small,batch,f32r,adjoint,euler,weno)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--workspace-candidates", type=int, default=8,
                    help="workspace allocations btd_plan_place_workspace tries for the headline")
    ap.add_argument("--batch-probe", default="",
                    help="PLACEMENT,GRID: only solve the 20000-system mixed-size batch 3 times at these launch "
                         "parameters and print the times (for profilers)")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="only copy one buffer of this many GiB to another (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_gib > 0:
        import torch
        nel = int(args.copy_probe_gib * 2**30) // 8
        src = torch.empty(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": nel * 8, "copy_probe_bytes_written": nel * 8}))
        return 0
    if args.batch_probe:
        import torch
        pl, gr = (int(v) for v in args.batch_probe.split(","))
        print(json.dumps(batch_probe(torch.device("cuda:0"), pl, gr)))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(independent systems shard across GPUs with no communication)",
              file=sys.stderr)
        return 2

    import torch

    import pascal_btdma_amd as pb

    checks = ParityChecks()
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

    # the in-place solve (the reference's contract: C' over C), then the workspace form (C' to a
    # plan-owned array whose placement btd_plan_place_workspace picks; C only read): both are
    # k_plain, the same arithmetic and the same bits; the faster form is the headline
    for _ in range(args.warmup):
        restore()
        pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
    torch.cuda.synchronize()
    t_in = event_time(lambda: pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n), restore, args.steps, stream)
    sing_in = plan.check()
    x_in = d.clone()
    restore()
    plan.set_form(pb.Form.WORKSPACE)
    cand_ms = plan.place_workspace(a, b, c, d, candidates=args.workspace_candidates)
    plan.check()  # the placement's trial solves are checked with the timed ones below
    for _ in range(args.warmup):
        d.copy_(d0)
        pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t_ws = event_time(lambda: pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n), lambda: d.copy_(d0),
                      args.steps, stream)
    wall = time.perf_counter() - t0
    singular = plan.check() or sing_in
    same_bits = bool(torch.equal(d, x_in))
    c_untouched = bool(torch.equal(c, c0))
    del x_in
    form = "workspace" if t_ws.mean() <= t_in.mean() else "in_place"
    times = t_ws if form == "workspace" else t_in
    ms = float(times.mean())
    rows = nsys * n
    value = rows / (ms * 1e-3)
    achieved = algorithmic_bytes(m, rows) / (ms * 1e-3) / 1e9
    ms_in = float(t_in.mean())
    forms = {"headline_form": form,
             "in_place": {"ms": ms_in, "ms_min": float(t_in.min()), "ms_max": float(t_in.max()),
                          "block_rows_per_s": rows / (ms_in * 1e-3),
                          "frac": algorithmic_bytes(m, rows) / (ms_in * 1e-3) / 1e9 / HBM_PEAK_GBS},
             "workspace": {"ms": float(t_ws.mean()), "ms_min": float(t_ws.min()), "ms_max": float(t_ws.max()),
                           "placement_candidates_ms": cand_ms,
                           "frac": algorithmic_bytes(m, rows) / (float(t_ws.mean()) * 1e-3) / 1e9 / HBM_PEAK_GBS,
                           "x_bit_identical_to_in_place": same_bits, "c_unchanged": c_untouched}}

    # measured copy bandwidth on this device (torch D2D copy of C: read + write)
    tc = event_time(lambda: c.copy_(c0), lambda: None, 3, stream)
    copy_gbs = 2 * nm * 8 / (float(np.median(tc)) * 1e-3) / 1e9
    d.copy_(d0)
    pb.solve_noncyclic(plan, a, b, c, d, m, nsys, n)
    torch.cuda.synchronize()

    # parity on a subset of systems: the first two and the last tiles
    ntiles = (nsys + 63) // 64
    tiles = sorted({0, 1, ntiles - 1})
    if lay == pb.Layout.TILED:
        tile_v = n * m * 64
        x = np.concatenate([tiled_vec_subset(d[t * tile_v:(t + 1) * tile_v].cpu().numpy(), n, m, [0])
                            for t in tiles], axis=1)
        sys_ids = [t * 64 + j for t in tiles for j in range(64)]
    else:
        full = d.view(m, n, nsys)
        sys_ids = [s for t in tiles for s in range(t * 64, min(nsys, t * 64 + 64))]
        x = full[:, :, sys_ids].permute(1, 2, 0).cpu().numpy()
    parity = {"max_rel_err_vs_oracle": None}
    for lo in range(0, len(sys_ids), 64):
        s0 = sys_ids[lo]
        checks.add(lambda s0=s0: pb.generate_host(w["kind"], w["seed"], m, nsys, n, s0, 64),
                   x[:, lo:lo + 64], max_sink(parity, "max_rel_err_vs_oracle", "bit_identical_to_oracle"))
    plan.destroy()
    del a, b, c, d, c0, d0
    torch.cuda.empty_cache()

    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_plain_{args.workload}_{args.layout}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")

    extras = {}
    if not args.no_extras:
        only = set(filter(None, args.only.split(",")))
        jobs = [("study", lambda: extra_study(dev)),
                ("study_large", lambda: extra_study_large(dev)),
                ("small", lambda: extra_small(dev, checks)),
                ("batch", lambda: extra_batch(dev, 3, checks)),
                ("f32r", lambda: extra_f32r(dev, 3, ms, checks)),
                ("adjoint", lambda: extra_adjoint(dev, 3)),
                ("euler", lambda: extra_euler(dev, 3)),
                ("weno", lambda: extra_weno(dev, 5))]
        for name, job in jobs:
            if only and name not in only:
                continue
            t_start = time.perf_counter()
            try:
                extras[name] = job()
            except Exception as e:  # report, keep the headline line
                extras[name] = {"error": f"{type(e).__name__}: {e}",
                                "trace": traceback.format_exc(limit=3)}
            extras[name]["bench_seconds"] = round(time.perf_counter() - t_start, 1)
            torch.cuda.empty_cache()
            print(f"[bench] extra {name} done in {extras[name]['bench_seconds']} s", file=sys.stderr,
                  flush=True)

    # the CPU leg: the oracle timed on the host and every recorded device result checked
    cpu = cpu_leg(w, args.cpu_seconds, checks, baseline=not args.no_cpu)
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
        "max_rel_err_vs_oracle": parity["max_rel_err_vs_oracle"],
        "bit_identical_to_oracle": parity.get("bit_identical_to_oracle"), "parity_systems": len(sys_ids),
        "singular": singular, "ms_min": float(times.min()), "ms_max": float(times.max()),
        "wall_s_timed_region_incl_restore": wall,
        "measured_copy_GBs": copy_gbs,
        "kernel": f"k_plain<{m},{'TILED' if lay == pb.Layout.TILED else 'REF_GPU'}>",
        "forms": forms,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
