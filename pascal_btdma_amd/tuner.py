"""Autotuner for the launch choices of the solver, by measurement on the device.

* ``tune_batch``: for a mixed-size batch, the placement (one lane per system with blocks
  in VGPRs or in LDS, m lanes per system from a block size up, or the same with the longest
  systems solved twisted from both ends) and the persistent grid size (64-lane workgroups
  per CU).
* ``tune_partitions``: for long systems, the number of row partitions P of the two-stage
  solve (P partitions per system become P lanes; the reduced system has 2P rows).

Each candidate runs on copies of the caller's inputs (the solves are in place), is timed
with HIP events over ``reps`` runs after one warm-up, and the fastest is applied/returned
with the whole table of timings.
"""
from __future__ import annotations

import torch

from . import BtdmaBatch, BtdmaGpuPlan, Layout, solve_noncyclic


def _cus() -> int:
    return torch.cuda.get_device_properties(0).multi_processor_count


def _time(fn, restore, reps: int) -> float:
    restore()
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        restore()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def tune_batch(batch: BtdmaBatch, a, b, c, d, reps: int = 3,
               placements=(0, 1, 2, 3, 4, 5, 6), grids_per_cu=(1, 2, 3, 4)) -> dict:
    """Pick (placement, grid) for ``batch``; a, b, c, d are its packed inputs (unchanged)."""
    cw, dw = c.clone(), d.clone()

    def restore():
        cw.copy_(c)
        dw.copy_(d)

    cus = _cus()
    table = {}
    for pl in placements:
        if pl == 1 and int(batch.m.max()) < 6:
            continue
        for g in grids_per_cu:
            grid = max(1, g * cus)
            batch.set_params(pl, grid)
            table[(pl, grid)] = _time(lambda: batch.solve(a, b, cw, dw), restore, reps)
    best = min(table, key=table.get)
    batch.set_params(*best)
    return {"placement": best[0], "grid": best[1], "ms": table[best],
            "table_ms": {f"placement={k[0]},grid={k[1]}": v for k, v in table.items()}}


def tune_partitions(m: int, nsys: int, nrow: int, a, b, c, d, layout: Layout = Layout.TILED,
                    candidates=(8, 16, 32), reps: int = 3) -> dict:
    """Pick the partition count P for long systems; a, b, c, d are the inputs (unchanged)."""
    aw, cw, dw = a.clone(), c.clone(), d.clone()

    def restore():
        aw.copy_(a)
        cw.copy_(c)
        dw.copy_(d)

    table = {}
    for P in candidates:
        if nrow < 3 * P:
            continue
        with BtdmaGpuPlan().create(m, nsys, nrow, P, layout) as plan:
            table[P] = _time(lambda: solve_noncyclic(plan, aw, b, cw, dw, m, nsys, nrow), restore, reps)
    best = min(table, key=table.get)
    return {"npart": best, "ms": table[best], "table_ms": {str(k): v for k, v in table.items()}}
