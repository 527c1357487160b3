"""Study-compatible profiling driver: the reference Study's measurement, written in its CSV
schemas so that results line up with its tables and its analysis script.

The reference driver (Study/example_cuda_cxx_btdma_profile.cu) solves the diagonal workload
(A = 0.25 I, B = 2 I, C = 0.25 I, d = 1) of an n1 x n2 x n3 grid: nsys = n1*n2 systems of n3
block rows, split over nranks MPI ranks along the rows. Here the ranks become row partitions of
one GPU (npart = nranks, the two-stage solve with in-HBM partitions). Per case it restores the
inputs, runs `iterations` profiled solves and writes

* one timing row per iteration (Study/README.md, timing CSV schema): the six phase fields of
  btd_solve_noncyclic_profiled in seconds (exchanges are 0: there are none), their compute and
  communication sums, and total_s_avg = total_s_max (one device);
* one solution-signature row from iteration 0 (signature CSV schema): sum, L2 norm and max-norm
  of the solution and x at system 0, element 0 of rows 0, n3/2 and n3-1.

`implementation` is "hip-mi355x" and `mpi_mode` "in-hbm", so the reference's analysis keeps
these rows apart from its cuda-cxx / fortran-original rows. With --npart auto every case runs
with the partition count the library chooses for one GPU (btd_auto_npart) instead of nranks;
those rows keep the case's nranks and carry mpi_mode "in-hbm-np<P>" and the partitions' row counts.
With --form twisted or lapack_order, cases solved in one partition run that form of the
single-partition solve (btd_plan_set_form: two groups of m lanes per system from both ends, or
one lane per system in the CPU path's operation order) instead of the library's choice, and
their mpi_mode ends in "-twisted" / "-lapack_order"; cases with more partitions are refused.

    python -m pascal_btdma_amd.study --profile-out P.csv --signature-out S.csv [--iterations 10]
        [--case nranks,n1,n2,n3,m ...] [--npart auto] [--form plain|twisted|lapack_order]
"""
from __future__ import annotations

import argparse
import math
import os
import sys

PROFILE_HEADER = ("solver,variant,implementation,nranks,n1,n2,n3,m,nsys,nrow_min,nrow_max,iter,"
                  "iterations,mpi_mode,total_s_max,total_s_avg,local_compute_s_max,"
                  "forward_exchange_s_max,reduced_compute_s_max,backward_exchange_s_max,"
                  "update_compute_s_max,compute_s_max,communication_s_max")
SIGNATURE_HEADER = ("solver,variant,implementation,nranks,n1,n2,n3,m,nsys,nrow_min,nrow_max,mpi_mode,"
                    "solution_sum,solution_l2,solution_linf,sample_z0,sample_zmid,sample_zlast")
IMPLEMENTATION = "hip-mi355x"
MPI_MODE = "in-hbm"

# The reference Study's case matrix (Study/result/btdma_full_case_list_260702_193138.csv):
# (nranks, n1, n2, n3, m), in its case order.
STUDY_CASES = [
    (1, 32, 32, 128, 2), (1, 32, 32, 128, 3), (1, 32, 32, 128, 5), (1, 32, 32, 128, 8),
    (2, 64, 64, 2048, 2), (4, 64, 64, 2048, 2), (8, 64, 64, 2048, 2),
    (2, 64, 64, 2048, 3), (4, 64, 64, 2048, 3), (8, 64, 64, 2048, 3),
    (2, 64, 64, 2048, 5), (4, 64, 64, 2048, 5), (8, 64, 64, 2048, 5),
    (2, 64, 64, 2048, 8), (4, 64, 64, 2048, 8), (8, 64, 64, 2048, 8),
    (2, 64, 64, 1024, 2), (8, 64, 64, 4096, 2), (2, 64, 64, 1024, 5), (8, 64, 64, 4096, 5),
    (2, 64, 64, 1024, 8), (8, 64, 64, 4096, 8),
    (8, 32, 32, 2048, 2), (8, 128, 128, 2048, 2), (8, 32, 32, 2048, 5), (8, 128, 128, 2048, 5),
    (8, 32, 32, 2048, 8), (8, 128, 128, 2048, 8),
]


def _fmt(x: float, digits: int) -> str:
    """A number as the reference driver's iostream with setprecision(digits) prints it."""
    return f"{x:.{digits}g}"


FORMS = ("plain", "twisted", "lapack_order")


def run_case(case, iterations: int = 10, device=None, auto_npart: bool = False, form: str = "plain"):
    """Profile one case; returns (timing rows, signature row) as lists of CSV strings. `form`
    (one of FORMS) selects the single-partition solve's form; other than "plain" it needs a
    case solved in one partition."""
    import torch
    import pascal_btdma_amd as pb
    nranks, n1, n2, n3, m = case
    nsys, nrow = n1 * n2, n3
    npart = pb._capi.lib().btd_auto_npart(m, nsys, nrow) if auto_npart else nranks
    if form not in FORMS:
        raise ValueError(f"form must be one of {FORMS}, got {form!r}")
    if form != "plain" and npart != 1:
        raise ValueError(f"form {form!r} needs a case solved in one partition (npart = {npart})")
    dev = device or torch.device("cuda:0")
    lay = pb.Layout.REF_GPU
    nm, nv = nsys * nrow * m * m, nsys * nrow * m
    a, b, c = (torch.empty(nm, dtype=torch.float64, device=dev) for _ in range(3))
    d = torch.empty(nv, dtype=torch.float64, device=dev)
    pb.generate(pb.GenKind.DIAG, 0, m, nsys, nrow, lay, a, b, c, d)
    a0, c0, d0 = a.clone(), c.clone(), d.clone()
    mode = f"{MPI_MODE}-np{npart}" if auto_npart else MPI_MODE
    if form != "plain":
        mode += f"-{form}"
    sizes = [pb.partition_1d(0, nrow - 1, npart, r) for r in range(npart)]
    rows = [last - first + 1 for first, last in sizes]
    head = f"btdma,noncyclic,{IMPLEMENTATION},{nranks},{n1},{n2},{n3},{m},{nsys},{min(rows)},{max(rows)}"
    timing, signature = [], None
    with pb.BtdmaGpuPlan().create(m, nsys, nrow, npart, lay) as plan:
        if form != "plain":
            plan.set_form(pb.Form.TWISTED if form == "twisted" else pb.Form.LAPACK_ORDER)
        for it in range(iterations):
            a.copy_(a0)
            c.copy_(c0)
            d.copy_(d0)
            torch.cuda.synchronize()
            t = pb.solve_noncyclic_profiled(plan, a, b, c, d, m, nsys, nrow)
            if it == 0:
                x = d.double()
                s, ss, linf = float(x.sum()), float((x * x).sum()), float(x.abs().max())
                samples = [float(x[pb.vector_index(0, z, 0, nsys, nrow)]) for z in (0, n3 // 2, n3 - 1)]
                signature = ",".join([head, mode] + [_fmt(v, 16) for v in
                                                        (s, math.sqrt(ss), linf, *samples)])
            fields = [t.total, t.total, t.local_compute, t.forward_exchange, t.reduced_compute,
                      t.backward_exchange, t.update_compute, t.computation(), t.communication()]
            timing.append(",".join([head, str(it), str(iterations), mode] + [_fmt(v, 12) for v in fields]))
    del a, b, c, d, a0, c0, d0
    torch.cuda.empty_cache()
    return timing, signature


def _append(path: str, header: str, lines) -> None:
    new = not os.path.exists(path) or os.path.getsize(path) == 0
    with open(path, "a") as f:
        if new:
            f.write(header + "\n")
        for ln in lines:
            f.write(ln + "\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--profile-out", required=True)
    ap.add_argument("--signature-out", required=True)
    ap.add_argument("--iterations", type=int, default=10)
    ap.add_argument("--case", action="append", default=[],
                    help="nranks,n1,n2,n3,m (repeatable; default: the reference Study's case matrix)")
    ap.add_argument("--npart", choices=["nranks", "auto"], default="nranks",
                    help="partitions per system: the case's rank count, or the library's choice")
    ap.add_argument("--form", choices=FORMS, default="plain",
                    help="single-partition solve form (cases in one partition only)")
    args = ap.parse_args(argv)
    cases = [tuple(int(v) for v in c.split(",")) for c in args.case] or STUDY_CASES
    for case in cases:
        if len(case) != 5:
            raise SystemExit(f"--case needs nranks,n1,n2,n3,m, got {case}")
        timing, signature = run_case(case, args.iterations, auto_npart=args.npart == "auto", form=args.form)
        _append(args.profile_out, PROFILE_HEADER, timing)
        _append(args.signature_out, SIGNATURE_HEADER, [signature])
        print(f"[study] case np={case[0]} {case[1]}x{case[2]}x{case[3]} m={case[4]} done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
