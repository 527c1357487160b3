"""Extracts the reference Study's solution signatures into tests/golden/study_signatures.csv.

Source (data file held by the reference): Study/result/btdma_solution_signature_260702_193138.csv,
written by Study/example_cuda_cxx_btdma_profile.cu:76-160 after the first solve of the
diagonal workload A = 0.25 I, B = 2 I, C = 0.25 I, D = 1 (Study/README.md:40-50), in the
reference GPU layout, on 1-8 H200 ranks (nranks = row partitions). Run in a container
that has /root/reference; the output is committed so tests never read the reference.
"""
import csv
import os
import sys

SRC = "/root/reference/Study/result/btdma_solution_signature_260702_193138.csv"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "study_signatures.csv")
KEEP = ["nranks", "n1", "n2", "n3", "m", "nsys", "mpi_mode", "solution_sum", "solution_l2",
        "solution_linf", "sample_z0", "sample_zmid", "sample_zlast"]


def main() -> int:
    with open(SRC, newline="") as f:
        rows = list(csv.DictReader(f))
    seen = set()
    out = []
    for r in rows:
        key = (r["nranks"], r["n1"], r["n2"], r["n3"], r["m"])
        if r["mpi_mode"] != "device" or key in seen:
            continue
        seen.add(key)
        out.append({k: r[k] for k in KEEP})
    with open(DST, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=KEEP)
        w.writeheader()
        w.writerows(out)
    print(f"wrote {len(out)} rows to {DST}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
