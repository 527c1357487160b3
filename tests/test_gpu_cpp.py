"""The reference's C++ example, ported to the C++ mirror header (tests/cpp), run on the
device against the Study's published signatures."""
import csv
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "bin", "ex_btdma_noncyclic")


def rows():
    with open(os.path.join(HERE, "golden", "study_signatures.csv")) as f:
        return [r for r in csv.DictReader(f) if r["nranks"] == "1"]


@pytest.mark.parametrize("row", rows(), ids=lambda r: f"m{r['m']}")
def test_cpp_example(cuda, row):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run python __graft_entry__.py)")
    r = subprocess.run([EXE, row["m"], row["solution_sum"], row["solution_l2"]],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
