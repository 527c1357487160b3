"""The Study-compatible driver (pascal_btdma_amd/study.py): its CSV schemas are the reference
Study's (Study/README.md, "Study Outputs"), its case matrix covers every case of the
reference's published signatures; on the device, its signature rows reproduce those
signatures and its timing rows are well formed."""
import csv
import io
import os

import pytest

from pascal_btdma_amd import study

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "study_signatures.csv")


def test_schemas_are_the_reference_study_schemas():
    assert study.PROFILE_HEADER == (
        "solver,variant,implementation,nranks,n1,n2,n3,m,nsys,nrow_min,nrow_max,iter,iterations,"
        "mpi_mode,total_s_max,total_s_avg,local_compute_s_max,forward_exchange_s_max,"
        "reduced_compute_s_max,backward_exchange_s_max,update_compute_s_max,compute_s_max,"
        "communication_s_max")
    assert study.SIGNATURE_HEADER == (
        "solver,variant,implementation,nranks,n1,n2,n3,m,nsys,nrow_min,nrow_max,mpi_mode,"
        "solution_sum,solution_l2,solution_linf,sample_z0,sample_zmid,sample_zlast")


def test_case_matrix_covers_the_published_signatures():
    with open(GOLDEN) as f:
        golden = {(int(r["nranks"]), int(r["n1"]), int(r["n2"]), int(r["n3"]), int(r["m"]))
                  for r in csv.DictReader(f)}
    assert golden <= set(study.STUDY_CASES)
    assert len(set(study.STUDY_CASES)) == len(study.STUDY_CASES) == 28


def test_number_format_is_iostream_setprecision():
    assert study._fmt(0.4508066615170332, 16) == "0.4508066615170332"
    assert study._fmt(2.5789e-05, 12) == "2.5789e-05"
    assert study._fmt(0.0, 12) == "0"


@pytest.mark.gpu
@pytest.mark.parametrize("auto", [False, True])
@pytest.mark.parametrize("case", [(1, 32, 32, 128, 5), (2, 64, 64, 1024, 2), (8, 32, 32, 2048, 8)])
def test_study_rows_on_device(cuda, case, auto, tmp_path):
    timing, signature = study.run_case(case, iterations=3, auto_npart=auto)
    prof = tmp_path / "p.csv"
    sig = tmp_path / "s.csv"
    study._append(str(prof), study.PROFILE_HEADER, timing)
    study._append(str(sig), study.SIGNATURE_HEADER, [signature])
    rows = list(csv.DictReader(io.StringIO(prof.read_text())))
    assert [int(r["iter"]) for r in rows] == [0, 1, 2]
    P = case[0]
    if auto:
        import pascal_btdma_amd as pb
        P = pb._capi.lib().btd_auto_npart(case[4], case[1] * case[2], case[3])
    for r in rows:
        assert r["implementation"] == "hip-mi355x"
        assert r["mpi_mode"] == (f"in-hbm-np{P}" if auto else "in-hbm")
        total, comp = float(r["total_s_max"]), float(r["compute_s_max"])
        parts = sum(float(r[k]) for k in ("local_compute_s_max", "reduced_compute_s_max",
                                          "update_compute_s_max"))
        assert float(r["communication_s_max"]) == 0.0
        assert abs(comp - parts) <= 1e-9 * max(1.0, parts) + 1e-11
        # the phases share their boundary events with the total; each event-pair time is read
        # back separately, which has put their sum a few ns above the total
        assert 0 < comp <= total * (1 + 1e-6) + 5e-8
    (srow,) = list(csv.DictReader(io.StringIO(sig.read_text())))
    with open(GOLDEN) as f:
        gold = [g for g in csv.DictReader(f)
                if (int(g["nranks"]), int(g["n1"]), int(g["n2"]), int(g["n3"]), int(g["m"])) == case]
    assert gold, case
    g = gold[0]
    assert abs(float(srow["solution_sum"]) / float(g["solution_sum"]) - 1) < 1e-9
    assert abs(float(srow["solution_l2"]) / float(g["solution_l2"]) - 1) < 1e-9
    for k in ("solution_linf", "sample_z0", "sample_zmid", "sample_zlast"):
        assert abs(float(srow[k]) - float(g[k])) < 1e-14
    nrows = [int(srow["nrow_min"]), int(srow["nrow_max"])]
    assert nrows == [case[3] // P] * 2


def test_form_needs_one_partition():
    with pytest.raises(ValueError, match="one partition"):
        study.run_case((2, 8, 8, 64, 3), iterations=1, form="twisted")
    with pytest.raises(ValueError, match="form must be"):
        study.run_case((1, 8, 8, 64, 3), iterations=1, form="coop")


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["twisted", "lapack_order"])
def test_study_forms_on_device(cuda, form):
    """The single-partition forms give the published signature of the np = 1 cases (m = 5: the
    default form is the twisted solve at 1024 systems, so lapack_order runs another kernel)."""
    case = (1, 32, 32, 128, 5)
    timing, signature = study.run_case(case, iterations=2, form=form)
    assert all(f",in-hbm-{form}," in t for t in timing)
    srow = dict(zip(study.SIGNATURE_HEADER.split(","), signature.split(",")))
    assert srow["mpi_mode"] == f"in-hbm-{form}"
    with open(GOLDEN) as f:
        (g,) = [g for g in csv.DictReader(f)
                if (int(g["nranks"]), int(g["n1"]), int(g["n2"]), int(g["n3"]), int(g["m"])) == case]
    assert abs(float(srow["solution_sum"]) / float(g["solution_sum"]) - 1) < 1e-9
    assert abs(float(srow["solution_l2"]) / float(g["solution_l2"]) - 1) < 1e-9
    for k in ("solution_linf", "sample_z0", "sample_zmid", "sample_zlast"):
        assert abs(float(srow[k]) - float(g[k])) < 1e-14
