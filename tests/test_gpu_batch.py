"""Mixed-size batches (persistent kernel), the autotuner and graph capture, on the device."""
import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout

pytestmark = pytest.mark.gpu
TOL = 1e-12


def make_batch(cuda, m, n, kind=pb.GenKind.DDPERM, seed=5):
    import torch
    bt = pb.BtdmaBatch(m, n)
    a, b, c = (torch.empty(bt.mat_elems, dtype=torch.float64, device=cuda) for _ in range(3))
    d = torch.empty(bt.vec_elems, dtype=torch.float64, device=cuda)
    bt.generate(kind, seed, a, b, c, d)
    torch.cuda.synchronize()
    return bt, a, b, c, d


def oracle_system(oracle, bt, kind, seed, s):
    m, n = int(bt.m[s]), int(bt.n[s])
    A, B, C, d = pb.generate_host(kind, seed, m, bt.nsys, n, s, 1)
    x, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
    assert info == 0
    return (A, B, C, d), x[:, 0]


def rand_sizes(rng, nsys, mset=range(1, 11), nmax=80):
    m = rng.choice(np.array(list(mset)), nsys)
    n = rng.integers(2, nmax, nsys)
    return m, n


def test_batch_generator_matches_host(cuda):
    rng = np.random.default_rng(0)
    m, n = rand_sizes(rng, 150)
    bt, a, b, c, d = make_batch(cuda, m, n, pb.GenKind.EULER, 8)
    ha, hb, hc, hd = (t.cpu().numpy() for t in (a, b, c, d))
    for s in range(0, 150, 7):
        A, B, C, dd = pb.generate_host(pb.GenKind.EULER, 8, int(m[s]), 150, int(n[s]), s, 1)
        np.testing.assert_array_equal(bt.extract(ha, s, vector=False), A[:, 0])
        np.testing.assert_array_equal(bt.extract(hb, s, vector=False), B[:, 0])
        np.testing.assert_array_equal(bt.extract(hc, s, vector=False), C[:, 0])
        np.testing.assert_array_equal(bt.extract(hd, s), dd[:, 0])


@pytest.mark.parametrize("placement,grid", [(0, 64), (1, 64), (2, 64), (3, 64), (4, 64), (5, 64),
                                            (5, 100000), (6, 64)])
def test_batch_matches_oracle(cuda, oracle, placement, grid):
    """Every placement against the oracle; placement 5 twists the items above total/grid, so
    grid 100000 makes every cooperative item twisted (n from 2: meeting rows 0 and 1)."""
    rng = np.random.default_rng(placement + 1)
    m, n = rand_sizes(rng, 300)
    bt, a, b, c, d = make_batch(cuda, m, n, pb.GenKind.DDPERM, 5)
    bt.set_params(placement, grid)
    bt.solve(a, b, c, d)
    assert not bt.check()
    hd = d.cpu().numpy()
    worst = 0.0
    for s in range(bt.nsys):
        _, xo = oracle_system(oracle, bt, pb.GenKind.DDPERM, 5, s)
        x = bt.extract(hd, s)
        worst = max(worst, np.abs(x - xo).max() / np.abs(xo).max())
    print(f"placement {placement} grid {grid}: {bt.npacks} packs, max rel err {worst:.2e}")
    assert worst <= TOL


@pytest.mark.parametrize("placement,grid", [(0, 64), (1, 64), (4, 64), (5, 100000), (6, 64)])
def test_batch_zero_pivot_flag(cuda, placement, grid):
    """A system with an all-zero block row is reported by btd_batch_check in every placement
    (twisted: the row falls in the bottom half), a clean batch is not, and the report clears."""
    import torch
    rng = np.random.default_rng(3)
    m, n = rand_sizes(rng, 120, mset=(2, 5, 8, 10))
    bt, a, b, c0, d0 = make_batch(cuda, m, n)
    bt.set_params(placement, grid)
    c, d = c0.clone(), d0.clone()
    bt.solve(a, b, c, d)
    assert not bt.check()
    s = int(np.argmax(n))
    mm, row = int(m[s]), 2 * int(n[s]) // 3
    mo, _, w = bt.locate(s)
    idx = torch.tensor([mo + (row * mm * mm + e) * w for e in range(mm * mm)], device=cuda)
    for t in (a, b, c0):
        t[idx] = 0.0
    c, d = c0.clone(), d0.clone()
    bt.solve(a, b, c, d)
    assert bt.check()
    assert not bt.check()


def test_graph_replays_current_item_list(cuda, oracle):
    """A graph captured before set_params changes the work items replays the current item
    table with its own count (both live on the device), so every system is still solved."""
    import torch
    rng = np.random.default_rng(21)
    m, n = rand_sizes(rng, 200, mset=(2, 5, 10), nmax=120)
    bt, a, b, c0, d0 = make_batch(cuda, m, n)
    bt.set_params(5, 64)
    c, d = c0.clone(), d0.clone()
    s1 = torch.cuda.Stream()
    g = pb.Graph()
    torch.cuda.synchronize()
    g.begin(s1)
    bt.solve(a, b, c, d, stream=s1)
    g.end(s1)
    bt.set_params(5, 100000)  # every cooperative item twisted: a longer item list
    c.copy_(c0)
    d.copy_(d0)
    torch.cuda.synchronize()
    g.launch(s1)
    s1.synchronize()
    g.destroy()
    assert not bt.check()
    hd = d.cpu().numpy()
    worst = 0.0
    for s in range(bt.nsys):
        _, xo = oracle_system(oracle, bt, pb.GenKind.DDPERM, 5, s)
        worst = max(worst, np.abs(bt.extract(hd, s) - xo).max() / np.abs(xo).max())
    assert worst <= TOL


def test_tuner_and_graph_replay(cuda, oracle):
    """Autotune the batch, then capture batch solve + a two-stage long-system solve (on a
    forked stream) into one graph; every replay must equal the eager result bit for bit."""
    import torch
    from pascal_btdma_amd.tuner import tune_batch, tune_partitions
    rng = np.random.default_rng(11)
    m, n = rand_sizes(rng, 400, mset=(2, 3, 5, 8, 10), nmax=200)
    bt, a, b, c0, d0 = make_batch(cuda, m, n)
    res = tune_batch(bt, a, b, c0, d0, reps=2)
    print("batch tuning:", res)
    # long systems: m = 5, n = 960, 64 systems
    lm, lnsys, ln = 5, 64, 960
    lay = pb.Layout.TILED
    la, lb, lc0 = (torch.empty(pb.layout_elems(lay, lnsys, ln, lm * lm), dtype=torch.float64, device=cuda)
                   for _ in range(3))
    ld0 = torch.empty(pb.layout_elems(lay, lnsys, ln, lm), dtype=torch.float64, device=cuda)
    pb.generate(pb.GenKind.DDPERM, 21, lm, lnsys, ln, lay, la, lb, lc0, ld0)
    resp = tune_partitions(lm, lnsys, ln, la, lb, lc0, ld0, lay, candidates=(8, 16, 32), reps=2)
    print("partition tuning:", resp)
    plan = pb.BtdmaGpuPlan().create(lm, lnsys, ln, resp["npart"], lay)
    la0 = la.clone()

    # eager reference
    c, d, lc, ld, la_w = c0.clone(), d0.clone(), lc0.clone(), ld0.clone(), la0.clone()
    bt.solve(a, b, c, d)
    pb.solve_noncyclic(plan, la_w, lb, lc, ld, lm, lnsys, ln)
    torch.cuda.synchronize()
    ref_d, ref_ld = d.clone(), ld.clone()

    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    g = pb.Graph()
    fork, join = torch.cuda.Event(), torch.cuda.Event()
    torch.cuda.synchronize()
    g.begin(s1)
    fork.record(s1)
    s2.wait_event(fork)
    bt.solve(a, b, c, d, stream=s1)
    pb.solve_noncyclic(plan, la_w, lb, lc, ld, lm, lnsys, ln, stream=s2)
    join.record(s2)
    s1.wait_event(join)
    g.end(s1)
    for rep in range(3):
        c.copy_(c0)
        d.copy_(d0)
        lc.copy_(lc0)
        ld.copy_(ld0)
        la_w.copy_(la0)
        torch.cuda.synchronize()
        g.launch(s1)
        s1.synchronize()
        assert torch.equal(d, ref_d), f"replay {rep}: batch differs"
        assert torch.equal(ld, ref_ld), f"replay {rep}: long systems differ"
        _ = (d * 2).sum().item()  # eager work between replays
    g.destroy()
    # the eager result itself matches the oracle on a sample of systems
    hd = ref_d.cpu().numpy()
    for s in range(0, bt.nsys, 37):
        _, xo = oracle_system(oracle, bt, pb.GenKind.DDPERM, 5, s)
        assert np.abs(bt.extract(hd, s) - xo).max() / np.abs(xo).max() <= TOL
    A, B, C, dd = pb.generate_host(pb.GenKind.DDPERM, 21, lm, lnsys, ln, 0, 4)
    xo, info = oracle.solve_two_stage(cpu_layout(A), cpu_layout(B), cpu_layout(C), dd, resp["npart"])
    flat = ref_ld.cpu().numpy()[: ln * lm * 64].reshape(ln, lm, 64)
    x = np.transpose(flat, (0, 2, 1))[:, :4]
    assert (np.abs(x - xo).max(axis=(0, 2)) / np.abs(xo).max(axis=(0, 2))).max() <= TOL
    plan.destroy()


@pytest.mark.parametrize("m,nsys,n,placement,form", [
    (7, 600, 40, 6, pb.Form.TWISTED), (8, 600, 41, 6, pb.Form.TWISTED), (10, 300, 40, 6, pb.Form.TWISTED),
    (9, 130, 2, 6, pb.Form.TWISTED), (7, 129, 3, 6, pb.Form.TWISTED), (10, 97, 4, 6, pb.Form.TWISTED),
    (8, 8200, 24, 4, pb.Form.PLAIN), (9, 4100, 30, 4, pb.Form.PLAIN), (10, 3100, 25, 4, pb.Form.PLAIN)])
def test_column_form_equals_row_form(cuda, m, nsys, n, placement, form):
    """From m = 7 the single-partition cooperative and twisted solves run the column form
    (lane j owns column j of every block); the mixed-size batch runs the row form (lane i owns
    row i). Both perform the same operations on the same operands, so on the same systems
    (equal m and n, so both generators give the same values) the plain plan's x equals the
    batch's bit for bit: twisted (placement 6, every item twisted) and cooperative
    (placement 4; PLAIN picks the cooperative kernel at these batch sizes for m >= 8)."""
    import torch
    bt, a, b, c, d = make_batch(cuda, np.full(nsys, m), np.full(nsys, n), pb.GenKind.DDPERM, 9)
    bt.set_params(placement, 512)
    bt.solve(a, b, c, d)
    lay = pb.Layout.REF_GPU
    t = [torch.empty(pb.layout_elems(lay, nsys, n, m * m if k < 3 else m), dtype=torch.float64, device=cuda)
         for k in range(4)]
    pb.generate(pb.GenKind.DDPERM, 9, m, nsys, n, lay, *t)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, lay) as plan:
        plan.set_form(form)
        pb.solve_noncyclic(plan, *t, m, nsys, n)
        torch.cuda.synchronize()
        assert not plan.check()
    assert not bt.check()
    hd = d.cpu().numpy()
    x = np.transpose(t[3].cpu().numpy().reshape(m, n, nsys), (1, 2, 0))  # (n, nsys, m)
    for s in range(nsys):
        np.testing.assert_array_equal(x[:, s, :], bt.extract(hd, s), err_msg=f"system {s}")
