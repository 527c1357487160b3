"""The one-lane-per-system solve at the batch shapes it serves, against the oracle.

Single-partition solves run one lane per system with the blocks in VGPRs (k_plain) at m = 1
at every size, at m = 2-4 from 131072 systems and at m = 5 from 98304 (the headline batch:
m = 5, n = 512, 131072 Euler-like systems); smaller batches run two lanes per system from both
ends or 2m lanes twisted, and m = 6, 7 two lanes per system at every size from 16384; the
BTD_FORM_LAPACK_ORDER form runs k_plain at any size and m. Inputs are generated on the device
(bit-identical to the host generator, pinned in test_gpu_parity.py), the whole batch is solved,
and a sample of systems is compared with the oracle's btdma_many restatement
(mod_btdma_cpu.f90:15-78): the first tiles, a middle tile and the partial last tile. The
kernel computes every block operation in the oracle's order and rounding, so the bar is
equality bit for bit (which implies the 1e-12 max per-system normwise relative error).
"""
import numpy as np
import pytest

import pascal_btdma_amd as pb
from helpers import cpu_layout, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12
# the batch size from which the default form runs one lane per system (None: never at m = 6, 7)
LANE_FROM = {1: 1, 2: 131072, 3: 131072, 4: 131072, 5: 98304, 6: None, 7: None}


def sample_ranges(nsys, first_tiles):
    mid = 64 * ((nsys // 64) // 2)
    last = 64 * (nsys // 64)
    return [(0, 64 * first_tiles), (mid, 64), (last, nsys - last)]


def gather_x(td, layout, nsys, n, m, systems):
    """x[row, k, i] of the given systems from the solved flat vector array (on the device)."""
    import torch
    dev = td.device
    s = torch.as_tensor(systems, dtype=torch.int64, device=dev)[None, :, None]
    row = torch.arange(n, dtype=torch.int64, device=dev)[:, None, None]
    e = torch.arange(m, dtype=torch.int64, device=dev)[None, None, :]
    if layout == pb.Layout.TILED:
        idx = (((s // 64) * n + row) * m + e) * 64 + s % 64
    else:
        idx = s + nsys * (row + n * e)
    return td[idx].cpu().numpy()


def solve_and_compare(cuda, oracle, kind, seed, m, nsys, n, layout, first_tiles, form=None):
    import torch
    nm = pb.layout_elems(layout, nsys, n, m * m)
    nv = pb.layout_elems(layout, nsys, n, m)
    t = [torch.empty(nm, dtype=torch.float64, device=cuda) for _ in range(3)]
    td = torch.empty(nv, dtype=torch.float64, device=cuda)
    pb.generate(kind, seed, m, nsys, n, layout, *t, td)
    with pb.BtdmaGpuPlan().create(m, nsys, n, 1, layout) as plan:
        assert plan.form == pb.Form.PLAIN
        if form is not None:
            plan.set_form(form)
        pb.solve_noncyclic(plan, t[0], t[1], t[2], td, m, nsys, n)
        torch.cuda.synchronize()
        assert not plan.check()
    worst = 0.0
    count = 0
    for s0, cnt in sample_ranges(nsys, first_tiles):
        A, B, C, d = pb.generate_host(kind, seed, m, nsys, n, s0, cnt)
        xo, info = oracle.solve_plain(cpu_layout(A), cpu_layout(B), cpu_layout(C), d)
        assert info == 0
        x = gather_x(td, layout, nsys, n, m, list(range(s0, s0 + cnt)))
        worst = max(worst, rel_err(x, xo).max())
        np.testing.assert_array_equal(x, xo)
        count += cnt
    del t, td
    torch.cuda.empty_cache()
    return worst, count


@pytest.mark.parametrize("layout", [pb.Layout.TILED, pb.Layout.REF_GPU])
def test_headline_shape_euler_m5(cuda, oracle, layout):
    """m = 5, n = 512, 98309 Euler-like systems (past the one-lane threshold, a partial last
    tile of 5 systems); 261 systems against the oracle."""
    nsys = LANE_FROM[5] + 5
    worst, count = solve_and_compare(cuda, oracle, pb.GenKind.EULER, 1, 5, nsys, 512, layout, 3)
    print(f"euler m=5 n=512 nsys={nsys} {layout.name}: {count} systems, max rel err {worst:.3e}")
    assert count >= 256
    assert worst <= TOL


@pytest.mark.parametrize("m", list(range(1, 8)))
@pytest.mark.parametrize("layout", [pb.Layout.TILED, pb.Layout.REF_GPU])
def test_one_lane_every_m_permuted_rows(cuda, oracle, m, layout):
    """m = 1..7 one lane per system on row-permuted diagonally dominant blocks (pivoting
    exercised in every block LU), n = 64: 5 systems past the default form's one-lane threshold
    (m <= 5), and 32773 systems in the BTD_FORM_LAPACK_ORDER form (every m)."""
    cases = [(32768 + 5, pb.Form.LAPACK_ORDER)]
    if LANE_FROM[m] is not None:
        cases.append((max(LANE_FROM[m], 64) + 5, None))
    for nsys, form in cases:
        worst, count = solve_and_compare(cuda, oracle, pb.GenKind.DDPERM, 300 + m, m, nsys, 64, layout, 1, form)
        print(f"ddperm m={m} n=64 nsys={nsys} form={form} {layout.name}: {count} systems, max rel err {worst:.3e}")
        assert worst <= TOL
