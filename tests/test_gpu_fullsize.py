"""GPU tests at the full size of the single-TOD workload (2048 detectors x 262144 samples of the
raster scan on the sparse 3x4096x8192 map), through properties that hold at any size: the
planned operators are adjoint (FP64 inner products), map2tod is linear, the planned and plan-free
operators agree, the plan is sorted by cell and has the preplan's size, and the first detectors
agree with the FP64 oracle."""

import numpy as np
import pytest
import torch

from gpu_mm_amd import synth

pytestmark = pytest.mark.gpu

F32 = np.float32
NDET, NT = 2048, 262144


def l1rel(a, b):
    a = a.double().reshape(-1)
    b = b.double().reshape(-1)
    return float((a - b).abs().sum() / (a.abs().sum() + b.abs().sum()))


@pytest.fixture(scope="module")
def full(gmm, dev):
    c = synth.CFG1
    print(f"\n[full size] {NDET} x {NT} raster on 3x{c['ny']}x{c['nx']}", flush=True)
    xp, _ = synth.raster_xpointing_torch(NDET, NT, seed=1, device=dev, **c["scan"])
    pre = gmm.PointingPrePlan(xp, c["ny"], c["nx"], periodic_xcoord=False)
    plan = gmm.PointingPlan(pre, xp)
    cells = np.unique(plan.get_plan_mt() & np.uint64(0xFFFFF))
    cy, cx = (cells >> np.uint64(10)).astype(np.int64), (cells & np.uint64(1023)).astype(np.int64)
    lp = gmm.LocalPixelization.make_cells(c["ny"], c["nx"], list(zip(cy.tolist(), cx.tolist())),
                                          periodic_xcoord=False)
    gen = torch.Generator(device=dev)
    gen.manual_seed(9)
    m1 = torch.randn(3 * lp.npix, generator=gen, device=dev)
    m2 = torch.randn(3 * lp.npix, generator=gen, device=dev)
    d = torch.randn((NDET, NT), generator=gen, device=dev)
    return dict(xp=xp, pre=pre, plan=plan, lp=lp, m1=m1, m2=m2, d=d, ny=c["ny"], nx=c["nx"])


def test_plan_structure(full):
    mt = full["plan"].get_plan_mt()
    assert len(mt) == full["pre"].plan_nmt
    keys = mt & np.uint64(0xFFFFF)
    assert np.all(keys[1:] >= keys[:-1])
    # Every line of the TOD appears at least once.
    lines = np.unique((mt >> np.uint64(20)) & np.uint64((1 << 26) - 1))
    assert len(lines) == NDET * NT // 32


def test_adjoint_and_linearity(full, gmm, dev):
    f = full
    lp, xp, plan = f["lp"], f["xp"], f["plan"]
    t1 = torch.empty((NDET, NT), dtype=torch.float32, device=dev)
    t2 = torch.empty_like(t1)
    gmm.map2tod(t1, gmm.LocalMap(lp, f["m1"]), xp, plan, debug=True)
    gmm.map2tod(t2, gmm.LocalMap(lp, f["m2"]), xp, plan, debug=True)
    t12 = torch.empty_like(t1)
    gmm.map2tod(t12, gmm.LocalMap(lp, f["m1"] + 2.0 * f["m2"]), xp, plan, debug=True)
    assert l1rel(t12, t1 + 2.0 * t2) < 1e-6
    out = gmm.LocalMap(lp, torch.zeros_like(f["m1"]))
    gmm.tod2map(out, f["d"], xp, plan, debug=True)
    lhs = float(torch.sum(t1.double() * f["d"].double()))
    rhs = float(torch.sum(f["m1"].double() * out.arr.double()))
    scale = float(torch.sum((t1.double() * f["d"].double()).abs()))
    print(f"  <A m, d> = {lhs:.6e}, <m, A^T d> = {rhs:.6e}", flush=True)
    assert abs(lhs - rhs) < 1e-6 * scale


def test_planned_equals_planfree(full, gmm, dev):
    f = full
    lp, xp, plan = f["lp"], f["xp"], f["plan"]
    mg = gmm.LocalMap(lp, f["m1"])
    tp = torch.empty((NDET, NT), dtype=torch.float32, device=dev)
    tu = torch.empty_like(tp)
    gmm.map2tod(tp, mg, xp, plan, debug=True)
    gmm.map2tod(tu, mg, xp, None)
    assert l1rel(tp, tu) < 1e-6
    # tod2map: the plan-free kernel adds 12 global float atomics per sample, so it is compared
    # on the first 128 detectors (with their own plan); the full-size tod2map is covered by the
    # adjoint test.
    k = 128
    xk = xp[:, :k].contiguous()
    pk = gmm.PointingPlan(gmm.PointingPrePlan(xk, f["ny"], f["nx"], periodic_xcoord=False), xk)
    dk = f["d"][:k].contiguous()
    op = gmm.LocalMap(lp, torch.zeros_like(f["m1"]))
    ou = gmm.LocalMap(lp, torch.zeros_like(f["m1"]))
    gmm.tod2map(op, dk, xk, pk, debug=True)
    gmm.tod2map(ou, dk, xk, None)
    assert l1rel(op.arr, ou.arr) < 1e-6


def test_first_detectors_vs_oracle(full, gmm, orc, dev):
    f = full
    lp, xp, plan = f["lp"], f["xp"], f["plan"]
    k = 8
    t = torch.empty((NDET, NT), dtype=torch.float32, device=dev)
    gmm.map2tod(t, gmm.LocalMap(lp, f["m1"]), xp, plan, debug=True)
    pix = orc.Pix(f["ny"], f["nx"], lp.cell_offsets_cpu, 64, 4096, False)
    xk = xp[:, :k].cpu().numpy()
    ref, err = orc.map2tod(f["m1"].cpu().numpy().astype(np.float64), xk, pix, f64=True)
    assert err == 0
    got = t[:k].cpu().numpy().astype(np.float64).reshape(-1)
    assert np.sum(np.abs(got - ref)) / np.sum(np.abs(ref)) < 1e-5


@pytest.mark.parametrize("d0", [0, 1000, NDET - 32])
def test_plan_bit_exact_on_detector_ranges(full, orc, d0):
    """Entries of 32 detectors' lines, taken from the full-size plan (1024 planner blocks of 16384
    lines, multi-pass radix sort over all 2e7 entries), against the oracle's plan of those 32
    detectors alone: a line's cells and flags depend on its own samples only, and the canonical
    (cell, line) order restricted to these lines is their own canonical order."""
    k = 32
    lpd = NT // 32                                   # lines per detector
    mt = full["plan"].get_plan_mt()
    icl = (mt >> np.uint64(20)) & np.uint64((1 << 26) - 1)
    lo, hi = np.uint64(d0 * lpd), np.uint64((d0 + k) * lpd)
    sub = mt[(icl >= lo) & (icl < hi)]
    xk = full["xp"][:, d0:d0 + k].cpu().numpy()
    ref, err = orc.plan_build(xk, full["ny"], full["nx"], False)
    assert err == 0
    ref = ref + (np.uint64(d0 * lpd) << np.uint64(20))  # the subset's line numbers start at d0's
    print(f"  detectors {d0}..{d0 + k - 1}: {len(sub)} entries", flush=True)
    assert np.array_equal(sub, ref)
    # nmt_cumsum of the planner blocks these detectors fill (2 detectors per block of 16384 lines)
    cum = full["pre"].get_nmt_cumsum()
    blk = (d0 * lpd) // 16384
    assert d0 * lpd % 16384 == 0 and full["pre"].ncl_per_threadblock == 16384
    per_block = np.bincount(((ref >> np.uint64(20)) & np.uint64((1 << 26) - 1)).astype(np.int64) // 16384,
                            minlength=blk + k // 2)[blk:blk + k // 2]
    before = 0 if blk == 0 else int(cum[blk - 1])
    assert np.array_equal(cum[blk:blk + k // 2].astype(np.int64), before + np.cumsum(per_block))
