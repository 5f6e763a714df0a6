"""GPU parity tests: the HIP planner and operators against the oracle on the same seeded inputs.

Bars: plans bit-exact (all 48 bits, canonical order) and nmt_cumsum equal; FP32 operators
within 1e-6 L1-relative of the FP32 oracle (the reference's own metric and threshold,
gpu_mm/tests.py:220-224, 277, 326) and within 1e-5 of the FP64 oracle; adjoint identity
<A m, d> = <m, A^T d> to 1e-6 relative.
"""

import numpy as np
import pytest
import torch

from gpu_mm_amd import synth

pytestmark = pytest.mark.gpu

F32 = np.float32
# Bar against the FP64 oracle (FP32 kernels; the FP32 rounding of periodic x +- nx alone reaches
# ~1e-6 against an exact evaluation).
F64_BAR = 1e-5


def l1rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.sum(np.abs(a - b)) / (np.sum(np.abs(a)) + np.sum(np.abs(b))))


def to_gpu(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def to_cpu(t):
    return t.detach().to("cpu").numpy()


def make_instance(kind, seed):
    """Returns (xpointing (3,ndet,nt) float32, ny, nx, periodic)."""
    rng = np.random.default_rng(seed)
    if kind == "cfg0":
        d = synth.make_cfg0(seed=0)
        return d["xpointing"], d["ny"], d["nx"], False
    if kind == "fast_periodic":
        # Fast scan with wrap-around: many multi-cell lines, cells on both ends of the map.
        ndet, nt, ny, nx = 8, 32 * 512, 300, 700
        t = np.arange(nt)
        xp = np.empty((3, ndet, nt), F32)
        for d in range(ndet):
            xp[0, d] = 150 + 140 * np.sin(t / 700.0 + d) + rng.uniform(-1, 1)
            xp[1, d] = (t * (1.7 + 0.3 * d)) % (2 * nx) - nx / 2
            xp[2, d] = rng.uniform(0, np.pi) + 0.01 * t
        return xp, ny, nx, True
    if kind == "mixed":
        # Detectors at different scan speeds and angle rates, so that the waves of one cell mix
        # lanes whose 4-sample groups fit one 3x3 tod2map window (below ~0.33 px/sample) with
        # lanes whose groups do not (deposited sample by sample, up to 2.5 px/sample), with angle
        # steps of the first-order, second-order and table regimes.
        ndet, nt, ny, nx = 12, 32 * 256, 256, 512
        t = np.arange(nt, dtype=np.float64)
        speeds = [0.05, 0.2, 0.3, 0.4, 0.55, 0.7, 0.9, 1.3, 0.12, 0.6, 2.5, 0.33]
        xp = np.empty((3, ndet, nt), F32)
        for d in range(ndet):
            v, phi = speeds[d], rng.uniform(0, 2 * np.pi)
            u = (v * t) % 360.0
            u = np.where(u < 180.0, u, 360.0 - u) - 90.0         # back and forth over 180 px
            xp[0, d] = 128 + 0.5 * u * np.sin(phi) + rng.uniform(-2, 2)
            xp[1, d] = 256 + u * np.cos(phi) + rng.uniform(-2, 2)
            rate = (1e-6, 1e-3, 0.0)[d % 3]
            xp[2, d] = rng.uniform(0, np.pi) + rate * t if d % 3 < 2 else rng.uniform(-10, 10, nt)
        return xp, ny, nx, False
    if kind == "toy":
        from gpu_mm_amd import ToyPointing
        tp = ToyPointing.make_random(1 << 20, noisy=False, rng=rng)
        xp = tp.xpointing_cpu.reshape(3, -1, tp.xpointing_cpu.shape[-1])
        return xp, tp.nypix_global, tp.nxpix_global, False
    raise ValueError(kind)


INSTANCES = [("cfg0", 0), ("fast_periodic", 1), ("toy", 2), ("toy", 3), ("mixed", 4)]


@pytest.fixture(scope="module", params=INSTANCES, ids=lambda p: f"{p[0]}-{p[1]}")
def inst(request, gmm, dev):
    kind, seed = request.param
    xp, ny, nx, periodic = make_instance(kind, seed)
    print(f"\n[instance {kind} seed={seed}] xpointing {xp.shape} ny={ny} nx={nx} periodic={periodic}", flush=True)
    xg = to_gpu(xp, dev)
    pre = gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=periodic)
    plan = gmm.PointingPlan(pre, xg)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=periodic)
    return dict(xp=xp, xg=xg, ny=ny, nx=nx, periodic=periodic, pre=pre, plan=plan, lp=lp, seed=seed)


def opix(orc, inst):
    return orc.Pix.rectangle(inst["ny"], inst["nx"], inst["periodic"])


def test_preplan_matches_oracle(inst, orc):
    cum, nmt, err = orc.plan_count(inst["xp"], inst["ny"], inst["nx"], inst["periodic"])
    assert err == 0
    pre = inst["pre"]
    assert pre.plan_nmt == nmt
    assert np.array_equal(pre.get_nmt_cumsum(), cum)
    per, nb = orc.preplan_params(pre.nsamp)
    assert (pre.ncl_per_threadblock, pre.planner_nblocks) == (per, nb)
    assert (pre.nmt_per_threadblock, pre.pointing_nblocks) == orc.pointing_params(nmt)
    assert pre.overhead == pytest.approx(32 * nmt / pre.nsamp - 1)


def test_plan_bit_exact(inst, orc):
    mt = inst["plan"].get_plan_mt()
    ref, err = orc.plan_build(inst["xp"], inst["ny"], inst["nx"], inst["periodic"])
    assert err == 0
    assert mt.dtype == np.uint64 and mt.shape == ref.shape
    assert np.array_equal(mt, ref)


def test_plan_matches_plan_tester(inst, gmm):
    """The reference's own criterion (gpu_mm/tests.py:227-254)."""
    tester = gmm.PointingPlanTester(inst["pre"], inst["xg"])
    assert np.array_equal(inst["pre"].get_nmt_cumsum(), tester.nmt_cumsum)
    mt = inst["plan"].get_plan_mt()
    low20 = mt & np.uint64((1 << 20) - 1)
    assert np.all(low20[1:] >= low20[:-1])
    assert np.array_equal(np.sort(mt & np.uint64((1 << 46) - 1)), tester.sorted_mt)


def test_plan_iterator_on_plan(inst, gmm):
    mt = inst["plan"].get_plan_mt()
    for nper, waves in ((32, 4), (1000, 8), (4096, 16)):
        gmm.test_plan_iterator(mt, nper, waves)


@pytest.mark.parametrize("accum", [False, True])
def test_map2tod(inst, gmm, orc, dev, accum):
    rng = np.random.default_rng(100 + inst["seed"] + 7 * accum)
    lp, plan = inst["lp"], inst["plan"]
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    tod0 = rng.standard_normal(inst["xp"].shape[1:]).astype(F32)
    pix = opix(orc, inst)
    ref32, err = orc.map2tod(m, inst["xp"], pix, tod=tod0, accum=accum)
    assert err == 0
    ref64, _ = orc.map2tod(m.astype(np.float64), inst["xp"], pix, tod=tod0, accum=accum, f64=True)
    mg = gmm.LocalMap(lp, to_gpu(m, dev))
    for p in (plan, None):
        t = to_gpu(tod0 if accum else rng.standard_normal(tod0.shape).astype(F32), dev)
        gmm.map2tod(t, mg, inst["xg"], p, accum=accum, debug=True)
        got = to_cpu(t)
        e32, e64 = l1rel(got, ref32), l1rel(got, ref64)
        print(f"  map2tod plan={'planned' if p is not None else 'unplanned'} accum={accum}: "
              f"l1rel vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}")
        assert e32 < 1e-6
        assert e64 < 1e-5


def test_tod2map(inst, gmm, orc, dev):
    rng = np.random.default_rng(200 + inst["seed"])
    lp, plan = inst["lp"], inst["plan"]
    tod = rng.standard_normal(inst["xp"].shape[1:]).astype(F32)
    m0 = rng.standard_normal(3 * lp.npix).astype(F32)
    pix = opix(orc, inst)
    ref32, err = orc.tod2map(tod, inst["xp"], pix, lmap=m0)
    assert err == 0
    ref64, _ = orc.tod2map(tod.astype(np.float64), inst["xp"], pix, lmap=m0.astype(np.float64), f64=True)
    tg = to_gpu(tod, dev)
    for p in (plan, None):
        mg = gmm.LocalMap(lp, to_gpu(m0, dev))
        gmm.tod2map(mg, tg, inst["xg"], p, debug=True)
        got = to_cpu(mg.arr)
        e32, e64 = l1rel(got, ref32), l1rel(got, ref64)
        print(f"  tod2map plan={'planned' if p is not None else 'unplanned'}: "
              f"l1rel vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}")
        assert e32 < 1e-6
        assert e64 < 1e-5


def test_adjoint(inst, gmm, dev):
    rng = np.random.default_rng(300 + inst["seed"])
    lp, plan = inst["lp"], inst["plan"]
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    d = rng.standard_normal(inst["xp"].shape[1:]).astype(F32)
    am = torch.zeros(d.shape, dtype=torch.float32, device=dev)
    gmm.map2tod(am, gmm.LocalMap(lp, to_gpu(m, dev)), inst["xg"], plan)
    atd = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(atd, to_gpu(d, dev), inst["xg"], plan)
    lhs = float(np.dot(to_cpu(am).astype(np.float64).ravel(), d.astype(np.float64).ravel()))
    rhs = float(np.dot(m.astype(np.float64), to_cpu(atd.arr).astype(np.float64)))
    scale = np.linalg.norm(to_cpu(am)) * np.linalg.norm(d)
    assert abs(lhs - rhs) < 1e-6 * scale


# ------------------------------------------------------------------------------------------
# Edge cases.

def test_partial_pixelization_gpu(gmm, orc, dev):
    rng = np.random.default_rng(7)
    ny, nx = 256, 320
    cells = [(cy, cx) for cy in range(4) for cx in range(5) if (cy + cx) % 3 != 0]
    lp = gmm.LocalPixelization.make_cells(ny, nx, cells, periodic_xcoord=False)
    pix = orc.Pix(ny, nx, lp.cell_offsets_cpu, 64, 4096, False)
    n = 32 * 2048
    t = np.arange(n)
    xp = np.stack([128 + 120 * np.sin(t / 900.0), 160 + 150 * np.sin(t / 1300.0 + 1), 0.002 * t]).astype(F32)
    xg = to_gpu(xp, dev)
    pre = gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False)
    plan = gmm.PointingPlan(pre, xg)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    mg = gmm.LocalMap(lp, to_gpu(m, dev))
    tg = torch.zeros(n, dtype=torch.float32, device=dev)
    for p in (plan, None):
        with pytest.raises(RuntimeError, match="not in LocalPixelization"):
            gmm.map2tod(tg, mg, xg, p)
        with pytest.raises(RuntimeError, match="not in LocalPixelization"):
            gmm.tod2map(gmm.LocalMap(lp, torch.zeros_like(mg.arr)), tg, xg, p)
    ref, err = orc.map2tod(m, xp, pix, partial=True)
    assert err == 0
    d = rng.standard_normal(n).astype(F32)
    mref, _ = orc.tod2map(d, xp, pix, partial=True)
    for p in (plan, None):
        gmm.map2tod(tg, mg, xg, p, partial_pixelization=True)
        assert l1rel(to_cpu(tg), ref) < 1e-6
        out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
        gmm.tod2map(out, to_gpu(d, dev), xg, p, partial_pixelization=True)
        assert l1rel(to_cpu(out.arr), mref) < 1e-6


def test_bad_pointing_raises(gmm, dev):
    ny, nx = 128, 128
    n = 32 * 64
    xp = np.zeros((3, n), F32)
    xp[0] = 10.0
    xp[1] = np.linspace(0, 120, n)
    xp[0, 77] = 200.0
    xg = to_gpu(xp, dev)
    with pytest.raises(RuntimeError, match="y-value is outside range"):
        gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False)
    xp[0, 77] = 10.0
    xp[1, 99] = 127.5
    with pytest.raises(RuntimeError, match="periodic_xcoord"):
        gmm.PointingPrePlan(to_gpu(xp, dev), ny, nx, periodic_xcoord=False)
    # Valid plan, then pointing goes bad: the operators report it.
    xp[1, 99] = 50.0
    xg = to_gpu(xp, dev)
    pre = gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False)
    plan = gmm.PointingPlan(pre, xg)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    mg = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    tg = torch.zeros(n, dtype=torch.float32, device=dev)
    xg[0, 5] = -3.0
    for p in (plan, None):
        with pytest.raises(RuntimeError, match="y-value"):
            gmm.map2tod(tg, mg, xg, p)
    # Plan-free calls have no plan to defer their errors to: they raise in deferred mode too.
    with gmm.deferred_errflags():
        with pytest.raises(RuntimeError, match="y-value"):
            gmm.map2tod(tg, mg, xg, None)
        with pytest.raises(RuntimeError, match="y-value"):
            gmm.tod2map(mg, tg, xg, None)
    # The flags were cleared by the failed check: a good call succeeds.
    xg[0, 5] = 10.0
    gmm.map2tod(tg, mg, xg, plan)
    gmm.map2tod(tg, mg, xg, None)


def test_plan_from_other_pointing_raises(gmm, dev):
    ny, nx = 256, 256
    n = 32 * 1024
    t = np.arange(n)
    a = np.stack([20 + 0.006 * t, 128 + 100 * np.sin(t / 300.0), 0 * t]).astype(F32)
    b = np.stack([20 + 0.006 * t, 128 + 100 * np.sin(t / 30.0), 0 * t]).astype(F32)
    pre = gmm.PointingPrePlan(to_gpu(a, dev), ny, nx, periodic_xcoord=False)
    with pytest.raises(RuntimeError, match="inconsistent value of nmt"):
        gmm.PointingPlan(pre, to_gpu(b, dev))


def test_small_buffers_raise(gmm, dev):
    xp = to_gpu(synth.make_cfg0(seed=0, ndet=2, nt=4096)["xpointing"], dev)
    pre = gmm.PointingPrePlan(xp, 256, 512, periodic_xcoord=False)
    small = torch.empty(pre.plan_nbytes - 8, dtype=torch.uint8, device=dev)
    with pytest.raises(RuntimeError, match="'buf'"):
        gmm.PointingPlan(pre, xp, buf=small)
    small = torch.empty(pre.plan_constructor_tmp_nbytes - 8, dtype=torch.uint8, device=dev)
    with pytest.raises(RuntimeError, match="'tmp_buf'"):
        gmm.PointingPlan(pre, xp, tmp_buf=small)


@pytest.mark.parametrize("ncells,lo,hi,nper,waves", [(100, 1, 1000, 37, 4), (700, 1, 50, 512, 8),
                                                     (1, 1, 16384, 4096, 16), (1000, 5, 5, 1, 4)])
def test_plan_iterator_random(gmm, ncells, lo, hi, nper, waves):
    mt = gmm.make_random_plan_mt(ncells, lo, hi, seed=ncells + nper)
    gmm.test_plan_iterator(mt, nper, waves)


def test_shared_plan_buffers(gmm, dev):
    """tmp_buf may be shared by several plans; buf may not (gpu_mm.py:1227-1232)."""
    d0 = synth.make_cfg0(seed=0, ndet=4, nt=8192)
    d1 = synth.make_cfg0(seed=1, ndet=4, nt=8192)
    x0, x1 = to_gpu(d0["xpointing"], dev), to_gpu(d1["xpointing"], dev)
    p0 = gmm.PointingPrePlan(x0, 256, 512, periodic_xcoord=False)
    p1 = gmm.PointingPrePlan(x1, 256, 512, periodic_xcoord=False)
    tmp = torch.empty(max(p0.plan_constructor_tmp_nbytes, p1.plan_constructor_tmp_nbytes), dtype=torch.uint8, device=dev)
    a = gmm.PointingPlan(p0, x0, tmp_buf=tmp)
    b = gmm.PointingPlan(p1, x1, tmp_buf=tmp)
    a2 = gmm.PointingPlan(p0, x0)
    assert np.array_equal(a.get_plan_mt(), a2.get_plan_mt())
    assert b.plan_nmt == p1.plan_nmt


def test_large_angles(gmm, orc, dev):
    """The device sin/cos of 2*alpha: polynomial path (|2a| <= 1e6) and double reduction beyond."""
    rng = np.random.default_rng(11)
    ny, nx = 128, 256
    n = 32 * 4096
    t = np.arange(n)
    alpha = np.concatenate([rng.uniform(-3, 3, n // 4), rng.uniform(-4e5, 4e5, n // 4),
                            rng.uniform(-6e6, 6e6, n // 4), rng.uniform(-1e9, 1e9, n // 4)])
    xp = np.stack([64 + 50 * np.sin(t / 500.0), 128 + 100 * np.sin(t / 700.0), alpha]).astype(F32)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    pix = orc.Pix.rectangle(ny, nx, False)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    m[:lp.npix] = 0.0          # T = 0: the TOD is pure Q cos + U sin, so angle errors show fully
    xg = to_gpu(xp, dev)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False), xg)
    ref32, err = orc.map2tod(m, xp, pix)
    assert err == 0
    ref64, _ = orc.map2tod(m.astype(np.float64), xp, pix, f64=True)
    tg = torch.zeros(n, dtype=torch.float32, device=dev)
    for p in (plan, None):
        gmm.map2tod(tg, gmm.LocalMap(lp, to_gpu(m, dev)), xg, p)
        got = to_cpu(tg)
        for sl in (slice(0, n // 4), slice(n // 4, n // 2), slice(n // 2, 3 * n // 4), slice(3 * n // 4, n)):
            assert l1rel(got[sl], ref32[sl]) < 1e-6
            assert l1rel(got[sl], ref64[sl]) < 1e-5


# ------------------------------------------------------------------------------------------
# Per-detector response (src_lib/map2tod_response.cu, tod2map_response.cu).

def test_response_ops(inst, gmm, orc, dev):
    rng = np.random.default_rng(400 + inst["seed"])
    lp, plan = inst["lp"], inst["plan"]
    ndet = inst["xp"].shape[1]
    pix = opix(orc, inst)
    resp = rng.uniform(-2, 2, size=(2, ndet)).astype(F32)
    rg = to_gpu(resp, dev)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    tod0 = rng.standard_normal(inst["xp"].shape[1:]).astype(F32)
    mg = gmm.LocalMap(lp, to_gpu(m, dev))
    for accum in (False, True):
        ref, err = orc.response_map2tod(m, inst["xp"], resp, pix, tod=tod0, accum=accum)
        assert err == 0
        t = to_gpu(tod0, dev)
        gmm.map2tod(t, mg, inst["xg"], plan, response=rg, accum=accum, debug=True)
        e = l1rel(to_cpu(t), ref)
        print(f"  response map2tod accum={accum}: l1rel vs f64 oracle {e:.2e}")
        assert e < F64_BAR
    m0 = rng.standard_normal(3 * lp.npix).astype(F32)
    ref, err = orc.response_tod2map(tod0, inst["xp"], resp, pix, lmap=m0)
    assert err == 0
    out = gmm.LocalMap(lp, to_gpu(m0, dev))
    gmm.tod2map(out, to_gpu(tod0, dev), inst["xg"], plan, response=rg, debug=True)
    e = l1rel(to_cpu(out.arr), ref)
    print(f"  response tod2map: l1rel vs f64 oracle {e:.2e}")
    assert e < F64_BAR


def test_response_reference_check(inst, gmm, orc, dev):
    """gpu_mm/tests.py:295-315: response (2, -1) with the map divided by (2, -1) gives the plain TOD."""
    rng = np.random.default_rng(500 + inst["seed"])
    lp, plan = inst["lp"], inst["plan"]
    ndet = inst["xp"].shape[1]
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    ref, _ = orc.map2tod(m, inst["xp"], opix(orc, inst))
    m2 = m.copy()
    m2[:lp.npix] /= 2
    m2[lp.npix:] /= -1
    resp = torch.zeros((2, ndet), dtype=torch.float32, device=dev)
    resp[0], resp[1] = 2, -1
    t = torch.zeros(inst["xp"].shape[1:], dtype=torch.float32, device=dev)
    gmm.map2tod(t, gmm.LocalMap(lp, to_gpu(m2, dev)), inst["xg"], plan, response=resp, debug=True)
    assert l1rel(to_cpu(t), ref) < 1e-6


@pytest.mark.parametrize("ndet", [2240, 20160, 630])
def test_response_detector_boundaries(gmm, orc, dev, ndet):
    """nsamp/ndet = 9, 1 and 32 samples per detector: detector boundaries inside a 4-sample group."""
    rng = np.random.default_rng(ndet)
    ny, nx, n = 200, 300, 20160
    t = np.arange(n)
    xp = np.stack([100 + 90 * np.sin(t / 300.0), 150 + 140 * np.sin(t / 410.0),
                   rng.uniform(-3, 3, n)]).astype(F32)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    pix = orc.Pix.rectangle(ny, nx, False)
    xg = to_gpu(xp, dev)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False), xg)
    resp = rng.uniform(-2, 2, size=(2, ndet)).astype(F32)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    d = rng.standard_normal(n).astype(F32)
    tg = torch.zeros(n, dtype=torch.float32, device=dev)
    gmm.map2tod(tg, gmm.LocalMap(lp, to_gpu(m, dev)), xg, plan, response=to_gpu(resp, dev))
    ref, _ = orc.response_map2tod(m, xp, resp, pix)
    assert l1rel(to_cpu(tg), ref) < F64_BAR
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, to_gpu(d, dev), xg, plan, response=to_gpu(resp, dev))
    ref, _ = orc.response_tod2map(d, xp, resp, pix)
    assert l1rel(to_cpu(out.arr), ref) < F64_BAR


def test_response_argument_errors(gmm, dev):
    ny, nx, n = 128, 128, 4096
    xp = np.stack([np.full(n, 60.0), np.linspace(10, 100, n), np.zeros(n)]).astype(F32)
    xg = to_gpu(xp, dev)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False), xg)
    mg = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    tg = torch.zeros(n, dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError, match="shape"):
        gmm.map2tod(tg, mg, xg, plan, response=torch.ones((3, 4), device=dev))
    with pytest.raises(RuntimeError, match="divide"):
        gmm.map2tod(tg, mg, xg, plan, response=torch.ones((2, 3), device=dev))
    with pytest.raises(RuntimeError, match="divide"):
        gmm.tod2map(mg, tg, xg, plan, response=torch.ones((2, 3), device=dev))
    with pytest.raises(RuntimeError, match="PointingPlan"):
        gmm.map2tod(tg, mg, xg, None, response=torch.ones((2, 4), device=dev))
    gmm.map2tod(tg, mg, xg, plan, response=torch.ones((2, 4), device=dev))


# ------------------------------------------------------------------------------------------
# 16-bit TODs (FP16 / BF16 values with one FP32 scale per row). Not in the reference.

TOD16 = [(torch.float16, 32752.0, 2.0 ** -11), (torch.bfloat16, float.fromhex("0x1.fep126"), 2.0 ** -8)]


def _decode(t16, scale):
    return (t16.float() * scale[:, None]).cpu().numpy()


@pytest.mark.parametrize("dtype,half_max,eps", TOD16, ids=["f16", "bf16"])
def test_tod16_encode(gmm, dev, dtype, half_max, eps):
    rng = np.random.default_rng(31)
    tod = rng.standard_normal((37, 1000)).astype(F32) * np.exp(rng.uniform(-20, 20, (37, 1)).astype(F32))
    tod[5] = 0.0
    tg = to_gpu(tod, dev)
    t16, sc = gmm.encode_tod16(tg, dtype)
    assert t16.dtype == dtype and sc.dtype == torch.float32 and sc.shape == (37,)
    expect = np.maximum(np.abs(tod).max(axis=1) / F32(half_max), np.finfo(F32).tiny).astype(F32)
    assert np.array_equal(to_cpu(sc), expect)
    dec = _decode(t16, sc)
    assert np.all(np.abs(dec - tod) <= eps * np.abs(tod) + 1e-45)
    assert np.all(dec[5] == 0.0)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("dtype,half_max,eps", TOD16, ids=["f16", "bf16"])
def test_tod16_encode_error_feedback(gmm, dev, dtype, half_max, eps, order):
    """With error feedback the running sum (order 1) or the running double sum (order 2) of the
    decoding error stays within a few rounding errors, independently of the row length: each
    thread's run of a row carries its rounding errors forward (256 runs per row)."""
    rng = np.random.default_rng(41 + order)
    nrows, row_len = 6, 256 * 97 + 13
    tod = rng.standard_normal((nrows, row_len)).astype(F32) * np.exp(rng.uniform(-5, 5, (nrows, 1))).astype(F32)
    t16, sc = gmm.encode_tod16(to_gpu(tod, dev), dtype, error_feedback=order)
    plain16, sc0 = gmm.encode_tod16(to_gpu(tod, dev), dtype)
    assert np.array_equal(to_cpu(sc), to_cpu(sc0))
    dec = (t16.double() * sc.double()[:, None]).cpu().numpy()
    err = dec - tod.astype(np.float64)
    run = -(-row_len // 256)
    for r in range(nrows):
        s = float(to_cpu(sc)[r])
        # one rounding error of the encoded scale (value / scale <= half_max) in decoded units
        h = eps * half_max * s
        for i0 in range(0, row_len, run):
            e = err[r, i0:i0 + run]
            c = np.cumsum(e) if order == 1 else np.cumsum(np.cumsum(e))
            assert np.max(np.abs(c)) <= 2.0 * h, (r, i0)
    # Independent rounding lets the running sum grow like sqrt(n) (a random walk over the row).
    e0 = (plain16.double() * sc0.double()[:, None]).cpu().numpy() - tod
    s0 = float(np.max(np.abs(np.cumsum(e0[0]))))
    assert s0 > 4.0 * float(np.max(np.abs(np.cumsum(err[0, :run]))))


def test_tod16_error_feedback_map_error(gmm, orc, dev):
    """FP16 TOD of a slice of the 2048x262144 raster (16 detectors): the map built by tod2map from
    the encoded TOD, against the FP64 oracle's map of the float32 TOD, RMS error relative to the
    map RMS over hit pixels: about 2.1e-4 with independent rounding, below 1e-4 with error
    feedback of order 1 and 2."""
    c = synth.CFG1
    ny, nx = c["ny"], c["nx"]
    xp, _ = synth.raster_xpointing(16, 262144, seed=1, **c["scan"])
    tod = np.random.default_rng(3).standard_normal(xp.shape[1:]).astype(F32)
    print(f"\n[error feedback] FP16 TOD, xpointing {xp.shape}", flush=True)
    xg = to_gpu(xp, dev)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False), xg)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    pix = orc.Pix.rectangle(ny, nx, False)
    ref, err = orc.tod2map(tod.astype(np.float64), xp, pix, f64=True)
    assert err == 0
    hit = ref[:pix.npix] != 0.0
    errs = {}
    for order in (0, 1, 2):
        t16, sc = gmm.encode_tod16(to_gpu(tod, dev), torch.float16, error_feedback=order)
        out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
        gmm.tod2map(out, t16, xg, plan, tod_scale=sc, debug=True)
        got = to_cpu(out.arr).astype(np.float64)
        e = []
        for comp in range(3):
            r = ref[comp * pix.npix:(comp + 1) * pix.npix][hit]
            g = got[comp * pix.npix:(comp + 1) * pix.npix][hit]
            e.append(float(np.sqrt(np.mean((g - r) ** 2)) / np.sqrt(np.mean(r ** 2))))
        errs[order] = max(e)
        print(f"  order {order}: map RMS error / map RMS {errs[order]:.3e}", flush=True)
    assert 1.5e-4 < errs[0] < 3e-4
    assert errs[1] < 1e-4 and errs[2] < 1e-4


@pytest.mark.parametrize("dtype,half_max,eps", TOD16, ids=["f16", "bf16"])
def test_tod16_ops(inst, gmm, orc, dev, dtype, half_max, eps):
    rng = np.random.default_rng(600 + inst["seed"])
    lp, plan = inst["lp"], inst["plan"]
    pix = opix(orc, inst)
    shape = inst["xp"].shape[1:]
    tod = rng.standard_normal(shape).astype(F32)
    t16, sc = gmm.encode_tod16(to_gpu(tod, dev), dtype)
    dec = _decode(t16, sc)

    # tod2map on the decoded samples: the same FP32 inputs as the float32 operator.
    m0 = rng.standard_normal(3 * lp.npix).astype(F32)
    ref, err = orc.tod2map(dec.astype(np.float64), inst["xp"], pix, lmap=m0.astype(np.float64), f64=True)
    assert err == 0
    out = gmm.LocalMap(lp, to_gpu(m0, dev))
    gmm.tod2map(out, t16, inst["xg"], plan, tod_scale=sc, debug=True)
    e = l1rel(to_cpu(out.arr), ref)
    print(f"  tod2map {dtype}: l1rel vs f64 oracle on decoded TOD {e:.2e}")
    assert e < F64_BAR

    # map2tod: outputs are value / scale rounded to the 16-bit type (twice on multi-cell lines).
    # Output scales with headroom for |T + cos Q + sin U| of an N(0,1) map, doubled by accum.
    osc = sc * 8
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    mg = gmm.LocalMap(lp, to_gpu(m, dev))
    ref, _ = orc.map2tod(m.astype(np.float64), inst["xp"], pix, f64=True)
    o16 = torch.empty_like(t16)
    gmm.map2tod(o16, mg, inst["xg"], plan, tod_scale=osc, debug=True)
    got = _decode(o16, osc).reshape(-1)
    assert np.all(np.isfinite(got))
    e = l1rel(got, ref)
    print(f"  map2tod {dtype}: l1rel vs f64 oracle {e:.2e}")
    assert e < eps

    # accum: adds into the existing 16-bit values.
    gmm.map2tod(o16, mg, inst["xg"], plan, tod_scale=osc, accum=True)
    got2 = _decode(o16, osc).reshape(-1)
    assert np.all(np.isfinite(got2))
    assert l1rel(got2, 2 * ref) < 2 * eps


@pytest.mark.parametrize("dtype,half_max,eps", TOD16, ids=["f16", "bf16"])
def test_tod16_response(gmm, orc, dev, dtype, half_max, eps):
    rng = np.random.default_rng(77)
    ny, nx, ndet, nt = 200, 300, 24, 840
    n = ndet * nt
    t = np.arange(n)
    xp = np.stack([100 + 90 * np.sin(t / 300.0), 150 + 140 * np.sin(t / 410.0),
                   rng.uniform(-3, 3, n)]).astype(F32).reshape(3, ndet, nt)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    pix = orc.Pix.rectangle(ny, nx, False)
    xg = to_gpu(xp, dev)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False), xg)
    resp = rng.uniform(-2, 2, size=(2, 2 * ndet)).astype(F32)   # 420 samples per response detector
    tod = rng.standard_normal((ndet, nt)).astype(F32)
    t16, sc = gmm.encode_tod16(to_gpu(tod, dev), dtype)
    dec = _decode(t16, sc)
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, t16, xg, plan, tod_scale=sc, response=to_gpu(resp, dev))
    ref, _ = orc.response_tod2map(dec, xp, resp, pix)
    assert l1rel(to_cpu(out.arr), ref) < F64_BAR
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    o16 = torch.empty_like(t16)
    osc = sc * 16          # headroom for |response| <= 2 times |T + cos Q + sin U|
    gmm.map2tod(o16, gmm.LocalMap(lp, to_gpu(m, dev)), xg, plan, tod_scale=osc, response=to_gpu(resp, dev))
    ref, _ = orc.response_map2tod(m, xp, resp, pix)
    got = _decode(o16, osc)
    assert np.all(np.isfinite(got))
    assert l1rel(got, ref) < eps


@pytest.mark.parametrize("kind", ["f32", "f16"])
def test_tod2map_nonfinite_fallback(gmm, orc, dev, kind):
    """A NaN sample: the touched pixels become NaN and every other pixel stays exact."""
    rng = np.random.default_rng(5)
    ny, nx, n = 128, 256, 32 * 512
    t = np.arange(n)
    xp = np.stack([64 + 50 * np.sin(t / 500.0), 128 + 100 * np.sin(t / 700.0),
                   rng.uniform(-3, 3, n)]).astype(F32)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    pix = orc.Pix.rectangle(ny, nx, False)
    xg = to_gpu(xp, dev)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False), xg)
    tod = rng.standard_normal(n).astype(F32)
    tod[1234] = np.nan
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    if kind == "f32":
        gmm.tod2map(out, to_gpu(tod, dev), xg, plan)
        dec = tod
    else:
        t16, sc = gmm.encode_tod16(to_gpu(tod, dev), torch.float16)
        gmm.tod2map(out, t16, xg, plan, tod_scale=sc)
        dec = (t16.float() * sc).cpu().numpy()
    ref, _ = orc.tod2map(dec.astype(np.float64), xp, pix, f64=True)
    got = to_cpu(out.arr)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = ~np.isnan(ref)
    assert np.isnan(ref).sum() > 0
    assert l1rel(got[fin], ref[fin]) < F64_BAR


# ------------------------------------------------------------------------------------------
# DynamicMap (gpu_mm/gpu_mm.py:586-806, src_lib/expand_dynamic_map.cu).

def _random_plan_mt(rng, nycells, nxcells):
    """Cell-sorted plan entries with junk high bits (gpu_mm/tests.py:536-584)."""
    ntot = nycells * nxcells
    nhit = int(rng.integers(1, ntot + 1))
    nmt = int(rng.integers(nhit, 200000))
    iy, ix = np.mgrid[:nycells, :nxcells]
    icell = rng.permutation(((iy << 10) + ix).ravel())[:nhit]
    mt = np.concatenate([icell, icell[rng.integers(0, nhit, size=nmt - nhit)]]).astype(np.uint64)
    mt = np.sort(mt)
    return mt | (rng.integers(0, 1 << 28, size=nmt).astype(np.uint64) << np.uint64(20))


@pytest.mark.parametrize("seed", range(6))
def test_expand_dynamic_map_kernel(gmm, dev, seed):
    """gpu_mm/tests.py:654-716, seeded; additionally the new indices follow cell-key order."""
    rng = np.random.default_rng(900 + seed)
    nycells, nxcells = int(rng.integers(1, 40)), int(rng.integers(1, 40))
    mt = _random_plan_mt(rng, nycells, nxcells)
    off = np.where(rng.uniform(size=(nycells, nxcells)) < rng.uniform(-0.2, 1.2), -1, -2).astype(np.int64)
    cur = rng.uniform(size=(nycells, nxcells)) < rng.uniform(-0.2, 1.2)
    ncur = int(cur.sum())
    off[cur] = rng.permutation(ncur) * 3 * 64 * 64
    og = to_gpu(off, dev)
    g = torch.full((1,), ncur, dtype=torch.uint32, device=dev)
    n = gmm.expand_dynamic_map(g, og, to_gpu(mt, dev))
    got = to_cpu(og)

    icells = np.unique(mt & np.uint64((1 << 20) - 1))
    iy, ix = (icells >> np.uint64(10)).astype(np.int64), (icells & np.uint64(1023)).astype(np.int64)
    keep = (iy < nycells) & (ix < nxcells)
    iy, ix = iy[keep], ix[keep]
    new = off[iy, ix] == -1
    iy, ix = iy[new], ix[new]
    assert n == ncur + len(iy) and int(to_cpu(g)[0]) == n
    assert np.array_equal(got[iy, ix], (ncur + np.arange(len(iy))) * 3 * 64 * 64)
    unchanged = np.ones_like(cur)
    unchanged[iy, ix] = False
    assert np.array_equal(got[unchanged], off[unchanged])


@pytest.mark.parametrize("masked", [False, True])
def test_dynamic_map(inst, gmm, orc, dev, masked):
    """gpu_mm/tests.py:356-411: tod2map over time chunks into a DynamicMap equals the reference."""
    rng = np.random.default_rng(700 + inst["seed"])
    ny, nx, periodic = inst["ny"], inst["nx"], inst["periodic"]
    xp = inst["xp"]
    nt = xp.shape[-1]
    tod = rng.standard_normal(xp.shape[1:]).astype(F32)
    nyc, nxc = (ny + 63) // 64, (nx + 63) // 64
    mask = (rng.uniform(size=(nyc, nxc)) < 0.7) if masked else None

    # Cells deposit independently, so the masked reference is the full one with the untargeted
    # cells zeroed.
    ref, _ = orc.tod2map(tod.astype(np.float64), xp, opix(orc, inst), f64=True)
    ref = ref.reshape(3, nyc, 64, nxc, 64)
    if masked:
        ref.transpose(0, 1, 3, 2, 4)[:, ~mask] = 0.0
    ref_global = ref.reshape(3, 64 * nyc, 64 * nxc)[:, :ny, :nx]

    chunks, done = [], 0
    while done < nt:
        n = int(min(32 * int(rng.integers(1, nt // 16 // 32 + 2)), nt - done))
        chunks.append((done, n))
        done += n
    dmap = gmm.DynamicMap(ny, nx, np.float32, cell_mask=mask, periodic_xcoord=periodic, initial_ncells=1)
    for (t0, n) in chunks:
        xc = to_gpu(np.ascontiguousarray(xp[..., t0:t0 + n]), dev)
        tc = to_gpu(np.ascontiguousarray(tod[..., t0:t0 + n]), dev)
        plan = gmm.PointingPlan(gmm.PointingPrePlan(xc, ny, nx, periodic_xcoord=periodic), xc)
        gmm.tod2map(dmap, tc, xc, plan, partial_pixelization=masked, debug=True)
    dmap.randomly_permute(rng)
    lmap = dmap.finalize()
    got = lmap.to_global()
    print(f"  dynamic map: {len(chunks)} chunks, {dmap.ncells_curr} cells")
    assert l1rel(got, ref_global) < F64_BAR
    if masked:
        assert np.all(lmap.pixelization.cell_offsets_cpu[~mask] < 0)


# ------------------------------------------------------------------------------------------
# Cell utilities (src_lib/cell_broadcast.cu, cell_reduce.cu, local_map_to_global.cu).

def test_cell_broadcast_reduce(gmm, dev):
    rng = np.random.default_rng(41)
    n = 3 * 64 * 64
    nsrc, nout = 37, 2500
    src = rng.standard_normal(nsrc * n).astype(F32)
    idx = rng.integers(0, nsrc, size=nout).astype(np.int64)
    dst = torch.empty(nout * n, dtype=torch.float32, device=dev)
    gmm.cell_broadcast(dst, to_gpu(src, dev), to_gpu(idx, dev))
    assert np.array_equal(to_cpu(dst).reshape(nout, n), src.reshape(nsrc, n)[idx])

    # Reduce back: duplicates accumulate in increasing source order (a sequential float32 sum).
    red = torch.full((nsrc * n,), 7.0, dtype=torch.float32, device=dev)
    parts = rng.standard_normal(nout * n).astype(F32)
    gmm.cell_reduce(red, to_gpu(parts, dev), to_gpu(idx, dev))
    expect = np.zeros((nsrc, n), dtype=F32)
    for i in range(nout):
        expect[idx[i]] += parts.reshape(nout, n)[i]
    assert np.array_equal(to_cpu(red).reshape(nsrc, n), expect)

    with pytest.raises(RuntimeError, match="out of range"):
        gmm.cell_broadcast(dst, to_gpu(src, dev), to_gpu(np.full(nout, nsrc, dtype=np.int64), dev))
    with pytest.raises(RuntimeError, match="out of range"):
        gmm.cell_reduce(red, to_gpu(parts, dev), to_gpu(np.full(nout, -1, dtype=np.int64), dev))


def test_local_map_to_global_gpu(gmm, dev):
    rng = np.random.default_rng(42)
    ny, nx = 200, 330
    cells = [(cy, cx) for cy in range(4) for cx in range(6) if (cy * 7 + cx) % 4 != 1]
    lp = gmm.LocalPixelization.make_cells(ny, nx, cells, periodic_xcoord=True)
    arr = rng.standard_normal(3 * lp.npix).astype(F32)
    expect = gmm.LocalMap(lp, arr).to_global()           # host path (numpy)
    got = gmm.LocalMap(lp, to_gpu(arr, dev)).to_global()  # device kernel
    assert got.shape == (3, ny, nx)
    assert np.array_equal(got, expect)
    # Independent check of a few pixels.
    off = lp.cell_offsets_cpu
    for (y, x) in [(0, 0), (63, 64), (199, 329), (130, 200)]:
        o = off[y // 64, x // 64]
        for p in range(3):
            v = 0.0 if o < 0 else arr[o + p * 4096 + (y % 64) * 64 + (x % 64)]
            assert got[p, y, x] == v


# ------------------------------------------------------------------------------------------
# TODs whose length is not a multiple of 32 (the reference requires one; here the last line of
# a plan may be partial). nt = 4093: nsamp % 4 != 0 (scalar kernel variant, a cut last group);
# nt = 4116: nsamp % 32 == 28 (vector variant; the last group of the last line is empty);
# nt = 4112: nsamp % 32 == 16 (vector variant; the last four groups are empty).

@pytest.mark.parametrize("nt", [4093, 4116, 4112], ids=["nsamp%4!=0", "nsamp%32==28", "nsamp%32==16"])
@pytest.mark.parametrize("periodic", [False, True], ids=["nonperiodic", "periodic"])
def test_partial_last_line(gmm, orc, dev, nt, periodic):
    rng = np.random.default_rng(nt + periodic)
    ndet, ny, nx = 3, 200, 320
    t = np.arange(nt)
    xp = np.empty((3, ndet, nt), F32)
    for d in range(ndet):
        xp[0, d] = 100 + 90 * np.sin(t / 300.0 + d)
        xp[1, d] = (150 + 140 * np.sin(t / 170.0 + 2 * d)) if not periodic else ((t * 1.9 + 100 * d) % (2 * nx) - nx / 2)
        xp[2, d] = rng.uniform(0, np.pi) + 0.003 * t
    nsamp = ndet * nt
    print(f"\n[partial last line] nsamp={nsamp} periodic={periodic}", flush=True)
    xg = to_gpu(xp, dev)
    pre = gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=periodic)
    plan = gmm.PointingPlan(pre, xg)
    cum, nmt, err = orc.plan_count(xp, ny, nx, periodic)
    assert err == 0 and pre.plan_nmt == nmt
    assert np.array_equal(pre.get_nmt_cumsum(), cum)
    ref_mt, err = orc.plan_build(xp, ny, nx, periodic)
    assert err == 0 and np.array_equal(plan.get_plan_mt(), ref_mt)
    # The brute-force tester stops at nsamp inside the partial last line (gpu_mm/tests.py:227-254).
    tester = gmm.PointingPlanTester(pre, xg)
    assert np.array_equal(tester.nmt_cumsum, cum)
    assert np.array_equal(tester.sorted_mt, np.sort(ref_mt & np.uint64((1 << 46) - 1)))

    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=periodic)
    pix = orc.Pix.rectangle(ny, nx, periodic)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    mg = gmm.LocalMap(lp, to_gpu(m, dev))
    ref32, _ = orc.map2tod(m, xp, pix)
    ref64, _ = orc.map2tod(m.astype(np.float64), xp, pix, f64=True)
    for p in (plan, None):
        tg = torch.full((ndet, nt), 5.0, dtype=torch.float32, device=dev)
        gmm.map2tod(tg, mg, xg, p, debug=True)
        got = to_cpu(tg)
        assert l1rel(got, ref32) < 1e-6 and l1rel(got, ref64) < 1e-5
    d = rng.standard_normal((ndet, nt)).astype(F32)
    mref32, _ = orc.tod2map(d, xp, pix)
    mref64, _ = orc.tod2map(d.astype(np.float64), xp, pix, f64=True)
    for p in (plan, None):
        out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
        gmm.tod2map(out, to_gpu(d, dev), xg, p, debug=True)
        assert l1rel(to_cpu(out.arr), mref32) < 1e-6 and l1rel(to_cpu(out.arr), mref64) < 1e-5
    # Per-detector response (rows of nt samples).
    resp = rng.uniform(0.5, 1.5, (2, ndet)).astype(F32)
    rg = to_gpu(resp, dev)
    rref, _ = orc.response_map2tod(m, xp, resp, pix)
    tg = torch.empty((ndet, nt), dtype=torch.float32, device=dev)
    gmm.map2tod(tg, mg, xg, plan, response=rg, debug=True)
    assert l1rel(to_cpu(tg), rref) < F64_BAR          # the response oracle is FP64
    mr, _ = orc.response_tod2map(d, xp, resp, pix)
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, to_gpu(d, dev), xg, plan, response=rg, debug=True)
    assert l1rel(to_cpu(out.arr), mr) < F64_BAR
    # 16-bit TODs and batches use 4-sample vector accesses: nsamp % 4 must be 0.
    if nsamp % 4 != 0:
        t16, sc = gmm.encode_tod16(to_gpu(d, dev))
        with pytest.raises(RuntimeError, match="divisible by 4"):
            gmm.tod2map(out, t16, xg, plan, tod_scale=sc)
        with pytest.raises(RuntimeError, match="divisible by 4"):
            gmm.PlanBatch(lp, [plan], [xg], [tg])
    else:
        b = gmm.PlanBatch(lp, [plan], [xg], [tg])
        b.map2tod(mg, debug=True)
        assert l1rel(to_cpu(tg), ref32) < 1e-6
        tg.copy_(to_gpu(d, dev))
        out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
        b.tod2map(out, debug=True)
        assert l1rel(to_cpu(out.arr), mref32) < 1e-6 and l1rel(to_cpu(out.arr), mref64) < 1e-5


# ------------------------------------------------------------------------------------------
# Halo ownership rule near the right edge of the map: slow x scans (lines handed to the
# left-hand cell's block) that cross into a last cell column narrower than the halo (nx % 64 =
# 5 or 11 columns) and, in periodic mode, wrap around to column 0.

@pytest.mark.parametrize("nx", [64 * 4 + 5, 64 * 3 + 11])
@pytest.mark.parametrize("periodic", [False, True], ids=["nonperiodic", "periodic"])
def test_halo_rule_right_edge(gmm, orc, dev, nx, periodic):
    rng = np.random.default_rng(nx + periodic)
    ndet, nt, ny = 16, 32 * 256, 130
    t = np.arange(nt)
    tri = 1.0 - 4.0 * np.abs((t / 600.0) % 1.0 - 0.5)           # period 600 samples, 0.27 px/sample
    xp = np.empty((3, ndet, nt), F32)
    for d in range(ndet):
        xp[0, d] = 30 + 4.5 * d + 0.002 * t
        if periodic:
            xp[1, d] = (nx - 60 + 80 * (0.5 + 0.5 * tri) + 0.37 * d) % nx   # crosses x = nx and wraps
        else:
            xp[1, d] = nx - 80 + 78 * (0.5 + 0.5 * tri) - 0.11 * d          # up to x = nx - 2
        xp[2, d] = rng.uniform(0, np.pi) + 0.001 * t
    print(f"\n[halo right edge] nx={nx} periodic={periodic}", flush=True)
    xg = to_gpu(xp, dev)
    pre = gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=periodic)
    plan = gmm.PointingPlan(pre, xg)
    ref_mt, err = orc.plan_build(xp, ny, nx, periodic)
    assert err == 0 and np.array_equal(plan.get_plan_mt(), ref_mt)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=periodic)
    pix = orc.Pix.rectangle(ny, nx, periodic)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    mg = gmm.LocalMap(lp, to_gpu(m, dev))
    ref32, _ = orc.map2tod(m, xp, pix)
    ref64, _ = orc.map2tod(m.astype(np.float64), xp, pix, f64=True)
    tg = torch.full((ndet, nt), 7.0, dtype=torch.float32, device=dev)
    gmm.map2tod(tg, mg, xg, plan, debug=True)
    assert l1rel(to_cpu(tg), ref32) < 1e-6 and l1rel(to_cpu(tg), ref64) < 1e-5
    d = rng.standard_normal((ndet, nt)).astype(F32)
    mref32, _ = orc.tod2map(d, xp, pix)
    mref64, _ = orc.tod2map(d.astype(np.float64), xp, pix, f64=True)
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, to_gpu(d, dev), xg, plan, debug=True)
    got = to_cpu(out.arr)
    assert l1rel(got, mref32) < 1e-6 and l1rel(got, mref64) < 1e-5
    # Batched forms use the same ownership masks.
    b = gmm.PlanBatch(lp, [plan], [xg], [tg])
    b.map2tod(mg, debug=True)
    assert l1rel(to_cpu(tg), ref32) < 1e-6
    tg.copy_(to_gpu(d, dev))
    out.arr.zero_()
    b.tod2map(out, debug=True)
    assert l1rel(to_cpu(out.arr), mref32) < 1e-6


# ------------------------------------------------------------------------------------------
# tod2map on TODs far from N(0,1): glitches of 1e4 .. 1e9 times the noise inside one cell, an
# offset TOD (every deposit of one sign), a TOD that is zero except for a short stretch. The
# fixed-point quantum follows the bulk of the samples, so the FP64 oracle bar holds pixel by
# pixel, also next to the glitches.

def _heavy_tail_tods(rng, shape, xp):
    base = rng.standard_normal(shape).astype(F32)
    ndet, nt = shape
    out = {}
    for amp in (1e4, 1e6, 1e9):
        d = base.copy()
        # four glitches of detector 5 within 200 samples (one or two cells), one elsewhere
        for k, s in enumerate((30000, 30050, 30111, 30190)):
            d[5, s] = amp * (1 + 0.25 * k) * (-1) ** k
        d[40, 1234] = -amp
        out[f"glitch{amp:.0e}"] = d
    out["offset"] = (3.0 + 0.1 * base).astype(F32)
    d = np.zeros(shape, F32)
    d[:, 20000:20800] = base[:, 20000:20800]
    out["sparse"] = d
    return out


@pytest.mark.parametrize("kind", ["glitch1e+04", "glitch1e+06", "glitch1e+09", "offset", "sparse"])
def test_tod2map_heavy_tailed_tod(gmm, orc, dev, kind):
    d = synth.make_cfg0(seed=0)
    xp, ny, nx = d["xpointing"], d["ny"], d["nx"]
    print(f"\n[heavy-tailed tod2map] {kind} xpointing {xp.shape}", flush=True)
    rng = np.random.default_rng(88)
    tod = _heavy_tail_tods(rng, xp.shape[1:], xp)[kind]
    xg = to_gpu(xp, dev)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False), xg)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    pix = orc.Pix.rectangle(ny, nx, False)
    ref, err = orc.tod2map(tod.astype(np.float64), xp, pix, f64=True)
    assert err == 0
    # Bulk reference scale: the map of the TOD without its glitches.
    bulk = tod.copy()
    bulk[np.abs(bulk) > 100] = 0.0
    rms = float(np.sqrt(np.mean(orc.tod2map(bulk.astype(np.float64), xp, pix, f64=True)[0] ** 2)))
    # Pixels holding a glitch's deposit: |value| well above the bulk (below 10 x rms the float
    # rounding of a value stays under 1e-6 rms).
    hot = np.abs(ref) > 10 * rms

    def check(got, what):
        e_bulk = float(np.max(np.abs(got - ref)[~hot])) / rms
        e_hot = float(np.max((np.abs(got - ref) / np.abs(ref))[hot])) if hot.any() else 0.0
        print(f"  {what}: max error / map rms (bulk pixels) {e_bulk:.2e}; max relative error (glitch"
              f" pixels, {int(hot.sum())}) {e_hot:.2e}; l1rel {l1rel(got, ref):.2e}", flush=True)
        assert e_bulk < 1e-5
        assert e_hot < 1e-6
        assert l1rel(got, ref) < 1e-6

    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, to_gpu(tod, dev), xg, plan, debug=True)
    check(to_cpu(out.arr), "planned")

    # The batched operator shares the accumulation (one TOD, and the same TOD twice).
    tg = to_gpu(tod, dev)
    b = gmm.PlanBatch(lp, [plan, plan], [xg, xg], [tg, tg])
    out2 = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    b.tod2map(out2, debug=True)
    got2 = to_cpu(out2.arr) / 2
    check(got2, "batched (2 x same TOD) / 2")


# ------------------------------------------------------------------------------------------
# Debug checks (debug=True): device-side invariants of plans and of the planned operators, the
# role of the reference's Debug asserts (include/plan_iterator.hpp:64-141,
# include/gpu_mm_internals.hpp:224-235, src_lib/map2tod.cu:105-109, src_lib/tod2map.cu:134-139).

def test_debug_checks_pass(inst, gmm, dev):
    """Real plans pass every plan invariant, and the instrumented operators walk every entry
    once, evaluate every sample once and stay inside the halo tile; their results equal the
    non-debug kernels' (map2tod bit for bit; tod2map to float rounding: the flush's atomics)."""
    xg, lp = inst["xg"], inst["lp"]
    plan = gmm.PointingPlan(inst["pre"], xg, debug=True)
    plan.check_invariants()
    rng = np.random.default_rng(71 + inst["seed"])
    shape = inst["xp"].shape[1:]
    m = gmm.LocalMap(lp, to_gpu(rng.standard_normal(3 * lp.npix).astype(F32), dev))
    tod = to_gpu(rng.standard_normal(shape).astype(F32), dev)
    ndet = shape[0]
    resp = to_gpu(rng.uniform(0.5, 1.5, (2, ndet)).astype(F32), dev)
    for response in (None, resp):
        t_dbg, t_ref = torch.empty_like(tod), torch.empty_like(tod)
        gmm.map2tod(t_dbg, m, xg, plan, response=response, debug=True)
        gmm.map2tod(t_ref, m, xg, plan, response=response)
        assert torch.equal(t_dbg, t_ref)
        o_dbg = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
        o_ref = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
        gmm.tod2map(o_dbg, tod, xg, plan, response=response, debug=True)
        gmm.tod2map(o_ref, tod, xg, plan, response=response)
        assert l1rel(to_cpu(o_dbg.arr), to_cpu(o_ref.arr)) < 1e-6


def test_debug_checks_partial_last_line(gmm, dev):
    """A TOD whose length is not a multiple of 32: the samples past the end are neither walked nor
    counted, every real sample is evaluated exactly once."""
    d = synth.make_cfg0(seed=0, ndet=4, nt=4096)
    xp = np.ascontiguousarray(d["xpointing"].reshape(3, -1)[:, :4 * 4096 - 13])
    xg = to_gpu(xp, dev)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, d["ny"], d["nx"], periodic_xcoord=False), xg, debug=True)
    lp = gmm.LocalPixelization.make_rectangle(d["ny"], d["nx"], periodic_xcoord=False)
    m = gmm.LocalMap(lp, to_gpu(d["map"].reshape(-1), dev))
    tod = torch.empty(xp.shape[1], dtype=torch.float32, device=dev)
    gmm.map2tod(tod, m, xg, plan, debug=True)
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, tod, xg, plan, debug=True)


def _corrupt_entry(plan, dev, i, new_entry):
    """Overwrites plan entry i (uint64, at byte 8*i of the plan buffer) on the device."""
    plan.buf[8 * i:8 * i + 8].copy_(torch.from_numpy(np.array([new_entry], dtype=np.uint64).view(np.uint8)).to(dev))


def test_debug_checks_detect_corrupted_plans(gmm, dev):
    """Entries changed on the device after construction (indices stay in range, so every kernel
    reads and writes inside its buffers): the plan check and the instrumented operators raise."""
    d = synth.make_cfg0(seed=0, ndet=8, nt=8192)
    xp, ny, nx = d["xpointing"], d["ny"], d["nx"]
    xg = to_gpu(xp, dev)
    pre = gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    m = gmm.LocalMap(lp, to_gpu(d["map"].reshape(-1), dev))
    tod = torch.empty(xp.shape[1:], dtype=torch.float32, device=dev)
    mt = gmm.PointingPlan(pre, xg).get_plan_mt()
    cell = mt & np.uint64((1 << 20) - 1)
    # An entry that repeats the line of the entry before it in the same cell: out of order, one
    # line's samples evaluated twice in that cell and another's not at all.
    i = int(np.nonzero(cell[1:] == cell[:-1])[0][10]) + 1
    plan = gmm.PointingPlan(pre, xg)
    _corrupt_entry(plan, dev, i, int(mt[i - 1]))
    with pytest.raises(RuntimeError, match="plan invariant violated"):
        plan.check_invariants()
    with pytest.raises(RuntimeError, match="operator invariant violated"):
        gmm.map2tod(tod, m, xg, plan, debug=True)
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    with pytest.raises(RuntimeError, match="samples not evaluated exactly once"):
        gmm.tod2map(out, tod, xg, plan, debug=True)
    # An entry moved to another cell: its chunk now spans two cells.
    plan = gmm.PointingPlan(pre, xg)
    j = int(np.nonzero(cell[1:] != cell[:-1])[0][3])
    _corrupt_entry(plan, dev, j, int((mt[j] & ~np.uint64((1 << 20) - 1)) | cell[j + 1]))
    with pytest.raises(RuntimeError, match="spanning cells"):
        plan.check_invariants()
    # The untouched plan passes.
    gmm.PointingPlan(pre, xg).check_invariants()


# ------------------------------------------------------------------------------------------
# The streaming-read probe behind the benchmark's bandwidth ceiling.
def test_stream_read_probe(gmm, dev):
    """Every element is read once: with 0/1 values all partial sums are exact in float32, so
    the probe's sum equals the count of ones whatever the order of its atomic additions. Lengths
    below, at and above one grid's worth; misaligned or ragged inputs raise."""
    g = torch.Generator(device="cpu").manual_seed(7)
    for n in (4, 4 * 1001, 1 << 22, (1 << 22) + 4 * 37):
        print(f"\n[stream_read] n={n}", flush=True)
        src = torch.randint(0, 2, (n,), generator=g).to(torch.float32).to(dev)
        acc = torch.zeros(1, dtype=torch.float32, device=dev)
        gmm.stream_read(src, acc)
        gmm.stream_read(src, acc)          # adds: two reads give twice the count
        assert acc.item() == 2 * int(src.sum().item())
    with pytest.raises(RuntimeError, match="16-byte aligned"):
        gmm.stream_read(src[1:9], acc)
    with pytest.raises(RuntimeError, match="multiple of 4"):
        gmm.stream_read(src[:6], acc)
