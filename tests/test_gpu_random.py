"""Randomized GPU parity on small instances: random map sizes, scans back and forth at 0.01-3
pixels per sample in any direction (drifting and wrapping around in periodic x), detector
offsets, angles that are constant, slowly or quickly turning or random per sample, and TOD lengths
that are not multiples of 32 or of 4. The planner (preplan counts, plan bit for bit), the planned
operators (FP32 and FP64 oracle bars, device-side invariant checks of the debug mode on; also with
per-detector gains and FP16/BF16 TODs), the plan-free operators and the batched ones against the
oracle, as in tests/test_gpu_ops.py."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

F32 = np.float32
F64_BAR = 1e-5


def l1rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.sum(np.abs(a - b)) / (np.sum(np.abs(a)) + np.sum(np.abs(b))))


def random_scan(rng, ndet, nt, ny, nx, periodic):
    """xpointing (3, ndet, nt) float32: per detector a scan back and forth at 0.01-3 pixels per
    sample in a random direction (plus a wrapped drift in periodic x), a small offset, and
    angles that are constant, slowly or quickly turning, or random per sample."""
    t = np.arange(nt, dtype=np.float64)
    amp = 0.4 * min(ny, nx)
    xp = np.empty((3, ndet, nt), F32)
    for d in range(ndet):
        v = float(np.exp(rng.uniform(np.log(0.01), np.log(3.0))))     # pixels per sample
        phi = rng.uniform(0.0, 2.0 * np.pi)
        u = (v * t + rng.uniform(0.0, 2.0 * amp)) % (2.0 * amp)
        u = np.where(u < amp, u, 2.0 * amp - u) - 0.5 * amp            # back and forth, |u| <= amp/2
        y = 0.5 * ny + u * np.sin(phi) + rng.uniform(-3.0, 3.0)
        x = 0.5 * nx + u * np.cos(phi) + rng.uniform(-3.0, 3.0)
        if periodic:
            x = x + rng.uniform(-0.5, 0.5) * t                          # drift, wrapped below
            x = np.mod(x + 0.5 * nx, 2.0 * nx) - 0.5 * nx
        kind = int(rng.integers(4))
        a0 = rng.uniform(0.0, np.pi)
        if kind == 3:
            a = rng.uniform(-10.0, 10.0, nt)
        else:
            a = a0 + (0.0, 1e-6, 2e-3)[kind] * t
        xp[0, d], xp[1, d], xp[2, d] = y, x, a
    return xp


def make_random_instance(seed):
    """(xpointing (3, ndet, nt) float32, ny, nx, periodic) from the seed."""
    rng = np.random.default_rng(5000 + seed)
    periodic = bool(rng.integers(2))
    ny, nx = int(rng.integers(70, 400)), int(rng.integers(70, 500))
    ndet, nt = int(rng.integers(1, 7)), int(rng.integers(100, 3000))
    return random_scan(rng, ndet, nt, ny, nx, periodic), ny, nx, periodic


def make_random_batch(seed):
    """(list of xpointings with nsamp % 4 == 0, ny, nx, periodic): 3-5 TODs on one map."""
    rng = np.random.default_rng(9000 + seed)
    periodic = bool(rng.integers(2))
    ny, nx = int(rng.integers(70, 400)), int(rng.integers(70, 500))
    xps = [random_scan(rng, int(rng.integers(1, 6)), 4 * int(rng.integers(25, 750)), ny, nx, periodic)
           for _ in range(int(rng.integers(3, 6)))]
    return xps, ny, nx, periodic


@pytest.mark.parametrize("seed", range(48))
def test_random_instance_parity(gmm, orc, dev, seed):
    xp, ny, nx, periodic = make_random_instance(seed)
    ndet, nt = xp.shape[1], xp.shape[2]
    print(f"\n[random instance seed={seed}] ndet={ndet} nt={nt} ny={ny} nx={nx} periodic={periodic}", flush=True)
    xg = torch.from_numpy(xp).to(dev)
    pre = gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=periodic)
    plan = gmm.PointingPlan(pre, xg, debug=True)
    cum, nmt, err = orc.plan_count(xp, ny, nx, periodic)
    assert err == 0 and pre.plan_nmt == nmt
    assert np.array_equal(pre.get_nmt_cumsum(), cum)
    ref_mt, err = orc.plan_build(xp, ny, nx, periodic)
    assert err == 0 and np.array_equal(plan.get_plan_mt(), ref_mt)

    rng = np.random.default_rng(seed)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=periodic)
    pix = orc.Pix.rectangle(ny, nx, periodic)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    mg = gmm.LocalMap(lp, torch.from_numpy(m).to(dev))
    ref32, _ = orc.map2tod(m, xp, pix)
    ref64, _ = orc.map2tod(m.astype(np.float64), xp, pix, f64=True)
    tg = torch.full((ndet, nt), 3.0, dtype=torch.float32, device=dev)
    gmm.map2tod(tg, mg, xg, plan, debug=True)
    got = tg.cpu().numpy()
    assert l1rel(got, ref32) < 1e-6 and l1rel(got, ref64) < F64_BAR

    d = rng.standard_normal((ndet, nt)).astype(F32)
    mref32, _ = orc.tod2map(d, xp, pix)
    mref64, _ = orc.tod2map(d.astype(np.float64), xp, pix, f64=True)
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, torch.from_numpy(d).to(dev), xg, plan, debug=True)
    got = out.arr.cpu().numpy()
    assert l1rel(got, mref32) < 1e-6 and l1rel(got, mref64) < F64_BAR

    # The plan-free kernels on the same inputs.
    tg.fill_(3.0)
    gmm.map2tod(tg, mg, xg, None)
    got = tg.cpu().numpy()
    assert l1rel(got, ref32) < 1e-6 and l1rel(got, ref64) < F64_BAR
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, torch.from_numpy(d).to(dev), xg, None)
    got = out.arr.cpu().numpy()
    assert l1rel(got, mref32) < 1e-6 and l1rel(got, mref64) < F64_BAR

    # Per-detector gains (rows of nt samples) against the FP64 response oracle.
    resp = rng.uniform(0.5, 1.5, (2, ndet)).astype(F32)
    rg = torch.from_numpy(resp).to(dev)
    rref, _ = orc.response_map2tod(m, xp, resp, pix)
    gmm.map2tod(tg, mg, xg, plan, response=rg, debug=True)
    assert l1rel(tg.cpu().numpy(), rref) < F64_BAR
    mr, _ = orc.response_tod2map(d, xp, resp, pix)
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, torch.from_numpy(d).to(dev), xg, plan, response=rg, debug=True)
    assert l1rel(out.arr.cpu().numpy(), mr) < F64_BAR


# Work-item shapes of the batched operators: the defaults, many small items, and a persistent
# grid with cells split across items (as in tests/test_gpu_batch.py).
BATCH_CONFIGS = [
    dict(),
    dict(m2t_item_entries=64, m2t_threads=256, m2t_blocks_per_cu=0, t2m_item_entries=64, t2m_blocks_per_cu=0),
    dict(m2t_item_entries=200, m2t_threads=512, m2t_blocks_per_cu=1, t2m_item_entries=300, t2m_blocks_per_cu=2),
]


@pytest.mark.parametrize("seed", range(12))
def test_random_batch_parity(gmm, orc, dev, seed):
    """Batched operators over 3-5 random TODs of uneven shapes on one map: map2tod of each TOD
    and tod2map summed over the TODs against the oracle, for each work-item shape."""
    xps, ny, nx, periodic = make_random_batch(seed)
    print(f"\n[random batch seed={seed}] shapes={[x.shape[1:] for x in xps]} ny={ny} nx={nx} "
          f"periodic={periodic}", flush=True)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=periodic)
    pix = orc.Pix.rectangle(ny, nx, periodic)
    xgs = [torch.from_numpy(x).to(dev) for x in xps]
    plans = [gmm.PointingPlan(gmm.PointingPrePlan(x, ny, nx, periodic_xcoord=periodic), x) for x in xgs]
    rng = np.random.default_rng(seed)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    mg = gmm.LocalMap(lp, torch.from_numpy(m).to(dev))
    ref_tods = [orc.map2tod(m, x, pix)[0] for x in xps]
    tods = [rng.standard_normal(x.shape[1:]).astype(F32) for x in xps]
    ref32 = np.zeros(3 * lp.npix, F32)
    ref64 = np.zeros(3 * lp.npix, np.float64)
    for x, d in zip(xps, tods):
        ref32, _ = orc.tod2map(d, x, pix, lmap=ref32)
        ref64, _ = orc.tod2map(d.astype(np.float64), x, pix, lmap=ref64, f64=True)
    for cfg in BATCH_CONFIGS:
        conf = gmm.BatchConfig()
        for k, v in cfg.items():
            setattr(conf, k, v)
        outs = [torch.full(x.shape[1:], 3.0, dtype=torch.float32, device=dev) for x in xps]
        b = gmm.PlanBatch(lp, plans, xgs, outs, config=conf)
        b.map2tod(mg, debug=True)
        for o, r in zip(outs, ref_tods):
            assert l1rel(o.cpu().numpy(), r) < 1e-6
        for o, d in zip(outs, tods):
            o.copy_(torch.from_numpy(d))
        out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
        b.tod2map(out, debug=True)
        got = out.arr.cpu().numpy()
        assert l1rel(got, ref32) < 1e-6 and l1rel(got, ref64) < F64_BAR, cfg


def make_random_instance16(seed):
    """As make_random_instance, with nt a multiple of 4 (16-bit TODs use 4-sample accesses)."""
    rng = np.random.default_rng(7000 + seed)
    periodic = bool(rng.integers(2))
    ny, nx = int(rng.integers(70, 400)), int(rng.integers(70, 500))
    ndet, nt = int(rng.integers(1, 7)), 4 * int(rng.integers(25, 750))
    return random_scan(rng, ndet, nt, ny, nx, periodic), ny, nx, periodic


# 16-bit TOD types and the relative rounding of one value (half an ulp at 1).
TOD16 = [(torch.float16, 2.0 ** -11), (torch.bfloat16, 2.0 ** -8)]


@pytest.mark.parametrize("dtype,eps", TOD16, ids=["f16", "bf16"])
@pytest.mark.parametrize("seed", range(12))
def test_random_tod16_parity(gmm, orc, dev, seed, dtype, eps):
    """16-bit TODs on random instances: tod2map of the encoded TOD equals the FP64 oracle's on the
    decoded samples (1e-5), and map2tod's 16-bit output is the FP64 oracle's to the type's
    rounding."""
    xp, ny, nx, periodic = make_random_instance16(seed)
    ndet, nt = xp.shape[1], xp.shape[2]
    print(f"\n[random tod16 seed={seed} {dtype}] ndet={ndet} nt={nt} ny={ny} nx={nx} periodic={periodic}",
          flush=True)
    xg = torch.from_numpy(xp).to(dev)
    plan = gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=periodic), xg)
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=periodic)
    pix = orc.Pix.rectangle(ny, nx, periodic)
    rng = np.random.default_rng(100 + seed)
    tod = rng.standard_normal((ndet, nt)).astype(F32)
    t16, sc = gmm.encode_tod16(torch.from_numpy(tod).to(dev), dtype)
    dec = (t16.float() * sc[:, None]).cpu().numpy()
    ref, err = orc.tod2map(dec.astype(np.float64), xp, pix, f64=True)
    assert err == 0
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(out, t16, xg, plan, tod_scale=sc, debug=True)
    assert l1rel(out.arr.cpu().numpy(), ref) < F64_BAR

    # Output scales with headroom for |T + cos Q + sin U| of an N(0,1) map.
    osc = sc * 8
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    ref, _ = orc.map2tod(m.astype(np.float64), xp, pix, f64=True)
    o16 = torch.empty_like(t16)
    gmm.map2tod(o16, gmm.LocalMap(lp, torch.from_numpy(m).to(dev)), xg, plan, tod_scale=osc, debug=True)
    got = (o16.float() * osc[:, None]).cpu().numpy()
    assert np.all(np.isfinite(got))
    assert l1rel(got, ref) < eps
