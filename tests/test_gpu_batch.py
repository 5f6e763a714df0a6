"""GPU tests of the batched operators (PlanBatch): several TODs of different sizes, each with its
own rotated scan and plan, against one local map in one launch.

Bars: batched map2tod bit-identical to the plain planned map2tod of each TOD (both evaluate every
sample with the same arithmetic, in the block that owns it); batched tod2map within 1e-6
L1-relative of the FP32 oracle summed over the TODs and within 1e-5 of the FP64 oracle (the
bars of the plain operators); HIP-graph replays equal to eager launches.
"""

import numpy as np
import pytest
import torch

from gpu_mm_amd import synth

pytestmark = pytest.mark.gpu

F32 = np.float32


def l1rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.sum(np.abs(a - b)) / (np.sum(np.abs(a)) + np.sum(np.abs(b))))


def to_cpu(t):
    return t.detach().to("cpu").numpy()


# Small uneven batch: 5 TODs, ndet in {8..40}, nt in {2048..12288}, on a 3x384x640 map.
NY, NX = 384, 640


@pytest.fixture(scope="module", params=[False, True], ids=["nonperiodic", "periodic"])
def batch_case(request, gmm, dev):
    periodic = request.param
    shapes = synth.batch_shapes(11, 5, tuple(range(8, 41, 8)), tuple(range(2048, 12289, 64)), NY, NX, (60.0, 200.0))
    print(f"\n[batch case periodic={periodic}] shapes {[(a, b) for a, b, _ in shapes]}", flush=True)
    xps, plans = [], []
    for i, (ndet, nt, scan) in enumerate(shapes):
        xg = synth.batch_xpointing_torch(ndet, nt, scan, seed=100 + i, device=dev)
        pre = gmm.PointingPrePlan(xg, NY, NX, periodic_xcoord=periodic)
        plans.append(gmm.PointingPlan(pre, xg))
        xps.append(xg)
    lp = gmm.LocalPixelization.make_rectangle(NY, NX, periodic_xcoord=periodic)
    rng = np.random.default_rng(5)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    tods = [rng.standard_normal(tuple(x.shape[1:])).astype(F32) for x in xps]
    return dict(shapes=shapes, xps=xps, plans=plans, lp=lp, m=m, tods=tods, periodic=periodic)


CONFIGS = [
    dict(),                                                                    # defaults
    dict(m2t_item_entries=64, m2t_threads=256, m2t_blocks_per_cu=0,           # many small items
         t2m_item_entries=64, t2m_blocks_per_cu=0),
    dict(m2t_item_entries=200, m2t_threads=512, m2t_blocks_per_cu=1,          # persistent, split cells
         t2m_item_entries=300, t2m_blocks_per_cu=2),
    dict(m2t_item_entries=1 << 20, m2t_threads=512, m2t_blocks_per_cu=4,      # one item per cell
         t2m_item_entries=1 << 20, t2m_blocks_per_cu=1),
]


def make_config(gmm, **kw):
    c = gmm.BatchConfig()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def oracle_pix(orc, case):
    return orc.Pix.rectangle(NY, NX, case["periodic"])


@pytest.mark.parametrize("cfg", CONFIGS, ids=["default", "small-items", "persistent-split", "whole-cells"])
def test_batch_map2tod_equals_plain(batch_case, gmm, dev, cfg):
    c = batch_case
    lp = c["lp"]
    mg = gmm.LocalMap(lp, torch.from_numpy(c["m"]).to(dev))
    outs = [torch.full(tuple(x.shape[1:]), 7.0, dtype=torch.float32, device=dev) for x in c["xps"]]
    b = gmm.PlanBatch(lp, c["plans"], c["xps"], outs, config=make_config(gmm, **cfg))
    print(f"  config {b.config} items {b.work_items()}", flush=True)
    for accum in (False, True):
        if accum:
            init = [torch.from_numpy(t).to(dev) for t in c["tods"]]
            for o, t in zip(outs, init):
                o.copy_(t)
        b.map2tod(mg, accum=accum, debug=True)
        for i, (x, p) in enumerate(zip(c["xps"], c["plans"])):
            ref = torch.from_numpy(c["tods"][i]).to(dev) if accum else torch.empty_like(outs[i])
            gmm.map2tod(ref, mg, x, p, accum=accum, debug=True)
            assert torch.equal(outs[i], ref), f"TOD {i} accum={accum}"


def test_batch_map2tod_oracle(batch_case, gmm, orc, dev):
    c = batch_case
    pix = oracle_pix(orc, c)
    mg = gmm.LocalMap(c["lp"], torch.from_numpy(c["m"]).to(dev))
    outs = [torch.zeros(tuple(x.shape[1:]), dtype=torch.float32, device=dev) for x in c["xps"]]
    b = gmm.PlanBatch(c["lp"], c["plans"], c["xps"], outs)
    b.map2tod(mg, debug=True)
    for i, x in enumerate(c["xps"]):
        ref64, err = orc.map2tod(c["m"].astype(np.float64), to_cpu(x), pix, f64=True)
        assert err == 0
        assert l1rel(to_cpu(outs[i]), ref64) < 1e-5


@pytest.mark.parametrize("cfg", CONFIGS, ids=["default", "small-items", "persistent-split", "whole-cells"])
def test_batch_tod2map(batch_case, gmm, orc, dev, cfg):
    c = batch_case
    pix = oracle_pix(orc, c)
    rng = np.random.default_rng(9)
    m0 = rng.standard_normal(3 * c["lp"].npix).astype(F32)
    ref32 = m0.copy()
    ref64 = m0.astype(np.float64)
    for i, x in enumerate(c["xps"]):
        xc = to_cpu(x)
        ref32, err = orc.tod2map(c["tods"][i], xc, pix, lmap=ref32)
        assert err == 0
        ref64, _ = orc.tod2map(c["tods"][i].astype(np.float64), xc, pix, lmap=ref64, f64=True)
    tg = [torch.from_numpy(t).to(dev) for t in c["tods"]]
    b = gmm.PlanBatch(c["lp"], c["plans"], c["xps"], tg, config=make_config(gmm, **cfg))
    mg = gmm.LocalMap(c["lp"], torch.from_numpy(m0).to(dev))
    b.tod2map(mg, debug=True)
    got = to_cpu(mg.arr)
    e32, e64 = l1rel(got, ref32), l1rel(got, ref64)
    print(f"  batch tod2map {b.config}: l1rel vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}")
    assert e32 < 1e-6
    assert e64 < 1e-5
    # Against the plain planned tod2map of each TOD in turn.
    mp = gmm.LocalMap(c["lp"], torch.from_numpy(m0).to(dev))
    for x, p, t in zip(c["xps"], c["plans"], tg):
        gmm.tod2map(mp, t, x, p, debug=True)
    assert l1rel(got, to_cpu(mp.arr)) < 1e-6


def test_batch_adjoint(batch_case, gmm, dev):
    c = batch_case
    rng = np.random.default_rng(13)
    mg = gmm.LocalMap(c["lp"], torch.from_numpy(rng.standard_normal(3 * c["lp"].npix).astype(F32)).to(dev))
    am = [torch.zeros(tuple(x.shape[1:]), dtype=torch.float32, device=dev) for x in c["xps"]]
    gmm.PlanBatch(c["lp"], c["plans"], c["xps"], am).map2tod(mg, debug=True)
    d = [torch.from_numpy(rng.standard_normal(tuple(x.shape[1:])).astype(F32)).to(dev) for x in c["xps"]]
    atd = gmm.LocalMap(c["lp"], torch.zeros(3 * c["lp"].npix, dtype=torch.float32, device=dev))
    gmm.PlanBatch(c["lp"], c["plans"], c["xps"], d).tod2map(atd, debug=True)
    lhs = sum(float(torch.dot(a.double().ravel(), y.double().ravel())) for a, y in zip(am, d))
    rhs = float(torch.dot(mg.arr.double(), atd.arr.double()))
    scale = np.sqrt(sum(float(a.double().norm()) ** 2 for a in am)) * np.sqrt(sum(float(y.double().norm()) ** 2 for y in d))
    assert abs(lhs - rhs) < 1e-6 * scale


def test_batch_graph_replays(batch_case, gmm, dev):
    """Captured launches, replayed several times with eager work and synchronizations between,
    equal the eager launches (map2tod bit for bit; tod2map to 1e-6: its flush adds atomically)."""
    c = batch_case
    lp = c["lp"]
    mg = gmm.LocalMap(lp, torch.from_numpy(c["m"]).to(dev))
    outs = [torch.zeros(tuple(x.shape[1:]), dtype=torch.float32, device=dev) for x in c["xps"]]
    b = gmm.PlanBatch(lp, c["plans"], c["xps"], outs, config=make_config(gmm, **CONFIGS[2]))
    b.map2tod(mg, debug=True)
    eager = [o.clone() for o in outs]
    g = b.capture("map2tod", mg)
    for rep in range(3):
        for o in outs:
            o.fill_(float(rep + 1))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e), f"map2tod replay {rep}"
        b.check_errflags()

    tods = [torch.from_numpy(t).to(dev) for t in c["tods"]]
    bt = gmm.PlanBatch(lp, c["plans"], c["xps"], tods, config=make_config(gmm, **CONFIGS[2]))
    acc = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    bt.tod2map(acc, debug=True)
    eager_map = to_cpu(acc.arr)
    gt = bt.capture("tod2map", acc)
    for rep in range(3):
        acc.arr.zero_()
        torch.cuda.synchronize()
        gt.launch()
        torch.cuda.synchronize()
        assert l1rel(to_cpu(acc.arr), eager_map) < 1e-6, f"tod2map replay {rep}"
        bt.check_errflags()


def test_batch_autotune(batch_case, gmm, dev):
    c = batch_case
    lp = c["lp"]
    mg = gmm.LocalMap(lp, torch.from_numpy(c["m"]).to(dev))
    outs = [torch.zeros(tuple(x.shape[1:]), dtype=torch.float32, device=dev) for x in c["xps"]]
    b = gmm.PlanBatch(lp, c["plans"], c["xps"], outs)
    g = b.capture("map2tod", mg)
    table = b.autotune(mg, reps=1)
    ops = [t[0] for t in table]
    assert ops.count("map2tod") == 50 and ops.count("tod2map") == 15
    best_m2t = min((t for t in table if t[0] == "map2tod"), key=lambda t: t[4])
    cfg = b.config
    assert (cfg.m2t_item_entries, cfg.m2t_threads, cfg.m2t_blocks_per_cu) == tuple(best_m2t[1:4])
    # The autotuned batch still gives the plain results; the old graph is refused.
    with pytest.raises(RuntimeError, match="reconfigured"):
        g.launch()
    b.map2tod(mg, debug=True)
    for x, p, o in zip(c["xps"], c["plans"], outs):
        ref = torch.empty_like(o)
        gmm.map2tod(ref, mg, x, p, debug=True)
        assert torch.equal(o, ref)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16], ids=["f16", "bf16"])
def test_batch_tod16(batch_case, gmm, dev, dtype):
    c = batch_case
    lp = c["lp"]
    mg = gmm.LocalMap(lp, torch.from_numpy(c["m"]).to(dev))
    enc = [gmm.encode_tod16(torch.from_numpy(t).to(dev), dtype) for t in c["tods"]]
    tods16 = [e[0] for e in enc]
    scales = [e[1] * 8 for e in enc]      # room for map2tod outputs larger than the random TOD
    b = gmm.PlanBatch(lp, c["plans"], c["xps"], tods16, scales=scales, config=make_config(gmm, **CONFIGS[2]))
    # tod2map from the 16-bit TODs against the plain 16-bit tod2map of each TOD.
    src = [t.clone() for t in tods16]
    mb = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    b.tod2map(mb, debug=True)
    mp = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    for x, p, t, s in zip(c["xps"], c["plans"], src, scales):
        gmm.tod2map(mp, t, x, p, tod_scale=s, debug=True)
    assert l1rel(to_cpu(mb.arr), to_cpu(mp.arr)) < 1e-6
    # map2tod into the 16-bit TODs: identical to the plain 16-bit map2tod.
    b.map2tod(mg, debug=True)
    for x, p, t, s in zip(c["xps"], c["plans"], tods16, scales):
        ref = torch.empty_like(t)
        gmm.map2tod(ref, mg, x, p, tod_scale=s, debug=True)
        assert torch.equal(t, ref)


def test_batch_partial_pixelization(gmm, orc, dev):
    """A batch whose TODs leave the local map raises unless partial_pixelization is set, and
    then matches the oracle with the missing cells read as zero / dropped."""
    ny, nx = 256, 384
    shapes = synth.batch_shapes(21, 3, (8, 16), tuple(range(2048, 4097, 64)), ny, nx, (60.0, 120.0))
    xps, plans = [], []
    for i, (ndet, nt, scan) in enumerate(shapes):
        xg = synth.batch_xpointing_torch(ndet, nt, scan, seed=200 + i, device=dev)
        plans.append(gmm.PointingPlan(gmm.PointingPrePlan(xg, ny, nx, periodic_xcoord=False), xg))
        xps.append(xg)
    cells = [(cy, cx) for cy in range(ny // 64) for cx in range(nx // 64) if (cy * 7 + cx) % 4 != 0]
    lp = gmm.LocalPixelization.make_cells(ny, nx, cells, periodic_xcoord=False)
    pix = orc.Pix(ny, nx, lp.cell_offsets_cpu, 64, 4096, False)
    rng = np.random.default_rng(3)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    mg = gmm.LocalMap(lp, torch.from_numpy(m).to(dev))
    outs = [torch.zeros(tuple(x.shape[1:]), dtype=torch.float32, device=dev) for x in xps]
    b = gmm.PlanBatch(lp, plans, xps, outs)
    with pytest.raises(RuntimeError, match="not in LocalPixelization"):
        b.map2tod(mg, debug=True)
    b.map2tod(mg, partial_pixelization=True, debug=True)
    for x, o in zip(xps, outs):
        ref, err = orc.map2tod(m, to_cpu(x), pix, partial=True)
        assert err == 0
        assert l1rel(to_cpu(o), ref) < 1e-6
    d = [torch.from_numpy(rng.standard_normal(tuple(x.shape[1:])).astype(F32)).to(dev) for x in xps]
    bd = gmm.PlanBatch(lp, plans, xps, d)
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    with pytest.raises(RuntimeError, match="not in LocalPixelization"):
        bd.tod2map(out, debug=True)
    out.arr.zero_()
    bd.tod2map(out, partial_pixelization=True, debug=True)
    ref = np.zeros(3 * lp.npix, F32)
    for x, t in zip(xps, d):
        ref, err = orc.tod2map(to_cpu(t), to_cpu(x), pix, lmap=ref, partial=True)
        assert err == 0
    assert l1rel(to_cpu(out.arr), ref) < 1e-6


def test_batch_argument_errors(batch_case, gmm, dev):
    c = batch_case
    lp = c["lp"]
    outs = [torch.zeros(tuple(x.shape[1:]), dtype=torch.float32, device=dev) for x in c["xps"]]
    with pytest.raises(RuntimeError, match="equal length"):
        gmm.PlanBatch(lp, c["plans"], c["xps"][:-1], outs)
    with pytest.raises(RuntimeError, match="samples but its xpointing"):
        gmm.PlanBatch(lp, c["plans"], c["xps"], outs[1:] + outs[:1])
    with pytest.raises(RuntimeError, match="plan was built for nsamp"):
        gmm.PlanBatch(lp, c["plans"][1:] + c["plans"][:1], c["xps"], outs)
    other = gmm.LocalPixelization.make_rectangle(NY, NX + 64, periodic_xcoord=c["periodic"])
    with pytest.raises(RuntimeError, match="disagree"):
        gmm.PlanBatch(other, c["plans"], c["xps"], outs)
    b = gmm.PlanBatch(lp, c["plans"], c["xps"], outs)
    with pytest.raises(RuntimeError, match="size 3\\*npix"):
        b.map2tod(torch.zeros(10, dtype=torch.float32, device=dev))
    with pytest.raises(RuntimeError, match="item_entries"):
        b.configure(make_config(gmm, m2t_item_entries=16))
    with pytest.raises(RuntimeError, match="m2t_threads"):
        b.configure(make_config(gmm, m2t_threads=1024))


@pytest.mark.parametrize("cfg", CONFIGS, ids=["default", "small-items", "persistent-split", "whole-cells"])
def test_batch_response(batch_case, gmm, orc, dev, cfg):
    """Per-detector gains (set_response): batched map2tod bit-identical to the plain response
    map2tod of each TOD; batched tod2map within the response bar of the FP64 response oracle
    summed over the TODs; set_response(None) restores the plain operators."""
    c = batch_case
    lp = c["lp"]
    rng = np.random.default_rng(23)
    resps = [rng.uniform(0.5, 1.5, (2, x.shape[1])).astype(F32) for x in c["xps"]]
    rg = [torch.from_numpy(r).to(dev) for r in resps]
    mg = gmm.LocalMap(lp, torch.from_numpy(c["m"]).to(dev))
    outs = [torch.full(tuple(x.shape[1:]), 7.0, dtype=torch.float32, device=dev) for x in c["xps"]]
    b = gmm.PlanBatch(lp, c["plans"], c["xps"], outs, config=make_config(gmm, **cfg))
    b.set_response(rg)
    b.map2tod(mg, debug=True)
    for x, p, r, o in zip(c["xps"], c["plans"], rg, outs):
        ref = torch.empty_like(o)
        gmm.map2tod(ref, mg, x, p, response=r, debug=True)
        assert torch.equal(o, ref)
    for o, t in zip(outs, c["tods"]):
        o.copy_(torch.from_numpy(t).to(dev))
    out = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    b.tod2map(out, debug=True)
    pix = oracle_pix(orc, c)
    mref = np.zeros(3 * lp.npix)
    for x, t, r in zip(c["xps"], c["tods"], resps):
        mr, err = orc.response_tod2map(t, x.cpu().numpy(), r, pix)
        assert err == 0
        mref += mr
    assert l1rel(to_cpu(out.arr), mref) < 1e-5       # the response oracle is FP64
    # Without gains again: the plain batched map2tod.
    b.set_response(None)
    b.map2tod(mg, debug=True)
    for x, p, o in zip(c["xps"], c["plans"], outs):
        ref = torch.empty_like(o)
        gmm.map2tod(ref, mg, x, p, debug=True)
        assert torch.equal(o, ref)


def test_batch_response_errors(batch_case, gmm, dev):
    c = batch_case
    lp = c["lp"]
    outs = [torch.zeros(tuple(x.shape[1:]), dtype=torch.float32, device=dev) for x in c["xps"]]
    b = gmm.PlanBatch(lp, c["plans"], c["xps"], outs)
    good = [torch.ones((2, x.shape[1]), dtype=torch.float32, device=dev) for x in c["xps"]]
    with pytest.raises(RuntimeError, match="one response array per TOD"):
        b.set_response(good[:-1])
    bad = list(good)
    bad[1] = torch.ones((2, c["xps"][1].shape[1] * c["xps"][1].shape[2] + 1), dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError, match="does not divide nsamp"):
        b.set_response(bad)
    with pytest.raises(RuntimeError, match=r"shape \(2, ndet\)"):
        b.set_response([torch.ones((3, x.shape[1]), dtype=torch.float32, device=dev) for x in c["xps"]])
    t16 = []
    sc = []
    for x in c["xps"]:
        t, s = gmm.encode_tod16(torch.zeros(tuple(x.shape[1:]), dtype=torch.float32, device=dev))
        t16.append(t)
        sc.append(s)
    b16 = gmm.PlanBatch(lp, c["plans"], c["xps"], t16, scales=sc)
    with pytest.raises(RuntimeError, match="float32 TODs only"):
        b16.set_response(good)
