"""GPU tests of the differentiable operators and the preconditioned CG map solver.

Gradients of the white-noise loss 1/2 sum_i ||N_i^-1/2 (A_i m - d_i)||^2 are checked against
FP64 central differences of the loss evaluated with the FP64 oracle (exact for a quadratic up to
rounding) along random directions supported on a 3x64x128 subregion, and against the FP64
oracle's transpose A^T N^-1 (A m - d). The 3x3 block preconditioner is checked against a
brute-force per-sample accumulation on the host.
"""

import numpy as np
import pytest
import torch

from gpu_mm_amd import synth

pytestmark = pytest.mark.gpu

F32 = np.float32
NY, NX, NDET, NT = 256, 512, 16, 8192


def l1rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.sum(np.abs(a - b)) / (np.sum(np.abs(a)) + np.sum(np.abs(b))))


@pytest.fixture(scope="module")
def problem(gmm, dev):
    from gpu_mm_amd import mapmaker
    scans = synth.crossing_scans(NY, NX, NT, (0.0, 45.0, 90.0, 135.0), 60.0)
    xps, plans, ninv, tods = [], [], [], []
    rng = np.random.default_rng(4)
    lp = gmm.LocalPixelization.make_rectangle(NY, NX, periodic_xcoord=False)
    m_true = torch.from_numpy(rng.standard_normal(3 * lp.npix).astype(F32)).to(dev)
    for i, sc in enumerate(scans):
        x = synth.batch_xpointing_torch(NDET, NT, sc, seed=50 + i, device=dev)
        p = gmm.PointingPlan(gmm.PointingPrePlan(x, NY, NX, periodic_xcoord=False), x)
        var = rng.uniform(0.5, 2.0, NDET)
        w = torch.from_numpy((1.0 / var).astype(F32)).to(dev)
        t = torch.empty((NDET, NT), dtype=torch.float32, device=dev)
        gmm.map2tod(t, gmm.LocalMap(lp, m_true), x, p)
        noise = rng.standard_normal((NDET, NT)) * np.sqrt(var)[:, None]
        t += torch.from_numpy(noise.astype(F32)).to(dev)
        xps.append(x)
        plans.append(p)
        ninv.append(w)
        tods.append(t)
    wp = mapmaker.WeightedPointing(lp, plans, xps, ninv)
    return dict(lp=lp, xps=xps, plans=plans, ninv=ninv, tods=tods, wp=wp, m_true=m_true, mm=mapmaker)


def oracle_loss(orc, pr, m64):
    pix = orc.Pix.rectangle(NY, NX, False)
    total = 0.0
    for x, w, d in zip(pr["xps"], pr["ninv"], pr["tods"]):
        am, err = orc.map2tod(m64, x.cpu().numpy(), pix, f64=True)
        assert err == 0
        r = am.reshape(NDET, NT) - d.cpu().numpy().astype(np.float64)
        total += 0.5 * float(np.sum(w.cpu().numpy().astype(np.float64)[:, None] * r * r))
    return total


def test_autograd_ops_are_the_kernels(problem, gmm, dev):
    pr = problem
    mm, lp, x, p = pr["mm"], pr["lp"], pr["xps"][0], pr["plans"][0]
    rng = np.random.default_rng(1)
    m = torch.from_numpy(rng.standard_normal(3 * lp.npix).astype(F32)).to(dev).requires_grad_(True)
    t = mm.map2tod_ad(m, x, p, lp)
    ref = torch.empty_like(t)
    gmm.map2tod(ref, gmm.LocalMap(lp, m.detach()), x, p)
    assert torch.equal(t, ref)
    g = torch.from_numpy(rng.standard_normal((NDET, NT)).astype(F32)).to(dev)
    (gm,) = torch.autograd.grad(t, m, grad_outputs=g)
    refm = gmm.LocalMap(lp, torch.zeros(3 * lp.npix, dtype=torch.float32, device=dev))
    gmm.tod2map(refm, g, x, p)
    # tod2map adds cell-boundary rows into the map atomically, so two calls agree to rounding.
    assert l1rel(gm.cpu().numpy(), refm.arr.cpu().numpy()) < 1e-6
    # tod2map_ad: forward A^T t, backward A g.
    tt = g.clone().requires_grad_(True)
    mt = mm.tod2map_ad(tt, x, p, lp)
    assert l1rel(mt.detach().cpu().numpy(), refm.arr.cpu().numpy()) < 1e-6
    (gt,) = torch.autograd.grad(mt, tt, grad_outputs=m.detach())
    assert torch.equal(gt, ref)
    # Dot-product test through the autograd functions.
    lhs = float(torch.dot(t.detach().double().ravel(), g.double().ravel()))
    rhs = float(torch.dot(m.detach().double(), mt.detach().double()))
    assert abs(lhs - rhs) < 1e-6 * float(t.detach().double().norm() * g.double().norm())


def test_gradient_finite_differences_fp64(problem, orc, dev):
    pr = problem
    wp, lp = pr["wp"], pr["lp"]
    rng = np.random.default_rng(2)
    m0 = rng.standard_normal(3 * lp.npix).astype(F32)
    m = torch.from_numpy(m0).to(dev).requires_grad_(True)
    loss = wp.loss(m, pr["tods"])
    loss.backward()
    g = m.grad.double().cpu().numpy()
    # FP64 loss from the oracle agrees with the GPU loss.
    l64 = oracle_loss(orc, pr, m0.astype(np.float64))
    assert abs(float(loss) - l64) < 1e-5 * l64
    # Central differences along directions supported on a 3x64x128 subregion at the map centre.
    y0, x0 = NY // 2 - 32, NX // 2 - 64
    mask = np.zeros((3, NY, NX), bool)
    mask[:, y0:y0 + 64, x0:x0 + 128] = True
    mask = mask.reshape(-1)
    for k in range(3):
        v = np.where(mask, rng.standard_normal(mask.size), 0.0)
        eps = 0.05
        lp_, lm_ = oracle_loss(orc, pr, m0 + eps * v), oracle_loss(orc, pr, m0 - eps * v)
        fd = (lp_ - lm_) / (2 * eps)
        an = float(np.dot(g, v))
        scale = np.linalg.norm(g[mask]) * np.linalg.norm(v)
        print(f"  direction {k}: finite difference {fd:.8e}, autograd {an:.8e}, rel {abs(fd - an) / scale:.2e}")
        assert abs(fd - an) < 1e-5 * scale


def test_gradient_matches_oracle_transpose(problem, orc, dev):
    pr = problem
    wp, lp = pr["wp"], pr["lp"]
    rng = np.random.default_rng(3)
    m0 = rng.standard_normal(3 * lp.npix).astype(F32)
    m = torch.from_numpy(m0).to(dev).requires_grad_(True)
    wp.loss(m, pr["tods"]).backward()
    pix = orc.Pix.rectangle(NY, NX, False)
    ref = np.zeros(3 * lp.npix)
    for x, w, d in zip(pr["xps"], pr["ninv"], pr["tods"]):
        xh = x.cpu().numpy()
        am, _ = orc.map2tod(m0.astype(np.float64), xh, pix, f64=True)
        r = (am.reshape(NDET, NT) - d.cpu().numpy().astype(np.float64)) * w.cpu().numpy().astype(np.float64)[:, None]
        ref, err = orc.tod2map(r, xh, pix, lmap=ref, f64=True)
        assert err == 0
    e = l1rel(m.grad.cpu().numpy(), ref)
    print(f"  gradient vs FP64 oracle A^T N^-1 (A m - d): l1rel {e:.2e}")
    assert e < 1e-5
    # The normal operator and right-hand side agree with the gradient: A^T N^-1 A m - rhs.
    nm = wp.apply_normal(m.detach()) - wp.rhs(pr["tods"])
    assert l1rel(nm.cpu().numpy(), ref) < 1e-5


def test_block_preconditioner_brute_force(gmm, orc, dev):
    """M_p = sum_s sum_k w_k ninv [1,c,s]^T [1,c,s], accumulated sample by sample on the host."""
    from gpu_mm_amd import mapmaker
    ny, nx, ndet, nt = 128, 192, 2, 1024
    lp = gmm.LocalPixelization.make_rectangle(ny, nx, periodic_xcoord=False)
    sc = synth.crossing_scans(ny, nx, nt, (30.0,), 40.0)[0]
    x = synth.batch_xpointing_torch(ndet, nt, sc, seed=9, device=dev)
    x[2] = torch.from_numpy(np.random.default_rng(0).uniform(0, 7, (ndet, nt)).astype(F32)).to(dev)
    p = gmm.PointingPlan(gmm.PointingPrePlan(x, ny, nx, periodic_xcoord=False), x)
    ninv = torch.tensor([0.7, 1.9], dtype=torch.float32, device=dev)
    wp = mapmaker.WeightedPointing(lp, [p], [x], [ninv])
    pc = mapmaker.BlockPreconditioner(wp)
    xh = x.cpu().numpy().reshape(3, -1)
    wh = np.repeat(ninv.cpu().numpy().astype(np.float64), nt)
    ref = np.zeros((ny, nx, 3, 3))
    for s in range(xh.shape[1]):
        err, idx, d = orc.locate(float(xh[0, s]), float(xh[1, s]), ny, nx, False, f64=True)
        assert err == 0
        iy0, iy1, ix0, ix1 = (int(v) for v in idx)
        dy, dx = float(d[0]), float(d[1])
        a = 2.0 * float(xh[2, s])
        u = np.array([1.0, np.cos(a), np.sin(a)])
        o = wh[s] * np.outer(u, u)
        for (yy, xx, wk) in ((iy0, ix0, (1 - dy) * (1 - dx)), (iy0, ix1, (1 - dy) * dx),
                             (iy1, ix0, dy * (1 - dx)), (iy1, ix1, dy * dx)):
            ref[yy, xx] += wk * o
    # The preconditioner stores pixels cell by cell; place them by their T offsets, which are
    # y * nx + x in a rectangle pixelization of whole cells.
    got = np.zeros((ny * nx, 3, 3))
    got[pc.idx[0].cpu().numpy()] = pc.m.cpu().numpy()
    got = got.reshape(ny, nx, 3, 3)
    scale = np.abs(ref).max()
    print(f"  preconditioner blocks: max abs error {np.abs(got - ref).max() / scale:.2e} of the largest entry")
    assert np.abs(got - ref).max() < 1e-5 * scale


def test_pcg_converges(problem, dev):
    pr = problem
    mm, wp = pr["mm"], pr["wp"]
    b = wp.rhs(pr["tods"])
    pc = mm.BlockPreconditioner(wp)
    print(f"  preconditioner: {pc.nfull} pixels with the full 3x3 inverse, {pc.ndiag} with the diagonal one")
    x, hist = mm.pcg(wp.apply_normal, b, precond=pc, niter=50)
    _, hist_cg = mm.pcg(wp.apply_normal, b, precond=None, niter=50)
    print("  PCG relative residual: " + " ".join(f"{h:.1e}" for h in hist[::5]))
    print("  CG relative residual:  " + " ".join(f"{h:.1e}" for h in hist_cg[::5]))
    assert hist[-1] < 0.5 * hist_cg[-1]
    assert hist[-1] < 0.02
    # The solution satisfies the normal equations to the reported residual.
    r = b - wp.apply_normal(x)
    rel = float(r.double().norm() / b.double().norm())
    print(f"  recomputed residual {rel:.2e}")
    assert abs(rel - hist[-1]) < 0.1 * hist[-1] + 1e-6


def test_pcg_recovers_smooth_map(gmm, dev):
    """Noiseless data from a Gaussian random field with k^-2 power: the solve recovers the map in
    the interior of the well-covered region (pixels at least 4 pixels inside it; at the edges of
    the coverage the bilinear weights leave nearly unconstrained modes that CG resolves slowly)."""
    from gpu_mm_amd import mapmaker as mm
    ndet, nt = 64, 16384
    lp = gmm.LocalPixelization.make_rectangle(NY, NX, periodic_xcoord=False)
    rng = np.random.default_rng(4)
    xps, plans, ninv = [], [], []
    for i, sc in enumerate(synth.crossing_scans(NY, NX, nt, (0.0, 45.0, 90.0, 135.0), 100.0)):
        x = synth.batch_xpointing_torch(ndet, nt, sc, seed=50 + i, device=dev)
        xps.append(x)
        plans.append(gmm.PointingPlan(gmm.PointingPrePlan(x, NY, NX, periodic_xcoord=False), x))
        ninv.append(torch.from_numpy((1.0 / rng.uniform(0.5, 2.0, ndet)).astype(F32)).to(dev))
    wp = mm.WeightedPointing(lp, plans, xps, ninv)
    m_grf = synth.grf_map_torch(NY, NX, seed=8, device=dev).reshape(-1).contiguous()
    d = []
    for x, p in zip(xps, plans):
        t = torch.empty((ndet, nt), dtype=torch.float32, device=dev)
        gmm.map2tod(t, gmm.LocalMap(lp, m_grf), x, p)
        d.append(t)
    b = wp.rhs(d)
    pc = mm.BlockPreconditioner(wp)
    good = ((pc.hits > 20.0) & (torch.linalg.eigvalsh(pc.m)[:, 0] > 0.05 * pc.hits)).cpu().numpy()
    mask = np.zeros(NY * NX, bool)
    mask[pc.idx[0].cpu().numpy()] = good        # T offsets are y * NX + x in this pixelization
    mask = mask.reshape(NY, NX)
    for _ in range(4):
        inner = mask.copy()
        inner[1:-1, 1:-1] &= mask[:-2, 1:-1] & mask[2:, 1:-1] & mask[1:-1, :-2] & mask[1:-1, 2:]
        inner[0, :] = inner[-1, :] = False
        inner[:, 0] = inner[:, -1] = False
        mask = inner
    sel = torch.from_numpy(np.flatnonzero(mask)).to(dev)
    sel = torch.stack([sel + k * lp.polstride for k in range(3)])
    errs = []

    def cb(k, xk, res):
        if (k + 1) % 25 == 0:
            errs.append(float((xk[sel] - m_grf[sel]).double().norm() / m_grf[sel].double().norm()))

    x, hist = mm.pcg(wp.apply_normal, b, precond=pc, niter=100, callback=cb)
    print(f"  {int(mask.sum())} interior pixels; relative map error after 25/50/75/100 iterations: "
          + " ".join(f"{e:.2e}" for e in errs) + f"; residual {hist[-1]:.2e}")
    assert int(mask.sum()) > 10000
    assert errs[-1] < errs[0]
    assert errs[-1] < 5e-3


def test_apply_normal_defers_errors(problem, gmm, dev):
    """apply_normal() does not block on the error flags; check_errflags() raises what its calls
    reported (here: a local map without the cells of the lower half of the sky)."""
    pr = problem
    cells = [(cy, cx) for cy in range(NY // 128) for cx in range(NX // 64)]
    lp = gmm.LocalPixelization.make_cells(NY, NX, cells, periodic_xcoord=False)
    wp = pr["mm"].WeightedPointing(lp, pr["plans"], pr["xps"], pr["ninv"])
    m = torch.ones(3 * lp.npix, dtype=torch.float32, device=dev)
    assert gmm.get_errflag_check()
    wp.apply_normal(m)                       # returns without raising
    assert gmm.get_errflag_check()
    with pytest.raises(RuntimeError, match="not in LocalPixelization"):
        wp.check_errflags()
    wp.check_errflags()                      # the check cleared the flags
