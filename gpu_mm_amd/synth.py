"""Synthetic scan pointing, maps and timestreams for tests and benchmarks.

Pointing is computed in float64 and rounded to float32. Random draws use numpy's PCG64
(np.random.default_rng(seed)) in a fixed order: detector offsets, detector angles, map, TOD.

  tri(u) = 1 - 4 |frac(u) - 1/2|      (period-1 triangle wave with range [-1, 1])
  detector offsets uniform in a disk of radius R: r = R sqrt(u1), phi = 2 pi u2,
  dy = r sin(phi), dx = r cos(phi); theta_det uniform in [0, pi).
"""

import math

import numpy as np


def tri(u):
    return 1.0 - 4.0 * np.abs((u - np.floor(u)) - 0.5)


def disk_offsets(rng, ndet, radius):
    u1 = rng.random(ndet)
    u2 = rng.random(ndet)
    r = radius * np.sqrt(u1)
    phi = 2 * np.pi * u2
    return r * np.sin(phi), r * np.cos(phi)


def raster_xpointing(ndet, nt, *, x0, xamp, xperiod, y0, yspeed, radius, alpha_rate, seed,
                     rng=None, t_chunk=None):
    """
    xpointing float32[3][ndet][nt] for a constant-elevation raster:
      x = x0 + xamp * tri(t / xperiod) + dx_det,  y = y0 + yspeed * t + dy_det,
      alpha = theta_det + alpha_rate * t.
    Returns (xpointing, rng) so that callers continue the same random stream.
    """
    rng = np.random.default_rng(seed) if rng is None else rng
    dy, dx = disk_offsets(rng, ndet, radius)
    theta = rng.uniform(0.0, np.pi, ndet)
    xp = np.empty((3, ndet, nt), dtype=np.float32)
    t = np.arange(nt, dtype=np.float64)
    xs = x0 + xamp * tri(t / xperiod)
    ys = y0 + yspeed * t
    al = alpha_rate * t
    for d in range(ndet):
        xp[0, d] = ys + dy[d]
        xp[1, d] = xs + dx[d]
        xp[2, d] = al + theta[d]
    return xp, rng


# ---------------------------------------------------------------------------------------------
# Benchmark configurations.

CFG0 = dict(ny=256, nx=512, ndet=64, nt=65536,
            scan=dict(x0=256.0, xamp=200.0, xperiod=8192.0, y0=60.0, yspeed=0.0025, radius=20.0,
                      alpha_rate=0.001))

CFG1 = dict(ny=4096, nx=8192, ndet=2048, nt=262144,
            scan=dict(x0=4096.0, xamp=3000.0, xperiod=40000.0, y0=500.0, yspeed=0.01, radius=300.0,
                      alpha_rate=2 * math.pi / 1.0e6))


def make_cfg0(seed=0, ndet=None, nt=None):
    """Config 0: map 3x256x512 N(0,1) (dense), TOD N(0,1), 64 x 65536 raster."""
    c = CFG0
    ndet = c["ndet"] if ndet is None else ndet
    nt = c["nt"] if nt is None else nt
    xp, rng = raster_xpointing(ndet, nt, seed=seed, **c["scan"])
    m = rng.standard_normal((3, c["ny"], c["nx"])).astype(np.float32)
    tod = rng.standard_normal((ndet, nt)).astype(np.float32)
    return dict(xpointing=xp, map=m, tod=tod, ny=c["ny"], nx=c["nx"])


def grf_map_torch(ny, nx, seed, device, ncomp=3, chunk=1):
    """Gaussian random field with P(k) ~ k^-2 per component (white noise times |k|^-1 in
    Fourier space, k=0 removed), normalized to unit variance; float32 (ncomp, ny, nx)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    ky = torch.fft.fftfreq(ny, device=device, dtype=torch.float64).reshape(-1, 1)
    kx = torch.fft.rfftfreq(nx, device=device, dtype=torch.float64).reshape(1, -1)
    k = torch.sqrt(ky * ky + kx * kx)
    filt = torch.where(k > 0, 1.0 / torch.clamp(k, min=1e-30), torch.zeros_like(k)).to(torch.float32)
    out = torch.empty((ncomp, ny, nx), dtype=torch.float32, device=device)
    for c in range(ncomp):
        w = torch.randn((ny, nx), generator=g, device=device, dtype=torch.float32)
        f = torch.fft.rfft2(w) * filt
        v = torch.fft.irfft2(f, s=(ny, nx))
        out[c] = v / v.std()
        del w, f, v
    return out


def raster_xpointing_torch(ndet, nt, *, x0, xamp, xperiod, y0, yspeed, radius, alpha_rate, seed, device):
    """GPU version of raster_xpointing (same formulas; detector parameters from the same numpy
    stream); returns a float32 torch tensor (3, ndet, nt)."""
    import torch
    rng = np.random.default_rng(seed)
    dy, dx = disk_offsets(rng, ndet, radius)
    theta = rng.uniform(0.0, np.pi, ndet)
    xp = torch.empty((3, ndet, nt), dtype=torch.float32, device=device)
    t = torch.arange(nt, dtype=torch.float64, device=device)
    u = t / xperiod
    xs = x0 + xamp * (1.0 - 4.0 * torch.abs((u - torch.floor(u)) - 0.5))
    ys = y0 + yspeed * t
    al = alpha_rate * t
    dyt = torch.from_numpy(dy).to(device)
    dxt = torch.from_numpy(dx).to(device)
    tht = torch.from_numpy(theta).to(device)
    step = max(1, (1 << 27) // nt)
    for d0 in range(0, ndet, step):
        d1 = min(ndet, d0 + step)
        xp[0, d0:d1] = (ys[None, :] + dyt[d0:d1, None]).to(torch.float32)
        xp[1, d0:d1] = (xs[None, :] + dxt[d0:d1, None]).to(torch.float32)
        xp[2, d0:d1] = (al[None, :] + tht[d0:d1, None]).to(torch.float32)
    return xp, rng
