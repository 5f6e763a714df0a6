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


# ---------------------------------------------------------------------------------------------
# Batched uneven workload: many TODs of different sizes, each with its own rotated raster scan,
# on one shared map.
#
# TOD i: a scan frame (u, v) with u = amp * tri(t / 8192) (the sweep) and v = 0.0025 t (the
# drift), rotated by phi: y = y0 + u sin(phi) + v cos(phi) + dy_det, x = x0 + u cos(phi) -
# v sin(phi) + dx_det; detector offsets uniform in a disk of radius 20 px and alpha = theta_det
# + 0.001 t (the offset/alpha model of the 64 x 65536 raster). (y0, x0) is drawn uniformly among
# the positions that keep the whole footprint (plus the offset radius and a 2-pixel margin)
# inside 0 <= y <= ny-1, 0 <= x <= nx-2; phi is redrawn (up to 64 times) while the footprint
# does not fit, and then the amplitude is halved until it does.

BATCH_SCAN = dict(xperiod=8192.0, speed=0.0025, radius=20.0, alpha_rate=0.001, margin=2.0)

CFG2 = dict(ny=2048, nx=4096, ntod=24, ndet_choices=tuple(range(96, 1537, 32)),
            nt_choices=tuple(range(16384, 262145, 64)), amp_range=(300.0, 1500.0))


def batch_scan_params(rng, ndet, nt, ny, nx, amp_range):
    """Draws one TOD's scan: returns dict(y0, x0, amp, phi) (float64)."""
    s = BATCH_SCAN
    amp = rng.uniform(*amp_range)
    vmax = s["speed"] * (nt - 1)
    pad = s["radius"] + s["margin"]
    for attempt in range(10000):
        phi = rng.uniform(0.0, np.pi)
        sn, cs = np.sin(phi), np.cos(phi)
        uu = np.array([-amp, amp, -amp, amp])
        vv = np.array([0.0, 0.0, vmax, vmax])
        ys = uu * sn + vv * cs
        xs = uu * cs - vv * sn
        ylo, yhi = pad - ys.min(), (ny - 1) - pad - ys.max()
        xlo, xhi = pad - xs.min(), (nx - 2) - pad - xs.max()
        if ylo <= yhi and xlo <= xhi:
            return dict(y0=rng.uniform(ylo, yhi), x0=rng.uniform(xlo, xhi), amp=amp, phi=phi)
        if attempt % 64 == 63:
            amp *= 0.5
    raise RuntimeError("batch_scan_params: the map is too small for this scan")


def batch_xpointing_torch(ndet, nt, scan, seed, device):
    """xpointing float32 (3, ndet, nt) on 'device' for one TOD of the batched workload
    (detector offsets and angles from np.random.default_rng(seed))."""
    import torch
    s = BATCH_SCAN
    rng = np.random.default_rng(seed)
    dy, dx = disk_offsets(rng, ndet, s["radius"])
    theta = rng.uniform(0.0, np.pi, ndet)
    t = torch.arange(nt, dtype=torch.float64, device=device)
    w = t / s["xperiod"]
    u = scan["amp"] * (1.0 - 4.0 * torch.abs((w - torch.floor(w)) - 0.5))
    v = s["speed"] * t
    sn, cs = float(np.sin(scan["phi"])), float(np.cos(scan["phi"]))
    ys = scan["y0"] + u * sn + v * cs
    xs = scan["x0"] + u * cs - v * sn
    al = s["alpha_rate"] * t
    xp = torch.empty((3, ndet, nt), dtype=torch.float32, device=device)
    dyt = torch.from_numpy(dy).to(device)
    dxt = torch.from_numpy(dx).to(device)
    tht = torch.from_numpy(theta).to(device)
    step = max(1, (1 << 26) // nt)
    for d0 in range(0, ndet, step):
        d1 = min(ndet, d0 + step)
        xp[0, d0:d1] = (ys[None, :] + dyt[d0:d1, None]).to(torch.float32)
        xp[1, d0:d1] = (xs[None, :] + dxt[d0:d1, None]).to(torch.float32)
        xp[2, d0:d1] = (al[None, :] + tht[d0:d1, None]).to(torch.float32)
    return xp


def batch_shapes(seed, ntod, ndet_choices, nt_choices, ny, nx, amp_range):
    """Per-TOD (ndet, nt, scan) of a batched workload, drawn from np.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ntod):
        ndet = int(rng.choice(ndet_choices))
        nt = int(rng.choice(nt_choices))
        out.append((ndet, nt, batch_scan_params(rng, ndet, nt, ny, nx, amp_range)))
    return out


# ---------------------------------------------------------------------------------------------
# Map-making workload: TODs whose scans cross at several angles, all centred on the map, so that
# the crossings constrain Q and U. Scan model as for the batched workload (sweep amplitude 'amp',
# drift 0.0025 px per sample), with the drift's midpoint at the map centre.

CFG3 = dict(ny=1024, nx=2048, ndet=512, nt=131072, angles_deg=(0.0, 45.0, 90.0, 135.0), amp=300.0,
            noise_var=(0.5, 2.0))


def crossing_scans(ny, nx, nt, angles_deg, amp):
    """Scan dicts (y0, x0, amp, phi) for batch_xpointing_torch, one per angle."""
    s = BATCH_SCAN
    vmax = s["speed"] * (nt - 1)
    out = []
    for a in angles_deg:
        phi = np.deg2rad(a)
        cy, cx = 0.5 * (ny - 1), 0.5 * (nx - 2)
        out.append(dict(y0=cy - 0.5 * vmax * np.cos(phi), x0=cx + 0.5 * vmax * np.sin(phi), amp=amp, phi=phi))
    return out
