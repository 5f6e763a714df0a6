"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker.

No golden vectors exist in the reference (its tests are unseeded same-run GPU-vs-CPU
comparisons), so each oracle function is pinned by hand-computed cases, closed forms,
invariants, and an independent numpy / pure-Python restatement on small inputs.
"""

import math

import numpy as np
import pytest

F32 = np.float32


# ------------------------------------------------------------------------------------------
# Independent numpy restatement of locate() (truncation-based, as the reference writes it).

def np_locate(y, x, ny, nx, periodic):
    y = np.asarray(y, dtype=F32)
    x = np.asarray(x, dtype=F32).copy()
    err = np.zeros(y.shape, dtype=np.uint32)
    err |= np.where((y >= 0) & (y <= F32(ny - 1)), 0, 1).astype(np.uint32)
    iy0 = np.clip(np.trunc(np.nan_to_num(y, nan=0.0, posinf=1e9, neginf=-1e9)), 0, ny - 2).astype(np.int64)
    dy = (y - iy0.astype(F32)).astype(F32)
    if periodic:
        err |= np.where((x >= F32(-nx)) & (x <= F32(2 * nx)), 0, 2).astype(np.uint32)
        x = np.where(x >= 0, x, (x + F32(nx)).astype(F32)).astype(F32)
        x = np.where(x <= F32(nx), x, (x - F32(nx)).astype(F32)).astype(F32)
        ix0 = np.maximum(np.trunc(np.nan_to_num(x, nan=0.0, posinf=float(nx), neginf=-1e9)), 0).astype(np.int64)
        ix0 = np.minimum(ix0, nx)
        dx = (x - ix0.astype(F32)).astype(F32)
        ix0 = np.where(ix0 < nx, ix0, 0)
        ix1 = np.where(ix0 + 1 < nx, ix0 + 1, 0)
    else:
        err |= np.where((x >= 0) & (x <= F32(nx - 2)), 0, 2).astype(np.uint32)
        ix0 = np.clip(np.trunc(np.nan_to_num(x, nan=0.0, posinf=1e9, neginf=-1e9)), 0, nx - 2).astype(np.int64)
        ix1 = ix0 + 1
        dx = (x - ix0.astype(F32)).astype(F32)
    return err, iy0, iy0 + 1, ix0, ix1, dy, dx


# Hand-computed cases: (y, x, ny, nx, periodic) -> (err, iy0, iy1, ix0, ix1, dy, dx)
LOCATE_CASES = [
    ((10.0, 20.0, 256, 256, False), (0, 10, 11, 20, 21, 0.0, 0.0)),
    ((10.25, 20.75, 256, 256, False), (0, 10, 11, 20, 21, 0.25, 0.75)),
    ((255.0, 3.5, 256, 256, False), (0, 254, 255, 3, 4, 1.0, 0.5)),        # y = ny-1 clamps, dy=1
    ((255.5, 3.5, 256, 256, False), (1, 254, 255, 3, 4, 1.5, 0.5)),        # y > ny-1
    ((-0.5, 3.5, 256, 256, False), (1, 0, 1, 3, 4, -0.5, 0.5)),            # y < 0
    ((5.0, 254.0, 256, 256, False), (0, 5, 6, 254, 255, 0.0, 0.0)),        # x = nx-2 is valid
    ((5.0, 254.5, 256, 256, False), (2, 5, 6, 254, 255, 0.0, 0.5)),        # x > nx-2
    ((5.0, -0.5, 256, 256, True), (0, 5, 6, 255, 0, 0.0, 0.5)),            # wrap: -0.5 -> 255.5
    ((5.0, 256.0, 256, 256, True), (0, 5, 6, 0, 1, 0.0, 0.0)),             # x = nx folds to 0
    ((5.0, 300.25, 256, 256, True), (0, 5, 6, 44, 45, 0.0, 0.25)),         # x > nx shifts by -nx
    ((5.0, -256.0, 256, 256, True), (0, 5, 6, 0, 1, 0.0, 0.0)),            # x = -nx valid
    ((5.0, 512.0, 256, 256, True), (0, 5, 6, 0, 1, 0.0, 0.0)),             # x = 2nx valid
    ((5.0, 512.5, 256, 256, True), (2, 5, 6, None, None, None, None)),      # x > 2nx
    ((5.0, -256.5, 256, 256, True), (2, 5, 6, None, None, None, None)),     # x < -nx
    ((5.0, 255.5, 256, 256, True), (0, 5, 6, 255, 0, 0.0, 0.5)),           # ix1 wraps
    # float32 rounding of x + nx: -1e-6 + 65536 rounds to 65536 in FP32, which folds to pixel 0
    # with dx = 0 (an exact evaluation would give ix0 = 65535, dx ~ 1).
    ((5.0, -1e-6, 256, 65536, True), (0, 5, 6, 0, 1, 0.0, 0.0)),
    ((float("nan"), 3.0, 256, 256, False), (1, 0, 1, 3, 4, None, 0.0)),
]


@pytest.mark.parametrize("args,expect", LOCATE_CASES)
def test_locate_hand_cases(orc, args, expect):
    y, x, ny, nx, periodic = args
    err, idx, d = orc.locate(y, x, ny, nx, periodic)
    e_err, e_iy0, e_iy1, e_ix0, e_ix1, e_dy, e_dx = expect
    assert err == e_err
    assert (idx[0], idx[1]) == (e_iy0, e_iy1)
    if e_ix0 is not None:
        assert (idx[2], idx[3]) == (e_ix0, e_ix1)
    if e_dy is not None:
        assert d[0] == F32(e_dy)
    if e_dx is not None:
        assert d[1] == F32(e_dx)


@pytest.mark.parametrize("periodic", [False, True])
def test_locate_matches_numpy_restatement(orc, periodic):
    rng = np.random.default_rng(1 + periodic)
    ny, nx = 300, 200
    n = 20000
    y = rng.uniform(-2, ny + 1, n).astype(F32)
    x = (rng.uniform(-nx - 3, 2 * nx + 3, n) if periodic else rng.uniform(-2, nx + 1, n)).astype(F32)
    # Include exact integers and boundary values.
    y[:50] = np.arange(50, dtype=F32) * 6
    x[:50] = (np.arange(50, dtype=F32) * 4 - (nx if periodic else 0))
    err, iy0, iy1, ix0, ix1, dy, dx = np_locate(y, x, ny, nx, periodic)
    for i in range(n):
        e, idx, d = orc.locate(float(y[i]), float(x[i]), ny, nx, periodic)
        assert e == err[i], i
        assert tuple(idx) == (iy0[i], iy1[i], ix0[i], ix1[i]), i
        assert d[0] == dy[i] and (e & 2 or d[1] == dx[i]), i


# ------------------------------------------------------------------------------------------
# map2tod / tod2map closed forms.

def random_pointing(rng, n, ny, nx, periodic, margin=0.0):
    y = rng.uniform(margin, ny - 1 - margin, n)
    if periodic:
        x = rng.uniform(-nx, 2 * nx, n)
    else:
        x = rng.uniform(margin, nx - 2 - margin, n)
    a = rng.uniform(-4, 4, n)
    return np.stack([y, x, a]).astype(F32)


def global_to_local(m, pix):
    """Copies a (3, ny_padded, nx_padded) global map into the pixelization's layout."""
    out = np.zeros(3 * pix.npix, dtype=m.dtype)
    for cy in range(pix.cell_offsets.shape[0]):
        for cx in range(pix.cell_offsets.shape[1]):
            o = pix.cell_offsets[cy, cx]
            if o < 0:
                continue
            for p in range(3):
                for r in range(64):
                    out[o + p * pix.polstride + r * pix.ystride: o + p * pix.polstride + r * pix.ystride + 64] = \
                        m[p, 64 * cy + r, 64 * cx: 64 * cx + 64]
    return out


@pytest.mark.parametrize("periodic", [False, True])
def test_map2tod_constant_map(orc, periodic):
    rng = np.random.default_rng(3)
    ny, nx = 128, 192
    pix = orc.Pix.rectangle(ny, nx, periodic)
    a, b, c = 0.7, -1.3, 2.1
    m = np.zeros((3, 128, 192))
    m[0], m[1], m[2] = a, b, c
    xp = random_pointing(rng, 4096, ny, nx, periodic)
    tod64, err = orc.map2tod(global_to_local(m, pix), xp, pix, f64=True)
    assert err == 0
    al = xp[2].astype(np.float64)
    expect = a + b * np.cos(2 * al) + c * np.sin(2 * al)
    assert np.max(np.abs(tod64 - expect)) < 1e-12
    tod32, err = orc.map2tod(global_to_local(m, pix).astype(F32), xp, pix)
    assert err == 0
    assert np.max(np.abs(tod32 - expect)) < 3e-6


def test_map2tod_affine_map_exact(orc):
    # Bilinear interpolation reproduces a function affine in (y, x) exactly; a swapped dy/dx,
    # a wrong corner or a wrong weight fails this.
    rng = np.random.default_rng(4)
    ny, nx = 128, 128
    pix = orc.Pix.rectangle(ny, nx, False)
    yy, xx = np.meshgrid(np.arange(128.0), np.arange(128.0), indexing="ij")
    p0, py, px = 0.25, 0.125, -0.0625
    m = np.zeros((3, 128, 128))
    m[0] = p0 + py * yy + px * xx
    m[1] = 2 * py * yy            # Q affine too
    xp = random_pointing(rng, 4096, ny, nx, False)
    tod, err = orc.map2tod(global_to_local(m, pix), xp, pix, f64=True)
    assert err == 0
    y, x, al = (xp[i].astype(np.float64) for i in range(3))
    expect = p0 + py * y + px * x + np.cos(2 * al) * (2 * py * y)
    assert np.max(np.abs(tod - expect)) < 1e-10


def test_map2tod_integer_pointing_t_plus_q(orc):
    rng = np.random.default_rng(5)
    ny, nx = 100, 130
    pix = orc.Pix.rectangle(ny, nx, False)
    m = rng.standard_normal((3, 128, 192)).astype(F32)
    n = 1024
    iy = rng.integers(0, ny - 1, n)
    ix = rng.integers(0, nx - 2, n)
    xp = np.stack([iy, ix, np.zeros(n)]).astype(F32)
    tod, err = orc.map2tod(global_to_local(m, pix), xp, pix)
    assert err == 0
    assert np.array_equal(tod, (m[0, iy, ix] + m[1, iy, ix]).astype(F32))


@pytest.mark.parametrize("periodic", [False, True])
def test_fp32_matches_fp64(orc, periodic):
    rng = np.random.default_rng(6)
    ny, nx = 192, 256
    pix = orc.Pix.rectangle(ny, nx, periodic)
    m = rng.standard_normal(3 * pix.npix)
    xp = random_pointing(rng, 8192, ny, nx, periodic)
    # With periodic_xcoord the FP32 path shifts x by +-nx in float32, which rounds the
    # coordinate (the FP64 path shifts exactly), so its agreement is ~1e-6 instead of ~1e-7.
    tol = 1e-5 if periodic else 1e-6
    t32, e1 = orc.map2tod(m.astype(F32), xp, pix)
    t64, e2 = orc.map2tod(m, xp, pix, f64=True)
    assert e1 == e2 == 0
    assert np.sum(np.abs(t32 - t64)) / np.sum(np.abs(t64)) < tol
    d = rng.standard_normal(8192)
    m32, _ = orc.tod2map(d.astype(F32), xp, pix)
    m64, _ = orc.tod2map(d, xp, pix, f64=True)
    assert np.sum(np.abs(m32 - m64)) / np.sum(np.abs(m64)) < tol


@pytest.mark.parametrize("periodic", [False, True])
def test_adjoint_identity(orc, periodic):
    rng = np.random.default_rng(7)
    ny, nx = 130, 250
    pix = orc.Pix.rectangle(ny, nx, periodic)
    m = rng.standard_normal(3 * pix.npix)
    d = rng.standard_normal(8192)
    xp = random_pointing(rng, 8192, ny, nx, periodic)
    am, _ = orc.map2tod(m, xp, pix, f64=True)
    atd, _ = orc.tod2map(d, xp, pix, f64=True)
    lhs, rhs = np.dot(am, d), np.dot(m, atd)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.linalg.norm(am) * np.linalg.norm(d))


def test_tod2map_sum_rules(orc):
    rng = np.random.default_rng(8)
    ny, nx = 128, 128
    pix = orc.Pix.rectangle(ny, nx, False)
    n = 4096
    xp = random_pointing(rng, n, ny, nx, False)
    xp[2] = 0.0
    mm, err = orc.tod2map(np.ones(n), xp, pix, f64=True)
    assert err == 0
    npix = pix.npix
    assert abs(mm[:npix].sum() - n) < 1e-9 and abs(mm[npix:2 * npix].sum() - n) < 1e-9
    assert abs(mm[2 * npix:].sum()) < 1e-9


def test_tod2map_single_deposit(orc):
    ny, nx = 128, 128
    pix = orc.Pix.rectangle(ny, nx, False)
    xp = np.zeros((3, 32), dtype=F32)
    xp[0], xp[1] = 17.0, 33.0
    xp[2] = 0.3
    t = np.zeros(32)
    t[0] = 2.0
    mm, err = orc.tod2map(t, xp, pix, f64=True)
    assert err == 0
    k = 17 * pix.ystride + 33
    npix = pix.npix
    assert mm[k] == 2.0
    assert abs(mm[npix + k] - 2 * math.cos(2 * float(F32(0.3)))) < 1e-15
    assert abs(mm[2 * npix + k] - 2 * math.sin(2 * float(F32(0.3)))) < 1e-15
    assert np.count_nonzero(mm) == 3


def test_partial_pixelization(orc):
    rng = np.random.default_rng(9)
    ny, nx = 128, 128
    full = orc.Pix.rectangle(ny, nx, False)
    off = full.cell_offsets.copy()
    # Drop cell (1,1): remaining cells must stay contiguous, so use a cell list layout.
    cells = [(0, 0), (0, 1), (1, 0)]
    off = -np.ones((2, 2), dtype=np.int64)
    for i, (cy, cx) in enumerate(cells):
        off[cy, cx] = 3 * 4096 * i
    pix = orc.Pix(ny, nx, off, 64, 4096, False)
    m = rng.standard_normal(3 * pix.npix)
    xp = random_pointing(rng, 2048, ny, nx, False)
    _, err = orc.map2tod(m, xp, pix, f64=True)
    assert err == orc.NOT_IN_PIX
    t, err = orc.map2tod(m, xp, pix, f64=True, partial=True)
    assert err == 0
    # Samples entirely inside cell (1,1) read 0.
    inside = (xp[0] >= 64) & (xp[0] < 127) & (xp[1] >= 64) & (xp[1] < 126)
    assert np.all(t[inside] == 0)


# ------------------------------------------------------------------------------------------
# Planner.

def test_preplan_params_closed_forms(orc):
    assert orc.preplan_params(512) == (16, 1)
    assert orc.preplan_params(4194304) == (128, 1024)
    assert orc.preplan_params(536870912) == (16384, 1024)
    assert orc.preplan_params(32 * 16400) == (20, 820)   # ceil(16400/1024)=17 -> 20
    assert orc.pointing_params(140000) == (384, 365)
    assert orc.pointing_params(1) == (32, 1)


def py_plan(xp, ny, nx, periodic):
    """Pure-Python brute-force planner (set semantics), for tiny inputs."""
    n = xp.shape[1]
    err, iy0, iy1, ix0, ix1, _, _ = np_locate(xp[0], xp[1], ny, nx, periodic)
    entries = []
    per, nb = None, None
    lines = []
    for l in range((n + 31) // 32):
        cells = set()
        first = None
        for s in range(32 * l, min(32 * l + 32, n)):
            for (a, b) in ((iy0[s], ix0[s]), (iy0[s], ix1[s]), (iy1[s], ix0[s]), (iy1[s], ix1[s])):
                k = ((int(a) >> 6) << 10) | (int(b) >> 6)
                if first is None:
                    first = k
                cells.add(k)
        lines.append(len(cells))
        for k in cells:
            e = k | (l << 20)
            if len(cells) > 1:
                e |= 1 << 46
                if k == first:
                    e |= 1 << 47
            entries.append(e)
    entries.sort(key=lambda e: ((e & 0xFFFFF), (e >> 20) & 0x3FFFFFF))
    return np.array(entries, dtype=np.uint64), lines, int(np.bitwise_or.reduce(err))


# n = 32 * 700 - 13 and 32 * 700 + 4: a partial last line (nsamp % 32 != 0, which the library
# accepts and the reference does not).
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("n", [32 * 700, 32 * 700 - 13, 32 * 700 + 4])
def test_plan_matches_python_bruteforce(orc, periodic, n):
    rng = np.random.default_rng(10 + periodic)
    ny, nx = 300, 400
    # Fast scan so lines often cross cells, plus wrap-around in the periodic case.
    t = np.arange(n)
    y = 150 + 140 * np.sin(t / 500.0) + rng.uniform(-1, 1, n)
    x = (t * 3.3) % (nx + (150 if periodic else -10)) - (50 if periodic else -5)
    a = rng.uniform(0, 3, n)
    xp = np.stack([y, x, a]).astype(F32)
    cum, nmt, err = orc.plan_count(xp, ny, nx, periodic)
    assert err == 0
    mt, err = orc.plan_build(xp, ny, nx, periodic, nmt)
    assert err == 0
    ref, lines, perr = py_plan(xp, ny, nx, periodic)
    assert perr == 0
    assert nmt == len(ref) == sum(lines)
    assert np.array_equal(mt, ref)
    # nmt_cumsum over planner blocks
    per, nb = orc.preplan_params(n)
    expect = np.cumsum([sum(lines[b * per:(b + 1) * per]) for b in range(nb)])
    assert np.array_equal(cum, expect.astype(np.uint32))


def test_plan_invariants(orc):
    from gpu_mm_amd import synth
    d = synth.make_cfg0(seed=0, ndet=4, nt=8192)
    xp = d["xpointing"].reshape(3, -1)
    mt, err = orc.plan_build(xp, d["ny"], d["nx"], False)
    assert err == 0
    cell = mt & np.uint64(0xFFFFF)
    line = (mt >> np.uint64(20)) & np.uint64(0x3FFFFFF)
    mflag = (mt >> np.uint64(46)) & np.uint64(1)
    zflag = (mt >> np.uint64(47)) & np.uint64(1)
    key = cell * np.uint64(1 << 26) + line
    assert np.all(key[1:] > key[:-1])                    # sorted by (cell, line), unique
    counts = np.bincount(line.astype(np.int64), minlength=xp.shape[1] // 32)
    assert np.array_equal(mflag, (counts[line.astype(np.int64)] > 1).astype(np.uint64))
    zl = line[zflag == 1].astype(np.int64)
    assert np.array_equal(np.sort(zl), np.nonzero(counts > 1)[0])   # one zflag per multi-cell line
    assert np.all(mflag[zflag == 1] == 1)


# ------------------------------------------------------------------------------------------
# LocalPixelization rules (src_lib/LocalPixelization.cu:33-101).

def test_lpix_rules(orc):
    rect = orc.Pix.rectangle(256, 512, False)
    assert orc.lpix_check(256, 512, rect.cell_offsets, rect.ystride, rect.polstride) == (0, 256 * 512)
    # cell list layout
    off = np.array([[0, 12288], [-1, 24576]], dtype=np.int64)
    assert orc.lpix_check(128, 128, off, 64, 4096) == (0, 3 * 4096)
    assert orc.lpix_check(0, 128, off, 64, 4096)[0] == 1
    assert orc.lpix_check(128, 65537, off, 64, 4096)[0] == 1
    assert orc.lpix_check(128, 128, np.zeros((0, 2), np.int64), 64, 4096)[0] == 2
    assert orc.lpix_check(128, 128, off, 32, 4096)[0] == 3
    assert orc.lpix_check(128, 128, off, 64, 4000)[0] == 4
    assert orc.lpix_check(128, 128, off, 64 * 3, 64 * 5)[0] == 5
    bad = off.copy(); bad[0, 1] = 12290
    assert orc.lpix_check(128, 128, bad, 64, 4096)[0] == 6
    gap = off.copy(); gap[1, 1] = 3 * 12288
    assert orc.lpix_check(128, 128, gap, 64, 4096)[0] == 7


# ------------------------------------------------------------------------------------------
# Response operators (per-detector T and P gains), pinned by decomposition into the plain
# operators (themselves pinned above): with the samples split by detector, T and QU parts of
# the map are scaled independently. Plus the reference's own check (gpu_mm/tests.py:295-315:
# response (2, -1) with the map divided by (2, -1) leaves the TOD unchanged) and adjointness.

def _resp_setup(orc, seed, periodic, ndet=6, nper=700):
    rng = np.random.default_rng(seed)
    ny, nx = 100, 190
    pix = orc.Pix.rectangle(ny, nx, periodic)
    xp = random_pointing(rng, ndet * nper, ny, nx, periodic)
    resp = rng.uniform(-2, 2, size=(2, ndet)).astype(F32)
    return rng, pix, xp, resp, ndet, nper


@pytest.mark.parametrize("periodic", [False, True])
def test_response_map2tod_decomposes(orc, periodic):
    rng, pix, xp, resp, ndet, nper = _resp_setup(orc, 21, periodic)
    m = rng.standard_normal(3 * pix.npix)
    npix = pix.npix
    m_t = m.copy(); m_t[npix:] = 0.0
    m_p = m.copy(); m_p[:npix] = 0.0
    tod_t, _ = orc.map2tod(m_t, xp, pix, f64=True)
    tod_p, _ = orc.map2tod(m_p, xp, pix, f64=True)
    det = np.arange(ndet * nper) // nper
    expect = resp[0, det] * tod_t + resp[1, det] * tod_p
    got, err = orc.response_map2tod(m, xp, resp, pix)
    assert err == 0
    np.testing.assert_allclose(got, expect, rtol=1e-12, atol=1e-12)
    base = rng.standard_normal(ndet * nper)
    got2, _ = orc.response_map2tod(m, xp, resp, pix, tod=base, accum=True)
    np.testing.assert_allclose(got2, base + expect, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("periodic", [False, True])
def test_response_tod2map_decomposes(orc, periodic):
    rng, pix, xp, resp, ndet, nper = _resp_setup(orc, 22, periodic)
    d = rng.standard_normal(ndet * nper)
    npix = pix.npix
    det = np.arange(ndet * nper) // nper
    expect = np.zeros(3 * npix)
    for i in range(ndet):
        di = np.where(det == i, d, 0.0)
        mi, _ = orc.tod2map(di, xp, pix, f64=True)
        expect[:npix] += resp[0, i] * mi[:npix]
        expect[npix:] += resp[1, i] * mi[npix:]
    got, err = orc.response_tod2map(d, xp, resp, pix)
    assert err == 0
    np.testing.assert_allclose(got, expect, rtol=1e-10, atol=1e-10)


def test_response_reference_check_and_adjoint(orc):
    rng, pix, xp, _, ndet, nper = _resp_setup(orc, 23, True)
    npix = pix.npix
    m = rng.standard_normal(3 * pix.npix)
    plain, _ = orc.map2tod(m, xp, pix, f64=True)
    resp = np.zeros((2, ndet), dtype=F32)
    resp[0], resp[1] = 2.0, -1.0
    m2 = m.copy(); m2[:npix] /= 2.0; m2[npix:] /= -1.0
    got, _ = orc.response_map2tod(m2, xp, resp, pix)
    np.testing.assert_allclose(got, plain, rtol=1e-12, atol=1e-12)
    resp = rng.uniform(-2, 2, size=(2, ndet)).astype(F32)
    d = rng.standard_normal(ndet * nper)
    am, _ = orc.response_map2tod(m, xp, resp, pix)
    atd, _ = orc.response_tod2map(d, xp, resp, pix)
    lhs, rhs = np.dot(am, d), np.dot(m, atd)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.linalg.norm(am) * np.linalg.norm(d))


# ------------------------------------------------------------------------------------------
# The OpenMP variants (the multi-core CPU baseline) against the single-thread oracle.

@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_multithreaded_variants_match_single_thread(orc, periodic, nthreads):
    rng = np.random.default_rng(40 + nthreads)
    ny, nx = 200, 300
    pix = orc.Pix.rectangle(ny, nx, periodic)
    n = 3 * 32 * 1000 + 17                      # not a multiple of the thread count or of 32
    xp = random_pointing(rng, n, ny, nx, periodic)
    m = rng.standard_normal(3 * pix.npix).astype(F32)
    t0 = rng.standard_normal(n).astype(F32)
    for accum in (False, True):
        ref, e1 = orc.map2tod(m, xp, pix, tod=t0, accum=accum)
        got, e2 = orc.map2tod(m, xp, pix, tod=t0, accum=accum, nthreads=nthreads)
        assert e1 == e2 == 0
        assert np.array_equal(got, ref)        # per-sample work: identical bits
    d = rng.standard_normal(n).astype(F32)
    m0 = rng.standard_normal(3 * pix.npix).astype(F32)
    ref, e1 = orc.tod2map(d, xp, pix, lmap=m0)
    got, e2 = orc.tod2map(d, xp, pix, lmap=m0, nthreads=nthreads)
    assert e1 == e2 == 0
    # private maps summed afterwards: a different float summation order
    assert np.sum(np.abs(got - ref)) / np.sum(np.abs(ref)) < 1e-6
    # and the FP64 evaluation on the same inputs
    r64, _ = orc.tod2map(d.astype(np.float64), xp, pix, lmap=m0.astype(np.float64), f64=True)
    assert np.sum(np.abs(got - r64)) / np.sum(np.abs(r64)) < 1e-6


def test_multithreaded_errors_and_partial(orc):
    rng = np.random.default_rng(44)
    ny, nx = 128, 128
    off = -np.ones((2, 2), dtype=np.int64)
    for i, (cy, cx) in enumerate([(0, 0), (0, 1), (1, 0)]):
        off[cy, cx] = 3 * 4096 * i
    pix = orc.Pix(ny, nx, off, 64, 4096, False)
    xp = random_pointing(rng, 4096, ny, nx, False)
    xp[0, 100] = ny + 5.0                       # one bad y
    m = rng.standard_normal(3 * pix.npix).astype(F32)
    for partial in (False, True):
        ref, e1 = orc.map2tod(m, xp, pix, partial=partial)
        got, e2 = orc.map2tod(m, xp, pix, partial=partial, nthreads=4)
        assert e1 == e2 and (e1 & orc.BAD_Y)
        assert bool(e1 & orc.NOT_IN_PIX) == (not partial)
        d = rng.standard_normal(4096).astype(F32)
        ref, e1 = orc.tod2map(d, xp, pix, partial=partial)
        got, e2 = orc.tod2map(d, xp, pix, partial=partial, nthreads=4)
        assert e1 == e2


# ------------------------------------------------------------------------------------------
# The wrappers refuse arrays the C loops would overrun.

def test_wrappers_check_lengths(orc):
    rng = np.random.default_rng(45)
    pix = orc.Pix.rectangle(128, 128, False)
    xp = random_pointing(rng, 256, 128, 128, False)
    m = np.zeros(3 * pix.npix, dtype=F32)
    resp = np.ones((2, 8), dtype=F32)
    with pytest.raises(ValueError, match="lmap"):
        orc.map2tod(m[:-1], xp, pix)
    with pytest.raises(ValueError, match="tod"):
        orc.map2tod(m, xp, pix, tod=np.zeros(255, F32))
    with pytest.raises(ValueError, match="tod"):
        orc.tod2map(np.zeros(257, F32), xp, pix)
    with pytest.raises(ValueError, match="lmap"):
        orc.tod2map(np.zeros(256, F32), xp, pix, lmap=np.zeros(3 * pix.npix - 64, F32))
    with pytest.raises(ValueError, match="lmap"):
        orc.tod2map(np.zeros(256, F32), xp, pix, lmap=m[:-1], nthreads=2)
    with pytest.raises(ValueError, match="lmap"):
        orc.response_map2tod(m[:-1], xp, resp, pix)
    with pytest.raises(ValueError, match="tod"):
        orc.response_tod2map(np.zeros(255), xp, resp, pix)
    with pytest.raises(ValueError, match="response"):
        orc.response_tod2map(np.zeros(256), xp, np.ones((2, 7), F32), pix)
    with pytest.raises(ValueError, match="beyond"):
        orc.Pix(128, 128, np.array([[0, 12288], [24576, 40960]]), 64, 4096, False)
