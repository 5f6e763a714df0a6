"""CPU-side tests of the product: the C-ABI library loads and exports every declared symbol,
the host parts (plan='reference' operators, pixelization checks, ToyPointing, random plans)
agree with the oracle, and argument errors surface as RuntimeError with the reference's texts."""

import ctypes
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F32 = np.float32


def declared_symbols():
    text = open(os.path.join(REPO, "include", "gpu_mm_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(gmm_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (load torch's HIP runtime first)
    from gpu_mm_amd import _native
    lib = ctypes.CDLL(_native.lib_path())
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.gmm_version.restype = ctypes.c_int
    assert lib.gmm_version() >= 100


def test_stream_read_argument_checks():
    """The bandwidth probe's argument checks run before any GPU call: a length that is not a
    multiple of 4, a source that is not 16-byte aligned and null pointers return an error with a
    message; n = 0 launches nothing."""
    import torch  # noqa: F401
    from gpu_mm_amd import _native
    lib = ctypes.CDLL(_native.lib_path())
    lib.gmm_stream_read.restype = ctypes.c_int
    lib.gmm_stream_read.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    lib.gmm_last_error.restype = ctypes.c_char_p
    # Host addresses stand in for device pointers: none of these calls reaches the GPU.
    buf = (ctypes.c_float * 16)()
    base = (ctypes.addressof(buf) + 15) & ~15
    acc = ctypes.addressof(buf)
    for src, n, sum_, text in ((base, 6, acc, b"multiple of 4"), (base, -4, acc, b"multiple of 4"),
                               (base + 4, 8, acc, b"16-byte aligned"), (None, 8, acc, b"src"),
                               (base, 8, None, b"sum")):
        assert lib.gmm_stream_read(src, n, sum_, None) != 0
        assert text in lib.gmm_last_error()
    assert lib.gmm_stream_read(base, 0, acc, None) == 0


def test_extension_imports(gmm):
    assert gmm.PointingPrePlan.preplan_size == 1024
    assert gmm.gpu_mm._ext._get_tsize() == 4


def rect_pix(gmm, orc, ny, nx, periodic):
    return gmm.LocalPixelization.make_rectangle(ny, nx, periodic), orc.Pix.rectangle(ny, nx, periodic)


# n = 32 * 256 - 5: nsamp % 32 != 0 (accepted here; the reference requires a multiple of 32).
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("accum", [False, True])
@pytest.mark.parametrize("n", [32 * 256, 32 * 256 - 5])
def test_reference_map2tod_matches_oracle(gmm, orc, periodic, accum, n):
    rng = np.random.default_rng(20 + 2 * periodic + accum)
    ny, nx = 200, 300
    lp, op = rect_pix(gmm, orc, ny, nx, periodic)
    xp = np.stack([rng.uniform(0, ny - 1, n), rng.uniform(-nx, 2 * nx, n) if periodic else rng.uniform(0, nx - 2, n),
                   rng.uniform(-3, 3, n)]).astype(F32)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    tod0 = rng.standard_normal(n).astype(F32)
    tod = tod0.copy()
    gmm.map2tod(tod, gmm.LocalMap(lp, m), xp, plan='reference', accum=accum)
    ref, err = orc.map2tod(m, xp, op, tod=tod0, accum=accum)
    assert err == 0
    # Same operation order in both: equal to the last bit up to sincosf and the compiler.
    assert np.max(np.abs(tod - ref)) <= 1e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("periodic", [False, True])
def test_reference_tod2map_matches_oracle(gmm, orc, periodic):
    rng = np.random.default_rng(30 + periodic)
    ny, nx = 200, 300
    lp, op = rect_pix(gmm, orc, ny, nx, periodic)
    n = 32 * 256
    xp = np.stack([rng.uniform(0, ny - 1, n), rng.uniform(-nx, 2 * nx, n) if periodic else rng.uniform(0, nx - 2, n),
                   rng.uniform(-3, 3, n)]).astype(F32)
    tod = rng.standard_normal((8, n // 8)).astype(F32)
    m0 = rng.standard_normal(3 * lp.npix).astype(F32)
    m = m0.copy()
    gmm.tod2map(gmm.LocalMap(lp, m), tod, xp.reshape(3, 8, n // 8), plan='reference')
    ref, err = orc.tod2map(tod, xp, op, lmap=m0)
    assert err == 0
    assert np.max(np.abs(m - ref)) <= 1e-6 * np.max(np.abs(ref))


def test_reference_errors(gmm):
    lp = gmm.LocalPixelization.make_rectangle(64, 64, False)
    m = np.zeros(3 * lp.npix, dtype=F32)
    xp = np.zeros((3, 32), dtype=F32)
    tod = np.zeros(32, dtype=F32)
    xp[0, 5] = 70.0
    with pytest.raises(RuntimeError, match="y-value is outside range"):
        gmm.map2tod(tod, gmm.LocalMap(lp, m), xp, plan='reference')
    xp[0, 5] = 3.0
    xp[1, 7] = 63.5
    with pytest.raises(RuntimeError, match="periodic_xcoord"):
        gmm.tod2map(gmm.LocalMap(lp, m), tod, xp, plan='reference')
    with pytest.raises(RuntimeError, match="expected nsamp > 0"):
        gmm.map2tod(np.zeros(0, F32), gmm.LocalMap(lp, m), np.zeros((3, 0), F32), plan='reference')
    with pytest.raises(RuntimeError, match="dtype float32"):
        gmm.map2tod(np.zeros(32, np.float64), gmm.LocalMap(lp, m), np.zeros((3, 32), F32), plan='reference')
    with pytest.raises(RuntimeError, match="Bad 'plan'"):
        gmm.map2tod(tod, gmm.LocalMap(lp, m), np.zeros((3, 32), F32), plan='bogus')


def test_reference_partial_pixelization(gmm, orc):
    rng = np.random.default_rng(40)
    ny, nx = 128, 128
    cells = [(0, 0), (1, 0), (1, 1)]
    lp = gmm.LocalPixelization.make_cells(ny, nx, cells, periodic_xcoord=False)
    op = orc.Pix(ny, nx, lp.cell_offsets_cpu, 64, 4096, False)
    n = 32 * 64
    xp = np.stack([rng.uniform(0, ny - 1, n), rng.uniform(0, nx - 2, n), rng.uniform(0, 3, n)]).astype(F32)
    m = rng.standard_normal(3 * lp.npix).astype(F32)
    tod = np.zeros(n, F32)
    with pytest.raises(RuntimeError, match="not in LocalPixelization"):
        gmm.map2tod(tod, gmm.LocalMap(lp, m), xp, plan='reference')
    gmm.map2tod(tod, gmm.LocalMap(lp, m), xp, plan='reference', partial_pixelization=True)
    ref, err = orc.map2tod(m, xp, op, partial=True)
    assert err == 0
    assert np.max(np.abs(tod - ref)) <= 1e-6 * np.max(np.abs(ref))


def test_lpix_validation_matches_oracle(gmm, orc):
    cases = [
        (128, 128, np.array([[0, 12288], [-1, 24576]]), 64, 4096),
        (128, 128, np.array([[0, 12288], [-1, 24576]]), 32, 4096),
        (128, 128, np.array([[0, 12288], [-1, 24576]]), 64, 4000),
        (128, 128, np.array([[0, 12288], [-1, 24576]]), 192, 320),
        (128, 128, np.array([[0, 12290], [-1, 24576]]), 64, 4096),
        (128, 128, np.array([[0, 12288], [-1, 36864]]), 64, 4096),
        (0, 128, np.array([[0, 12288], [-1, 24576]]), 64, 4096),
    ]
    for ny, nx, off, ys, ps in cases:
        code, npix = orc.lpix_check(ny, nx, off, ys, ps)
        if code == 0:
            lp = gmm.LocalPixelization(ny, nx, off, ys, ps)
            assert lp.npix == npix
        else:
            with pytest.raises(RuntimeError):
                gmm.LocalPixelization(ny, nx, off, ys, ps)


def test_make_rectangle_and_to_global(gmm):
    lp = gmm.LocalPixelization.make_rectangle(100, 130, periodic_xcoord=False)
    assert (lp.nycells, lp.nxcells, lp.ystride, lp.polstride) == (2, 3, 192, 64 * 64 * 6)
    assert lp.is_simple_rectangle()
    arr = np.arange(3 * lp.npix, dtype=F32)
    g = gmm.LocalMap(lp, arr).to_global()
    assert g.shape == (3, 100, 130)
    assert np.array_equal(g, arr.reshape(3, 128, 192)[:, :100, :130])
    assert lp == gmm.LocalPixelization.make_rectangle(100, 130, periodic_xcoord=False)
    assert not (lp == gmm.LocalPixelization.make_rectangle(100, 130, periodic_xcoord=True))


def test_toy_pointing(gmm):
    tp = gmm.ToyPointing(4, 1024, 100, 300, 0.4, 50.0, noisy=False)
    xp = tp.xpointing_cpu
    assert xp.shape == (3, 4, 1024)
    assert tp.drift_speed == pytest.approx(50.0 / (9 * 4096))
    y = xp[0].reshape(-1).astype(np.float64)
    # 45-degree bounce between 1.1*v and ny-2-1.1*v, speed 0.4 px/sample
    assert np.all(np.abs(np.diff(y)) <= 0.4 + 1e-5)
    assert y.min() >= 0.0 and y.max() <= 98.0
    assert np.all(xp[2] == 1.0)


def test_random_plan_is_sorted_and_seeded(gmm):
    a = gmm.make_random_plan_mt(50, 1, 20, seed=3)
    b = gmm.make_random_plan_mt(50, 1, 20, seed=3)
    assert np.array_equal(a, b)
    cells = a & np.uint64(0xFFFFF)
    assert np.all(cells[1:] >= cells[:-1])
    assert len(np.unique(cells)) == 50
