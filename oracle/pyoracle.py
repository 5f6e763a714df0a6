"""ctypes wrapper of the oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker / CPU baseline. The product never imports it.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
# ORACLE_LIB selects another build of the same library (the AddressSanitizer one: make asan).
_LIB_OVERRIDE = os.environ.get("ORACLE_LIB")

_lib = None

_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_vp = ctypes.c_void_p


class _Pix(ctypes.Structure):
    _fields_ = [("ny", _int), ("nx", _int), ("periodic", _int), ("cell_offsets", _vp),
                ("nycells", _int), ("nxcells", _int), ("ystride", _i64), ("polstride", _i64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if _LIB_OVERRIDE:
            L = ctypes.CDLL(os.path.abspath(_LIB_OVERRIDE))
        else:
            if not os.path.exists(LIB_PATH):
                build()
            L = ctypes.CDLL(LIB_PATH)
        L.orc_locate_f32.argtypes = [ctypes.c_float, ctypes.c_float, _int, _int, _int, _vp, _vp]
        L.orc_locate_f32.restype = _u32
        L.orc_locate_f64.argtypes = [ctypes.c_double, ctypes.c_double, _int, _int, _int, _vp, _vp]
        L.orc_locate_f64.restype = _u32
        for name in ("orc_map2tod_f32", "orc_map2tod_f64"):
            getattr(L, name).argtypes = [_vp, _vp, _vp, _i64, ctypes.POINTER(_Pix), _int, _int]
            getattr(L, name).restype = _u32
        for name in ("orc_tod2map_f32", "orc_tod2map_f64"):
            getattr(L, name).argtypes = [_vp, _vp, _vp, _i64, ctypes.POINTER(_Pix), _int]
            getattr(L, name).restype = _u32
        L.orc_response_map2tod_f64.argtypes = [_vp, _vp, _vp, _vp, _i64, _i64, ctypes.POINTER(_Pix), _int, _int]
        L.orc_response_map2tod_f64.restype = _u32
        L.orc_response_tod2map_f64.argtypes = [_vp, _vp, _vp, _vp, _i64, _i64, ctypes.POINTER(_Pix), _int]
        L.orc_response_tod2map_f64.restype = _u32
        L.orc_map2tod_f32_mt.argtypes = [_vp, _vp, _vp, _i64, ctypes.POINTER(_Pix), _int, _int, _int]
        L.orc_map2tod_f32_mt.restype = _u32
        L.orc_tod2map_f32_mt.argtypes = [_vp, _i64, _vp, _vp, _i64, ctypes.POINTER(_Pix), _int, _int]
        L.orc_tod2map_f32_mt.restype = _u32
        L.orc_max_threads.restype = _int
        L.orc_preplan_params.argtypes = [_i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        L.orc_pointing_params.argtypes = [_i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        L.orc_plan_count.argtypes = [_vp, _i64, _int, _int, _int, _vp, ctypes.POINTER(_i64)]
        L.orc_plan_count.restype = _u32
        L.orc_plan_build.argtypes = [_vp, _i64, _int, _int, _int, _vp, _i64]
        L.orc_plan_build.restype = _u32
        L.orc_lpix_check.argtypes = [_i64, _i64, _vp, _i64, _i64, _i64, _i64, ctypes.POINTER(_i64)]
        L.orc_lpix_check.restype = _int
        _lib = L
    return _lib


BAD_Y, BAD_X, NOT_IN_PIX = 1, 2, 8


class Pix:
    """Local pixelization (cell_offsets int64[nycells][nxcells], ystride, polstride)."""

    def __init__(self, ny, nx, cell_offsets, ystride, polstride, periodic):
        self.ny, self.nx, self.periodic = int(ny), int(nx), bool(periodic)
        self.cell_offsets = np.ascontiguousarray(cell_offsets, dtype=np.int64)
        self.ystride, self.polstride = int(ystride), int(polstride)
        self.npix = 4096 * int(np.sum(self.cell_offsets >= 0))
        present = self.cell_offsets[self.cell_offsets >= 0]
        if present.size and int(present.max()) + 2 * self.polstride + 63 * self.ystride + 63 >= 3 * self.npix:
            # The C functions index the map through these offsets without bounds checks.
            raise ValueError("Pix: cell offsets reach beyond the 3*npix map")

    @classmethod
    def rectangle(cls, ny, nx, periodic):
        nyc, nxc = (ny + 63) // 64, (nx + 63) // 64
        off = 64 * np.arange(nxc, dtype=np.int64)[None, :] + 4096 * nxc * np.arange(nyc, dtype=np.int64)[:, None]
        return cls(ny, nx, off, 64 * nxc, 4096 * nyc * nxc, periodic)

    def c(self):
        return _Pix(self.ny, self.nx, int(self.periodic), self.cell_offsets.ctypes.data,
                    self.cell_offsets.shape[0], self.cell_offsets.shape[1], self.ystride, self.polstride)


def _ptr(a):
    return a.ctypes.data


def _check_len(name, a, n):
    """The C loops index caller arrays without bounds checks: every length is checked here."""
    if a.size != n:
        raise ValueError(f"pyoracle: {name} has {a.size} elements, expected {n}")


def _check_map(lmap, pix):
    _check_len("lmap", lmap, 3 * pix.npix)


def locate(y, x, ny, nx, periodic, f64=False):
    out = np.zeros(4, dtype=np.int32)
    d = np.zeros(2, dtype=np.float64 if f64 else np.float32)
    fn = lib().orc_locate_f64 if f64 else lib().orc_locate_f32
    err = fn(y, x, ny, nx, int(periodic), _ptr(out), _ptr(d))
    return int(err), out, d


def map2tod(lmap, xp, pix, partial=False, tod=None, accum=False, f64=False, nthreads=None):
    xp = np.ascontiguousarray(xp, dtype=np.float32).reshape(3, -1)
    n = xp.shape[1]
    dt = np.float64 if f64 else np.float32
    lmap = np.ascontiguousarray(lmap, dtype=dt).reshape(-1)
    out = np.zeros(n, dtype=dt) if tod is None else np.array(tod, dtype=dt).reshape(-1)
    _check_map(lmap, pix)
    _check_len("tod", out, n)
    p = pix.c()
    if nthreads is not None:
        assert not f64
        err = lib().orc_map2tod_f32_mt(_ptr(out), _ptr(lmap), _ptr(xp), n, ctypes.byref(p), int(partial), int(accum), int(nthreads))
    elif f64:
        err = lib().orc_map2tod_f64(_ptr(out), _ptr(lmap), _ptr(xp), n, ctypes.byref(p), int(partial), int(accum))
    else:
        err = lib().orc_map2tod_f32(_ptr(out), _ptr(lmap), _ptr(xp), n, ctypes.byref(p), int(partial), int(accum))
    return out, int(err)


def tod2map(tod, xp, pix, lmap=None, partial=False, f64=False, nthreads=None):
    xp = np.ascontiguousarray(xp, dtype=np.float32).reshape(3, -1)
    n = xp.shape[1]
    dt = np.float64 if f64 else np.float32
    tod = np.ascontiguousarray(tod, dtype=dt).reshape(-1)
    out = np.zeros(3 * pix.npix, dtype=dt) if lmap is None else np.array(lmap, dtype=dt).reshape(-1)
    _check_len("tod", tod, n)
    _check_map(out, pix)
    p = pix.c()
    if nthreads is not None:
        assert not f64
        err = lib().orc_tod2map_f32_mt(_ptr(out), out.size, _ptr(tod), _ptr(xp), n, ctypes.byref(p), int(partial), int(nthreads))
    elif f64:
        err = lib().orc_tod2map_f64(_ptr(out), _ptr(tod), _ptr(xp), n, ctypes.byref(p), int(partial))
    else:
        err = lib().orc_tod2map_f32(_ptr(out), _ptr(tod), _ptr(xp), n, ctypes.byref(p), int(partial))
    return out, int(err)


def response_map2tod(lmap, xp, resp, pix, partial=False, tod=None, accum=False):
    """FP64 response map2tod; resp has shape (2, ndet)."""
    xp = np.ascontiguousarray(xp, dtype=np.float32).reshape(3, -1)
    n = xp.shape[1]
    resp = np.ascontiguousarray(resp, dtype=np.float32)
    if not (resp.ndim == 2 and resp.shape[0] == 2 and resp.shape[1] > 0 and n % resp.shape[1] == 0):
        raise ValueError(f"pyoracle: response must have shape (2, ndet) with ndet dividing {n}")
    lmap = np.ascontiguousarray(lmap, dtype=np.float64).reshape(-1)
    out = np.zeros(n) if tod is None else np.array(tod, dtype=np.float64).reshape(-1)
    _check_map(lmap, pix)
    _check_len("tod", out, n)
    p = pix.c()
    err = lib().orc_response_map2tod_f64(_ptr(out), _ptr(lmap), _ptr(xp), _ptr(resp), resp.shape[1], n,
                                         ctypes.byref(p), int(partial), int(accum))
    return out, int(err)


def response_tod2map(tod, xp, resp, pix, lmap=None, partial=False):
    """FP64 response tod2map; resp has shape (2, ndet)."""
    xp = np.ascontiguousarray(xp, dtype=np.float32).reshape(3, -1)
    n = xp.shape[1]
    resp = np.ascontiguousarray(resp, dtype=np.float32)
    if not (resp.ndim == 2 and resp.shape[0] == 2 and resp.shape[1] > 0 and n % resp.shape[1] == 0):
        raise ValueError(f"pyoracle: response must have shape (2, ndet) with ndet dividing {n}")
    tod = np.ascontiguousarray(tod, dtype=np.float64).reshape(-1)
    out = np.zeros(3 * pix.npix) if lmap is None else np.array(lmap, dtype=np.float64).reshape(-1)
    _check_len("tod", tod, n)
    _check_map(out, pix)
    p = pix.c()
    err = lib().orc_response_tod2map_f64(_ptr(out), _ptr(tod), _ptr(xp), _ptr(resp), resp.shape[1], n,
                                         ctypes.byref(p), int(partial))
    return out, int(err)


def max_threads():
    return int(lib().orc_max_threads())


def preplan_params(nsamp):
    a, b = _i64(), _i64()
    lib().orc_preplan_params(nsamp, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def pointing_params(nmt):
    a, b = _i64(), _i64()
    lib().orc_pointing_params(nmt, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def plan_count(xp, ny, nx, periodic):
    xp = np.ascontiguousarray(xp, dtype=np.float32).reshape(3, -1)
    n = xp.shape[1]
    _, nb = preplan_params(n)
    cum = np.zeros(nb, dtype=np.uint32)
    nmt = _i64()
    err = lib().orc_plan_count(_ptr(xp), n, ny, nx, int(periodic), _ptr(cum), ctypes.byref(nmt))
    return cum, nmt.value, int(err)


def plan_build(xp, ny, nx, periodic, nmt=None):
    xp = np.ascontiguousarray(xp, dtype=np.float32).reshape(3, -1)
    if nmt is None:
        _, nmt, _ = plan_count(xp, ny, nx, periodic)
    mt = np.zeros(nmt, dtype=np.uint64)
    err = lib().orc_plan_build(_ptr(xp), xp.shape[1], ny, nx, int(periodic), _ptr(mt), nmt)
    return mt, int(err)


def lpix_check(ny, nx, offsets, ystride, polstride):
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    npix = _i64()
    code = lib().orc_lpix_check(ny, nx, _ptr(off), off.shape[0], off.shape[1], ystride, polstride, ctypes.byref(npix))
    return int(code), npix.value
