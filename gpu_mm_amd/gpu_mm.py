"""
gpu_mm_amd: planned pointing operators for CMB map-making on AMD MI355X (gfx950).

Drop-in for the pointing-operator part of kmsmith137/gpu_mm (gpu_mm/gpu_mm.py). The data model
is the reference's; only the array library changes: GPU arrays are torch tensors (on the ROCm
device), CPU arrays are numpy arrays.

  map2tod(tod, local_map, xpointing, plan, partial_pixelization=False, response=None, debug=False, accum=False)
  tod2map(local_map, tod, xpointing, plan, partial_pixelization=False, response=None, debug=False)

map2tod() overwrites its output TOD (or adds to it, accum=True); tod2map() always adds into
the existing map.

Global maps are float32[3][nypix_global][nxpix_global] with components {T,Q,U}, indexed
map[y][x]. A timestream is float32[nsamp] or float32[ndet][nt]. Pointing is "exploded":
xpointing = float32[3][nsamp] (or [3][ndet][nt]) with rows {y, x, alpha}; detectors measure
T + cos(2 alpha) Q + sin(2 alpha) U. Valid pointing: 0 <= y <= nypix_global-1, and
0 <= x <= nxpix_global-2 (periodic_xcoord=False) or -nxpix_global <= x <= 2*nxpix_global
(periodic_xcoord=True).

A local map stores a set of 64x64 "cells" of the global map; element (p, iy, ix) of cell
(iycell, ixcell) lives at cell_offsets[iycell, ixcell] + p*polstride + iy*ystride + ix, and a
negative cell offset means "cell not stored". (Reference: gpu_mm/gpu_mm.py:1-240.)

Plans: PointingPrePlan(xpointing, ...) counts the (32-sample line, cell) pairs; PointingPlan
(preplan, xpointing) lists them, sorted by cell, together with the per-block work list used by
the planned kernels. A preplan is small and can be kept per TOD; a plan is large and can be
rebuilt on the fly.

The 'plan' argument of map2tod/tod2map selects the implementation:
  PointingPlan  -> planned HIP kernels (cells staged in LDS)
  None          -> plan-free HIP kernels (global gathers / atomics)
  'reference'   -> single-threaded CPU implementation (host arrays), for testing
"""

import contextlib

import numpy as np
import torch

from . import _native

_ext = _native.ext()

# The planned and unplanned operators check the GPU error flags synchronously after every
# call (as the reference does). set_errflag_check(False) makes planned calls asynchronous;
# errors are then accumulated on the plan and raised by PointingPlan.check_errflags().
# Plan-free calls (plan=None) always check synchronously.
_errflag_check = True

mm_dtype = np.float32


def set_errflag_check(flag):
    """Enable (default) or disable the synchronous error-flag check after each GPU operator."""
    global _errflag_check
    _errflag_check = bool(flag)


def get_errflag_check():
    return _errflag_check


@contextlib.contextmanager
def deferred_errflags():
    """
    Context manager: inside the block the GPU operators do not block on their error flags (as
    after set_errflag_check(False)); errors accumulate on the plans and are raised by
    PointingPlan.check_errflags(). The previous setting is restored on exit.
    """
    global _errflag_check
    old = _errflag_check
    _errflag_check = False
    try:
        yield
    finally:
        _errflag_check = old


def _stream():
    if torch.cuda.is_available():
        return torch.cuda.current_stream().cuda_stream
    return 0


def _device():
    return torch.device("cuda", torch.cuda.current_device())


def _is_gpu_array(a):
    return isinstance(a, torch.Tensor) and a.is_cuda


def _is_tod16(tod):
    return isinstance(tod, torch.Tensor) and tod.dtype in (torch.float16, torch.bfloat16)


def encode_tod16(tod, dtype=torch.float16, error_feedback=0):
    """
    Encodes a float32 GPU TOD of shape (nsamp,) or (ndet, nt) as a 16-bit TOD (float16 or
    bfloat16) with one float32 scale per row: scale = max(max|row| / (0.5 * largest finite
    value of dtype), FLT_MIN), tod16 = tod / scale rounded to nearest even.

    error_feedback = 1 or 2 carries the rounding error of each sample (first or second order)
    into the next sample of its row before rounding: the TOD's encoding error becomes the first
    (second) difference of the rounding errors, which tod2map's smoothly varying weights average
    out, so maps built from the encoded TOD lose about half of the encoding error (the TOD's own
    per-sample error grows). 0 rounds every sample independently.

    Returns (tod16, scale); decode with tod16.float() * scale[:, None] (or * scale for 1-d).
    Not in the reference.
    """
    if error_feedback not in (0, 1, 2):
        raise RuntimeError(f"encode_tod16(): error_feedback must be 0, 1 or 2, got {error_feedback!r}")
    if dtype not in (torch.float16, torch.bfloat16):
        raise RuntimeError(f"encode_tod16(): dtype must be torch.float16 or torch.bfloat16, got {dtype}")
    nrows = tod.shape[0] if tod.ndim == 2 else 1
    out = torch.empty(tod.shape, dtype=dtype, device=tod.device)
    scale = torch.empty(nrows, dtype=torch.float32, device=tod.device)
    _ext.tod16_encode(out, scale, tod, error_feedback=int(error_feedback), stream=_stream())
    return out, scale


def kernel_timer_enable(enable=True):
    """Starts (or stops) recording HIP events around every planned map2tod/tod2map kernel."""
    _ext.kernel_timer_enable(bool(enable))


def kernel_timer_read(kernel):
    """(total_ms, launches) of 'map2tod' or 'tod2map' kernels since the last enable/read."""
    return _ext.kernel_timer_read({"map2tod": 0, "tod2map": 1}[kernel])


def stream_read(src, out):
    """Bandwidth probe (a measurement aid): reads the float32 GPU tensor 'src' (contiguous,
    numel % 4 == 0, 16-byte aligned) once with non-temporal 16-byte loads, the access of the
    planned operators' input streams, and adds its sum to out[0]. Asynchronous on torch's
    current stream."""
    _ext.stream_read(src, out, stream=_stream())


def map2tod(tod, local_map, xpointing, plan, partial_pixelization=False, response=None, debug=False, accum=False,
            tod_scale=None):
    """
    tod (=, or += if accum) A * local_map.

      - tod: float32 array of shape (nsamp,) or (ndet, nt); output. With a PointingPlan it may
        also be a float16 / bfloat16 GPU tensor that stores value / tod_scale[row] (see
        encode_tod16); then tod_scale (float32, one entry per row) is required.
      - local_map: LocalMap.
      - xpointing: float32 array of shape (3, nsamp) or (3, ndet, nt).
      - plan: PointingPlan, None (plan-free GPU kernels) or 'reference' (CPU, host arrays).
      - partial_pixelization: if True, pointing outside the local map reads zeros instead of
        raising.
      - response: optional float32 array of shape (2, ndet) on the GPU (PointingPlan only): row 0
        multiplies T, row 1 multiplies Q and U, for every sample of detector i = s // (nsamp/ndet).
      - accum: add into 'tod' instead of overwriting it.

    Reference: gpu_mm/gpu_mm.py:262-312.
    """
    if not isinstance(local_map, LocalMap):
        raise RuntimeError(f"map2tod(): expected LocalMap, got {local_map!r}")
    lpix = local_map.pixelization
    larr = local_map.arr
    if (response is not None) and not isinstance(plan, PointingPlan):
        # The reference silently ignores 'response' here (gpu_mm.py:305-310); an error is safer
        # than a result without the response applied.
        raise RuntimeError("map2tod(): 'response' requires a PointingPlan")

    if _is_tod16(tod) or (tod_scale is not None):
        if not (isinstance(plan, PointingPlan) and _is_tod16(tod) and tod_scale is not None):
            raise RuntimeError("map2tod(): a 16-bit TOD needs a PointingPlan and tod_scale")
        _ext.planned_map2tod_q16(tod, tod_scale, larr, xpointing, lpix, plan, partial_pixelization, debug, accum,
                                 response=response, check_errors=_errflag_check, stream=_stream())
    elif isinstance(plan, PointingPlan):
        if response is None:
            _ext.planned_map2tod(tod, larr, xpointing, lpix, plan, partial_pixelization, debug, accum,
                                 check_errors=_errflag_check, stream=_stream())
        else:
            _ext.response_map2tod(tod, larr, xpointing, response, lpix, plan, partial_pixelization, debug,
                                  accum, check_errors=_errflag_check, stream=_stream())
    elif plan is None:
        # Plan-free calls always check their error flags synchronously: there is no plan to
        # accumulate them on.
        _ext.unplanned_map2tod(tod, larr, xpointing, lpix, _unplanned_errflags(), partial_pixelization, accum,
                               check_errors=True, stream=_stream())
    elif isinstance(plan, str) and plan == 'reference':
        _ext.reference_map2tod(tod, larr, xpointing, lpix, partial_pixelization, accum)
    else:
        raise RuntimeError(f"Bad 'plan' argument to map2tod(): {plan}")


def tod2map(local_map, tod, xpointing, plan, partial_pixelization=False, response=None, debug=False,
            tod_scale=None):
    """
    local_map += A^T * tod.

    Arguments as in map2tod() (including 16-bit TODs with tod_scale); the output map is always
    accumulated into.
    Reference: gpu_mm/gpu_mm.py:314-385.
    """
    if isinstance(local_map, DynamicMap):
        if not isinstance(plan, PointingPlan):
            raise RuntimeError("tod2map(): if output is a DynamicMap, then the 'plan' argument"
                               + f" must be a PointingPlan (got: {plan})'")
        if partial_pixelization != local_map.have_cell_mask:
            raise RuntimeError(f"tod2map(): {partial_pixelization=} was specified,"
                               + f" but DynamicMap.have_cell_mask={local_map.have_cell_mask}"
                               + " (expected these to be equal)")
        local_map.expand(plan)
        lpix, larr = local_map._target()
    elif isinstance(local_map, LocalMap):
        lpix = local_map.pixelization
        larr = local_map.arr
    else:
        raise RuntimeError(f"tod2map(): Bad 'local_map' argument to tod2map(): expected LocalMap or DynamicMap, got: {local_map}")
    if (response is not None) and not isinstance(plan, PointingPlan):
        raise RuntimeError("tod2map(): 'response' requires a PointingPlan")

    if _is_tod16(tod) or (tod_scale is not None):
        if not (isinstance(plan, PointingPlan) and _is_tod16(tod) and tod_scale is not None):
            raise RuntimeError("tod2map(): a 16-bit TOD needs a PointingPlan and tod_scale")
        _ext.planned_tod2map_q16(larr, tod, tod_scale, xpointing, lpix, plan, partial_pixelization, debug,
                                 response=response, check_errors=_errflag_check, stream=_stream())
    elif isinstance(plan, PointingPlan):
        if response is None:
            _ext.planned_tod2map(larr, tod, xpointing, lpix, plan, partial_pixelization, debug,
                                 check_errors=_errflag_check, stream=_stream())
        else:
            _ext.response_tod2map(larr, tod, xpointing, response, lpix, plan, partial_pixelization, debug,
                                  check_errors=_errflag_check, stream=_stream())
    elif plan is None:
        _ext.unplanned_tod2map(larr, tod, xpointing, lpix, _unplanned_errflags(), partial_pixelization,
                               check_errors=True, stream=_stream())
    elif isinstance(plan, str) and plan == 'reference':
        _ext.reference_tod2map(larr, tod, xpointing, lpix, partial_pixelization)
    else:
        raise RuntimeError(f"Bad 'plan' argument to tod2map(): {plan}")


# Block budget of the plan-free kernels (gpu_mm.py:306 uses 2048 as well). Every block of a
# plan-free launch writes its own error-flag word, so the buffer needs no clearing.
_UNPLANNED_MAX_BLOCKS = 2048


def _unplanned_errflags():
    return torch.empty(_UNPLANNED_MAX_BLOCKS, dtype=torch.uint32, device=_device())


####################################################################################################


class LocalPixelization(_ext.LocalPixelization):
    """
    Global pixelization (nypix_global, nxpix_global, periodic_xcoord) plus the set of 64x64
    cells held locally and their memory layout (cell_offsets, ystride, polstride).

    Members: nypix_global, nxpix_global, periodic_xcoord, cell_offsets_cpu (numpy int64),
    cell_offsets_gpu (torch int64 on the GPU, when a GPU is present), ystride, polstride,
    nycells, nxcells, npix (local pixels, without the factor 3).

    Reference: gpu_mm/gpu_mm.py:421-526; checks from src_lib/LocalPixelization.cu:33-101.
    """

    def __init__(self, nypix_global, nxpix_global, cell_offsets, ystride, polstride, periodic_xcoord=True):
        if isinstance(cell_offsets, torch.Tensor):
            cpu = cell_offsets.detach().to("cpu").numpy()
        else:
            cpu = np.asarray(cell_offsets)
        self.cell_offsets_cpu = np.ascontiguousarray(cpu, dtype=np.int64)
        if torch.cuda.is_available():
            self.cell_offsets_gpu = torch.from_numpy(self.cell_offsets_cpu).to(_device())
        else:
            self.cell_offsets_gpu = None
        _ext.LocalPixelization.__init__(self, int(nypix_global), int(nxpix_global), self.cell_offsets_cpu,
                                        self.cell_offsets_gpu, int(ystride), int(polstride), bool(periodic_xcoord))

    def is_simple_rectangle(self):
        """True if the local map is a plain contiguous (3, 64*nycells, 64*nxcells) array."""
        rect = self._make_rectangular_cell_offsets(self.nycells, self.nxcells)
        return (self.ystride == 64 * self.nxcells) and (self.polstride == 64 * 64 * self.nycells * self.nxcells) \
            and np.array_equal(self.cell_offsets_cpu, rect)

    def __eq__(self, x):
        if not isinstance(x, LocalPixelization):
            return NotImplemented
        if self is x:
            return True
        if (self.nypix_global, self.nxpix_global) != (x.nypix_global, x.nxpix_global):
            return False
        if (self.periodic_xcoord, self.ystride, self.polstride) != (x.periodic_xcoord, x.ystride, x.polstride):
            return False
        if (self.nycells, self.nxcells, self.npix) != (x.nycells, x.nxcells, x.npix):
            return False
        return bool(np.all(np.maximum(self.cell_offsets_cpu, -1) == np.maximum(x.cell_offsets_cpu, -1)))

    __hash__ = None

    @classmethod
    def _make_rectangular_cell_offsets(cls, nycells, nxcells):
        off = np.empty((nycells, nxcells), dtype=np.int64)
        off[:, :] = 64 * np.arange(nxcells, dtype=np.int64).reshape((1, -1))
        off[:, :] += 64 * 64 * nxcells * np.arange(nycells, dtype=np.int64).reshape((-1, 1))
        return off

    @classmethod
    def make_rectangle(cls, nypix_global, nxpix_global, periodic_xcoord=True):
        """
        LocalPixelization covering the whole global map, stored as a contiguous array of
        shape (3, nypix_padded, nxpix_padded), with n_padded = n rounded up to a multiple of 64.
        """
        assert nypix_global > 0
        assert nxpix_global > 0
        nycells = (nypix_global + 63) // 64
        nxcells = (nxpix_global + 63) // 64
        return LocalPixelization(nypix_global, nxpix_global,
                                 cell_offsets=cls._make_rectangular_cell_offsets(nycells, nxcells),
                                 ystride=64 * nxcells, polstride=64 * 64 * nycells * nxcells,
                                 periodic_xcoord=periodic_xcoord)

    @classmethod
    def make_cells(cls, nypix_global, nxpix_global, cells, periodic_xcoord=True):
        """
        LocalPixelization holding the listed cells (an iterable of (iycell, ixcell)) as a
        float32[ncells][3][64][64] array, in the order given (ystride=64, polstride=4096).
        """
        nycells = (nypix_global + 63) // 64
        nxcells = (nxpix_global + 63) // 64
        off = np.full((nycells, nxcells), -1, dtype=np.int64)
        for i, (cy, cx) in enumerate(cells):
            if off[cy, cx] >= 0:
                raise RuntimeError(f"LocalPixelization.make_cells(): cell {(cy, cx)} listed twice")
            off[cy, cx] = 3 * 64 * 64 * i
        return LocalPixelization(nypix_global, nxpix_global, off, ystride=64, polstride=64 * 64,
                                 periodic_xcoord=periodic_xcoord)


class LocalMap:
    """
    A LocalPixelization plus the array that holds the map: a float32, C-contiguous numpy
    array (CPU) or torch tensor (GPU) with 3*npix elements. Reference: gpu_mm/gpu_mm.py:529-580.
    """

    def __init__(self, pixelization, arr):
        assert isinstance(pixelization, LocalPixelization)
        if isinstance(arr, torch.Tensor):
            assert arr.dtype == torch.float32, "LocalMap: expected float32 array"
            assert arr.is_contiguous(), "LocalMap: expected contiguous array"
            self.on_gpu = arr.is_cuda
            self.dtype = np.float32
        else:
            assert isinstance(arr, np.ndarray), "LocalMap: expected numpy array or torch tensor"
            assert arr.dtype == np.float32, "LocalMap: expected float32 array"
            assert arr.flags['C_CONTIGUOUS'], "LocalMap: expected C-contiguous array"
            self.on_gpu = False
            self.dtype = arr.dtype

        self.pixelization = pixelization
        self.nypix_global = pixelization.nypix_global
        self.nxpix_global = pixelization.nxpix_global
        self.periodic_xcoord = pixelization.periodic_xcoord
        self.npix = pixelization.npix
        self.arr = arr

    def to_global(self):
        """
        Returns the map as a numpy array of shape (3, nypix_global, nxpix_global); cells that
        are not in the local map are zero. Pixels of boundary cells beyond the global map are
        ignored (no wrapping). Reference: gpu_mm/gpu_mm.py:565-580, src_lib/local_map_to_global.cu.
        """
        p = self.pixelization
        if isinstance(self.arr, torch.Tensor) and self.arr.is_cuda and p.cell_offsets_gpu is not None:
            dst = torch.empty((3, p.nypix_global, p.nxpix_global), dtype=torch.float32, device=self.arr.device)
            _ext.local_map_to_global(p, dst, self.arr, stream=_stream())
            return dst.cpu().numpy()
        if isinstance(self.arr, torch.Tensor):
            src = self.arr.detach().to("cpu").numpy().reshape(-1)
        else:
            src = np.asarray(self.arr).reshape(-1)
        ny, nx = p.nypix_global, p.nxpix_global
        dst = np.zeros((3, ny, nx), dtype=np.float32)
        off = p.cell_offsets_cpu
        iy = np.arange(64).reshape(64, 1)
        ix = np.arange(64).reshape(1, 64)
        for cy in range(p.nycells):
            for cx in range(p.nxcells):
                o = int(off[cy, cx])
                if o < 0:
                    continue
                y0, x0 = 64 * cy, 64 * cx
                if y0 >= ny or x0 >= nx:
                    continue
                hy, hx = min(64, ny - y0), min(64, nx - x0)
                for pol in range(3):
                    idx = o + pol * p.polstride + iy * p.ystride + ix
                    dst[pol, y0:y0 + hy, x0:x0 + hx] = src[idx][:hy, :hx]
        return dst


####################################################################################################


class DynamicMap:
    """
    Local map whose cells are created on demand by tod2map(): it starts empty, each
    tod2map(dmap, tod, xpointing, plan) call adds (zero-filled) the cells the plan touches, and
    finalize() hands the result over as a LocalMap with layout float32[ncells][3][64][64]
    (ystride 64, polstride 4096; cell i at offset 3*64*64*i of the array).

    cell_mask (bool, shape (nycells, nxcells)) limits the cells that may be created to a target
    region; tod2map() then needs partial_pixelization=True, and samples outside the region
    deposit nothing.

    How it is held here:
      - the cell store: one float32 GPU tensor of shape (capacity, 3, 64, 64), whose first
        ncells_curr cells are the map; when a plan needs more cells the capacity doubles (or
        jumps to what is needed), and the filled cells move over;
      - the cell table: an int64 GPU array (nycells, nxcells) inside a LocalPixelization: the
        store index times 3*64*64 for present cells, -1 for cells that may still be created,
        -2 for cells outside the target region. It is updated on the GPU (cell creation is a
        device kernel, numbering new cells in cell-key order, so layouts are reproducible); its
        host copy is brought up to date only when the layout is read (permute, finalize);
      - a lifecycle: "open" (cells may be added) -> "fixed" (after permute: the layout no
        longer changes) -> "done" (after finalize: the map belongs to the returned LocalMap).

    Only tod2map() accepts a DynamicMap. API (constructor, expand, permute, finalize,
    randomly_permute, have_cell_mask): gpu_mm/gpu_mm.py:586-806; cell creation:
    src_lib/expand_dynamic_map.cu.
    """

    _CELL_FLOATS = 3 * 64 * 64

    def __init__(self, nypix_global, nxpix_global, dtype, cell_mask=None, periodic_xcoord=True, initial_ncells=1024):
        if not (0 < nypix_global <= 64 * 1024 and 0 < nxpix_global <= 64 * 1024):
            raise RuntimeError(f"DynamicMap: expected 0 < nypix_global, nxpix_global <= 65536, got"
                               f" {(nypix_global, nxpix_global)}")
        if dtype not in (np.float32, torch.float32, 'float32'):
            raise RuntimeError(f"DynamicMap: only float32 maps are supported (got {dtype})")
        if initial_ncells <= 0:
            raise RuntimeError(f"DynamicMap: expected initial_ncells > 0, got {initial_ncells}")

        self.nypix_global, self.nxpix_global = int(nypix_global), int(nxpix_global)
        self.periodic_xcoord = bool(periodic_xcoord)
        self.dtype = np.float32
        shape = ((self.nypix_global + 63) // 64, (self.nxpix_global + 63) // 64)

        if cell_mask is None:
            self._mask = None
            table = np.full(shape, -1, dtype=np.int64)
        else:
            self._mask = np.asarray(cell_mask)
            if self._mask.dtype != bool or self._mask.shape != shape:
                raise RuntimeError(f"DynamicMap: cell_mask must be a bool array of shape {shape}")
            table = np.where(self._mask, -1, -2).astype(np.int64)

        self._table = LocalPixelization(self.nypix_global, self.nxpix_global, table, ystride=64,
                                        polstride=64 * 64, periodic_xcoord=self.periodic_xcoord)
        self._store = torch.zeros((int(initial_ncells), 3, 64, 64), dtype=torch.float32, device=_device())
        self._count = torch.zeros(1, dtype=torch.uint32, device=_device())   # cells in use (device side)
        self._ncells = 0
        self._state = "open"

    # Read-only state (names as in the reference).
    have_cell_mask = property(lambda self: self._mask is not None)
    ncells_curr = property(lambda self: self._ncells)
    ncells_allocated = property(lambda self: 0 if self._store is None else int(self._store.shape[0]))
    pixelization_is_finalized = property(lambda self: self._state != "open")
    map_is_finalized = property(lambda self: self._state == "done")

    def _target(self):
        """(LocalPixelization, flat float32 map) for tod2map(); valid until the next expansion."""
        return self._table, self._store[:self._ncells].view(-1)

    def _reserve(self, ncells):
        cap = int(self._store.shape[0])
        if ncells <= cap:
            return
        grown = torch.zeros((max(ncells, 2 * cap), 3, 64, 64), dtype=torch.float32, device=self._store.device)
        grown[:self._ncells].copy_(self._store[:self._ncells])
        self._store = grown

    def expand(self, plan):
        """Creates (zero-filled) every cell the plan touches that the map does not hold yet."""
        if self._state != "open":
            raise RuntimeError("Can't currently call tod2map() [or DynamicMap.expand()]"
                               + " after DynamicMap.permute() or DynamicMap.finalize().")
        if not isinstance(plan, PointingPlan):
            raise RuntimeError(f"DynamicMap.expand(): expected a PointingPlan, got {plan!r}")
        if (plan.nypix_global, plan.nxpix_global, plan.periodic_xcoord) != \
                (self.nypix_global, self.nxpix_global, self.periodic_xcoord):
            raise RuntimeError("DynamicMap.expand(): the plan's global pixelization (nypix_global, nxpix_global,"
                               " periodic_xcoord) differs from the map's")
        n = int(_ext.expand_dynamic_map2(self._count, self._table, plan, stream=_stream()))
        self._reserve(n)
        self._ncells = n
        self._table.npix = n * 64 * 64

    def _layout(self):
        """Host copy of the cell table, refreshed from the GPU while the layout is open."""
        if self._state == "open":
            self._table.copy_gpu_offsets_to_cpu()
        return self._table.cell_offsets_cpu

    def permute(self, iycells, ixcells):
        """
        New cell order: cell i of the result is (iycells[i], ixcells[i]). The two integer arrays
        (host, length ncells_curr) must name every present cell once. Fixes the layout: no cells
        can be added afterwards.
        """
        if self._state != "open":
            raise RuntimeError("Can't call DynamicMap.permute() after DynamicMap.permute() or DynamicMap.finalize()")
        table = self._layout()
        n = self._ncells
        iy = np.asarray(iycells)
        ix = np.asarray(ixcells)
        if iy.shape != (n,) or ix.shape != (n,) or not (np.issubdtype(iy.dtype, np.integer)
                                                      and np.issubdtype(ix.dtype, np.integer)):
            raise RuntimeError(f"DynamicMap.permute(): expected two integer arrays of length {n}")
        inside = np.all((iy >= 0) & (iy < table.shape[0]) & (ix >= 0) & (ix < table.shape[1]))
        src = table[iy, ix] if (inside and n > 0) else np.zeros(0, dtype=np.int64)
        if not inside or np.any(src < 0) or np.unique(src).size != n:
            raise RuntimeError("DynamicMap.permute(): specified (iycells,ixcells) do not match footprint of current pixelization")

        # Gather the cells into the new order on the GPU (one kernel), then rewrite the table.
        order = torch.from_numpy(np.ascontiguousarray(src // self._CELL_FLOATS)).to(self._store.device)
        moved = torch.empty((n, 3, 64, 64), dtype=torch.float32, device=self._store.device)
        if n > 0:
            _ext.cell_broadcast(moved.view(-1), self._store[:n].reshape(-1), order, stream=_stream())
        table[iy, ix] = np.arange(n, dtype=np.int64) * self._CELL_FLOATS
        self._table.copy_cpu_offsets_to_gpu()
        self._store = moved
        self._state = "fixed"

    def finalize(self):
        """Ends accumulation and returns the LocalMap (the DynamicMap gives up its storage)."""
        if self._state == "done":
            raise RuntimeError("Double call to DynamicMap.finalize() [not currently allowed]")
        self._layout()
        arr = self._store[:self._ncells].reshape(-1)
        if self._store.shape[0] > self._ncells:
            arr = arr.clone()          # drop the unused capacity
        self._store = None
        self._state = "done"
        return LocalMap(self._table, arr)

    def randomly_permute(self, rng=None):
        """Test aid: puts the present cells in a random order (see permute())."""
        rng = np.random.default_rng() if rng is None else rng
        iy, ix = np.nonzero(self._layout() >= 0)
        p = rng.permutation(iy.size)
        self.permute(iy[p], ix[p])


def cell_broadcast(dst, src, index_map):
    """
    dst cell i = src cell index_map[i], for float32 GPU arrays of whole 3x64x64 cells (flattened)
    and an int64 GPU index map. Reference: src_lib/cell_broadcast.cu:19-36 (host loop).
    """
    _ext.cell_broadcast(dst, src, index_map, stream=_stream())


def cell_reduce(dst, src, index_map):
    """
    dst cell d = sum of src cells i with index_map[i] == d (zero if none), deterministic.
    Reference: src_lib/cell_reduce.cu:19-44 (host loop).
    """
    _ext.cell_reduce(dst, src, index_map, stream=_stream())


def expand_dynamic_map(global_ncells, cell_offsets, plan_mt):
    """
    Low-level expansion over a cell-sorted plan_mt array (uint64, GPU): every targeted (-1)
    entry of cell_offsets (int64 (nycells, nxcells), GPU) named by plan_mt gets the next index.
    Returns the new cell count (also left in global_ncells, uint32 (1,) on the GPU).
    Reference: src_lib/expand_dynamic_map.cu:107-141.
    """
    return int(_ext.expand_dynamic_map(global_ncells, cell_offsets, plan_mt, stream=_stream()))


####################################################################################################


class PointingPrePlan(_ext.PointingPrePlan):
    """
    First planning step: counts the (32-sample line, 64x64 cell) pairs of an xpointing array.

    Arguments: xpointing_gpu (float32 (3,nsamp) or (3,ndet,nt) on the GPU), nypix_global,
    nxpix_global, buf and tmp_buf (uint32[preplan_size] on the GPU, allocated when None; the
    preplan keeps 'buf', 'tmp_buf' is used only during construction), periodic_xcoord, debug.

    Attributes: nsamp, nypix_global, nxpix_global, periodic_xcoord, plan_nbytes,
    plan_constructor_tmp_nbytes, overhead, ncl_per_threadblock, planner_nblocks,
    nmt_per_threadblock, pointing_nblocks, plan_nmt, cub_nbytes (radix-sort scratch bytes),
    get_nmt_cumsum().

    Reference: gpu_mm/gpu_mm.py:1208-1259; src_lib/PointingPrePlan.cu:129-251.
    """

    preplan_size = _ext.PointingPrePlan._get_preplan_size()

    def __init__(self, xpointing_gpu, nypix_global, nxpix_global, buf=None, tmp_buf=None, periodic_xcoord=True, debug=False):
        if buf is None:
            buf = torch.empty(self.preplan_size, dtype=torch.uint32, device=_device())
        if tmp_buf is None:
            tmp_buf = torch.empty(self.preplan_size, dtype=torch.uint32, device=_device())
        _ext.PointingPrePlan.__init__(self, xpointing_gpu, int(nypix_global), int(nxpix_global), buf, tmp_buf,
                                      bool(periodic_xcoord), bool(debug), _stream())


class PointingPlan(_ext.PointingPlan):
    """
    Second planning step: the sorted list of (line, cell) entries plus the work list of the
    planned kernels.

    Arguments: preplan, xpointing_gpu (the same pointing as the preplan), buf (uint8, at least
    preplan.plan_nbytes; kept by the plan), tmp_buf (uint8, at least
    preplan.plan_constructor_tmp_nbytes; used only during construction), debug.

    Attributes: nsamp, nypix_global, nxpix_global, periodic_xcoord, plan_nmt, nchunks,
    nmlines; get_plan_mt() returns the entries (uint64, bit layout of the reference).

    With debug=True the constructor also checks the plan's invariants on the device
    (check_invariants()), and raises RuntimeError if one is broken.

    Reference: gpu_mm/gpu_mm.py:1262-1299; src_lib/PointingPlan.cu:237-312.
    """

    def __init__(self, preplan, xpointing_gpu, buf=None, tmp_buf=None, debug=False):
        if buf is None:
            buf = torch.empty(preplan.plan_nbytes, dtype=torch.uint8, device=_device())
        if tmp_buf is None:
            tmp_buf = torch.empty(preplan.plan_constructor_tmp_nbytes, dtype=torch.uint8, device=_device())
        _ext.PointingPlan.__init__(self, preplan, xpointing_gpu, buf, tmp_buf, bool(debug), _stream())
        self.buf = buf

    def check_invariants(self):
        """
        Checks the plan on the device and raises RuntimeError, counting each kind of violation,
        unless: entries are in (cell, line) order with each pair once; every line index is in
        range and every line has an entry; mflag marks exactly the entries of multi-cell lines
        and each such line has exactly one zflag; the operators' work list covers every entry
        exactly once, each chunk within one cell. (The reference asserts its iterator and cell
        enumerator invariants on the device under Debug, include/plan_iterator.hpp:64-141.)
        """
        self._check_invariants(_stream())

    def check_errflags(self, where="PointingPlan"):
        """Raises if an operator call since the last check reported a pointing error."""
        self._check_errflags(where)


####################################################################################################


class PlanBatch:
    """
    Several TODs, each with its own PointingPlan and xpointing, bound to one local map, so that
    one kernel launch runs map2tod (or tod2map) for all of them. Not in the reference, which
    launches one TOD at a time.

      PlanBatch(local_pixelization, plans, xpointings, tods, scales=None, config=None)

      - plans, xpointings, tods: lists of equal length; TOD i has the samples of xpointing i and
        plan i. TODs are float32 GPU tensors, or (with 'scales') float16 / bfloat16 ones, all of
        one dtype, each with a float32 scale per row (see encode_tod16).
      - config: BatchConfig (work-item size, block size and persistent grid per operator), or
        None for the defaults; autotune() picks one by measurement.

    Work items are runs of about item_entries plan entries of one map cell, gathered across the
    TODs: the cell's halo tile is staged (map2tod) or accumulated and flushed (tod2map) once for
    all TODs that touch it. map2tod writes every bound TOD (bit for bit what the plain per-TOD
    map2tod writes); tod2map adds all of them into the map (equal to the per-TOD calls up to the
    order of the float additions). The batch holds references to everything it points into.
    Calls on one batch must not overlap in time.
    """

    def __init__(self, local_pixelization, plans, xpointings, tods, scales=None, config=None):
        if not isinstance(local_pixelization, LocalPixelization):
            raise RuntimeError(f"PlanBatch: expected LocalPixelization, got {local_pixelization!r}")
        for p in plans:
            if not isinstance(p, PointingPlan):
                raise RuntimeError(f"PlanBatch: expected a list of PointingPlans, got {p!r}")
        self.pixelization = local_pixelization
        self.tods = list(tods)
        self.scales = None if scales is None else list(scales)
        self._b = _ext.Batch(local_pixelization, list(plans), list(xpointings), self.tods, self.scales, config)

    @property
    def ntod(self):
        return self._b.ntod

    @property
    def config(self):
        return self._b.config

    @property
    def nentries(self):
        return self._b.nentries

    def work_items(self):
        """Work-item counts, segment counts and grid sizes of both operators."""
        return self._b.work_items()

    def configure(self, config):
        self._b.configure(config)

    def set_response(self, responses):
        """
        Per-detector gains for every bound TOD (float32 TODs only): responses[i] is a float32
        (2, ndet_i) GPU array as in map2tod(..., response=...), with ndet_i dividing TOD i's
        sample count; the batched operators then apply them as the per-TOD calls do. None removes
        them. Graphs captured before the call must be captured again.
        """
        self._b.set_response(None if responses is None else list(responses))

    def _arr(self, local_map):
        if isinstance(local_map, LocalMap):
            if not (local_map.pixelization == self.pixelization):
                raise RuntimeError("PlanBatch: local_map has a different LocalPixelization")
            return local_map.arr
        return local_map

    def map2tod(self, local_map, partial_pixelization=False, accum=False, debug=False):
        """Every bound TOD (=, or += if accum) A_i * local_map."""
        self._b.map2tod(self._arr(local_map), partial_pixelization, accum,
                        check_errors=(_errflag_check or debug), stream=_stream())

    def tod2map(self, local_map, partial_pixelization=False, debug=False):
        """local_map += sum_i A_i^T * tod_i."""
        self._b.tod2map(self._arr(local_map), partial_pixelization,
                        check_errors=(_errflag_check or debug), stream=_stream())

    def check_errflags(self, where="PlanBatch"):
        self._b.check_errflags(where)

    def autotune(self, local_map, reps=3):
        """
        Times every candidate configuration of both operators on this batch (work-item size
        1024..16384 entries; map2tod blocks of 256 or 512 threads; one block per item or a
        persistent grid of 1..4 blocks per CU) and keeps the fastest. map2tod runs on
        'local_map' and overwrites the bound TODs with map2tod(local_map); tod2map runs on a
        scratch map. Returns the measurements: (op, item_entries, threads, blocks_per_cu, ms).
        """
        arr = self._arr(local_map)
        scratch = torch.zeros_like(arr)
        return self._b.autotune(arr, scratch, reps, stream=_stream())

    def capture(self, op, local_map, partial_pixelization=False, accum=False):
        """
        HIP graph of one batched launch ('map2tod' or 'tod2map') with these arguments; replay it
        with .launch() (on torch's current stream). The graph keeps the batch and the map alive,
        and is refused after configure() or autotune().
        """
        g = self._b.capture(op, self._arr(local_map), partial_pixelization, accum)
        return _BatchGraph(g)


class _BatchGraph:
    def __init__(self, g):
        self._g = g

    def launch(self):
        self._g.launch(stream=_stream())


BatchConfig = _ext.BatchConfig


def read_xpointing_npzfile(filename):
    """
    Reads an xpointing array stored in the reference's Nov-2023 .npz format: key 'xpointing',
    shape (3, ndet, nt) with rows ordered (x, y, alpha). Returns a host array of shape
    (3, ndet, nt_padded), same dtype, with rows in this library's order (y, x, alpha) and nt
    padded up to a multiple of 32 by repeating each detector's last sample.
    Reference: gpu_mm/gpu_mm.py:1303-1329.
    """
    with np.load(filename, allow_pickle=False) as f:
        xp = f['xpointing']
    if xp.ndim != 3 or xp.shape[0] != 3 or xp.shape[2] == 0:
        raise RuntimeError(f"read_xpointing_npzfile({filename!r}): expected 'xpointing' of shape (3, ndet, nt>0),"
                           f" got {xp.shape}")
    pad = (-xp.shape[2]) % 32
    # (x, y, alpha) -> (y, x, alpha); edge padding repeats the last sample of every row.
    return np.pad(xp[[1, 0, 2]], ((0, 0), (0, 0), (0, pad)), mode="edge")


class ToyPointing:
    """
    ToyPointing(ndet, nt, nypix_global, nxpix_global, scan_speed, total_drift, noisy=True):
    a 45-degree scan bouncing over the full y range with a slow drift in x (test aid).

    Members: xpointing_cpu (numpy), xpointing_gpu (torch, on the GPU when available), ndet, nt,
    nypix_global, nxpix_global, scan_speed, total_drift, drift_speed.
    Reference: gpu_mm/gpu_mm.py:1332-1385; src_lib/ToyPointing.cu:41-109.
    """

    def __init__(self, ndet, nt, nypix_global, nxpix_global, scan_speed, total_drift, noisy=True):
        shape = (3, ndet, nt) if (ndet is not None) else (3, nt)
        self.ndet = ndet
        self.nt = nt
        self.nypix_global = nypix_global
        self.nxpix_global = nxpix_global
        self.scan_speed = scan_speed
        self.total_drift = total_drift
        self.xpointing_cpu = np.zeros(shape, dtype=np.float32)
        if noisy:
            print(f"{self}: constructor called")
        self.drift_speed = _ext.toy_pointing_fill(self.xpointing_cpu, nypix_global, nxpix_global, scan_speed, total_drift)
        self.xpointing_gpu = torch.from_numpy(self.xpointing_cpu).to(_device()) if torch.cuda.is_available() else None

    def __str__(self):
        shape = (self.ndet, self.nt) if self.ndet is not None else (self.nt,)
        return (f"ToyPointing(tod_shape={shape}, nypix_global={self.nypix_global}, nxpix_global={self.nxpix_global},"
                f" scan_speed={self.scan_speed}, total_drift={self.total_drift})")

    @staticmethod
    def make_random(nsamp_max, noisy=True, rng=None):
        """Random toy instance with at most nsamp_max samples (gpu_mm.py:1364-1385)."""
        assert nsamp_max >= 64 * 1024
        rng = np.random.default_rng() if rng is None else rng
        if rng.integers(0, 2):
            ndet = None
            nt = 32 * int(rng.integers(nsamp_max // 64, nsamp_max // 32))
        else:
            r = int(rng.integers(0, 6))
            s = int((nsamp_max / 32.) ** 0.5)
            ndet = 2 ** r * int(rng.integers(s // 2, s + 1))
            nt = 2 ** (5 - r) * int(rng.integers(s // 2, s + 1))
        npix_max = min(nsamp_max // 128, 16384)
        nypix_global = int(rng.integers(npix_max / 8, npix_max // 2 + 1))
        nxpix_global = int(rng.integers(nypix_global + 2, npix_max + 1))
        scan_speed = float(rng.uniform(0.1, 0.5))
        total_drift = float(rng.uniform(0.1 * (nxpix_global - nypix_global), (nxpix_global - nypix_global) - 2))
        return ToyPointing(ndet, nt, nypix_global, nxpix_global, scan_speed, total_drift, noisy=noisy)


class PointingPlanTester:
    """
    Brute-force plan for tests: iypix/ixpix (the (iy0, ix0) pixel of every sample, computed on
    the GPU), nmt_cumsum (per planner block) and sorted_mt (46-bit entries without flags,
    sorted as integers). Reference: gpu_mm/gpu_mm.py:1388-1415; src_lib/PointingPlanTester.cu.
    """

    def __init__(self, preplan, xpointing_gpu):
        assert isinstance(preplan, PointingPrePlan)
        tmp = torch.empty(_ext.plan_tester_tmp_nbytes(preplan), dtype=torch.uint8, device=_device())
        self.iypix, self.ixpix, self.nmt_cumsum, self.sorted_mt = _ext.plan_tester_run(preplan, xpointing_gpu, tmp, _stream())
        self.nsamp = preplan.nsamp
        self.nypix_global = preplan.nypix_global
        self.nxpix_global = preplan.nxpix_global
        self.plan_nmt = preplan.plan_nmt
        self.ncl_per_threadblock = preplan.ncl_per_threadblock
        self.planner_nblocks = preplan.planner_nblocks


def test_plan_iterator(plan_mt, nmt_per_block, warps_per_threadblock):
    """Walks the planned kernels' work list for a cell-sorted plan_mt and checks coverage."""
    _ext.test_plan_iterator(np.ascontiguousarray(plan_mt, dtype=np.uint64), int(nmt_per_block), int(warps_per_threadblock))


def make_random_plan_mt(ncells, min_nmt_per_cell, max_nmt_per_cell, seed=0):
    """Random cell-sorted plan (uint64 entries) with junk in the bits above 46."""
    return _ext.make_random_plan_mt(int(ncells), int(min_nmt_per_cell), int(max_nmt_per_cell), int(seed))
