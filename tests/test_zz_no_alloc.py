"""The allocation-free dynamic-map expansion (gmm_expand_dynamic_map_noalloc) against the
allocating one (gmm_expand_dynamic_map), both called through the C ABI with ctypes: identical
cell tables, counts and returned counts on two calls in a row, the second on a smaller cell grid
and plan (the library's static scratch array then holds the first call's flags)."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

c_void_p, c_int64 = ctypes.c_void_p, ctypes.c_int64


@pytest.fixture(scope="module")
def lib(dev):
    from gpu_mm_amd import _native
    lib = ctypes.CDLL(_native.lib_path())
    for name in ("gmm_expand_dynamic_map", "gmm_expand_dynamic_map_noalloc"):
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_int64, ctypes.POINTER(ctypes.c_uint32),
                      c_void_p]
    lib.gmm_last_error.restype = ctypes.c_char_p
    return lib


def _plan_mt(rng, nycells, nxcells, nmt_max):
    """Cell-sorted plan entries with junk line bits; some cells lie outside the grid."""
    ntot = nycells * nxcells
    nhit = int(rng.integers(1, ntot + 1))
    nmt = int(rng.integers(nhit, nmt_max))
    iy, ix = np.mgrid[:nycells, :nxcells]
    keys = rng.permutation(((iy << 10) + ix).ravel())[:nhit]
    outside = np.array([(nycells << 10) + 1, 5 * 1024 + nxcells, (1023 << 10) + 1023])
    keys = np.concatenate([keys, outside])
    mt = np.sort(np.concatenate([keys, keys[rng.integers(0, len(keys), size=nmt - len(keys))]]).astype(np.uint64))
    return mt | (rng.integers(0, 1 << 28, size=len(mt)).astype(np.uint64) << np.uint64(20))


def _table(rng, nycells, nxcells):
    off = np.where(rng.uniform(size=(nycells, nxcells)) < 0.7, -1, -2).astype(np.int64)
    cur = rng.uniform(size=(nycells, nxcells)) < 0.2
    ncur = int(cur.sum())
    off[cur] = rng.permutation(ncur) * 3 * 64 * 64
    return off, ncur


def _call(lib, name, g, og, nycells, nxcells, mtg, want_count, stream):
    n = ctypes.c_uint32(0)
    rc = getattr(lib, name)(g.data_ptr(), og.data_ptr(), nycells, nxcells, mtg.data_ptr(), mtg.numel(),
                            ctypes.byref(n) if want_count else None, stream)
    assert rc == 0, lib.gmm_last_error().decode()
    return n.value


@pytest.mark.parametrize("want_count", [True, False])
def test_noalloc_matches_allocating_entry_point(lib, dev, want_count):
    rng = np.random.default_rng(4242 + want_count)
    stream = torch.cuda.current_stream().cuda_stream
    for nycells, nxcells, nmt_max in ((48, 96, 300000), (7, 13, 600)):
        print(f"\n[no-alloc expand] grid {nycells}x{nxcells}, nmt < {nmt_max}, count={want_count}", flush=True)
        mt = _plan_mt(rng, nycells, nxcells, nmt_max)
        off, ncur = _table(rng, nycells, nxcells)
        mtg = torch.from_numpy(mt.view(np.int64)).to(dev)
        og_a, og_b = torch.from_numpy(off).to(dev), torch.from_numpy(off).to(dev)
        g_a = torch.full((1,), ncur, dtype=torch.int32, device=dev)
        g_b = g_a.clone()
        torch.cuda.synchronize()
        n_a = _call(lib, "gmm_expand_dynamic_map", g_a, og_a, nycells, nxcells, mtg, True, stream)
        n_b = _call(lib, "gmm_expand_dynamic_map_noalloc", g_b, og_b, nycells, nxcells, mtg, want_count, stream)
        torch.cuda.synchronize()
        got_a, got_b = og_a.cpu().numpy(), og_b.cpu().numpy()
        assert np.array_equal(got_a, got_b)
        assert int(g_a.cpu()[0]) == int(g_b.cpu()[0]) == n_a
        if want_count:
            assert n_b == n_a
        # Something happened: new cells were added, in cell-key order.
        new = (off == -1) & (got_a >= 0)
        assert n_a > ncur and int(new.sum()) == n_a - ncur
        assert np.array_equal(got_a[new], (ncur + np.arange(n_a - ncur)) * 3 * 64 * 64)


def test_noalloc_rejects_a_grid_beyond_its_capacity(lib, dev):
    og = torch.full((1,), -1, dtype=torch.int64, device=dev)
    g = torch.zeros(1, dtype=torch.int32, device=dev)
    mtg = torch.zeros(1, dtype=torch.int64, device=dev)
    rc = lib.gmm_expand_dynamic_map_noalloc(g.data_ptr(), og.data_ptr(), 1025, 1, mtg.data_ptr(), 1, None, None)
    assert rc != 0 and b"nycells, nxcells <= 1024" in lib.gmm_last_error()
