"""Host-side I/O formats: read_xpointing_npzfile (gpu_mm/gpu_mm.py:1303-1329)."""

import numpy as np
import pytest


@pytest.mark.parametrize("nt", [1, 31, 32, 100, 4096])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_read_xpointing_npzfile_round_trip(gmm, tmp_path, nt, dtype):
    rng = np.random.default_rng(nt)
    ndet = 5
    x = rng.uniform(0, 500, (ndet, nt)).astype(dtype)
    y = rng.uniform(0, 300, (ndet, nt)).astype(dtype)
    a = rng.uniform(-3, 3, (ndet, nt)).astype(dtype)
    fn = tmp_path / "xp.npz"
    np.savez(fn, xpointing=np.stack([x, y, a]))        # the file's row order is (x, y, alpha)

    got = gmm.read_xpointing_npzfile(str(fn))
    ntp = 32 * ((nt + 31) // 32)
    assert got.shape == (3, ndet, ntp) and got.dtype == dtype
    # rows swapped to (y, x, alpha), values untouched
    assert np.array_equal(got[0, :, :nt], y)
    assert np.array_equal(got[1, :, :nt], x)
    assert np.array_equal(got[2, :, :nt], a)
    # padding repeats each detector's last sample
    for r, src in enumerate((y, x, a)):
        assert np.array_equal(got[r, :, nt:], np.repeat(src[:, -1:], ntp - nt, axis=1))


def test_read_xpointing_npzfile_rejects_bad_shapes(gmm, tmp_path):
    for shape in [(2, 4, 64), (3, 64), (3, 4, 0)]:
        fn = tmp_path / f"bad{len(shape)}{shape[0]}.npz"
        np.savez(fn, xpointing=np.zeros(shape, np.float32))
        with pytest.raises(RuntimeError, match="expected 'xpointing'"):
            gmm.read_xpointing_npzfile(str(fn))
