"""CPU tests of the conjugate-gradient solver of gpu_mm_amd.mapmaker (generic torch code):
convergence to the exact solution of a random symmetric positive definite system, with and
without a Jacobi preconditioner, and FP64 scalars."""

import numpy as np
import pytest
import torch

from gpu_mm_amd import mapmaker


def _spd(n, cond, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.geomspace(1.0, cond, n)
    d = rng.uniform(0.5, 20.0, n)             # badly scaled rows: what a Jacobi preconditioner fixes
    a = (q * ev) @ q.T
    return (d[:, None] * a * d[None, :]), rng


def test_pcg_solves_spd_system():
    n = 300
    a64, rng = _spd(n, 50.0, 0)
    x_true = rng.standard_normal(n)
    b64 = a64 @ x_true
    a = torch.from_numpy(a64.astype(np.float32))
    b = torch.from_numpy(b64.astype(np.float32))

    def apply_a(x, out):
        torch.mv(a, x, out=out)
        return out

    dinv = torch.from_numpy((1.0 / np.diag(a64)).astype(np.float32))

    def jacobi(r, out=None):
        out = torch.empty_like(r) if out is None else out
        torch.mul(dinv, r, out=out)
        return out

    x, hist = mapmaker.pcg(apply_a, b, precond=jacobi, niter=200, rtol=1e-6)
    _, hist_cg = mapmaker.pcg(apply_a, b, precond=None, niter=len(hist) - 1)
    err = np.linalg.norm(x.numpy().astype(np.float64) - x_true) / np.linalg.norm(x_true)
    assert hist[-1] < 1e-6
    assert err < 1e-4
    assert hist_cg[-1] > hist[-1]             # the preconditioner helps on a badly scaled system
    assert all(isinstance(h, float) for h in hist)


def test_pcg_zero_rhs_and_initial_guess():
    a64, rng = _spd(50, 10.0, 1)
    a = torch.from_numpy(a64.astype(np.float32))
    x_true = torch.from_numpy(rng.standard_normal(50).astype(np.float32))
    b = a @ x_true

    def apply_a(x, out):
        torch.mv(a, x, out=out)
        return out

    x, hist = mapmaker.pcg(apply_a, b, x0=x_true, niter=5)
    assert hist[0] < 1e-5
    assert torch.allclose(x, x_true, atol=1e-4)


def test_pcg_breakdown_without_host_checks():
    """A = 2 I: the first step solves the system exactly (r = 0); the following iterations have
    zero curvature and, run without host checks, leave x unchanged (no NaN in x or history)."""
    n = 16
    b = torch.from_numpy(np.random.default_rng(5).standard_normal(n).astype(np.float32))

    def apply_a(x, out):
        torch.mul(x, 2.0, out=out)
        return out

    x, hist = mapmaker.pcg(apply_a, b, niter=6)
    assert len(hist) == 7
    assert all(np.isfinite(h) for h in hist)
    assert hist[1] < 1e-7 and hist[-1] == hist[1]
    assert torch.equal(x, b * 0.5)


def test_deferred_errflags_restores_setting():
    import gpu_mm_amd as gmm
    assert gmm.get_errflag_check()
    with gmm.deferred_errflags():
        assert not gmm.get_errflag_check()
    assert gmm.get_errflag_check()
    with pytest.raises(ValueError):
        with gmm.deferred_errflags():
            raise ValueError("inside")
    assert gmm.get_errflag_check()
