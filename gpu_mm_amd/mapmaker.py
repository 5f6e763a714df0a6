"""
Differentiable pointing operator and a preconditioned conjugate-gradient map solver.

Not in the reference, which provides the operators only (gpu_mm/gpu_mm.py:262-385); this module
builds the map-maker around them on the GPU:

  map2tod_ad(map_arr, xpointing, plan, lpix)   torch.autograd.Function: A m, backward A^T g
  tod2map_ad(tod, xpointing, plan, lpix)       torch.autograd.Function: A^T t, backward A g

Both backwards are the other planned kernel (tod2map for map2tod and vice versa), so gradients
are exact transposes of the forward operators.

  WeightedPointing(lpix, plans, xpointings, ninv)   the normal operator A^T N^-1 A over several
      TODs with white noise of inverse variance ninv[i][det] per detector; N^-1/2 is applied
      inside the kernels as a per-detector gain (the response arguments of map2tod/tod2map),
      so no separate weighting pass touches the TODs.
  block_preconditioner(...)   per-pixel 3x3 inverse of the hit-count / polarization-angle matrix
      M_p = sum_s sum_k w_k(s) ninv(s) [1, c, s]^T [1, c, s] (c, s = cos, sin 2 alpha; w_k the
      bilinear weights of the pixel's corner), built from two tod2map passes.
  pcg(...)   preconditioned CG with FP32 vectors and FP64 inner products, step scalars and
      residual norms.

Maps are local-map arrays (float32, 3*npix elements) in the layout of a LocalPixelization.
"""

import numpy as np
import torch

from . import gpu_mm as _g


# ------------------------------------------------------------------------------------------
# Autograd functions.

class _Map2Tod(torch.autograd.Function):
    @staticmethod
    def forward(ctx, map_arr, xpointing, plan, lpix):
        ctx.save_for_backward(xpointing)
        ctx.plan, ctx.lpix = plan, lpix
        tod = torch.empty(tuple(xpointing.shape[1:]), dtype=torch.float32, device=map_arr.device)
        _g.map2tod(tod, _g.LocalMap(lpix, map_arr.detach().contiguous()), xpointing, plan)
        return tod

    @staticmethod
    def backward(ctx, grad_tod):
        (xpointing,) = ctx.saved_tensors
        g = torch.zeros(3 * ctx.lpix.npix, dtype=torch.float32, device=grad_tod.device)
        _g.tod2map(_g.LocalMap(ctx.lpix, g), grad_tod.contiguous(), xpointing, ctx.plan)
        return g, None, None, None


class _Tod2Map(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tod, xpointing, plan, lpix):
        ctx.save_for_backward(xpointing)
        ctx.plan, ctx.lpix = plan, lpix
        m = torch.zeros(3 * lpix.npix, dtype=torch.float32, device=tod.device)
        _g.tod2map(_g.LocalMap(lpix, m), tod.detach().contiguous(), xpointing, plan)
        return m

    @staticmethod
    def backward(ctx, grad_map):
        (xpointing,) = ctx.saved_tensors
        g = torch.empty(tuple(xpointing.shape[1:]), dtype=torch.float32, device=grad_map.device)
        _g.map2tod(g, _g.LocalMap(ctx.lpix, grad_map.contiguous()), xpointing, ctx.plan)
        return g, None, None, None


def map2tod_ad(map_arr, xpointing, plan, lpix):
    """A m as a differentiable function of the local-map array 'map_arr' (float32, 3*npix); the
    result has the shape of one xpointing row. Backward: the planned tod2map kernel."""
    return _Map2Tod.apply(map_arr, xpointing, plan, lpix)


def tod2map_ad(tod, xpointing, plan, lpix):
    """A^T t as a differentiable function of the TOD; returns a new local-map array. Backward:
    the planned map2tod kernel."""
    return _Tod2Map.apply(tod, xpointing, plan, lpix)


# ------------------------------------------------------------------------------------------
# Local-map pixel indexing.

def pixel_index(lpix, device):
    """int64 (3, npix) tensor: element [p, i] is the offset of component p of local pixel i in
    the local-map array (cells in cell_offsets order, pixels row-major inside a cell)."""
    off = np.asarray(lpix.cell_offsets_cpu, dtype=np.int64)
    cells = off[off >= 0]
    iy, ix = np.meshgrid(np.arange(64, dtype=np.int64), np.arange(64, dtype=np.int64), indexing="ij")
    inner = (iy * lpix.ystride + ix).reshape(-1)
    base = (cells[:, None] + inner[None, :]).reshape(-1)
    idx = np.stack([base + p * lpix.polstride for p in range(3)])
    return torch.from_numpy(idx).to(device)


# ------------------------------------------------------------------------------------------
# Weighted normal operator over several TODs.

class WeightedPointing:
    """
    The operators of a white-noise map-maker over several TODs sharing one LocalPixelization:

      apply_normal(m)  = A^T N^-1 A m
      rhs(d)           = A^T N^-1 d
      loss(m, d)       = 1/2 sum_i || N_i^-1/2 (A_i m - d_i) ||^2   (FP64 sum)

    ninv[i] is a float32 GPU tensor of per-detector inverse variances of TOD i (TOD i has shape
    (ndet_i, nt_i)). The weights enter the kernels as per-detector gains (response rows (g, g)):
    A_g m = g A m with g = sqrt(ninv), so A^T N^-1 A = A_g^T A_g, evaluated with one batched
    map2tod and one batched tod2map over all TODs and no pass over the TODs in between.
    autotune=True times the batch configurations once at construction (PlanBatch.autotune) and
    keeps the fastest.
    """

    def __init__(self, lpix, plans, xpointings, ninv, autotune=False):
        if not (len(plans) == len(xpointings) == len(ninv)):
            raise RuntimeError("WeightedPointing: plans, xpointings and ninv must have equal length")
        self.lpix = lpix
        self.plans = list(plans)
        self.xps = list(xpointings)
        self.ninv = [w.to(torch.float32).contiguous() for w in ninv]
        for x, w in zip(self.xps, self.ninv):
            if x.ndim != 3 or w.shape != (x.shape[1],):
                raise RuntimeError("WeightedPointing: xpointing must be (3, ndet, nt) with ninv of shape (ndet,)")
        self.g = [torch.stack([torch.sqrt(w), torch.sqrt(w)]).contiguous() for w in self.ninv]
        self.w2 = [torch.stack([w, w]).contiguous() for w in self.ninv]
        self.work = [torch.empty(tuple(x.shape[1:]), dtype=torch.float32, device=x.device) for x in self.xps]
        # The normal operator runs as one batched map2tod and one batched tod2map over all TODs
        # (the gains as per-detector responses); TODs that cannot be batched (a sample count
        # not divisible by 4) are run one by one.
        try:
            self.batch = _g.PlanBatch(lpix, self.plans, self.xps, self.work)
            self.batch.set_response(self.g)
        except RuntimeError:
            self.batch = None
        if autotune and self.batch is not None:
            scratch = torch.zeros(self.nmap, dtype=torch.float32, device=self.xps[0].device)
            with _g.deferred_errflags():
                self.batch.autotune(_g.LocalMap(lpix, scratch))
            self.batch.check_errflags("WeightedPointing")

    @property
    def nmap(self):
        return 3 * self.lpix.npix

    def _lmap(self, arr):
        return _g.LocalMap(self.lpix, arr)

    def apply_normal(self, m, out=None):
        """
        out = A^T N^-1 A m (out is zeroed first; a new array if None). The operator calls do not
        block on their error flags (an iterative solver calls this many times): errors
        accumulate on the plans and check_errflags() raises them.
        """
        out = torch.zeros_like(m) if out is None else out.zero_()
        lm, lo = self._lmap(m), self._lmap(out)
        with _g.deferred_errflags():
            if self.batch is not None:
                self.batch.map2tod(lm)
                self.batch.tod2map(lo)
            else:
                for x, p, g, t in zip(self.xps, self.plans, self.g, self.work):
                    _g.map2tod(t, lm, x, p, response=g)
                    _g.tod2map(lo, t, x, p, response=g)
        return out

    def check_errflags(self):
        """Raises if an operator call of apply_normal() reported an error since the last check
        (every plan's flags are read and cleared; the first error is raised)."""
        first = None
        checks = [p.check_errflags for p in self.plans]
        if self.batch is not None:
            checks.append(self.batch.check_errflags)
        for c in checks:
            try:
                c("WeightedPointing")
            except RuntimeError as e:
                first = e if first is None else first
        if first is not None:
            raise first

    def rhs(self, tods):
        """A^T N^-1 d for the list of TODs d."""
        out = torch.zeros(self.nmap, dtype=torch.float32, device=self.xps[0].device)
        lo = self._lmap(out)
        for x, p, w, d in zip(self.xps, self.plans, self.w2, tods):
            _g.tod2map(lo, d.contiguous(), x, p, response=w)
        return out

    def loss(self, m, tods):
        """1/2 sum_i ||N_i^-1/2 (A_i m - d_i)||^2 (differentiable in m; FP64 reduction)."""
        total = torch.zeros((), dtype=torch.float64, device=m.device)
        for x, p, w, d in zip(self.xps, self.plans, self.ninv, tods):
            r = map2tod_ad(m, x, p, self.lpix) - d
            total = total + 0.5 * torch.sum(w.double()[:, None] * r.double() ** 2)
        return total


# ------------------------------------------------------------------------------------------
# 3x3 block preconditioner.

class BlockPreconditioner:
    """
    z = M^-1 r with M_p the per-pixel 3x3 matrix sum_s sum_k w_k ninv [1, c, s]^T [1, c, s]
    (c = cos 2 alpha, s = sin 2 alpha, w_k the bilinear weight of pixel p for sample s). With
    c^2 = (1 + cos 4a)/2, s^2 = (1 - cos 4a)/2 and cs = sin(4a)/2, the six distinct entries come
    from two tod2map passes of the inverse-variance weights: one with the pointing as it is
    ([H, C, S] = sum w ninv [1, cos 2a, sin 2a]) and one with every angle doubled ([H, C4, S4]):

        M = [[H, C, S], [C, (H + C4)/2, S4/2], [S, S4/2, (H - C4)/2]].

    Pixels whose M has reciprocal condition number below rcond_min (poor angle coverage) get the
    inverse of its diagonal, which keeps the preconditioner positive definite on every hit pixel;
    unhit pixels get 0. The inverse is computed in FP64 and applied in FP32.
    """

    def __init__(self, wp, rcond_min=1e-3):
        dev = wp.xps[0].device
        a = torch.zeros(wp.nmap, dtype=torch.float32, device=dev)
        b = torch.zeros(wp.nmap, dtype=torch.float32, device=dev)
        la, lb = _g.LocalMap(wp.lpix, a), _g.LocalMap(wp.lpix, b)
        for x, p, w in zip(wp.xps, wp.plans, wp.w2):
            ones = torch.ones(tuple(x.shape[1:]), dtype=torch.float32, device=dev)
            _g.tod2map(la, ones, x, p, response=w)
            x2 = x.clone()
            x2[2] *= 2.0
            _g.tod2map(lb, ones, x2, p, response=w)
            del x2, ones
        self.idx = pixel_index(wp.lpix, dev)
        va = a[self.idx].double()      # (3, npix): H, C, S
        vb = b[self.idx].double()      # (3, npix): H, C4, S4
        h, c, s = va[0], va[1], va[2]
        c4, s4 = vb[1], vb[2]
        m = torch.empty((h.shape[0], 3, 3), dtype=torch.float64, device=dev)
        m[:, 0, 0] = h
        m[:, 0, 1] = m[:, 1, 0] = c
        m[:, 0, 2] = m[:, 2, 0] = s
        m[:, 1, 1] = 0.5 * (h + c4)
        m[:, 2, 2] = 0.5 * (h - c4)
        m[:, 1, 2] = m[:, 2, 1] = 0.5 * s4
        self.hits = h
        self.m = m
        hit = h > 0
        ev = torch.linalg.eigvalsh(m)
        rc = torch.where(ev[:, 2] > 0, ev[:, 0] / ev[:, 2].clamp(min=1e-300), torch.zeros_like(h))
        full = hit & (rc >= rcond_min)
        inv = torch.zeros_like(m)
        if bool(full.any()):
            inv[full] = torch.linalg.inv(m[full])
        diag = hit & ~full
        for k in range(3):
            dk = m[:, k, k]
            inv[:, k, k] = torch.where(diag & (dk > 0), 1.0 / dk.clamp(min=1e-300), inv[:, k, k])
        self.nfull, self.ndiag = int(full.sum()), int(diag.sum())
        minv = inv.permute(1, 2, 0).to(torch.float32)    # (3, 3, npix), pixels in cell order
        # When component p of every pixel sits at p * polstride + j with j covering
        # [0, polstride) exactly (the dense rectangle), the inverse is stored in that order and
        # applied to the (3, polstride) view of the map: no gather or scatter per application.
        ps = int(wp.lpix.polstride)
        self.planar = (3 * ps == wp.nmap and self.idx.shape[1] == ps
                       and int(self.idx[0].min()) >= 0 and int(self.idx[0].max()) < ps)
        if self.planar:
            self.minv = torch.empty_like(minv)
            self.minv[:, :, self.idx[0]] = minv
        else:
            self.minv = minv.contiguous()

    def __call__(self, r, out=None):
        # The pixel index covers every element of the local map (the cells tile it exactly).
        out = torch.empty_like(r) if out is None else out
        if self.planar:
            v, o = r.view(3, -1), out.view(3, -1)
        else:
            v, o = r[self.idx], torch.empty((3, self.idx.shape[1]), dtype=r.dtype, device=r.device)
        mi = self.minv
        for i in range(3):
            torch.mul(mi[i, 0], v[0], out=o[i])
            o[i].addcmul_(mi[i, 1], v[1])
            o[i].addcmul_(mi[i, 2], v[2])
        if not self.planar:
            out[self.idx] = o
        return out


# ------------------------------------------------------------------------------------------
# Preconditioned conjugate gradient.

def dot64(a, b):
    """Inner product of two FP32 vectors: FP32 products summed in FP64; a Python float."""
    return float(torch.sum(a * b, dtype=torch.float64))


def _dot64_dev(a, b):
    """Inner product of two FP32 vectors (FP32 products summed in FP64), as a 0-d FP64 tensor."""
    return torch.sum(a * b, dtype=torch.float64)


def pcg(apply_a, b, precond=None, x0=None, niter=50, rtol=0.0, callback=None, check=None):
    """
    Solves A x = b (A symmetric positive definite, given as apply_a(x, out)) by preconditioned
    conjugate gradients. Vectors are FP32; inner products, step scalars and residual norms are
    FP64. Returns (x, history) with history[k] = ||r_k|| / ||b|| (k = 0..iterations run); stops
    after niter iterations or when the relative residual falls below rtol.

    The scalars stay on the device: without rtol and callback, the iterations run without a
    host synchronization and the history is read once at the end. A step whose curvature p.Ap
    is not positive (breakdown, e.g. an exactly zero residual) leaves x unchanged: with rtol or
    a callback the iteration stops there, otherwise the remaining iterations keep x as it is.
    check, if given, is called once after the iterations (e.g. WeightedPointing.check_errflags).
    """
    host_checks = (rtol > 0.0) or (callback is not None)
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    r = b.clone()
    ap = torch.empty_like(b)
    if x0 is not None:
        apply_a(x, ap)
        r -= ap
    z = precond(r) if precond is not None else r.clone()
    p = z.clone()
    rz = _dot64_dev(r, z)
    bnorm = float(torch.sqrt(_dot64_dev(b, b)))
    rr = [_dot64_dev(r, r)]
    for k in range(niter):
        apply_a(p, ap)
        pap = _dot64_dev(p, ap)
        ok = pap > 0.0
        alpha = torch.where(ok, rz / torch.where(ok, pap, 1.0), 0.0)
        a32 = alpha.to(torch.float32)
        x.addcmul_(p, a32)
        r.addcmul_(ap, -a32)
        rr.append(_dot64_dev(r, r))
        if host_checks:
            if not bool(ok):
                rr.pop()
                break
            res = float(torch.sqrt(rr[-1])) / bnorm
            if callback is not None:
                callback(k, x, res)
            if res < rtol:
                break
        if precond is not None:
            precond(r, out=z)
        else:
            z.copy_(r)
        rz_new = _dot64_dev(r, z)
        beta = torch.where(rz > 0.0, rz_new / torch.where(rz > 0.0, rz, 1.0), 0.0)
        rz = rz_new
        p.mul_(beta.to(torch.float32)).add_(z)
    hist = (torch.sqrt(torch.stack(rr)) / bnorm).cpu().tolist()
    if check is not None:
        check()
    return x, hist
