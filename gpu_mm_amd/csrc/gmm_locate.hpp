// Pixel location of a pointing sample on the device, shared by the planner and the operators:
// the (iy0, ix0) corner of its bilinear footprint (clamped as in include/gpu_mm_internals.hpp:
// 124-164) with range checks deferred to one per lane. Checks keeps, per lane, the largest bit
// pattern of y + 0 and of x + 0 (non-periodic x): as unsigned integers the patterns of floats in
// [0, hi] are exactly those <= bits(hi), so one integer max per coordinate covers the range and
// NaN checks (+ 0 turns -0.0 into +0.0). Periodic x keeps the largest excess over [-nx, 2nx] and
// a running sum that a NaN turns into NaN.

#pragma once

#include "gmm_internal.hpp"

#include <cmath>

namespace gmm {

struct Geom
{
    int ny, nx;
    float ymax;     // ny - 1
    float xlo, xhi; // valid x range of the periodic case: [-nx, 2nx]
    float fnx;
    float fny2, fnx2; // ny - 2, nx - 2
    uint32_t uymax, uxmax;  // bit patterns of ny - 1 and nx - 2
};

__device__ inline Geom make_geom(int ny, int nx, bool periodic)
{
    Geom g;
    g.ny = ny;
    g.nx = nx;
    g.ymax = float(ny - 1);
    g.fnx = float(nx);
    g.fny2 = float(ny - 2);
    g.fnx2 = float(nx - 2);
    g.xlo = periodic ? -float(nx) : 0.0f;
    g.xhi = periodic ? float(2 * nx) : float(nx - 2);
    g.uymax = __float_as_uint(g.ymax);
    g.uxmax = __float_as_uint(g.fnx2);
    return g;
}

struct Checks
{
    uint32_t uy = 0, ux = 0;
    float px = -INFINITY, pn = 0.0f;

    __device__ uint32_t bits(const Geom &g) const
    {
        uint32_t e = 0;
        if (uy > g.uymax)
            e |= ERR_BAD_YPIX;
        if (ux > g.uxmax || !(px <= 0.0f) || !(pn == pn))
            e |= ERR_BAD_XPIX;
        return e;
    }
};

// Clamp to [0, hi] in one instruction (v_med3_f32). For a NaN input the hardware returns the
// minimum of the operands, which is 0 here or NaN; either converts to index 0.
__device__ inline float clamp0(float v, float hi)
{
    return __builtin_amdgcn_fmed3f(v, 0.0f, hi);
}

template<bool Periodic>
__device__ inline void locate_op(float y, float x, const Geom &g, Checks &c, int &iy0, int &ix0, float &dy, float &dx)
{
    c.uy = max(c.uy, __float_as_uint(y + 0.0f));
    const int iy = int(clamp0(y, g.fny2));
    dy = y - float(iy);
    iy0 = iy;

    if (Periodic) {
        c.px = fmaxf(c.px, fmaxf(g.xlo - x, x - g.xhi));
        c.pn += x;
        x = (x >= 0.0f) ? x : (x + g.fnx);
        x = (x <= g.fnx) ? x : (x - g.fnx);
        const int ix = int(clamp0(x, g.fnx));
        dx = __builtin_amdgcn_fractf(x);                  // x - ix for x in [0, nx]
        ix0 = (ix < g.nx) ? ix : 0;
    }
    else {
        c.ux = max(c.ux, __float_as_uint(x + 0.0f));
        const int ix = int(clamp0(x, g.fnx2));
        dx = __builtin_amdgcn_fractf(x);                  // x - ix for x in [0, nx - 2]
        ix0 = ix;
    }
}

}  // namespace gmm
