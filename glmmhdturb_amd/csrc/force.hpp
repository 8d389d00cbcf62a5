// force.hpp -- OU forcing interpolation with a per-column cache, shared by the GLM and CT marches.
//
// The kick's acceleration at a cell is the trilinear interpolation of the 3 x g^3 forcing field
// (glmmhd_turb.jl:483-505 turb_interp).  A z-march crosses a forcing plane only every n/g cells,
// so each lane keeps the (x, y) bilinear values on the two forcing planes that bracket its
// current cell and redoes the 24 corner loads only at a crossing.  The cache lives in LDS (one
// slot per lane and value, stride STRIDE = the workgroup width): it is touched once per cell only
// when the kick is on, and keeping it out of registers leaves the unforced march unchanged.
#pragma once
#include "gpu_util.hpp"

namespace glm {

// Bilinear (x,y) part of the trilinear interpolation at forcing plane kw for one component, in
// the reference's association order.
template <typename T>
GLM_DEV T force_bilinear(const T* af, int g, int i0w, int i1w, int j0w, int j1w, int kw, T wx, T wy)
{
    auto at = [&](int i, int j) { return gload(af + ((size_t)i + (size_t)g * ((size_t)j + (size_t)g * kw))); };
    return (at(i0w, j0w) * (T(1) - wx) + at(i1w, j0w) * wx) * (T(1) - wy) +
           (at(i0w, j1w) * (T(1) - wx) + at(i1w, j1w) * wx) * wy;
}

template <typename T, int STRIDE>
struct ForceCache {
    int k0;        // lower forcing plane of b0, or a sentinel when empty
    T* b;          // this lane's 6 LDS slots: b0[3], b1[3], stride STRIDE
    GLM_DEV T& b0(int c) { return b[(size_t)c * STRIDE]; }
    GLM_DEV T& b1(int c) { return b[(size_t)(3 + c) * STRIDE]; }
};

// ramp * (interpolated acceleration) at cell (gx, gy, gz).  A is any argument block with the
// forcing field (afield, tgs), the cell size and box lengths (dx, lx, ly, lz) and the ramp.
template <typename T, int STRIDE, typename A>
GLM_DEV void force_accel(ForceCache<T, STRIDE>& fc, const A& a, int gx, int gy, int gz, T acc3[3])
{
    const int g = a.tgs;
    const T gg = T(g);
    T cz = ((T(gz) + T(0.5)) * a.dx) / a.lz * gg;
    T fz = floor(cz);
    int k0 = (int)fz;
    T wz = cz - fz;
    if (k0 != fc.k0) {   // the column's (x, y) position is needed only when a forcing plane is crossed
        T cx = ((T(gx) + T(0.5)) * a.dx) / a.lx * gg;
        T cy = ((T(gy) + T(0.5)) * a.dx) / a.ly * gg;
        T fx = floor(cx), fy = floor(cy);
        int i0 = (int)fx, j0 = (int)fy;
        T wx = cx - fx, wy = cy - fy;
        int i0w = wrapi(i0, g), i1w = wrapi(i0 + 1, g), j0w = wrapi(j0, g), j1w = wrapi(j0 + 1, g);
        const size_t g3 = (size_t)g * g * g;
        const bool shift = k0 == fc.k0 + 1;
        // all corner loads are issued before the first result is stored
        T n0[3], n1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            n1[c] = force_bilinear(a.afield + c * g3, g, i0w, i1w, j0w, j1w, wrapi(k0 + 1, g), wx, wy);
        if (shift) {
#pragma unroll
            for (int c = 0; c < 3; ++c) n0[c] = fc.b1(c);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                n0[c] = force_bilinear(a.afield + c * g3, g, i0w, i1w, j0w, j1w, wrapi(k0, g), wx, wy);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) { fc.b0(c) = n0[c]; fc.b1(c) = n1[c]; }
        fc.k0 = k0;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) acc3[c] = a.ramp * (fc.b0(c) * (T(1) - wz) + fc.b1(c) * wz);
}

}  // namespace glm
