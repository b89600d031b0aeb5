// Render-side particle splats over the simulation's particle state: trilinear u32 fixed-point
// sums of particle mass on a dense grid.
//   SPLAT_DENSITY  calculateDensity (src/lib/gpu/marchingCubes/mcDensity.cs.wgsl:14-63): cell of
//                  the floor of the cell-centred coordinate, nodes outside the grid skipped.
//   SPLAT_MASS     calculateGridMass (src/lib/gpu/volumetric/calculateGridMass.wgsl:6-51): cell by
//                  u32 truncation (a coordinate in [-0.5, 0) maps to cell 0 with a negative
//                  weight, whose contribution u32() clamps to 0), linear weights
//                  (volumetric/prelude.wgsl:6-8), index x + rx (y + ry z) (mpmOps.wgsl:24-26).
// Contribution per node: u32((mass weight) S), truncated and clamped to [0, 2^32 - 1] like WGSL u32().
#pragma once
#include "mpm_kernels.h"

namespace mpm {

enum : int { SPLAT_DENSITY = 0, SPLAT_MASS = 1 };

// WGSL u32(f32): truncation toward zero, clamped to [0, 2^32 - 1]; NaN -> 0.
MPM_DEV uint32_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_splat(const float* __restrict__ s, uint32_t n, uint32_t stride, float gx0, float gy0,
                                               float gz0, float gx1, float gy1, float gz1, float S, uint32_t rx, uint32_t ry,
                                               uint32_t rz, uint32_t* __restrict__ grid) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const v3 pos = ld3(s, stride, PX, i);
    const float mass = ld(s, stride, PM, i);
    const v3 gmin = mk(gx0, gy0, gz0), gmax = mk(gx1, gy1, gz1);
    // any(pos < min) || any(pos >= max) returns; a NaN coordinate is skipped as well
    if (!(pos.x >= gmin.x) || !(pos.y >= gmin.y) || !(pos.z >= gmin.z)) return;
    if (pos.x >= gmax.x || pos.y >= gmax.y || pos.z >= gmax.z) return;
    const v3 cell = cdiv(gmax - gmin, mk((float)rx, (float)ry, (float)rz));
    const v3 pc = cdiv(pos - gmin, cell) - mk(0.5f, 0.5f, 0.5f);
    int st[3];
    v3 f;
    if (KIND == SPLAT_DENSITY) {
        st[0] = (int)__builtin_floorf(pc.x); st[1] = (int)__builtin_floorf(pc.y); st[2] = (int)__builtin_floorf(pc.z);
    } else {
        st[0] = (int)f2u_sat(pc.x); st[1] = (int)f2u_sat(pc.y); st[2] = (int)f2u_sat(pc.z);
    }
    f = pc - mk((float)st[0], (float)st[1], (float)st[2]);
    const v3 w0 = mk(1.0f - f.x, 1.0f - f.y, 1.0f - f.z);
#pragma unroll
    for (int z = 0; z < 2; ++z)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const int cx = st[0] + x, cy = st[1] + y, cz = st[2] + z;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)rx || cy >= (int)ry || cz >= (int)rz) continue;
        const float weight = ((x ? f.x : w0.x) * (y ? f.y : w0.y)) * (z ? f.z : w0.z);
        const uint32_t c = f2u_sat((mass * weight) * S);
        if (c) atomicAdd(&grid[(uint32_t)cx + rx * ((uint32_t)cy + ry * (uint32_t)cz)], c);
    }
}

}  // namespace mpm
