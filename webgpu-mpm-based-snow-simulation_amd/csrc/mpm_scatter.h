// Particle scatter init: scatterParticles (src/lib/gpu/particleInitialize/particleScatter.cs.wgsl:
// 7-84) with noise3 / fbm (shaderPrelude/randomNumberGenerators.wgsl:58-94) and the even-odd
// inside test by Moller-Trumbore ray casts (shaderPrelude/triangleOps.wgsl:1-38).
// Each particle draws up to 512 candidates uniformly in the mesh AABB (three hash1 steps of a
// hash4 seed per candidate), keeps the first inside the mesh (or the last), jitters it by fbm
// noise and gets mass (hash1(i) / 2^32)^2. f32(u32) rounds to nearest even and f32(0xFFFFFFFF) is
// 2^32, so u32 / f32(0xFFFFFFFF) is an exact scaling; mix(a, b, t) = a (1 - t) + b t.
#pragma once
#include "mpm_kernels.h"

namespace mpm {

MPM_DEV float unit_of(uint32_t h) { return (float)h * 2.3283064365386963e-10f; }   // h / 2^32
MPM_DEV float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
MPM_DEV float rand_unit(uint32_t& seed) { seed = hash1(seed); return unit_of(seed); }

MPM_DEV bool ray_hits_triangle(v3 o, v3 d, v3 v0, v3 v1, v3 v2) {
    const float EPS = 1e-6f;
    const v3 e1 = v1 - v0, e2 = v2 - v0;
    const v3 dxe2 = cross(d, e2);
    const float det = dot(e1, dxe2);
    if (fabs_(det) < EPS) return false;
    const float inv = 1.0f / det;
    const v3 ov = o - v0;
    const float b0 = dot(ov, dxe2) * inv;
    if (b0 < 0.0f || 1.0f < b0) return false;
    const v3 ovxe1 = cross(ov, e1);
    const float b1 = dot(d, ovxe1) * inv;
    if (b1 < 0.0f || 1.0f < b1 || b0 + b1 > 1.0f) return false;
    return dot(e2, ovxe1) * inv > EPS;
}

MPM_DEV bool inside_mesh(v3 p, const float4* __restrict__ mesh, uint32_t ntri) {
    const v3 d = mk(1.0f, 0.0f, 0.0f);
    bool in = false;
    for (uint32_t i = 0; i < ntri; ++i) {
        const float4 a = mesh[3 * i], b = mesh[3 * i + 1], c = mesh[3 * i + 2];
        if (ray_hits_triangle(p, d, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z))) in = !in;
    }
    return in;
}

MPM_DEV float noise3(v3 x) {
    const v3 p = mk(__builtin_floorf(x.x), __builtin_floorf(x.y), __builtin_floorf(x.z));
    const v3 f = x - p;
    const v3 u = mk((f.x * f.x) * (3.0f - 2.0f * f.x), (f.y * f.y) * (3.0f - 2.0f * f.y), (f.z * f.z) * (3.0f - 2.0f * f.z));
    const uint32_t nx = (uint32_t)(int32_t)p.x, ny = (uint32_t)(int32_t)p.y, nz = (uint32_t)(int32_t)p.z;
    const float a = unit_of(hash3(nx, ny, nz));
    const float b = unit_of(hash3(nx + 1u, ny, nz));
    const float c = unit_of(hash3(nx, ny + 1u, nz));
    const float d = unit_of(hash3(nx + 1u, ny + 1u, nz));
    const float e = unit_of(hash3(nx, ny, nz + 1u));
    const float fv = unit_of(hash3(nx + 1u, ny, nz + 1u));
    const float g = unit_of(hash3(nx, ny + 1u, nz + 1u));
    const float h = unit_of(hash3(nx + 1u, ny + 1u, nz + 1u));
    return mixf(mixf(mixf(a, b, u.x), mixf(c, d, u.x), u.y), mixf(mixf(e, fv, u.x), mixf(g, h, u.x), u.y), u.z);
}

MPM_DEV float fbm(v3 x) {
    float v = 0.0f, a = 0.5f;
    v3 p = x;
#pragma unroll 1
    for (int i = 0; i < 6; ++i) {
        v += a * noise3(p);
        p = p * 2.0f;
        a *= 0.5f;
    }
    return v;
}

// One particle record per thread, written to SoA row i (original order, id i).
__global__ __launch_bounds__(256) void k_scatter_init(const float4* __restrict__ mesh, uint32_t ntri, float mn_x, float mn_y,
                                                      float mn_z, float mx_x, float mx_y, float mx_z, float* __restrict__ s,
                                                      uint32_t* __restrict__ pid, float4* __restrict__ fpid, uint32_t n, uint32_t stride) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t seed = hash4(i, 100u, 200u, 300u);
    v3 c = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t a = 0; a < 512u; ++a) {
        const float rx = rand_unit(seed);
        const float ry = rand_unit(seed);
        const float rz = rand_unit(seed);
        c = mk(mixf(mn_x, mx_x, rx), mixf(mn_y, mx_y, ry), mixf(mn_z, mx_z, rz));
        if (inside_mesh(c, mesh, ntri)) break;
    }
    const v3 q = c * 4.0f;
    const float dx = fbm(q);
    const float dy = fbm(q + mk(100.0f, 200.0f, 300.0f));
    const float dz = fbm(q + mk(-300.0f, -200.0f, -100.0f));
    c = c + (mk(dx, dy, dz) - mk(0.5f, 0.5f, 0.5f)) * 0.1f;
    const float m1 = unit_of(hash1(i));
    st3(s, stride, PX, i, c);
    st3(s, stride, VX, i, mk(0.0f, 0.0f, 0.0f));
    st(s, stride, PM, i, m1 * m1);
    stm(s, stride, FE, i, ident());
    st_fp(fpid, i, ident());
    st(s, stride, JP, i, det3(ident()));
    st3(s, stride, DXX, i, mk(0.0f, 0.0f, 0.0f));
    stm(s, stride, DD, i, zero3());
    pid[i] = i;
}

}  // namespace mpm
