// Oracle of the particle scatter init and the render-side splats — TEST INFRASTRUCTURE ONLY
// (tests/ and bench.py load it).
//
// CPU restatement, FP32 with no contraction, of
//   src/lib/gpu/particleInitialize/particleScatter.cs.wgsl:7-84   randFloat, randVec3,
//       pointInsideMesh (even-odd count along +x), scatterParticles
//   src/lib/gpu/shaderPrelude/randomNumberGenerators.wgsl:58-94  noise3, fbm
//   src/lib/gpu/shaderPrelude/triangleOps.wgsl:1-38               rayIntersectsTriangle (Moller-Trumbore)
// Operation order follows the WGSL text left to right; mix(a, b, t) = a (1 - t) + b t (WGSL
// definition), dot = (x x + y y) + z z. f32(u32) rounds to nearest even and f32(0xFFFFFFFF) = 2^32,
// so every u32 / f32(0xFFFFFFFF) is an exact scaling by 2^-32.
//   src/lib/gpu/marchingCubes/mcDensity.cs.wgsl:14-63             calculateDensity
//   src/lib/gpu/volumetric/calculateGridMass.wgsl:6-51             calculateGridMass
// Pinned (tests/test_scatter_oracle.py) by: noise3 at lattice points equals the hash value,
// fbm's octave sum at lattice points, the mass closed form (hash1(i) 2^-32)^2, the first
// candidate of every particle from the integer hash stream, the inside test of an axis-aligned
// box against its closed form, and the jitter bound |d - 0.5| * 0.1.
#include <cmath>
#include <cstdint>
#include <cstring>
#include "orc_math.h"

using namespace orc;

namespace {

typedef V3<float> F3;

static inline float u2f_unit(uint32_t h) { return (float)h / 4294967296.0f; }     // f32(h) / f32(0xFFFFFFFF)
static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline float rand_float(uint32_t* seed) { *seed = hash1(*seed); return u2f_unit(*seed); }

static F3 rand_vec3(uint32_t* seed, F3 mn, F3 mx) {
    const float x = mixf(mn.x, mx.x, rand_float(seed));
    const float y = mixf(mn.y, mx.y, rand_float(seed));
    const float z = mixf(mn.z, mx.z, rand_float(seed));
    return {x, y, z};
}

static bool ray_hits_triangle(F3 o, F3 d, F3 v0, F3 v1, F3 v2) {
    const float EPS = 1e-6f;
    const F3 e1 = v1 - v0, e2 = v2 - v0;
    const F3 dxe2 = cross(d, e2);
    const float det = dot(e1, dxe2);
    if (std::fabs(det) < EPS) return false;
    const float inv = 1.0f / det;
    const F3 ov = o - v0;
    const float b0 = dot(ov, dxe2) * inv;
    if (b0 < 0.0f || 1.0f < b0) return false;
    const F3 ovxe1 = cross(ov, e1);
    const float b1 = dot(d, ovxe1) * inv;
    if (b1 < 0.0f || 1.0f < b1 || b0 + b1 > 1.0f) return false;
    const float t = dot(e2, ovxe1) * inv;
    return t > EPS;
}

static bool inside_mesh(F3 p, const float* v, uint32_t ntri) {
    const F3 d = {1.0f, 0.0f, 0.0f};
    bool in = false;
    for (uint32_t i = 0; i < ntri; ++i) {
        const float* a = v + 12 * i;
        if (ray_hits_triangle(p, d, {a[0], a[1], a[2]}, {a[4], a[5], a[6]}, {a[8], a[9], a[10]})) in = !in;
    }
    return in;
}

static float noise3(F3 x) {
    const F3 p = {std::floor(x.x), std::floor(x.y), std::floor(x.z)};
    const F3 f = x - p;
    const F3 u = {(f.x * f.x) * (3.0f - 2.0f * f.x), (f.y * f.y) * (3.0f - 2.0f * f.y), (f.z * f.z) * (3.0f - 2.0f * f.z)};
    const uint32_t nx = (uint32_t)(int32_t)p.x, ny = (uint32_t)(int32_t)p.y, nz = (uint32_t)(int32_t)p.z;
    const float a = u2f_unit(hash3(nx, ny, nz));
    const float b = u2f_unit(hash3(nx + 1u, ny, nz));
    const float c = u2f_unit(hash3(nx, ny + 1u, nz));
    const float d = u2f_unit(hash3(nx + 1u, ny + 1u, nz));
    const float e = u2f_unit(hash3(nx, ny, nz + 1u));
    const float fv = u2f_unit(hash3(nx + 1u, ny, nz + 1u));
    const float g = u2f_unit(hash3(nx, ny + 1u, nz + 1u));
    const float h = u2f_unit(hash3(nx + 1u, ny + 1u, nz + 1u));
    return mixf(mixf(mixf(a, b, u.x), mixf(c, d, u.x), u.y), mixf(mixf(e, fv, u.x), mixf(g, h, u.x), u.y), u.z);
}

static float fbm(F3 x) {
    float v = 0.0f, a = 0.5f;
    F3 p = x;
    for (int i = 0; i < 6; ++i) {
        v += a * noise3(p);
        p = p * 2.0f;
        a *= 0.5f;
    }
    return v;
}

}  // namespace

extern "C" {

float orc_noise3(float x, float y, float z) { return noise3({x, y, z}); }
float orc_fbm(float x, float y, float z) { return fbm({x, y, z}); }
int orc_inside_mesh(const float* p, const float* mesh_xyzw, uint32_t n_vertices) {
    return inside_mesh({p[0], p[1], p[2]}, mesh_xyzw, n_vertices / 3) ? 1 : 0;
}

// scatterParticles for particle indices [0, n): 192-byte records into aos (48 floats each).
void orc_scatter(const float* mesh_xyzw, uint32_t n_vertices, const float* mesh_min, const float* mesh_max,
                 uint32_t n, float* aos) {
    const uint32_t ntri = n_vertices / 3;
    const F3 mn = {mesh_min[0], mesh_min[1], mesh_min[2]}, mx = {mesh_max[0], mesh_max[1], mesh_max[2]};
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t seed = hash4(i, 100u, 200u, 300u);
        F3 c = {0.0f, 0.0f, 0.0f};
        for (uint32_t a = 0; a < 512u; ++a) {
            c = rand_vec3(&seed, mn, mx);
            if (inside_mesh(c, mesh_xyzw, ntri)) break;
        }
        const F3 q = c * 4.0f;
        const float dx = fbm(q);
        const float dy = fbm(q + F3{100.0f, 200.0f, 300.0f});
        const float dz = fbm(q + F3{-300.0f, -200.0f, -100.0f});
        c = c + (F3{dx, dy, dz} - F3{0.5f, 0.5f, 0.5f}) * 0.1f;
        const float m1 = u2f_unit(hash1(i));
        float* r = aos + (size_t)i * 48;
        std::memset(r, 0, 192);
        r[0] = c.x; r[1] = c.y; r[2] = c.z; r[3] = 1.0f;
        r[7] = m1 * m1;
        r[8] = 1.0f; r[13] = 1.0f; r[18] = 1.0f;       // F_e = I (columns at 8, 12, 16)
        r[20] = 1.0f; r[25] = 1.0f; r[30] = 1.0f;      // F_p = I (columns at 20, 24, 28)
    }
}

// WGSL u32(f32): truncation toward zero clamped to [0, 2^32 - 1].
static uint32_t u32_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// kind 0: calculateDensity on a res grid; kind 1: calculateGridMass. aos: 48 floats per particle.
void orc_splat(int kind, const float* aos, uint32_t n, const float* gmin, const float* gmax, float S,
               const uint32_t* res, uint32_t* out) {
    const uint64_t cells = (uint64_t)res[0] * res[1] * res[2];
    std::memset(out, 0, cells * 4);
    const F3 g0 = {gmin[0], gmin[1], gmin[2]}, g1 = {gmax[0], gmax[1], gmax[2]};
    const F3 cell = cdiv(g1 - g0, F3{(float)res[0], (float)res[1], (float)res[2]});
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = aos + (size_t)i * 48;
        const F3 pos = {r[0], r[1], r[2]};
        const float mass = r[7];
        if (!(pos.x >= g0.x) || !(pos.y >= g0.y) || !(pos.z >= g0.z)) continue;
        if (pos.x >= g1.x || pos.y >= g1.y || pos.z >= g1.z) continue;
        const F3 pc = cdiv(pos - g0, cell) - F3{0.5f, 0.5f, 0.5f};
        int st[3];
        if (kind == 0) { st[0] = (int)std::floor(pc.x); st[1] = (int)std::floor(pc.y); st[2] = (int)std::floor(pc.z); }
        else { st[0] = (int)u32_sat(pc.x); st[1] = (int)u32_sat(pc.y); st[2] = (int)u32_sat(pc.z); }
        const F3 f = pc - F3{(float)st[0], (float)st[1], (float)st[2]};
        const float w[2][3] = {{1.0f - f.x, 1.0f - f.y, 1.0f - f.z}, {f.x, f.y, f.z}};
        for (int z = 0; z < 2; ++z) for (int y = 0; y < 2; ++y) for (int x = 0; x < 2; ++x) {
            const int cx = st[0] + x, cy = st[1] + y, cz = st[2] + z;
            if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)res[0] || cy >= (int)res[1] || cz >= (int)res[2]) continue;
            const float weight = (w[x][0] * w[y][1]) * w[z][2];
            out[(uint64_t)cx + res[0] * ((uint64_t)cy + res[1] * (uint64_t)cz)] += u32_sat((mass * weight) * S);
        }
    }
}

}  // extern "C"
