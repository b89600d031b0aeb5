// Oracle — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's MPM substep (explicit MPM and PBMPM on the
// 4^3-block hashed sparse grid) used by tests/ and bench.py's cpu_baseline leg as the
// checker. It is never linked into, loaded by, or called from the product library.
//
// Two instantiations of one template:
//   mode 0  float,  i32 grid  : FP32-faithful restatement — same operation order as the WGSL
//                               source, -ffp-contract=off, IEEE division/sqrt, i32 fixed-point
//                               grid accumulators with truncating saturating conversion.
//   mode 1  double, i64 grid  : FP64 twin of the same algorithm (same fixed-point truncation).
//   mode 2  double, f64 grid  : FP64 physics of the same step with no fixed-point quantization.
//
// Reference functions restated (paths relative to src/lib/gpu/):
//   mpm/sparseGridPrelude.wgsl:26-64   block / cell indexing, hash lookup
//   mpm/clearHashMap.wgsl:5-19         hash map reset
//   mpm/mapAffectedBlocks.wgsl:7-86    block allocation (serial schedule of the atomics)
//   mpm/clearBlockParticleCounts.wgsl, countParticlesPerBlock.wgsl:7-23,
//   mpm/computeBlockOffsets.wgsl:5-15, binParticles.wgsl:8-25   counting sort by block
//   mpm/clearMappedBlocks.wgsl:9-26    grid clear (restricted to allocated blocks: same values)
//   mpm/particleToGrid.cs.wgsl:27-79   explicit P2G (corotated PK1 stress, hardening)
//   mpm/gridUpdate.cs.wgsl:35-53       explicit grid update
//   mpm/gridToParticle.cs.wgsl:28-130  G2P, explicit and PBMPM forms
//   mpm/pbmpm.wgsl:16-178              PBMPM constraint solve + P2G + gravity/pointer scatter
//   mpm/integrateParticles.wgsl:12-72  x/F update, plasticity, walls, collider
//   mpm/colliderPrelude.wgsl:37-102,153-356  closest point, ray/triangle, BVH traversal, response
//   shaderPrelude/stressTensorOps.wgsl:9-22,39-56, plasticityOps.wgsl:7-48
//   mpm/GpuMpmPipelineManager.ts:327-521     dispatch order of one substep
// Threads (mode 0 only, orc_set_threads): the passes run over particles / blocks with OpenMP, the
// grid sums as relaxed i32 atomic adds (integer sums are order-free, so every value equals the
// serial run), the block table with a CAS claim of an EMPTY slot (RESERVED while the claimant
// publishes, as mapAffectedBlocks.wgsl:7-62 does) and the binning cursors as atomic counters.
// Block numbering and the order within a block then vary; no result depends on either.
// Deliberate, documented departures (they change no value the reference defines):
//   * particles whose own cell is outside the grid are not binned and take no part in
//     P2G/G2P (the reference then re-processes a stale sortedParticleIndices entry);
//   * blocks beyond max_blocks are dropped and counted, as in mapAffectedBlocks.wgsl:27-31;
//   * integrateParticles is bounds-guarded.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <numeric>
#include <type_traits>
#include "orc_math.h"
#include "orc_svd.h"

using namespace orc;

extern "C" {

// Parameters that the reference hard-codes (defaults = reference values).
typedef struct orc_params {
    int32_t  pbmpm_iterations;      // GpuMpmPipelineManager.ts:445 (3)
    float    pbmpm_force_divisor;   // pbmpm.wgsl:137 (3.0)
    float    gravity[3];            // gridUpdate.cs.wgsl:43, pbmpm.wgsl:101 ((0,0,-9.81))
    double   youngs_modulus;        // stressTensorOps.wgsl:1 (1.4e5)
    double   poissons_ratio;        // stressTensorOps.wgsl:2 (0.2)
    float    hardening_explicit;    // stressTensorOps.wgsl:44 (5)
    float    hardening_pbmpm;       // pbmpm.wgsl:35 (20)
    float    pbmpm_blend;           // pbmpm.wgsl:26 (0.95)
    float    pbmpm_volume_exponent; // pbmpm.wgsl:23 (-0.3333333)
    double   theta_c;               // plasticityOps.wgsl:18 (2.5e-2)
    double   theta_s;               // plasticityOps.wgsl:19 (7.5e-3)
    double   density;               // particleToGrid.cs.wgsl:36 (400)
    float    grad_scale;            // multiplies D written by G2P (1 = reference)
    float    collider_margin;       // colliderPrelude.wgsl:158,297,330,351 (0.05)
    float    friction_ccd;          // colliderPrelude.wgsl:308 (0.1)
    float    friction_proximity;    // colliderPrelude.wgsl:339 (0.0)
    float    collider_velocity_factor; // colliderPrelude.wgsl:294 (0.2)
    uint32_t wall_mode;             // 0 = reference clamp+restitution; 1 = + sticky grid border
    uint32_t wall_cells;            // sticky border thickness in cells
    uint32_t hash_map_size;         // sparseGridPrelude.wgsl:6
    uint32_t max_blocks;            // sparseGridPrelude.wgsl:7
    uint32_t max_probes;            // sparseGridPrelude.wgsl:8 (100)
    uint32_t plasticity_mode;       // 0 reference polar-diag check, 1 SVD clamp, 2 SVD + FP64 fallback (orc_svd.h)
} orc_params;

typedef struct orc_stats {
    uint64_t blocks_allocated;
    uint64_t blocks_dropped;
    uint64_t particles_unbinned;
    uint64_t lookups_missed;
    uint64_t svd_fp64;              // SVD plasticity decompositions taken in FP64 (fallback)
} orc_stats;

}  // extern "C"

namespace {

static const uint32_t EMPTY = 0xFFFFFFFFu;
static const uint32_t RESERVED = 0xFFFFFFFEu;

static inline uint32_t ld_acq(const uint32_t* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
static inline void st_rel(uint32_t* p, uint32_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
static inline uint32_t fetch_add_u32(uint32_t* p, uint32_t v) { return __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
static inline void add_u64(uint64_t* p, uint64_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
static const int UNIFORM_BYTES = 37376;

struct Uniforms {   // byte offsets: GpuUniformsBufferManager.ts:54-182
    float dt, S;
    float grid_min[3], grid_max[3], cell_dims[3];
    uint32_t res[3];
    float coll_min[3], coll_max[3];
    float coll_T[16], coll_Tinv[16];
    float coll_vel[3];
    uint32_t coll_num_indices, coll_num_objects;
    float ip[3], istrength, iradius;
    uint32_t interacting, imode;
};

template <class T> static T rd(const uint8_t* b, int off) { T v; std::memcpy(&v, b + off, sizeof(T)); return v; }

static Uniforms parse_uniforms(const uint8_t* b) {
    Uniforms u;
    u.dt = rd<float>(b, 0);
    u.S = rd<float>(b, 4);
    for (int i = 0; i < 3; ++i) {
        u.grid_min[i] = rd<float>(b, 16 + 4 * i);
        u.grid_max[i] = rd<float>(b, 32 + 4 * i);
        u.res[i] = rd<uint32_t>(b, 208 + 4 * i);
        u.coll_min[i] = rd<float>(b, 224 + 4 * i);
        u.coll_max[i] = rd<float>(b, 240 + 4 * i);
        u.coll_vel[i] = rd<float>(b, 320 + 4 * i);
        u.cell_dims[i] = rd<float>(b, 352 + 4 * i);
        u.ip[i] = rd<float>(b, 496 + 4 * i);
    }
    for (int i = 0; i < 16; ++i) { u.coll_T[i] = rd<float>(b, 256 + 4 * i); u.coll_Tinv[i] = rd<float>(b, 432 + 4 * i); }
    u.coll_num_indices = rd<uint32_t>(b, 348);
    u.coll_num_objects = rd<uint32_t>(b, 364);
    u.istrength = rd<float>(b, 508);
    u.iradius = rd<float>(b, 512);
    u.interacting = rd<uint32_t>(b, 516);
    u.imode = rd<uint32_t>(b, 520);
    return u;
}

template <class R> struct Particle {
    V3<R> pos, vel; R mass;
    M3<R> Fe, Fp;
    V3<R> dx;
    M3<R> D;
    bool own_lame = false;  // per-particle Lame parameters (explicit stress) instead of the constants
    R mu0 = 0, la0 = 0;
};

// Grid accumulator policies.
//   AccI32: the reference's i32 fixed point, i32((x) * S) truncating + wrapping atomics (FP32 mode).
//   AccI64: the same fixed-point truncation evaluated in FP64 (FP64 twin of the algorithm).
//   AccF64: no quantization (FP64 physics of the same step).
struct AccI32 {
    typedef int32_t T;
    static T conv(float x, float S) { return f2i(x * S); }
    static float val(T a, float S) { return (float)a / S; }
    static T add(T a, T b) { return (T)((uint32_t)a + (uint32_t)b); }
    static T rewrite(float mS) { return f2i(mS); }
    static constexpr bool scaled = true;
};
struct AccI64 {
    typedef int64_t T;
    static T conv(double x, float S) { double y = x * (double)S; return (y != y) ? 0 : (T)std::trunc(y); }
    static double val(T a, float S) { return (double)a / (double)S; }
    static T add(T a, T b) { return a + b; }
    static T rewrite(double mS) { return (mS != mS) ? 0 : (T)std::trunc(mS); }
    static constexpr bool scaled = true;
};
struct AccF64 {
    typedef double T;
    static T conv(double x, float) { return x; }
    static double val(T a, float) { return a; }
    static T add(T a, T b) { return a + b; }
    static T rewrite(double m) { return m; }
    static constexpr bool scaled = false;
};

template <class R> static V3<R> mv4(const float* T, V3<R> p) {   // (mat4 * vec4(p,1)).xyz
    V3<R> c0 = {(R)T[0], (R)T[1], (R)T[2]}, c1 = {(R)T[4], (R)T[5], (R)T[6]};
    V3<R> c2 = {(R)T[8], (R)T[9], (R)T[10]}, c3 = {(R)T[12], (R)T[13], (R)T[14]};
    return ((c0 * p.x + c1 * p.y) + c2 * p.z) + c3 * (R)1;
}

template <class R, class AP> struct Sim {
    typedef typename AP::T A;
    orc_params P;
    Uniforms U;
    std::vector<Particle<R>> parts;
    // hashed sparse grid (sparseGridPrelude.wgsl:12-24)
    std::vector<int32_t> hm_bx, hm_by, hm_bz;
    std::vector<uint32_t> hm_index;
    uint32_t n_alloc = 0;
    std::vector<uint32_t> mapped, counts, offsets, sorted;
    uint32_t n_sorted = 0;
    std::vector<A> gm, gx, gy, gz;
    std::vector<uint32_t> coll_data;
    std::vector<uint32_t> bvh;   // 8 words per node
    orc_stats st{};
    int threads = 1;             // > 1: OpenMP passes with atomic grid sums (i32 grid only)
    bool par() const { return threads > 1; }

    R dt() const { return (R)U.dt; }
    V3<R> gmin() const { return {(R)U.grid_min[0], (R)U.grid_min[1], (R)U.grid_min[2]}; }
    V3<R> gmax() const { return {(R)U.grid_max[0], (R)U.grid_max[1], (R)U.grid_max[2]}; }
    V3<R> cdims() const { return {(R)U.cell_dims[0], (R)U.cell_dims[1], (R)U.cell_dims[2]}; }

    void resize_grid() {
        hm_bx.assign(P.hash_map_size, 0); hm_by.assign(P.hash_map_size, 0); hm_bz.assign(P.hash_map_size, 0);
        hm_index.assign(P.hash_map_size, EMPTY);
        mapped.assign(P.max_blocks, 0); counts.assign(P.max_blocks, 0); offsets.assign(P.max_blocks, 0);
        size_t cells = (size_t)P.max_blocks * 64;
        gm.assign(cells, A(0)); gx.assign(cells, A(0)); gy.assign(cells, A(0)); gz.assign(cells, A(0));
    }

    // ------------------------------------------------ mpmOps.wgsl:14-53
    void cell_number(V3<R> pos, int32_t c[3]) const {
        V3<R> p = pos - gmin();
        V3<R> q = cdiv(p, cdims());
        c[0] = f2i(q.x); c[1] = f2i(q.y); c[2] = f2i(q.z);
    }
    bool in_range(const int32_t c[3]) const {
        return c[0] >= 0 && c[1] >= 0 && c[2] >= 0 &&
               c[0] < (int32_t)U.res[0] && c[1] < (int32_t)U.res[1] && c[2] < (int32_t)U.res[2];
    }
    V3<R> frac_pos(V3<R> pos, const int32_t c[3]) const {
        V3<R> cf = {(R)c[0], (R)c[1], (R)c[2]};
        V3<R> minPos = gmin() + cmul(cdims(), cf);
        return cdiv(pos - minPos, cdims());
    }
    static void weights(V3<R> f, V3<R> w[3]) {
        const V3<R> one = {1, 1, 1}, half = {(R)0.5, (R)0.5, (R)0.5};
        V3<R> a = one - f;
        w[0] = cmul((R)0.5 * a, a);
        V3<R> b = f - half;
        w[1] = V3<R>{(R)0.75, (R)0.75, (R)0.75} - cmul(b, b);
        w[2] = cmul((R)0.5 * f, f);
    }
    static void weight_derivs(V3<R> f, V3<R> d[3]) {
        const V3<R> one = {1, 1, 1}, half = {(R)0.5, (R)0.5, (R)0.5};
        d[0] = f - one;
        d[1] = (R)-2 * (f - half);
        d[2] = f;
    }

    // ------------------------------------------------ sparseGridPrelude.wgsl:35-64
    uint32_t lookup_block(int32_t bx, int32_t by, int32_t bz) {
        uint32_t h = hash3((uint32_t)bx, (uint32_t)by, (uint32_t)bz);
        for (uint32_t i = 0; i < P.max_probes; ++i) {
            uint32_t ci = (h + i) % P.hash_map_size;
            uint32_t bi = hm_index[ci];
            if (bi == EMPTY) return EMPTY;
            if (hm_bx[ci] == bx && hm_by[ci] == by && hm_bz[ci] == bz) return bi;
        }
        return EMPTY;
    }
    uint32_t cell_index(const int32_t c[3]) {
        uint32_t bi = lookup_block(c[0] >> 2, c[1] >> 2, c[2] >> 2);
        if (bi == EMPTY) { if (par()) add_u64(&st.lookups_missed, 1); else st.lookups_missed++; return EMPTY; }
        uint32_t local = (uint32_t)((c[0] & 3) + (c[1] & 3) * 4 + (c[2] & 3) * 16);
        return bi * 64u + local;
    }

    // ------------------------------------------------ clearHashMap + mapAffectedBlocks
    void allocate_block(int32_t bx, int32_t by, int32_t bz) {
        uint32_t h = hash3((uint32_t)bx, (uint32_t)by, (uint32_t)bz);
        uint32_t cand = h % P.hash_map_size;
        for (uint32_t i = 0; i < P.max_probes; ++i) {
            uint32_t cur = hm_index[cand];
            if (cur == EMPTY) {
                uint32_t nb = n_alloc++;
                if (nb >= P.max_blocks) { n_alloc = P.max_blocks; st.blocks_dropped++; return; }
                hm_bx[cand] = bx; hm_by[cand] = by; hm_bz[cand] = bz;
                mapped[nb] = cand;
                hm_index[cand] = nb;
                return;
            }
            if (hm_bx[cand] == bx && hm_by[cand] == by && hm_bz[cand] == bz) return;
            cand = (cand + 1 >= P.hash_map_size) ? 0 : cand + 1;
        }
    }
    // The threaded insert: claim an EMPTY slot by CAS to RESERVED, take a block number, publish it
    // (a claim beyond max_blocks is undone to EMPTY, mapAffectedBlocks.wgsl:27-31); a RESERVED
    // slot is waited on until its claimant publishes; an undone slot is probed again.
    void allocate_block_par(int32_t bx, int32_t by, int32_t bz) {
        uint32_t h = hash3((uint32_t)bx, (uint32_t)by, (uint32_t)bz);
        uint32_t cand = h % P.hash_map_size;
        for (uint32_t i = 0; i < P.max_probes;) {
            uint32_t cur = ld_acq(&hm_index[cand]);
            if (cur == EMPTY) {
                uint32_t expect = EMPTY;
                if (__atomic_compare_exchange_n(&hm_index[cand], &expect, RESERVED, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
                    const uint32_t nb = fetch_add_u32(&n_alloc, 1u);
                    if (nb >= P.max_blocks) { add_u64(&st.blocks_dropped, 1); st_rel(&hm_index[cand], EMPTY); return; }
                    hm_bx[cand] = bx; hm_by[cand] = by; hm_bz[cand] = bz;
                    mapped[nb] = cand;
                    st_rel(&hm_index[cand], nb);
                    return;
                }
                cur = expect;
            }
            while (cur == RESERVED) cur = ld_acq(&hm_index[cand]);
            if (cur == EMPTY) continue;                       // undone claim: the same slot again
            if (hm_bx[cand] == bx && hm_by[cand] == by && hm_bz[cand] == bz) return;
            cand = (cand + 1 >= P.hash_map_size) ? 0 : cand + 1;
            ++i;
        }
    }
    void bin() {
        std::fill(hm_index.begin(), hm_index.end(), EMPTY);
        n_alloc = 0;
        const size_t N = parts.size();
        const bool pr = par();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads) if (pr)
        for (size_t p = 0; p < N; ++p) {
            int32_t c0[3]; cell_number(parts[p].pos, c0);
            for (int oz = -1; oz <= 1; ++oz) for (int oy = -1; oy <= 1; ++oy) for (int ox = -1; ox <= 1; ++ox) {
                int32_t c[3] = {wadd(c0[0], ox), wadd(c0[1], oy), wadd(c0[2], oz)};
                if (!in_range(c)) continue;
                if (pr) allocate_block_par(c[0] >> 2, c[1] >> 2, c[2] >> 2);
                else allocate_block(c[0] >> 2, c[1] >> 2, c[2] >> 2);
            }
        }
        if (n_alloc > P.max_blocks) n_alloc = P.max_blocks;
        st.blocks_allocated = n_alloc;
        std::fill(counts.begin(), counts.end(), 0u);
        std::vector<uint32_t> own(N, EMPTY);
#pragma omp parallel for schedule(static) num_threads(threads) if (pr)
        for (size_t p = 0; p < N; ++p) {
            int32_t c[3]; cell_number(parts[p].pos, c);
            // countParticlesPerBlock.wgsl:16-18: own block looked up with no range check
            uint32_t bi = lookup_block(c[0] >> 2, c[1] >> 2, c[2] >> 2);
            own[p] = bi;
            if (bi != EMPTY) { if (pr) fetch_add_u32(&counts[bi], 1u); else counts[bi]++; }
        }
        uint32_t off = 0;
        for (uint32_t b = 0; b < n_alloc; ++b) { offsets[b] = off; off += counts[b]; }
        sorted.assign(N, 0);
        n_sorted = off;
        st.particles_unbinned = N - off;
        std::vector<uint32_t> cur(offsets.begin(), offsets.begin() + n_alloc);
#pragma omp parallel for schedule(static) num_threads(threads) if (pr)
        for (size_t p = 0; p < N; ++p)
            if (own[p] != EMPTY) sorted[pr ? fetch_add_u32(&cur[own[p]], 1u) : cur[own[p]]++] = (uint32_t)p;
    }
    void clear_grid() {
        size_t cells = (size_t)n_alloc * 64;
        std::fill(gm.begin(), gm.begin() + cells, A(0));
        std::fill(gx.begin(), gx.begin() + cells, A(0));
        std::fill(gy.begin(), gy.begin() + cells, A(0));
        std::fill(gz.begin(), gz.begin() + cells, A(0));
    }
    void dep(uint32_t ci, A* g, A v) {
        if constexpr (std::is_same<A, int32_t>::value) {
            if (par()) { fetch_add_u32((uint32_t*)&g[ci], (uint32_t)v); return; }   // wrapping, like the i32 atomics
        }
        g[ci] = AP::add(g[ci], v);
    }
    A conv_acc(R x, float S) const { return AP::conv(x, S); }

    // ------------------------------------------------ particleToGrid.cs.wgsl:27-79
    // stressTensorOps.wgsl:39-56 (hardened Lame parameters) and :9-22 (corotated PK1 stress)
    M3<R> pk1_explicit(const Particle<R>& pt) const {
        R mu = (R)(float)(P.youngs_modulus / (2.0 * (1.0 + P.poissons_ratio)));
        R la = (R)(float)(P.youngs_modulus * P.poissons_ratio / ((1.0 + P.poissons_ratio) * (1.0 - 2.0 * P.poissons_ratio)));
        if (pt.own_lame) { mu = pt.mu0; la = pt.la0; }
        R Jp = det3(pt.Fp);
        R ef = r_exp((R)P.hardening_explicit * ((R)1 - Jp));
        mu = mu * ef; la = la * ef;
        R J = det3(pt.Fe);
        M3<R> Rr = polar_rotation(pt.Fe);
        return madd(mscale((R)2 * mu, msub(pt.Fe, Rr)), mscale((la * (J - (R)1)) * J, transpose(inverse3(pt.Fe))));
    }
    void p2g_explicit(const Particle<R>& pt) {
        const R S = (R)U.S;
        int32_t c0[3]; cell_number(pt.pos, c0);
        V3<R> f = frac_pos(pt.pos, c0);
        V3<R> w[3], d[3]; weights(f, w); weight_derivs(f, d);
        M3<R> PT = transpose(pk1_explicit(pt));
        R V = pt.mass * (R)(float)(1.0 / P.density);
        M3<R> negVPT = mscale(-V, PT);
        V3<R> pm = pt.mass * pt.vel;
        for (int oz = -1; oz <= 1; ++oz) for (int oy = -1; oy <= 1; ++oy) for (int ox = -1; ox <= 1; ++ox) {
            int32_t c[3] = {wadd(c0[0], ox), wadd(c0[1], oy), wadd(c0[2], oz)};
            if (!in_range(c)) continue;
            uint32_t ci = cell_index(c);
            if (ci == EMPTY) continue;
            R cw = (w[ox + 1].x * w[oy + 1].y) * w[oz + 1].z;
            V3<R> g = {(d[ox + 1].x * w[oy + 1].y) * w[oz + 1].z,
                       (w[ox + 1].x * d[oy + 1].y) * w[oz + 1].z,
                       (w[ox + 1].x * w[oy + 1].y) * d[oz + 1].z};
            g = cdiv(g, cdims());
            V3<R> sf = mvmul(negVPT, g);
            V3<R> sm = sf * dt();
            V3<R> mom = cw * pm + sm;
            dep(ci, gx.data(), AP::conv(mom.x, S));
            dep(ci, gy.data(), AP::conv(mom.y, S));
            dep(ci, gz.data(), AP::conv(mom.z, S));
            dep(ci, gm.data(), AP::conv((cw * pt.mass), S));
        }
    }

    bool sticky(const int32_t c[3]) const {
        if (!(P.wall_mode & 1u)) return false;
        const int32_t b = (int32_t)P.wall_cells;
        for (int a = 0; a < 3; ++a) if (c[a] < b || c[a] >= (int32_t)U.res[a] - b) return true;
        return false;
    }

    // ------------------------------------------------ gridUpdate.cs.wgsl:35-53
    void grid_update_explicit() {
        const R S = (R)U.S;
        const V3<R> g = {(R)P.gravity[0], (R)P.gravity[1], (R)P.gravity[2]};
#pragma omp parallel for schedule(static) num_threads(threads) if (par())
        for (uint32_t b = 0; b < n_alloc; ++b) {
            uint32_t slot = mapped[b];
            for (uint32_t l = 0; l < 64; ++l) {
                uint32_t ci = b * 64 + l;
                R m = AP::val(gm[ci], S);
                if (m <= (R)0) continue;
                V3<R> mom = {AP::val(gx[ci], S), AP::val(gy[ci], S), AP::val(gz[ci], S)};
                V3<R> v = mom / m;
                v = v + g * dt();
                int32_t c[3] = {hm_bx[slot] * 4 + (int32_t)(l & 3), hm_by[slot] * 4 + (int32_t)((l >> 2) & 3), hm_bz[slot] * 4 + (int32_t)(l >> 4)};
                if (sticky(c)) v = V3<R>{0, 0, 0};
                V3<R> nm = AP::scaled ? (v * m) * S : v * m;
                gx[ci] = AP::rewrite(nm.x); gy[ci] = AP::rewrite(nm.y); gz[ci] = AP::rewrite(nm.z);
            }
        }
    }

    // grid node velocity as read by G2P (gridToParticle.cs.wgsl:43-49 / :102-108)
    bool node_velocity(const int32_t c[3], V3<R>& v) {
        if (!in_range(c)) return false;
        uint32_t ci = cell_index(c);
        if (ci == EMPTY) return false;
        const R S = (R)U.S;
        R m = AP::val(gm[ci], S);
        if (m <= (R)0) return false;
        V3<R> mom = {AP::val(gx[ci], S), AP::val(gy[ci], S), AP::val(gz[ci], S)};
        v = mom / m;
        if ((P.wall_mode & 1u) && sticky(c)) v = V3<R>{0, 0, 0};
        return true;
    }

    // ------------------------------------------------ gridToParticle.cs.wgsl:28-80
    void g2p_explicit(Particle<R>& pt) {
        int32_t c0[3]; cell_number(pt.pos, c0);
        V3<R> f = frac_pos(pt.pos, c0);
        V3<R> w[3], d[3]; weights(f, w); weight_derivs(f, d);
        V3<R> nv = {0, 0, 0};
        M3<R> G = zero3<R>();
        for (int oz = -1; oz <= 1; ++oz) for (int oy = -1; oy <= 1; ++oy) for (int ox = -1; ox <= 1; ++ox) {
            int32_t c[3] = {wadd(c0[0], ox), wadd(c0[1], oy), wadd(c0[2], oz)};
            V3<R> v;
            if (!node_velocity(c, v)) continue;
            R cw = (w[ox + 1].x * w[oy + 1].y) * w[oz + 1].z;
            nv = nv + cw * v;
            V3<R> g = {(d[ox + 1].x * w[oy + 1].y) * w[oz + 1].z,
                       (w[ox + 1].x * d[oy + 1].y) * w[oz + 1].z,
                       (w[ox + 1].x * w[oy + 1].y) * d[oz + 1].z};
            M3<R> o; o.c[0] = g.x * v; o.c[1] = g.y * v; o.c[2] = g.z * v;
            G = madd(G, o);
        }
        pt.vel = nv;
        pt.dx = nv * dt();
        pt.D = mscale_r(mscale_r(G, dt()), (R)P.grad_scale);
    }

    // ------------------------------------------------ gridToParticle.cs.wgsl:82-130
    void g2p_pbmpm(Particle<R>& pt) {
        int32_t c0[3]; cell_number(pt.pos, c0);
        V3<R> f = frac_pos(pt.pos, c0);
        V3<R> w[3]; weights(f, w);
        V3<R> pcp = (V3<R>{(R)c0[0], (R)c0[1], (R)c0[2]} + f) - V3<R>{(R)0.5, (R)0.5, (R)0.5};
        V3<R> nv = {0, 0, 0};
        M3<R> B = zero3<R>();
        for (int oz = -1; oz <= 1; ++oz) for (int oy = -1; oy <= 1; ++oy) for (int ox = -1; ox <= 1; ++ox) {
            int32_t c[3] = {wadd(c0[0], ox), wadd(c0[1], oy), wadd(c0[2], oz)};
            V3<R> v;
            if (!node_velocity(c, v)) continue;
            R cw = (w[ox + 1].x * w[oy + 1].y) * w[oz + 1].z;
            nv = nv + cw * v;
            V3<R> dist = V3<R>{(R)c[0], (R)c[1], (R)c[2]} - pcp;
            B = madd(B, outer(cw * v, dist));
        }
        pt.dx = nv * dt();
        pt.D = mscale_r(mscale_r(mscale_r(B, dt()), (R)4), (R)P.grad_scale);
        pt.vel = nv;
    }

    // ------------------------------------------------ pbmpm.wgsl:16-39
    void solve_constraints(Particle<R>& pt) {
        M3<R> trial = mmul(madd(ident<R>(), pt.D), pt.Fe);
        M3<R> rot = polar_rotation(trial);
        R tv = det3(trial);
        R vs = r_pow(std::fabs(tv), (R)P.pbmpm_volume_exponent);
        M3<R> ts = mscale_r(rot, vs);
        R bf = (R)P.pbmpm_blend;
        M3<R> tb = madd(mscale(bf, ts), mscale((R)1 - bf, rot));
        M3<R> corr = msub(mmul(tb, inverse3(pt.Fe)), ident<R>());
        R Jp = det3(pt.Fp);
        M3<R> diff = msub(corr, pt.D);
        R relax = (R)0.5 + (R)0.5 * ((R)1 - r_exp(-(R)P.hardening_pbmpm / Jp));
        pt.D = madd(pt.D, mscale(relax, diff));
    }
    // ------------------------------------------------ pbmpm.wgsl:41-76 and :79-147
    void p2g_pbmpm(const Particle<R>& pt) {
        const R S = (R)U.S;
        int32_t c0[3]; cell_number(pt.pos, c0);
        V3<R> f = frac_pos(pt.pos, c0);
        V3<R> w[3]; weights(f, w);
        V3<R> pcp = (V3<R>{(R)c0[0], (R)c0[1], (R)c0[2]} + f) - V3<R>{(R)0.5, (R)0.5, (R)0.5};
        for (int oz = -1; oz <= 1; ++oz) for (int oy = -1; oy <= 1; ++oy) for (int ox = -1; ox <= 1; ++ox) {
            int32_t c[3] = {wadd(c0[0], ox), wadd(c0[1], oy), wadd(c0[2], oz)};
            if (!in_range(c)) continue;
            uint32_t ci = cell_index(c);
            if (ci == EMPTY) continue;
            R cw = (w[ox + 1].x * w[oy + 1].y) * w[oz + 1].z;
            R wm = cw * pt.mass;
            V3<R> off = V3<R>{(R)c[0], (R)c[1], (R)c[2]} - pcp;
            V3<R> aff = mvmul(pt.D, off);
            V3<R> mom = (wm * (pt.dx + aff)) / dt();
            dep(ci, gx.data(), AP::conv(mom.x, S));
            dep(ci, gy.data(), AP::conv(mom.y, S));
            dep(ci, gz.data(), AP::conv(mom.z, S));
            dep(ci, gm.data(), AP::conv(wm, S));
        }
        for (int oz = -1; oz <= 1; ++oz) for (int oy = -1; oy <= 1; ++oy) for (int ox = -1; ox <= 1; ++ox) {
            int32_t c[3] = {wadd(c0[0], ox), wadd(c0[1], oy), wadd(c0[2], oz)};
            if (!in_range(c)) continue;
            uint32_t ci = cell_index(c);
            if (ci == EMPTY) continue;
            R cw = (w[ox + 1].x * w[oy + 1].y) * w[oz + 1].z;
            R wm = cw * pt.mass;
            V3<R> forces = {(R)P.gravity[0], (R)P.gravity[1], (R)P.gravity[2]};
            if (U.interacting != 0u) forces = forces + pointer_force(c);
            forces = forces / (R)P.pbmpm_force_divisor;
            V3<R> mc = (forces * wm) * dt();
            dep(ci, gx.data(), AP::conv(mc.x, S));
            dep(ci, gy.data(), AP::conv(mc.y, S));
            dep(ci, gz.data(), AP::conv(mc.z, S));
        }
    }
    // pbmpm.wgsl:104-134 (node position in grid-cell units)
    V3<R> pointer_force(const int32_t c[3]) const {
        V3<R> node = {(R)c[0], (R)c[1], (R)c[2]};
        V3<R> ip = {(R)U.ip[0], (R)U.ip[1], (R)U.ip[2]};
        R dist = length(node - ip);
        V3<R> applied = {0, 0, 0};
        if (dist < (R)U.iradius) {
            V3<R> off = node - ip;
            V3<R> dir = {0, 0, 1};
            if (length(off) > (R)0.001f) dir = normalize(off);
            R fl = dist / (R)U.iradius;
            R fall = (R)1 - fl * fl;
            if (U.imode == 0u) applied = (dir * (R)U.istrength) * fall;
            else applied = ((((R)-1 * dir) * (R)U.istrength) * fall) * ((R)1 - fall);
        }
        return applied;
    }

    // ------------------------------------------------ plasticityOps.wgsl:7-48
    void plasticity(Particle<R>& pt) {
        if (P.plasticity_mode != 0u) { plasticity_svd_option(pt); return; }
        M3<R> dE = pt.Fe, dP = pt.Fp;
        M3<R> rot = polar_rotation(dE);
        M3<R> sc = mmul(transpose(rot), dE);
        const R lo = (R)(float)(1.0 - P.theta_c), hi = (R)(float)(1.0 + P.theta_s);
        V3<R> s = {sc.c[0].x, sc.c[1].y, sc.c[2].z};
        if (lo <= s.x && lo <= s.y && lo <= s.z && s.x <= hi && s.y <= hi && s.z <= hi) return;
        auto cl = [&](R v) { R t = (v < lo) ? lo : v; return (hi < t) ? hi : t; };
        M3<R> ns = zero3<R>();
        ns.c[0].x = cl(s.x); ns.c[1].y = cl(s.y); ns.c[2].z = cl(s.z);
        M3<R> nFe = mmul(rot, ns);
        M3<R> nFeInv = inverse3(nFe);
        M3<R> F = mmul(dE, dP);
        pt.Fe = nFe;
        pt.Fp = mmul(nFeInv, F);
    }

    // The SVD plasticity option (orc_svd.h). The FP64 modes of the oracle decompose in double.
    void plasticity_svd_option(Particle<R>& pt) {
        const R lo = (R)(float)(1.0 - P.theta_c), hi = (R)(float)(1.0 + P.theta_s);
        if constexpr (std::is_same<R, float>::value) {
            bool fp64 = false;
            plasticity_svd(pt.Fe, pt.Fp, lo, hi, P.plasticity_mode == 2u, &fp64);
            if (fp64) { if (par()) add_u64(&st.svd_fp64, 1); else st.svd_fp64++; }
        } else {
            const Svd3<R> d = svd3(pt.Fe);
            if (lo <= d.sig[0] && lo <= d.sig[1] && lo <= d.sig[2] && d.sig[0] <= hi && d.sig[1] <= hi && d.sig[2] <= hi) return;
            const M3<R> nFe = svd_clamped(d, lo, hi);
            pt.Fp = mmul(inverse3(nFe), mmul(pt.Fe, pt.Fp));
            pt.Fe = nFe;
        }
    }

    // ------------------------------------------------ colliderPrelude.wgsl:37-102
    static V3<R> closest_point_triangle(V3<R> p, V3<R> a, V3<R> b, V3<R> c) {
        V3<R> ab = b - a, ac = c - a, ap = p - a;
        R d1 = dot(ab, ap), d2 = dot(ac, ap);
        if (d1 <= (R)0 && d2 <= (R)0) return a;
        V3<R> bp = p - b;
        R d3 = dot(ab, bp), d4 = dot(ac, bp);
        if (d3 >= (R)0 && d4 <= d3) return b;
        R vc = d1 * d4 - d3 * d2;
        if (vc <= (R)0 && d1 >= (R)0 && d3 <= (R)0) { R v = d1 / (d1 - d3); return a + v * ab; }
        V3<R> cp = p - c;
        R d5 = dot(ab, cp), d6 = dot(ac, cp);
        if (d6 >= (R)0 && d5 <= d6) return c;
        R vb = d5 * d2 - d1 * d6;
        if (vb <= (R)0 && d2 >= (R)0 && d6 <= (R)0) { R w = d2 / (d2 - d6); return a + w * ac; }
        R va = d3 * d6 - d5 * d4;
        if (va <= (R)0 && (d4 - d3) >= (R)0 && (d5 - d6) >= (R)0) {
            R w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            return b + w * (c - b);
        }
        R denom = (R)1 / ((va + vb) + vc);
        R v = vb * denom, w = vc * denom;
        return (a + v * ab) + w * ac;
    }
    static R intersect_ray_triangle(V3<R> o, V3<R> dir, V3<R> a, V3<R> b, V3<R> c) {
        const R EPS = (R)1e-6f;
        V3<R> e1 = b - a, e2 = c - a;
        V3<R> h = cross(dir, e2);
        R aa = dot(e1, h);
        if (aa > -EPS && aa < EPS) return (R)-1;
        R f = (R)1 / aa;
        V3<R> s = o - a;
        R u = f * dot(s, h);
        if (u < (R)0 || u > (R)1) return (R)-1;
        V3<R> q = cross(s, e1);
        R v = f * dot(dir, q);
        if (v < (R)0 || u + v > (R)1) return (R)-1;
        R t = f * dot(e2, q);
        if (t > EPS) return t;
        return (R)-1;
    }
    V3<R> collider_vertex(uint32_t idx) const {
        uint32_t base = U.coll_num_indices + idx * 3u;
        return {(R)bits_f(coll_data[base]), (R)bits_f(coll_data[base + 1]), (R)bits_f(coll_data[base + 2])};
    }
    static R rsign(R x) { return x > (R)0 ? (R)1 : (x < (R)0 ? (R)-1 : (R)0); }

    // colliderPrelude.wgsl:153-356
    void resolve_collision(Particle<R>& pt) {
        if (bvh.empty() || coll_data.empty()) return;
        V3<R> cmin = {(R)U.coll_min[0], (R)U.coll_min[1], (R)U.coll_min[2]};
        V3<R> cmax = {(R)U.coll_max[0], (R)U.coll_max[1], (R)U.coll_max[2]};
        V3<R> min_b = mv4(U.coll_T, cmin), max_b = mv4(U.coll_T, cmax);
        const R margin = (R)P.collider_margin;
        V3<R> mg = {margin, margin, margin};
        V3<R> smin = vmin(min_b, max_b) - mg, smax = vmax(min_b, max_b) + mg;
        V3<R> cur = pt.pos;
        V3<R> prev = cur - pt.dx;
        V3<R> minp = vmin(cur, prev), maxp = vmax(cur, prev);
        if (minp.x > smax.x || minp.y > smax.y || minp.z > smax.z || maxp.x < smin.x || maxp.y < smin.y || maxp.z < smin.z) return;
        if (U.coll_num_objects == 0u) return;
        R min_d2 = (R)1e20f;
        V3<R> cl_n = {0, 0, 1}, cl_p = cur;
        R min_t = (R)1;
        V3<R> hit_n = {0, 0, 1}, hit_p = cur;
        bool has_hit = false;
        V3<R> ray = pt.dx;
        R rl = length(ray);
        bool do_ccd = rl > (R)1e-4f;
        V3<R> cur_l = mv4(U.coll_Tinv, cur), prev_l = mv4(U.coll_Tinv, prev);
        uint32_t stack[32]; uint32_t sp = 0; stack[sp++] = 0;
        V3<R> pmin = vmin(cur_l, prev_l), pmax = vmax(cur_l, prev_l);
        const size_t n_nodes = bvh.size() / 8;
        while (sp > 0) {
            uint32_t ni = stack[--sp];
            if (ni >= n_nodes) continue;   // bounds guard (reference relies on robust access)
            const uint32_t* nd = &bvh[(size_t)ni * 8];
            V3<R> bmin = V3<R>{(R)bits_f(nd[0]), (R)bits_f(nd[1]), (R)bits_f(nd[2])} - mg;
            V3<R> bmax = V3<R>{(R)bits_f(nd[4]), (R)bits_f(nd[5]), (R)bits_f(nd[6])} + mg;
            if (pmin.x > bmax.x || pmin.y > bmax.y || pmin.z > bmax.z || pmax.x < bmin.x || pmax.y < bmin.y || pmax.z < bmin.z) continue;
            if (nd[7] & 0x80000000u) {
                uint32_t start = nd[3], end = start + (nd[7] & 0x7FFFFFFFu);
                for (uint32_t i = start; i < end; ++i) {
                    V3<R> v0 = mv4(U.coll_T, collider_vertex(coll_data[i * 3u]));
                    V3<R> v1 = mv4(U.coll_T, collider_vertex(coll_data[i * 3u + 1u]));
                    V3<R> v2 = mv4(U.coll_T, collider_vertex(coll_data[i * 3u + 2u]));
                    V3<R> cp = closest_point_triangle(cur, v0, v1, v2);
                    V3<R> df = cur - cp;
                    R d2 = dot(df, df);
                    if (d2 < min_d2) {
                        min_d2 = d2; cl_p = cp;
                        V3<R> n = cross(v1 - v0, v2 - v0);
                        R len = length(n);
                        if (len > (R)1e-6f) cl_n = n / len;
                    }
                    if (do_ccd) {
                        R t = intersect_ray_triangle(prev, ray, v0, v1, v2);
                        if (t > (R)0 && t < min_t) {
                            min_t = t; has_hit = true;
                            hit_p = prev + ray * t;
                            V3<R> fn = normalize(cross(v1 - v0, v2 - v0));
                            hit_n = fn * rsign(dot(ray, fn));
                        }
                    }
                }
            } else {
                if (sp < 32u - 1u) { stack[sp++] = nd[3]; stack[sp++] = nd[7]; }
            }
        }
        const V3<R> cvel = {(R)U.coll_vel[0], (R)U.coll_vel[1], (R)U.coll_vel[2]};
        R vsf = (R)P.collider_velocity_factor / dt();
        if (has_hit) {
            pt.pos = hit_p + hit_n * margin;
            V3<R> ov = pt.dx / dt();
            V3<R> vr = ov - cvel;
            R vn = dot(vr, hit_n);
            V3<R> vnv = vn * hit_n, vt = vr - vnv;
            V3<R> nv;
            if (vn < (R)0) nv = vt * ((R)1 - (R)P.friction_ccd) + cvel * vsf;
            else nv = vr + cvel * vsf;
            pt.vel = nv;
            pt.dx = nv * dt();
            return;
        }
        R dist = r_sqrt(min_d2);
        V3<R> df = cur - cl_p;
        V3<R> pd = cl_n;
        R ld = length(df);
        if (ld > (R)1e-6f) pd = df / ld;
        if (dist < margin) {
            V3<R> ov = pt.dx / dt();
            V3<R> vr = ov - cvel;
            R vn = dot(vr, pd);
            V3<R> vnv = vn * pd, vt = vr - vnv;
            V3<R> nv;
            if (vn < (R)0) nv = vt * ((R)1 - (R)P.friction_proximity) + cvel * vsf;
            else nv = vr + cvel * vsf;
            pt.vel = nv;
            pt.dx = nv * dt();
            if (dist < margin) pt.pos = cl_p + pd * margin;
        }
    }

    // ------------------------------------------------ integrateParticles.wgsl:12-72
    void integrate(Particle<R>& pt) {
        pt.pos = pt.pos + pt.dx;
        pt.Fe = mmul(madd(ident<R>(), pt.D), pt.Fe);
        plasticity(pt);
        V3<R> mn = gmin(), mx = gmax();
        R* p = &pt.pos.x; R* d = &pt.dx.x; R* v = &pt.vel.x;
        const R* lo = &mn.x; const R* hi = &mx.x;
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a]) { d[a] = d[a] * (R)-0.5; p[a] = lo[a]; v[a] = v[a] * (R)-0.5; }
            if (p[a] >= hi[a]) { d[a] = d[a] * (R)-0.5; p[a] = hi[a]; v[a] = v[a] * (R)-0.5; }
        }
        resolve_collision(pt);
    }

    // ------------------------------------------------ GpuMpmPipelineManager.ts:327-428
    // stage: 0 binning, 1 +P2G, 2 +grid update, >=3 full substep
    void integrate_all() {
        const size_t N = parts.size();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads) if (par())
        for (size_t p = 0; p < N; ++p) integrate(parts[p]);
    }
    void explicit_substep(int stage) {
        bin();
        if (stage < 1) return;
        clear_grid();
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads) if (par())
        for (uint32_t i = 0; i < n_sorted; ++i) p2g_explicit(parts[sorted[i]]);
        if (stage < 2) return;
        grid_update_explicit();
        if (stage < 3) return;
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads) if (par())
        for (uint32_t i = 0; i < n_sorted; ++i) g2p_explicit(parts[sorted[i]]);
        integrate_all();
    }
    // ------------------------------------------------ GpuMpmPipelineManager.ts:431-521
    // stage: 0 binning, s in 1..k: grid after the pbmpm dispatch of iteration s-1, > k full
    std::vector<float>* trace = nullptr;   // when set: dx of every particle (by id) after each G2P
    void pbmpm_substep(int stage) {
        bin();
        const int k = P.pbmpm_iterations;
        for (int it = 0; it < k; ++it) {
            if (stage < it + 1) return;
            clear_grid();
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads) if (par())
            for (uint32_t i = 0; i < n_sorted; ++i) {
                Particle<R>& pt = parts[sorted[i]];
                solve_constraints(pt);
                p2g_pbmpm(pt);
            }
            if (stage < it + 2 && stage <= k) return;
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads) if (par())
            for (uint32_t i = 0; i < n_sorted; ++i) g2p_pbmpm(parts[sorted[i]]);
            if (trace) {
                const size_t N = parts.size();
                for (size_t p = 0; p < N; ++p) {
                    float* o = trace->data() + ((size_t)it * N + p) * 12;
                    o[0] = (float)parts[p].dx.x; o[1] = (float)parts[p].dx.y; o[2] = (float)parts[p].dx.z;
                    for (int c = 0; c < 3; ++c) { o[3 + 3 * c] = (float)parts[p].D.c[c].x; o[4 + 3 * c] = (float)parts[p].D.c[c].y; o[5 + 3 * c] = (float)parts[p].D.c[c].z; }
                }
            }
        }
        integrate_all();
    }
};

template <class R> static void load_aos(std::vector<Particle<R>>& ps, const float* a, uint32_t n) {
    ps.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = a + (size_t)i * 48;
        Particle<R>& p = ps[i];
        p.pos = {(R)q[0], (R)q[1], (R)q[2]};
        p.vel = {(R)q[4], (R)q[5], (R)q[6]};
        p.mass = (R)q[7];
        for (int c = 0; c < 3; ++c) {
            p.Fe.c[c] = {(R)q[8 + 4 * c], (R)q[9 + 4 * c], (R)q[10 + 4 * c]};
            p.Fp.c[c] = {(R)q[20 + 4 * c], (R)q[21 + 4 * c], (R)q[22 + 4 * c]};
            p.D.c[c] = {(R)q[36 + 4 * c], (R)q[37 + 4 * c], (R)q[38 + 4 * c]};
        }
        p.dx = {(R)q[32], (R)q[33], (R)q[34]};
    }
}
// 37 values per particle: pos3 vel3 mass Fe9 Fp9 dx3 D9 (column-major)
template <class R> static void store_flat(const std::vector<Particle<R>>& ps, double* o) {
    for (size_t i = 0; i < ps.size(); ++i) {
        const Particle<R>& p = ps[i];
        double* q = o + i * 37;
        q[0] = p.pos.x; q[1] = p.pos.y; q[2] = p.pos.z;
        q[3] = p.vel.x; q[4] = p.vel.y; q[5] = p.vel.z; q[6] = p.mass;
        for (int c = 0; c < 3; ++c) {
            q[7 + 3 * c] = p.Fe.c[c].x; q[8 + 3 * c] = p.Fe.c[c].y; q[9 + 3 * c] = p.Fe.c[c].z;
            q[16 + 3 * c] = p.Fp.c[c].x; q[17 + 3 * c] = p.Fp.c[c].y; q[18 + 3 * c] = p.Fp.c[c].z;
            q[28 + 3 * c] = p.D.c[c].x; q[29 + 3 * c] = p.D.c[c].y; q[30 + 3 * c] = p.D.c[c].z;
        }
        q[25] = p.dx.x; q[26] = p.dx.y; q[27] = p.dx.z;
    }
}
template <class R> static void store_aos(const std::vector<Particle<R>>& ps, float* o) {
    for (size_t i = 0; i < ps.size(); ++i) {
        const Particle<R>& p = ps[i];
        float* q = o + i * 48;
        std::memset(q, 0, 192);
        q[0] = (float)p.pos.x; q[1] = (float)p.pos.y; q[2] = (float)p.pos.z; q[3] = 1.0f;
        q[4] = (float)p.vel.x; q[5] = (float)p.vel.y; q[6] = (float)p.vel.z; q[7] = (float)p.mass;
        for (int c = 0; c < 3; ++c) {
            q[8 + 4 * c] = (float)p.Fe.c[c].x; q[9 + 4 * c] = (float)p.Fe.c[c].y; q[10 + 4 * c] = (float)p.Fe.c[c].z;
            q[20 + 4 * c] = (float)p.Fp.c[c].x; q[21 + 4 * c] = (float)p.Fp.c[c].y; q[22 + 4 * c] = (float)p.Fp.c[c].z;
            q[36 + 4 * c] = (float)p.D.c[c].x; q[37 + 4 * c] = (float)p.D.c[c].y; q[38 + 4 * c] = (float)p.D.c[c].z;
        }
        q[32] = (float)p.dx.x; q[33] = (float)p.dx.y; q[34] = (float)p.dx.z;
    }
}

typedef Sim<float, AccI32> Sim32;
typedef Sim<double, AccI64> Sim64q;
typedef Sim<double, AccF64> Sim64;

struct Handle {
    int mode;   // 0 FP32-faithful, 1 FP64 with the fixed-point grid, 2 FP64 without quantization
    Sim32* s32;
    Sim64q* s64q;
    Sim64* s64;
    template <class F> auto visit(F f) { return s32 ? f(s32) : (s64q ? f(s64q) : f(s64)); }
};

}  // namespace

extern "C" {

void orc_default_params(orc_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->pbmpm_iterations = 3;
    p->pbmpm_force_divisor = 3.0f;
    p->gravity[0] = 0.0f; p->gravity[1] = 0.0f; p->gravity[2] = -9.81f;
    p->youngs_modulus = 1.4e5;
    p->poissons_ratio = 0.2;
    p->hardening_explicit = 5.0f;
    p->hardening_pbmpm = 20.0f;
    p->pbmpm_blend = 0.95f;
    p->pbmpm_volume_exponent = -0.3333333f;
    p->theta_c = 2.5e-2;
    p->theta_s = 7.5e-3;
    p->density = 400.0;
    p->grad_scale = 1.0f;
    p->collider_margin = 0.05f;
    p->friction_ccd = 0.1f;
    p->friction_proximity = 0.0f;
    p->collider_velocity_factor = 0.2f;
    p->wall_mode = 0;
    p->wall_cells = 3;
    p->hash_map_size = 200003u;
    p->max_blocks = 100000u;
    p->max_probes = 100u;
}

size_t orc_params_size(void) { return sizeof(orc_params); }

void* orc_create(int mode, const orc_params* prm, const uint8_t* uniforms, uint32_t n, const float* aos192) {
    Handle* h = new Handle{mode, nullptr, nullptr, nullptr};
    if (mode == 1) h->s64q = new Sim64q(); else if (mode == 2) h->s64 = new Sim64(); else h->s32 = new Sim32();
    h->visit([&](auto* s) { s->P = *prm; s->U = parse_uniforms(uniforms); s->resize_grid(); load_aos(s->parts, aos192, n); return 0; });
    return h;
}
void orc_destroy(void* hp) { Handle* h = (Handle*)hp; delete h->s32; delete h->s64q; delete h->s64; delete h; }

void orc_set_uniforms(void* hp, const uint8_t* u) {
    ((Handle*)hp)->visit([&](auto* s) { s->U = parse_uniforms(u); return 0; });
}
void orc_set_collider(void* hp, const uint32_t* data, size_t n_words, const uint32_t* bvh_words, size_t n_nodes) {
    std::vector<uint32_t> d(data, data + n_words), b(bvh_words, bvh_words + n_nodes * 8);
    ((Handle*)hp)->visit([&](auto* s) { s->coll_data = d; s->bvh = b; return 0; });
}
// method: 0 explicit, 1 PBMPM
void orc_step(void* hp, int method, int n_substeps) {
    ((Handle*)hp)->visit([&](auto* s) {
        for (int i = 0; i < n_substeps; ++i) { if (method) s->pbmpm_substep(1 << 30); else s->explicit_substep(1 << 30); }
        return 0;
    });
}
void orc_run_stage(void* hp, int method, int stage) {
    ((Handle*)hp)->visit([&](auto* s) { if (method) s->pbmpm_substep(stage); else s->explicit_substep(stage); return 0; });
}
void orc_get_particles(void* hp, float* aos192) { ((Handle*)hp)->visit([&](auto* s) { store_aos(s->parts, aos192); return 0; }); }
void orc_get_particles_f64(void* hp, double* out37) { ((Handle*)hp)->visit([&](auto* s) { store_flat(s->parts, out37); return 0; }); }
void orc_set_particles(void* hp, const float* aos192, uint32_t n) { ((Handle*)hp)->visit([&](auto* s) { load_aos(s->parts, aos192, n); return 0; }); }
// Per-particle Young's modulus (explicit stress): Lame parameters in double, rounded to FP32,
// the same forms as the scene-wide constants. E == NULL clears them.
void orc_set_particle_youngs(void* hp, const float* E, uint32_t n) {
    ((Handle*)hp)->visit([&](auto* s) {
        const double nu = s->P.poissons_ratio;
        for (uint32_t i = 0; i < n && i < s->parts.size(); ++i) {
            auto& p = s->parts[i];
            p.own_lame = E != nullptr;
            if (E) {
                p.mu0 = (decltype(p.mu0))(float)((double)E[i] / (2.0 * (1.0 + nu)));
                p.la0 = (decltype(p.la0))(float)((double)E[i] * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
            }
        }
        return 0;
    });
}
// Cell-keyed dump of every allocated cell: xyz[3*i], vals[4*i] = (mass, mx, my, mz) as raw
// accumulator values (fixed-point integers, or unquantized sums in mode 2).
size_t orc_dump_grid(void* hp, int32_t* xyz, double* vals, size_t cap) {
    return ((Handle*)hp)->visit([&](auto* s) -> size_t {
        size_t n = 0;
        for (uint32_t b = 0; b < s->n_alloc; ++b) {
            uint32_t slot = s->mapped[b];
            for (uint32_t l = 0; l < 64; ++l) {
                if (n < cap) {
                    xyz[3 * n] = s->hm_bx[slot] * 4 + (int32_t)(l & 3);
                    xyz[3 * n + 1] = s->hm_by[slot] * 4 + (int32_t)((l >> 2) & 3);
                    xyz[3 * n + 2] = s->hm_bz[slot] * 4 + (int32_t)(l >> 4);
                    size_t ci = (size_t)b * 64 + l;
                    vals[4 * n] = (double)s->gm[ci]; vals[4 * n + 1] = (double)s->gx[ci];
                    vals[4 * n + 2] = (double)s->gy[ci]; vals[4 * n + 3] = (double)s->gz[ci];
                }
                ++n;
            }
        }
        return n;
    });
}
// Allocated block coordinates and the per-block binned-particle histogram.
size_t orc_active_blocks(void* hp, int32_t* xyz, uint32_t* counts, size_t cap) {
    return ((Handle*)hp)->visit([&](auto* s) -> size_t {
        for (uint32_t b = 0; b < s->n_alloc && b < cap; ++b) {
            uint32_t slot = s->mapped[b];
            xyz[3 * b] = s->hm_bx[slot]; xyz[3 * b + 1] = s->hm_by[slot]; xyz[3 * b + 2] = s->hm_bz[slot];
            counts[b] = s->counts[b];
        }
        return (size_t)s->n_alloc;
    });
}
void orc_get_stats(void* hp, orc_stats* out) { ((Handle*)hp)->visit([&](auto* s) { *out = s->st; return 0; }); }

// ---------------------------------------------------------------- unit-level entry points
uint32_t orc_hash1(uint32_t n) { return hash1(n); }
uint32_t orc_hash3(uint32_t x, uint32_t y, uint32_t z) { return hash3(x, y, z); }
uint32_t orc_hash4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return hash4(x, y, z, w); }
float orc_exp(float x) { return orc_expf(x); }
float orc_log(float x) { return orc_logf(x); }
float orc_pow(float x, float y) { return orc_powf(x, y); }
int32_t orc_f2i(float x) { return f2i(x); }

static M3<float> ld3(const float* m) { M3<float> r; for (int c = 0; c < 3; ++c) r.c[c] = {m[3 * c], m[3 * c + 1], m[3 * c + 2]}; return r; }
static void st3(const M3<float>& r, float* m) { for (int c = 0; c < 3; ++c) { m[3 * c] = r.c[c].x; m[3 * c + 1] = r.c[c].y; m[3 * c + 2] = r.c[c].z; } }
float orc_det3(const float* m) { return det3(ld3(m)); }
void orc_inverse3(const float* m, float* out) { st3(inverse3(ld3(m)), out); }
void orc_polar3(const float* m, float* out) { st3(polar_rotation(ld3(m)), out); }
void orc_mmul3(const float* a, const float* b, float* out) { st3(mmul(ld3(a), ld3(b)), out); }
// B-spline weights/derivatives of one fractional coordinate triple: w[9] = w0xyz w1xyz w2xyz
void orc_bspline(const float* f, float* w, float* d) {
    V3<float> ww[3], dd[3];
    V3<float> ff = {f[0], f[1], f[2]};
    Sim32::weights(ff, ww); Sim32::weight_derivs(ff, dd);
    for (int i = 0; i < 3; ++i) { w[3 * i] = ww[i].x; w[3 * i + 1] = ww[i].y; w[3 * i + 2] = ww[i].z; d[3 * i] = dd[i].x; d[3 * i + 1] = dd[i].y; d[3 * i + 2] = dd[i].z; }
}
void orc_closest_point_triangle(const float* p, const float* a, const float* b, const float* c, float* out) {
    V3<float> r = Sim32::closest_point_triangle({p[0], p[1], p[2]}, {a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float orc_intersect_ray_triangle(const float* o, const float* d, const float* a, const float* b, const float* c) {
    return Sim32::intersect_ray_triangle({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, {a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]});
}

// One particle through the PBMPM constraint solve (pbmpm.wgsl:16-39), FP32-faithful: D9 is
// updated in place. Matrices column-major.
void orc_solve_constraints1(const orc_params* prm, const float* Fe9, const float* Fp9, float* D9) {
    Sim32 s; s.P = *prm;
    Particle<float> pt{};
    pt.Fe = ld3(Fe9); pt.Fp = ld3(Fp9); pt.D = ld3(D9);
    s.solve_constraints(pt);
    st3(pt.D, D9);
}
// The explicit method's first Piola-Kirchhoff stress of one particle (stressTensorOps.wgsl:9-22
// with the hardening of :39-56), FP32-faithful.
void orc_pk1_explicit1(const orc_params* prm, const float* Fe9, const float* Fp9, float* P9) {
    Sim32 s; s.P = *prm;
    Particle<float> pt{};
    pt.Fe = ld3(Fe9); pt.Fp = ld3(Fp9);
    st3(s.pk1_explicit(pt), P9);
}
// The SVD of the plasticity option for one FP32 matrix (column-major): sig[3], V (9, column-major),
// B = A V (9); use_fp64 decomposes the matrix in double (results rounded to float).
void orc_svd3(const float* A9, int use_fp64, float* sig, float* V9, float* B9) {
    if (use_fp64) {
        M3<double> A;
        for (int c = 0; c < 3; ++c) A.c[c] = {(double)A9[3 * c], (double)A9[3 * c + 1], (double)A9[3 * c + 2]};
        const Svd3<double> d = svd3(A);
        for (int c = 0; c < 3; ++c) {
            sig[c] = (float)d.sig[c];
            V9[3 * c] = (float)d.V.c[c].x; V9[3 * c + 1] = (float)d.V.c[c].y; V9[3 * c + 2] = (float)d.V.c[c].z;
            B9[3 * c] = (float)d.B.c[c].x; B9[3 * c + 1] = (float)d.B.c[c].y; B9[3 * c + 2] = (float)d.B.c[c].z;
        }
        return;
    }
    const Svd3<float> d = svd3(ld3(A9));
    for (int c = 0; c < 3; ++c) sig[c] = d.sig[c];
    st3(d.V, V9); st3(d.B, B9);
}
// The SVD plasticity update of one particle (mode 1 or 2 of plasticity_mode): returns 1 when F_e
// changed; *fp64 = whether the FP64 path was taken.
int orc_plasticity_svd1(const float* Fe9, const float* Fp9, float lo, float hi, int fallback, float* Fe_out, float* Fp_out, int* fp64) {
    M3<float> Fe = ld3(Fe9), Fp = ld3(Fp9);
    bool f = false;
    const bool ch = plasticity_svd(Fe, Fp, lo, hi, fallback != 0, &f);
    st3(Fe, Fe_out); st3(Fp, Fp_out);
    *fp64 = f ? 1 : 0;
    return ch ? 1 : 0;
}
// integrateParticles (integrateParticles.wgsl:12-72) on every particle, with no transfer before.
void orc_integrate_only(void* hp) {
    ((Handle*)hp)->visit([&](auto* s) { s->integrate_all(); return 0; });
}
// One full PBMPM substep recording every particle's dx and D (by particle id) after the G2P of each
// of the P.pbmpm_iterations iterations: out[it][n][12] = dx (3), D (9, column-major); unbinned
// particles keep theirs.
void orc_pbmpm_trace(void* hp, float* out) {
    ((Handle*)hp)->visit([&](auto* s) {
        std::vector<float> tr((size_t)s->P.pbmpm_iterations * s->parts.size() * 12, 0.0f);
        s->trace = &tr;
        s->pbmpm_substep(1 << 30);
        s->trace = nullptr;
        std::memcpy(out, tr.data(), tr.size() * sizeof(float));
        return 0;
    });
}
// OpenMP threads of the passes (FP32-faithful mode only; the FP64 modes stay serial). Returns the
// count in effect.
int orc_set_threads(void* hp, int n) {
    Handle* h = (Handle*)hp;
    if (!h->s32) return 1;
    h->s32->threads = n > 1 ? n : 1;
    return h->s32->threads;
}
// Binning, then every allocated cell set to mass m and momentum m (a + M c) (c = the cell's
// integer coordinates, M column-major), then the G2P of every binned particle (method 0
// explicit, gridToParticle.cs.wgsl:28-80; 1 PBMPM, :82-130).
void orc_g2p_affine(void* hp, int method, const float* a, const float* M9, float m) {
    ((Handle*)hp)->visit([&](auto* s) {
        typedef typename std::remove_reference<decltype(s->parts[0].pos.x)>::type R;
        s->bin();
        M3<R> M;
        for (int c = 0; c < 3; ++c) M.c[c] = {(R)M9[3 * c], (R)M9[3 * c + 1], (R)M9[3 * c + 2]};
        const V3<R> av = {(R)a[0], (R)a[1], (R)a[2]};
        const float S = s->U.S;
        for (uint32_t b = 0; b < s->n_alloc; ++b) {
            uint32_t slot = s->mapped[b];
            for (uint32_t l = 0; l < 64; ++l) {
                V3<R> c = {(R)(s->hm_bx[slot] * 4 + (int32_t)(l & 3)), (R)(s->hm_by[slot] * 4 + (int32_t)((l >> 2) & 3)),
                           (R)(s->hm_bz[slot] * 4 + (int32_t)(l >> 4))};
                V3<R> v = av + mvmul(M, c);
                size_t ci = (size_t)b * 64 + l;
                s->gm[ci] = s->conv_acc((R)m, S);
                s->gx[ci] = s->conv_acc((R)m * v.x, S);
                s->gy[ci] = s->conv_acc((R)m * v.y, S);
                s->gz[ci] = s->conv_acc((R)m * v.z, S);
            }
        }
        for (uint32_t i = 0; i < s->n_sorted; ++i) {
            if (method) s->g2p_pbmpm(s->parts[s->sorted[i]]); else s->g2p_explicit(s->parts[s->sorted[i]]);
        }
        return 0;
    });
}

}  // extern "C"
