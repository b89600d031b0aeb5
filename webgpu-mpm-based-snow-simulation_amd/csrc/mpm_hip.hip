// libmpm_hip — C ABI and host orchestration of the MPM substep on one MI355X.
//
// Host-side counterpart of src/lib/gpu/mpm/GpuMpmBufferManager.ts (buffer sizing),
// GpuMpmPipelineManager.ts:327-521 (per-substep dispatch schedule),
// src/lib/gpu/uniforms/GpuUniformsBufferManager.ts (uniform blob) and
// src/lib/gpu/collider/GpuColliderBufferManager.ts:79-350 (collider packing, BVH build).
// One substep is a fixed sequence of kernels on the context's stream; with use_graph it is
// captured once per (method, buffer parity, parameters) into a hipGraph and replayed, and runs of
// GRAPH_SUBSTEPS substeps are replayed from one graph holding that many (a graph launch costs ~8 us
// between replays; inside a graph the kernels follow each other without gaps).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <cmath>
#include <string>
#include <vector>
#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include "mpm_kernels.h"
#include "mpm_mlp.h"
#include "mpm_ensemble.h"
#include "mpm_scatter.h"
#include "mpm_splat.h"
#include "mpm_gltf.h"
#include "mpm_bvh.h"
#include "../../include/mpm_hip.h"

using namespace mpm;

// Accumulation grid of mpm_splat_noalloc: a file-scope device array, so that entry point calls no
// allocator. Every GPU the library runs on holds its own copy until the process ends. In host
// code `scratch` names a host-side stand-in, never the device memory: host code uses the
// per-device address from hipGetSymbolAddress (scratch_ptr below) for kernel arguments and copies.
static __device__ __align__(256) uint32_t scratch[MPM_SPLAT_NOALLOC_MAX_CELLS];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
hipError_t g_scratch_err[SCRATCH_MAX_DEVICES] = {};
bool g_scratch_done[SCRATCH_MAX_DEVICES] = {};

// Once per device, with that device current (context set-up): the device address of `scratch`.
void scratch_resolve(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* p = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_scratch_err[device] == hipSuccess) g_scratch_ptr[device] = (uint32_t*)p;
    else (void)hipGetLastError();
    g_scratch_done[device] = true;
}
uint32_t* scratch_ptr(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}
}  // namespace

struct mpm_ctx {
    mpm_config cfg;
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    uint8_t uni[MPM_UNIFORMS_BYTES];
    std::string err;
    uint32_t n = 0, stride = 0, K = 0, nb[3] = {0, 0, 0}, cap = 0, max_blocks = 0, max_runs = 0, max_chunks = 0, ntiles = 0, scan_kpt = 1;
    uint32_t chunk = 512;
    float* planes[2] = {nullptr, nullptr};
    uint32_t* pid[2] = {nullptr, nullptr};
    float4* fpid = nullptr;                   // F_p by particle id (3 columns each)
    int cur = 0;
    float* relax = nullptr;
    uint32_t *key = nullptr, *mask = nullptr, *perm = nullptr, *kcount = nullptr, *kmask = nullptr, *kcursor = nullptr;
    uint32_t* lofs = nullptr;
    uint32_t *tile_sums = nullptr, *tile_ofs = nullptr;
    uint32_t *kalloc = nullptr, *kslot = nullptr;
    bool direct_blocks = false;       // block set by direct key addressing (hash_map_size == 0)
    uint32_t *run_key = nullptr, *run_start = nullptr, *run_count = nullptr, *run_mask = nullptr;
    uint2* chunks = nullptr;
    uint32_t *table = nullptr, *slot_val = nullptr, *nbr = nullptr, *block_key = nullptr;
    int4* grid[3] = {nullptr, nullptr, nullptr};
    int* gmass = nullptr;
    bool last_pbmpm = false;
    // fused keys pass (SimParams::fuse_keys): the buffer parity whose rows the stored keys and key-space
    // counts describe (-1: none), and the grid they were made for
    int keys_parity = -1;
    float keys_grid[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t* counters = nullptr;
    uint32_t* cqueue = nullptr;
    uint32_t* coll_data = nullptr;
    float4* crec = nullptr;           // threaded BVH records (2 float4 each)
    float4* ctri = nullptr;           // world-space triangles (CT_WORDS float4 each)
    size_t n_coll_words = 0, n_coll_rec = 0, n_coll_tri = 0;
    uint32_t coll_num_indices = 0;
    int last_grid = 0;
    bool has_particles = false;
    // [method][buffer parity][0: one substep, 1: GRAPH_SUBSTEPS substeps] (batched scenes: [0][parity][0])
    hipGraphExec_t gexec[2][2][2] = {};
    hipGraph_t graph[2][2][2] = {};
    SimParams gparams[2][2][2];
    Bufs gbufs[2][2][2];              // device addresses a captured graph's kernels were given
    MlpBufs gmlp[2][2][2];
    bool timing = false;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};   // phases of a substep; [4], [5] the whole call
    mpm_timings tm{};
    uint64_t substeps = 0;
    std::vector<void*> allocs;
    int k_cur = 3;                 // PBMPM iterations of the following substeps
    float force_div_cur = 3.0f;
    // learned warm start: MLP parameters, Adam state, standardisation, logged dataset
    float *mlp_params = nullptr, *mlp_grads = nullptr, *mlp_m1 = nullptr, *mlp_m2 = nullptr, *mlp_norm = nullptr;
    short* mlp_wbf = nullptr;
    float* mlp_colsum = nullptr;              // per-workgroup partial column moments (1024 x 64)
    float* mlp_loss = nullptr;
    float *ds_feats = nullptr, *ds_targs = nullptr;
    uint64_t ds_cap = 0, ds_n = 0;
    bool mlp_ready = false, learn_on = false;
    float learn_limit = 8.0f;
    uint64_t adam_t = 0;
    // per-iteration residual tracking
    float* resid_dx = nullptr;
    uint32_t *resid_max = nullptr, *resid_peak = nullptr;
    float* resid_sum = nullptr;
    bool resid_on = false;
    float* learn_dprev = nullptr;            // learned warm start input: last per-iteration change of dx, by row
    bool collecting = false;                 // inside mpm_learn_collect
    uint64_t resid_substeps = 0;
    std::map<uint64_t, bool> div_verified;   // (divisor, scale) bits -> fixed_div exact for all inputs
    std::map<uint64_t, bool> quot_verified;  // divisor bits -> quot_i32 exact for all |x| <= 2^100
    std::map<uint64_t, bool> qdiv_verified;  // divisor bits -> qdiv bit-exact for all |x| <= 2^100
    bool per_particle_lame = false;
    uint32_t* d_scratch = nullptr;           // one counter for k_verify_div
    // per-kernel event timing (eager launches, one event after every kernel)
    bool ktiming = false;
    std::vector<hipEvent_t> kt_ev;
    std::vector<int> kt_id;
    size_t kt_n = 0;
    double kt_ms[MPM_KERNEL_COUNT] = {};
    uint64_t kt_cnt[MPM_KERNEL_COUNT] = {};
    // batched scenes (ens > 0: this context advances ens independent scenes together)
    uint32_t ens = 0, ens_R = 0;
    std::vector<mpm_config> ens_cfg;
    std::vector<int> ens_method;
    std::vector<uint32_t> ens_part_base, ens_key_base, ens_blk_base, ens_nphase;
    std::vector<SimParams> ens_sp;
    std::vector<bool> ens_loaded;
    SimParams ens_P0;
    SimParams* d_sp = nullptr;
    uint32_t *d_part_base = nullptr, *d_key_base = nullptr, *d_nphase = nullptr, *d_method = nullptr, *d_nblk = nullptr;
    uint32_t* d_rbase = nullptr;      // batched scenes, direct addressing: first block-list entry per scene
    uint32_t *d_cbeg = nullptr, *d_seg = nullptr, *d_done = nullptr, *d_ctl = nullptr;
    int ens_persistent = 0;
    uint32_t ens_wg_per_cu = 4;
    bool ens_graph_ok[2] = {false, false};
    float* snap_planes = nullptr;
    uint32_t* snap_pid = nullptr;
    float4* snap_fpid = nullptr;
    // render-side splat grid
    uint32_t* splat = nullptr;
    size_t splat_cells = 0;
};

namespace {

const char* k_version = "libmpm_hip 0.1 (gfx950)";
// substeps per graph for runs of substeps (even: the buffer parity after the graph is the one before)
enum { GRAPH_SUBSTEPS = 8 };
const char* k_kernel_names[MPM_KERNEL_COUNT] = {
    "k_keys_count", "k_scan_tiles", "k_scan_top", "k_scan_apply", "k_scatter", "k_insert", "k_compact", "k_nbr",
    "k_pbmpm<FIRST>", "k_pbmpm<MID>", "k_pbmpm<LAST>", "k_explicit_p2g", "k_grid_update", "k_explicit_g2p", "k_straggle",
    "k_collide", "k_mlp_warmstart", "k_ens_keys_count", "k_ens_insert", "k_ens_compact", "k_ens_nbr", "k_ens_tasks",
    "k_ens_transfer", "k_ens_straggle", "k_plastic_svd"};
enum { KID_KEYS = 0, KID_SCAN_TILES, KID_SCAN_TOP, KID_SCAN_APPLY, KID_SCATTER, KID_INSERT, KID_COMPACT, KID_NBR,
       KID_PB_FIRST, KID_PB_MID, KID_PB_LAST, KID_EX_P2G, KID_EX_GRID, KID_EX_G2P, KID_STRAGGLE, KID_COLLIDE, KID_MLP,
       KID_ENS_KEYS, KID_ENS_INSERT, KID_ENS_COMPACT, KID_ENS_NBR, KID_ENS_TASKS, KID_ENS_TRANSFER, KID_ENS_STRAGGLE,
       KID_PLASTIC_SVD };

// Records an event after a kernel launch when per-kernel timing is on (the first event of a
// step is the start mark, id -1).
void kt(mpm_ctx* c, int id) {
    if (!c->ktiming) return;
    if (c->kt_n == c->kt_ev.size()) { hipEvent_t e; hipEventCreate(&e); c->kt_ev.push_back(e); c->kt_id.push_back(-1); }
    c->kt_id[c->kt_n] = id;
    hipEventRecord(c->kt_ev[c->kt_n], c->stream);
    c->kt_n++;
}
// Scan tiling over K keys: one key per thread while that keeps the tile count within SCAN_OWN_MAX,
// four per thread beyond (fewer tiles for k_scan_top to combine).
void set_scan_tiles(mpm_ctx* c, uint64_t K) {
    c->scan_kpt = (K + 255u) / 256u <= SCAN_OWN_MAX ? 1u : 4u;
    c->ntiles = (uint32_t)((K + 256u * c->scan_kpt - 1) / (256u * c->scan_kpt));
}
// k_scan_tiles, k_scan_top when there are many tiles, k_scan_apply.
void enqueue_scan(mpm_ctx* c, const SimParams& P, const Bufs& B, hipStream_t s) {
    const int own_ofs = c->ntiles <= SCAN_OWN_MAX ? 1 : 0;
    if (c->scan_kpt == 1) k_scan_tiles<1><<<c->ntiles, 256, 0, s>>>(P, B);
    else k_scan_tiles<4><<<c->ntiles, 256, 0, s>>>(P, B);
    kt(c, KID_SCAN_TILES);
    if (!own_ofs) { k_scan_top<<<1, 256, 0, s>>>(P, B, c->ntiles); kt(c, KID_SCAN_TOP); }
    if (c->scan_kpt == 1) k_scan_apply<1><<<c->ntiles, 256, 0, s>>>(P, B, c->ntiles, own_ofs);
    else k_scan_apply<4><<<c->ntiles, 256, 0, s>>>(P, B, c->ntiles, own_ofs);
    kt(c, KID_SCAN_APPLY);
}
void kt_collect(mpm_ctx* c) {
    if (!c->ktiming || c->kt_n == 0) return;
    hipEventSynchronize(c->kt_ev[c->kt_n - 1]);
    for (size_t i = 1; i < c->kt_n; ++i) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, c->kt_ev[i - 1], c->kt_ev[i]);
        const int id = c->kt_id[i];
        if (id >= 0) { c->kt_ms[id] += ms; c->kt_cnt[id] += 1; }
    }
    c->kt_n = 0;
}

int fail(mpm_ctx* c, int code, const std::string& m) { if (c) c->err = m; return code; }
#define HIPCHK(c, x) do { hipError_t _e = (x); if (_e != hipSuccess) return fail((c), MPM_E_HIP, std::string(#x) + ": " + hipGetErrorString(_e)); } while (0)

// After a step with timing on: the whole call's device time between events recorded before its
// first and after its last substep (the reference's compute-pass timestamps, slots 0 and 1,
// GpuSnowPipelineRunner.svelte.ts:411-420,588-611), and the phases of its last substep, which
// was launched eagerly with phase events (binning, transfer passes, integrate + collider).
int finish_timing(mpm_ctx* c, int n_substeps) {
    HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[5]));
    float pass = 0, a = 0, b = 0, d = 0;
    HIPCHK(c, hipEventElapsedTime(&pass, c->ev[4], c->ev[5]));
    if (n_substeps > 0) {
        HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        HIPCHK(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        HIPCHK(c, hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
    }
    c->tm.compute_simulation_step_ns = (double)pass * 1e6;
    c->tm.binning_ns = (double)a * 1e6; c->tm.transfer_ns = (double)b * 1e6; c->tm.integrate_ns = (double)d * 1e6;
    c->tm.substeps = c->substeps;
    c->tm.pass_substeps = (uint64_t)n_substeps;
    return MPM_OK;
}

template <class T> int dalloc(mpm_ctx* c, T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (hipMalloc(&q, bytes) != hipSuccess) { (void)hipGetLastError(); return fail(c, MPM_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"); }
    c->allocs.push_back(q);
    *p = (T*)q;
    return MPM_OK;
}

uint32_t next_prime(uint32_t v) {
    if (v < 3) return 3;
    if (!(v & 1u)) ++v;
    for (;; v += 2) {
        bool prime = true;
        for (uint32_t d = 3; (uint64_t)d * d <= v; d += 2) if (v % d == 0) { prime = false; break; }
        if (prime) return v;
    }
}

template <class T> T rdu(const uint8_t* b, size_t off) { T v; std::memcpy(&v, b + off, sizeof(T)); return v; }
template <class T> void wru(uint8_t* b, size_t off, T v) { std::memcpy(b + off, &v, sizeof(T)); }

MlpBufs mlp_bufs(mpm_ctx* c) {
    MlpBufs M;
    std::memset(&M, 0, sizeof(M));   // padding too: graph keys compare the bytes
    M.params = c->mlp_params; M.wbf = c->mlp_wbf; M.grads = c->mlp_grads; M.norm = c->mlp_norm;
    return M;
}

Bufs make_bufs(mpm_ctx* c) {
    Bufs B;
    std::memset(&B, 0, sizeof(B));   // padding too: graph keys compare the bytes
    B.cur = c->planes[c->cur]; B.nxt = c->planes[c->cur ^ 1];
    B.pid_cur = c->pid[c->cur]; B.pid_nxt = c->pid[c->cur ^ 1]; B.fpid = c->fpid;
    B.stride = c->stride;
    B.relax = c->relax; B.key = c->key; B.mask = c->mask; B.perm = c->perm; B.lofs = c->lofs;
    B.kcount = c->kcount; B.kmask = c->kmask; B.kcursor = c->kcursor;
    B.tile_sums = c->tile_sums; B.tile_ofs = c->tile_ofs;
    B.kalloc = c->kalloc; B.kslot = c->kslot;
    B.run_key = c->run_key; B.run_start = c->run_start; B.run_count = c->run_count; B.run_mask = c->run_mask;
    B.chunks = c->chunks;
    B.table = c->table; B.slot_val = c->slot_val; B.nbr = c->nbr; B.block_key = c->block_key;
    B.grid0 = c->grid[0]; B.grid1 = c->grid[1]; B.grid2 = c->grid[2];
    B.gmass = c->gmass;
    B.counters = c->counters;
    B.cqueue = c->cqueue;
    B.coll_data = c->coll_data; B.crec = c->crec; B.ctri = c->ctri;
    // the per-iteration dx bookkeeping runs for residual tracking and for the learned warm start
    // (which reads the last iteration's change of dx, dprev, as an input of the next substep)
    const bool learn = (c->learn_on && c->mlp_ready) || c->collecting;
    B.resid_dx = (c->resid_on || learn) ? c->resid_dx : nullptr;
    B.dprev = learn ? c->learn_dprev : nullptr;
    B.resid_max = c->resid_max;
    B.resid_sum = c->resid_sum;
    return B;
}

// Checks on the device, once per (divisor, fixed-point scale) pair, that fixed_div's FMA form
// gives i32((x / d) * S) for every float x (mode 0), or once per divisor that quot_i32 gives
// i32(x / d) (mode 1) or qdiv gives x / d bit for bit (mode 2) for every float with |x| <= 2^100;
// the kernels use a form only where it passed.
int verify_divisor(mpm_ctx* c, float d, float S, int mode, uint32_t* ok) {
    if (mode != 0) S = 1.0f;
    uint32_t db, sb; std::memcpy(&db, &d, 4); std::memcpy(&sb, &S, 4);
    const uint64_t key = ((uint64_t)db << 32) | sb;
    std::map<uint64_t, bool>& seen = mode == 2 ? c->qdiv_verified : (mode == 1 ? c->quot_verified : c->div_verified);
    auto it = seen.find(key);
    if (it != seen.end()) { *ok = it->second ? 1u : 0u; return MPM_OK; }
    *ok = 0;
    if (!(d > 0.0f) || d > 1e30f || d < 1e-30f || !(S > 0.0f) || S > 1e30f) { seen[key] = false; return MPM_OK; }
    HIPCHK(c, hipMemsetAsync(c->d_scratch, 0, 4, c->stream));
    k_verify_fixed_div<<<8192, 256, 0, c->stream>>>(d, 1.0f / d, S, mode, c->d_scratch);
    uint32_t bad = 1;
    HIPCHK(c, hipMemcpyAsync(&bad, c->d_scratch, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    seen[key] = bad == 0;
    *ok = bad == 0 ? 1u : 0u;
    return MPM_OK;
}

// Parse a uniform blob (GpuUniformsBufferManager.ts:54-182 offsets) + creation config.
int params_from(mpm_ctx* c, const uint8_t* u, const mpm_config& g, int k_iter, float force_div, bool scene_only, SimParams& P) {
    std::memset(&P, 0, sizeof(P));
    P.dt = rdu<float>(u, 0);
    P.S = rdu<float>(u, 4);
    for (int a = 0; a < 3; ++a) {
        P.gmin[a] = rdu<float>(u, 16 + 4 * a);
        P.gmax[a] = rdu<float>(u, 32 + 4 * a);
        P.res[a] = (int32_t)rdu<uint32_t>(u, 208 + 4 * a);
        P.cd[a] = rdu<float>(u, 352 + 4 * a);
        P.cmin[a] = rdu<float>(u, 224 + 4 * a);
        P.cmax[a] = rdu<float>(u, 240 + 4 * a);
        P.cvel[a] = rdu<float>(u, 320 + 4 * a);
        P.ip[a] = rdu<float>(u, 496 + 4 * a);
        P.grav[a] = g.gravity[a];
        if (P.res[a] <= 0) return fail(c, MPM_E_RANGE, "grid resolution must be positive");
        P.nb[a] = (P.res[a] + 3) / 4;
    }
    uint64_t K = (uint64_t)P.nb[0] * P.nb[1] * P.nb[2];
    if (!scene_only && K > c->K) return fail(c, MPM_E_RANGE, "uniform grid resolution exceeds the resolution the context was created for");
    P.K = (uint32_t)K;
    for (int i = 0; i < 16; ++i) { P.T[i] = rdu<float>(u, 256 + 4 * i); P.Tinv[i] = rdu<float>(u, 432 + 4 * i); }
    P.coll_num_indices = rdu<uint32_t>(u, 348);
    P.coll_num_objects = rdu<uint32_t>(u, 364);
    P.istrength = rdu<float>(u, 508);
    P.iradius = rdu<float>(u, 512);
    P.interacting = rdu<uint32_t>(u, 516);
    P.imode = rdu<uint32_t>(u, 520);
    P.n = c->n;
    P.chunk = c->chunk;
    P.cap = c->cap;
    P.max_blocks = c->max_blocks;
    P.max_probes = g.max_probes ? g.max_probes : 100u;
    P.mu0 = (float)(g.youngs_modulus / (2.0 * (1.0 + g.poissons_ratio)));
    P.la0 = (float)(g.youngs_modulus * g.poissons_ratio / ((1.0 + g.poissons_ratio) * (1.0 - 2.0 * g.poissons_ratio)));
    P.xi_e = g.hardening_explicit;
    P.xi_p = g.hardening_pbmpm;
    P.blend = g.pbmpm_blend;
    P.one_minus_blend = 1.0f - g.pbmpm_blend;
    P.vol_exp = g.pbmpm_volume_exponent;
    P.plast_lo = (float)(1.0 - g.theta_c);
    P.plast_hi = (float)(1.0 + g.theta_s);
    P.plast_mode = g.plasticity_mode;
    {
        // squared inner band of band_clear, rounded toward its inside
        const double blo = (double)P.plast_lo + BAND_MARGIN, bhi = (double)P.plast_hi - BAND_MARGIN;
        float lo2 = (float)(blo * blo), hi2 = (float)(bhi * bhi);
        if ((double)lo2 < blo * blo) lo2 = std::nextafter(lo2, 2.0f);
        if ((double)hi2 > bhi * bhi) hi2 = std::nextafter(hi2, 0.0f);
        P.band_lo2 = lo2; P.band_hi2 = hi2;
        P.band_skip = (blo >= 0.8 && bhi <= 1.25 && blo < bhi) ? 1u : 0u;
    }
    P.inv_density = (float)(1.0 / g.density);
    P.grad_scale = g.grad_scale;
    P.force_div = force_div;
    P.k = k_iter;
    P.learn_flags = scene_only ? 0u : ((c->learn_on && c->mlp_ready ? 1u : 0u) | (c->resid_on ? 2u : 0u));
    P.learn_limit = c->learn_limit;
    {
        float ext = 0.0f;
        for (int a = 0; a < 3; ++a) ext = std::max(ext, P.gmax[a] - P.gmin[a]);
        P.inv_domain = 1.0f / ext;
    }
    for (int a = 0; a < 3; ++a) P.grav_div[a] = P.grav[a] / P.force_div;
    for (int a = 0; a < 3; ++a) P.grav_nz[a] = P.grav_div[a] != 0.0f ? 1u : 0u;
    P.rdt = 1.0f / P.dt;
    int vr;
    if ((vr = verify_divisor(c, P.dt, P.S, 0, &P.fast_dt))) return vr;
    {
        int e2 = 0;
        const bool pow2 = std::frexp(P.S, &e2) == 0.5f && P.S >= 1.0f && P.S <= 65536.0f;
        const bool g_ok = P.grav_div[0] == 0.0f && P.grav_div[2] == 0.0f && std::fabs(P.grav_div[1]) <= 0x1p40f;
        uint32_t q_ok = 0;
        if (pow2 && g_ok && !P.interacting && P.dt >= 0x1p-20f && P.dt <= 0x1p20f)
            if ((vr = verify_divisor(c, P.dt, P.S, 1, &q_ok))) return vr;
        P.p2g_fold = q_ok;
        P.dtS = P.dt * P.S;
        P.pow2_S = pow2 ? 1u : 0u;
        P.rS = 1.0f / P.S;
        P.ex_fold = (pow2 && P.dtS <= 0x1p20f && P.cd[0] >= 0x1p-20f && P.cd[1] >= 0x1p-20f && P.cd[2] >= 0x1p-20f) ? 1u : 0u;
        uint32_t cd_ok = 1;
        for (int a = 0; a < 3; ++a) P.rcd[a] = 1.0f / P.cd[a];
        for (int a = 0; a < 3 && cd_ok; ++a) {
            uint32_t ok = 0;
            if ((vr = verify_divisor(c, P.cd[a], 1.0f, 2, &ok))) return vr;
            cd_ok = ok;
        }
        P.fast_cd = cd_ok;
        // a power-of-two cell of at most 1 divides exactly as a product by 2^k (|x| <= 2^100 and
        // k <= 126 cannot overflow, and scaling up leaves no rounding)
        P.pow2_cd = 1u;
        for (int a = 0; a < 3; ++a) {
            int ea = 0;
            if (!(std::frexp(P.cd[a], &ea) == 0.5f && P.cd[a] <= 1.0f && P.cd[a] >= 0x1p-26f)) P.pow2_cd = 0u;
        }
    }
    P.wall_mode = g.wall_mode;
    P.wall_cells = g.wall_cells;
    P.direct_blocks = (!scene_only && c->direct_blocks) ? 1u : 0u;
    P.n_coll_rec = scene_only ? 0u : (uint32_t)c->n_coll_rec;
    P.n_coll_tri = scene_only ? 0u : (uint32_t)c->n_coll_tri;
    P.has_collider = (!scene_only && c->n_coll_rec > 0 && c->n_coll_words > 0) ? 1u : 0u;
    if (P.has_collider && P.coll_num_indices != c->coll_num_indices)
        return fail(c, MPM_E_RANGE, "uniform colliderNumIndices differs from the uploaded collider");
    P.margin = g.collider_margin;
    P.fr_ccd = g.friction_ccd;
    P.fr_prox = g.friction_proximity;
    P.vel_factor = g.collider_velocity_factor;
    if (!(P.dt > 0.0f)) return fail(c, MPM_E_RANGE, "simulationTimestep must be > 0");
    if (k_iter < 1 || k_iter > 64) return fail(c, MPM_E_RANGE, "pbmpm iterations must be in 1..64");
    return MPM_OK;
}
int make_params(mpm_ctx* c, SimParams& P) {
    if (c->ens) return fail(c, MPM_E_STATE, "this entry point is for single-scene contexts; use mpm_ensemble_*");
    const int rc = params_from(c, c->uni, c->cfg, c->k_cur, c->force_div_cur, false, P);
    P.per_particle_lame = c->per_particle_lame ? 1u : 0u;
    // one scene: chunks longest first (a pass hands them out in list order and ends on the short ones)
    P.chunk_lpt = 1u;
    P.fuse_keys = 1u;
    P.keys_ready = 0u;       // set per substep by keys_state
    return rc;
}

// Whether the stored keys (made by the previous substep's fused keys pass) describe the rows the next
// substep bins with these parameters: same buffer parity, same grid.
static void grid_of(const SimParams& P, float g[9]) {
    for (int a = 0; a < 3; ++a) { g[a] = P.gmin[a]; g[3 + a] = P.cd[a]; g[6 + a] = (float)P.res[a]; }
}
void keys_state(mpm_ctx* c, SimParams& P) {
    float g[9];
    grid_of(P, g);
    P.keys_ready = (P.fuse_keys && c->keys_parity == c->cur && std::memcmp(g, c->keys_grid, sizeof g) == 0) ? 1u : 0u;
}
// After a substep: with the fused keys pass its last passes made the keys of the rows that are now
// current (`cur` flipped already); a partial substep leaves none.
void keys_after(mpm_ctx* c, const SimParams& P, bool full) {
    if (full && P.fuse_keys) { c->keys_parity = c->cur; grid_of(P, c->keys_grid); }
    else c->keys_parity = -1;
}

uint32_t hash_host(uint32_t n) {
    uint32_t h = n + 374761393u;
    h = 668265263u * ((h << 17) | (h >> 15));
    h = 2246822519u * (h ^ (h >> 15));
    h = 3266489917u * (h ^ (h >> 13));
    return h ^ (h >> 16);
}

uint32_t grid_for(uint64_t work, uint32_t cap_wg) {
    uint64_t g = (work + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap_wg) g = cap_wg;
    return (uint32_t)g;
}

// One substep (or a debug prefix of it) enqueued on the stream. Returns the stage reached.
void enqueue_binning(mpm_ctx* c, const SimParams& P, const Bufs& B) {
    hipStream_t s = c->stream;
    // all workgroups resident at once (8 per CU at 256 threads): measured faster than one particle
    // per thread in two dispatch waves (1M particles: 30.3 -> 26.7 us)
    const uint32_t gk = grid_for(P.direct_blocks ? (uint64_t)c->n : std::max<uint64_t>(c->n, c->cap), (uint32_t)c->n_cu * 8u);
    kt(c, -1);
    if (!P.keys_ready) {
        if (P.fuse_keys) {
            // the key space may hold the counts of a fused keys pass whose keys are not used (new
            // particles, another grid, a discarded substep): clear it and the non-finite count first
            k_fill_u32<<<grid_for(c->K, 4096), 256, 0, s>>>(B.kcount, 0u, c->K);
            k_fill_u32<<<grid_for(c->K, 4096), 256, 0, s>>>(B.kmask, 0u, c->K);
            k_fill_u32<<<grid_for(c->K, 4096), 256, 0, s>>>(B.kalloc, 0u, c->K);
            k_fill_u32<<<1, 256, 0, s>>>(B.counters + C_NONFIN, 0u, 1u);
        }
        k_keys_count<<<gk, 256, 0, s>>>(P, B); kt(c, KID_KEYS);
    }
    enqueue_scan(c, P, B, s);
    const uint32_t gi = grid_for((uint64_t)c->max_runs * 27u, 4096);
    if (P.direct_blocks) {
        // the neighbour table needs only the scan here: one launch with the scatter
        k_scatter<true><<<std::max(grid_for(c->n, 1u << 30), gi), 256, 0, s>>>(P, B); kt(c, KID_SCATTER);
    } else {
        k_scatter<<<grid_for(c->n, 1u << 30), 256, 0, s>>>(P, B); kt(c, KID_SCATTER);
        k_insert<<<gi, 256, 0, s>>>(P, B); kt(c, KID_INSERT);
        k_compact<<<grid_for(c->cap, 1u << 30), 256, 0, s>>>(P, B); kt(c, KID_COMPACT);
        k_nbr<<<gi, 256, 0, s>>>(P, B); kt(c, KID_NBR);
    }
}

// Workgroups of the chunk kernels: four times the resident count (4 per CU at 128 VGPRs), so the
// dispatcher hands the rest out as workgroups finish and balances uneven chunks.
uint32_t chunk_grid(mpm_ctx* c) {
    uint32_t g = c->cfg.grid_workgroups ? c->cfg.grid_workgroups : (uint32_t)c->n_cu * 16u;
    return std::min<uint32_t>(g, std::max<uint32_t>(c->max_chunks, 1u));
}

// stage < 0: full substep. Returns true if the substep completed (buffers swap).
// `full` records the phase events of mpm_get_timings.
bool enqueue_substep(mpm_ctx* c, int method, const SimParams& P, int stage, bool full = false) {
    hipStream_t s = c->stream;
    Bufs B = make_bufs(c);
    if (full) hipEventRecord(c->ev[0], s);
    enqueue_binning(c, P, B);
    if (full) hipEventRecord(c->ev[1], s);
    if (stage == 0) return false;
    const uint32_t G = chunk_grid(c);
    const uint32_t gs = 256;
    if (method == MPM_METHOD_PBMPM) {
        const int k = P.k;
        if (P.learn_flags & 1u) {
            k_mlp_warmstart<<<std::min<uint32_t>(grid_for(c->n, 1u << 30), (uint32_t)c->n_cu * 4u), 256, 0, s>>>(
                mlp_bufs(c), B, c->n, P.learn_limit, c->counters + C_FALLBACK);
            kt(c, KID_MLP);
        }
        if (stage == MPM_STAGE_WARM_START) return false;
        k_pbmpm<PB_FIRST><<<G, 256, 0, s>>>(P, B, -1, 0, 1, -1); kt(c, KID_PB_FIRST);
        c->last_grid = 0;
        c->last_pbmpm = true;
        if (stage == 1) return false;
        for (int i = 1; i < k; ++i) {
            // the residual of every iteration when it is tracked; for the warm start's dprev alone,
            // only the dx of the G2P before LAST's (the last MID pass)
            const bool track = B.resid_dx && ((P.learn_flags & 2u) || i == k - 1);
            if (track) k_pbmpm<PB_MID, true><<<G, 256, 0, s>>>(P, B, (i - 1) % 3, i % 3, (i + 1) % 3, i - 1);
            else k_pbmpm<PB_MID><<<G, 256, 0, s>>>(P, B, (i - 1) % 3, i % 3, (i + 1) % 3, i - 1);
            kt(c, KID_PB_MID);
            c->last_grid = i % 3;
            if (stage == i + 1) return false;
        }
        if (full) hipEventRecord(c->ev[2], s);
        if (B.resid_dx) k_pbmpm<PB_LAST, true><<<G, 256, 0, s>>>(P, B, (k - 1) % 3, -1, -1, k - 1);
        else k_pbmpm<PB_LAST><<<G, 256, 0, s>>>(P, B, (k - 1) % 3, -1, -1, k - 1);
        kt(c, KID_PB_LAST);
        if (P.learn_flags & 2u) k_resid_roll<<<1, 64, 0, s>>>(B, k, c->resid_peak);
    } else {
        k_explicit_p2g<0><<<G, 256, 0, s>>>(P, B); kt(c, KID_EX_P2G);
        c->last_grid = 0;
        c->last_pbmpm = false;
        if (stage == 1) return false;
        k_grid_update<<<G, 256, 0, s>>>(P, B); kt(c, KID_EX_GRID);
        if (stage == 2) return false;
        if (full) hipEventRecord(c->ev[2], s);
        k_explicit_g2p<0><<<G, 256, 0, s>>>(P, B); kt(c, KID_EX_G2P);
    }
    if (P.has_collider) {
        const uint32_t gc = std::min<uint32_t>(grid_for(c->n, 1u << 30), (uint32_t)c->n_cu * 8u);
        if (P.n_coll_rec <= COLL_LDS_REC && P.n_coll_tri <= COLL_LDS_TRI) k_collide<true><<<gc, 256, 0, s>>>(P, B);
        else k_collide<false><<<gc, 256, 0, s>>>(P, B);
        kt(c, KID_COLLIDE);     // also the unbinned particles
    } else {
        k_straggle<<<gs, 256, 0, s>>>(P, B); kt(c, KID_STRAGGLE);
    }
    if (P.plast_mode != 0u) { k_plastic_svd<<<grid_for(c->n, 1u << 30), 256, 0, s>>>(P, B); kt(c, KID_PLASTIC_SVD); }
    if (full) hipEventRecord(c->ev[3], s);
    return true;
}

// A captured substep holds its kernels' arguments by value: the parameters and every device
// address. It is replayed only while both are unchanged; a buffer freed and reallocated since
// (a new collider, residual tracking or network buffers switched on) forces a new capture.
// A captured graph for (method, current parity, many) that these parameters and buffers can replay.
bool graph_valid(mpm_ctx* c, int method, const SimParams& P, int many) {
    const int par = c->cur;
    const Bufs B = make_bufs(c);
    const MlpBufs M = mlp_bufs(c);
    return c->gexec[method][par][many] && std::memcmp(&c->gparams[method][par][many], &P, sizeof(SimParams)) == 0 &&
           std::memcmp(&c->gbufs[method][par][many], &B, sizeof(Bufs)) == 0 &&
           std::memcmp(&c->gmlp[method][par][many], &M, sizeof(MlpBufs)) == 0;
}
// many = 0: a graph of one substep from the current buffer parity; 1: of GRAPH_SUBSTEPS substeps
// (an even count, so the parity after it is the one before).
int ensure_graph(mpm_ctx* c, int method, const SimParams& P, int many = 0) {
    if (graph_valid(c, method, P, many)) return MPM_OK;
    const int par = c->cur;
    const Bufs B = make_bufs(c);
    const MlpBufs M = mlp_bufs(c);
    if (c->gexec[method][par][many]) { hipGraphExecDestroy(c->gexec[method][par][many]); c->gexec[method][par][many] = nullptr; }
    if (c->graph[method][par][many]) { hipGraphDestroy(c->graph[method][par][many]); c->graph[method][par][many] = nullptr; }
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int n = many ? GRAPH_SUBSTEPS : 1;
    for (int i = 0; i < n; ++i) {
        enqueue_substep(c, method, P, -1, false);
        if (many) c->cur ^= 1;        // the next substep of the graph reads what this one wrote
    }
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t e = nullptr;
    hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (r != hipSuccess) { hipGraphDestroy(g); return fail(c, MPM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(r)); }
    c->graph[method][par][many] = g;
    c->gexec[method][par][many] = e;
    c->gparams[method][par][many] = P;
    c->gbufs[method][par][many] = B;
    c->gmlp[method][par][many] = M;
    return MPM_OK;
}

// Device buffers of a context, sized by n, stride, K, cap, max_blocks, max_runs, max_chunks, ntiles.
int alloc_buffers(mpm_ctx* c) {
    const uint64_t K = c->K;
    int rc = MPM_OK;
    const size_t ps = (size_t)NPLANES * c->stride;
    if ((rc = dalloc(c, &c->planes[0], ps)) || (rc = dalloc(c, &c->planes[1], ps)) ||
        (rc = dalloc(c, &c->pid[0], c->stride)) || (rc = dalloc(c, &c->pid[1], c->stride)) ||
        (rc = dalloc(c, &c->fpid, 3 * (size_t)c->stride)) ||
        (rc = dalloc(c, &c->relax, c->stride)) || (rc = dalloc(c, &c->key, c->stride)) || (rc = dalloc(c, &c->lofs, c->stride)) ||
        (rc = dalloc(c, &c->mask, c->stride)) || (rc = dalloc(c, &c->perm, c->stride)) ||
        (rc = dalloc(c, &c->kcount, K)) || (rc = dalloc(c, &c->kmask, K)) || (rc = dalloc(c, &c->kcursor, K)) ||
        (rc = dalloc(c, &c->tile_sums, SCAN_NS * (size_t)c->ntiles)) || (rc = dalloc(c, &c->tile_ofs, SCAN_NS * (size_t)c->ntiles)) ||
        (rc = dalloc(c, &c->kalloc, K)) || (rc = dalloc(c, &c->kslot, K)) ||
        (rc = dalloc(c, &c->run_key, c->max_runs)) || (rc = dalloc(c, &c->run_start, c->max_runs)) ||
        (rc = dalloc(c, &c->run_count, c->max_runs)) || (rc = dalloc(c, &c->run_mask, c->max_runs)) ||
        (rc = dalloc(c, &c->chunks, c->max_chunks)) ||
        (rc = dalloc(c, &c->table, c->direct_blocks ? 1 : c->cap)) || (rc = dalloc(c, &c->slot_val, c->direct_blocks ? 1 : c->cap)) ||
        (rc = dalloc(c, &c->nbr, 27 * (size_t)c->max_runs)) ||
        (rc = dalloc(c, &c->block_key, c->max_blocks)) ||
        (rc = dalloc(c, &c->grid[0], 64 * (size_t)c->max_blocks)) || (rc = dalloc(c, &c->grid[1], 64 * (size_t)c->max_blocks)) ||
        (rc = dalloc(c, &c->grid[2], 64 * (size_t)c->max_blocks)) || (rc = dalloc(c, &c->gmass, 64 * (size_t)c->max_blocks)) ||
        (rc = dalloc(c, &c->counters, C_NCOUNTERS)) || (rc = dalloc(c, &c->cqueue, c->stride)) || (rc = dalloc(c, &c->d_scratch, 4))) {
        return rc;
    }
    // zero-initialised key-space accumulators and counters (kernels keep them zero afterwards)
    k_fill_u32<<<grid_for(K, 4096), 256, 0, c->stream>>>(c->kcount, 0u, (uint32_t)K);
    k_fill_u32<<<grid_for(K, 4096), 256, 0, c->stream>>>(c->kmask, 0u, (uint32_t)K);
    k_fill_u32<<<grid_for(K, 4096), 256, 0, c->stream>>>(c->kalloc, 0u, (uint32_t)K);
    k_iota<<<grid_for(c->stride, 4096), 256, 0, c->stream>>>(c->pid[0], c->stride);
    k_fill_u32<<<1, 256, 0, c->stream>>>(c->counters, 0u, C_NCOUNTERS);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { (void)hipGetLastError(); return MPM_E_HIP; }
    for (int i = 0; i < 6; ++i) hipEventCreate(&c->ev[i]);
    return MPM_OK;
}


// ------------------------------------------------------------------ batched scenes
Ens ens_bufs(mpm_ctx* c) {
    Ens E;
    E.sp = c->d_sp; E.part_base = c->d_part_base; E.key_base = c->d_key_base; E.nphase = c->d_nphase; E.method = c->d_method;
    E.nblk = c->d_nblk; E.cbeg = c->d_cbeg; E.seg = c->d_seg; E.done = c->d_done; E.ctl = c->d_ctl;
    E.rbase = c->direct_blocks ? c->d_rbase : nullptr;
    E.S = c->ens; E.R = c->ens_R;
    return E;
}

// One substep of every scene: shared binning, task table, the transfer phases (one persistent
// launch, or one launch per phase), unbinned particles.
void enqueue_ens_substep(mpm_ctx* c, bool full) {
    hipStream_t s = c->stream;
    const SimParams& P0 = c->ens_P0;
    Bufs B = make_bufs(c);
    Ens E = ens_bufs(c);
    if (full) hipEventRecord(c->ev[0], s);
    const uint32_t gk = grid_for(c->direct_blocks ? (uint64_t)c->n : std::max<uint64_t>(c->n, c->cap), 8192);
    kt(c, -1);
    k_ens_keys_count<<<gk, 256, 0, s>>>(P0, B, E); kt(c, KID_ENS_KEYS);
    enqueue_scan(c, P0, B, s);
    k_scatter<<<grid_for(c->n, 1u << 30), 256, 0, s>>>(P0, B); kt(c, KID_SCATTER);
    const uint32_t gi = grid_for((uint64_t)c->max_runs * 27u, 8192);
    if (!c->direct_blocks) {
        k_ens_insert<<<gi, 256, 0, s>>>(P0, B, E); kt(c, KID_ENS_INSERT);
        k_ens_compact<<<grid_for(c->cap, 1u << 30), 256, 0, s>>>(P0, B, E); kt(c, KID_ENS_COMPACT);
    }
    k_ens_nbr<<<gi, 256, 0, s>>>(P0, B, E); kt(c, KID_ENS_NBR);
    k_ens_tasks<<<1, 256, 0, s>>>(B, E); kt(c, KID_ENS_TASKS);
    if (full) hipEventRecord(c->ev[1], s);
    const uint32_t G = (uint32_t)c->n_cu * c->ens_wg_per_cu;
    const int nl = c->ens_persistent ? 1 : (int)c->ens_R;
    for (int r = 0; r < nl; ++r) {
        const int ph = c->ens_persistent ? -1 : r;
        if (c->ens_wg_per_cu >= 4) k_ens_transfer<4><<<G, 256, 0, s>>>(B, E, ph);
        else if (c->ens_wg_per_cu == 3) k_ens_transfer<3><<<G, 256, 0, s>>>(B, E, ph);
        else k_ens_transfer<2><<<G, 256, 0, s>>>(B, E, ph);
        kt(c, KID_ENS_TRANSFER);
    }
    if (full) hipEventRecord(c->ev[2], s);
    k_ens_straggle<<<256, 256, 0, s>>>(P0, B, E); kt(c, KID_ENS_STRAGGLE);
    if (full) hipEventRecord(c->ev[3], s);
}

int ens_graph(mpm_ctx* c) {
    const int par = c->cur;
    if (c->gexec[0][par][0] && c->ens_graph_ok[par]) return MPM_OK;
    if (c->gexec[0][par][0]) { hipGraphExecDestroy(c->gexec[0][par][0]); c->gexec[0][par][0] = nullptr; }
    if (c->graph[0][par][0]) { hipGraphDestroy(c->graph[0][par][0]); c->graph[0][par][0] = nullptr; }
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_ens_substep(c, false);
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t e = nullptr;
    hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (r != hipSuccess) { hipGraphDestroy(g); return fail(c, MPM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(r)); }
    c->graph[0][par][0] = g;
    c->gexec[0][par][0] = e;
    c->ens_graph_ok[par] = true;
    return MPM_OK;
}

// Per-scene parameters with the shared chunk size; uploads them.
int ens_set_chunk(mpm_ctx* c, uint32_t chunk) {
    c->chunk = chunk;
    c->ens_P0.chunk = chunk;
    for (auto& P : c->ens_sp) P.chunk = chunk;
    HIPCHK(c, hipMemcpy(c->d_sp, c->ens_sp.data(), c->ens_sp.size() * sizeof(SimParams), hipMemcpyHostToDevice));
    c->ens_graph_ok[0] = c->ens_graph_ok[1] = false;
    return MPM_OK;
}

int ens_check_flags(mpm_ctx* c) {
    uint32_t f = 0;
    HIPCHK(c, hipMemcpy(&f, c->d_ctl + 2, 4, hipMemcpyDeviceToHost));
    if (f & ENS_ERR_WAIT) {
        const uint32_t z = 0;
        HIPCHK(c, hipMemcpy(c->d_ctl + 2, &z, 4, hipMemcpyHostToDevice));
        return fail(c, MPM_E_HIP, "batched step: a phase dependency wait hit its spin limit (results of that substep are invalid)");
    }
    return MPM_OK;
}
}  // namespace

extern "C" {

const char* mpm_version(void) { return k_version; }

void mpm_config_defaults(mpm_config* g) {
    if (!g) return;
    std::memset(g, 0, sizeof(*g));
    g->n_particles = 0;
    g->grid_res[0] = g->grid_res[1] = g->grid_res[2] = 256;   // SimulationState.svelte.ts:22-24
    for (int a = 0; a < 3; ++a) { g->grid_min[a] = -5.0f; g->grid_max[a] = 5.0f; }   // GpuSnowPipelineRunner.svelte.ts:133-134
    g->fixed_point_scale = 65536.0f;
    g->dt = 1.0f / 384.0f;                  // SimulationState.svelte.ts:26 (PBMPM)
    g->max_blocks = 0;
    g->hash_map_size = 0;
    g->max_probes = 100;
    g->pbmpm_iterations = 3;
    g->pbmpm_force_divisor = 3.0f;
    g->gravity[0] = 0.0f; g->gravity[1] = 0.0f; g->gravity[2] = -9.81f;
    g->youngs_modulus = 1.4e5;
    g->poissons_ratio = 0.2;
    g->hardening_explicit = 5.0f;
    g->hardening_pbmpm = 20.0f;
    g->pbmpm_blend = 0.95f;
    g->pbmpm_volume_exponent = -0.3333333f;
    g->theta_c = 2.5e-2;
    g->theta_s = 7.5e-3;
    g->density = 400.0;
    g->grad_scale = 1.0f;
    g->collider_margin = 0.05f;
    g->friction_ccd = 0.1f;
    g->friction_proximity = 0.0f;
    g->collider_velocity_factor = 0.2f;
    g->wall_mode = 0;
    g->wall_cells = 3;
    g->chunk_particles = 0;
    g->grid_workgroups = 0;
    g->device = 0;
    g->use_graph = 1;
}

static void release(mpm_ctx* c) {
    for (int m = 0; m < 2; ++m) for (int p = 0; p < 2; ++p) for (int k = 0; k < 2; ++k) {
        if (c->gexec[m][p][k]) hipGraphExecDestroy(c->gexec[m][p][k]);
        if (c->graph[m][p][k]) hipGraphDestroy(c->graph[m][p][k]);
    }
    for (int i = 0; i < 6; ++i) if (c->ev[i]) hipEventDestroy(c->ev[i]);
    for (hipEvent_t e : c->kt_ev) hipEventDestroy(e);
    for (void* p : c->allocs) hipFree(p);
    if (c->stream) hipStreamDestroy(c->stream);
}

int mpm_create(const mpm_config* cfg, mpm_ctx** out) {
    if (!cfg || !out) return MPM_E_INVALID;
    *out = nullptr;
    if (cfg->n_particles == 0) return MPM_E_INVALID;
    for (int a = 0; a < 3; ++a) if (cfg->grid_res[a] == 0 || cfg->grid_res[a] > 65536 || !(cfg->grid_max[a] > cfg->grid_min[a])) return MPM_E_INVALID;
    if (cfg->pbmpm_iterations < 1 || cfg->pbmpm_iterations > 64) return MPM_E_INVALID;
    // row byte offsets are 32-bit (the transfer passes address a plane's row i at base + 4 i), and
    // chunk counts are formed as (count + chunk - 1) / chunk in u32
    if (cfg->n_particles > MPM_MAX_PARTICLES) return MPM_E_RANGE;
    if (cfg->chunk_particles && (cfg->chunk_particles < 256 || cfg->chunk_particles > 65536)) return MPM_E_RANGE;
    if (cfg->plasticity_mode > 2) return MPM_E_INVALID;
    mpm_ctx* c = new mpm_ctx();
    c->cfg = *cfg;
    c->device = cfg->device;
    if (hipSetDevice(c->device) != hipSuccess) { (void)hipGetLastError(); delete c; return MPM_E_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { (void)hipGetLastError(); delete c; return MPM_E_HIP; }
    scratch_resolve(c->device);
    c->n = cfg->n_particles;
    c->stride = (c->n + 63u) & ~63u;
    for (int a = 0; a < 3; ++a) c->nb[a] = (cfg->grid_res[a] + 3u) / 4u;
    uint64_t K = (uint64_t)c->nb[0] * c->nb[1] * c->nb[2];
    if (K >= 0xFFFFFFF0ull) { release(c); delete c; return MPM_E_RANGE; }
    c->K = (uint32_t)K;
    c->chunk = cfg->chunk_particles ? cfg->chunk_particles : 512u;
    c->k_cur = cfg->pbmpm_iterations;
    c->force_div_cur = cfg->pbmpm_force_divisor;
    uint64_t mb = cfg->max_blocks ? cfg->max_blocks : std::min<uint64_t>(K, std::max<uint64_t>(8ull * c->n, 64));
    c->max_blocks = (uint32_t)mb;
    c->cap = cfg->hash_map_size ? cfg->hash_map_size : next_prime((uint32_t)std::min<uint64_t>(2ull * mb + 1, 0xFFFFFF00ull));
    c->direct_blocks = cfg->hash_map_size == 0;
    c->max_runs = (uint32_t)std::min<uint64_t>(c->n, K);
    c->max_chunks = (c->n + c->chunk - 1) / c->chunk + c->max_runs;
    set_scan_tiles(c, K);
    int rc = alloc_buffers(c);
    if (rc) { release(c); delete c; return rc; }
    // uniform fields the runner writes at construction (GpuSnowPipelineRunner.svelte.ts:136-144,171-177)
    std::memset(c->uni, 0, sizeof(c->uni));
    wru<float>(c->uni, 0, cfg->dt);
    wru<float>(c->uni, 4, cfg->fixed_point_scale);
    for (int a = 0; a < 3; ++a) {
        wru<float>(c->uni, 16 + 4 * a, cfg->grid_min[a]);
        wru<float>(c->uni, 32 + 4 * a, cfg->grid_max[a]);
        wru<uint32_t>(c->uni, 208 + 4 * a, cfg->grid_res[a]);
        wru<float>(c->uni, 352 + 4 * a, (float)(((double)cfg->grid_max[a] - (double)cfg->grid_min[a]) / (double)cfg->grid_res[a]));
    }
    for (int i = 0; i < 4; ++i) { wru<float>(c->uni, 256 + 20 * i, 1.0f); wru<float>(c->uni, 432 + 20 * i, 1.0f); }
    *out = c;
    return MPM_OK;
}

void mpm_destroy(mpm_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    release(c);
    delete c;
}

const char* mpm_last_error(mpm_ctx* c) { return c ? c->err.c_str() : "null context"; }
void* mpm_stream(mpm_ctx* c) { return c ? (void*)c->stream : nullptr; }

int mpm_write_uniforms(mpm_ctx* c, size_t offset, const void* bytes, size_t len) {
    if (!c || (!bytes && len)) return MPM_E_INVALID;
    if (offset > MPM_UNIFORMS_BYTES || len > MPM_UNIFORMS_BYTES - offset) return fail(c, MPM_E_RANGE, "uniform write out of the 37376-byte buffer");
    std::memcpy(c->uni + offset, bytes, len);
    return MPM_OK;
}
int mpm_read_uniforms(mpm_ctx* c, size_t offset, void* bytes, size_t len) {
    if (!c || (!bytes && len)) return MPM_E_INVALID;
    if (offset > MPM_UNIFORMS_BYTES || len > MPM_UNIFORMS_BYTES - offset) return fail(c, MPM_E_RANGE, "uniform read out of the 37376-byte buffer");
    std::memcpy(bytes, c->uni + offset, len);
    return MPM_OK;
}

int mpm_set_collider(mpm_ctx* c, const uint32_t* data, size_t n_words, uint32_t n_indices, const void* nodes, size_t n_nodes) {
    if (!c) return MPM_E_INVALID;
    if ((n_words && !data) || (n_nodes && !nodes)) return fail(c, MPM_E_INVALID, "null collider pointer");
    if (n_indices > n_words || n_indices % 3) return fail(c, MPM_E_INVALID, "collider index count must be a multiple of 3 within the data");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // validate the node array before it can steer a device traversal
    const uint32_t* w = (const uint32_t*)nodes;
    const size_t n_tris = n_indices / 3;
    for (size_t i = 0; i < n_nodes; ++i) {
        const uint32_t l = w[8 * i + 3], r = w[8 * i + 7];
        if (r & 0x80000000u) { if ((uint64_t)l + (r & 0x7FFFFFFFu) > n_tris) return fail(c, MPM_E_INVALID, "BVH leaf range outside the triangle list"); }
        else if (l >= n_nodes || r >= n_nodes) return fail(c, MPM_E_INVALID, "BVH child index out of range");
    }
    const size_t n_vert_words = n_words - n_indices;
    for (size_t i = 0; i < n_indices; ++i) if ((uint64_t)data[i] * 3 + 2 >= n_vert_words) return fail(c, MPM_E_INVALID, "collider index beyond the vertex data");
    if (n_nodes > REC_INDEX || n_tris > REC_INDEX) return fail(c, MPM_E_RANGE, "collider too large (2^30 nodes or triangles)");
    // Threaded records in the visiting order of the reference's traversal (colliderPrelude.wgsl
    // :174-356: pop a node; cull by box; leaf -> triangles; else, while fewer than 31 entries are
    // pending, push left then right). The right subtree is visited first; a node popped with s
    // pending entries has its right child popped with s + 1 and its left child with s.
    std::vector<uint32_t> rec;   // 8 words per record
    if (n_words && n_nodes) {
        std::vector<uint8_t> seen(n_nodes, 0);
        struct Task { uint32_t node, s, fix; };   // fix != ~0u: set the skip index of record `fix`
        std::vector<Task> todo{{0u, 0u, ~0u}};
        while (!todo.empty()) {
            const Task t = todo.back();
            todo.pop_back();
            if (t.fix != ~0u) { rec[8 * (size_t)t.fix + 7] |= (uint32_t)(rec.size() / 8); continue; }
            if (seen[t.node]) return fail(c, MPM_E_INVALID, "BVH nodes do not form a tree (a node is reached twice)");
            seen[t.node] = 1;
            const uint32_t* nw = w + 8 * (size_t)t.node;
            const uint32_t l = nw[3], r = nw[7];
            const uint32_t idx = (uint32_t)(rec.size() / 8);
            uint32_t lo_w = 0, hi_w;
            if (r & 0x80000000u) { lo_w = l; hi_w = REC_LEAF | (l + (r & 0x7FFFFFFFu)); }
            else if (t.s < 31u) {
                hi_w = REC_DESCEND;
                todo.push_back({0u, 0u, idx});
                todo.push_back({l, t.s, ~0u});
                todo.push_back({r, t.s + 1u, ~0u});
            } else hi_w = idx + 1u;
            rec.insert(rec.end(), {nw[0], nw[1], nw[2], lo_w, nw[4], nw[5], nw[6], hi_w});
        }
    }
    if (c->coll_data) { hipFree(c->coll_data); c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->coll_data)); c->coll_data = nullptr; }
    if (c->crec) { hipFree(c->crec); c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->crec)); c->crec = nullptr; }
    if (c->ctri) { hipFree(c->ctri); c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->ctri)); c->ctri = nullptr; }
    c->n_coll_words = 0; c->n_coll_rec = 0; c->n_coll_tri = 0;
    int rc;
    if (n_words && n_nodes) {
        if ((rc = dalloc(c, &c->coll_data, n_words))) return rc;
        if ((rc = dalloc(c, &c->crec, rec.size() / 4))) return rc;
        if ((rc = dalloc(c, &c->ctri, (size_t)CT_WORDS * n_tris))) return rc;
        HIPCHK(c, hipMemcpy(c->coll_data, data, n_words * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->crec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice));
        c->n_coll_words = n_words; c->n_coll_rec = rec.size() / 8; c->n_coll_tri = n_tris;
    }
    c->coll_num_indices = n_indices;
    wru<uint32_t>(c->uni, 348, n_indices);
    return MPM_OK;
}

int mpm_build_bvh(const float* pos, size_t n_vertices, const uint32_t* indices, size_t n_indices,
                  void* nodes_out, size_t nodes_cap, size_t* n_nodes, uint32_t* tri_order) {
    if (!n_nodes || (n_indices && (!pos || !indices)) || n_indices % 3) return MPM_E_INVALID;
    std::vector<mpm_bvh::BNode> nodes;
    std::vector<uint32_t> order;
    try {
        if (!mpm_bvh::build(pos, n_vertices, indices, n_indices, nodes, order)) return MPM_E_INVALID;
    } catch (const std::bad_alloc&) {
        return MPM_E_NOMEM;
    }
    *n_nodes = nodes.size();
    if (!nodes_out || nodes_cap < nodes.size()) return MPM_E_RANGE;
    mpm_bvh::pack(nodes, (uint8_t*)nodes_out);
    if (tri_order) std::memcpy(tri_order, order.data(), order.size() * 4);
    return MPM_OK;
}

int mpm_upload_particles(mpm_ctx* c, const void* aos, size_t n) {
    if (!c || !aos) return MPM_E_INVALID;
    if (c->ens) return fail(c, MPM_E_INVALID, "batched context: use mpm_ensemble_upload");
    if (n != c->n) return fail(c, MPM_E_INVALID, "particle count differs from the context's n_particles");
    HIPCHK(c, hipSetDevice(c->device));
    float* tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, n * MPM_PARTICLE_BYTES));
    hipError_t e = hipMemcpyAsync(tmp, aos, n * MPM_PARTICLE_BYTES, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        k_upload<<<grid_for(n, 1u << 30), 256, 0, c->stream>>>(tmp, c->planes[c->cur], c->pid[c->cur], c->fpid, c->n, c->stride);
        e = hipStreamSynchronize(c->stream);
    }
    hipFree(tmp);
    if (e != hipSuccess) return fail(c, MPM_E_HIP, std::string("particle upload: ") + hipGetErrorString(e));
    c->has_particles = true;
    c->keys_parity = -1;
    return MPM_OK;
}

int mpm_scatter_particles(mpm_ctx* c, const float* mesh_xyzw, size_t n_vertices) {
    if (!c || !mesh_xyzw) return MPM_E_INVALID;
    if (c->ens) return fail(c, MPM_E_INVALID, "batched context: scatter into a single-scene context");
    if (n_vertices < 3 || n_vertices % 3 || n_vertices > (1ull << 30)) return fail(c, MPM_E_INVALID, "mesh vertex count must be a positive multiple of 3");
    HIPCHK(c, hipSetDevice(c->device));
    float mn[3], mx[3];
    for (int a = 0; a < 3; ++a) { mn[a] = rdu<float>(c->uni, 176 + 4 * a); mx[a] = rdu<float>(c->uni, 192 + 4 * a); }
    float4* dm = nullptr;
    HIPCHK(c, hipMalloc(&dm, n_vertices * 16));
    hipError_t e = hipMemcpyAsync(dm, mesh_xyzw, n_vertices * 16, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        k_scatter_init<<<grid_for(c->n, 1u << 30), 256, 0, c->stream>>>(dm, (uint32_t)(n_vertices / 3), mn[0], mn[1], mn[2], mx[0], mx[1], mx[2],
                                                                        c->planes[c->cur], c->pid[c->cur], c->fpid, c->n, c->stride);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    hipFree(dm);
    if (e != hipSuccess) return fail(c, MPM_E_HIP, std::string("particle scatter: ") + hipGetErrorString(e));
    c->has_particles = true;
    c->keys_parity = -1;
    return MPM_OK;
}

int mpm_splat(mpm_ctx* c, int kind, const uint32_t* res_in, uint32_t* host_out, size_t cap) {
    if (!c) return MPM_E_INVALID;
    if (c->ens) return fail(c, MPM_E_INVALID, "batched context: splat a single-scene context");
    if (kind != MPM_SPLAT_DENSITY && kind != MPM_SPLAT_MASS) return fail(c, MPM_E_INVALID, "unknown splat kind");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "no particles uploaded");
    uint32_t res[3];
    for (int a = 0; a < 3; ++a) res[a] = kind == MPM_SPLAT_MASS ? rdu<uint32_t>(c->uni, 208 + 4 * a) : (res_in ? res_in[a] : 0u);
    const uint64_t cells = (uint64_t)res[0] * res[1] * res[2];
    if (cells == 0 || cells > (1ull << 32)) return fail(c, MPM_E_RANGE, "splat grid resolution must be positive and hold at most 2^32 cells");
    if (host_out && cap < cells) return fail(c, MPM_E_RANGE, "splat output capacity too small");
    HIPCHK(c, hipSetDevice(c->device));
    if (cells > c->splat_cells) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->splat) { hipFree(c->splat); c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->splat)); c->splat = nullptr; }
        int rc;
        if ((rc = dalloc(c, &c->splat, (size_t)cells))) return rc;
        c->splat_cells = cells;
    }
    const float S = rdu<float>(c->uni, 4);
    float g0[3], g1[3];
    for (int a = 0; a < 3; ++a) { g0[a] = rdu<float>(c->uni, 16 + 4 * a); g1[a] = rdu<float>(c->uni, 32 + 4 * a); }
    HIPCHK(c, hipMemsetAsync(c->splat, 0, cells * 4, c->stream));
    const uint32_t G = grid_for(c->n, 1u << 30);
    if (kind == MPM_SPLAT_DENSITY)
        k_splat<SPLAT_DENSITY><<<G, 256, 0, c->stream>>>(c->planes[c->cur], c->n, c->stride, g0[0], g0[1], g0[2], g1[0], g1[1], g1[2], S, res[0], res[1], res[2], c->splat);
    else
        k_splat<SPLAT_MASS><<<G, 256, 0, c->stream>>>(c->planes[c->cur], c->n, c->stride, g0[0], g0[1], g0[2], g1[0], g1[1], g1[2], S, res[0], res[1], res[2], c->splat);
    HIPCHK(c, hipGetLastError());
    if (host_out) {
        HIPCHK(c, hipMemcpyAsync(host_out, c->splat, cells * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return MPM_OK;
}
void* mpm_splat_grid(mpm_ctx* c) { return c ? (void*)c->splat : nullptr; }

// mpm_splat with the accumulation grid in the file-scope `scratch` array instead of the context's
// lazily grown grid: no hipMalloc / hipFree, the grid cleared by a fill kernel (no memset node
// for a capturing caller), copied to host_out. Calls that overlap on one GPU share the array.
int mpm_splat_noalloc(mpm_ctx* c, int kind, const uint32_t* res_in, uint32_t* host_out, size_t cap) {
    if (!c) return MPM_E_INVALID;
    if (c->ens) return fail(c, MPM_E_INVALID, "batched context: splat a single-scene context");
    if (kind != MPM_SPLAT_DENSITY && kind != MPM_SPLAT_MASS) return fail(c, MPM_E_INVALID, "unknown splat kind");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "no particles uploaded");
    if (!host_out) return fail(c, MPM_E_INVALID, "the allocation-free splat needs host_out");
    uint32_t res[3];
    for (int a = 0; a < 3; ++a) res[a] = kind == MPM_SPLAT_MASS ? rdu<uint32_t>(c->uni, 208 + 4 * a) : (res_in ? res_in[a] : 0u);
    const uint64_t cells = (uint64_t)res[0] * res[1] * res[2];
    if (cells == 0) return fail(c, MPM_E_RANGE, "splat grid resolution must be positive");
    if (cells > MPM_SPLAT_NOALLOC_MAX_CELLS)
        return fail(c, MPM_E_RANGE, "allocation-free splat grid holds at most 2^21 cells (" + std::to_string(cells) + " asked); use mpm_splat");
    if (cap < cells) return fail(c, MPM_E_RANGE, "splat output capacity too small");
    HIPCHK(c, hipSetDevice(c->device));
    uint32_t* grid = scratch_ptr(c->device);
    if (!grid) {
        const hipError_t e = (c->device >= 0 && c->device < SCRATCH_MAX_DEVICES) ? g_scratch_err[c->device] : hipErrorInvalidDevice;
        return fail(c, MPM_E_HIP, std::string("no device address for the splat scratch array: ") + hipGetErrorString(e));
    }
    const float S = rdu<float>(c->uni, 4);
    float g0[3], g1[3];
    for (int a = 0; a < 3; ++a) { g0[a] = rdu<float>(c->uni, 16 + 4 * a); g1[a] = rdu<float>(c->uni, 32 + 4 * a); }
    k_fill_u32<<<grid_for(cells, 4096), 256, 0, c->stream>>>(grid, 0u, (uint32_t)cells);
    const uint32_t G = grid_for(c->n, 1u << 30);
    if (kind == MPM_SPLAT_DENSITY)
        k_splat<SPLAT_DENSITY><<<G, 256, 0, c->stream>>>(c->planes[c->cur], c->n, c->stride, g0[0], g0[1], g0[2], g1[0], g1[1], g1[2], S, res[0], res[1], res[2], grid);
    else
        k_splat<SPLAT_MASS><<<G, 256, 0, c->stream>>>(c->planes[c->cur], c->n, c->stride, g0[0], g0[1], g0[2], g1[0], g1[1], g1[2], S, res[0], res[1], res[2], grid);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(host_out, grid, cells * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MPM_OK;
}

int mpm_load_glb(const void* glb, size_t len, mpm_scene_arrays* a, char* err, size_t err_cap) {
    auto seterr = [&](const std::string& m) { if (err && err_cap) { std::strncpy(err, m.c_str(), err_cap - 1); err[err_cap - 1] = 0; } };
    if (!glb || !a) { seterr("null argument"); return MPM_E_INVALID; }
    mpm_gltf::Scene sc;
    mpm_gltf::Loader ld;
    // no exception crosses the C ABI: an allocation the file's sizes make impossible is an error
    try {
        if (!ld.load((const uint8_t*)glb, len, sc)) { seterr(ld.err); return MPM_E_INVALID; }
    } catch (const std::bad_alloc&) {
        seterr("out of host memory while reading the GLB");
        return MPM_E_NOMEM;
    } catch (const std::exception& ex) {
        seterr(std::string("GLB reader: ") + ex.what());
        return MPM_E_INVALID;
    }
    const size_t nv = sc.positions.size() / 3, nn = sc.normals.size() / 3, ni = sc.indices.size(), no = sc.objects.size(), ns = sc.soup.size() / 4;
    const bool room = (!a->positions || a->n_vertices >= nv) && (!a->uvs || a->n_vertices >= nv) &&
                      (!a->material_indices || a->n_vertices >= nv) && (!a->normals || a->n_normals >= nn) &&
                      (!a->indices || a->n_indices >= ni) && (!a->objects || a->n_objects >= no) && (!a->soup_xyzw || a->n_soup >= ns);
    a->n_vertices = nv; a->n_normals = nn; a->n_indices = ni; a->n_objects = no; a->n_soup = ns;
    if (!room) { seterr("output capacity too small"); return MPM_E_RANGE; }
    if (a->positions) std::memcpy(a->positions, sc.positions.data(), nv * 12);
    if (a->normals) std::memcpy(a->normals, sc.normals.data(), nn * 12);
    if (a->uvs) std::memcpy(a->uvs, sc.uvs.data(), nv * 8);
    if (a->material_indices) std::memcpy(a->material_indices, sc.material_indices.data(), nv * 4);
    if (a->indices) std::memcpy(a->indices, sc.indices.data(), ni * 4);
    if (a->soup_xyzw) std::memcpy(a->soup_xyzw, sc.soup.data(), ns * 16);
    if (a->objects)
        for (size_t o = 0; o < no; ++o) {
            uint32_t* w = a->objects + 8 * o;
            std::memcpy(w, sc.objects[o].mn, 12); w[3] = sc.objects[o].start;
            std::memcpy(w + 4, sc.objects[o].mx, 12); w[7] = sc.objects[o].count;
        }
    return MPM_OK;
}

int mpm_set_particle_youngs(mpm_ctx* c, const float* E, size_t n) {
    if (!c) return MPM_E_INVALID;
    if (c->ens) return fail(c, MPM_E_INVALID, "batched context: per-particle moduli are for single-scene contexts");
    if (!E) { c->per_particle_lame = false; return MPM_OK; }
    if (n != c->n) return fail(c, MPM_E_INVALID, "one modulus per particle");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "upload the particles first");
    const double nu = c->cfg.poissons_ratio;
    std::vector<float> mu(n), la(n);
    for (size_t i = 0; i < n; ++i) {
        if (!(E[i] > 0.0f) || !std::isfinite(E[i])) return fail(c, MPM_E_RANGE, "moduli must be positive and finite");
        // the same double-precision forms as the scene-wide constants
        mu[i] = (float)((double)E[i] / (2.0 * (1.0 + nu)));
        la[i] = (float)((double)E[i] * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float* tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, 2 * n * sizeof(float)));
    hipError_t e = hipMemcpy(tmp, mu.data(), n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(tmp + n, la.data(), n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_set_lame<<<grid_for(n, 1u << 30), 256, 0, c->stream>>>(tmp, tmp + n, c->planes[c->cur], c->pid[c->cur], c->n, c->stride);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    hipFree(tmp);
    if (e != hipSuccess) return fail(c, MPM_E_HIP, std::string("per-particle moduli: ") + hipGetErrorString(e));
    c->per_particle_lame = true;
    return MPM_OK;
}

int mpm_download_particles(mpm_ctx* c, void* aos, size_t n) {
    if (!c || !aos) return MPM_E_INVALID;
    if (c->ens) return fail(c, MPM_E_INVALID, "batched context: use mpm_ensemble_download");
    if (n != c->n) return fail(c, MPM_E_INVALID, "particle count differs from the context's n_particles");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "no particles uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    float* tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, n * MPM_PARTICLE_BYTES));
    k_download<<<grid_for(n, 1u << 30), 256, 0, c->stream>>>(c->planes[c->cur], c->pid[c->cur], c->fpid, tmp, c->n, c->stride);
    hipError_t e = hipMemcpyAsync(aos, tmp, n * MPM_PARTICLE_BYTES, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(tmp);
    if (e != hipSuccess) return fail(c, MPM_E_HIP, std::string("particle download: ") + hipGetErrorString(e));
    return MPM_OK;
}

int mpm_step(mpm_ctx* c, int method, int n_substeps) {
    if (!c) return MPM_E_INVALID;
    if (method != MPM_METHOD_EXPLICIT && method != MPM_METHOD_PBMPM) return fail(c, MPM_E_INVALID, "unknown method");
    if (n_substeps < 0) return fail(c, MPM_E_INVALID, "negative substep count");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "no particles uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    SimParams P;
    int rc = make_params(c, P);
    if (rc) return rc;
    wru<uint32_t>(c->uni, 8, method == MPM_METHOD_PBMPM ? 1u : 0u);   // writeUsePbmpm
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    for (int s = 0; s < n_substeps; ++s) {
        const bool timed = c->timing && s == n_substeps - 1;
        keys_state(c, P);
        // graphs hold the steady state only: with the fused keys pass, a substep whose keys the
        // previous one did not make (after an upload or a grid change) runs eagerly
        const bool steady = !P.fuse_keys || P.keys_ready;
        // the substeps before the timed last one; a run of GRAPH_SUBSTEPS replays the graph holding
        // them when it is captured already (mpm_prepare_graphs, or an earlier call) or when the call
        // is long enough to pay for capturing it here
        const int untimed_left = n_substeps - s - (c->timing ? 1 : 0);
        if (c->cfg.use_graph && !c->ktiming && steady && untimed_left >= GRAPH_SUBSTEPS &&
            (untimed_left >= 4 * GRAPH_SUBSTEPS || graph_valid(c, method, P, 1))) {
            if ((rc = ensure_graph(c, method, P, 1))) return rc;
            HIPCHK(c, hipGraphLaunch(c->gexec[method][c->cur][1], c->stream));
            c->substeps += GRAPH_SUBSTEPS;
            s += GRAPH_SUBSTEPS - 1;          // the parity is back where it was
            keys_after(c, P, true);
            continue;
        }
        if (c->cfg.use_graph && !timed && !c->ktiming && steady) {
            if ((rc = ensure_graph(c, method, P))) return rc;
            HIPCHK(c, hipGraphLaunch(c->gexec[method][c->cur][0], c->stream));
        } else {
            enqueue_substep(c, method, P, -1, timed);
            HIPCHK(c, hipGetLastError());
        }
        c->cur ^= 1;
        c->substeps++;
        keys_after(c, P, true);
        if (c->ktiming && c->kt_n > 512) kt_collect(c);
    }
    if (c->timing && (rc = finish_timing(c, n_substeps))) return rc;
    kt_collect(c);
    return MPM_OK;
}

int mpm_prepare_graphs(mpm_ctx* c, int method) {
    if (!c) return MPM_E_INVALID;
    if (method != MPM_METHOD_EXPLICIT && method != MPM_METHOD_PBMPM) return fail(c, MPM_E_INVALID, "unknown method");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "no particles uploaded");
    if (!c->cfg.use_graph) return MPM_OK;
    HIPCHK(c, hipSetDevice(c->device));
    SimParams P;
    int rc = make_params(c, P);
    if (rc) return rc;
    wru<uint32_t>(c->uni, 8, method == MPM_METHOD_PBMPM ? 1u : 0u);   // writeUsePbmpm, as mpm_step
    P.keys_ready = P.fuse_keys;      // the steady state: keys made by the previous substep
    for (int p = 0; p < 2; ++p) {
        for (int many = 0; many < 2 && !rc; ++many) rc = ensure_graph(c, method, P, many);
        c->cur ^= 1;
        if (rc) { if (p == 0) c->cur ^= 1; return rc; }
    }
    return MPM_OK;
}

int mpm_sync(mpm_ctx* c) {
    if (!c) return MPM_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return c->ens ? ens_check_flags(c) : MPM_OK;
}

int mpm_debug_stage(mpm_ctx* c, int method, int stage) {
    if (!c) return MPM_E_INVALID;
    if (method != MPM_METHOD_EXPLICIT && method != MPM_METHOD_PBMPM) return fail(c, MPM_E_INVALID, "unknown method");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "no particles uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (stage == MPM_STAGE_WARM_START && (method != MPM_METHOD_PBMPM || !c->learn_on))
        return fail(c, MPM_E_STATE, "the warm-start stage needs PBMPM with the learned warm start enabled");
    SimParams P;
    int rc = make_params(c, P);
    if (rc) return rc;
    keys_state(c, P);
    bool done = enqueue_substep(c, method, P, stage == MPM_STAGE_WARM_START ? stage : stage < 0 ? 1 << 30 : stage);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (done) { c->cur ^= 1; c->substeps++; }
    keys_after(c, P, done);
    return MPM_OK;
}

static int read_counters(mpm_ctx* c, uint32_t* h) {
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(h, c->counters, C_NCOUNTERS * 4, hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_dump_grid(mpm_ctx* c, int32_t* xyz, int32_t* vals, size_t cap, size_t* n_out) {
    if (!c || !n_out) return MPM_E_INVALID;
    uint32_t h[C_NCOUNTERS];
    int rc = read_counters(c, h);
    if (rc) return rc;
    const uint32_t nb = std::min(h[C_NBLOCKS], c->max_blocks);
    *n_out = (size_t)nb * 64;
    if (!xyz || !vals) return MPM_OK;
    if (cap < (size_t)nb * 64) return fail(c, MPM_E_RANGE, "dump capacity too small");
    std::vector<uint32_t> keys(nb);
    std::vector<int4> g((size_t)nb * 64);
    if (nb) {
        HIPCHK(c, hipMemcpy(keys.data(), c->block_key, nb * 4, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(g.data(), c->grid[c->last_grid], (size_t)nb * 64 * sizeof(int4), hipMemcpyDeviceToHost));
        if (c->last_pbmpm) {   // PBMPM keeps the node mass in its own array (deposited once per substep)
            std::vector<int> gm((size_t)nb * 64);
            HIPCHK(c, hipMemcpy(gm.data(), c->gmass, (size_t)nb * 64 * sizeof(int), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < gm.size(); ++i) g[i].x = gm[i];
        }
    }
    SimParams P;
    if ((rc = make_params(c, P))) return rc;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t k = keys[b];
        const int bx = (int)(k % (uint32_t)P.nb[0]), by = (int)((k / (uint32_t)P.nb[0]) % (uint32_t)P.nb[1]), bz = (int)(k / ((uint32_t)P.nb[0] * (uint32_t)P.nb[1]));
        for (uint32_t l = 0; l < 64; ++l) {
            const size_t i = (size_t)b * 64 + l;
            xyz[3 * i] = 4 * bx + (int)(l & 3u); xyz[3 * i + 1] = 4 * by + (int)((l >> 2) & 3u); xyz[3 * i + 2] = 4 * bz + (int)(l >> 4);
            vals[4 * i] = g[i].x; vals[4 * i + 1] = g[i].y; vals[4 * i + 2] = g[i].z; vals[4 * i + 3] = g[i].w;
        }
    }
    return MPM_OK;
}

int mpm_dump_blocks(mpm_ctx* c, int32_t* xyz, size_t cap, size_t* n_out) {
    if (!c || !n_out) return MPM_E_INVALID;
    uint32_t h[C_NCOUNTERS];
    int rc = read_counters(c, h);
    if (rc) return rc;
    const uint32_t nb = std::min(h[C_NBLOCKS], c->max_blocks);
    *n_out = nb;
    if (!xyz) return MPM_OK;
    if (cap < nb) return fail(c, MPM_E_RANGE, "dump capacity too small");
    std::vector<uint32_t> keys(nb);
    if (nb) HIPCHK(c, hipMemcpy(keys.data(), c->block_key, nb * 4, hipMemcpyDeviceToHost));
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t k = keys[b];
        xyz[3 * b] = (int)(k % c->nb[0]); xyz[3 * b + 1] = (int)((k / c->nb[0]) % c->nb[1]); xyz[3 * b + 2] = (int)(k / (c->nb[0] * c->nb[1]));
    }
    return MPM_OK;
}

int mpm_dump_particle_blocks(mpm_ctx* c, int32_t* xyz, uint32_t* counts, size_t cap, size_t* n_out) {
    if (!c || !n_out) return MPM_E_INVALID;
    uint32_t h[C_NCOUNTERS];
    int rc = read_counters(c, h);
    if (rc) return rc;
    const uint32_t nr = h[C_NRUNS];
    *n_out = nr;
    if (!xyz || !counts) return MPM_OK;
    if (cap < nr) return fail(c, MPM_E_RANGE, "dump capacity too small");
    std::vector<uint32_t> keys(nr), cnt(nr);
    if (nr) {
        HIPCHK(c, hipMemcpy(keys.data(), c->run_key, nr * 4, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(cnt.data(), c->run_count, nr * 4, hipMemcpyDeviceToHost));
    }
    for (uint32_t r = 0; r < nr; ++r) {
        const uint32_t k = keys[r];
        xyz[3 * r] = (int)(k % c->nb[0]); xyz[3 * r + 1] = (int)((k / c->nb[0]) % c->nb[1]); xyz[3 * r + 2] = (int)(k / (c->nb[0] * c->nb[1]));
        counts[r] = cnt[r];
    }
    return MPM_OK;
}

int mpm_get_stats(mpm_ctx* c, mpm_stats* out) {
    if (!c || !out) return MPM_E_INVALID;
    uint32_t h[C_NCOUNTERS];
    int rc = read_counters(c, h);
    if (rc) return rc;
    std::memset(out, 0, sizeof(*out));
    // a particle block (run) whose own block was not allocated (dropped beyond max_blocks, or its
    // table insert exhausted max_probes) is not transferred: its particles count as unbinned
    uint64_t orphans = 0;
    const uint32_t nr = h[C_NRUNS];
    if (nr && nr <= c->max_runs) {
        std::vector<uint32_t> cnt(nr), nbr(27 * (size_t)nr);
        HIPCHK(c, hipMemcpy(cnt.data(), c->run_count, nr * 4, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(nbr.data(), c->nbr, nbr.size() * 4, hipMemcpyDeviceToHost));
        for (uint32_t r = 0; r < nr; ++r) if (nbr[27 * (size_t)r + 13] == EMPTY) orphans += cnt[r];
    }
    out->n_particles = c->n;
    out->particles_binned = h[C_NVALID] - orphans;
    out->particles_unbinned = c->n - out->particles_binned;
    out->particle_blocks = h[C_NRUNS];
    out->work_chunks = h[C_NCHUNKS];
    out->blocks_allocated = std::min(h[C_NBLOCKS], c->max_blocks);
    out->blocks_dropped = h[C_DROPS];
    out->probe_failures = h[C_PROBEFAIL];
    out->substeps = c->substeps;
    out->mlp_fallbacks = h[C_FALLBACK];
    out->nonfinite_particles = h[C_NONFIN_LAST];
    out->svd_fp64_particles = h[C_SVD64];
    out->collider_queued = h[C_NCOLL];
    return MPM_OK;
}

int mpm_divisor_fast(mpm_ctx* c, float d, float scale, int* fast) {
    if (!c || !fast) return MPM_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    uint32_t ok = 0;
    int rc = verify_divisor(c, d, scale, 0, &ok);
    *fast = (int)ok;
    return rc;
}

int mpm_selftest(int kind, const float* a, const float* b, float* out, size_t n) {
    if (kind == 11) {
        // LDS atomic throughput: n iterations of a conflict-free ds_add_u32 per wave, 8 workgroups
        // of 256 per CU; out = {elapsed ms, wave-instructions, CUs}
        if (!out || n == 0 || n > (1u << 20)) return MPM_E_INVALID;
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return MPM_E_HIP;
        const uint32_t blocks = (uint32_t)ncu * 8u;
        uint32_t* dout = nullptr;
        if (hipMalloc(&dout, blocks * 4) != hipSuccess) return MPM_E_NOMEM;
        hipEvent_t e0, e1;
        hipEventCreate(&e0); hipEventCreate(&e1);
        float ms = 0.0f;
        hipError_t e = hipSuccess;
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {   // the second launch is the one timed
            hipEventRecord(e0, 0);
            k_lds_rate<<<blocks, 256, 0, 0>>>((uint32_t)n, dout);
            hipEventRecord(e1, 0);
            e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        }
        hipEventDestroy(e0); hipEventDestroy(e1); hipFree(dout);
        out[0] = ms; out[1] = (float)((double)blocks * 4.0 * (double)n); out[2] = (float)ncu;
        return e == hipSuccess ? MPM_OK : MPM_E_HIP;
    }
    if (kind < 0 || kind > 10 || !a || !out || n == 0 || n > (1u << 25)) return MPM_E_INVALID;
    if ((kind == 2 || kind == 4 || kind == 5) && !b) return MPM_E_INVALID;
    const size_t in_w = (kind >= 6 && kind <= 8) ? 9 : 1;
    const size_t out_w = (kind == 7 || kind == 8) ? 9 : 1;
    float *da = nullptr, *db = nullptr, *dout = nullptr;
    if (hipMalloc(&da, n * in_w * 4) != hipSuccess) return MPM_E_NOMEM;
    if (b && hipMalloc(&db, n * 4) != hipSuccess) { hipFree(da); return MPM_E_NOMEM; }
    if (hipMalloc(&dout, n * out_w * 4) != hipSuccess) { hipFree(da); hipFree(db); return MPM_E_NOMEM; }
    hipError_t e = hipMemcpy(da, a, n * in_w * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && b) e = hipMemcpy(db, b, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_selftest<<<(uint32_t)((n + 255) / 256), 256>>>(kind, da, db, dout, (uint32_t)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, n * out_w * 4, hipMemcpyDeviceToHost);
    hipFree(da); if (db) hipFree(db); hipFree(dout);
    return e == hipSuccess ? MPM_OK : MPM_E_HIP;
}


// ------------------------------------------------------------------ learned warm start / residuals
int mpm_set_iterations(mpm_ctx* c, int k, float force_divisor) {
    if (!c) return MPM_E_INVALID;
    if (k < 1 || k > 64 || !(force_divisor > 0.0f)) return fail(c, MPM_E_RANGE, "iterations must be in 1..64, divisor > 0");
    c->k_cur = k;
    c->force_div_cur = force_divisor;
    return MPM_OK;
}

static int ensure_mlp(mpm_ctx* c) {
    if (c->mlp_params) return MPM_OK;
    int rc;
    if ((rc = dalloc(c, &c->mlp_params, MLP_NPARAM)) || (rc = dalloc(c, &c->mlp_grads, MLP_NPARAM)) ||
        (rc = dalloc(c, &c->mlp_m1, MLP_NPARAM)) || (rc = dalloc(c, &c->mlp_m2, MLP_NPARAM)) ||
        (rc = dalloc(c, &c->mlp_norm, 64)) || (rc = dalloc(c, &c->mlp_wbf, 16 * 64 + 64 * 64 + 64 * 16)) ||
        (rc = dalloc(c, &c->mlp_colsum, 1024 * 64)) || (rc = dalloc(c, &c->mlp_loss, 1)))
        return rc;
    HIPCHK(c, hipMemsetAsync(c->mlp_grads, 0, MLP_NPARAM * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mlp_m1, 0, MLP_NPARAM * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mlp_m2, 0, MLP_NPARAM * 4, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MPM_OK;
}

// Standard-normal initialisation scaled by 1/sqrt(fan_in) from a seeded host generator.
static void init_params(std::vector<float>& w, uint32_t seed) {
    w.assign(MLP_NPARAM, 0.0f);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
    auto uni = [&]() { st = st * 6364136223846793005ull + 1442695040888963407ull; return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
    auto nrm = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
    for (int f = 0; f < MLP_NIN; ++f) for (int h = 0; h < 64; ++h) w[MLP_W1 + f * 64 + h] = (float)(nrm() / std::sqrt((double)MLP_NIN));
    for (int a = 0; a < 64; ++a) for (int b = 0; b < 64; ++b) w[MLP_W2 + a * 64 + b] = (float)(nrm() / 8.0);
    for (int a = 0; a < 64; ++a) for (int o = 0; o < MLP_NOUT; ++o) w[MLP_W3 + a * 16 + o] = (float)(nrm() / 8.0 * 0.1);
}

// The per-iteration dx planes (resid_dx, and the residual maxima) and the warm start's input
// planes (dprev, zeroed: no particle has been iterated with them yet).
int ensure_learn_buffers(mpm_ctx* c) {
    int rc;
    if (!c->resid_dx) {
        if ((rc = dalloc(c, &c->resid_dx, 3 * (size_t)c->stride)) || (rc = dalloc(c, &c->resid_max, 64 * RESID_SLOTS)) ||
            (rc = dalloc(c, &c->resid_peak, 64)) || (rc = dalloc(c, &c->resid_sum, 64)))
            return rc;
        HIPCHK(c, hipMemsetAsync(c->resid_max, 0, 64 * RESID_SLOTS * 4, c->stream));
        HIPCHK(c, hipMemsetAsync(c->resid_peak, 0, 64 * 4, c->stream));
        HIPCHK(c, hipMemsetAsync(c->resid_sum, 0, 64 * 4, c->stream));
    }
    if (!c->learn_dprev) {
        if ((rc = dalloc(c, &c->learn_dprev, 3 * (size_t)c->stride))) return rc;
        HIPCHK(c, hipMemsetAsync(c->learn_dprev, 0, 3 * (size_t)c->stride * 4, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MPM_OK;
}

int mpm_learn_collect(mpm_ctx* c, int n_substeps, int k_run, int k_ref, int every, uint64_t max_pairs, uint64_t* n_pairs) {
    if (!c || n_substeps < 0 || k_run < 1 || k_run > 64 || k_ref < 1 || k_ref > 64 || every < 1) return MPM_E_INVALID;
    if (!c->has_particles) return fail(c, MPM_E_STATE, "no particles uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (max_pairs > c->ds_cap) {
        if (c->ds_n) return fail(c, MPM_E_STATE, "dataset capacity is fixed by the first collection");
        int rc;
        if ((rc = dalloc(c, &c->ds_feats, max_pairs * 16)) || (rc = dalloc(c, &c->ds_targs, max_pairs * 16))) return rc;
        c->ds_cap = max_pairs;
    }
    int rc;
    if ((rc = ensure_learn_buffers(c))) return rc;
    if (k_run < 2 || k_ref < 2) return fail(c, MPM_E_INVALID, "collection needs k_run, k_ref >= 2 (the features use the last iteration's change)");
    const int k_save = c->k_cur; const float fd_save = c->force_div_cur; const bool learn_save = c->learn_on;
    c->learn_on = false;
    c->collecting = true;
    SimParams Prun, Pref;
    c->k_cur = k_run; c->force_div_cur = (float)k_run;
    rc = make_params(c, Prun);
    c->k_cur = k_ref; c->force_div_cur = (float)k_ref;
    if (rc == MPM_OK) rc = make_params(c, Pref);
    if (rc == MPM_OK) {
        // what the iterations' gravity shares add to dx over the substep (k_ref x grav_div dt^2)
        const float gdt2[3] = {Pref.grav_div[0] * (Pref.dt * Pref.dt) * (float)k_ref, Pref.grav_div[1] * (Pref.dt * Pref.dt) * (float)k_ref,
                               Pref.grav_div[2] * (Pref.dt * Pref.dt) * (float)k_ref};
        for (int s = 0; s < n_substeps; ++s) {
            // substep 0 only produces dprev (the features' last-iteration change) for the ones after it
            const bool log = s > 0 && (s % every == 0) && c->ds_n + c->n <= c->ds_cap;
            if (log) {
                Bufs B = make_bufs(c);
                k_mlp_log_features<<<grid_for(c->n, 1u << 30), 256, 0, c->stream>>>(B, c->n, c->ds_feats, c->ds_targs, c->ds_n);
                if (k_ref != k_run) {
                    // the target substep: k_ref iterations from the same start state into the other
                    // planes (a substep reads `cur` and writes only the other set), then the logged
                    // targets from there; the trajectory's own substep below overwrites them
                    keys_state(c, Pref);
                    enqueue_substep(c, MPM_METHOD_PBMPM, Pref, -1, false);
                    c->cur ^= 1;
                    B = make_bufs(c);
                    c->cur ^= 1;
                    keys_after(c, Pref, false);      // its keys are of rows the trajectory discards
                    k_mlp_log_targets<<<grid_for(c->n, 1u << 30), 256, 0, c->stream>>>(B, c->n, c->ds_feats, c->ds_targs, c->ds_n,
                                                                                     gdt2[0], gdt2[1], gdt2[2]);
                }
            }
            SimParams& Pt = k_ref != k_run ? Prun : Pref;
            keys_state(c, Pt);
            enqueue_substep(c, MPM_METHOD_PBMPM, Pt, -1, false);
            c->cur ^= 1; c->substeps++;
            keys_after(c, Pt, true);
            if (log) {
                if (k_ref == k_run) {
                    Bufs B = make_bufs(c);
                    k_mlp_log_targets<<<grid_for(c->n, 1u << 30), 256, 0, c->stream>>>(B, c->n, c->ds_feats, c->ds_targs, c->ds_n,
                                                                                     gdt2[0], gdt2[1], gdt2[2]);
                }
                c->ds_n += c->n;
            }
        }
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(c, MPM_E_HIP, std::string("collect: ") + hipGetErrorString(e));
    }
    c->collecting = false;
    c->k_cur = k_save; c->force_div_cur = fd_save; c->learn_on = learn_save;
    if (n_pairs) *n_pairs = c->ds_n;
    return rc;
}

int mpm_learn_train(mpm_ctx* c, int steps, int batch, float lr, uint32_t seed, float* mean_loss) {
    if (!c || steps < 0 || batch < 256 || batch % 256 || !(lr > 0.0f)) return MPM_E_INVALID;
    if (c->ds_n < 16) return fail(c, MPM_E_STATE, "no training pairs collected");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ensure_mlp(c))) return rc;
    if (!c->mlp_ready) {
        // standardisation from the collected pairs, then initial weights
        k_mlp_colstats<<<1024, 256, 0, c->stream>>>(c->ds_feats, c->ds_targs, c->ds_n, c->mlp_colsum);
        std::vector<float> partial(1024 * 64);
        HIPCHK(c, hipMemcpyAsync(partial.data(), c->mlp_colsum, partial.size() * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double cs[64] = {};
        for (size_t b = 0; b < 1024; ++b) for (int k = 0; k < 64; ++k) cs[k] += partial[b * 64 + k];
        float norm[64];
        for (int k = 0; k < 16; ++k) {
            const double n = (double)c->ds_n;
            const double mx = cs[k] / n, vx = std::max(cs[16 + k] / n - mx * mx, 0.0);
            const double mt = cs[32 + k] / n, vt = std::max(cs[48 + k] / n - mt * mt, 0.0);
            const bool fx = k < MLP_NIN, ft = k < MLP_NOUT;
            norm[k] = fx ? (float)mx : 0.0f;
            norm[16 + k] = fx ? (float)std::max(std::sqrt(vx), 1e-12) : 1.0f;
            norm[32 + k] = ft ? (float)mt : 0.0f;
            norm[48 + k] = ft ? (float)std::max(std::sqrt(vt), 1e-20) : 1.0f;
        }
        HIPCHK(c, hipMemcpy(c->mlp_norm, norm, sizeof(norm), hipMemcpyHostToDevice));
        std::vector<float> w;
        init_params(w, seed);
        HIPCHK(c, hipMemcpy(c->mlp_params, w.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice));
        k_mlp_refresh_bf16<<<(MLP_NPARAM + 255) / 256, 256, 0, c->stream>>>(c->mlp_params, c->mlp_wbf);
        c->adam_t = 0;
        c->mlp_ready = true;
    }
    const uint32_t tiles = 4, nwg = (uint32_t)batch / (4u * 16u * tiles);
    MlpTrainArgs A;
    A.feats = c->ds_feats; A.targs = c->ds_targs; A.n = c->ds_n; A.tiles = tiles; A.inv_batch = 1.0f / (float)batch;
    A.loss_sum = c->mlp_loss;
    HIPCHK(c, hipMemsetAsync(c->mlp_loss, 0, 4, c->stream));
    const int tail = std::min(steps, 50);
    for (int it = 0; it < steps; ++it) {
        if (it == steps - tail) HIPCHK(c, hipMemsetAsync(c->mlp_loss, 0, 4, c->stream));
        A.seed = hash_host(seed * 2654435761u + (uint32_t)c->adam_t);
        k_mlp_train<<<nwg, 256, 0, c->stream>>>(mlp_bufs(c), A);
        c->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - (float)std::pow((double)b1, (double)c->adam_t), c2 = 1.0f - (float)std::pow((double)b2, (double)c->adam_t);
        k_mlp_adam<<<(MLP_NPARAM + 255) / 256, 256, 0, c->stream>>>(c->mlp_params, c->mlp_grads, c->mlp_m1, c->mlp_m2, c->mlp_wbf,
                                                                    lr, b1, b2, 1e-8f, c1, c2);
    }
    float loss = 0.0f;
    HIPCHK(c, hipMemcpyAsync(&loss, c->mlp_loss, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (mean_loss) *mean_loss = tail ? loss / ((float)tail * (float)batch * (float)MLP_NOUT) : 0.0f;
    return MPM_OK;
}

int mpm_learn_get_params(mpm_ctx* c, float* params, float* norm) {
    if (!c || !params || !norm) return MPM_E_INVALID;
    if (!c->mlp_ready) return fail(c, MPM_E_STATE, "no trained network");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(params, c->mlp_params, MLP_NPARAM * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(norm, c->mlp_norm, 64 * 4, hipMemcpyDeviceToHost));
    return MPM_OK;
}
int mpm_learn_set_params(mpm_ctx* c, const float* params, const float* norm) {
    if (!c || !params || !norm) return MPM_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ensure_mlp(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->mlp_params, params, MLP_NPARAM * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->mlp_norm, norm, 64 * 4, hipMemcpyHostToDevice));
    k_mlp_refresh_bf16<<<(MLP_NPARAM + 255) / 256, 256, 0, c->stream>>>(c->mlp_params, c->mlp_wbf);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->mlp_ready = true;
    c->adam_t = 0;
    HIPCHK(c, hipMemsetAsync(c->mlp_m1, 0, MLP_NPARAM * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mlp_m2, 0, MLP_NPARAM * 4, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MPM_OK;
}
// The warm start's per-particle input dprev, by particle id (upload order): set from in_xyz and/or
// read into out_xyz (3 floats per particle each).
int mpm_learn_dprev(mpm_ctx* c, const float* in_xyz, float* out_xyz, size_t n) {
    if (!c || (!in_xyz && !out_xyz)) return MPM_E_INVALID;
    if (c->ens) return fail(c, MPM_E_INVALID, "batched context: the warm start is for single-scene contexts");
    if (n != c->n) return fail(c, MPM_E_INVALID, "one dprev per particle");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc;
    if ((rc = ensure_learn_buffers(c))) return rc;
    std::vector<uint32_t> pid(c->n);
    std::vector<float> planes(3 * (size_t)c->stride);
    HIPCHK(c, hipMemcpy(pid.data(), c->pid[c->cur], c->n * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(planes.data(), c->learn_dprev, planes.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t r = 0; r < c->n; ++r) {
        const uint32_t id = pid[r];
        if (id >= c->n) continue;
        for (int a = 0; a < 3; ++a) {
            if (out_xyz) out_xyz[3 * (size_t)id + a] = planes[(size_t)a * c->stride + r];
            if (in_xyz) planes[(size_t)a * c->stride + r] = in_xyz[3 * (size_t)id + a];
        }
    }
    if (in_xyz) HIPCHK(c, hipMemcpy(c->learn_dprev, planes.data(), planes.size() * 4, hipMemcpyHostToDevice));
    return MPM_OK;
}
int mpm_learn_enable(mpm_ctx* c, int on, float limit) {
    if (!c || !(limit > 0.0f)) return MPM_E_INVALID;
    if (on && !c->mlp_ready) return fail(c, MPM_E_STATE, "no trained network");
    if (on) {
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int rc;
        if ((rc = ensure_learn_buffers(c))) return rc;
        // the first warm-started substep has no previous iteration change: every particle keeps
        // plain PBMPM there (and counts as a fallback); the ones after use the network
        if (!c->learn_on) {
            HIPCHK(c, hipMemsetAsync(c->learn_dprev, 0, 3 * (size_t)c->stride * 4, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
    }
    c->learn_on = on != 0;
    c->learn_limit = limit;
    return MPM_OK;
}

int mpm_set_residual_tracking(mpm_ctx* c, int on) {
    if (!c) return MPM_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    if (on && !c->resid_dx) {
        int rc;
        if ((rc = dalloc(c, &c->resid_dx, 3 * (size_t)c->stride)) || (rc = dalloc(c, &c->resid_max, 64 * RESID_SLOTS)) ||
            (rc = dalloc(c, &c->resid_peak, 64)) || (rc = dalloc(c, &c->resid_sum, 64)))
            return rc;
    }
    if (c->resid_dx) {
        // on the context's stream (non-blocking: a null-stream memset would not order with it)
        HIPCHK(c, hipMemsetAsync(c->resid_max, 0, 64 * RESID_SLOTS * 4, c->stream));
        HIPCHK(c, hipMemsetAsync(c->resid_peak, 0, 64 * 4, c->stream));
        HIPCHK(c, hipMemsetAsync(c->resid_sum, 0, 64 * 4, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->resid_on = on != 0;
    c->resid_substeps = c->substeps;
    return MPM_OK;
}
int mpm_get_residuals(mpm_ctx* c, float* mean, float* peak, int cap, uint64_t* substeps) {
    if (!c || !mean || !peak || cap < 1) return MPM_E_INVALID;
    if (!c->resid_dx) return fail(c, MPM_E_STATE, "residual tracking was never enabled");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float sum[64]; uint32_t pk[64];
    HIPCHK(c, hipMemcpy(sum, c->resid_sum, sizeof(sum), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(pk, c->resid_peak, sizeof(pk), hipMemcpyDeviceToHost));
    const uint64_t ns = c->substeps - c->resid_substeps;
    for (int i = 0; i < cap && i < 64; ++i) { mean[i] = ns ? sum[i] / (float)ns : 0.0f; std::memcpy(&peak[i], &pk[i], 4); }
    if (substeps) *substeps = ns;
    return MPM_OK;
}

int mpm_mlp_grad(const float* params, const float* norm, const float* feats, const float* targs, uint32_t n,
                 float* grads, float* y, float* loss) {
    if (!params || !norm || !feats || !targs || !grads || !y || !loss || n == 0 || n % 16 || n > (1u << 20)) return MPM_E_INVALID;
    float *dp = nullptr, *dg = nullptr, *dn = nullptr, *df = nullptr, *dt = nullptr, *dy = nullptr, *dl = nullptr;
    short* dw = nullptr;
    hipError_t e = hipSuccess;
    auto al = [&](void** p, size_t b) { if (e == hipSuccess) e = hipMalloc(p, b); };
    al((void**)&dp, MLP_NPARAM * 4); al((void**)&dg, MLP_NPARAM * 4); al((void**)&dn, 64 * 4);
    al((void**)&df, (size_t)n * 64); al((void**)&dt, (size_t)n * 64); al((void**)&dy, (size_t)n * 64); al((void**)&dl, 4);
    al((void**)&dw, (16 * 64 + 64 * 64 + 64 * 16) * 2);
    if (e == hipSuccess) e = hipMemcpy(dp, params, MLP_NPARAM * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn, norm, 64 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(df, feats, (size_t)n * 64, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, targs, (size_t)n * 64, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dg, 0, MLP_NPARAM * 4);
    if (e == hipSuccess) e = hipMemset(dl, 0, 4);
    if (e == hipSuccess) {
        k_mlp_refresh_bf16<<<(MLP_NPARAM + 255) / 256, 256>>>(dp, dw);
        MlpBufs M; M.params = dp; M.wbf = dw; M.grads = dg; M.norm = dn;
        k_mlp_grad_fixed<<<(n / 16 + 3) / 4, 256>>>(M, df, dt, n, dy, dl);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(grads, dg, MLP_NPARAM * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(y, dy, (size_t)n * 64, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(loss, dl, 4, hipMemcpyDeviceToHost);
    hipFree(dp); hipFree(dg); hipFree(dn); hipFree(df); hipFree(dt); hipFree(dy); hipFree(dl); hipFree(dw);
    return e == hipSuccess ? MPM_OK : MPM_E_HIP;
}


// ------------------------------------------------------------------ batched scenes (C ABI)
int mpm_ensemble_create(const mpm_config* cfgs, const int* methods, int n_scenes, mpm_ctx** out) {
    if (!cfgs || !methods || !out || n_scenes < 1 || n_scenes > (int)ENS_MAX_SCENES) return MPM_E_INVALID;
    *out = nullptr;
    uint64_t n = 0, K = 0;
    uint32_t R = 0;
    for (int s = 0; s < n_scenes; ++s) {
        const mpm_config& g = cfgs[s];
        if (g.n_particles == 0 || (methods[s] != MPM_METHOD_EXPLICIT && methods[s] != MPM_METHOD_PBMPM)) return MPM_E_INVALID;
        for (int a = 0; a < 3; ++a) if (g.grid_res[a] == 0 || g.grid_res[a] > 65536 || !(g.grid_max[a] > g.grid_min[a])) return MPM_E_INVALID;
        if (g.pbmpm_iterations < 1 || g.pbmpm_iterations > 64) return MPM_E_INVALID;
        // the chunk list is sized for chunks of at least 256 particles (and the tuning's range)
        if (g.chunk_particles && (g.chunk_particles < 256 || g.chunk_particles > 65536)) return MPM_E_RANGE;
        if (g.plasticity_mode != 0) return MPM_E_INVALID;   // the SVD pass is for single-scene contexts
        n += g.n_particles;
        K += (uint64_t)((g.grid_res[0] + 3u) / 4u) * ((g.grid_res[1] + 3u) / 4u) * ((g.grid_res[2] + 3u) / 4u);
        R = std::max<uint32_t>(R, methods[s] == MPM_METHOD_PBMPM ? (uint32_t)g.pbmpm_iterations + 1u : 2u);
    }
    if (n > MPM_MAX_PARTICLES || K >= 0xFFFFFF00ull || (uint64_t)R * n_scenes > ENS_MAX_SEG) return MPM_E_RANGE;
    mpm_ctx* c = new mpm_ctx();
    c->cfg = cfgs[0];
    c->device = cfgs[0].device;
    if (hipSetDevice(c->device) != hipSuccess) { (void)hipGetLastError(); delete c; return MPM_E_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { (void)hipGetLastError(); delete c; return MPM_E_HIP; }
    c->ens = (uint32_t)n_scenes;
    c->ens_R = R;
    c->ens_cfg.assign(cfgs, cfgs + n_scenes);
    c->ens_method.assign(methods, methods + n_scenes);
    c->n = (uint32_t)n;
    c->stride = (c->n + 63u) & ~63u;
    c->K = (uint32_t)K;
    c->chunk = cfgs[0].chunk_particles ? cfgs[0].chunk_particles : 1024u;   // the autotuner's usual choice
    c->max_blocks = (uint32_t)K;
    c->cap = cfgs[0].hash_map_size ? cfgs[0].hash_map_size
                                   : next_prime((uint32_t)std::min<uint64_t>(2ull * std::min<uint64_t>(K, 8ull * n) + 1, 0xFFFFFF00ull));
    c->direct_blocks = cfgs[0].hash_map_size == 0;
    c->max_runs = (uint32_t)std::min<uint64_t>(n, K);
    c->max_chunks = (c->n + 255u) / 256u + c->max_runs;      // room for chunks down to 256 particles
    set_scan_tiles(c, K);
    int rc = alloc_buffers(c);
    if (!rc) {
        const size_t S = (size_t)n_scenes;
        if ((rc = dalloc(c, &c->d_sp, S)) || (rc = dalloc(c, &c->d_part_base, S + 1)) || (rc = dalloc(c, &c->d_key_base, S + 1)) ||
            (rc = dalloc(c, &c->d_nphase, S)) || (rc = dalloc(c, &c->d_method, S)) || (rc = dalloc(c, &c->d_nblk, S)) ||
            (rc = dalloc(c, &c->d_rbase, S)) ||
            (rc = dalloc(c, &c->d_cbeg, S + 1)) || (rc = dalloc(c, &c->d_seg, (size_t)R * S + 1)) ||
            (rc = dalloc(c, &c->d_done, (size_t)R * S)) || (rc = dalloc(c, &c->d_ctl, ENS_CTL_PHASE + 128))) {}
    }
    if (!rc) {
        // per-scene parameters from each configuration (no collider, no pointer interaction)
        uint32_t pb = 0, kb = 0;
        c->ens_sp.resize(n_scenes);
        for (int s = 0; s < n_scenes && !rc; ++s) {
            const mpm_config& g = cfgs[s];
            uint8_t uni[MPM_UNIFORMS_BYTES];
            std::memset(uni, 0, sizeof(uni));
            wru<float>(uni, 0, g.dt);
            wru<float>(uni, 4, g.fixed_point_scale);
            for (int a = 0; a < 3; ++a) {
                wru<float>(uni, 16 + 4 * a, g.grid_min[a]);
                wru<float>(uni, 32 + 4 * a, g.grid_max[a]);
                wru<uint32_t>(uni, 208 + 4 * a, g.grid_res[a]);
                wru<float>(uni, 352 + 4 * a, (float)(((double)g.grid_max[a] - (double)g.grid_min[a]) / (double)g.grid_res[a]));
            }
            SimParams& P = c->ens_sp[s];
            rc = params_from(c, uni, g, g.pbmpm_iterations, g.pbmpm_force_divisor, true, P);
            if (rc) break;
            P.n = g.n_particles;
            P.chunk = c->chunk;
            P.cap = c->cap;
            P.max_probes = c->cfg.max_probes ? c->cfg.max_probes : 100u;
            P.kbase = kb;
            P.scene = (uint32_t)s;
            P.blk_base = kb;            // a scene's dense block range is its key range
            P.blk_cap = P.K;
            P.max_blocks = P.K;
            c->ens_part_base.push_back(pb);
            c->ens_key_base.push_back(kb);
            c->ens_nphase.push_back(methods[s] == MPM_METHOD_PBMPM ? (uint32_t)g.pbmpm_iterations + 1u : 2u);
            pb += g.n_particles;
            kb += P.K;
        }
        c->ens_part_base.push_back(pb);
        c->ens_key_base.push_back(kb);
    }
    if (!rc) {
        std::memset(&c->ens_P0, 0, sizeof(SimParams));
        c->ens_P0.n = c->n; c->ens_P0.K = c->K; c->ens_P0.chunk = c->chunk; c->ens_P0.cap = c->cap;
        c->ens_P0.max_blocks = c->max_blocks; c->ens_P0.max_probes = c->cfg.max_probes ? c->cfg.max_probes : 100u;
        c->ens_P0.direct_blocks = c->direct_blocks ? 1u : 0u;
        c->ens_P0.block_identity = 1u;     // a scene's dense block range is its key range

        std::vector<uint32_t> meth(n_scenes);
        for (int s = 0; s < n_scenes; ++s) meth[s] = (uint32_t)methods[s];
        hipError_t e = hipMemcpy(c->d_sp, c->ens_sp.data(), n_scenes * sizeof(SimParams), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->d_part_base, c->ens_part_base.data(), (n_scenes + 1) * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->d_key_base, c->ens_key_base.data(), (n_scenes + 1) * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->d_nphase, c->ens_nphase.data(), n_scenes * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->d_method, meth.data(), n_scenes * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemsetAsync(c->d_ctl, 0, (ENS_CTL_PHASE + 128) * 4, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = MPM_E_HIP;
    }
    if (rc) { release(c); delete c; return rc; }
    c->ens_loaded.assign(n_scenes, false);
    c->ens_wg_per_cu = cfgs[0].grid_workgroups ? std::max<uint32_t>(1u, cfgs[0].grid_workgroups / (uint32_t)c->n_cu) : 4u;
    *out = c;
    return MPM_OK;
}

int mpm_ensemble_upload(mpm_ctx* c, int scene, const void* aos, size_t n) {
    if (!c || !aos) return MPM_E_INVALID;
    if (!c->ens || scene < 0 || scene >= (int)c->ens) return fail(c, MPM_E_INVALID, "not a batched context, or scene out of range");
    const uint32_t base = c->ens_part_base[scene], cnt = c->ens_part_base[scene + 1] - base;
    if (n != cnt) return fail(c, MPM_E_INVALID, "particle count differs from the scene's n_particles");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float* tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, n * MPM_PARTICLE_BYTES));
    hipError_t e = hipMemcpyAsync(tmp, aos, n * MPM_PARTICLE_BYTES, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        k_upload_range<<<grid_for(c->n, 1u << 30), 256, 0, c->stream>>>(tmp, c->planes[c->cur], c->pid[c->cur], c->fpid, c->n, c->stride, base, cnt);
        e = hipStreamSynchronize(c->stream);
    }
    hipFree(tmp);
    if (e != hipSuccess) return fail(c, MPM_E_HIP, std::string("scene upload: ") + hipGetErrorString(e));
    c->ens_loaded[scene] = true;
    c->has_particles = std::all_of(c->ens_loaded.begin(), c->ens_loaded.end(), [](bool b) { return b; });
    return MPM_OK;
}

int mpm_ensemble_download(mpm_ctx* c, int scene, void* aos, size_t n) {
    if (!c || !aos) return MPM_E_INVALID;
    if (!c->ens || scene < 0 || scene >= (int)c->ens) return fail(c, MPM_E_INVALID, "not a batched context, or scene out of range");
    const uint32_t base = c->ens_part_base[scene], cnt = c->ens_part_base[scene + 1] - base;
    if (n != cnt) return fail(c, MPM_E_INVALID, "particle count differs from the scene's n_particles");
    HIPCHK(c, hipSetDevice(c->device));
    float* tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, n * MPM_PARTICLE_BYTES));
    k_download_range<<<grid_for(c->n, 1u << 30), 256, 0, c->stream>>>(c->planes[c->cur], c->pid[c->cur], c->fpid, tmp, c->n, c->stride, base, cnt);
    hipError_t e = hipMemcpyAsync(aos, tmp, n * MPM_PARTICLE_BYTES, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(tmp);
    if (e != hipSuccess) return fail(c, MPM_E_HIP, std::string("scene download: ") + hipGetErrorString(e));
    return ens_check_flags(c);
}

int mpm_ensemble_step(mpm_ctx* c, int n_substeps) {
    if (!c) return MPM_E_INVALID;
    if (!c->ens) return fail(c, MPM_E_INVALID, "not a batched context");
    if (n_substeps < 0) return fail(c, MPM_E_INVALID, "negative substep count");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "every scene needs its particles uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    for (int s = 0; s < n_substeps; ++s) {
        const bool timed = c->timing && s == n_substeps - 1;
        if (c->cfg.use_graph && !timed && !c->ktiming) {
            if ((rc = ens_graph(c))) return rc;
            HIPCHK(c, hipGraphLaunch(c->gexec[0][c->cur][0], c->stream));
        } else {
            enqueue_ens_substep(c, timed);
            HIPCHK(c, hipGetLastError());
        }
        c->cur ^= 1;
        c->substeps++;
        if (c->ktiming && c->kt_n > 512) kt_collect(c);
    }
    if (c->timing && (rc = finish_timing(c, n_substeps))) return rc;
    kt_collect(c);
    return MPM_OK;
}

int mpm_ensemble_set_tuning(mpm_ctx* c, const mpm_tuning* t) {
    if (!c || !t) return MPM_E_INVALID;
    if (!c->ens) return fail(c, MPM_E_INVALID, "not a batched context");
    if (t->chunk_particles < 256 || t->chunk_particles > 65536 || t->workgroups_per_cu < 1 || t->workgroups_per_cu > 8)
        return fail(c, MPM_E_RANGE, "chunk_particles must be in 256..65536 and workgroups_per_cu in 1..8");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc;
    if ((rc = ens_set_chunk(c, t->chunk_particles))) return rc;
    c->ens_wg_per_cu = t->workgroups_per_cu;
    c->ens_persistent = t->persistent ? 1 : 0;
    c->ens_graph_ok[0] = c->ens_graph_ok[1] = false;
    return MPM_OK;
}
int mpm_ensemble_get_tuning(mpm_ctx* c, mpm_tuning* t) {
    if (!c || !t) return MPM_E_INVALID;
    if (!c->ens) return fail(c, MPM_E_INVALID, "not a batched context");
    t->chunk_particles = c->chunk;
    t->workgroups_per_cu = c->ens_wg_per_cu;
    t->persistent = c->ens_persistent;
    t->ms_per_substep = 0.0f;
    return MPM_OK;
}

// Times every candidate on the live scenes (state restored after each trial and at the end) and
// keeps the fastest. out receives up to cap candidates with their times; the chosen one is set.
int mpm_ensemble_autotune(mpm_ctx* c, int trial_substeps, mpm_tuning* out, int cap, int* n_out, mpm_tuning* best) {
    if (!c || trial_substeps < 1 || (cap > 0 && !out)) return MPM_E_INVALID;
    if (!c->ens) return fail(c, MPM_E_INVALID, "not a batched context");
    if (!c->has_particles) return fail(c, MPM_E_STATE, "every scene needs its particles uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc;
    const size_t ps = (size_t)NPLANES * c->stride;
    if (!c->snap_planes) {
        if ((rc = dalloc(c, &c->snap_planes, ps)) || (rc = dalloc(c, &c->snap_pid, c->stride)) ||
            (rc = dalloc(c, &c->snap_fpid, 3 * (size_t)c->stride))) return rc;
    }
    HIPCHK(c, hipMemcpy(c->snap_planes, c->planes[c->cur], ps * 4, hipMemcpyDeviceToDevice));
    HIPCHK(c, hipMemcpy(c->snap_pid, c->pid[c->cur], (size_t)c->stride * 4, hipMemcpyDeviceToDevice));
    HIPCHK(c, hipMemcpy(c->snap_fpid, c->fpid, (size_t)c->stride * 48, hipMemcpyDeviceToDevice));
    const int cur0 = c->cur;
    const uint64_t sub0 = c->substeps;
    // one launch per phase only: the persistent schedule (selectable with mpm_ensemble_set_tuning)
    // measured 3.8-7 ms per substep of the 64-scene workload against 2.0 ms, its agent-scope
    // release / acquire per dependency costing more than the launches it saves
    const uint32_t chunks[] = {512, 1024, 2048};
    const uint32_t wgs[] = {2, 3, 4};
    std::vector<mpm_tuning> cand;
    for (uint32_t ch : chunks)
        for (uint32_t w : wgs) { mpm_tuning t{}; t.chunk_particles = ch; t.workgroups_per_cu = w; t.persistent = 0; cand.push_back(t); }
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    int bi = -1;
    for (size_t i = 0; i < cand.size(); ++i) {
        if ((rc = mpm_ensemble_set_tuning(c, &cand[i]))) break;
        HIPCHK(c, hipMemcpy(c->planes[cur0], c->snap_planes, ps * 4, hipMemcpyDeviceToDevice));
        HIPCHK(c, hipMemcpy(c->pid[cur0], c->snap_pid, (size_t)c->stride * 4, hipMemcpyDeviceToDevice));
        HIPCHK(c, hipMemcpy(c->fpid, c->snap_fpid, (size_t)c->stride * 48, hipMemcpyDeviceToDevice));
        c->cur = cur0;
        if ((rc = mpm_ensemble_step(c, 2))) break;                   // builds both graph parities
        HIPCHK(c, hipMemcpy(c->planes[cur0], c->snap_planes, ps * 4, hipMemcpyDeviceToDevice));
        HIPCHK(c, hipMemcpy(c->pid[cur0], c->snap_pid, (size_t)c->stride * 4, hipMemcpyDeviceToDevice));
        HIPCHK(c, hipMemcpy(c->fpid, c->snap_fpid, (size_t)c->stride * 48, hipMemcpyDeviceToDevice));
        c->cur = cur0;
        HIPCHK(c, hipEventRecord(e0, c->stream));
        if ((rc = mpm_ensemble_step(c, trial_substeps))) break;
        HIPCHK(c, hipEventRecord(e1, c->stream));
        HIPCHK(c, hipEventSynchronize(e1));
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        cand[i].ms_per_substep = ms / (float)trial_substeps;
        if ((rc = ens_check_flags(c))) break;
        if (bi < 0 || cand[i].ms_per_substep < cand[bi].ms_per_substep) bi = (int)i;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->planes[cur0], c->snap_planes, ps * 4, hipMemcpyDeviceToDevice));
    HIPCHK(c, hipMemcpy(c->pid[cur0], c->snap_pid, (size_t)c->stride * 4, hipMemcpyDeviceToDevice));
    HIPCHK(c, hipMemcpy(c->fpid, c->snap_fpid, (size_t)c->stride * 48, hipMemcpyDeviceToDevice));
    c->cur = cur0;
    c->substeps = sub0;
    if (rc) return rc;
    if ((rc = mpm_ensemble_set_tuning(c, &cand[bi]))) return rc;
    for (int i = 0; i < cap && i < (int)cand.size(); ++i) out[i] = cand[i];
    if (n_out) *n_out = (int)cand.size();
    if (best) *best = cand[bi];
    return MPM_OK;
}

int mpm_set_timing(mpm_ctx* c, int enabled) {
    if (!c) return MPM_E_INVALID;
    c->timing = enabled != 0;
    return MPM_OK;
}
int mpm_set_kernel_timing(mpm_ctx* c, int enabled) {
    if (!c) return MPM_E_INVALID;
    kt_collect(c);
    c->ktiming = enabled != 0;
    for (int i = 0; i < MPM_KERNEL_COUNT; ++i) { c->kt_ms[i] = 0.0; c->kt_cnt[i] = 0; }
    return MPM_OK;
}
int mpm_get_kernel_times(mpm_ctx* c, double* total_ms, uint64_t* counts) {
    if (!c || !total_ms || !counts) return MPM_E_INVALID;
    for (int i = 0; i < MPM_KERNEL_COUNT; ++i) { total_ms[i] = c->kt_ms[i]; counts[i] = c->kt_cnt[i]; }
    return MPM_OK;
}
const char* mpm_kernel_name(int id) { return (id >= 0 && id < MPM_KERNEL_COUNT) ? k_kernel_names[id] : ""; }

int mpm_get_timings(mpm_ctx* c, mpm_timings* out) {
    if (!c || !out) return MPM_E_INVALID;
    *out = c->tm;
    return MPM_OK;
}

}  // extern "C"
