// Batched scenes: many independent simulations of uneven particle count, grid size, method and
// iteration count advance together, one substep of every scene per call.
//
// Layout: the scenes share every buffer of the single-scene path. Particles of scene s have the
// original ids [part_base[s], part_base[s+1]); its block keys are key_base[s] + the scene's own
// linearised block coordinate, so one scan over the concatenated key space bins all scenes and
// leaves each scene's runs and work chunks contiguous (scene order = key order). Scene s owns the
// dense block range [blk_base, blk_base + blk_cap) of the shared grids (blk_cap = its whole key
// space, so no scene can starve another of blocks). The block hash is salted with the scene index.
//
// Schedule: binning runs once over all particles (k_ens_keys_count .. k_ens_nbr, the single-scene
// scans in between), k_ens_tasks builds the task table, then ONE persistent kernel,
// k_ens_transfer, runs every transfer phase of every scene:
//   PBMPM scene, k iterations: phase 0 FIRST, phases 1..k-1 MID, phase k LAST
//   explicit scene:            phase 0 P2G, phase 1 grid update fused into G2P + integrate
// A task is (phase r, scene s, chunk j). Tasks are dequeued from one atomic counter in
// phase-major order; a task of phase r > 0 waits until all phase r-1 tasks of its scene have
// finished (a per-(phase, scene) counter). Every task it waits on was dequeued before it by a
// workgroup that is already running, and phase-0 tasks wait on nothing, so the wait always ends
// whatever the number of resident workgroups. A spin limit still bounds it: on expiry the kernel
// sets an error flag (reported by the host) and goes on. Ordering across workgroups (and the
// L2 of different XCDs) is agent-scope: release fence before a finished-task count, acquire fence
// after the wait. Each task also clears its slice of the scene's blocks in the grid buffer the
// scene's next phase writes (PBMPM triple buffering).
#pragma once
#include "mpm_kernels.h"

namespace mpm {

enum : uint32_t { ENS_MAX_SCENES = 1024, ENS_MAX_SEG = 4096, ENS_CTL_PHASE = 8, ENS_SPIN_LIMIT = 1u << 24 };
enum : uint32_t { ENS_ERR_WAIT = 1u };

struct Ens {
    const SimParams* sp;          // [S] per-scene parameters (kbase, blk_base, blk_cap, scene set)
    const uint32_t* part_base;    // [S+1] first original particle id of each scene
    const uint32_t* key_base;     // [S+1] first global block key of each scene
    const uint32_t* nphase;       // [S] transfer phases per substep
    const uint32_t* method;       // [S] 0 explicit, 1 PBMPM
    uint32_t* nblk;               // [S] blocks allocated this substep
    uint32_t* rbase;              // [S] direct addressing: first entry of the scene's blocks in the
                                  // block list; nullptr: open-addressing block table
    uint32_t* cbeg;               // [S+1] first work chunk of each scene
    uint32_t* seg;                // [R*S+1] first task of each (phase, scene) segment, phase-major
    uint32_t* done;               // [R*S] finished tasks per segment
    uint32_t* ctl;                // [0] next task (persistent), [1] task count, [2] error flags,
                                  // [ENS_CTL_PHASE + r] next task of phase r (one launch per phase)
    uint32_t S, R;
};

// Largest i in [0, n) with base[i] <= v (base nondecreasing, base[0] <= v).
MPM_DEV uint32_t seg_find(const uint32_t* base, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

// Runs f(params of scene sc). Items are in scene order (particles after the first sort, runs and
// keys always), so a wave is nearly always inside one scene: then the scene index is made
// wave-uniform and the parameters are read with scalar loads instead of per-lane loads.
template <class F>
MPM_DEV void with_scene(const Ens& E, uint32_t sc, F&& f) {
    const uint32_t s0 = __builtin_amdgcn_readfirstlane(sc);
    if (__ballot(sc == s0) == __ballot(true)) f(E.sp[s0]);
    else f(E.sp[sc]);
}

// ------------------------------------------------------------------ binning over all scenes
// One particle of a lane in the batched keys pass: its scene, key, mask and block coordinates (the
// particle's own scene's parameters), and its place in the wave's key groups.
struct EnsKey {
    bool valid;
    uint32_t key, mask, sc;
    int b[3];
    int leader;                 // lowest lane of the wave with the same key
    uint32_t rank;              // lanes before this one with the same key
    uint32_t lead_cnt, lead_m;  // on a leader: its key's lanes and the OR of their masks
};
MPM_DEV EnsKey ens_key_of(const SimParams& P0, const Bufs& B, const Ens& E, const uint32_t* pb, bool active, uint32_t i) {
    EnsKey k; k.valid = false; k.key = EMPTY; k.mask = 0u; k.sc = 0u; k.b[0] = k.b[1] = k.b[2] = 0;
    v3 pos = {0.0f, 0.0f, 0.0f};
    if (active) pos = ld3(B.cur, B.stride, PX, i);
    count_nonfinite(B, active, pos);
    if (active) {
        k.sc = seg_find(pb, E.S, B.pid_cur[i]);
        with_scene(E, k.sc, [&](const SimParams& P) {
            int c[3]; cell_number(P, pos, c);
            k.mask = touched_mask(P, c);
            const int bx = c[0] >> 2, by = c[1] >> 2, bz = c[2] >> 2;
            k.b[0] = bx; k.b[1] = by; k.b[2] = bz;
            if (bx >= 0 && by >= 0 && bz >= 0 && bx < P.nb[0] && by < P.nb[1] && bz < P.nb[2]) {
                k.key = P.kbase + (uint32_t)bx + (uint32_t)P.nb[0] * ((uint32_t)by + (uint32_t)P.nb[1] * (uint32_t)bz);
                k.valid = true;
            } else if (P0.direct_blocks && k.mask) {
                for (int jj = 0; jj < 27; ++jj)
                    if ((k.mask >> jj) & 1u) B.kalloc[block_key_at(P, bx, by, bz, jj)] = 1u;
            }
        });
        B.key[i] = k.key; B.mask[i] = k.mask;
    }
    return k;
}
// Groups the wave's lanes by key (as group_keys): the lowest lane of each key leads it.
MPM_DEV void ens_group(EnsKey& k) {
    uint64_t active = __ballot(k.valid);
    k.leader = 0; k.rank = 0; k.lead_cnt = 0; k.lead_m = 0;
    while (active) {
        const int leader = __ffsll((unsigned long long)active) - 1;
        const uint32_t lk = __shfl(k.key, leader, 64);
        const bool peer = k.valid && k.key == lk;
        const uint64_t peers = __ballot(peer);
        const uint32_t m = wave_or(peer ? k.mask : 0u);
        if (peer) { k.leader = leader; k.rank = (uint32_t)__popcll(peers & lanemask_lt()); }
        if (lane_id() == leader) { k.lead_cnt = (uint32_t)__popcll(peers); k.lead_m = m; }
        active &= ~peers;
    }
}
// Flags each touched block whose (key, neighbour) bit this wave's OR set first: the leader's new
// bits, scene and block coordinates are read from its lane once per key (scalar), one lane per
// neighbour.
MPM_DEV void ens_flag_fresh(const Bufs& B, const Ens& E, const EnsKey& k, uint32_t fresh) {
    uint64_t fl = __ballot(fresh != 0u);
    while (fl) {
        const int l = __ffsll((unsigned long long)fl) - 1;
        const int j = lane_id();
        const uint32_t um = (uint32_t)__builtin_amdgcn_readlane((int)fresh, l);
        const SimParams& P = E.sp[(uint32_t)__builtin_amdgcn_readlane((int)k.sc, l)];
        const int b0 = __builtin_amdgcn_readlane(k.b[0], l), b1 = __builtin_amdgcn_readlane(k.b[1], l),
                  b2 = __builtin_amdgcn_readlane(k.b[2], l);
        if (j < 27 && ((um >> j) & 1u)) B.kalloc[block_key_at(P, b0, b1, b2, j)] = 1u;
        fl &= fl - 1ull;
    }
}
// Direct addressing (P0.direct_blocks): a scene's dense block range is its key range, so a block's
// index is its key (kslot[key] = key). The touched blocks are flagged here (per wave and key, one
// lane per neighbour; unbinned particles per lane), listed in key order by the key-space scan
// (block_key), and each scene's part of the list is found in k_ens_tasks.
__global__ __launch_bounds__(256) void k_ens_keys_count(SimParams P0, Bufs B, Ens E) {
    __shared__ uint32_t pb[ENS_MAX_SCENES + 1];
    for (uint32_t s = threadIdx.x; s <= E.S; s += blockDim.x) pb[s] = E.part_base[s];
    const uint32_t gsz = gridDim.x * blockDim.x;
    const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (!P0.direct_blocks)
        for (uint32_t s = gi; s < P0.cap; s += gsz) B.table[s] = EMPTY;
    for (uint32_t s = gi; s < E.S; s += gsz) E.nblk[s] = 0;
    if (gi == 0) { B.counters[C_STRAG] = 0; B.counters[C_NBLOCKS] = 0; B.counters[C_DROPS] = 0; B.counters[C_PROBEFAIL] = 0; B.counters[C_NCOLL] = 0; B.counters[C_FALLBACK] = 0; B.counters[C_SVD64] = 0; }
    __syncthreads();
    const uint32_t n = P0.n;
    // two particles per lane per iteration (i0 and i0 + gsz), their key groups' atomics issued
    // together; every lane runs every iteration (the wave-wide groupings need all lanes)
    const uint32_t nloop = (n + 2u * gsz - 1u) / (2u * gsz);
    for (uint32_t it = 0; it < nloop; ++it) {
        const uint32_t i0 = it * 2u * gsz + gi, i1 = i0 + gsz;
        EnsKey a = ens_key_of(P0, B, E, pb, i0 < n, i0);
        EnsKey c = ens_key_of(P0, B, E, pb, i1 < n, i1);
        ens_group(a);
        ens_group(c);
        uint32_t ba = 0, bc = 0, fa = 0, fc = 0;
        if (a.valid && a.leader == lane_id()) { ba = atomicAdd(&B.kcount[a.key], a.lead_cnt); fa = a.lead_m & ~atomicOr(&B.kmask[a.key], a.lead_m); }
        if (c.valid && c.leader == lane_id()) { bc = atomicAdd(&B.kcount[c.key], c.lead_cnt); fc = c.lead_m & ~atomicOr(&B.kmask[c.key], c.lead_m); }
        ba = __shfl(ba, a.leader, 64);
        bc = __shfl(bc, c.leader, 64);
        if (a.valid) B.lofs[i0] = ba + a.rank;
        if (c.valid) B.lofs[i1] = bc + c.rank;
        if (P0.direct_blocks) { ens_flag_fresh(B, E, a, fa); ens_flag_fresh(B, E, c, fc); }
    }
}

__global__ __launch_bounds__(256) void k_ens_insert(SimParams P0, Bufs B, Ens E) {
    __shared__ uint32_t pb[ENS_MAX_SCENES + 1], kb[ENS_MAX_SCENES + 1];
    for (uint32_t s = threadIdx.x; s <= E.S; s += blockDim.x) { pb[s] = E.part_base[s]; kb[s] = E.key_base[s]; }
    __syncthreads();
    const uint32_t nruns = B.counters[C_NRUNS];
    const uint32_t nvalid = B.counters[C_NVALID];
    const uint32_t gsz = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nruns * 27u; t += gsz) {
        const uint32_t r = t / 27u, j = t % 27u;
        if ((B.run_mask[r] >> j) & 1u) {
            const uint32_t key = B.run_key[r];
            with_scene(E, seg_find(kb, E.S, key), [&](const SimParams& P) {
                int b[3]; block_coords(P, key, b);
                table_insert(P, B, b[0] + (int)(j % 3u) - 1, b[1] + (int)((j / 3u) % 3u) - 1, b[2] + (int)(j / 9u) - 1);
            });
        }
    }
    for (uint32_t p = nvalid + blockIdx.x * blockDim.x + threadIdx.x; p < P0.n; p += gsz) {
        const uint32_t old = B.perm[p];
        const uint32_t m = B.mask[old];
        if (!m) continue;
        const SimParams& P = E.sp[seg_find(pb, E.S, B.pid_cur[old])];
        int c[3]; cell_number(P, ld3(B.cur, B.stride, PX, old), c);
        for (int j = 0; j < 27; ++j)
            if ((m >> j) & 1u) table_insert(P, B, (c[0] >> 2) + j % 3 - 1, (c[1] >> 2) + (j / 3) % 3 - 1, (c[2] >> 2) + j / 9 - 1);
    }
}

// Dense block indices per scene: scene s numbers its blocks from blk_base. Each workgroup counts
// its occupied slots per scene in LDS and reserves them with one global atomic per scene present.
__global__ __launch_bounds__(256) void k_ens_compact(SimParams P0, Bufs B, Ens E) {
    __shared__ uint32_t kb[ENS_MAX_SCENES + 1], hist[ENS_MAX_SCENES];
    __shared__ uint32_t total;
    for (uint32_t s = threadIdx.x; s <= E.S; s += blockDim.x) kb[s] = E.key_base[s];
    for (uint32_t s = threadIdx.x; s < E.S; s += blockDim.x) hist[s] = 0;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t key = s < P0.cap ? B.table[s] : EMPTY;
    const bool occ = key != EMPTY;
    const uint64_t om = __ballot(occ);
    if (om && lane_id() == __ffsll((unsigned long long)om) - 1) atomicAdd(&total, (uint32_t)__popcll(om));
    uint32_t sc = 0, local = 0;
    if (occ) { sc = seg_find(kb, E.S, key); local = atomicAdd(&hist[sc], 1u); }
    __syncthreads();
    if (threadIdx.x == 0 && total) atomicAdd(&B.counters[C_NBLOCKS], total);
    for (uint32_t t = threadIdx.x; t < E.S; t += blockDim.x)
        if (hist[t]) hist[t] = atomicAdd(&E.nblk[t], hist[t]);
    __syncthreads();
    if (occ) {
        const SimParams& P = E.sp[sc];
        const uint32_t b = hist[sc] + local;
        if (b < P.blk_cap) { B.slot_val[s] = P.blk_base + b; B.block_key[P.blk_base + b] = key; }
        else { B.slot_val[s] = EMPTY; atomicAdd(&B.counters[C_DROPS], 1u); }
    }
}

// Neighbour-block table of every run; clear of grid buffer 0 over every scene's live blocks.
__global__ __launch_bounds__(256) void k_ens_nbr(SimParams P0, Bufs B, Ens E) {
    __shared__ uint32_t kb[ENS_MAX_SCENES + 1];
    for (uint32_t s = threadIdx.x; s <= E.S; s += blockDim.x) kb[s] = E.key_base[s];
    __syncthreads();
    const uint32_t nruns = B.counters[C_NRUNS];
    const uint32_t gsz = gridDim.x * blockDim.x;
    const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t t = gi; t < nruns * 27u; t += gsz) {
        const uint32_t r = t / 27u, j = t % 27u;
        const uint32_t key = B.run_key[r];
        with_scene(E, seg_find(kb, E.S, key), [&](const SimParams& P) {
            int b[3]; block_coords(P, key, b);
            const int x = b[0] + (int)(j % 3u) - 1, y = b[1] + (int)((j / 3u) % 3u) - 1, z = b[2] + (int)(j / 9u) - 1;
            if (P0.direct_blocks) {
                const bool in = x >= 0 && y >= 0 && z >= 0 && x < P.nb[0] && y < P.nb[1] && z < P.nb[2];
                B.nbr[t] = in ? B.kslot[P.kbase + (uint32_t)x + (uint32_t)P.nb[0] * ((uint32_t)y + (uint32_t)P.nb[1] * (uint32_t)z)] : EMPTY;
            } else {
                B.nbr[t] = table_lookup(P, B, x, y, z);
            }
        });
    }
    if (P0.direct_blocks) {
        const uint32_t ncell = min(B.counters[C_NBLOCKS], P0.max_blocks) * 64u;
        for (uint32_t c = gi; c < ncell; c += gsz) {
            const size_t at = (size_t)B.block_key[c >> 6] * 64u + (c & 63u);
            B.grid0[at] = make_int4(0, 0, 0, 0); B.gmass[at] = 0;
        }
        return;
    }
    for (uint32_t sc = 0; sc < E.S; ++sc) {
        const uint32_t nb = min(E.nblk[sc], E.sp[sc].blk_cap);
        int4* g = B.grid0 + (size_t)E.sp[sc].blk_base * 64u;
        int* gm = B.gmass + (size_t)E.sp[sc].blk_base * 64u;
        for (uint32_t c = gi; c < nb * 64u; c += gsz) { g[c] = make_int4(0, 0, 0, 0); gm[c] = 0; }
    }
}

// Task table of one substep (one workgroup): first chunk of every scene, the phase-major
// segments of (phase, scene) tasks, cleared finished-task counters and dequeue counters.
__global__ __launch_bounds__(256) void k_ens_tasks(Bufs B, Ens E) {
    __shared__ uint32_t cb[ENS_MAX_SCENES + 1], kb[ENS_MAX_SCENES + 1];
    __shared__ uint32_t part[256];
    for (uint32_t s = threadIdx.x; s <= E.S; s += blockDim.x) kb[s] = E.key_base[s];
    __syncthreads();
    const uint32_t nch = B.counters[C_NCHUNKS];
    if (E.rbase) {
        // direct addressing: scene s's blocks are the block-list entries with keys in its key range
        const uint32_t nbt = B.counters[C_NBLOCKS];
        for (uint32_t s = threadIdx.x; s < E.S; s += blockDim.x) {
            uint32_t b[2];
            for (int e = 0; e < 2; ++e) {
                uint32_t lo = 0, hi = nbt;
                const uint32_t kv = kb[s + e];
                while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (B.block_key[mid] < kv) lo = mid + 1; else hi = mid; }
                b[e] = lo;
            }
            E.rbase[s] = b[0]; E.nblk[s] = b[1] - b[0];
        }
    }
    for (uint32_t s = threadIdx.x; s <= E.S; s += blockDim.x) {
        // first chunk whose scene is >= s (chunks are in key order, keys in scene order)
        uint32_t lo = 0, hi = nch;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (seg_find(kb, E.S, B.run_key[B.chunks[mid].x]) < s) lo = mid + 1; else hi = mid;
        }
        cb[s] = s == E.S ? nch : lo;
        E.cbeg[s] = cb[s];
    }
    __syncthreads();
    const uint32_t nseg = E.R * E.S;
    const uint32_t per = (nseg + blockDim.x - 1) / blockDim.x;
    const uint32_t g0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t g = g0; g < g0 + per && g < nseg; ++g) {
        const uint32_t r = g / E.S, s = g % E.S;
        sum += r < E.nphase[s] ? cb[s + 1] - cb[s] : 0u;
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t t = 0; t < blockDim.x; ++t) { const uint32_t v = part[t]; part[t] = acc; acc += v; }
        E.seg[nseg] = acc;
        E.ctl[0] = 0; E.ctl[1] = acc;
        for (uint32_t r = 0; r < E.R; ++r) E.ctl[ENS_CTL_PHASE + r] = 0;
    }
    __syncthreads();
    uint32_t acc = part[threadIdx.x];
    for (uint32_t g = g0; g < g0 + per && g < nseg; ++g) {
        const uint32_t r = g / E.S, s = g % E.S;
        E.seg[g] = acc;
        E.done[g] = 0;
        acc += r < E.nphase[s] ? cb[s + 1] - cb[s] : 0u;
    }
}

// Clear of slice j of n of scene sc's live blocks in grid buffer g: a dense range from blk_base
// (block table), or the scene's entries of the block list, each block at its key (direct).
MPM_DEV void ens_clear_slice(const SimParams& P, const Bufs& B, const Ens& E, uint32_t sc, int4* g, uint32_t j, uint32_t n) {
    const uint32_t nb = min(E.nblk[sc], P.blk_cap);
    const uint32_t b0 = (uint32_t)(((uint64_t)nb * j) / n), b1 = (uint32_t)(((uint64_t)nb * (j + 1)) / n);
    if (E.rbase) {
        const uint32_t* list = B.block_key + E.rbase[sc] + b0;
        for (uint32_t c = threadIdx.x; c < (b1 - b0) * 64u; c += blockDim.x)
            g[(size_t)list[c >> 6] * 64u + (c & 63u)] = make_int4(0, 0, 0, 0);
        return;
    }
    int4* base = g + (size_t)(P.blk_base + b0) * 64u;
    for (uint32_t c = threadIdx.x; c < (b1 - b0) * 64u; c += blockDim.x) base[c] = make_int4(0, 0, 0, 0);
}

// The transfer phases of every scene. phase < 0: persistent, all phases with dependency waits;
// phase >= 0: only that phase's tasks (the launch boundary orders the phases). Tasks are dequeued
// from an atomic counter by the resident workgroups, which load the task tables into LDS once (an
// oversubscribed grid with one task per workgroup, reloading them, measured 4% slower). OCC:
// resident workgroups per CU the register allocation targets (2, 3 or 4; the autotuner measures
// which pays).
template <int OCC>
__global__ __launch_bounds__(256, OCC) void k_ens_transfer(Bufs B, Ens E, int phase) {
    __shared__ Tile T;
    __shared__ uint32_t sseg[ENS_MAX_SEG + 1];
    __shared__ uint32_t scb[ENS_MAX_SCENES + 1];
    __shared__ uint32_t task[3];
    const uint32_t S = E.S, nseg = E.R * S;
    for (uint32_t i = threadIdx.x; i <= nseg; i += blockDim.x) sseg[i] = E.seg[i];
    for (uint32_t i = threadIdx.x; i <= S; i += blockDim.x) scb[i] = E.cbeg[i];
    __syncthreads();
    const uint32_t first = phase < 0 ? 0u : sseg[(uint32_t)phase * S];
    const uint32_t last = phase < 0 ? sseg[nseg] : sseg[((uint32_t)phase + 1u) * S];
    uint32_t* counter = phase < 0 ? &E.ctl[0] : &E.ctl[ENS_CTL_PHASE + phase];
    for (;;) {
        if (threadIdx.x == 0) {
            const uint32_t t = first + atomicAdd(counter, 1u);
            uint32_t g = 0;
            if (t < last) {
                g = seg_find(sseg, nseg, t);
                const uint32_t r = g / S, sc = g % S;
                if (phase < 0 && r > 0) {
                    const uint32_t need = scb[sc + 1] - scb[sc];
                    uint32_t spins = 0;
                    while (__hip_atomic_load(&E.done[g - S], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                        __builtin_amdgcn_s_sleep(2);
                        if (++spins > ENS_SPIN_LIMIT) { atomicOr(&E.ctl[2], ENS_ERR_WAIT); break; }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                }
            }
            task[0] = t; task[1] = g; task[2] = t < last ? t - sseg[g] : 0u;
        }
        __syncthreads();
        const uint32_t t = __builtin_amdgcn_readfirstlane(task[0]);
        if (t >= last) break;
        const uint32_t g = __builtin_amdgcn_readfirstlane(task[1]);
        const uint32_t j = __builtin_amdgcn_readfirstlane(task[2]);
        const uint32_t r = g / S, sc = g % S;
        const SimParams& P = E.sp[sc];
        const uint2 ch = B.chunks[scb[sc] + j];
        const uint32_t ntask = scb[sc + 1] - scb[sc];
        if (E.method[sc] != 0u) {
            const int k = P.k;
            if (r == 0) {
                if (k > 1) ens_clear_slice(P, B, E, sc, B.grid1, j, ntask);
                pbmpm_chunk<PB_FIRST, false, false>(P, B, T, ch, nullptr, B.grid0, -1);
            } else if ((int)r < k) {
                ens_clear_slice(P, B, E, sc, grid_sel(B, (int)(r + 1) % 3), j, ntask);
                pbmpm_chunk<PB_MID, false, false>(P, B, T, ch, grid_sel(B, (int)(r - 1) % 3), grid_sel(B, (int)r % 3), (int)r - 1);
            } else {
                pbmpm_chunk<PB_LAST, false, false>(P, B, T, ch, grid_sel(B, (k - 1) % 3), nullptr, k - 1);
            }
        } else {
            if (r == 0) explicit_p2g_chunk<false>(P, B, T, ch, B.grid0);
            else explicit_g2p_chunk<true, false>(P, B, T, ch, B.grid0);
        }
        __syncthreads();
        if (threadIdx.x == 0 && phase < 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            atomicAdd(&E.done[g], 1u);
        }
    }
}

// Unbinned particles of every scene, integrated with their scene's parameters.
__global__ __launch_bounds__(256) void k_ens_straggle(SimParams P0, Bufs B, Ens E) {
    __shared__ uint32_t pb[ENS_MAX_SCENES + 1];
    for (uint32_t s = threadIdx.x; s <= E.S; s += blockDim.x) pb[s] = E.part_base[s];
    __syncthreads();
    const uint32_t nvalid = B.counters[C_NVALID];
    for (uint32_t q = nvalid + blockIdx.x * blockDim.x + threadIdx.x; q < P0.n; q += gridDim.x * blockDim.x) {
        const uint32_t old = B.perm[q];
        straggle_one(E.sp[seg_find(pb, E.S, B.pid_cur[old])], B, q, old);
    }
}

// Per-scene particle exchange: rows whose original id lies in [base, base + cnt).
__global__ void k_upload_range(const float* __restrict__ aos, float* __restrict__ s, const uint32_t* __restrict__ pid,
                               float4* __restrict__ fpid, uint32_t nrows, uint32_t stride, uint32_t base, uint32_t cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const uint32_t id = pid[i] - base;
    if (id >= cnt) return;
    put_record(aos + (size_t)id * 48, s, stride, fpid, i, pid[i]);
}
__global__ void k_download_range(const float* __restrict__ s, const uint32_t* __restrict__ pid, const float4* __restrict__ fpid,
                                 float* __restrict__ aos, uint32_t nrows, uint32_t stride, uint32_t base, uint32_t cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const uint32_t id = pid[i] - base;
    if (id >= cnt) return;
    get_record(aos + (size_t)id * 48, s, stride, fpid, i, pid[i]);
}
__global__ void k_iota(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

}  // namespace mpm
