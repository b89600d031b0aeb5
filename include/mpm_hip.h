/*
 * libmpm_hip — C ABI of the MI355X (gfx950) MPM snow-simulation step.
 *
 * Drop-in for the simulation path of luoluobuli/WebGPU-MPM-based-Snow-Simulation: the
 * TypeScript managers that own the WGSL compute passes (src/lib/gpu/mpm/GpuMpmBufferManager.ts,
 * GpuMpmPipelineManager.ts, src/lib/gpu/uniforms/GpuUniformsBufferManager.ts,
 * src/lib/gpu/collider/GpuColliderBufferManager.ts) bind these entry points instead of
 * creating GPU buffers and pipelines. Plain pointers and sizes only; every entry returns
 * 0 on success or a negative MPM_E_* code, and mpm_last_error() gives the message.
 *
 * Threading: a context is not reentrant (one caller at a time) but may be called from
 * different OS threads; each entry binds the context's device first.
 */
#ifndef MPM_HIP_H
#define MPM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_UNIFORMS_BYTES 37376   /* GpuUniformsBufferManager.ts:17 */
#define MPM_PARTICLE_BYTES 192     /* shaderPrelude/mpmOps.wgsl:1-12, GpuMpmBufferManager.ts:24 */
#define MPM_BVH_NODE_BYTES 32      /* mpm/colliderPrelude.wgsl:6-11 */
#define MPM_KERNEL_COUNT 25        /* kernels of one substep, see mpm_kernel_name() */

enum {
    MPM_OK = 0,
    MPM_E_INVALID = -1,   /* bad argument / handle / size mismatch */
    MPM_E_HIP = -2,       /* HIP runtime error */
    MPM_E_NOMEM = -3,     /* device allocation failed */
    MPM_E_RANGE = -4,     /* value outside what the context was created for */
    MPM_E_STATE = -5      /* call not valid in the current state (e.g. no particles yet) */
};

enum { MPM_METHOD_EXPLICIT = 0, MPM_METHOD_PBMPM = 1 };   /* GpuSimulationMethod.ts:1-4 */

/* Creation-time configuration. Every field that the reference hard-codes in WGSL or TS has
 * a runtime value here; mpm_config_defaults() fills the reference values. */
/* Largest particle count of a context (the sum over the scenes of a batch): rows are addressed
 * with 32-bit byte offsets. mpm_create / mpm_ensemble_create return MPM_E_RANGE above it. */
#define MPM_MAX_PARTICLES 0x3FFFFF00u

typedef struct mpm_config {
    uint32_t n_particles;          /* GpuMpmBufferManager.ts:15-21 nParticles */
    uint32_t grid_res[3];          /* uniforms gridResolution @208 */
    float    grid_min[3];          /* uniforms gridMinCoords @16 (reference: -5) */
    float    grid_max[3];          /* uniforms gridMaxCoords @32 (reference: +5) */
    float    fixed_point_scale;    /* uniforms fixedPointScale @4 (GpuSnowPipelineRunner.svelte.ts:28, 65536) */
    float    dt;                   /* uniforms simulationTimestep @0 */
    uint32_t max_blocks;           /* sparseGridPrelude.wgsl:7; 0 = derived from grid and N */
    uint32_t hash_map_size;        /* sparseGridPrelude.wgsl:6; nonzero: open-addressing block table of
                                      this size with the reference's probing (probe failures possible);
                                      0: blocks addressed directly by block key (same allocated set,
                                      no table, no probe failures) */
    uint32_t max_probes;           /* sparseGridPrelude.wgsl:8 (100) */
    int32_t  pbmpm_iterations;     /* GpuMpmPipelineManager.ts:445 (3) */
    float    pbmpm_force_divisor;  /* pbmpm.wgsl:137 (3.0) */
    float    gravity[3];           /* gridUpdate.cs.wgsl:43, pbmpm.wgsl:101 ((0,0,-9.81)) */
    double   youngs_modulus;       /* stressTensorOps.wgsl:1 (1.4e5) */
    double   poissons_ratio;       /* stressTensorOps.wgsl:2 (0.2) */
    float    hardening_explicit;   /* stressTensorOps.wgsl:44 (5) */
    float    hardening_pbmpm;      /* pbmpm.wgsl:35 (20) */
    float    pbmpm_blend;          /* pbmpm.wgsl:26 (0.95) */
    float    pbmpm_volume_exponent;/* pbmpm.wgsl:23 (-0.3333333) */
    double   theta_c;              /* plasticityOps.wgsl:18 (2.5e-2) */
    double   theta_s;              /* plasticityOps.wgsl:19 (7.5e-3) */
    double   density;              /* particleToGrid.cs.wgsl:36 (400) */
    float    grad_scale;           /* scale of D written by G2P (1 = reference, 1/h = physical) */
    float    collider_margin;      /* colliderPrelude.wgsl:158,297,330,351 (0.05) */
    float    friction_ccd;         /* colliderPrelude.wgsl:308 (0.1) */
    float    friction_proximity;   /* colliderPrelude.wgsl:339 (0.0) */
    float    collider_velocity_factor; /* colliderPrelude.wgsl:294 (0.2) */
    uint32_t wall_mode;            /* 0 = reference clamp+restitution; bit0 = also sticky grid border */
    uint32_t wall_cells;           /* sticky border thickness in cells (3) */
    uint32_t chunk_particles;      /* particles per LDS-tiled work item: 0 = default 512, else 256..65536
                                      (MPM_E_RANGE otherwise); a batch takes scene 0's value */
    uint32_t grid_workgroups;      /* workgroups of the chunk kernels (0 = derived from CU count) */
    int32_t  device;               /* HIP device ordinal */
    uint32_t use_graph;            /* capture one substep in a hipGraph and replay it */
    uint32_t plasticity_mode;      /* MPM_PLASTICITY_*: 0 = the reference's check of diag(R^T F_e) after the
                                      4-step polar decomposition (plasticityOps.wgsl:7-48); 1 = 3x3 SVD
                                      F_e = U S V^T, S clamped to [1 - theta_c, 1 + theta_s] (no reference
                                      counterpart; Stomakhin et al. 2013); 2 = SVD with an FP64 SVD on the
                                      GPU for near-degenerate F_e (sigma_min < 1e-2 sigma_max) */
} mpm_config;

enum { MPM_PLASTICITY_POLAR = 0, MPM_PLASTICITY_SVD = 1, MPM_PLASTICITY_SVD_FP64 = 2 };

typedef struct mpm_ctx mpm_ctx;

/* Counters for conditions the reference handles silently (mapAffectedBlocks.wgsl:27-31,52;
 * particleToGrid.cs.wgsl:47-48; countParticlesPerBlock.wgsl:20). Values are of the last substep. */
typedef struct mpm_stats {
    uint64_t n_particles;
    uint64_t particles_binned;
    uint64_t particles_unbinned;     /* own block not allocated / outside the block grid */
    uint64_t particle_blocks;        /* distinct own blocks (runs of the sort) */
    uint64_t work_chunks;
    uint64_t blocks_allocated;       /* n_allocated_blocks */
    uint64_t blocks_dropped;         /* beyond max_blocks */
    uint64_t probe_failures;         /* hash insert exhausted max_probes */
    uint64_t substeps;
    uint64_t mlp_fallbacks;          /* particles whose warm-start prediction was dropped (last substep) */
    uint64_t nonfinite_particles;    /* particles whose position was NaN or infinite at binning (binned at
                                        cell 0 with zero grid contributions, as the reference's i32(NaN) = 0) */
    uint64_t svd_fp64_particles;     /* MPM_PLASTICITY_SVD_FP64: particles decomposed in FP64 (last substep) */
    uint64_t collider_queued;        /* collider broad-phase candidates sent to the narrow phase (last substep) */
} mpm_stats;

/* Device time of the last mpm_step / mpm_ensemble_step made with timing on (mpm_set_timing).
 * compute_simulation_step_ns spans the whole call, all pass_substeps substeps, as the reference's
 * computeSimulationStepNs spans the simulation compute pass of a frame (timestamp slots 0/1,
 * GpuSnowPipelineRunner.svelte.ts:411-420,588-611). The three phases are those of the call's last
 * substep, which is launched eagerly with events between its phases. */
typedef struct mpm_timings {
    double compute_simulation_step_ns;   /* GpuSnowPipelineRunner.svelte.ts:520-524 computeSimulationStepNs */
    double binning_ns;                   /* last substep: keys, scan, scatter (and table) */
    double transfer_ns;                  /* last substep: P2G / grid update / G2P / constraint passes */
    double integrate_ns;                 /* last substep: integrate, collider, unbinned particles */
    uint64_t substeps;                   /* substeps stepped by the context so far */
    uint64_t pass_substeps;              /* substeps inside compute_simulation_step_ns */
} mpm_timings;

void mpm_config_defaults(mpm_config* cfg);

/* new GpuMpmBufferManager({device, nParticles}) + new GpuMpmPipelineManager({...})
 * (GpuMpmBufferManager.ts:15-80, GpuMpmPipelineManager.ts:42-66). Also writes the uniform
 * fields the runner's constructor writes (GpuSnowPipelineRunner.svelte.ts:136-144). */
int  mpm_create(const mpm_config* cfg, mpm_ctx** out);
void mpm_destroy(mpm_ctx* ctx);

/* device.queue.writeBuffer(uniformsBuffer, offset, bytes) — byte-compatible with the
 * reference Uniforms layout (shaderPrelude/uniforms.wgsl:8-56, GpuUniformsBufferManager.ts:54-182). */
int  mpm_write_uniforms(mpm_ctx* ctx, size_t offset, const void* bytes, size_t len);
int  mpm_read_uniforms(mpm_ctx* ctx, size_t offset, void* bytes, size_t len);

/* colliderDataBuffer + colliderBvhBuffer (GpuColliderBufferManager.ts:96-219): data_words is
 * the packed u32 collider data (indices in BVH leaf order, then f32 xyz vertices, ...),
 * bvh_nodes are 32-byte BvhNode records. Also writes uniforms colliderNumIndices @348. */
int  mpm_set_collider(mpm_ctx* ctx, const uint32_t* data_words, size_t n_words, uint32_t n_indices,
                      const void* bvh_nodes, size_t n_nodes);

/* GpuColliderBufferManager.buildBvh / buildBvhRecursive (GpuColliderBufferManager.ts:222-350)
 * on the host: median split along the longest axis by centroid, leaf <= 4 triangles, 32-byte
 * nodes, left child = node + 1. tri_order receives the leaf-order triangle permutation
 * (reorderedIndices, :211-219). nodes_cap >= 2*n_tris (or 1 when empty). */
int  mpm_build_bvh(const float* positions, size_t n_vertices, const uint32_t* indices, size_t n_indices,
                   void* nodes_out, size_t nodes_cap, size_t* n_nodes, uint32_t* tri_order);

/* queue.writeBuffer(particleDataBuffer, ...) / readback of the 192-byte ParticleData records. */
int  mpm_upload_particles(mpm_ctx* ctx, const void* aos192, size_t n);
int  mpm_download_particles(mpm_ctx* ctx, void* aos192, size_t n);

/* Per-particle Young's modulus for the explicit method (no reference counterpart: the reference
 * uses one E, stressTensorOps.wgsl:1-6). E is given by particle (upload order); the Lame
 * parameters are formed per particle exactly as the scene-wide ones (double precision, then FP32)
 * and travel with the particle. E = NULL returns to the scene-wide constants. Call after the
 * particles are uploaded or scattered. */
int  mpm_set_particle_youngs(mpm_ctx* ctx, const float* E, size_t n);

/* scatterParticlesInMeshVolume (GpuSnowPipelineRunner.svelte.ts:368-381) with the scatterParticles
 * pass (particleScatter.cs.wgsl:40-84): fills every particle record of the context by rejection
 * sampling inside a closed triangle mesh. mesh_xyzw holds consecutive triangle vertices, 16 bytes
 * each (the GpuMeshBufferManager.ts:29-43 buffer); the sampling box is the uniforms'
 * meshMinCoords / meshMaxCoords (@176 / @192, GpuUniformsBufferManager.ts:86-92). */
int  mpm_scatter_particles(mpm_ctx* ctx, const float* mesh_xyzw, size_t n_vertices);

/* loadGltfScene (src/lib/components/simulationViewer/loadScene.ts:31-251) in C++ for a glTF binary
 * (.glb): world-space positions with the y/z swap, normals, uvs, per-vertex texture indices,
 * offset indices, one collider object per mesh and the scatter mesh (the vertex of every index).
 * On entry the n_* fields are capacities (a NULL array is not written); on return they are the
 * counts. MPM_E_RANGE when a non-NULL array is too small (counts still set); MPM_E_INVALID with a
 * message in err for unsupported or malformed files. Context-free. */
typedef struct mpm_scene_arrays {
    float*    positions;         /* 3 per vertex */
    float*    uvs;               /* 2 per vertex */
    uint32_t* material_indices;  /* 1 per vertex */
    size_t    n_vertices;
    float*    normals;           /* 3 per normal; meshes without normals add none (as the reference) */
    size_t    n_normals;
    uint32_t* indices;
    size_t    n_indices;
    uint32_t* objects;           /* 8 words per mesh: min xyz (f32 bits), first index, max xyz, index count
                                    (the uniforms' ColliderObject layout, uniforms.wgsl:1-6) */
    size_t    n_objects;
    float*    soup_xyzw;         /* 4 per index: the scatter mesh (GpuMeshBufferManager.ts layout) */
    size_t    n_soup;
} mpm_scene_arrays;
int  mpm_load_glb(const void* glb, size_t len, mpm_scene_arrays* out, char* err, size_t err_cap);

/* Render-side particle splats: trilinear u32 fixed-point mass sums on a dense grid, cleared first.
 *  MPM_SPLAT_DENSITY  calculateDensity (marchingCubes/mcDensity.cs.wgsl:14-63) on a res[0..2] grid
 *                     (the reference sizes it floor(grid resolution / 2), GpuMarchingCubesBufferManager.ts:44-48)
 *  MPM_SPLAT_MASS     calculateGridMass (volumetric/calculateGridMass.wgsl:6-51) on the uniforms'
 *                     gridResolution (res ignored)
 * The grid stays on the device (mpm_splat_grid, valid until the next larger splat or destroy) and is
 * copied to host_out (cap cells) when host_out is non-null. Enqueued on the context's stream. */
#define MPM_SPLAT_DENSITY 0
#define MPM_SPLAT_MASS 1
int  mpm_splat(mpm_ctx* ctx, int kind, const uint32_t* res, uint32_t* host_out, size_t cap);
void* mpm_splat_grid(mpm_ctx* ctx);

/* Allocation-free variant of mpm_splat (no counterpart in the reference). For callers that cannot
 * allocate device memory: render loops that splat every frame, serving loops, code run under
 * stream capture. The same splat kernels accumulate into a file-scope device array of
 * MPM_SPLAT_NOALLOC_MAX_CELLS u32 cells (8 MiB, held on each GPU the library runs on until the
 * process ends) instead of the context's grid, which mpm_splat grows with hipMalloc; the array is
 * cleared by a kernel at every call and copied to host_out (required, cap >= cells). Grids of more
 * cells return MPM_E_RANGE. mpm_splat_grid is not changed. Calls that overlap on one GPU (from
 * several contexts or threads) share the array and race; the last call's values remain in it. */
#define MPM_SPLAT_NOALLOC_MAX_CELLS (1u << 21)
int  mpm_splat_noalloc(mpm_ctx* ctx, int kind, const uint32_t* res, uint32_t* host_out, size_t cap);

/* addExplicitMpmDispatches / addPbmpmDispatches x n_substeps, then queue.submit
 * (GpuSnowPipelineRunner.svelte.ts:392-445, GpuMpmPipelineManager.ts:327-521). Enqueues on the
 * context's stream and returns; mpm_sync waits. Each substep also makes the next one's block keys;
 * the first substep after an upload, a scatter or a change of the grid uniforms makes its own and
 * runs without a graph. */
int  mpm_step(mpm_ctx* ctx, int method, int n_substeps);
int  mpm_sync(mpm_ctx* ctx);
/* No reference counterpart (the reference creates its compute pipelines up front,
 * GpuMpmPipelineManager.ts:42-66): captures, without running anything, the hipGraphs mpm_step
 * replays for `method` with the current uniforms and buffers: one substep, and a run of 8 substeps
 * (mpm_step replays a run from one graph when it is captured, or when a call has 32 or more
 * substeps to go), for both buffer parities. Lets a caller keep graph capture out of a timed or
 * latency-bound call. A no-op without use_graph. */
int  mpm_prepare_graphs(mpm_ctx* ctx, int method);

/* Test hooks (no reference counterpart). mpm_debug_stage runs one substep prefix:
 * explicit: 0 binning, 1 +P2G, 2 +grid update, >=3 full; PBMPM: 0 binning, s in 1..k the grid
 * after the pbmpm dispatch of iteration s-1, > k full, MPM_STAGE_WARM_START binning and the learned
 * warm start only (the particle state then holds the warm-started D). mpm_dump_grid returns every cell of every
 * allocated block keyed by cell coordinate with its (mass, mx, my, mz) i32 accumulators. */
#define MPM_STAGE_WARM_START (-2)
int  mpm_debug_stage(mpm_ctx* ctx, int method, int stage);
int  mpm_dump_grid(mpm_ctx* ctx, int32_t* cell_xyz, int32_t* vals4, size_t cap, size_t* n_out);
int  mpm_dump_blocks(mpm_ctx* ctx, int32_t* block_xyz, size_t cap, size_t* n_out);
int  mpm_dump_particle_blocks(mpm_ctx* ctx, int32_t* block_xyz, uint32_t* counts, size_t cap, size_t* n_out);

/* Device evaluation of the step's scalar / 3x3 arithmetic for bitwise comparison with a host
 * restatement (no reference counterpart). kind: 0 exp(a) 1 log(a) 2 pow(a,b) 3 sqrt(a) 4 a/b
 * 5 hash3(bits a, bits b, index) 6 det3 7 inverse3 8 polar rotation 9 i32(a) 10 copy (a plain
 * device copy of n floats, for counter calibration) 11 LDS atomic throughput (n iterations of a
 * conflict-free wave-wide LDS add on every wave of 8 workgroups per CU; a and b unused, out = {ms,
 * wave-instructions, CUs}); matrix kinds take and return 9 column-major floats per item. */
int  mpm_selftest(int kind, const float* a, const float* b, float* out, size_t n);
/* Whether P2G forms i32((x / d) * scale) with the reciprocal + FMA quotient: it does only after
 * a device check over all 2^32 float x found it equal to the IEEE-division form (no reference
 * counterpart; test hook). */
int  mpm_divisor_fast(mpm_ctx* ctx, float d, float scale, int* fast);

/* PBMPM iteration count and per-iteration force divisor of the following substeps
 * (GpuMpmPipelineManager.ts:445 nSolveConstraintIterations, pbmpm.wgsl:137 "/ 3.0"). */
int  mpm_set_iterations(mpm_ctx* ctx, int k, float force_divisor);

/* Learned PBMPM warm start (no reference counterpart: an addition to the reference's PBMPM
 * solver). A 16-64-64-12 GELU MLP with BF16 weights on the matrix cores maps a particle's state at
 * the start of a substep (singular values of F_e, det F_p, D, and dprev: the previous substep's last
 * per-iteration change of dx) to the change of D and of dx that k_ref iterations make.
 *  - mpm_learn_collect: n_substeps PBMPM substeps with k_run iterations (force divisor k_run),
 *    logging (features, target) pairs of every particle every `every` substeps, up to max_pairs.
 *    Features come from that trajectory; targets from a k_ref-iteration substep (force divisor
 *    k_ref) run from the same start state, which the trajectory then discards (k_ref > k_run
 *    teaches the network where more iterations than the run uses would take the substep).
 *  - mpm_learn_train: Adam on minibatches of `batch` pairs (multiple of 256) drawn with replacement;
 *    mean_loss = mean squared error of the standardised targets over the last 50 steps.
 *  - mpm_learn_get/set_params: 6,288 FP32 parameters (W1[16][64] b1 W2[64][64] b2 W3[64][16] b3)
 *    and 64 standardisation values (feature mean, std, target mean, std; 16 each).
 *  - mpm_learn_enable: apply the warm start before every PBMPM substep; a particle whose
 *    standardised prediction exceeds `limit` (or is not finite), or that no iteration reached in the
 *    previous substep, keeps the plain D and dx (fallback). */
int  mpm_learn_collect(mpm_ctx* ctx, int n_substeps, int k_run, int k_ref, int every, uint64_t max_pairs, uint64_t* n_pairs);
int  mpm_learn_train(mpm_ctx* ctx, int steps, int batch, float lr, uint32_t seed, float* mean_loss);
int  mpm_learn_get_params(mpm_ctx* ctx, float* params, float* norm);
int  mpm_learn_set_params(mpm_ctx* ctx, const float* params, const float* norm);
int  mpm_learn_enable(mpm_ctx* ctx, int on, float limit);
/* The warm start's input dprev by particle id (3 floats each, n = n_particles): set from in_xyz and/or
 * read into out_xyz (either may be NULL). Enabling the warm start clears it; LAST writes it. */
int  mpm_learn_dprev(mpm_ctx* ctx, const float* in_xyz, float* out_xyz, size_t n);
/* Gradient of the MSE loss over n samples (n multiple of 16) computed by the MFMA forward and
 * backward kernels, for checking against a host computation (test hook). */
int  mpm_mlp_grad(const float* params, const float* norm, const float* feats, const float* targs, uint32_t n,
                  float* grads, float* y, float* loss);

/* Per-iteration PBMPM residual, r_i = max_particles |dx_i - dx_{i-1} - g dt^2 / force_divisor|_2
 * / largest grid extent, dx_i from the G2P of iteration i. Every iteration deposits g dt /
 * force_divisor of gravity (pbmpm.wgsl:137), so a free particle's dx grows by that times dt per
 * iteration; the residual subtracts it and measures the constraint solve alone. mean / peak over the substeps since tracking was enabled. */
int  mpm_set_residual_tracking(mpm_ctx* ctx, int on);
int  mpm_get_residuals(mpm_ctx* ctx, float* mean, float* peak, int cap, uint64_t* substeps);

/* Batched scenes (no reference counterpart; the reference simulates one scene). n_scenes
 * independent scenes, each with its own configuration (particle count, grid, dt, material,
 * PBMPM iterations) and method, advance together: one call steps every scene by one of its own
 * substeps. Binning runs once over all particles, and every transfer phase of every scene runs
 * in ONE persistent kernel launch per substep (a phase-ordered task queue; optionally one
 * launch per phase), captured in a hipGraph. Scenes have no collider and no pointer
 * interaction; walls as configured. Results per scene are identical to the same scene stepped
 * alone in its own context. The returned context works with mpm_sync, mpm_destroy,
 * mpm_last_error, mpm_get_stats (totals), timing and kernel timing; particles go through
 * mpm_ensemble_upload / _download (scene-local order). Limits: n_scenes <= 1024 and
 * n_scenes * (largest phase count: PBMPM k + 1, explicit 2) <= 4096. */
typedef struct mpm_tuning {
    uint32_t chunk_particles;      /* particles per work chunk (LDS tile), 256..65536 */
    uint32_t workgroups_per_cu;    /* resident workgroups of the transfer kernel per CU, 1..8 */
    int32_t  persistent;           /* 1: one launch per substep; 0: one launch per phase */
    float    ms_per_substep;       /* autotuner: measured time (0 otherwise) */
} mpm_tuning;
int  mpm_ensemble_create(const mpm_config* scene_cfgs, const int* methods, int n_scenes, mpm_ctx** out);
int  mpm_ensemble_upload(mpm_ctx* ctx, int scene, const void* aos192, size_t n);
int  mpm_ensemble_download(mpm_ctx* ctx, int scene, void* aos192, size_t n);
int  mpm_ensemble_step(mpm_ctx* ctx, int n_substeps);
int  mpm_ensemble_set_tuning(mpm_ctx* ctx, const mpm_tuning* t);
int  mpm_ensemble_get_tuning(mpm_ctx* ctx, mpm_tuning* t);
/* Autotuner: times each candidate of the one-launch-per-phase schedule (chunk size {512, 1024, 2048}
 * x workgroups per CU {2, 3, 4}, 9 in all; the persistent schedule stays selectable by
 * mpm_ensemble_set_tuning) for
 * trial_substeps substeps on the live scenes, restores the state after every trial, and keeps
 * the fastest. Candidates and their times go to out[0..cap), their count to *n_out. */
int  mpm_ensemble_autotune(mpm_ctx* ctx, int trial_substeps, mpm_tuning* out, int cap, int* n_out, mpm_tuning* best);

int  mpm_get_stats(mpm_ctx* ctx, mpm_stats* out);
int  mpm_set_timing(mpm_ctx* ctx, int enabled);
int  mpm_get_timings(mpm_ctx* ctx, mpm_timings* out);
/* Per-kernel device time from HIP events recorded on the context's stream after every kernel
 * (replaces the reference's whole-pass timestamp slots 0/1, GpuSnowPipelineRunner.svelte.ts:411-420).
 * While enabled, substeps are launched eagerly instead of through the hipGraph. total_ms and
 * counts have MPM_KERNEL_COUNT entries; enabling clears them. */
int  mpm_set_kernel_timing(mpm_ctx* ctx, int enabled);
int  mpm_get_kernel_times(mpm_ctx* ctx, double* total_ms, uint64_t* counts);
const char* mpm_kernel_name(int id);

const char* mpm_last_error(mpm_ctx* ctx);
const char* mpm_version(void);

/* In-process consumers: the device stream the context enqueues on (a hipStream_t). */
void* mpm_stream(mpm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* MPM_HIP_H */
