// Deno FFI shim: the simulation managers of the reference, backed by libmpm_hip.so.
//
// Drop-in for src/lib/gpu/mpm/GpuMpmBufferManager.ts, GpuMpmPipelineManager.ts,
// src/lib/gpu/uniforms/GpuUniformsBufferManager.ts and src/lib/gpu/collider/GpuColliderBufferManager.ts
// on the simulation path, with the same class and method names; MpmHipBackend holds the context and
// the simulation half of GpuSnowPipelineRunner.svelte.ts (addSimulationStepsComputePass,
// updateColliderTransformMat / Vel, scatterParticlesInMeshVolume, the GPU-time callback data).
// WebGPU arguments (devices, encoders) have no counterpart: work is enqueued on the library's HIP
// stream. The symbols and their C signatures are declared in include/mpm_hip.h.
// Needs `deno run --allow-ffi`.

const SYMBOLS = {
  mpm_config_defaults: { parameters: ["buffer"], result: "void" },
  mpm_create: { parameters: ["buffer", "buffer"], result: "i32" },
  mpm_destroy: { parameters: ["pointer"], result: "void" },
  mpm_write_uniforms: { parameters: ["pointer", "usize", "buffer", "usize"], result: "i32" },
  mpm_set_collider: { parameters: ["pointer", "buffer", "usize", "u32", "buffer", "usize"], result: "i32" },
  mpm_build_bvh: { parameters: ["buffer", "usize", "buffer", "usize", "buffer", "usize", "buffer", "buffer"], result: "i32" },
  mpm_upload_particles: { parameters: ["pointer", "buffer", "usize"], result: "i32" },
  mpm_download_particles: { parameters: ["pointer", "buffer", "usize"], result: "i32" },
  mpm_step: { parameters: ["pointer", "i32", "i32"], result: "i32" },
  mpm_step_async: { name: "mpm_step", parameters: ["pointer", "i32", "i32"], result: "i32", nonblocking: true },
  mpm_prepare_graphs: { parameters: ["pointer", "i32"], result: "i32" },
  mpm_sync: { parameters: ["pointer"], result: "i32" },
  mpm_set_timing: { parameters: ["pointer", "i32"], result: "i32" },
  mpm_get_timings: { parameters: ["pointer", "buffer"], result: "i32" },
  mpm_last_error: { parameters: ["pointer"], result: "pointer" },
  mpm_scatter_particles: { parameters: ["pointer", "buffer", "usize"], result: "i32" },
  mpm_splat: { parameters: ["pointer", "i32", "buffer", "buffer", "usize"], result: "i32" },
  mpm_load_glb: { parameters: ["buffer", "usize", "buffer", "buffer", "usize"], result: "i32" },
  mpm_get_stats: { parameters: ["pointer", "buffer"], result: "i32" },
} as const;

// mpm_config layout (include/mpm_hip.h), 184 bytes, natural alignment.
const CFG = {
  n_particles: 0, grid_res: 4, grid_min: 16, grid_max: 28, fixed_point_scale: 40, dt: 44,
  max_blocks: 48, hash_map_size: 52, max_probes: 56, pbmpm_iterations: 60, pbmpm_force_divisor: 64,
  gravity: 68, youngs_modulus: 80, poissons_ratio: 88, use_graph: 176, plasticity_mode: 180, size: 184,
} as const;

export enum GpuSimulationMethodType { ExplicitMpm = 0, Pbmpm = 1 }   // GpuSimulationMethod.ts:1-4

type Lib = Deno.DynamicLibrary<typeof SYMBOLS>;
type Vec3 = [number, number, number];

function check(lib: Lib, rc: number, ctx: Deno.PointerValue | null) {
  if (rc === 0) return;
  let msg = "";
  if (ctx) {
    const p = lib.symbols.mpm_last_error(ctx);
    if (p) msg = new Deno.UnsafePointerView(p).getCString();
  }
  throw new Error(`libmpm_hip error ${rc}: ${msg}`);
}

// GpuUniformsBufferManager.ts:54-182 — the same writers and byte offsets; every writer funnels
// through mpm_write_uniforms (the library parses the blob at every step).
export class GpuUniformsBufferManager {
  constructor(private lib: Lib, private ctx: Deno.PointerValue) {}
  private write(offset: number, data: ArrayBufferView) {
    const bytes = new Uint8Array(data.buffer, data.byteOffset, data.byteLength);
    check(this.lib, this.lib.symbols.mpm_write_uniforms(this.ctx, BigInt(offset), bytes, BigInt(bytes.byteLength)), this.ctx);
  }
  writeSimulationTimestepS(t: number) { this.write(0, new Float32Array([t])); }
  writeFixedPointScale(s: number) { this.write(4, new Float32Array([s])); }
  writeUsePbmpm(b: boolean) { this.write(8, new Uint32Array([b ? 1 : 0])); }
  writeTime(t: number) { this.write(12, new Uint32Array([t >>> 0])); }
  writeGridMinCoords(v: Vec3) { this.write(16, new Float32Array(v)); }
  writeGridMaxCoords(v: Vec3) { this.write(32, new Float32Array(v)); }
  writeViewProjMat(m: Float32Array) { this.write(48, m); }
  writeViewProjInvMat(m: Float32Array) { this.write(112, m); }
  writeMeshMinCoords(v: Vec3) { this.write(176, new Float32Array(v)); }
  writeMeshMaxCoords(v: Vec3) { this.write(192, new Float32Array(v)); }
  writeGridResolution(v: Vec3) { this.write(208, new Uint32Array(v)); }
  writeColliderMinCoords(v: Vec3) { this.write(224, new Float32Array(v)); }
  writeColliderMaxCoords(v: Vec3) { this.write(240, new Float32Array(v)); }
  writeColliderTransformMat(m: Float32Array) { this.write(256, m); }
  writeColliderVel(v: Vec3) { this.write(320, new Float32Array(v)); }
  writeCameraPos(v: Vec3) { this.write(336, new Float32Array(v)); }
  writeColliderNumIndices(n: number) { this.write(348, new Uint32Array([n])); }
  writeGridCellDims(v: Vec3) { this.write(352, new Float32Array(v)); }
  writeColliderNumObjects(n: number) { this.write(364, new Uint32Array([n])); }
  writeLightViewProjMat(m: Float32Array) { this.write(368, m); }
  writeColliderTransformInv(m: Float32Array) { this.write(432, m); }
  writeInteractionPos(v: Vec3) { this.write(496, new Float32Array(v)); }
  writeInteractionStrength(s: number) { this.write(508, new Float32Array([s])); }
  writeInteractionRadius(r: number) { this.write(512, new Float32Array([r])); }
  writeIsInteracting(b: boolean) { this.write(516, new Uint32Array([b ? 1 : 0])); }
  writeInteractionMode(m: number) { this.write(520, new Uint32Array([m])); }
  writeInteractionDir(v: Vec3) { this.write(528, new Float32Array(v)); }
  writeColliderObjects(objects: { min: Vec3, max: Vec3, startIndex: number, countIndices: number }[]) {
    const n = Math.min(objects.length, 1024);
    const buf = new ArrayBuffer(32 * n);
    const f = new Float32Array(buf), u = new Uint32Array(buf);
    objects.slice(0, n).forEach((o, i) => {
      f.set(o.min, 8 * i); u[8 * i + 3] = o.startIndex; f.set(o.max, 8 * i + 4); u[8 * i + 7] = o.countIndices;
    });
    if (n) this.write(544, u);
  }
}

// GpuMpmBufferManager.ts:15-80: the particle buffer (192-byte ParticleData records) lives in the
// library as SoA planes; these two calls are its writeBuffer / readback.
export class GpuMpmBufferManager {
  constructor(private lib: Lib, private ctx: Deno.PointerValue, readonly nParticles: number) {}
  writeParticles(aos: ArrayBuffer) {
    check(this.lib, this.lib.symbols.mpm_upload_particles(this.ctx, new Uint8Array(aos), BigInt(this.nParticles)), this.ctx);
  }
  readParticles(): ArrayBuffer {
    const out = new ArrayBuffer(this.nParticles * 192);
    check(this.lib, this.lib.symbols.mpm_download_particles(this.ctx, new Uint8Array(out), BigInt(this.nParticles)), this.ctx);
    return out;
  }
}

// GpuColliderBufferManager.ts:61-350: BVH built on the host by mpm_build_bvh (median split,
// leaf <= 4, 32-byte nodes), indices reordered to leaf order, data packed as
// [indices][xyz vertices][normals][uvs][material ids], uploaded with mpm_set_collider.
export class GpuColliderBufferManager {
  readonly colliderData: Uint32Array;
  readonly bvhNodes: Uint32Array;
  readonly numIndices: number;
  readonly minCoords: Vec3;
  readonly maxCoords: Vec3;
  constructor(private lib: Lib, { positions, indices, normals = new Float32Array(), uvs = new Float32Array(), materialIndices = new Uint32Array() }: {
    positions: Float32Array, indices: Uint32Array, normals?: Float32Array, uvs?: Float32Array, materialIndices?: Uint32Array,
  }) {
    const nTris = indices.length / 3;
    const nodes = new Uint32Array(8 * Math.max(2 * nTris, 1));
    const nNodes = new BigUint64Array(1);
    const order = new Uint32Array(Math.max(nTris, 1));
    check(lib, lib.symbols.mpm_build_bvh(new Uint8Array(positions.buffer), BigInt(positions.length / 3), new Uint8Array(indices.buffer),
      BigInt(indices.length), new Uint8Array(nodes.buffer), BigInt(nodes.length / 8), new Uint8Array(nNodes.buffer),
      new Uint8Array(order.buffer)), null);
    this.bvhNodes = nodes.slice(0, 8 * Number(nNodes[0]));
    const reordered = new Uint32Array(indices.length);
    for (let t = 0; t < nTris; t++) reordered.set(indices.subarray(3 * order[t], 3 * order[t] + 3), 3 * t);
    const parts = [reordered, new Uint32Array(positions.buffer, positions.byteOffset, positions.length),
      new Uint32Array(normals.buffer, normals.byteOffset, normals.length), new Uint32Array(uvs.buffer, uvs.byteOffset, uvs.length),
      materialIndices];
    this.colliderData = new Uint32Array(parts.reduce((a, p) => a + p.length, 0));
    let o = 0;
    for (const p of parts) { this.colliderData.set(p, o); o += p.length; }
    this.numIndices = indices.length;
    const mn: Vec3 = [Infinity, Infinity, Infinity], mx: Vec3 = [-Infinity, -Infinity, -Infinity];
    for (let i = 0; i < positions.length; i++) { mn[i % 3] = Math.min(mn[i % 3], positions[i]); mx[i % 3] = Math.max(mx[i % 3], positions[i]); }
    this.minCoords = mn; this.maxCoords = mx;
  }
  upload(ctx: Deno.PointerValue) {
    check(this.lib, this.lib.symbols.mpm_set_collider(ctx, new Uint8Array(this.colliderData.buffer), BigInt(this.colliderData.length),
      this.numIndices, new Uint8Array(this.bvhNodes.buffer), BigInt(this.bvhNodes.length / 8)), ctx);
  }
}

// GpuMpmPipelineManager.ts:327-521: one call enqueues one substep of the method's schedule on the
// library's stream (the reference records its dispatches into the given compute pass).
export class GpuMpmPipelineManager {
  constructor(private lib: Lib, private ctx: Deno.PointerValue) {}
  addExplicitMpmDispatches(_opts: { nParticles?: number, nBlocksInHashMap?: number, hashMapSize?: number } = {}) {
    check(this.lib, this.lib.symbols.mpm_step(this.ctx, GpuSimulationMethodType.ExplicitMpm, 1), this.ctx);
  }
  addPbmpmDispatches(_opts: { nParticles?: number, nBlocksInHashMap?: number, hashMapSize?: number } = {}) {
    check(this.lib, this.lib.symbols.mpm_step(this.ctx, GpuSimulationMethodType.Pbmpm, 1), this.ctx);
  }
}

export class MpmHipBackend {
  readonly lib: Lib;
  readonly ctx: Deno.PointerValue;
  readonly uniformsManager: GpuUniformsBufferManager;
  readonly mpmManager: GpuMpmBufferManager;
  readonly mpmPipelineManager: GpuMpmPipelineManager;
  colliderManager: GpuColliderBufferManager | null = null;

  constructor(libPath: string, opts: {
    nParticles: number, gridResolution: Vec3, gridMinCoords: Vec3, gridMaxCoords: Vec3,
    simulationTimestepS: number, pbmpmIterations?: number,
  }) {
    this.lib = Deno.dlopen(libPath, SYMBOLS);
    const cfg = new ArrayBuffer(CFG.size);
    this.lib.symbols.mpm_config_defaults(new Uint8Array(cfg));
    const dv = new DataView(cfg);
    dv.setUint32(CFG.n_particles, opts.nParticles, true);
    opts.gridResolution.forEach((r, i) => dv.setUint32(CFG.grid_res + 4 * i, r, true));
    opts.gridMinCoords.forEach((v, i) => dv.setFloat32(CFG.grid_min + 4 * i, v, true));
    opts.gridMaxCoords.forEach((v, i) => dv.setFloat32(CFG.grid_max + 4 * i, v, true));
    dv.setFloat32(CFG.dt, opts.simulationTimestepS, true);
    if (opts.pbmpmIterations !== undefined) {
      dv.setInt32(CFG.pbmpm_iterations, opts.pbmpmIterations, true);
      dv.setFloat32(CFG.pbmpm_force_divisor, opts.pbmpmIterations, true);
    }
    const out = new BigUint64Array(1);
    check(this.lib, this.lib.symbols.mpm_create(new Uint8Array(cfg), new Uint8Array(out.buffer)), null);
    this.ctx = Deno.UnsafePointer.create(out[0]);
    this.uniformsManager = new GpuUniformsBufferManager(this.lib, this.ctx);
    this.mpmManager = new GpuMpmBufferManager(this.lib, this.ctx, opts.nParticles);
    this.mpmPipelineManager = new GpuMpmPipelineManager(this.lib, this.ctx);
  }

  // GpuSnowPipelineRunner.svelte.ts:162-177 (collider set-up) and :383-390
  setCollider(mesh: { positions: Float32Array, indices: Uint32Array, normals?: Float32Array, uvs?: Float32Array, materialIndices?: Uint32Array }) {
    const cm = new GpuColliderBufferManager(this.lib, mesh);
    cm.upload(this.ctx);
    const um = this.uniformsManager;
    um.writeColliderMinCoords(cm.minCoords);
    um.writeColliderMaxCoords(cm.maxCoords);
    um.writeColliderNumObjects(1);
    um.writeColliderObjects([{ min: cm.minCoords, max: cm.maxCoords, startIndex: 0, countIndices: cm.numIndices }]);
    this.colliderManager = cm;
  }
  // column-major 4x4 (wgpu-matrix layout) and its inverse
  updateColliderTransformMat(transformMat: Float32Array, transformInv: Float32Array) {
    this.uniformsManager.writeColliderTransformMat(transformMat);
    this.uniformsManager.writeColliderTransformInv(transformInv);
  }
  updateColliderVel(v: Vec3) { this.uniformsManager.writeColliderVel(v); }

  // scatterParticlesInMeshVolume (:368-381): the GpuMeshBufferManager vertex list (consecutive
  // triangles, 4 floats per vertex); its AABB must already be in the uniforms (writeMeshMin/MaxCoords).
  scatterParticlesInMeshVolume(meshXyzw: Float32Array) {
    check(this.lib, this.lib.symbols.mpm_scatter_particles(this.ctx, new Uint8Array(meshXyzw.buffer), BigInt(meshXyzw.length / 4)), this.ctx);
  }

  // addSimulationStepsComputePass (:392-445): writeUsePbmpm, then nSimulationSteps substeps of the
  // method in one call (enqueue-only, like queue.submit); the async form runs on a Deno worker thread.
  addSimulationStepsComputePass({ nSimulationSteps, simulationMethodType }: { nSimulationSteps: number, simulationMethodType: GpuSimulationMethodType }) {
    this.uniformsManager.writeUsePbmpm(simulationMethodType === GpuSimulationMethodType.Pbmpm);
    check(this.lib, this.lib.symbols.mpm_step(this.ctx, simulationMethodType, nSimulationSteps), this.ctx);
  }
  async addSimulationStepsComputePassAsync({ nSimulationSteps, simulationMethodType }: { nSimulationSteps: number, simulationMethodType: GpuSimulationMethodType }) {
    this.uniformsManager.writeUsePbmpm(simulationMethodType === GpuSimulationMethodType.Pbmpm);
    check(this.lib, await this.lib.symbols.mpm_step_async(this.ctx, simulationMethodType, nSimulationSteps), this.ctx);
  }
  sync() { check(this.lib, this.lib.symbols.mpm_sync(this.ctx), this.ctx); }

  // the timestamp query behind onGpuTimeUpdate (:588-611): computeSimulationStepNs of the last
  // addSimulationStepsComputePass made with timing on
  setGpuTiming(on: boolean) { check(this.lib, this.lib.symbols.mpm_set_timing(this.ctx, on ? 1 : 0), this.ctx); }
  gpuTimes(): { computeSimulationStepNs: bigint } {
    const t = new Float64Array(6);
    check(this.lib, this.lib.symbols.mpm_get_timings(this.ctx, new Uint8Array(t.buffer)), this.ctx);
    return { computeSimulationStepNs: BigInt(Math.round(t[0])) };
  }

  // marching-cubes density (kind 0, res = the density grid) / volumetric mass grid (kind 1)
  splat(kind: 0 | 1, res: Vec3): Uint32Array {
    const cells = res[0] * res[1] * res[2];
    const out = new Uint32Array(cells);
    check(this.lib, this.lib.symbols.mpm_splat(this.ctx, kind, new Uint8Array(new Uint32Array(res).buffer),
      new Uint8Array(out.buffer), BigInt(cells)), this.ctx);
    return out;
  }

  destroy() { this.lib.symbols.mpm_destroy(this.ctx); this.lib.close(); }
}
