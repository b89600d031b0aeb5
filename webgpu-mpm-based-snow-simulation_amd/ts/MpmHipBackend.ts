// Deno FFI shim: the simulation managers of the reference, backed by libmpm_hip.so.
//
// Drop-in for src/lib/gpu/mpm/GpuMpmBufferManager.ts, GpuMpmPipelineManager.ts and
// src/lib/gpu/uniforms/GpuUniformsBufferManager.ts on the simulation path. The symbols and their
// C signatures are declared in include/mpm_hip.h. Needs `deno run --allow-ffi`.

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
  mpm_sync: { parameters: ["pointer"], result: "i32" },
  mpm_set_timing: { parameters: ["pointer", "i32"], result: "i32" },
  mpm_get_timings: { parameters: ["pointer", "buffer"], result: "i32" },
  mpm_last_error: { parameters: ["pointer"], result: "pointer" },
  mpm_scatter_particles: { parameters: ["pointer", "buffer", "usize"], result: "i32" },
  mpm_splat: { parameters: ["pointer", "i32", "buffer", "buffer", "usize"], result: "i32" },
  mpm_load_glb: { parameters: ["buffer", "usize", "buffer", "buffer", "usize"], result: "i32" },
} as const;

// mpm_config layout (include/mpm_hip.h), 184 bytes, natural alignment.
const CFG = {
  n_particles: 0, grid_res: 4, grid_min: 16, grid_max: 28, fixed_point_scale: 40, dt: 44,
  max_blocks: 48, hash_map_size: 52, max_probes: 56, pbmpm_iterations: 60, pbmpm_force_divisor: 64,
  gravity: 68, youngs_modulus: 80, poissons_ratio: 88, use_graph: 176, plasticity_mode: 180, size: 184,
} as const;

export enum GpuSimulationMethodType { ExplicitMpm = 0, Pbmpm = 1 }

export class MpmHipBackend {
  readonly lib: Deno.DynamicLibrary<typeof SYMBOLS>;
  readonly ctx: Deno.PointerValue;
  readonly nParticles: number;

  constructor(libPath: string, opts: {
    nParticles: number, gridResolution: [number, number, number],
    gridMinCoords: [number, number, number], gridMaxCoords: [number, number, number],
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
    this.check(this.lib.symbols.mpm_create(new Uint8Array(cfg), new Uint8Array(out.buffer)), null);
    this.ctx = Deno.UnsafePointer.create(out[0]);
    this.nParticles = opts.nParticles;
  }

  private check(rc: number, ctx: Deno.PointerValue | null) {
    if (rc === 0) return;
    let msg = "";
    if (ctx) {
      const p = this.lib.symbols.mpm_last_error(ctx);
      if (p) msg = new Deno.UnsafePointerView(p).getCString();
    }
    throw new Error(`libmpm_hip error ${rc}: ${msg}`);
  }

  // GpuUniformsBufferManager.ts:54-182 — same byte offsets; every writer funnels through here.
  writeBuffer(offset: number, data: ArrayBufferView) {
    const bytes = new Uint8Array(data.buffer, data.byteOffset, data.byteLength);
    this.check(this.lib.symbols.mpm_write_uniforms(this.ctx, BigInt(offset), bytes, BigInt(bytes.byteLength)), this.ctx);
  }
  writeSimulationTimestepS(t: number) { this.writeBuffer(0, new Float32Array([t])); }
  writeFixedPointScale(s: number) { this.writeBuffer(4, new Float32Array([s])); }
  writeUsePbmpm(b: boolean) { this.writeBuffer(8, new Uint32Array([b ? 1 : 0])); }
  writeGridMinCoords(v: [number, number, number]) { this.writeBuffer(16, new Float32Array(v)); }
  writeGridMaxCoords(v: [number, number, number]) { this.writeBuffer(32, new Float32Array(v)); }
  writeGridResolution(v: [number, number, number]) { this.writeBuffer(208, new Uint32Array(v)); }
  writeMeshMinCoords(v: [number, number, number]) { this.writeBuffer(176, new Float32Array(v)); }
  writeMeshMaxCoords(v: [number, number, number]) { this.writeBuffer(192, new Float32Array(v)); }
  writeColliderMinCoords(v: [number, number, number]) { this.writeBuffer(224, new Float32Array(v)); }
  writeColliderMaxCoords(v: [number, number, number]) { this.writeBuffer(240, new Float32Array(v)); }
  writeColliderTransformMat(m: Float32Array) { this.writeBuffer(256, m); }
  writeColliderVel(v: [number, number, number]) { this.writeBuffer(320, new Float32Array(v)); }
  writeColliderNumIndices(n: number) { this.writeBuffer(348, new Uint32Array([n])); }
  writeGridCellDims(v: [number, number, number]) { this.writeBuffer(352, new Float32Array(v)); }
  writeColliderNumObjects(n: number) { this.writeBuffer(364, new Uint32Array([n])); }
  writeColliderTransformInv(m: Float32Array) { this.writeBuffer(432, m); }
  writeInteractionPos(v: [number, number, number]) { this.writeBuffer(496, new Float32Array(v)); }
  writeInteractionStrength(s: number) { this.writeBuffer(508, new Float32Array([s])); }
  writeInteractionRadius(r: number) { this.writeBuffer(512, new Float32Array([r])); }
  writeIsInteracting(b: boolean) { this.writeBuffer(516, new Uint32Array([b ? 1 : 0])); }
  writeInteractionMode(m: number) { this.writeBuffer(520, new Uint32Array([m])); }

  // GpuColliderBufferManager: packed data (indices in BVH leaf order, then vertices) + BVH nodes.
  setCollider(colliderData: Uint32Array, numIndices: number, bvhNodes: Uint32Array) {
    this.check(this.lib.symbols.mpm_set_collider(this.ctx, new Uint8Array(colliderData.buffer), BigInt(colliderData.length),
      numIndices, new Uint8Array(bvhNodes.buffer), BigInt(bvhNodes.length / 8)), this.ctx);
  }

  // scatterParticlesInMeshVolume (GpuSnowPipelineRunner.svelte.ts:368-381): the mesh is the
  // GpuMeshBufferManager vertex list (consecutive triangles, 4 floats per vertex); its AABB must
  // already be in the uniforms (writeMeshMinCoords / writeMeshMaxCoords).
  scatterParticlesInMeshVolume(meshXyzw: Float32Array) {
    this.check(this.lib.symbols.mpm_scatter_particles(this.ctx, new Uint8Array(meshXyzw.buffer), BigInt(meshXyzw.length / 4)), this.ctx);
  }

  // marching-cubes density (kind 0, res = the density grid) / volumetric mass grid (kind 1)
  splat(kind: 0 | 1, res: [number, number, number]): Uint32Array {
    const cells = res[0] * res[1] * res[2];
    const out = new Uint32Array(cells);
    this.check(this.lib.symbols.mpm_splat(this.ctx, kind, new Uint8Array(new Uint32Array(res).buffer),
      new Uint8Array(out.buffer), BigInt(cells)), this.ctx);
    return out;
  }

  // particleDataBuffer (192-byte ParticleData records)
  uploadParticles(aos: ArrayBuffer) {
    this.check(this.lib.symbols.mpm_upload_particles(this.ctx, new Uint8Array(aos), BigInt(this.nParticles)), this.ctx);
  }
  downloadParticles(): ArrayBuffer {
    const out = new ArrayBuffer(this.nParticles * 192);
    this.check(this.lib.symbols.mpm_download_particles(this.ctx, new Uint8Array(out), BigInt(this.nParticles)), this.ctx);
    return out;
  }

  // GpuMpmPipelineManager.addExplicitMpmDispatches / addPbmpmDispatches, batched like
  // addSimulationStepsComputePass (GpuSnowPipelineRunner.svelte.ts:392-445).
  addSimulationSteps(method: GpuSimulationMethodType, nSimulationSteps: number) {
    this.check(this.lib.symbols.mpm_step(this.ctx, method, nSimulationSteps), this.ctx);
  }
  async addSimulationStepsAsync(method: GpuSimulationMethodType, nSimulationSteps: number) {
    this.check(await this.lib.symbols.mpm_step_async(this.ctx, method, nSimulationSteps), this.ctx);
  }
  sync() { this.check(this.lib.symbols.mpm_sync(this.ctx), this.ctx); }

  destroy() { this.lib.symbols.mpm_destroy(this.ctx); this.lib.close(); }
}
