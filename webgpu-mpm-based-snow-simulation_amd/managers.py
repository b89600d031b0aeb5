"""Host-side mirror of the reference's simulation managers over libmpm_hip.so.

Same class and method names as the TypeScript the drop-in replaces, so a caller of the
reference finds the same surface:
  GpuUniformsBufferManager   src/lib/gpu/uniforms/GpuUniformsBufferManager.ts:3-183
  GpuMpmBufferManager        src/lib/gpu/mpm/GpuMpmBufferManager.ts:2-81
  GpuColliderBufferManager   src/lib/gpu/collider/GpuColliderBufferManager.ts:42-351
  GpuMpmPipelineManager      src/lib/gpu/mpm/GpuMpmPipelineManager.ts:20-521
  MpmSimulation              the simulation half of GpuSnowPipelineRunner.svelte.ts:68-445
All GPU work happens in the HIP library; these classes pack bytes and call the C ABI.
"""
import ctypes
import numpy as np

from . import _lib
from ._lib import check, METHOD_EXPLICIT, METHOD_PBMPM, UNIFORMS_BYTES
from .particles import PARTICLE_DTYPE


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class GpuUniformsBufferManager:
    """Byte-offset writers of the 37,376-byte uniform buffer (GpuUniformsBufferManager.ts:54-182)."""

    def __init__(self, ctx_handle, lib):
        self._ctx = ctx_handle
        self._lib = lib

    def _write(self, offset, arr):
        a = np.ascontiguousarray(arr)
        check(self._lib, self._ctx, self._lib.mpm_write_uniforms(self._ctx, offset, _ptr(a), a.nbytes))

    def _f(self, offset, vals):
        self._write(offset, np.asarray(vals, dtype=np.float32).ravel())

    def _u(self, offset, vals):
        self._write(offset, np.asarray(vals, dtype=np.uint32).ravel())

    def read(self, offset, nbytes):
        out = np.zeros(nbytes, np.uint8)
        check(self._lib, self._ctx, self._lib.mpm_read_uniforms(self._ctx, offset, _ptr(out), nbytes))
        return out

    def writeSimulationTimestepS(self, timestep): self._f(0, [timestep])
    def writeFixedPointScale(self, scale): self._f(4, [scale])
    def writeUsePbmpm(self, use_pbmpm): self._u(8, [1 if use_pbmpm else 0])
    def writeTime(self, time): self._u(12, [int(time) & 0xFFFFFFFF])
    def writeGridMinCoords(self, mn): self._f(16, mn)
    def writeGridMaxCoords(self, mx): self._f(32, mx)
    def writeViewProjMat(self, m): self._f(48, m)
    def writeViewProjInvMat(self, m): self._f(112, m)
    def writeMeshMinCoords(self, mn): self._f(176, mn)
    def writeMeshMaxCoords(self, mx): self._f(192, mx)
    def writeGridResolution(self, res): self._u(208, res)
    def writeColliderMinCoords(self, mn): self._f(224, mn)
    def writeColliderMaxCoords(self, mx): self._f(240, mx)
    def writeColliderTransformMat(self, m): self._f(256, m)
    def writeColliderVel(self, v): self._f(320, v)
    def writeCameraPos(self, p): self._f(336, p)
    def writeGridCellDims(self, d): self._f(352, d)
    def writeColliderNumIndices(self, n): self._u(348, [n])
    def writeColliderNumObjects(self, n): self._u(364, [n])
    def writeLightViewProjMat(self, m): self._f(368, m)
    def writeColliderTransformInv(self, m): self._f(432, m)
    def writeInteractionPos(self, p): self._f(496, p)
    def writeInteractionStrength(self, s): self._f(508, [s])
    def writeInteractionRadius(self, r): self._f(512, [r])
    def writeIsInteracting(self, b): self._u(516, [1 if b else 0])
    def writeInteractionDir(self, d): self._f(528, d)
    def writeInteractionMode(self, m): self._u(520, [m])

    def writeColliderObjects(self, objects):
        count = min(len(objects), 1024)
        buf = np.zeros(count * 8, np.uint32)
        f = buf.view(np.float32)
        for i, o in enumerate(objects[:count]):
            f[8 * i:8 * i + 3] = o["min"]
            buf[8 * i + 3] = o["startIndex"]
            f[8 * i + 4:8 * i + 7] = o["max"]
            buf[8 * i + 7] = o["countIndices"]
        if count:
            self._write(544, buf)


def build_bvh(positions, indices):
    """buildBvh on the host (GpuColliderBufferManager.ts:222-350) via mpm_build_bvh.
    Returns (nodes uint32 (n,8), triangle leaf order)."""
    lib = _lib.load_library()
    pos = np.ascontiguousarray(positions, dtype=np.float32).ravel()
    idx = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
    nt = idx.size // 3
    cap = max(2 * nt, 1)
    nodes = np.zeros((cap, 8), np.uint32)
    order = np.zeros(max(nt, 1), np.uint32)
    n = ctypes.c_size_t(0)
    rc = lib.mpm_build_bvh(_ptr(pos), pos.size // 3, _ptr(idx), idx.size, _ptr(nodes), cap, ctypes.byref(n), _ptr(order))
    if rc != 0:
        raise _lib.MpmError(rc, "mpm_build_bvh")
    return nodes[:n.value].copy(), order[:nt].copy()


class GpuColliderBufferManager:
    """Packs collider data as colliderDataBuffer does ([indices][vertices][normals][uvs][materials],
    GpuColliderBufferManager.ts:96-126), builds the BVH and uploads both."""

    def __init__(self, vertices, indices, normals=(), uvs=(), materialIndices=()):
        v = np.asarray(vertices, np.float32).reshape(-1, 3)
        self.minCoords = v.min(axis=0).tolist() if len(v) else [np.inf] * 3
        self.maxCoords = v.max(axis=0).tolist() if len(v) else [-np.inf] * 3
        idx = np.asarray(indices, np.uint32).ravel()
        self.numIndices = int(idx.size)
        self.bvhNodes, tri_order = build_bvh(v, idx)
        self.numBvhNodes = int(self.bvhNodes.shape[0])
        reordered = idx.reshape(-1, 3)[tri_order].ravel() if idx.size else idx
        parts = [reordered.astype(np.uint32), v.ravel().view(np.uint32),
                 np.asarray(normals, np.float32).ravel().view(np.uint32),
                 np.asarray(uvs, np.float32).ravel().view(np.uint32),
                 np.asarray(materialIndices, np.uint32).ravel()]
        self.colliderData = np.concatenate(parts).astype(np.uint32)
        self.indicesOffset = 0
        self.verticesOffset = 4 * idx.size

    def upload(self, sim):
        lib, ctx = sim.lib, sim.ctx
        check(lib, ctx, lib.mpm_set_collider(ctx, _ptr(self.colliderData), self.colliderData.size, self.numIndices,
                                             _ptr(self.bvhNodes), self.numBvhNodes))


class GpuMeshBufferManager:
    """src/lib/gpu/particleInitialize/GpuMeshBufferManager.ts:7-45: the scatter mesh as
    consecutive triangle vertices, 16 bytes each, and its bounding box."""

    def __init__(self, vertices):
        v = np.asarray(vertices, np.float64).reshape(-1, 3)
        self.numVertices = v.shape[0]
        self.minCoords = v.min(axis=0).tolist() if len(v) else [np.inf] * 3
        self.maxCoords = v.max(axis=0).tolist() if len(v) else [-np.inf] * 3
        self.meshVertices = np.zeros((v.shape[0], 4), np.float32)
        self.meshVertices[:, :3] = v


def load_glb(data):
    """loadGltfScene (loadScene.ts:31-251) for a .glb file's bytes, in the library's C++ reader:
    dict(positions (V,3), normals (Nn,3), uvs (V,2), materialIndices (V,), indices (I,),
    objects [dict(min, max, startIndex, countIndices)], vertices (I,3) = the scatter mesh)."""
    lib = _lib.load_library()
    buf = np.frombuffer(bytes(data), np.uint8)
    a = _lib.MpmSceneArrays()
    err = ctypes.create_string_buffer(512)
    rc = lib.mpm_load_glb(_ptr(buf), buf.size, ctypes.byref(a), err, 512)
    if rc != 0:
        raise _lib.MpmError(rc, err.value.decode())
    pos = np.zeros((a.n_vertices, 3), np.float32)
    uv = np.zeros((a.n_vertices, 2), np.float32)
    mat = np.zeros(a.n_vertices, np.uint32)
    nor = np.zeros((a.n_normals, 3), np.float32)
    idx = np.zeros(a.n_indices, np.uint32)
    obj = np.zeros((a.n_objects, 8), np.uint32)
    soup = np.zeros((a.n_soup, 4), np.float32)
    b = _lib.MpmSceneArrays(_ptr(pos), _ptr(uv), _ptr(mat), a.n_vertices, _ptr(nor), a.n_normals, _ptr(idx), a.n_indices,
                            _ptr(obj), a.n_objects, _ptr(soup), a.n_soup)
    rc = lib.mpm_load_glb(_ptr(buf), buf.size, ctypes.byref(b), err, 512)
    if rc != 0:
        raise _lib.MpmError(rc, err.value.decode())
    objects = [dict(min=obj[o, 0:3].view(np.float32).tolist(), startIndex=int(obj[o, 3]),
                    max=obj[o, 4:7].view(np.float32).tolist(), countIndices=int(obj[o, 7])) for o in range(a.n_objects)]
    return dict(positions=pos, normals=nor, uvs=uv, materialIndices=mat, indices=idx, objects=objects, vertices=soup[:, :3].copy())


def triangle_soup(vertices, indices):
    """Indexed mesh -> consecutive triangle vertices (the scatter mesh layout)."""
    v = np.asarray(vertices, np.float64).reshape(-1, 3)
    return v[np.asarray(indices, np.int64).ravel()]


class GpuMpmBufferManager:
    """Creation-time sizing (GpuMpmBufferManager.ts:12-13): nParticles plus the sparse-grid
    capacities the library derived."""

    def __init__(self, nParticles, max_blocks, hash_map_size):
        self.nParticles = nParticles
        self.nMaxBlocksInHashMap = max_blocks
        self.hashMapSize = hash_map_size


class GpuMpmPipelineManager:
    """Per-substep dispatch schedules (GpuMpmPipelineManager.ts:327-521), one mpm_step each."""

    def __init__(self, sim):
        self._sim = sim

    def addExplicitMpmDispatches(self, nParticles=None, nBlocksInHashMap=None, hashMapSize=None):
        self._sim.step(METHOD_EXPLICIT, 1)

    def addPbmpmDispatches(self, nParticles=None, nBlocksInHashMap=None, hashMapSize=None):
        self._sim.step(METHOD_PBMPM, 1)


class MpmSimulation:
    """Headless simulation runner over one libmpm_hip context.

    Mirrors the simulation half of GpuSnowPipelineRunner (constructor uniform writes
    :133-144/:171-177, updateColliderTransformMat/Vel :383-390, addSimulationStepsComputePass
    :392-445) without any rendering."""

    def __init__(self, n_particles, grid_res, grid_min=(-5, -5, -5), grid_max=(5, 5, 5), dt=1 / 384, **cfg):
        self.lib = _lib.load_library()
        self.config = _lib.default_config(n_particles=n_particles, grid_res=list(grid_res),
                                          grid_min=list(grid_min), grid_max=list(grid_max), dt=dt, **cfg)
        h = ctypes.c_void_p()
        rc = self.lib.mpm_create(ctypes.byref(self.config), ctypes.byref(h))
        if rc != 0:
            raise _lib.MpmError(rc, "mpm_create")
        self.ctx = h
        self.nParticles = n_particles
        self.uniformsManager = GpuUniformsBufferManager(self.ctx, self.lib)
        self.pipelineManager = GpuMpmPipelineManager(self)
        self.colliderManager = None

    def close(self):
        if self.ctx:
            self.lib.mpm_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- collider (GpuSnowPipelineRunner.svelte.ts:162-177, 383-390)
    def set_collider(self, vertices, indices, transform=None):
        cm = GpuColliderBufferManager(vertices, indices)
        cm.upload(self)
        um = self.uniformsManager
        um.writeColliderMinCoords(cm.minCoords)
        um.writeColliderMaxCoords(cm.maxCoords)
        um.writeColliderNumObjects(1)
        um.writeColliderObjects([{"min": cm.minCoords, "max": cm.maxCoords, "startIndex": 0, "countIndices": cm.numIndices}])
        self.updateColliderTransformMat(np.eye(4, dtype=np.float32) if transform is None else transform)
        um.writeColliderVel([0.0, 0.0, 0.0])
        self.colliderManager = cm
        return cm

    def updateColliderTransformMat(self, mat4):
        m = np.asarray(mat4, np.float64).reshape(4, 4)
        # column-major storage like wgpu-matrix: element (row r, col c) at [4c + r]
        self.uniformsManager.writeColliderTransformMat(m.T.ravel().astype(np.float32))
        self.uniformsManager.writeColliderTransformInv(np.linalg.inv(m).T.ravel().astype(np.float32))

    def updateColliderVel(self, v):
        self.uniformsManager.writeColliderVel(v)

    def set_particle_youngs(self, E):
        """Per-particle Young's modulus for the explicit method (None: back to the constant)."""
        if E is None:
            check(self.lib, self.ctx, self.lib.mpm_set_particle_youngs(self.ctx, None, 0))
            return
        e = np.ascontiguousarray(E, np.float32)
        check(self.lib, self.ctx, self.lib.mpm_set_particle_youngs(self.ctx, _ptr(e), e.size))

    # ---- particle init (GpuSnowPipelineRunner.svelte.ts:158-160, 368-381)
    def set_mesh(self, vertices):
        """The scatter mesh: consecutive triangle vertices (see triangle_soup)."""
        self.meshManager = GpuMeshBufferManager(vertices)
        self.uniformsManager.writeMeshMinCoords(self.meshManager.minCoords)
        self.uniformsManager.writeMeshMaxCoords(self.meshManager.maxCoords)
        return self.meshManager

    def scatterParticlesInMeshVolume(self):
        mv = np.ascontiguousarray(self.meshManager.meshVertices)
        check(self.lib, self.ctx, self.lib.mpm_scatter_particles(self.ctx, _ptr(mv), mv.shape[0]))

    # ---- render-side splats (mcDensity.cs.wgsl, calculateGridMass.wgsl)
    def splat_density(self, res):
        """Marching-cubes density grid, shape (rz, ry, rx) of u32 fixed-point mass."""
        r = np.asarray(res, np.uint32)
        out = np.zeros(int(r[0]) * int(r[1]) * int(r[2]), np.uint32)
        check(self.lib, self.ctx, self.lib.mpm_splat(self.ctx, 0, _ptr(r), _ptr(out), out.size))
        return out.reshape(int(r[2]), int(r[1]), int(r[0]))

    def splat_mass(self):
        """Volumetric mass grid on the simulation grid resolution, shape (rz, ry, rx)."""
        r = self.uniformsManager.read(208, 12).view(np.uint32)
        out = np.zeros(int(r[0]) * int(r[1]) * int(r[2]), np.uint32)
        check(self.lib, self.ctx, self.lib.mpm_splat(self.ctx, 1, None, _ptr(out), out.size))
        return out.reshape(int(r[2]), int(r[1]), int(r[0]))

    # ---- particles
    def upload(self, particles):
        p = np.ascontiguousarray(particles)
        assert p.dtype == PARTICLE_DTYPE and p.shape[0] == self.nParticles
        check(self.lib, self.ctx, self.lib.mpm_upload_particles(self.ctx, _ptr(p), p.shape[0]))

    def download(self):
        out = np.zeros(self.nParticles, PARTICLE_DTYPE)
        check(self.lib, self.ctx, self.lib.mpm_download_particles(self.ctx, _ptr(out), self.nParticles))
        return out

    # ---- stepping
    def step(self, method, n_substeps=1):
        check(self.lib, self.ctx, self.lib.mpm_step(self.ctx, int(method), int(n_substeps)))

    def prepare_graphs(self, method):
        """Captures the step's graphs (one substep and a run of 8, both buffer parities) now."""
        check(self.lib, self.ctx, self.lib.mpm_prepare_graphs(self.ctx, int(method)))

    def sync(self):
        check(self.lib, self.ctx, self.lib.mpm_sync(self.ctx))

    def debug_stage(self, method, stage):
        check(self.lib, self.ctx, self.lib.mpm_debug_stage(self.ctx, int(method), int(stage)))

    def dump_grid(self):
        n = ctypes.c_size_t(0)
        check(self.lib, self.ctx, self.lib.mpm_dump_grid(self.ctx, None, None, 0, ctypes.byref(n)))
        xyz = np.zeros((n.value, 3), np.int32)
        vals = np.zeros((n.value, 4), np.int32)
        check(self.lib, self.ctx, self.lib.mpm_dump_grid(self.ctx, _ptr(xyz), _ptr(vals), n.value, ctypes.byref(n)))
        return xyz, vals

    def dump_blocks(self):
        n = ctypes.c_size_t(0)
        check(self.lib, self.ctx, self.lib.mpm_dump_blocks(self.ctx, None, 0, ctypes.byref(n)))
        xyz = np.zeros((n.value, 3), np.int32)
        check(self.lib, self.ctx, self.lib.mpm_dump_blocks(self.ctx, _ptr(xyz), n.value, ctypes.byref(n)))
        return xyz

    def dump_particle_blocks(self):
        n = ctypes.c_size_t(0)
        check(self.lib, self.ctx, self.lib.mpm_dump_particle_blocks(self.ctx, None, None, 0, ctypes.byref(n)))
        xyz = np.zeros((n.value, 3), np.int32)
        cnt = np.zeros(n.value, np.uint32)
        check(self.lib, self.ctx, self.lib.mpm_dump_particle_blocks(self.ctx, _ptr(xyz), _ptr(cnt), n.value, ctypes.byref(n)))
        return xyz, cnt

    def stats(self):
        s = _lib.MpmStats()
        check(self.lib, self.ctx, self.lib.mpm_get_stats(self.ctx, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def set_timing(self, on=True):
        check(self.lib, self.ctx, self.lib.mpm_set_timing(self.ctx, 1 if on else 0))

    def timings(self):
        t = _lib.MpmTimings()
        check(self.lib, self.ctx, self.lib.mpm_get_timings(self.ctx, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_kernel_timing(self, on=True):
        check(self.lib, self.ctx, self.lib.mpm_set_kernel_timing(self.ctx, 1 if on else 0))

    def kernel_times(self):
        """{kernel name: (total ms, launches)} accumulated since set_kernel_timing(True)."""
        ms = np.zeros(_lib.KERNEL_COUNT, np.float64)
        cnt = np.zeros(_lib.KERNEL_COUNT, np.uint64)
        check(self.lib, self.ctx, self.lib.mpm_get_kernel_times(self.ctx, _ptr(ms), _ptr(cnt)))
        return {self.lib.mpm_kernel_name(i).decode(): (float(ms[i]), int(cnt[i]))
                for i in range(_lib.KERNEL_COUNT) if cnt[i]}

    def uniform_bytes(self):
        return self.uniformsManager.read(0, UNIFORMS_BYTES)

    # ---- PBMPM iteration count and the learned warm start (no reference counterpart) ----
    def set_iterations(self, k, force_divisor=None):
        """nSolveConstraintIterations (GpuMpmPipelineManager.ts:445) and the per-iteration force
        divisor (pbmpm.wgsl:137), which the reference keeps equal to it."""
        fd = float(k if force_divisor is None else force_divisor)
        check(self.lib, self.ctx, self.lib.mpm_set_iterations(self.ctx, int(k), fd))

    def learn_collect(self, n_substeps, k_ref, every=1, max_pairs=None, k_run=None):
        """Advances n_substeps PBMPM substeps with k_run iterations (default k_ref), logging
        (features, target) pairs every `every` substeps; targets from k_ref-iteration substeps
        run from the same start states."""
        n = ctypes.c_uint64()
        mp = self.nParticles * max(1, (n_substeps + every - 1) // every) if max_pairs is None else max_pairs
        kr = int(k_ref if k_run is None else k_run)
        check(self.lib, self.ctx, self.lib.mpm_learn_collect(self.ctx, int(n_substeps), kr, int(k_ref), int(every), int(mp), ctypes.byref(n)))
        return n.value

    def learn_train(self, steps, batch=4096, lr=1e-3, seed=1):
        loss = ctypes.c_float()
        check(self.lib, self.ctx, self.lib.mpm_learn_train(self.ctx, int(steps), int(batch), float(lr), int(seed), ctypes.byref(loss)))
        return loss.value

    def learn_params(self):
        p = np.zeros(_lib.MLP_NPARAM, np.float32)
        nm = np.zeros(64, np.float32)
        check(self.lib, self.ctx, self.lib.mpm_learn_get_params(self.ctx, _ptr(p), _ptr(nm)))
        return p, nm

    def set_learn_params(self, params, norm):
        p = np.ascontiguousarray(params, np.float32)
        nm = np.ascontiguousarray(norm, np.float32)
        if p.size != _lib.MLP_NPARAM or nm.size != 64:
            raise ValueError("expected 6288 parameters and 64 standardisation values")
        check(self.lib, self.ctx, self.lib.mpm_learn_set_params(self.ctx, _ptr(p), _ptr(nm)))

    def learn_dprev(self, dprev=None):
        """The warm start's input dprev (previous substep's last per-iteration change of dx) by
        particle id: sets it when given, returns the current values (n, 3)."""
        out = np.zeros((self.nParticles, 3), np.float32)
        src = None if dprev is None else np.ascontiguousarray(dprev, np.float32).reshape(self.nParticles, 3)
        check(self.lib, self.ctx, self.lib.mpm_learn_dprev(self.ctx, None if src is None else _ptr(src), _ptr(out), self.nParticles))
        return out if src is None else src.copy()

    def learn_enable(self, on=True, limit=8.0):
        check(self.lib, self.ctx, self.lib.mpm_learn_enable(self.ctx, 1 if on else 0, float(limit)))

    def set_residual_tracking(self, on=True):
        check(self.lib, self.ctx, self.lib.mpm_set_residual_tracking(self.ctx, 1 if on else 0))

    def residuals(self):
        """(mean, peak) per PBMPM iteration of max_particles |dx_i - dx_{i-1} - g dt^2 / k|_2 / domain extent,
        over the substeps since tracking was enabled (entry 0 is unused), and that substep count."""
        mean = np.zeros(64, np.float32)
        peak = np.zeros(64, np.float32)
        ns = ctypes.c_uint64()
        check(self.lib, self.ctx, self.lib.mpm_get_residuals(self.ctx, _ptr(mean), _ptr(peak), 64, ctypes.byref(ns)))
        return mean, peak, ns.value


class MpmEnsemble:
    """Many independent scenes stepped together by one libmpm_hip context (no reference
    counterpart: the reference runs one scene). Each scene is a dict of MpmSimulation's keyword
    arguments plus n_particles and method (METHOD_EXPLICIT / METHOD_PBMPM). Scenes have no collider
    and no pointer interaction. step(n) advances every scene by n of its own substeps; each scene's
    particles come out exactly as if that scene were stepped alone."""

    def __init__(self, scenes, **common):
        self.lib = _lib.load_library()
        S = len(scenes)
        cfgs = (_lib.MpmConfig * S)()
        methods = (ctypes.c_int * S)()
        self.sizes = []
        for i, sc in enumerate(scenes):
            kw = dict(sc)
            methods[i] = int(kw.pop("method"))
            n = int(kw.pop("n_particles"))
            kw.setdefault("grid_min", (-5, -5, -5))
            kw.setdefault("grid_max", (5, 5, 5))
            kw.update(common)
            c = _lib.default_config(n_particles=n, grid_res=list(kw.pop("grid_res")), grid_min=list(kw.pop("grid_min")),
                                    grid_max=list(kw.pop("grid_max")), **kw)
            cfgs[i] = c
            self.sizes.append(n)
        h = ctypes.c_void_p()
        rc = self.lib.mpm_ensemble_create(ctypes.cast(cfgs, ctypes.c_void_p), ctypes.cast(methods, ctypes.c_void_p), S, ctypes.byref(h))
        if rc != 0:
            raise _lib.MpmError(rc, "mpm_ensemble_create")
        self.ctx = h
        self.n_scenes = S

    def close(self):
        if self.ctx:
            self.lib.mpm_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene, particles):
        p = np.ascontiguousarray(particles)
        assert p.dtype == PARTICLE_DTYPE and p.shape[0] == self.sizes[scene]
        check(self.lib, self.ctx, self.lib.mpm_ensemble_upload(self.ctx, int(scene), _ptr(p), p.shape[0]))

    def download(self, scene):
        out = np.zeros(self.sizes[scene], PARTICLE_DTYPE)
        check(self.lib, self.ctx, self.lib.mpm_ensemble_download(self.ctx, int(scene), _ptr(out), out.shape[0]))
        return out

    def step(self, n_substeps=1):
        check(self.lib, self.ctx, self.lib.mpm_ensemble_step(self.ctx, int(n_substeps)))

    def sync(self):
        check(self.lib, self.ctx, self.lib.mpm_sync(self.ctx))

    def set_tuning(self, chunk_particles=1024, workgroups_per_cu=4, persistent=False):
        t = _lib.MpmTuning(chunk_particles, workgroups_per_cu, 1 if persistent else 0, 0.0)
        check(self.lib, self.ctx, self.lib.mpm_ensemble_set_tuning(self.ctx, ctypes.byref(t)))

    def tuning(self):
        t = _lib.MpmTuning()
        check(self.lib, self.ctx, self.lib.mpm_ensemble_get_tuning(self.ctx, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def autotune(self, trial_substeps=3):
        """Times every candidate tuning on the live scenes (state restored) and keeps the
        fastest: (best, [all candidates with ms_per_substep])."""
        cap = 64
        arr = (_lib.MpmTuning * cap)()
        n = ctypes.c_int()
        best = _lib.MpmTuning()
        check(self.lib, self.ctx, self.lib.mpm_ensemble_autotune(self.ctx, int(trial_substeps), ctypes.cast(arr, ctypes.c_void_p),
                                                                 cap, ctypes.byref(n), ctypes.byref(best)))
        as_dict = lambda t: {k: getattr(t, k) for k, _ in t._fields_}
        return as_dict(best), [as_dict(arr[i]) for i in range(min(n.value, cap))]

    def stats(self):
        s = _lib.MpmStats()
        check(self.lib, self.ctx, self.lib.mpm_get_stats(self.ctx, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def set_kernel_timing(self, on=True):
        check(self.lib, self.ctx, self.lib.mpm_set_kernel_timing(self.ctx, 1 if on else 0))

    def kernel_times(self):
        ms = np.zeros(_lib.KERNEL_COUNT, np.float64)
        cnt = np.zeros(_lib.KERNEL_COUNT, np.uint64)
        check(self.lib, self.ctx, self.lib.mpm_get_kernel_times(self.ctx, _ptr(ms), _ptr(cnt)))
        return {self.lib.mpm_kernel_name(i).decode(): (float(ms[i]), int(cnt[i]))
                for i in range(_lib.KERNEL_COUNT) if cnt[i]}


def mlp_grad(params, norm, feats, targs):
    """MFMA forward + backward of the warm-start network over the given samples (count a
    multiple of 16): (gradient [6288], standardised outputs [n][16], sum of squared errors)."""
    lib = _lib.load_library()
    p = np.ascontiguousarray(params, np.float32)
    nm = np.ascontiguousarray(norm, np.float32)
    f = np.ascontiguousarray(feats, np.float32).reshape(-1, 16)
    t = np.ascontiguousarray(targs, np.float32).reshape(-1, 16)
    n = f.shape[0]
    g = np.zeros(_lib.MLP_NPARAM, np.float32)
    y = np.zeros((n, 16), np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib.mpm_mlp_grad(_ptr(p), _ptr(nm), _ptr(f), _ptr(t), n, _ptr(g), _ptr(y), _ptr(loss))
    if rc != 0:
        raise RuntimeError(f"mpm_mlp_grad failed ({rc})")
    return g, y, float(loss[0])
