"""Oracle — TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

ctypes front end of oracle/mpm_oracle.cpp: the CPU restatement of the reference's MPM
substep (FP32-faithful, and an FP64 physics variant). Never imported by the product package.
"""
import ctypes
import os
import pathlib
import subprocess

import numpy as np

ORACLE_DIR = pathlib.Path(__file__).resolve().parent
# MPM_ORACLE_LIB points the checker at another build of these sources (the mutation test
# compiles copies with planted mistakes and checks that the pins reject them).
LIB = pathlib.Path(os.environ.get("MPM_ORACLE_LIB", str(ORACLE_DIR / "_build" / "liborc.so")))


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("pbmpm_iterations", ctypes.c_int32),
        ("pbmpm_force_divisor", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("youngs_modulus", ctypes.c_double),
        ("poissons_ratio", ctypes.c_double),
        ("hardening_explicit", ctypes.c_float),
        ("hardening_pbmpm", ctypes.c_float),
        ("pbmpm_blend", ctypes.c_float),
        ("pbmpm_volume_exponent", ctypes.c_float),
        ("theta_c", ctypes.c_double),
        ("theta_s", ctypes.c_double),
        ("density", ctypes.c_double),
        ("grad_scale", ctypes.c_float),
        ("collider_margin", ctypes.c_float),
        ("friction_ccd", ctypes.c_float),
        ("friction_proximity", ctypes.c_float),
        ("collider_velocity_factor", ctypes.c_float),
        ("wall_mode", ctypes.c_uint32),
        ("wall_cells", ctypes.c_uint32),
        ("hash_map_size", ctypes.c_uint32),
        ("max_blocks", ctypes.c_uint32),
        ("max_probes", ctypes.c_uint32),
        ("plasticity_mode", ctypes.c_uint32),
    ]


class OrcStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("blocks_allocated", "blocks_dropped", "particles_unbinned", "lookups_missed", "svd_fp64")]


_lib = None
_P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        sig = {
            "orc_default_params": (None, [ctypes.POINTER(OrcParams)]),
            "orc_params_size": (ctypes.c_size_t, []),
            "orc_create": (_P, [ctypes.c_int, ctypes.POINTER(OrcParams), _P, ctypes.c_uint32, _P]),
            "orc_destroy": (None, [_P]),
            "orc_set_uniforms": (None, [_P, _P]),
            "orc_set_collider": (None, [_P, _P, ctypes.c_size_t, _P, ctypes.c_size_t]),
            "orc_step": (None, [_P, ctypes.c_int, ctypes.c_int]),
            "orc_run_stage": (None, [_P, ctypes.c_int, ctypes.c_int]),
            "orc_get_particles": (None, [_P, _P]),
            "orc_get_particles_f64": (None, [_P, _P]),
            "orc_set_particles": (None, [_P, _P, ctypes.c_uint32]),
            "orc_set_particle_youngs": (None, [_P, _P, ctypes.c_uint32]),
            "orc_dump_grid": (ctypes.c_size_t, [_P, _P, _P, ctypes.c_size_t]),
            "orc_active_blocks": (ctypes.c_size_t, [_P, _P, _P, ctypes.c_size_t]),
            "orc_get_stats": (None, [_P, ctypes.POINTER(OrcStats)]),
            "orc_hash1": (ctypes.c_uint32, [ctypes.c_uint32]),
            "orc_hash3": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
            "orc_hash4": (ctypes.c_uint32, [ctypes.c_uint32] * 4),
            "orc_exp": (ctypes.c_float, [ctypes.c_float]),
            "orc_log": (ctypes.c_float, [ctypes.c_float]),
            "orc_pow": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
            "orc_f2i": (ctypes.c_int32, [ctypes.c_float]),
            "orc_det3": (ctypes.c_float, [_P]),
            "orc_inverse3": (None, [_P, _P]),
            "orc_polar3": (None, [_P, _P]),
            "orc_mmul3": (None, [_P, _P, _P]),
            "orc_bspline": (None, [_P, _P, _P]),
            "orc_closest_point_triangle": (None, [_P, _P, _P, _P, _P]),
            "orc_intersect_ray_triangle": (ctypes.c_float, [_P, _P, _P, _P, _P]),
            "orc_solve_constraints1": (None, [ctypes.POINTER(OrcParams), _P, _P, _P]),
            "orc_pk1_explicit1": (None, [ctypes.POINTER(OrcParams), _P, _P, _P]),
            "orc_integrate_only": (None, [_P]),
            "orc_g2p_affine": (None, [_P, ctypes.c_int, _P, _P, ctypes.c_float]),
            "orc_set_threads": (ctypes.c_int, [_P, ctypes.c_int]),
            "orc_pbmpm_trace": (None, [_P, _P]),
            "orc_svd3": (None, [_P, ctypes.c_int, _P, _P, _P]),
            "orc_plasticity_svd1": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_int, _P, _P, _P]),
            "orc_noise3": (ctypes.c_float, [ctypes.c_float] * 3),
            "orc_fbm": (ctypes.c_float, [ctypes.c_float] * 3),
            "orc_inside_mesh": (ctypes.c_int, [_P, _P, ctypes.c_uint32]),
            "orc_scatter": (None, [_P, ctypes.c_uint32, _P, _P, ctypes.c_uint32, _P]),
            "orc_splat": (None, [ctypes.c_int, _P, ctypes.c_uint32, _P, _P, ctypes.c_float, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        assert L.orc_params_size() == ctypes.sizeof(OrcParams)
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def params(**kw):
    p = OrcParams()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        cur = getattr(p, k)
        if isinstance(cur, ctypes.Array):
            for i, x in enumerate(v):
                cur[i] = x
        else:
            setattr(p, k, v)
    return p


class OracleSim:
    """One CPU simulation. mode 0: FP32-faithful restatement; mode 1: FP64 twin with the same
    fixed-point grid; mode 2: FP64 physics without fixed-point quantization."""

    def __init__(self, particles, uniforms, fp64=False, mode=None, **param_kw):
        if mode is None:
            mode = 1 if fp64 else 0
        self.n = particles.shape[0]
        self._prm = params(**param_kw)
        self._uni = np.ascontiguousarray(uniforms, np.uint8)
        p = np.ascontiguousarray(particles)
        self.h = lib().orc_create(int(mode), ctypes.byref(self._prm), _ptr(self._uni), self.n, _ptr(p))
        self._dtype = particles.dtype

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particle_youngs(self, E):
        if E is None:
            lib().orc_set_particle_youngs(self.h, None, 0)
            return
        e = np.ascontiguousarray(E, np.float32)
        lib().orc_set_particle_youngs(self.h, _ptr(e), e.size)

    def set_uniforms(self, uniforms):
        self._uni = np.ascontiguousarray(uniforms, np.uint8)
        lib().orc_set_uniforms(self.h, _ptr(self._uni))

    def set_collider(self, data_words, bvh_nodes):
        d = np.ascontiguousarray(data_words, np.uint32)
        b = np.ascontiguousarray(bvh_nodes, np.uint32)
        lib().orc_set_collider(self.h, _ptr(d), d.size, _ptr(b), b.size // 8)

    def step(self, method, n=1):
        lib().orc_step(self.h, int(method), int(n))

    def run_stage(self, method, stage):
        lib().orc_run_stage(self.h, int(method), int(stage))

    def particles(self):
        out = np.zeros(self.n, self._dtype)
        lib().orc_get_particles(self.h, _ptr(out))
        return out

    def state_f64(self):
        out = np.zeros((self.n, 37), np.float64)
        lib().orc_get_particles_f64(self.h, _ptr(out))
        return out

    def set_threads(self, n):
        """OpenMP threads of the passes (FP32-faithful mode only); returns the count in effect."""
        return lib().orc_set_threads(self.h, int(n))

    def pbmpm_trace(self, k):
        """One PBMPM substep (k iterations, as created); dx and D (column-major) of every particle after
        each G2P: [k, n, 12]."""
        out = np.zeros((k, self.n, 12), np.float32)
        lib().orc_pbmpm_trace(self.h, _ptr(out))
        return out

    def integrate_only(self):
        """integrateParticles on every particle with no transfer before it."""
        lib().orc_integrate_only(self.h)

    def g2p_affine(self, method, a, M, mass=1000.0):
        """Binning, every allocated cell set to the affine velocity field a + M c (c = integer cell
        coordinates, M a numpy row-major 3x3) with node mass `mass`, then G2P of every binned particle."""
        av = np.ascontiguousarray(a, np.float32)
        Mc = np.ascontiguousarray(np.asarray(M, np.float32).T.reshape(9))
        lib().orc_g2p_affine(self.h, int(method), _ptr(av), _ptr(Mc), float(mass))

    def dump_grid(self):
        n = lib().orc_dump_grid(self.h, None, None, 0)
        xyz = np.zeros((n, 3), np.int32)
        vals = np.zeros((n, 4), np.float64)
        lib().orc_dump_grid(self.h, _ptr(xyz), _ptr(vals), n)
        return xyz, vals

    def active_blocks(self):
        n = lib().orc_active_blocks(self.h, None, None, 0)
        xyz = np.zeros((n, 3), np.int32)
        cnt = np.zeros(n, np.uint32)
        lib().orc_active_blocks(self.h, _ptr(xyz), _ptr(cnt), n)
        return xyz, cnt

    def stats(self):
        s = OrcStats()
        lib().orc_get_stats(self.h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}


def _colmajor(m):
    return np.ascontiguousarray(np.asarray(m, np.float32).T.reshape(9))


def _rowmajor(c):
    return c.reshape(3, 3).T.astype(np.float64)


def solve_constraints(Fe, Fp, D, **param_kw):
    """One particle through the PBMPM constraint solve (pbmpm.wgsl:16-39), FP32: the new D.
    Matrices as numpy row-major 3x3."""
    prm = params(**param_kw)
    d = _colmajor(D)
    lib().orc_solve_constraints1(ctypes.byref(prm), _ptr(_colmajor(Fe)), _ptr(_colmajor(Fp)), _ptr(d))
    return _rowmajor(d)


def pk1_explicit(Fe, Fp, **param_kw):
    """The explicit method's hardened corotated first Piola-Kirchhoff stress of one particle."""
    prm = params(**param_kw)
    out = np.zeros(9, np.float32)
    lib().orc_pk1_explicit1(ctypes.byref(prm), _ptr(_colmajor(Fe)), _ptr(_colmajor(Fp)), _ptr(out))
    return _rowmajor(out)


def svd3(A, fp64=False):
    """The SVD plasticity option's 3x3 SVD of one matrix (numpy row-major): (sigma[3], V, B = A V),
    V and B row-major with the singular vectors / sigma_j u_j as columns."""
    sig = np.zeros(3, np.float32)
    V = np.zeros(9, np.float32)
    B = np.zeros(9, np.float32)
    lib().orc_svd3(_ptr(_colmajor(A)), 1 if fp64 else 0, _ptr(sig), _ptr(V), _ptr(B))
    return sig.astype(np.float64), _rowmajor(V), _rowmajor(B)


def plasticity_svd(Fe, Fp, lo, hi, fallback=False):
    """The option's plasticity update of one particle: (changed, F_e, F_p, took_fp64), row-major."""
    fe, fp = np.zeros(9, np.float32), np.zeros(9, np.float32)
    flag = ctypes.c_int(0)
    ch = lib().orc_plasticity_svd1(_ptr(_colmajor(Fe)), _ptr(_colmajor(Fp)), float(lo), float(hi), 1 if fallback else 0,
                                   _ptr(fe), _ptr(fp), ctypes.byref(flag))
    return bool(ch), _rowmajor(fe), _rowmajor(fp), bool(flag.value)


def f32_vec(fn, *arrays):
    """Apply a scalar oracle function elementwise over float32 arrays."""
    arrays = [np.asarray(a, np.float32).ravel() for a in arrays]
    out = np.empty(arrays[0].size, np.float32)
    for i in range(out.size):
        out[i] = fn(*(float(a[i]) for a in arrays))
    return out


def scatter(mesh_xyzw, mesh_min, mesh_max, n):
    """scatterParticles (particleScatter.cs.wgsl:40-84) for particle indices [0, n): 192-byte
    records as a float32 array of shape (n, 48). mesh_xyzw: consecutive triangle vertices,
    16 bytes each (the WGSL array<vec3f> stride)."""
    v = np.ascontiguousarray(mesh_xyzw, np.float32).reshape(-1, 4)
    mn = np.ascontiguousarray(mesh_min, np.float32)
    mx = np.ascontiguousarray(mesh_max, np.float32)
    out = np.zeros((n, 48), np.float32)
    lib().orc_scatter(_ptr(v), v.shape[0], _ptr(mn), _ptr(mx), n, _ptr(out))
    return out


def splat(kind, particles, gmin, gmax, S, res):
    """calculateDensity (kind 0) / calculateGridMass (kind 1) of 192-byte records: u32 grid of
    shape (rz, ry, rx)."""
    a = np.ascontiguousarray(particles).view(np.float32).reshape(-1, 48)
    g0 = np.ascontiguousarray(gmin, np.float32)
    g1 = np.ascontiguousarray(gmax, np.float32)
    r = np.ascontiguousarray(res, np.uint32)
    out = np.zeros(int(r[0]) * int(r[1]) * int(r[2]), np.uint32)
    lib().orc_splat(int(kind), _ptr(a), a.shape[0], _ptr(g0), _ptr(g1), float(S), _ptr(r), _ptr(out))
    return out.reshape(int(r[2]), int(r[1]), int(r[0]))
