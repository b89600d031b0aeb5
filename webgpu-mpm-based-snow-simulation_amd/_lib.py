"""ctypes binding of libmpm_hip.so (declarations in include/mpm_hip.h).

The product path is the HIP library only: if the shared object is missing this module
raises instead of falling back to anything else.
"""
import ctypes
import os
import pathlib

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libmpm_hip.so"

UNIFORMS_BYTES = 37376
PARTICLE_BYTES = 192
BVH_NODE_BYTES = 32
KERNEL_COUNT = 25
MLP_NPARAM = 6288

MPM_OK = 0
MPM_E_INVALID = -1
MPM_E_HIP = -2
MPM_E_NOMEM = -3
MPM_E_RANGE = -4
MPM_E_STATE = -5

METHOD_EXPLICIT = 0
METHOD_PBMPM = 1


class MpmConfig(ctypes.Structure):
    _fields_ = [
        ("n_particles", ctypes.c_uint32),
        ("grid_res", ctypes.c_uint32 * 3),
        ("grid_min", ctypes.c_float * 3),
        ("grid_max", ctypes.c_float * 3),
        ("fixed_point_scale", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("max_blocks", ctypes.c_uint32),
        ("hash_map_size", ctypes.c_uint32),
        ("max_probes", ctypes.c_uint32),
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
        ("chunk_particles", ctypes.c_uint32),
        ("grid_workgroups", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("use_graph", ctypes.c_uint32),
        ("plasticity_mode", ctypes.c_uint32),
    ]


class MpmStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "n_particles", "particles_binned", "particles_unbinned", "particle_blocks", "work_chunks",
        "blocks_allocated", "blocks_dropped", "probe_failures", "substeps", "mlp_fallbacks", "nonfinite_particles", "svd_fp64_particles", "collider_queued")]


class MpmTimings(ctypes.Structure):
    _fields_ = [
        ("compute_simulation_step_ns", ctypes.c_double),
        ("binning_ns", ctypes.c_double),
        ("transfer_ns", ctypes.c_double),
        ("integrate_ns", ctypes.c_double),
        ("substeps", ctypes.c_uint64),
        ("pass_substeps", ctypes.c_uint64),
    ]


class MpmSceneArrays(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_void_p), ("uvs", ctypes.c_void_p), ("material_indices", ctypes.c_void_p),
        ("n_vertices", ctypes.c_size_t),
        ("normals", ctypes.c_void_p), ("n_normals", ctypes.c_size_t),
        ("indices", ctypes.c_void_p), ("n_indices", ctypes.c_size_t),
        ("objects", ctypes.c_void_p), ("n_objects", ctypes.c_size_t),
        ("soup_xyzw", ctypes.c_void_p), ("n_soup", ctypes.c_size_t),
    ]


class MpmTuning(ctypes.Structure):
    _fields_ = [
        ("chunk_particles", ctypes.c_uint32),
        ("workgroups_per_cu", ctypes.c_uint32),
        ("persistent", ctypes.c_int32),
        ("ms_per_substep", ctypes.c_float),
    ]


# (name, restype, argtypes) of every entry point declared in include/mpm_hip.h
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_SIGS = [
    ("mpm_config_defaults", None, [ctypes.POINTER(MpmConfig)]),
    ("mpm_create", ctypes.c_int, [ctypes.POINTER(MpmConfig), ctypes.POINTER(_P)]),
    ("mpm_destroy", None, [_P]),
    ("mpm_write_uniforms", ctypes.c_int, [_P, _SZ, _P, _SZ]),
    ("mpm_read_uniforms", ctypes.c_int, [_P, _SZ, _P, _SZ]),
    ("mpm_set_collider", ctypes.c_int, [_P, _P, _SZ, ctypes.c_uint32, _P, _SZ]),
    ("mpm_build_bvh", ctypes.c_int, [_P, _SZ, _P, _SZ, _P, _SZ, ctypes.POINTER(_SZ), _P]),
    ("mpm_upload_particles", ctypes.c_int, [_P, _P, _SZ]),
    ("mpm_download_particles", ctypes.c_int, [_P, _P, _SZ]),
    ("mpm_step", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    ("mpm_sync", ctypes.c_int, [_P]),
    ("mpm_prepare_graphs", ctypes.c_int, [_P, ctypes.c_int]),
    ("mpm_debug_stage", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    ("mpm_dump_grid", ctypes.c_int, [_P, _P, _P, _SZ, ctypes.POINTER(_SZ)]),
    ("mpm_dump_blocks", ctypes.c_int, [_P, _P, _SZ, ctypes.POINTER(_SZ)]),
    ("mpm_dump_particle_blocks", ctypes.c_int, [_P, _P, _P, _SZ, ctypes.POINTER(_SZ)]),
    ("mpm_selftest", ctypes.c_int, [ctypes.c_int, _P, _P, _P, _SZ]),
    ("mpm_divisor_fast", ctypes.c_int, [_P, ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_int)]),
    ("mpm_set_iterations", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_float]),
    ("mpm_learn_collect", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    ("mpm_learn_train", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]),
    ("mpm_learn_get_params", ctypes.c_int, [_P, _P, _P]),
    ("mpm_learn_set_params", ctypes.c_int, [_P, _P, _P]),
    ("mpm_learn_enable", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_float]),
    ("mpm_learn_dprev", ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    ("mpm_mlp_grad", ctypes.c_int, [_P, _P, _P, _P, ctypes.c_uint32, _P, _P, _P]),
    ("mpm_set_residual_tracking", ctypes.c_int, [_P, ctypes.c_int]),
    ("mpm_get_residuals", ctypes.c_int, [_P, _P, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
    ("mpm_set_particle_youngs", ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    ("mpm_scatter_particles", ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    ("mpm_load_glb", ctypes.c_int, [_P, ctypes.c_size_t, ctypes.POINTER(MpmSceneArrays), ctypes.c_char_p, ctypes.c_size_t]),
    ("mpm_splat", ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_size_t]),
    ("mpm_splat_grid", _P, [_P]),
    ("mpm_splat_noalloc", ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_size_t]),
    ("mpm_ensemble_create", ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.POINTER(_P)]),
    ("mpm_ensemble_upload", ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_size_t]),
    ("mpm_ensemble_download", ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_size_t]),
    ("mpm_ensemble_step", ctypes.c_int, [_P, ctypes.c_int]),
    ("mpm_ensemble_set_tuning", ctypes.c_int, [_P, ctypes.POINTER(MpmTuning)]),
    ("mpm_ensemble_get_tuning", ctypes.c_int, [_P, ctypes.POINTER(MpmTuning)]),
    ("mpm_ensemble_autotune", ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(MpmTuning)]),
    ("mpm_get_stats", ctypes.c_int, [_P, ctypes.POINTER(MpmStats)]),
    ("mpm_set_timing", ctypes.c_int, [_P, ctypes.c_int]),
    ("mpm_get_timings", ctypes.c_int, [_P, ctypes.POINTER(MpmTimings)]),
    ("mpm_set_kernel_timing", ctypes.c_int, [_P, ctypes.c_int]),
    ("mpm_get_kernel_times", ctypes.c_int, [_P, _P, _P]),
    ("mpm_kernel_name", ctypes.c_char_p, [ctypes.c_int]),
    ("mpm_last_error", ctypes.c_char_p, [_P]),
    ("mpm_version", ctypes.c_char_p, []),
    ("mpm_stream", _P, [_P]),
]
EXPORTED = [s[0] for s in _SIGS]

_lib = None


class MpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmpm_hip error {code}: {msg}")
        self.code = code


def load_library(path=None):
    """Load libmpm_hip.so (built in-tree by build.py). Raises if it is not there."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is missing: build it with __graft_entry__.build() "
                                f"(hipcc --offload-arch=gfx950); there is no fallback path")
    lib = ctypes.CDLL(str(p), mode=os.RTLD_LOCAL if hasattr(os, "RTLD_LOCAL") else 0)
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def default_config(**overrides):
    lib = load_library()
    cfg = MpmConfig()
    lib.mpm_config_defaults(ctypes.byref(cfg))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        cur = getattr(cfg, k)
        if isinstance(cur, ctypes.Array):
            for i, x in enumerate(v):
                cur[i] = x
        else:
            setattr(cfg, k, v)
    return cfg


def check(lib, ctx, rc):
    if rc != MPM_OK:
        msg = lib.mpm_last_error(ctx) if ctx else b""
        raise MpmError(rc, msg.decode() if msg else "")
    return rc
