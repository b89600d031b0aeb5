"""The 192-byte ParticleData record (src/lib/gpu/shaderPrelude/mpmOps.wgsl:1-12) as a numpy dtype.

Offsets: pos @0, _hom @12, vel @16, mass @28, deformationElastic @32 (3 columns x vec4),
deformationPlastic @80, pos_displacement @128 (vec3 + pad), deformation_displacement @144.
Matrices are column-major: field[c, r] is column c, row r; field[c, 3] is padding.
"""
import numpy as np

PARTICLE_DTYPE = np.dtype([
    ("pos", "<f4", (3,)),
    ("_hom", "<f4"),
    ("vel", "<f4", (3,)),
    ("mass", "<f4"),
    ("Fe", "<f4", (3, 4)),
    ("Fp", "<f4", (3, 4)),
    ("dx", "<f4", (4,)),
    ("D", "<f4", (3, 4)),
])
assert PARTICLE_DTYPE.itemsize == 192


def make_particles(pos, vel=None, mass=None):
    """Particle records with F_e = F_p = I and zero displacements (particleScatter.cs.wgsl:77-83)."""
    pos = np.asarray(pos, dtype=np.float32)
    n = pos.shape[0]
    p = np.zeros(n, dtype=PARTICLE_DTYPE)
    p["pos"] = pos
    p["_hom"] = 1.0
    if vel is not None:
        p["vel"] = np.broadcast_to(np.asarray(vel, dtype=np.float32), (n, 3))
    p["mass"] = 1.0 if mass is None else np.asarray(mass, dtype=np.float32)
    eye = np.zeros((3, 4), np.float32)
    eye[0, 0] = eye[1, 1] = eye[2, 2] = 1.0
    p["Fe"] = eye
    p["Fp"] = eye
    return p


def matrix(field):
    """(n, 3, 4) column-major storage -> (n, 3, 3) with [i, r, c] = row r, column c."""
    return np.ascontiguousarray(np.swapaxes(field[..., :3], -1, -2))


def state_columns(p):
    """37 FP32 values per particle: pos3 vel3 mass Fe9 Fp9 dx3 D9 (matrices column-major)."""
    n = p.shape[0]
    out = np.empty((n, 37), np.float64)
    out[:, 0:3] = p["pos"]
    out[:, 3:6] = p["vel"]
    out[:, 6] = p["mass"]
    out[:, 7:16] = p["Fe"][:, :, :3].reshape(n, 9)
    out[:, 16:25] = p["Fp"][:, :, :3].reshape(n, 9)
    out[:, 25:28] = p["dx"][:, :3]
    out[:, 28:37] = p["D"][:, :, :3].reshape(n, 9)
    return out
