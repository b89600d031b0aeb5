"""GPU parity with colliders that come through the GLB reader (loadScene.ts:27-251 restated in
C++, mpm_load_glb): the reference's own box asset (static/box.glb, stored as the reader's output)
scaled to side 0.2 at the domain centre as the two-sphere scene's collider, and a GLB written here
whose node rotates, scales and translates a cube (the reader's world transform and y/z swap).
PBMPM substeps through the C ABI match the oracle bit for bit with either collider."""
import json
import pathlib
import struct

import numpy as np
import pytest

from helpers import make_pair, assert_particles_equal

pytestmark = pytest.mark.gpu


def _run(mpm, orc, verts, idx, n=6000, seed=31, substeps=4):
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=n, seed=seed)
    sc = dict(sc, collider=(np.asarray(verts, np.float32), np.asarray(idx, np.uint32)))
    sim, o = make_pair(mpm, orc, sc)
    for s in range(substeps):
        sim.step(1, 1)
        o.step(1, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"substep {s + 1}")
    st = sim.stats()
    sim.close()
    return st


def test_reference_box_asset_as_collider_bitexact(mpm, orc):
    z = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "box_glb.npz")
    verts = z["positions"].astype(np.float32) * np.float32(0.2) + np.float32(0.5)
    print("collider: static/box.glb x 0.2 at (0.5, 0.5, 0.5)", flush=True)
    _run(mpm, orc, verts, z["indices"])


def _cube_glb():
    v = np.array([[x, y, z] for z in (-0.5, 0.5) for y in (-0.5, 0.5) for x in (-0.5, 0.5)], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    i = np.array([t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))], np.uint16).ravel()
    binb = v.tobytes() + i.tobytes()
    q = np.array([0.0, np.sin(0.15), 0.0, np.cos(0.15)])    # 0.3 rad about glTF y (the world z after the swap)
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0]}],
           "nodes": [{"mesh": 0, "translation": [0.5, 0.48, 0.52], "rotation": q.tolist(), "scale": [0.22, 0.18, 0.2]}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}],
           "buffers": [{"byteLength": len(binb)}],
           "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": v.nbytes},
                           {"buffer": 0, "byteOffset": v.nbytes, "byteLength": i.nbytes}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 8, "type": "VEC3"},
                         {"bufferView": 1, "componentType": 5123, "count": 36, "type": "SCALAR"}]}
    js = json.dumps(doc).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    bb = binb + b"\0" * ((4 - len(binb) % 4) % 4)
    total = 12 + 8 + len(js) + 8 + len(bb)
    return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js
            + struct.pack("<II", len(bb), 0x004E4942) + bb)


def test_transformed_glb_collider_bitexact(mpm, orc):
    d = mpm.load_glb(_cube_glb())
    pos = d["positions"].astype(np.float64)
    # the node's transform in glTF (y-up) coordinates, then the reader's (x, z, y) swap
    assert np.allclose(pos.mean(0), [0.5, 0.52, 0.48], atol=1e-6)
    print("collider: test-written GLB, rotated / scaled / translated node", flush=True)
    _run(mpm, orc, d["positions"], d["indices"], seed=32)
