"""The committed golden fixtures (tests/golden/, written by make_golden.py from the oracle) against
the oracle and the scene generators: the oracle still produces every stored value bit for bit, and
the seeded generators still produce the stored inputs. The GPU side of the same fixtures is
tests/test_gpu_golden.py."""
import hashlib

import numpy as np
import pytest

from helpers import GOLDEN_WHOLE, load_golden, golden_digests, particles_from_state


def _oracle(orc, g, p):
    meta = g["meta"]
    o = orc.OracleSim(p, g["uniforms"], **{k: (tuple(v) if isinstance(v, list) else v) for k, v in meta["oracle_params"].items()})
    if g["collider_words"].size:
        o.set_collider(g["collider_words"], g["bvh_nodes"])
    return o


@pytest.mark.parametrize("name", GOLDEN_WHOLE)
def test_golden_fixture_reproduced_by_oracle(mpm, orc, name):
    g = load_golden(name)
    meta = g["meta"]
    sc = mpm.scenes.scene(meta["scene"], n_particles=meta["n_particles"], seed=meta["seed"])
    gen = mpm.state_columns(sc["particles"]).astype(np.float32)
    assert np.array_equal(gen.view(np.uint32), g["input_state"].view(np.uint32)), "scene generator changed"
    assert hashlib.sha256(gen.tobytes()).hexdigest() == meta["input_sha256"]
    p = particles_from_state(mpm, g["input_state"])
    o = _oracle(orc, g, p)
    o.step(meta["method"], meta["substeps"])
    out = mpm.state_columns(o.particles()).astype(np.float32)
    assert np.array_equal(out.view(np.uint32), g["out_state"].view(np.uint32))
    o = _oracle(orc, g, p)
    o.run_stage(meta["method"], meta["grid_stage"])
    xyz, vals = o.dump_grid()
    blocks, _ = o.active_blocks()
    assert sorted(map(tuple, blocks.tolist())) == sorted(map(tuple, g["blocks"].tolist()))
    vals = vals.astype(np.int64).astype(np.int32)
    nz = (vals != 0).any(axis=1)
    order = np.lexsort((xyz[nz, 0], xyz[nz, 1], xyz[nz, 2]))
    assert np.array_equal(xyz[nz][order], g["cells"].astype(np.int32))
    assert np.array_equal(vals[nz][order], g["cell_vals"])


@pytest.mark.parametrize("name", sorted(golden_digests()))
def test_golden_digest_reproduced_by_oracle(mpm, orc, name):
    import sys, pathlib
    sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent / "golden"))
    import make_golden as mg
    d = golden_digests()[name]
    spec = next(v for k, v in {**mg.WHOLE, **mg.DIGEST}.items() if k == name)
    p, uni, words, nodes, okw, meta = mg.setup(name, spec)
    assert mg.digest(mg.state37(p)) == d["input_sha256"], "scene generator changed"
    out, xyz, vals, _ = mg.run_oracle(p, uni, words, nodes, okw, d["method"], d["substeps"], d["grid_stage"])
    assert mg.digest(mg.state37(out)) == d["out_state_sha256"]
    assert mg.grid_digest(xyz, vals) == d["grid_sha256"]


def test_box_glb_fixture(mpm):
    """The reference's box collider asset (static/box.glb) as the GLB reader returns it: the closed
    unit cube, 12 triangles, total area 6, every edge shared by two triangles, signed volume -1: the
    asset winds outward and the reader's (x, z, y) swap (loadScene.ts:27-30) is a reflection, which
    reverses the winding. Where the reference tree is present the asset is read again and must give
    the stored arrays bit for bit."""
    import pathlib
    z = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "box_glb.npz")
    v, idx = z["positions"].astype(np.float64), z["indices"].astype(np.int64)
    tri = v[idx.reshape(-1, 3)]
    cr = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert idx.size == 36 and np.allclose(0.5 * np.linalg.norm(cr, axis=1).sum(), 6.0)
    assert np.isclose((tri[:, 0] * cr).sum() / 6.0, -1.0)
    assert np.allclose(np.abs(v), 0.5)
    key = lambda p: tuple(np.round(p, 6))
    edges = {}
    for t in tri:
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = tuple(sorted((key(t[a]), key(t[b]))))
            edges[e] = edges.get(e, 0) + 1
    assert set(edges.values()) == {2}
    src = pathlib.Path("/root/reference/static/box.glb")
    if src.exists():
        d = mpm.load_glb(src.read_bytes())
        assert np.array_equal(d["positions"], z["positions"]) and np.array_equal(d["indices"], z["indices"])
        assert np.array_equal(d["normals"], z["normals"])
