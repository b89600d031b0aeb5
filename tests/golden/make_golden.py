"""Writes the golden fixtures in tests/golden/ from the oracle (the FP32-faithful CPU restatement
of the reference substep). Run from the repository root:

    python tests/golden/make_golden.py

Two kinds of fixture:
  * <case>.npz, small seeded scenes (4,096 particles) stored whole: the input particle state, the
    uniform bytes, the collider words and BVH nodes, the oracle's parameters, the final particle
    state after the substeps (37 FP32 words per particle: pos vel mass F_e F_p dx D), and the
    cell-keyed grid of the first substep's last transfer (explicit: after the grid update; PBMPM:
    after the last iteration's P2G), as allocated block coordinates and every cell with a nonzero
    i32 accumulator.
  * box_glb.npz, the reference's box collider asset (static/box.glb) as the library's GLB reader
    returns it (positions, indices, normals), for GPU runs, where the reference tree is absent.
  * digests.json, larger seeded scenes (65,536 particles) stored as SHA-256 digests of the final
    particle state and of the cell-keyed grid, with the inputs regenerated from their seeds
    (the input digest is stored too, so a changed generator is detected).

The scene generators are the package's seeded ones (scenes.py); the collider words and BVH come
from the host-side BVH builder (mpm_build_bvh, no GPU needed).
"""
import hashlib
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import __graft_entry__  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from helpers import uniforms_blob, SHARED_KEYS  # noqa: E402

HASH_MAP_SIZE, MAX_BLOCKS = 200_003, 70_000

# name -> (scene, particles, seed, grid override, substeps)
WHOLE = {
    "explicit_cube_4096": ("explicit_small", 4096, 0, None, 2),
    "pbmpm_spheres_4096": ("pbmpm_two_spheres", 4096, 7, (32, 32, 32), 2),
}
DIGEST = {
    "explicit_cube_65536": ("explicit_small", None, 0, None, 2),
    "pbmpm_spheres_65536": ("pbmpm_two_spheres", 65536, 5, (64, 64, 64), 2),
}


def setup(name, spec):
    """(particles, uniform bytes, collider words, BVH node words, n_indices, oracle kwargs, meta)."""
    m = __graft_entry__.load_package()
    scene, n, seed, grid, substeps = spec
    sc = m.scenes.scene(scene, n_particles=n, seed=seed)
    sim = dict(sc["sim"])
    if grid is not None:
        sim["grid_res"] = grid
    p = sc["particles"]
    words = np.zeros(0, np.uint32)
    nodes = np.zeros((0, 8), np.uint32)
    coll = None
    if sc["collider"] is not None:
        cm = m.GpuColliderBufferManager(*sc["collider"])
        words, nodes = cm.colliderData, cm.bvhNodes
        coll = (cm.minCoords, cm.maxCoords, cm.numIndices, None)
    uni = uniforms_blob(np.float32(sim["dt"]), sim["grid_min"], sim["grid_max"], sim["grid_res"], collider=coll)
    okw = {k: v for k, v in sim.items() if k in SHARED_KEYS}
    okw.update(hash_map_size=HASH_MAP_SIZE, max_blocks=MAX_BLOCKS)
    meta = dict(case=name, scene=scene, n_particles=int(p.shape[0]), seed=seed, method=sc["method"], substeps=substeps,
                sim={k: (list(v) if isinstance(v, tuple) else v) for k, v in sim.items()},
                oracle_params={k: (list(v) if isinstance(v, tuple) else v) for k, v in okw.items()},
                n_indices=int(coll[2]) if coll else 0)
    return p, uni, words, nodes, okw, meta


def state37(p):
    m = __graft_entry__.load_package()
    return m.state_columns(p).astype(np.float32)


def grid_stage(meta):
    """Substep prefix whose grid is stored: explicit 2 (P2G + grid update), PBMPM k (the grid
    after the last iteration's P2G)."""
    return 2 if meta["method"] == 0 else int(meta["oracle_params"]["pbmpm_iterations"])


def run_oracle(p, uni, words, nodes, okw, method, substeps, stage):
    def make():
        o = orc.OracleSim(p, uni, **okw)
        if words.size:
            o.set_collider(words, nodes)
        return o
    o = make()
    o.step(method, substeps)
    out = o.particles()
    o.close()
    o = make()
    o.run_stage(method, stage)
    xyz, vals = o.dump_grid()
    blocks, _ = o.active_blocks()
    o.close()
    return out, xyz, vals.astype(np.int64).astype(np.int32), blocks


def keyed_grid(xyz, vals):
    order = np.lexsort((xyz[:, 0], xyz[:, 1], xyz[:, 2]))
    return xyz[order].astype(np.int32), vals[order].astype(np.int32)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def grid_digest(xyz, vals):
    kx, kv = keyed_grid(xyz, vals)
    return digest(np.concatenate([kx, kv], axis=1))


def box_glb_fixture():
    """The reference's static/box.glb (the box collider asset) read by the library's GLB reader:
    positions, indices and normals, stored so GPU runs use the asset without the reference tree."""
    src = pathlib.Path("/root/reference/static/box.glb")
    if not src.exists():
        print("box_glb: reference asset not present, fixture kept as is")
        return
    m = __graft_entry__.load_package()
    d = m.load_glb(src.read_bytes())
    np.savez_compressed(HERE / "box_glb.npz", positions=d["positions"], indices=d["indices"], normals=d["normals"],
                        source_sha256=np.frombuffer(hashlib.sha256(src.read_bytes()).hexdigest().encode(), np.uint8))
    print(f"box_glb: {d['positions'].shape[0]} vertices, {d['indices'].size // 3} triangles")


def main():
    box_glb_fixture()
    for name, spec in WHOLE.items():
        p, uni, words, nodes, okw, meta = setup(name, spec)
        meta["grid_stage"] = grid_stage(meta)
        out, xyz, vals, blocks = run_oracle(p, uni, words, nodes, okw, meta["method"], meta["substeps"], meta["grid_stage"])
        kx, kv = keyed_grid(xyz, vals)
        nz = (kv != 0).any(axis=1)
        b = blocks[np.lexsort((blocks[:, 0], blocks[:, 1], blocks[:, 2]))]
        meta["input_sha256"] = digest(state37(p))
        np.savez_compressed(HERE / f"{name}.npz", meta=np.frombuffer(json.dumps(meta).encode(), np.uint8),
                            input_state=state37(p), uniforms=uni, collider_words=words, bvh_nodes=nodes,
                            out_state=state37(out), blocks=b.astype(np.int32),
                            cells=kx[nz].astype(np.int16), cell_vals=kv[nz])
        print(f"{name}: {p.shape[0]} particles, {b.shape[0]} blocks, {int(nz.sum())} nonzero cells")
    digests = {}
    for name, spec in DIGEST.items():
        p, uni, words, nodes, okw, meta = setup(name, spec)
        meta["grid_stage"] = grid_stage(meta)
        out, xyz, vals, blocks = run_oracle(p, uni, words, nodes, okw, meta["method"], meta["substeps"], meta["grid_stage"])
        meta.update(input_sha256=digest(state37(p)), out_state_sha256=digest(state37(out)),
                    grid_sha256=grid_digest(xyz, vals), allocated_cells=int(xyz.shape[0]))
        digests[name] = meta
        print(f"{name}: {p.shape[0]} particles, {xyz.shape[0]} allocated cells")
    (HERE / "digests.json").write_text(json.dumps(digests, indent=1) + "\n")


if __name__ == "__main__":
    main()
