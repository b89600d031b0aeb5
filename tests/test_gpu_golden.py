"""GPU parity against the committed golden fixtures (tests/golden/), with no live oracle: the HIP
path through the C ABI reproduces every stored particle word and every stored grid accumulator of
the 4,096-particle scenes bit for bit, and the SHA-256 digests of the 65,536-particle scenes."""
import ctypes
import pathlib
import sys

import numpy as np
import pytest

from helpers import GOLDEN_WHOLE, load_golden, golden_digests, particles_from_state, keyed

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent / "golden"))
import make_golden as mg  # noqa: E402


def _sim(mpm, meta, p, uniforms, words, nodes):
    kw = {k: (tuple(v) if isinstance(v, list) else v) for k, v in meta["sim"].items()}
    sim = mpm.MpmSimulation(p.shape[0], **kw)
    lib = sim.lib
    if words.size:
        w = np.ascontiguousarray(words, np.uint32)
        nd = np.ascontiguousarray(nodes, np.uint32)
        rc = lib.mpm_set_collider(sim.ctx, w.ctypes.data_as(ctypes.c_void_p), w.size, meta["n_indices"],
                                  nd.ctypes.data_as(ctypes.c_void_p), nd.shape[0])
        assert rc == 0, lib.mpm_last_error(sim.ctx)
    sim.uniformsManager._write(0, np.ascontiguousarray(uniforms, np.uint8))
    sim.upload(p)
    return sim


@pytest.mark.parametrize("name", GOLDEN_WHOLE)
def test_gpu_matches_golden_fixture(mpm, name):
    g = load_golden(name)
    meta = g["meta"]
    p = particles_from_state(mpm, g["input_state"])
    print(f"golden {name}: {p.shape[0]} particles, method {meta['method']}, {meta['substeps']} substeps", flush=True)
    with _sim(mpm, meta, p, g["uniforms"], g["collider_words"], g["bvh_nodes"]) as sim:
        sim.step(meta["method"], meta["substeps"])
        out = mpm.state_columns(sim.download()).astype(np.float32)
    bad = np.argwhere(out.view(np.uint32) != g["out_state"].view(np.uint32))
    assert bad.size == 0, f"{bad.shape[0]} particle words differ from the fixture, first {bad[0]}"
    with _sim(mpm, meta, p, g["uniforms"], g["collider_words"], g["bvh_nodes"]) as sim:
        sim.debug_stage(meta["method"], meta["grid_stage"])
        blocks = sim.dump_blocks()
        xyz, vals = keyed(*sim.dump_grid())
    assert sorted(map(tuple, blocks.tolist())) == sorted(map(tuple, g["blocks"].tolist())), "allocated block sets differ"
    nz = (vals != 0).any(axis=1)
    assert np.array_equal(xyz[nz], g["cells"].astype(np.int32)), "nonzero cell sets differ"
    assert np.array_equal(vals[nz], g["cell_vals"]), "grid accumulators differ"


@pytest.mark.parametrize("name", sorted(golden_digests()))
def test_gpu_matches_golden_digest(mpm, name):
    d = golden_digests()[name]
    spec = {**mg.WHOLE, **mg.DIGEST}[name]
    p, uni, words, nodes, okw, meta = mg.setup(name, spec)
    assert mg.digest(mg.state37(p)) == d["input_sha256"], "scene generator changed"
    print(f"golden digest {name}: {p.shape[0]} particles", flush=True)
    with _sim(mpm, d, p, uni, words, nodes) as sim:
        sim.step(d["method"], d["substeps"])
        assert mg.digest(mg.state37(sim.download())) == d["out_state_sha256"]
    with _sim(mpm, d, p, uni, words, nodes) as sim:
        sim.debug_stage(d["method"], d["grid_stage"])
        xyz, vals = sim.dump_grid()
        assert xyz.shape[0] == d["allocated_cells"]
        assert mg.grid_digest(xyz, vals) == d["grid_sha256"]
