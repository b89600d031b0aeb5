"""CPU checks of the glTF binary ingest (loadGltfScene restated in C++, mpm_load_glb): a GLB
written here with a known node hierarchy (TRS and matrix nodes, a two-primitive mesh, u8 / u16 /
u32 indices, normals, float and normalised texture coordinates, textures) must give the arrays the
glTF node semantics and loadScene.ts's conventions fix (world = parent x local, (x, z, y) swap,
unswapped normals by the normal matrix, offset indices, one object per mesh, the scatter mesh)."""
import json
import struct

import numpy as np
import pytest


def _glb(doc, bin_bytes):
    js = json.dumps(doc).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    bb = bin_bytes + b"\0" * ((4 - len(bin_bytes) % 4) % 4)
    total = 12 + 8 + len(js) + 8 + len(bb)
    return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js
            + struct.pack("<II", len(bb), 0x004E4942) + bb)


class _Builder:
    def __init__(self):
        self.bin = bytearray()
        self.views, self.accs = [], []

    def acc(self, arr, ctype, type_, normalized=False):
        a = np.ascontiguousarray(arr)
        while len(self.bin) % 4:
            self.bin.append(0)
        off = len(self.bin)
        self.bin += a.tobytes()
        self.views.append({"buffer": 0, "byteOffset": off, "byteLength": a.nbytes})
        d = {"bufferView": len(self.views) - 1, "componentType": ctype, "count": int(a.shape[0]), "type": type_}
        if normalized:
            d["normalized"] = True
        self.accs.append(d)
        return len(self.accs) - 1


def _quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _scene():
    rng = np.random.default_rng(7)
    b = _Builder()
    box_v = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    box_i = np.array([t for a, bb, c, d in quads for t in ((a, bb, c), (a, c, d))], np.uint16).ravel()
    box_n = (box_v - 0.5) / np.linalg.norm(box_v - 0.5, axis=1, keepdims=True)
    box_uv = rng.random((8, 2)).astype(np.float32)
    tri_v = np.array([[0, 0, 0], [2, 0, 0], [0, 3, 0]], np.float32)
    tri_i = np.array([0, 1, 2], np.uint32)
    tet_v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    tet_i = np.array([0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3], np.uint8)
    tet_n = rng.normal(size=(4, 3)).astype(np.float32)
    tet_uv = rng.integers(0, 256, (4, 2)).astype(np.uint8)
    p0 = {"attributes": {"POSITION": b.acc(box_v, 5126, "VEC3"), "NORMAL": b.acc(box_n.astype(np.float32), 5126, "VEC3"),
                         "TEXCOORD_0": b.acc(box_uv, 5126, "VEC2")}, "indices": b.acc(box_i, 5123, "SCALAR"), "material": 0}
    p1 = {"attributes": {"POSITION": b.acc(tri_v, 5126, "VEC3")}, "indices": b.acc(tri_i, 5125, "SCALAR"), "material": 1}
    p2 = {"attributes": {"POSITION": b.acc(tet_v, 5126, "VEC3"), "NORMAL": b.acc(tet_n, 5126, "VEC3"),
                         "TEXCOORD_0": b.acc(tet_uv, 5121, "VEC2", normalized=True)}, "indices": b.acc(tet_i, 5121, "SCALAR"), "material": 2}
    q = np.array([0.1, 0.2, 0.3, 0.9]); q /= np.linalg.norm(q)
    th = 0.7
    M1 = np.eye(4)
    M1[:3, :3] = [[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]]
    M1[:3, 3] = [0.5, -0.25, 1.0]
    doc = {
        "asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0, 2]}],
        "nodes": [
            {"mesh": 0, "translation": [0.1, 0.2, 0.3], "rotation": q.tolist(), "scale": [1.5, 0.5, 2.0], "children": [1]},
            {"mesh": 1, "matrix": M1.T.ravel().tolist()},
            {"mesh": 1, "scale": [2.0, 3.0, 4.0]},
        ],
        "meshes": [{"primitives": [p0, p1]}, {"primitives": [p2]}],
        "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 1}}}, {},
                      {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}],
        "textures": [{}, {}], "buffers": [{"byteLength": len(b.bin)}], "bufferViews": b.views, "accessors": b.accs,
    }
    W0 = np.eye(4)
    W0[:3, :3] = _quat_matrix(q) @ np.diag([1.5, 0.5, 2.0])
    W0[:3, 3] = [0.1, 0.2, 0.3]
    W1 = W0 @ M1
    W2 = np.diag([2.0, 3.0, 4.0, 1.0])
    meshes = [(box_v, box_i, box_n, box_uv, W0, 0), (tri_v, tri_i, None, None, W0, 0),
              (tet_v, tet_i, tet_n, tet_uv, W1, 1), (tet_v, tet_i, tet_n, tet_uv, W2, 1)]
    return _glb(doc, bytes(b.bin)), meshes


def _expected(meshes):
    pos, nor, uv, mat, idx, objs, soup = [], [], [], [], [], [], []
    off = 0
    for v, i, n, t, W, m in meshes:
        w = (W[:3, :3] @ v.astype(np.float64).T).T + W[:3, 3]
        sw = w[:, [0, 2, 1]]
        pos.append(sw)
        if n is not None:
            nm = np.linalg.inv(W[:3, :3]).T
            nn = (nm @ n.astype(np.float64).T).T
            nor.append(nn / np.linalg.norm(nn, axis=1, keepdims=True))
        uv.append(np.zeros((len(v), 2)) if t is None else t.astype(np.float64))
        mat.append(np.full(len(v), m))
        objs.append(dict(min=sw[i.astype(np.int64)].min(0), max=sw[i.astype(np.int64)].max(0), startIndex=sum(len(x) for x in idx),
                         countIndices=len(i)))
        idx.append(i.astype(np.int64) + off)
        soup.append(sw[i.astype(np.int64)])
        off += len(v)
    return dict(positions=np.concatenate(pos), normals=np.concatenate(nor), uvs=np.concatenate(uv), materialIndices=np.concatenate(mat),
                indices=np.concatenate(idx), objects=objs, vertices=np.concatenate(soup))


def test_glb_scene_arrays(mpm):
    data, meshes = _scene()
    got = mpm.load_glb(data)
    want = _expected(meshes)
    assert np.allclose(got["positions"], want["positions"], rtol=3e-7, atol=1e-7)
    assert np.allclose(got["vertices"], want["vertices"], rtol=3e-7, atol=1e-7)
    assert np.allclose(got["normals"], want["normals"], rtol=1e-6, atol=1e-7)
    assert np.array_equal(got["uvs"], want["uvs"].astype(np.float32))
    assert np.array_equal(got["materialIndices"], want["materialIndices"])
    assert np.array_equal(got["indices"], want["indices"])
    assert len(got["objects"]) == 4
    for g, w in zip(got["objects"], want["objects"]):
        assert g["startIndex"] == w["startIndex"] and g["countIndices"] == w["countIndices"]
        assert np.allclose(g["min"], w["min"], rtol=3e-7, atol=1e-7) and np.allclose(g["max"], w["max"], rtol=3e-7, atol=1e-7)
    # the soup is exactly the indexed positions
    assert np.array_equal(got["vertices"], got["positions"][got["indices"].astype(np.int64)])


def test_glb_feeds_collider_and_scatter(mpm):
    """The loaded arrays drive the collider BVH (set_collider) and the scatter mesh."""
    data, _ = _scene()
    sc = mpm.load_glb(data)
    nodes, order = mpm.build_bvh(sc["positions"], sc["indices"])
    assert len(nodes) > 0 and sorted(order.tolist()) == list(range(len(sc["indices"]) // 3))
    mm = mpm.GpuMeshBufferManager(sc["vertices"])
    assert mm.meshVertices.shape == (len(sc["indices"]), 4)


@pytest.mark.parametrize("bad", ["magic", "truncated", "nonindexed", "sparse", "uri"])
def test_glb_rejects_unsupported(mpm, bad):
    data, _ = _scene()
    if bad == "magic":
        data = b"XXXX" + data[4:]
    elif bad == "truncated":
        data = data[: len(data) // 2]
    else:
        jl = struct.unpack("<I", data[12:16])[0]
        doc = json.loads(data[20:20 + jl])
        binb = data[20 + jl + 8:]
        if bad == "nonindexed":
            del doc["meshes"][0]["primitives"][1]["indices"]
        elif bad == "sparse":
            doc["accessors"][0]["sparse"] = {"count": 1}
        else:
            doc["buffers"][0]["uri"] = "x.bin"
        data = _glb(doc, binb)
    with pytest.raises(mpm.MpmError):
        mpm.load_glb(data)


def _edit(data, fn):
    jl = struct.unpack("<I", data[12:16])[0]
    doc = json.loads(data[20:20 + jl])
    binb = data[20 + jl + 8:]
    fn(doc)
    return _glb(doc, binb)


# Counts, offsets and strides that would wrap a size_t range check or an allocation, indices that
# are fractions / negative / huge, attributes shorter than POSITION, nodes that are not a tree.
CRAFTED = {
    "count 2^63+2048, stride 8": lambda d: (d["accessors"][2].update(count=2 ** 63 + 2048),
                                            d["bufferViews"][d["accessors"][2]["bufferView"]].update(byteStride=8)),
    "count 2^64": lambda d: d["accessors"][0].update(count=2.0 ** 64),
    "count 1e300": lambda d: d["accessors"][0].update(count=1e300),
    "negative count": lambda d: d["accessors"][0].update(count=-1),
    "fractional count": lambda d: d["accessors"][0].update(count=2.5),
    "count beyond view": lambda d: d["accessors"][0].update(count=d["accessors"][0]["count"] + 1),
    "byteOffset -8": lambda d: d["accessors"][0].update(byteOffset=-8),
    "byteOffset 2^64-8": lambda d: d["accessors"][0].update(byteOffset=2 ** 64 - 8),
    "view byteOffset huge": lambda d: d["bufferViews"][0].update(byteOffset=2 ** 62),
    "view byteLength huge": lambda d: d["bufferViews"][0].update(byteLength=2 ** 63),
    "stride 2": lambda d: d["bufferViews"][0].update(byteStride=2),
    "stride 6": lambda d: d["bufferViews"][0].update(byteStride=6),
    "stride 256": lambda d: d["bufferViews"][0].update(byteStride=256),
    "stride 4 < element": lambda d: d["bufferViews"][0].update(byteStride=4),
    "accessor index -1": lambda d: d["meshes"][0]["primitives"][0]["attributes"].update(POSITION=-1),
    "accessor index 1e12": lambda d: d["meshes"][0]["primitives"][0]["attributes"].update(POSITION=1e12),
    "indices index 0.5": lambda d: d["meshes"][0]["primitives"][0].update(indices=0.5),
    "bufferView 99": lambda d: d["accessors"][0].update(bufferView=99),
    "mesh index -3": lambda d: d["nodes"][0].update(mesh=-3),
    "child index 7": lambda d: d["nodes"][0].update(children=[7]),
    "node listed twice": lambda d: d["nodes"][0].update(children=[1, 1]),
    "node cycle": lambda d: d["nodes"][1].update(children=[0]),
    "scene index 5": lambda d: d.update(scene=5),
    "texture index 2^40": lambda d: d["materials"][0]["pbrMetallicRoughness"]["baseColorTexture"].update(index=2 ** 40),
    "material 9": lambda d: d["meshes"][0]["primitives"][0].update(material=9),
    "normals shorter than positions": lambda d: d["accessors"][d["meshes"][0]["primitives"][0]["attributes"]["NORMAL"]].update(count=3),
    "uvs shorter than positions": lambda d: d["accessors"][d["meshes"][0]["primitives"][0]["attributes"]["TEXCOORD_0"]].update(count=2),
}


@pytest.mark.parametrize("case", sorted(CRAFTED))
def test_glb_rejects_crafted_sizes_and_indices(mpm, case):
    data, _ = _scene()
    bad = _edit(data, CRAFTED[case])
    with pytest.raises(mpm.MpmError):
        mpm.load_glb(bad)


def test_glb_accepts_valid_stride(mpm):
    """An interleaved-style byteStride (16 for a VEC3 float attribute over a copy of the data
    spread to 16-byte rows) reads the same positions."""
    data, meshes = _scene()
    jl = struct.unpack("<I", data[12:16])[0]
    doc = json.loads(data[20:20 + jl])
    binb = bytearray(data[20 + jl + 8:])
    acc = doc["accessors"][0]
    view = doc["bufferViews"][acc["bufferView"]]
    src = np.frombuffer(bytes(binb[view["byteOffset"]:view["byteOffset"] + view["byteLength"]]), np.float32).reshape(-1, 3)
    rows = np.zeros((src.shape[0], 4), np.float32)
    rows[:, :3] = src
    off = len(binb)
    binb += rows.tobytes()
    doc["bufferViews"].append({"buffer": 0, "byteOffset": off, "byteLength": rows.nbytes, "byteStride": 16})
    acc["bufferView"] = len(doc["bufferViews"]) - 1
    got = mpm.load_glb(_glb(doc, bytes(binb)))
    ref = mpm.load_glb(data)
    assert np.array_equal(got["positions"], ref["positions"])
