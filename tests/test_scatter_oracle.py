"""CPU checks that pin the scatter-init oracle (oracle/orc_scatter.cpp) with facts that do not
come from its own formulas: noise3 on the integer lattice is the (separately pinned) hash3 value,
fbm there is a closed-form octave sum, the mass is (hash1(i) / 2^32)^2, every particle's first
candidate follows from the integer hash stream, the even-odd inside test agrees with closed forms
for a box and an octahedron, and the jitter stays within 0.05 per axis."""
import numpy as np
import pytest

from oracle import oracle as orc


def _box_soup(lo=0.2, hi=0.8):
    import __graft_entry__
    m = __graft_entry__.load_package()
    v, idx = m.scenes.box_mesh(((lo + hi) / 2,) * 3, hi - lo)
    return m.triangle_soup(v, idx)


def _octahedron_soup(c=0.5, r=0.3):
    P = np.array([[c + r, c, c], [c - r, c, c], [c, c + r, c], [c, c - r, c], [c, c, c + r], [c, c, c - r]])
    tris = []
    for sx in (0, 1):
        for sy in (2, 3):
            for sz in (4, 5):
                tris.append((sx, sy, sz))
    return P[np.asarray(tris).ravel()]


def _xyzw(soup):
    out = np.zeros((soup.shape[0], 4), np.float32)
    out[:, :3] = soup
    return out


def test_noise3_on_lattice_is_the_hash():
    L = orc.lib()
    rng = np.random.default_rng(1)
    for p in rng.integers(-50, 400, (200, 3)):
        h = L.orc_hash3(*(int(x) & 0xFFFFFFFF for x in p))
        assert L.orc_noise3(*(float(x) for x in p)) == np.float32(h) / np.float32(2.0 ** 32)
    # cell centres: u = 1/2 in every axis, the mean of the 8 corner hashes
    for p in rng.integers(0, 100, (50, 3)):
        corners = [L.orc_hash3((int(p[0]) + dx) & 0xFFFFFFFF, (int(p[1]) + dy) & 0xFFFFFFFF, (int(p[2]) + dz) & 0xFFFFFFFF)
                   for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)]
        want = np.mean(np.asarray(corners, np.float64) / 2.0 ** 32)
        assert abs(L.orc_noise3(*(float(x) + 0.5 for x in p)) - want) < 1e-6


def test_fbm_on_lattice_is_the_octave_sum():
    L = orc.lib()
    for p in np.random.default_rng(2).integers(0, 60, (100, 3)):
        want = sum(0.5 ** (i + 1) * L.orc_hash3(*((int(x) << i) & 0xFFFFFFFF for x in p)) / 2.0 ** 32 for i in range(6))
        assert abs(L.orc_fbm(*(float(x) for x in p)) - want) < 2e-7


def test_inside_test_matches_closed_forms():
    rng = np.random.default_rng(3)
    box = _xyzw(_box_soup())
    octa = _xyzw(_octahedron_soup())
    L = orc.lib()
    pts = rng.uniform(0.0, 1.0, (4000, 3)).astype(np.float32)
    for p in pts:
        if np.min(np.abs(np.concatenate([p - 0.2, p - 0.8]))) > 1e-4:
            assert L.orc_inside_mesh(orc._ptr(p), orc._ptr(box), box.shape[0]) == int(np.all((p > 0.2) & (p < 0.8)))
        s = np.abs(p.astype(np.float64) - 0.5).sum()
        if abs(s - 0.3) > 1e-4 and np.min(np.abs(p[1:] - 0.5)) > 1e-4:
            assert L.orc_inside_mesh(orc._ptr(p), orc._ptr(octa), octa.shape[0]) == int(s < 0.3)


def _first_candidate(i, mn, mx):
    L = orc.lib()
    seed = L.orc_hash4(i, 100, 200, 300)
    out = []
    for a in range(3):
        seed = L.orc_hash1(seed)
        t = np.float32(seed) / np.float32(2.0 ** 32)
        out.append(np.float32(mn[a]) * (np.float32(1) - t) + np.float32(mx[a]) * t)
    return np.asarray(out, np.float32)


def test_scatter_box_first_candidate_mass_and_jitter():
    soup = _box_soup()
    n = 3000
    mn, mx = soup.min(0).astype(np.float32), soup.max(0).astype(np.float32)
    recs = orc.scatter(_xyzw(soup), mn, mx, n)
    L = orc.lib()
    for i in range(0, n, 7):
        c = _first_candidate(i, mn, mx)
        assert np.all((c > 0.2) & (c < 0.8))
        q = (c * np.float32(4)).astype(np.float32)
        d = np.array([L.orc_fbm(*q), L.orc_fbm(*(q + np.float32([100, 200, 300]))), L.orc_fbm(*(q + np.float32([-300, -200, -100])))], np.float32)
        want = c + (d - np.float32(0.5)) * np.float32(0.1)
        assert np.array_equal(recs[i, :3], want.astype(np.float32))
        m1 = np.float32(L.orc_hash1(i)) / np.float32(2.0 ** 32)
        assert recs[i, 7] == m1 * m1
    assert np.all(np.abs(recs[:, :3] - 0.5) <= 0.3 + 0.05 + 1e-6)
    # identity F_e, F_p; zero velocity, displacements; _hom = 1
    eye = np.zeros((3, 4), np.float32); eye[0, 0] = eye[1, 1] = eye[2, 2] = 1
    assert np.all(recs[:, 8:20].reshape(n, 3, 4) == eye) and np.all(recs[:, 20:32].reshape(n, 3, 4) == eye)
    assert np.all(recs[:, 3] == 1) and np.all(recs[:, 4:7] == 0) and np.all(recs[:, 32:48] == 0)


def test_scatter_rejection_keeps_candidates_inside_an_octahedron():
    soup = _octahedron_soup()
    n = 2000
    mn, mx = soup.min(0).astype(np.float32), soup.max(0).astype(np.float32)
    recs = orc.scatter(_xyzw(soup), mn, mx, n)
    # each final point lies within 0.05 per axis of a candidate inside the octahedron
    s = np.abs(recs[:, :3].astype(np.float64) - 0.5).sum(1)
    assert np.all(s < 0.3 + 3 * 0.05 + 1e-6)
    # the octahedron fills 1/6 of its box, so most first candidates are rejected
    moved = sum(1 for i in range(200) if np.abs(_first_candidate(i, mn, mx).astype(np.float64) - 0.5).sum() >= 0.3)
    assert moved > 140


# ------------------------------------------------------------------ render-side splats
def _recs(pos, mass):
    a = np.zeros((len(pos), 48), np.float32)
    a[:, :3] = pos
    a[:, 7] = mass
    return a


def test_splat_density_cell_centre_and_conservation():
    gmin, gmax, res, S = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (16, 8, 4), 65536.0
    # a particle at the centre of cell (5, 3, 1): all of its mass on that node
    p = np.array([[(5 + 0.5) / 16, (3 + 0.5) / 8, (1 + 0.5) / 4]], np.float32)
    g = orc.splat(0, _recs(p, [0.75]), gmin, gmax, S, res)
    assert g[1, 3, 5] == 49152 and g.sum() == 49152
    # interior particles: the node sum loses less than one unit per node to truncation
    rng = np.random.default_rng(4)
    pos = rng.uniform(0.2, 0.8, (5000, 3)).astype(np.float32)
    m = rng.uniform(0.1, 1.0, 5000).astype(np.float32)
    g = orc.splat(0, _recs(pos, m), gmin, gmax, S, res)
    tot = float((m.astype(np.float64) * S).sum())
    assert tot - 8 * 5000 <= float(g.sum()) <= tot


def test_splat_mass_lower_boundary_quirk():
    """calculateGridMass truncates the cell-centred coordinate with u32(): at pos = gridMin the
    coordinate is -0.5, the start cell 0 and the weights (1.5, -0.5) per axis; products with one or
    three negative factors clamp to 0."""
    gmin, gmax, res, S = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (8, 8, 8), 65536.0
    g = orc.splat(1, _recs(np.zeros((1, 3), np.float32), [1.0]), gmin, gmax, S, res)
    assert g[0, 0, 0] == int(np.float32(3.375) * S)
    for (z, y, x) in ((0, 1, 1), (1, 0, 1), (1, 1, 0)):
        assert g[z, y, x] == int(np.float32(0.375) * S)
    assert g.sum() == int(np.float32(3.375) * S) + 3 * int(np.float32(0.375) * S)
