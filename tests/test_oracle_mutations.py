"""The oracle's pins reject plausible mistakes: each case below plants one slip in a copy of
oracle/mpm_oracle.cpp or oracle/orc_svd.h (a swapped blend weight, a flipped exponent sign, a transposed product, a
dropped factor ...), builds it with the oracle's own flags in a temporary directory, and runs the
pin tests against that build. Every planted mistake must make at least one pin fail; the unmodified
sources must pass. CPU only; the copies never leave the temporary directory."""
import concurrent.futures
import os
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fno-finite-math-only", "-shared"]
PINS = ["tests/test_oracle_pins.py", "tests/test_oracle.py", "tests/test_oracle_svd.py"]

# (what the slip is, text in mpm_oracle.cpp, replacement); reference lines in the pins' docstrings
MUTATIONS = [
    ("PBMPM blend weights swapped", "madd(mscale(bf, ts), mscale((R)1 - bf, rot))", "madd(mscale((R)1 - bf, ts), mscale(bf, rot))"),
    ("PBMPM volume exponent sign", "r_pow(std::fabs(tv), (R)P.pbmpm_volume_exponent)", "r_pow(std::fabs(tv), -(R)P.pbmpm_volume_exponent)"),
    ("PBMPM relax exponent -h/Jp -> -h Jp", "r_exp(-(R)P.hardening_pbmpm / Jp)", "r_exp(-(R)P.hardening_pbmpm * Jp)"),
    ("PBMPM relax factor dropped", "pt.D = madd(pt.D, mscale(relax, diff));", "pt.D = madd(pt.D, diff);"),
    ("PBMPM correction F_e^-1 on the left", "mmul(tb, inverse3(pt.Fe))", "mmul(inverse3(pt.Fe), tb)"),
    ("PBMPM trial F_e (I + D)", "M3<R> trial = mmul(madd(ident<R>(), pt.D), pt.Fe);", "M3<R> trial = mmul(pt.Fe, madd(ident<R>(), pt.D));"),
    ("integrate F_e (I + D)", "pt.Fe = mmul(madd(ident<R>(), pt.D), pt.Fe);", "pt.Fe = mmul(pt.Fe, madd(ident<R>(), pt.D));"),
    ("hardening exponent sign", "r_exp((R)P.hardening_explicit * ((R)1 - Jp))", "r_exp((R)P.hardening_explicit * (Jp - (R)1))"),
    ("P2G uses P instead of P^T", "M3<R> PT = transpose(pk1_explicit(pt));", "M3<R> PT = pk1_explicit(pt);"),
    ("PK1 volume term loses a factor J", "(la * (J - (R)1)) * J", "(la * (J - (R)1))"),
    ("PK1 F^-T -> F^-1", "transpose(inverse3(pt.Fe))))", "inverse3(pt.Fe)))"),
    ("PK1 shear factor 2 dropped", "mscale((R)2 * mu, msub(pt.Fe, Rr))", "mscale(mu, msub(pt.Fe, Rr))"),
    ("P2G stress sign", "M3<R> negVPT = mscale(-V, PT);", "M3<R> negVPT = mscale(V, PT);"),
    ("wall restitution sign (min wall)", "if (p[a] < lo[a]) { d[a] = d[a] * (R)-0.5;", "if (p[a] < lo[a]) { d[a] = d[a] * (R)0.5;"),
    ("wall velocity sign (max wall)", "p[a] = hi[a]; v[a] = v[a] * (R)-0.5;", "p[a] = hi[a]; v[a] = v[a] * (R)0.5;"),
    ("collider CCD snap side", "pt.pos = hit_p + hit_n * margin;", "pt.pos = hit_p - hit_n * margin;"),
    ("collider CCD friction constant", "nv = vt * ((R)1 - (R)P.friction_ccd) + cvel * vsf;", "nv = vt * ((R)1 - (R)P.friction_proximity) + cvel * vsf;"),
    ("collider velocity factor 0.2/dt -> 0.2 dt", "R vsf = (R)P.collider_velocity_factor / dt();", "R vsf = (R)P.collider_velocity_factor * dt();"),
    ("collider proximity snap from the particle", "if (dist < margin) pt.pos = cl_p + pd * margin;", "if (dist < margin) pt.pos = cur + pd * margin;"),
    ("collider proximity keeps the normal velocity", "if (vn < (R)0) nv = vt * ((R)1 - (R)P.friction_proximity) + cvel * vsf;",
     "if (vn < (R)0) nv = vr * ((R)1 - (R)P.friction_proximity) + cvel * vsf;"),
    ("explicit G2P gradient transposed", "M3<R> o; o.c[0] = g.x * v; o.c[1] = g.y * v; o.c[2] = g.z * v;",
     "M3<R> o; o.c[0] = v.x * g; o.c[1] = v.y * g; o.c[2] = v.z * g;"),
    ("explicit G2P grad_scale dropped", "pt.D = mscale_r(mscale_r(G, dt()), (R)P.grad_scale);", "pt.D = mscale_r(G, dt());"),
    ("PBMPM G2P factor 4 -> 2", "pt.D = mscale_r(mscale_r(mscale_r(B, dt()), (R)4), (R)P.grad_scale);",
     "pt.D = mscale_r(mscale_r(mscale_r(B, dt()), (R)2), (R)P.grad_scale);"),
    ("PBMPM G2P B transposed", "B = madd(B, outer(cw * v, dist));", "B = madd(B, outer(dist, cw * v));"),
]
# slips in the SVD plasticity option (oracle/orc_svd.h)
SVD_MUTATIONS = [
    ("Jacobi V update sign", "out.V.c[p] = c * vp - s * vq;", "out.V.c[p] = c * vp + s * vq;"),
    ("Jacobi tau sign", "const R tau = (S[q][q] - S[p][p]) / ((R)2 * apq);", "const R tau = (S[p][p] - S[q][q]) / ((R)2 * apq);"),
    ("Jacobi diagonal update sign", "const R spp = S[p][p] - t * apq, sqq = S[q][q] + t * apq;",
     "const R spp = S[p][p] + t * apq, sqq = S[q][q] - t * apq;"),
    ("clamp ratio dropped", "f[j] = t / d.sig[j];", "f[j] = t;"),
    ("clamp with V transposed", "((f[0] * mat_el(d.V, k, 0)) * mat_el(d.B, row, 0) + (f[1] * mat_el(d.V, k, 1)) * mat_el(d.B, row, 1)) +\n                                  (f[2] * mat_el(d.V, k, 2)) * mat_el(d.B, row, 2)",
     "((f[0] * mat_el(d.V, 0, k)) * mat_el(d.B, row, 0) + (f[1] * mat_el(d.V, 1, k)) * mat_el(d.B, row, 1)) +\n                                  (f[2] * mat_el(d.V, 2, k)) * mat_el(d.B, row, 2)"),
    ("F_p update order", "const M3<float> nFp = mmul(inverse3(nFe), mmul(Fe, Fp));", "const M3<float> nFp = mmul(mmul(Fe, Fp), inverse3(nFe));"),
    ("degenerate threshold 1e-2 -> 1e-4", "return mn < (R)1e-2 * mx;", "return mn < (R)1e-4 * mx;"),
    ("one sweep instead of six", "for (int sweep = 0; sweep < 6; ++sweep) {", "for (int sweep = 0; sweep < 1; ++sweep) {"),
]


def _build(src_dir, texts, out):
    for name, text in texts.items():
        (src_dir / name).write_text(text)
    r = subprocess.run(["g++"] + FLAGS + ["-o", str(out), str(src_dir / "mpm_oracle.cpp"), str(src_dir / "orc_scatter.cpp")],
                       capture_output=True, text=True)
    return r.returncode, r.stderr


def _pins_pass(lib):
    env = dict(os.environ, MPM_ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + PINS,
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return r.returncode == 0, r.stdout[-400:]


def test_pins_reject_planted_mistakes(tmp_path):
    files = ("mpm_oracle.cpp", "orc_math.h", "orc_svd.h", "orc_scatter.cpp")
    base = {f: (ROOT / "oracle" / f).read_text() for f in files}
    cases = [("unmodified", dict(base))]
    for target, muts in (("mpm_oracle.cpp", MUTATIONS), ("orc_svd.h", SVD_MUTATIONS)):
        for name, a, b in muts:
            assert base[target].count(a) == 1, f"mutation site not unique or missing: {name}"
            cases.append((name, dict(base, **{target: base[target].replace(a, b)})))

    def run(i_case):
        i, (name, texts) = i_case
        d = tmp_path / f"m{i}"
        d.mkdir()
        rc, err = _build(d, texts, d / "liborc.so")
        assert rc == 0, f"{name}: build failed {err[-300:]}"
        ok, tail = _pins_pass(d / "liborc.so")
        return name, ok, tail

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(run, enumerate(cases)))
    first = results[0]
    assert first[1], f"unmodified oracle fails its pins: {first[2]}"
    survivors = [name for name, ok, _ in results[1:] if ok]
    assert not survivors, f"planted mistakes not caught by any pin: {survivors}"
