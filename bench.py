"""Benchmark: particle-substeps/s of the 1,048,576-particle PBMPM two-sphere scene on one MI355X.

A step is one substep of the whole scene (binning, k = 4 constraint iterations of
solve + P2G + G2P, integrate with the box collider), inputs resident in HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

roofline: the kernel with the largest device time per substep, measured with HIP events on the
library's stream over a separate instrumented pass (eager launches, one event after every kernel).
achieved = the reference formulation's algorithmic bytes for the work of one launch / its average
duration: a PBMPM iteration re-streams {x, m, F_e, F_p, D} and writes {D, v}, 172 B per particle
(the SURVEY's per-unit figure; 268 + 172 k B per particle-substep in all), so one k_pbmpm<MID>
launch is 172 B x binned particles. The implemented layout's own compulsory bytes for the kernel
are reported beside it. bound names what limits the kernel by its counters (VALU issue for the
transfer passes); fields read from committed rocprofv3 summaries carry their source file, and the
LDS atomic peak the atomics fields are priced on is measured in the run (mpm_selftest kind 11).
cpu_baseline: the oracle (FP32-faithful CPU restatement of the reference substep) with OpenMP
over the host's cores, 10 warm-up substeps, then the median of 5 timed substeps of the same 1M scene.
"""
import argparse
import importlib.util
import json
import os
import pathlib
import platform
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
PKG_DIR = ROOT / "webgpu-mpm-based-snow-simulation_amd"

HBM_PEAK_GBS = 8000.0
# Reference formulation, bytes per particle for the work of one launch of each transfer kernel:
#   PBMPM iteration: read x, m, F_e, F_p, D (124 B), write D, v (48 B) = 172 B (the solve, P2G and
#   G2P of one iteration); FIRST also carries the substep's first read of the state and LAST its
#   final write: per substep 268 + 172 k B (136 B read + 132 B written beyond the iterations).
#   Explicit substep: 268 B (P2G reads x, v, m, F_e, F_p; G2P + integrate rewrite x, v, F_e, F_p, D, dx).
REFERENCE_BYTES_PER_PARTICLE = {
    "k_pbmpm<FIRST>": 172, "k_pbmpm<MID>": 172, "k_pbmpm<LAST>": 172,
    "k_explicit_p2g": 136, "k_explicit_g2p": 132,
}
STEP_ALGO_BYTES = lambda k: 268 + 172 * k
# Compulsory particle-state bytes per particle of the implemented layout (FP32 SoA planes; F_p lives
# per particle id and is touched only on a plastic clamp, so the passes carry det F_p instead):
#   k_pbmpm<FIRST>:    read perm, id, x, m, F_e, det F_p, D, dx (112 B); write id, x, m, F_e, det F_p, relax (64 B)
#   k_pbmpm<MID>:      read x, m, F_e, relax (56 B)
#   k_pbmpm<LAST>:     read x, m, F_e (52 B); write x, v, F_e, dx, D (108 B)
#   k_explicit_p2g:    read perm, id, x, m, v, F_e, det F_p (76 B); write id, x, m, F_e, det F_p (60 B)
#   k_explicit_g2p:    as k_pbmpm<LAST>
OWN_LAYOUT_BYTES_PER_PARTICLE = {
    "k_pbmpm<FIRST>": 176, "k_pbmpm<MID>": 56, "k_pbmpm<LAST>": 160,
    "k_explicit_p2g": 136, "k_explicit_g2p": 160,
}
SIMDS, CLOCK_HZ = 1024, 2.4e9       # MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock


def load_package():
    spec = importlib.util.spec_from_file_location("mpm_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mpm_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def host_threads():
    """Threads for the CPU baseline: the host's cores, bounded by OMP_NUM_THREADS when set."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(mpm, sc, sim, warmup=10, runs=5):
    """The oracle on the host: the FP32-faithful restatement of the reference substep, OpenMP over
    host_threads() threads, `warmup` substeps, then the median of `runs` single timed substeps."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    p = sc["particles"]
    kw = sc["sim"]
    o = orc.OracleSim(p, sim.uniform_bytes(), pbmpm_iterations=kw["pbmpm_iterations"],
                      pbmpm_force_divisor=kw["pbmpm_force_divisor"], gravity=kw["gravity"])
    o.set_collider(sim.colliderManager.colliderData, sim.colliderManager.bvhNodes)
    threads = o.set_threads(host_threads())
    t0 = time.perf_counter()
    o.step(sc["method"], warmup)
    t_warm = time.perf_counter() - t0
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        o.step(sc["method"], 1)
        times.append(time.perf_counter() - t0)
    o.close()
    med = float(np.median(times))
    return {"value": p.shape[0] / med, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"full {p.shape[0]}-particle scene: {warmup} warm-up substeps ({t_warm:.1f} s), then the median of "
                      f"{runs} timed substeps ({', '.join(f'{t:.2f}' for t in times)} s); oracle FP32 restatement, "
                      f"OpenMP {threads} threads on {cpu_model()} ({os.cpu_count()} logical CPUs)"}


def learned_vs_plain(mpm, n, k=16, k_ref=64, substeps=600):
    """The learned warm start against plain PBMPM on the bench scene.

    PBMPM's k iterations relax the particle constraints towards a fixed point (the force
    divisor splits the external forces over them), so few iterations leave the material too
    soft; the reference for accuracy is the most-iterated run the library allows, k_ref = 64.
    Training pairs come from another sampling of the scene (seed 7): a trajectory at k
    iterations, every 10th substep logged with the k_ref-iteration result from the same start
    state as the target. The network is trained on the GPU (Adam, 6,000 steps).

    Measured on the bench scene (seed 1), `substeps` substeps each:
      * the rms position error after the run against plain PBMPM at k_ref, for the warm start at
        k and plain PBMPM at a sweep of iteration counts; iteration_ratio = (the fewest plain
        iterations whose error is at most the warm start's) / k;
      * the per-iteration residual r_i = mean over substeps of max_particles
        |dx_i - dx_{i-1} - g dt^2 / k|_2 / domain along each method's own run at k, and the
        iterations each needs to reach plain's r at k / 2;
      * the time per substep of each run, and the warm start's cost on one state: a run at k
        after 100 substeps, 64 substeps with the warm start off, then on, three times (median)."""
    def make(seed):
        sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=n, seed=seed)
        sim = mpm.MpmSimulation(n, **sc["sim"])
        sim.set_collider(*sc["collider"])
        sim.upload(sc["particles"])
        return sim

    t0 = time.perf_counter()
    tr = make(7)
    pairs = tr.learn_collect(substeps, k_ref=k_ref, every=10, k_run=k)
    loss_a = tr.learn_train(4000, batch=65536, lr=1e-3, seed=1)
    loss_b = tr.learn_train(2000, batch=65536, lr=2e-4, seed=2)
    net = tr.learn_params()
    tr.close()
    t_train = time.perf_counter() - t0

    def run(kk, learned, track=False):
        """Final positions, residuals (track) and the stepping time per substep: one substep
        first (graph capture), then the rest timed between synchronisations."""
        sim = make(1)
        sim.set_iterations(kk)
        if learned:
            sim.set_learn_params(*net)
            sim.learn_enable(True)
        if track:
            sim.set_residual_tracking(True)
        sim.step(1, 1)
        sim.prepare_graphs(1)
        sim.sync()
        t = time.perf_counter()
        sim.step(1, substeps - 1)
        sim.sync()
        ms = (time.perf_counter() - t) / (substeps - 1) * 1e3
        pos = sim.download()["pos"][:, :3].astype(np.float64)
        res = sim.residuals()[0][:kk].astype(np.float64) if track else None
        st = sim.stats()
        sim.close()
        return pos, res, st, ms

    ref, _, _, ms_ref = run(k_ref, False)

    def err(pos):
        e = np.linalg.norm(pos - ref, axis=1)
        return float(np.sqrt((e * e).mean()))

    pos_l, _, st, ms_learn = run(k, True)
    pos_p, _, _, ms_plain = run(k, False)
    _, r_learn, _, _ = run(k, True, True)
    _, r_plain, _, _ = run(k, False, True)
    e_learn = err(pos_l)
    sweep = {k: err(pos_p)}
    for kk in (k_ref // 2, 3 * k_ref // 4, k_ref - 8, k_ref - 4, k_ref - 2, k_ref - 1):
        if kk > k and kk not in sweep:
            sweep[kk] = err(run(kk, False)[0])
    match = [kk for kk in sorted(sweep) if sweep[kk] <= e_learn]

    def iters_to(r, thr):
        hit = [i for i in range(1, len(r)) if r[i] <= thr]
        return hit[0] if hit else None

    thr = float(r_plain[k // 2])
    it_p, it_l = iters_to(r_plain, thr), iters_to(r_learn, thr)

    # Cost on one state: the runs above follow different trajectories whose passes cost different
    # amounts, so the warm start's price is timed by switching it off and on along one run at k.
    sim = make(1)
    sim.set_iterations(k)
    sim.set_learn_params(*net)
    sim.step(1, 100)
    same = {"learned": [], "plain": []}
    for _ in range(3):
        for on in (False, True):
            sim.learn_enable(on)
            sim.step(1, 8)
            sim.prepare_graphs(1)
            sim.sync()
            t = time.perf_counter()
            sim.step(1, 64)
            sim.sync()
            same["learned" if on else "plain"].append((time.perf_counter() - t) / 64 * 1e3)
    sim.close()
    return {
        "k": k, "k_ref": k_ref, "substeps": substeps, "training_pairs": pairs,
        "train_loss": [loss_a, loss_b], "train_and_collect_s": round(t_train, 2),
        "rms_error_vs_k_ref": {"learned": e_learn, **{f"plain_k{kk}": v for kk, v in sorted(sweep.items())}},
        "plain_iterations_to_match": match[0] if match else None,
        "iteration_ratio": (match[0] / k) if match else k_ref / k,
        "iteration_ratio_is_bound": not match,
        "ms_per_substep": {"learned": round(ms_learn, 4), "plain": round(ms_plain, 4), "plain_k_ref": round(ms_ref, 4)},
        "ms_per_substep_same_state": {m: round(float(np.median(v)), 4) for m, v in same.items()},
        "residual_plain": [float(x) for x in r_plain[1:]], "residual_learned": [float(x) for x in r_learn[1:]],
        "residual_threshold": thr, "residual_iters_plain": it_p, "residual_iters_learned": it_l,
        "residual_iteration_ratio": (it_p / it_l) if (it_p and it_l) else None,
        "iters_to_1e-6": {"plain": iters_to(r_plain, 1e-6), "learned": iters_to(r_learn, 1e-6)},
        "fallbacks_last_substep": st["mlp_fallbacks"],
    }


def ensemble_bench(mpm, steps=30, warmup=5):
    """The 64-scene batched workload (scenes.ensemble_scenes): autotuned, then timed. Throughput
    counts every scene's particles once per substep."""
    scenes = mpm.scenes.ensemble_scenes()
    n = sum(sc["sim"]["n_particles"] for sc in scenes)
    with mpm.MpmEnsemble([sc["sim"] for sc in scenes]) as ens:
        for i, sc in enumerate(scenes):
            ens.upload(i, sc["particles"])
        best, cands = ens.autotune(trial_substeps=3)
        ens.step(warmup)
        ens.sync()
        t0 = time.perf_counter()
        ens.step(steps)
        ens.sync()
        dt = time.perf_counter() - t0
        ens.set_kernel_timing(True)
        ens.step(5)
        ens.sync()
        kt = ens.kernel_times()
        ens.set_kernel_timing(False)
    return {"workload": "ensemble_64_scenes", "scenes": len(scenes), "particles": n,
            "explicit_scenes": sum(1 for sc in scenes if sc["sim"]["method"] == 0),
            "value": n * steps / dt, "unit": "particle-substeps/s", "ms_per_step": 1e3 * dt / steps,
            "tuning": best, "autotune_candidates": cands,
            "kernel_ms_per_substep": {k: round(v[0] / 5, 4) for k, v in sorted(kt.items(), key=lambda x: -x[1][0])}}


def terrain_bench(mpm, steps=100, warmup=10):
    """The 4,194,304-particle explicit slab over the 7,938-triangle terrain (256^3 grid,
    per-particle Young's modulus): particle-substeps/s and per-kernel device time."""
    sc = mpm.scenes.scene("explicit_slab_terrain")
    n = sc["particles"].shape[0]
    with mpm.MpmSimulation(n, **sc["sim"]) as sim:
        sim.set_collider(*sc["collider"])
        sim.upload(sc["particles"])
        sim.set_particle_youngs(sc["youngs"])
        sim.step(0, warmup)
        sim.prepare_graphs(0)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(0, steps)
        sim.sync()
        dt = time.perf_counter() - t0
        sim.set_kernel_timing(True)
        sim.step(0, 10)
        sim.sync()
        kt = sim.kernel_times()
        st = sim.stats()
    return {"workload": "explicit_slab_terrain", "particles": n, "grid": "256^3 sparse, direct block keys", "collider_triangles": len(sc["collider"][1]) // 3,
            "value": n * steps / dt, "unit": "particle-substeps/s", "ms_per_step": 1e3 * dt / steps,
            "blocks_allocated": st["blocks_allocated"],
            "kernel_ms_per_substep": {k: round(v[0] / 10, 4) for k, v in sorted(kt.items(), key=lambda x: -x[1][0])}}


def copy_calibration(mpm, nbytes):
    """One plain device copy kernel of nbytes (k_selftest kind 10): what FETCH_SIZE / WRITE_SIZE
    report for it under rocprofv3 --pmc calibrates the counters against a copy of known size."""
    import ctypes
    lib = mpm.load_library()
    n = nbytes // 4
    a = np.ones(n, np.float32)
    out = np.zeros(n, np.float32)
    rc = lib.mpm_selftest(10, a.ctypes.data_as(ctypes.c_void_p), None, out.ctypes.data_as(ctypes.c_void_p), n)
    assert rc == 0 and out[-1] == 1.0


def lds_atomic_rate(mpm, iters=65536):
    """Conflict-free LDS atomic throughput of this GPU (mpm_selftest kind 11): wave-wide ds_add_u32
    instructions per second over all CUs, and clocks per instruction per CU at the 2.4 GHz peak clock."""
    import ctypes
    lib = mpm.load_library()
    out = np.zeros(3, np.float32)
    rc = lib.mpm_selftest(11, None, None, out.ctypes.data_as(ctypes.c_void_p), iters)
    if rc != 0 or out[0] <= 0:
        return None
    rate = float(out[1]) / (float(out[0]) * 1e-3)
    return {"wave_instr_per_s": rate, "clocks_per_wave_instr_per_cu": float(out[2]) * CLOCK_HZ / rate,
            "measured": "mpm_selftest kind 11 (conflict-free ds_add_u32 loop, 8 workgroups per CU)"}


def profile_entry(name, kernel):
    """(entry, source) of a kernel in a committed rocprofv3 summary under profiles/ (None if absent):
    these numbers come from separate counter runs, not from this run."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None, None
    try:
        d = json.loads(f.read_text())
    except (ValueError, OSError):
        return None, None
    return d.get(kernel), d.get("_source", f"profiles/{name}")


def profiled_fields(kernel, avg_s, lds_rate=None):
    """Counter-derived figures of the dominant kernel, each with the committed summary it is read from
    (the LDS atomic peak is measured in this run when lds_rate is given)."""
    out = {}
    tr, src = profile_entry("traffic.json", kernel)
    if tr:
        out["traffic_source"] = src
        out["traffic_fetch_bytes"] = tr["fetch_bytes"]
        out["traffic_write_bytes"] = tr["write_bytes"]
    va, src = profile_entry("valu.json", kernel)
    if va:
        # VALU issue utilisation = wave-instructions x 4 cycles / (1,024 SIMDs x kernel cycles), both
        # from the counter run; the same instruction count over this run's launch time beside it
        out["valu"] = {"source": src, "wave_instructions_per_launch": va["valu_wave_instructions_per_launch"],
                       "issue_utilisation_profiled": va.get("issue_utilisation"),
                       "issue_utilisation_at_this_run_duration": va["valu_wave_instructions_per_launch"] * 4 / (SIMDS * CLOCK_HZ * avg_s)}
    at, src = profile_entry("atomics.json", kernel)
    if at:
        pk = dict(profile_entry("atomics.json", "_peaks")[0] or {})
        if lds_rate:
            pk["lds_atomic_wave_instr_per_s"] = lds_rate["wave_instr_per_s"]
            pk["lds_measured_in_this_run"] = lds_rate
        g, l = at["tcc_atomic_requests"] / avg_s, at["lds_atomic_wave_instr"] / avg_s
        out["atomics"] = {"source": src, "global_requests_per_launch": at["tcc_atomic_requests"], "global_per_s": g,
                          "global_frac": g / pk["global_atomic_requests_per_s"] if pk else None,
                          "lds_wave_instr_per_launch": at["lds_atomic_wave_instr"], "lds_per_s": l,
                          "lds_frac": l / pk["lds_atomic_wave_instr_per_s"] if pk else None, "peaks_assumed": pk}
    return out


def hash_table_bench(mpm, steps=100, warmup=10):
    """The same 1M-particle scene with the reference's open-addressing block table (hash_map_size
    200,003 = sparseGridPrelude.wgsl:6, atomicCAS insert + compaction, mapAffectedBlocks.wgsl:7-62)
    instead of direct block keys: particle-substeps/s and the table kernels' device time."""
    sc = mpm.scenes.scene("pbmpm_two_spheres")
    n = sc["particles"].shape[0]
    with mpm.MpmSimulation(n, **dict(sc["sim"], hash_map_size=200003)) as sim:
        sim.set_collider(*sc["collider"])
        sim.upload(sc["particles"])
        sim.step(1, warmup)
        sim.prepare_graphs(1)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(1, steps)
        sim.sync()
        dt = time.perf_counter() - t0
        sim.set_kernel_timing(True)
        sim.step(1, 10)
        sim.sync()
        kt = sim.kernel_times()
        st = sim.stats()
    return {"workload": "pbmpm_two_spheres, open-addressing block table", "particles": n, "hash_map_size": 200003,
            "value": n * steps / dt, "unit": "particle-substeps/s", "ms_per_step": 1e3 * dt / steps,
            "blocks_allocated": st["blocks_allocated"], "probe_failures": st["probe_failures"],
            "kernel_ms_per_substep": {k: round(v[0] / 10, 4) for k, v in sorted(kt.items(), key=lambda x: -x[1][0])}}


def svd_plasticity_bench(mpm, steps=100, warmup=10):
    """The same 1M-particle scene with the SVD plasticity option (mode 1 FP32, mode 2 with the FP64
    fallback): particle-substeps/s and the SVD pass's device time per substep."""
    out = {}
    sc = mpm.scenes.scene("pbmpm_two_spheres")
    n = sc["particles"].shape[0]
    for mode in (1, 2):
        with mpm.MpmSimulation(n, **dict(sc["sim"], plasticity_mode=mode)) as sim:
            sim.set_collider(*sc["collider"])
            sim.upload(sc["particles"])
            sim.step(1, warmup)
            sim.prepare_graphs(1)
            sim.sync()
            t0 = time.perf_counter()
            sim.step(1, steps)
            sim.sync()
            dt = time.perf_counter() - t0
            sim.set_kernel_timing(True)
            sim.step(1, 10)
            sim.sync()
            kt = sim.kernel_times()
            st = sim.stats()
        out[f"mode_{mode}"] = {"value": n * steps / dt, "unit": "particle-substeps/s", "ms_per_step": 1e3 * dt / steps,
                               "k_plastic_svd_ms_per_substep": round(kt.get("k_plastic_svd", (0.0, 1))[0] / 10, 4),
                               "svd_fp64_particles_last_substep": st["svd_fp64_particles"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--kernel-steps", type=int, default=60)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched 64-scene workload")
    ap.add_argument("--no-terrain", action="store_true", help="skip the 4M-particle explicit terrain workload")
    ap.add_argument("--no-hash-table", action="store_true", help="skip the open-addressing block table run")
    ap.add_argument("--no-svd", action="store_true", help="skip the SVD plasticity option runs")
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: libmpm_hip steps a scene on exactly one GPU; --gpus N > 1 is not supported "
              "(the multi-GPU sharding of the step is design-only)", file=sys.stderr)
        sys.exit(2)

    mpm = load_package()
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=args.particles)
    p = sc["particles"]
    n = p.shape[0]
    k = sc["sim"]["pbmpm_iterations"]
    sim = mpm.MpmSimulation(n, **sc["sim"])
    v, idx = sc["collider"]
    sim.set_collider(v, idx)
    sim.upload(p)
    print(f"bench: {n} particles, PBMPM k={k}, warmup {args.warmup}", file=sys.stderr, flush=True)

    sim.step(1, args.warmup)
    sim.prepare_graphs(1)          # graph capture (one substep, a run of 8) outside the timed region
    sim.sync()
    t0 = time.perf_counter()
    sim.step(1, args.steps)
    sim.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = n * args.steps / elapsed
    print(f"bench: timed {args.steps} substeps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    # per-kernel device time (HIP events on the library's stream)
    sim.set_kernel_timing(True)
    sim.step(1, args.kernel_steps)
    sim.sync()
    kt = sim.kernel_times()
    sim.set_kernel_timing(False)
    per_substep = {name: ms / args.kernel_steps for name, (ms, cnt) in kt.items()}
    dom = max((nm for nm in per_substep if nm in REFERENCE_BYTES_PER_PARTICLE), key=lambda nm: per_substep[nm])
    dom_ms, dom_cnt = kt[dom]
    avg_s = dom_ms / dom_cnt * 1e-3
    stats = sim.stats()
    binned = stats["particles_binned"]
    algo_bytes = REFERENCE_BYTES_PER_PARTICLE[dom] * binned
    achieved = algo_bytes / avg_s / 1e9
    own_bytes = OWN_LAYOUT_BYTES_PER_PARTICLE[dom] * binned
    eager_ms = sum(per_substep.values())
    prof = profiled_fields(dom, avg_s, lds_atomic_rate(mpm))
    traffic = (prof["traffic_fetch_bytes"] + prof["traffic_write_bytes"]) if "traffic_fetch_bytes" in prof else None

    line = {
        "metric": "particle-substeps/s, 1M-particle PBMPM",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-sphere scene, 1 mass unit per particle)",
        "config": {"workload": "pbmpm_two_spheres", "particles": n, "grid": "128^3 sparse, 4^3 blocks addressed directly by block key",
                   "pbmpm_iterations": k, "dt": sc["sim"]["dt"], "collider": "12-triangle box, BVH",
                   "parallelism": "single GPU"},
        # the counters show the kernel's limiter is VALU issue ("valu": issue utilisation); achieved /
        # peak / frac price the reference formulation's HBM bytes for the same work against 8 TB/s
        "roofline": {"bound": "valu", "priced_on": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dom, "avg_launch_us": avg_s * 1e6, "launches_timed": dom_cnt,
                     "algorithmic_bytes_per_particle": REFERENCE_BYTES_PER_PARTICLE[dom], "particles_per_launch": binned,
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "own_layout_bytes_per_launch": own_bytes, "own_layout_achieved": own_bytes / avg_s / 1e9,
                     "own_layout_frac": own_bytes / avg_s / 1e9 / HBM_PEAK_GBS,
                     "step_frac_reference_bytes": (STEP_ALGO_BYTES(k) * value / 1e9) / HBM_PEAK_GBS,
                     **prof},
        "kernel_ms_per_substep": {nm: round(v, 5) for nm, v in sorted(per_substep.items(), key=lambda x: -x[1])},
        "eager_ms_per_substep": eager_ms,
    }
    if args.calibrate_copy_mb:
        copy_calibration(mpm, args.calibrate_copy_mb << 20)
    if not args.no_terrain:
        print("bench: 4M explicit terrain workload", file=sys.stderr, flush=True)
        line["explicit_terrain"] = terrain_bench(mpm)
    if not args.no_hash_table:
        print("bench: open-addressing block table", file=sys.stderr, flush=True)
        line["hash_table"] = hash_table_bench(mpm)
    if not args.no_svd:
        print("bench: SVD plasticity option", file=sys.stderr, flush=True)
        line["svd_plasticity"] = svd_plasticity_bench(mpm)
    if not args.no_ensemble:
        print("bench: batched 64-scene workload", file=sys.stderr, flush=True)
        line["ensemble"] = ensemble_bench(mpm)
    if not args.no_learned:
        print("bench: learned warm start vs plain PBMPM", file=sys.stderr, flush=True)
        line["learned_vs_plain"] = learned_vs_plain(mpm, n)
    if not args.no_cpu_baseline:
        print("bench: cpu baseline (oracle, OpenMP)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(mpm, sc, sim)
    else:
        line["cpu_baseline"] = None
    sim.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
