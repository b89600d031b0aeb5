"""The timing callback's source (GpuSnowPipelineRunner.svelte.ts:411-420,588-611): with timing on,
mpm_get_timings reports the device time of the whole mpm_step call, as computeSimulationStepNs spans
the frame's simulation compute pass, plus the phases of its last substep. Checked against the
phase sum, the substep count and the per-kernel event total of the same work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sim(mpm, n=65536):
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=n, seed=51)
    sim = mpm.MpmSimulation(n, **sc["sim"])
    sim.set_collider(*sc["collider"])
    sim.upload(sc["particles"])
    return sim


def test_timings_off_by_default_then_per_call(mpm):
    with _sim(mpm) as sim:
        sim.step(1, 2)
        sim.sync()
        t = sim.timings()
        assert t["compute_simulation_step_ns"] == 0 and t["pass_substeps"] == 0
        sim.set_timing(True)
        sim.step(1, 3)            # warm the graph of both buffer parities
        sim.step(1, 1)
        one = sim.timings()
        assert one["pass_substeps"] == 1 and one["substeps"] == 6
        phases = one["binning_ns"] + one["transfer_ns"] + one["integrate_ns"]
        assert min(one["binning_ns"], one["transfer_ns"], one["integrate_ns"]) > 0
        # one substep: the call's events bracket the same kernels as the phase events
        assert one["compute_simulation_step_ns"] >= 0.95 * phases
        assert one["compute_simulation_step_ns"] <= phases + 50e3 + 0.1 * phases
        sim.step(1, 10)
        ten = sim.timings()
        assert ten["pass_substeps"] == 10 and ten["substeps"] == 16
        per = ten["compute_simulation_step_ns"] / 10
        assert 0.3 * one["compute_simulation_step_ns"] < per < 3.0 * one["compute_simulation_step_ns"]
        # the same work with an event after every kernel: the kernels fill most of the call
        sim.set_kernel_timing(True)
        sim.step(1, 10)
        kt = sim.kernel_times()
        sim.set_kernel_timing(False)
        kernel_ns = 1e6 * sum(ms for ms, _ in kt.values())
        call_ns = sim.timings()["compute_simulation_step_ns"]
        assert 0.5 * call_ns < kernel_ns <= 1.05 * call_ns, (kernel_ns, call_ns)
        sim.set_timing(False)
        sim.step(1, 1)
        assert sim.timings()["pass_substeps"] == 10     # last timed call kept


def test_timings_of_a_batched_step(mpm):
    scenes = mpm.scenes.ensemble_scenes(n_scenes=4, n_min=8192, n_max=20000)
    with mpm.MpmEnsemble([sc["sim"] for sc in scenes]) as ens:
        for i, sc in enumerate(scenes):
            ens.upload(i, sc["particles"])
        ens.step(2)
        ens.sync()
        import ctypes
        lib = ens.lib
        assert lib.mpm_set_timing(ens.ctx, 1) == 0
        ens.step(3)
        t = mpm._lib.MpmTimings()
        assert lib.mpm_get_timings(ens.ctx, ctypes.byref(t)) == 0
        assert t.pass_substeps == 3 and t.compute_simulation_step_ns > 0
        assert t.compute_simulation_step_ns >= t.binning_ns + t.transfer_ns + t.integrate_ns
