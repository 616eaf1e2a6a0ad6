"""Batched sweeps (fullmag_fdm_batch_*): systems that share launches give
the same results as the same systems stepped alone, each with its own
adaptive dt sequence or ABM3 history (bit-identical: the batched launch only
adds a grid row per system)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plans(precision, integrator="dp45"):
    from fullmag_amd.backend import FdmPlan
    from fullmag_amd.workloads import batch_sweep
    shapes = ((32, 16, 1), (16, 16, 2), (24, 20, 1))
    out = [FdmPlan(**kw) for kw in batch_sweep(count=7, seed=9, shapes=shapes, precision=precision)]
    for p in out:
        p.integrator = integrator
    return out


@pytest.mark.parametrize("precision,integrator", [("single", "dp45"), ("double", "dp45"), ("double", "abm3"),
                                                  ("single", "rk23")])
def test_batch_equals_individual_runs(precision, integrator):
    from fullmag_amd.backend import FdmBackend, FdmBatch
    plans = _plans(precision, integrator)
    t_end = 2e-12
    with FdmBatch(plans) as b:
        assert b.groups() == len({(p.nx, p.ny, p.nz) for p in plans})
        wall, steps = b.advance(t_end, 100000, 1e-13)
        got = [s.field("m") for s in b.systems]
        stats = [s.stats() for s in b.systems]
    assert np.all(wall > 0)
    for i, p in enumerate(plans):
        with FdmBackend(p) as be:
            r = be.advance(t_end, 100000, 1e-13)
            assert r["steps_taken"] == steps[i], i
            assert np.array_equal(be.field("m"), got[i]), i
            assert r["time_seconds"] == stats[i]["time_seconds"]
            assert r["suggested_next_dt"] == stats[i]["suggested_next_dt"]


def test_batch_second_advance_continues_and_max_steps_caps():
    from fullmag_amd.backend import FdmBackend, FdmBatch
    plans = _plans("single")
    with FdmBatch(plans) as b:
        _, s1 = b.advance(1e-12, 100000, 1e-13)
        _, s2 = b.advance(2e-12, 3, 1e-13)
        assert np.all(s2 <= 3)
        got = [s.field("m") for s in b.systems]
    for i, p in enumerate(plans):
        with FdmBackend(p) as be:
            be.advance(1e-12, 100000, 1e-13)
            r = be.advance(2e-12, 3, 1e-13)
            assert r["steps_taken"] == s2[i]
            assert np.array_equal(be.field("m"), got[i])


def test_batch_reports_bad_plan():
    from fullmag_amd.backend import FdmBatch, FdmError
    plans = _plans("single")
    plans[3].extra = {"boundary_correction": 1}
    with pytest.raises(FdmError) as e:
        FdmBatch(plans)
    assert "system 3" in str(e.value)


def _one_per_shape(plans):
    seen = {}
    for i, p in enumerate(plans):
        seen.setdefault((p.nx, p.ny, p.nz), i)
    return sorted(seen.values())


def test_config4_sweep_equals_individual_runs():
    # the benchmark's sweep itself: 96 systems drawn from 64x64x1, 128x64x2,
    # 256x128x1 and 256x256x4, FP32, each with its own adaptive dt
    from fullmag_amd.backend import FdmBackend, FdmBatch, FdmPlan
    from fullmag_amd.workloads import batch_sweep
    plans = [FdmPlan(**kw) for kw in batch_sweep()]
    assert len(plans) == 96
    t_end = 2e-12
    pick = _one_per_shape(plans)
    assert len(pick) == 4
    with FdmBatch(plans) as b:
        assert b.groups() == 4
        wall, steps = b.advance(t_end, 10 ** 6, 1e-13)
        assert np.all(steps > 0) and np.all(wall > 0)
        got = {i: b.systems[i].field("m") for i in pick}
        for i in range(len(plans)):
            assert b.systems[i].stats()["time_seconds"] == pytest.approx(t_end, rel=1e-12)
    for i in pick:
        with FdmBackend(plans[i]) as be:
            r = be.advance(t_end, 10 ** 6, 1e-13)
            assert r["steps_taken"] == steps[i], i
            assert np.array_equal(be.field("m"), got[i]), i


def test_config4_fp64_systems_match_oracle():
    # the sweep in FP64; its 64x64x1 and 128x64x2 systems against the
    # oracle's runner loop to the same end time
    from fullmag_amd.backend import FdmBatch, FdmPlan
    from fullmag_amd.workloads import batch_sweep
    from oracle.oracle import Oracle, Problem
    kws = batch_sweep(precision="double")
    plans = [FdmPlan(**kw) for kw in kws]
    t_end = 1.5e-12
    pick = [i for i in _one_per_shape(plans) if (plans[i].nx, plans[i].ny) in ((64, 64), (128, 64))]
    assert len(pick) == 2
    with FdmBatch(plans) as b:
        _, steps = b.advance(t_end, 10 ** 6, 1e-13)
        got = {i: b.systems[i].field("m") for i in pick}
    for i in pick:
        kw = kws[i]
        orc = Oracle(Problem(kw["nx"], kw["ny"], kw["nz"], kw["dx"], kw["dy"], kw["dz"], kw["ms"], kw["a_ex"],
                             kw["alpha"], atol=kw["atol"], dt_min=kw["dt_min"], dt_max=kw["dt_max"],
                             headroom=kw["headroom"]))
        m = orc.prepare(kw["m0"])
        r = orc.run("rk45", m, 0.0, t_end, 1e-13, 10 ** 6)
        err = float(np.abs(got[i] - m).max())
        print(f"system {i} {kw['nx']}x{kw['ny']}x{kw['nz']}: {r['steps']} steps, |dm| vs oracle {err:.3e}",
              flush=True)
        assert r["steps"] == steps[i]
        assert err <= 1e-10


def test_stack_and_plain_grid_of_one_shape_form_two_groups():
    # a multilayer stack and a plain 3-D grid (both with the further-term
    # kernels) share nx, ny, nz and the integrator but not the y/z stage
    from fullmag_amd.backend import FdmBackend, FdmBatch
    from tests.cases import random_case
    stack = random_case(16, 12, 3, seed=1, alpha=0.1, dz=3e-9, terms=dict(layer_z=np.array([0.0, 7e-9, 15e-9])))
    plain = random_case(16, 12, 3, seed=2, alpha=0.1, dz=3e-9, terms=dict(dmi_i=1e-3))
    plans = [stack.plan(), plain.plan(), stack.plan(), plain.plan()]
    with FdmBatch(plans) as b:
        assert b.groups() == 2
        _, steps = b.advance(1e-12, 100000, 5e-14)
        got = [s.field("m") for s in b.systems]
    for i, p in enumerate(plans):
        with FdmBackend(p) as be:
            r = be.advance(1e-12, 100000, 5e-14)
            assert r["steps_taken"] == steps[i], i
            assert np.array_equal(be.field("m"), got[i]), i


def test_fp32_three_step_fft_shapes_batched():
    # FP32 films with px = 1024 (x passes with the three-step FFT and the
    # prefetching joined pass) and py = 1024 (the three-step y pass), two
    # systems each with their own Ms and m0: the batch equals each alone
    from fullmag_amd.backend import FdmBackend, FdmBatch, FdmPlan
    from fullmag_amd.workloads import batch_sweep
    shapes = ((512, 8, 1), (16, 512, 1))
    plans = [FdmPlan(**kw) for kw in batch_sweep(count=4, seed=5, shapes=shapes, precision="single")]
    t_end = 5e-13
    assert sorted((p.nx, p.ny) for p in plans) == [(16, 512), (16, 512), (512, 8), (512, 8)]
    with FdmBatch(plans) as b:
        assert b.groups() == 2
        _, steps = b.advance(t_end, 100000, 1e-13)
        got = [s.field("m") for s in b.systems]
    assert np.all(steps > 0)
    for i, p in enumerate(plans):
        with FdmBackend(p) as be:
            r = be.advance(t_end, 100000, 1e-13)
            assert r["steps_taken"] == steps[i], i
            assert np.array_equal(be.field("m"), got[i]), i
