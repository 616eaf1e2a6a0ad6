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
