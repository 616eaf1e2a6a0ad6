"""GPU parity of periodic boundaries (the CPU engine's FdmBoundaryPolicy and
truncated-image demag kernel, crates/fullmag-engine/src/fdm_types.rs:94-132,
fdm_fft.rs:140-229,611-717) against the FP64 oracle, through the C ABI.

The oracle convolves circularly on the unpadded periodic grid as the
reference does; the engine keeps its zero-padded transforms and places the
image-summed kernel so that the padded convolution equals the circular one
(complex spectrum).  Tolerances as the other step tests.
"""
import numpy as np
import pytest

from tests.cases import random_case

pytestmark = pytest.mark.gpu


def _case(nx, ny, nz, per, images, seed=3, hole=None, **terms):
    t = dict(periodic=per, periodic_images=images, **terms)
    return random_case(nx, ny, nz, seed=seed, hole_radius_cells=hole, alpha=0.1, dx=2e-9, dy=3e-9, dz=4e-9, terms=t)


CASES = [
    ((16, 12, 1), (True, False, False), (3, 0, 0), 3),
    ((16, 12, 1), (True, True, False), (2, 2, 0), 3),
    ((12, 10, 1), (False, True, True), (0, 2, 4), None),
    ((10, 8, 4), (True, True, False), (2, 1, 0), None),
    ((8, 6, 6), (True, False, True), (1, 0, 2), None),
    ((12, 10, 1), (True, True, False), (0, 0, 0), 2),   # default images (10)
]


def _fields(case, precision="double"):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    fo = orc.fields(m)
    with FdmBackend(case.plan(precision=precision)) as be:
        return {k: (be.field(k), fo[k]) for k in ("h_ex", "h_demag", "h_eff")}


@pytest.mark.parametrize("shape,per,images,hole", CASES)
def test_fp64_fields_match_oracle(shape, per, images, hole):
    out = _fields(_case(*shape, per, images, hole=hole))
    for k, (got, want) in out.items():
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max(), k


@pytest.mark.parametrize("integrator,terms", [("dp45", {}), ("rk4", dict(dmi_i=1.5e-3)),
                                              ("dp45", dict(dmi_b=1e-3, j=(3e11, -2e11, 0.0), stt_degree=0.5,
                                                            stt_beta=0.04))])
def test_fp64_steps_match_oracle(integrator, terms):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = _case(14, 12, 2, (True, True, False), (2, 1, 0), **terms)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    with FdmBackend(case.plan(integrator=integrator)) as be:
        dt_e = dt_o = 5e-14
        for i in range(5):
            st = be.step(dt_e)
            t, rep = orc.step("rk45" if integrator == "dp45" else integrator, m, t, dt_o)
            assert np.abs(be.field("m") - m).max() < 1e-11, i
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-9 * abs(rep["e_total"]), i
            assert abs(st["max_demag_field_amplitude"] - rep["max_hdemag"]) <= 1e-9 * rep["max_hdemag"], i
            if integrator == "dp45":
                dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]


def test_fp32_within_tier_b_band():
    out = _fields(_case(32, 16, 1, (True, True, False), (3, 3, 0), hole=4), precision="single")
    for k, (got, want) in out.items():
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max(), k


def test_batched_periodic_systems_equal_individual_runs():
    from fullmag_amd.backend import FdmBackend, FdmBatch
    plans = [_case(16, 12, 1, per, img, seed=s).plan()
             for per, img, s in (((True, False, False), (2, 0, 0), 1), ((True, True, False), (1, 1, 0), 2),
                                 ((False, False, False), (0, 0, 0), 3))]
    with FdmBatch(plans) as b:
        _, steps = b.advance(1e-12, 100000, 5e-14)
        got = [s.field("m") for s in b.systems]
    for i, p in enumerate(plans):
        with FdmBackend(p) as be:
            r = be.advance(1e-12, 100000, 5e-14)
            assert r["steps_taken"] == steps[i], i
            assert np.array_equal(be.field("m"), got[i]), i


def test_unusable_periodic_plans_are_rejected():
    from fullmag_amd.backend import FdmBackend, FdmError
    case = _case(16, 12, 1, (True, False, False), (2, 0, 0))
    bad = case.plan(extra={"boundary_correction": 1})
    with pytest.raises(FdmError) as e:
        FdmBackend(bad)
    assert "periodic" in str(e.value)
    p = case.plan()
    p.periodic_images = (2000, 0, 0)
    with pytest.raises(FdmError):
        FdmBackend(p)
