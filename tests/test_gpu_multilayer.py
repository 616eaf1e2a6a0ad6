"""GPU parity of the multilayer stack (single-cell layers at arbitrary
heights coupled by demag through shifted pair kernels, the reference's
crates/fullmag-engine/src/multilayer.rs and crates/fullmag-fdm-demag/src/
shifted_kernel.rs) against the FP64 oracle, through the C ABI.

Tolerances as the other step tests: FP64 fields 1e-11 relative, states
1e-11, energies 1e-9 relative; FP32 within the Tier-B band (1e-5).
"""
import numpy as np
import pytest

from tests.cases import random_case

pytestmark = pytest.mark.gpu


def _case(nx, ny, z, seed=4, hole=None, **terms):
    t = dict(layer_z=np.asarray(z, dtype=np.float64), **terms)
    return random_case(nx, ny, len(z), seed=seed, hole_radius_cells=hole, alpha=0.1, dx=2e-9, dy=3e-9, dz=4e-9,
                       terms=t)


STACKS = [
    (20, 16, [0.0, 9.3e-9], 3),
    (16, 12, [0.0, 6e-9, 17.5e-9], None),
    (24, 12, [-20e-9, -3e-9, 4e-9, 30e-9, 52e-9], None),
    (12, 10, [0.0, 4e-9, 8e-9, 12e-9, 16e-9, 20e-9, 24e-9, 28e-9], None),
]


def _fields(case, precision="double"):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    fo = orc.fields(m)
    with FdmBackend(case.plan(precision=precision)) as be:
        return {k: (be.field(k), fo[k]) for k in ("h_ex", "h_demag", "h_eff")}


@pytest.mark.parametrize("nx,ny,z,hole", STACKS)
def test_fp64_fields_match_oracle(nx, ny, z, hole):
    for k, (got, want) in _fields(_case(nx, ny, z, hole=hole)).items():
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max(), k


def test_layers_one_cell_apart_match_the_uniform_grid_on_the_device():
    from fullmag_amd.backend import FdmBackend
    case = _case(16, 12, [0.0, 4e-9, 8e-9])
    flat = random_case(16, 12, 3, seed=4, alpha=0.1, dx=2e-9, dy=3e-9, dz=4e-9)
    with FdmBackend(case.plan()) as a, FdmBackend(flat.plan()) as b:
        ha, hb = a.field("h_demag"), b.field("h_demag")
    assert np.abs(ha - hb).max() <= 1e-12 * np.abs(hb).max()


@pytest.mark.parametrize("integrator,terms", [("dp45", {}), ("rk4", dict(dmi_i=1.2e-3)),
                                              ("dp45", dict(j=(2e11, -3e11, 5e11), stt_degree=0.5, stt_beta=0.04))])
def test_fp64_steps_match_oracle(integrator, terms):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = _case(16, 14, [0.0, 7.5e-9, 19e-9], hole=3, **terms)
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
            assert abs(st["demag_energy_joules"] - rep["e_demag"]) <= 1e-9 * abs(rep["e_total"]), i
            if integrator == "dp45":
                dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]


def test_fp32_within_tier_b_band():
    for k, (got, want) in _fields(_case(32, 32, [0.0, 8.5e-9, 21e-9], hole=5), precision="single").items():
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max(), k


def test_unusable_layer_plans_are_rejected():
    from fullmag_amd.backend import FdmBackend, FdmError
    bad = [random_case(8, 8, 1, terms=dict(layer_z=np.zeros(1))),
           random_case(8, 8, 9, terms=dict(layer_z=np.arange(9) * 5e-9)),
           _case(8, 8, [0.0, np.nan]),
           _case(8, 8, [0.0, 5e-9], periodic=(False, False, True))]
    for c in bad:
        with pytest.raises(FdmError) as e:
            FdmBackend(c.plan())
        assert "layer_z" in str(e.value)
