"""GPU parity of the further field and torque terms (cubic anisotropy with
per-cell constants and Kc3, interfacial and bulk DMI, magnetoelastic
coupling, Oersted cylinder with a current envelope, Zhang-Li, Slonczewski
and spin-orbit torques) against the FP64 oracle, through the C ABI.

Tolerances as tests/test_gpu_steps.py: FP64 fields and states within 1e-11,
energies 1e-9 relative, dt sequences 1e-9; FP32 within the Tier-B band
(1e-5) over a few steps.
"""
import numpy as np
import pytest

from tests.cases import random_case

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi

TERMS = {
    "cubic": dict(cubic=(2.5e4, -1.2e4, 6e3, (1.0, 1.0, 0.0), (-1.0, 1.0, 0.3))),
    "dmi_i": dict(dmi_i=2.0e-3),
    "dmi_b": dict(dmi_b=1.5e-3),
    "mel": dict(mel=(-1.5e7, 2.8e7, (2e-3, -1e-3, 5e-4, 6e-4, -4e-4, 8e-4))),
    "zhang_li": dict(j=(4e11, -1e11, 0.0), stt_degree=0.6, stt_beta=0.05),
    "slonczewski": dict(j=(0.0, 0.0, 8e10), stt_degree=0.5, slon_p=(0.0, 0.6, 0.8), slon_lambda=1.7,
                        slon_eps=0.15),
    "sot": dict(sot=(3e11, 0.2, -0.05, (0.0, 1.0, 0.0), 1.2e-9)),
    "oersted": dict(oersted=(4e-3, 12e-9, (24e-9, 20e-9, 0.0), (0.0, 0.0, 1.0), 1, 5e9, 0.3, 0.2, 0.0, 0.0)),
}


def _case(name_or_terms, nx=24, ny=20, nz=1, seed=5, **kw):
    terms = TERMS[name_or_terms] if isinstance(name_or_terms, str) else name_or_terms
    return random_case(nx, ny, nz, seed=seed, hole_radius_cells=4 if nz == 1 else None, alpha=0.1, terms=dict(terms),
                       **kw)


def _steps(case, integrator, precision, nsteps, dt, poll=False):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    out = []
    with FdmBackend(case.plan(precision=precision, integrator=integrator)) as be:
        if poll:
            be.set_interrupt_poll(lambda: False)
        dt_e = dt_o = dt
        for _ in range(nsteps):
            st = be.step(dt_e)
            t, rep = orc.step("rk45" if integrator == "dp45" else integrator, m, t, dt_o)
            if integrator in ("dp45", "rk23"):
                dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]
            out.append((st, rep, be.field("m"), m.copy()))
    return out


@pytest.mark.parametrize("name", list(TERMS))
def test_fp64_steps_match_oracle(name):
    case = _case(name)
    for st, rep, me, mo in _steps(case, "dp45", "double", 5, 1e-13):
        assert np.abs(me - mo).max() < 1e-11
        assert abs(st["dt_seconds"] - rep["dt_used"]) <= 1e-9 * rep["dt_used"]
        assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-9 * abs(rep["e_total"])
        assert abs(st["max_rhs_amplitude"] - rep["max_rhs"]) <= 1e-9 * rep["max_rhs"]
        assert abs(st["max_effective_field_amplitude"] - rep["max_heff"]) <= 1e-9 * rep["max_heff"]
        for k, o in (("cubic_energy_joules", "e_cub"), ("dmi_energy_joules", "e_dmi")):
            assert abs(st[k] - rep[o]) <= 1e-9 * abs(rep["e_total"]) + 1e-30


def test_all_terms_together_3d_and_fields():
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    terms = {}
    for nm in ("cubic", "dmi_b", "mel", "zhang_li", "sot"):
        terms.update(TERMS[nm])
    terms["dmi_i"] = 1.0e-3
    terms["oersted"] = (2e-3, 8e-9, (12e-9, 10e-9, 6e-9), (0.3, 0.0, 1.0))
    rng = np.random.default_rng(2)
    n = 12 * 10 * 4
    terms["kc1_cell"] = rng.uniform(1e4, 4e4, n)
    terms["kc3_cell"] = rng.uniform(-5e3, 5e3, n)
    case = random_case(12, 10, 4, seed=9, dx=2e-9, dy=2e-9, dz=3e-9, alpha=0.1, terms=terms)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    with FdmBackend(case.plan(precision="double")) as be:
        h = be.field("h_eff")
        ho = orc.fields(m)["h_eff"]
        assert np.abs(h - ho).max() <= 1e-11 * np.abs(ho).max()
    for st, rep, me, mo in _steps(case, "rk4", "double", 4, 5e-14):
        assert np.abs(me - mo).max() < 1e-11
        assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-9 * abs(rep["e_total"])


@pytest.mark.parametrize("name", ["dmi_i", "zhang_li", "cubic"])
def test_without_demag_and_per_stage_launches(name):
    # the no-demag local kernel and the separate per-stage launches (poll) use
    # the same term code as the joined x passes
    case = _case(name, demag=False)
    for st, rep, me, mo in _steps(case, "dp45", "double", 4, 1e-13, poll=True):
        assert np.abs(me - mo).max() < 1e-11
    case = _case(name)
    a = _steps(case, "dp45", "double", 4, 1e-13, poll=True)
    b = _steps(case, "dp45", "double", 4, 1e-13, poll=False)
    assert np.abs(a[-1][2] - b[-1][2]).max() <= 1e-14


@pytest.mark.parametrize("name", ["cubic", "zhang_li", "oersted"])
def test_fp32_within_tier_b_band(name):
    case = _case(name, nx=32, ny=32)
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t, dt = 0.0, 1e-13
    with FdmBackend(case.plan(precision="single")) as be:
        for _ in range(4):
            st = be.step(dt)
            t, rep = orc.step("rk45", m, t, st["dt_seconds"])
            dt = st["suggested_next_dt"]
            assert np.abs(be.field("m") - m).max() <= 1e-5
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-4 * abs(rep["e_total"])


def test_plans_with_unusable_terms_are_rejected():
    from fullmag_amd.backend import FdmBackend, FdmError
    bad = [dict(sot=(1e11, 0.1, 0.0, (0.0, 0.0, 1.0), 0.0)), dict(oersted=(1e-3, 0.0, (0.0, 0.0, 0.0), (0, 0, 1)))]
    for t in bad:
        with pytest.raises(FdmError):
            FdmBackend(_case(t).plan())
    with pytest.raises(FdmError) as e:
        FdmBackend(_case("cubic").plan(extra={"boundary_correction": 2}))
    assert "volume_fraction" in str(e.value)


def test_adjoint_rejects_further_terms():
    from fullmag_amd.backend import FdmBackend, FdmError
    case = _case("dmi_i")
    with FdmBackend(case.plan(integrator="rk4")) as be:
        with pytest.raises(FdmError):
            be.adjoint_forward(4, 1e-13, np.full(case.n, 1.0 / case.n))
