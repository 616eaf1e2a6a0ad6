"""GPU parity of the thermal field (fdm_fields.rs:818-905) against the FP64
oracle, through the C ABI.

The noise is a pure function of (seed, step counter, cell, stream), so the
device and the oracle draw the same Gaussians and the trajectories agree to
rounding: FP64 states within 1e-11 and energies within 1e-9 relative, as the
other FP64 step tests; FP32 within the Tier-B band (1e-5) over a few steps.
The cases cover where k1 is evaluated with a thermal field: every attempt
(Heun, RK4, RK23), the ABM3 start-up only, and DP45's FSAL stage, whose
noise belongs to the previous counter, with fresh k1 after a rejection.
"""
import numpy as np
import pytest

from tests.cases import random_case

pytestmark = pytest.mark.gpu


def _case(temperature=300.0, nx=20, ny=16, nz=1, seed=3, demag=True, **terms):
    t = dict(temperature=temperature, **terms)
    return random_case(nx, ny, nz, seed=seed, hole_radius_cells=3 if nz == 1 else None, alpha=0.1, demag=demag,
                       terms=t)


def _orc_name(integrator):
    return "rk45" if integrator == "dp45" else integrator


def _compare_steps(case, integrator, nsteps, dt, precision="double", tol_m=1e-11, tol_e=1e-9):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    rejected = 0
    with FdmBackend(case.plan(precision=precision, integrator=integrator)) as be:
        dt_e = dt_o = dt
        for i in range(nsteps):
            st = be.step(dt_e)
            t, rep = orc.step(_orc_name(integrator), m, t, dt_o)
            assert np.abs(be.field("m") - m).max() <= tol_m, (integrator, i)
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= tol_e * abs(rep["e_total"]), (integrator, i)
            assert abs(st["dt_seconds"] - rep["dt_used"]) <= 1e-9 * rep["dt_used"], (integrator, i)
            if integrator in ("dp45", "rk23"):
                dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]
            rejected += rep["rejected"]
        assert orc.thermal_step() == nsteps
    return rejected


@pytest.mark.parametrize("integrator", ["heun", "rk4", "rk23", "abm3"])
def test_fp64_steps_match_oracle(integrator):
    # ABM3: three Heun start-up steps, then predictor-corrector steps whose
    # history values carry the noise of the step that pushed them
    _compare_steps(_case(), integrator, 7, 2e-14)


def test_fp64_dp45_fsal_and_retries_match_oracle():
    # the FSAL stage carries the previous counter's noise, so the first
    # attempts are rejected and retried with a fresh k1; the dt sequence and
    # the states follow the oracle through both paths
    rejected = _compare_steps(_case(), "dp45", 8, 1e-13)
    assert rejected > 0


def test_fixed_noise_dt_and_seed():
    # thermal_dt fixes the noise amplitude independently of the step; a
    # different seed gives a different trajectory
    _compare_steps(_case(thermal_dt=5e-14, thermal_seed=1234), "rk4", 4, 2e-14)
    _compare_steps(_case(thermal_dt=5e-14, thermal_seed=99, nx=12, ny=10, nz=3), "dp45", 4, 5e-14)


def test_advance_matches_oracle_run():
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    for integrator, until in (("dp45", 2e-12), ("rk4", 4.1e-13)):
        case = _case(temperature=50.0, seed=8)
        orc = Oracle(case.problem())
        m = orc.prepare(case.m0)
        ro = orc.run(_orc_name(integrator), m, 0.0, until, 2e-14, 100000)
        with FdmBackend(case.plan(integrator=integrator)) as be:
            r = be.advance(until, 100000, 2e-14)
            assert r["steps_taken"] == ro["steps"], integrator
            assert abs(r["time_seconds"] - ro["t"]) <= 1e-20
            assert np.abs(be.field("m") - m).max() < 1e-10, integrator


def test_observation_and_tuning_keep_the_trajectory():
    # h_eff between steps holds the noise of the current counter at the last
    # accepted dt (the oracle's fields()); observing it or timing the kernels
    # must not change the FSAL stage of the next DP45 step
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = _case(seed=11, demag=False)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    h0 = orc.fields(m)["h_eff"]
    t, dt_e, dt_o = 0.0, 5e-14, 5e-14
    with FdmBackend(case.plan(integrator="dp45")) as be:
        # before any step the noise dt is zero: no thermal field
        assert np.abs(be.field("h_eff") - h0).max() <= 1e-11 * np.abs(h0).max()
        for i in range(5):
            st = be.step(dt_e)
            t, rep = orc.step("rk45", m, t, dt_o)
            dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]
            assert abs(dt_e - dt_o) <= 1e-9 * dt_o
            ho = orc.fields(m)["h_eff"]
            assert np.abs(be.field("h_eff") - ho).max() <= 1e-11 * np.abs(ho).max(), i
            if i == 2:
                be.time_kernels(1)
            assert np.abs(be.field("m") - m).max() < 1e-11, i


def test_noise_statistics_on_device():
    # no other field: h_eff after a step is the noise of the next counter,
    # zero-mean with variance 2 alpha kB T / (gamma (1 + alpha^2) mu0 Ms V dt)
    from fullmag_amd.backend import FdmBackend
    case = random_case(64, 64, 4, seed=4, alpha=0.1, h_ext=(0.0, 0.0, 0.0),
                       terms=dict(temperature=300.0, thermal_dt=1e-13), exchange=False, demag=False)
    with FdmBackend(case.plan(integrator="rk4")) as be:
        be.step(1e-14)
        h = be.field("h_eff").reshape(-1)
    kb, mu0 = 1.380649e-23, 4e-7 * np.pi
    var = 2 * case.alpha * kb * 300.0 / (case.gamma * (1 + case.alpha ** 2) * mu0 * case.ms * case.dx * case.dy
                                         * case.dz * 1e-13)
    sig = np.sqrt(var)
    assert abs(h.mean()) <= 5 * sig / np.sqrt(h.size)
    assert abs(h.std() / sig - 1.0) <= 0.03
    # fourth moment of a Gaussian
    assert abs(np.mean(h ** 4) / var ** 2 - 3.0) <= 0.25


def test_fp32_within_tier_b_band():
    _compare_steps(_case(temperature=20.0, nx=32, ny=32), "rk4", 4, 2e-14, precision="single", tol_m=1e-5,
                   tol_e=1e-4)


def test_fp32_dp45_with_the_oracle_on_the_device_dt():
    # FP32 error norms of a random state at a few 1e-14 s carry rounding of
    # several percent, so the adaptive dt sequences drift apart; the oracle
    # takes the device's accepted dt instead, with a fresh k1 (no FSAL stage)
    # where the device retried, and must accept it as the device did
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = _case(temperature=20.0, nx=32, ny=32)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t, dt = 0.0, 2e-14
    retried = 0
    with FdmBackend(case.plan(precision="single", integrator="dp45")) as be:
        for i in range(4):
            r0 = be.counters()["rejected"]
            st = be.step(dt)
            if be.counters()["rejected"] > r0:
                retried += 1
                orc.invalidate_fsal()
            t, rep = orc.step("rk45", m, t, st["dt_seconds"])
            assert rep["rejected"] == 0 and rep["dt_used"] == st["dt_seconds"], i
            assert np.abs(be.field("m") - m).max() <= 1e-5, i
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-4 * abs(rep["e_total"]), i
            dt = st["suggested_next_dt"]
    assert retried > 0


def test_batched_thermal_systems_equal_individual_runs():
    from fullmag_amd.backend import FdmBackend, FdmBatch, FdmPlan
    from fullmag_amd.workloads import batch_sweep
    plans = [FdmPlan(**kw) for kw in batch_sweep(count=5, seed=9, shapes=((24, 16, 1), (16, 16, 2)),
                                                 precision="double")]
    for i, p in enumerate(plans):
        if i != 2:
            p.temperature = 30.0 + 20.0 * i
            p.thermal_seed = 7 + i
    t_end = 1e-12
    with FdmBatch(plans) as b:
        _, steps = b.advance(t_end, 100000, 5e-14)
        got = [s.field("m") for s in b.systems]
    for i, p in enumerate(plans):
        with FdmBackend(p) as be:
            r = be.advance(t_end, 100000, 5e-14)
            assert r["steps_taken"] == steps[i], i
            assert np.array_equal(be.field("m"), got[i]), i
