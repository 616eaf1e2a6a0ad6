"""GPU parity of the plan extensions that have no reference counterpart
(per-cell Ms and anisotropy axis, the reference ABI's per-cell Ku1/Ku2,
time-dependent Zeeman waveforms, mixed atol/rtol step control) against the
FP64 oracle extended with the same definitions.  The oracle side is pinned
by reduction: uniform maps give the scalar oracle's results exactly."""
import numpy as np
import pytest

from tests.cases import random_case, rel_err

pytestmark = pytest.mark.gpu


def _granular(nx=10, ny=8, nz=4, seed=3, grains=6):
    """Voronoi-like grains: per-cell Ku1, Ms and axis by nearest seed."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, (grains, 3)) * [nx, ny, nz]
    axes = rng.normal(size=(grains, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = np.stack([x.ravel() + 0.5, y.ravel() + 0.5, z.ravel() + 0.5], axis=1)
    g = np.argmin(((c[:, None, :] - pts[None]) ** 2).sum(-1), axis=1)
    case = random_case(nx, ny, nz, seed=seed, dx=3e-9, dy=3e-9, dz=3e-9, ms=1e6, a_ex=1.5e-11, alpha=0.1,
                       axis=(0.0, 0.0, 1.0), ku1=5e5)
    case.axis_cell = axes[g]
    case.ku1_cell = 5e5 * (1.0 + 0.2 * (g % 3))
    case.ku2_cell = 1e4 * (g % 2)
    case.ms_cell = 1e6 * (1.0 - 0.05 * (g % 4))
    return case


def _fields_vs_oracle(case, precision, t=0.0):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    orc.set_time(t)
    m = orc.prepare(case.m0)
    ref = orc.fields(m)
    with FdmBackend(case.plan(precision=precision)) as be:
        return {k: rel_err(be.field(k), ref[k]) for k in ("h_ex", "h_demag", "h_eff")}


@pytest.mark.parametrize("precision,tol", [("double", 1e-11), ("single", 2e-5)])
def test_granular_per_cell_maps_fields(precision, tol):
    errs = _fields_vs_oracle(_granular(), precision)
    assert max(errs.values()) < tol, errs


def _steps_vs_oracle(case, integ, nsteps, dt0, tol):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    with FdmBackend(case.plan(integrator=integ)) as be:
        dt_e = dt_o = dt0
        for _ in range(nsteps):
            st = be.step(dt_e)
            t, rep = orc.step("rk45" if integ == "dp45" else integ, m, t, dt_o)
            if integ == "dp45":
                dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]
            assert abs(st["time_seconds"] - t) <= 1e-9 * t
            assert np.abs(be.field("m") - m).max() < tol
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-9 * abs(rep["e_total"])


def test_granular_dp45_steps_fp64():
    _steps_vs_oracle(_granular(), "dp45", 5, 1e-13, 1e-11)


@pytest.mark.parametrize("integ", ["dp45", "rk4", "heun"])
def test_field_ramp_stage_times(integ):
    # ramp 0 -> 1 T along z over 20 ps: every stage sees its own t + c_s dt
    case = random_case(16, 12, 2, seed=5, dx=3e-9, dy=3e-9, dz=3e-9, h_ext=None)
    case.wave_kind, case.wave_amp, case.wave_t0, case.wave_t1 = 1, (0.0, 0.0, 1.0 / (4e-7 * np.pi)), 0.0, 2e-11
    _steps_vs_oracle(case, integ, 6, 2e-13, 1e-11)


def test_sinc_pulse_on_masked_columns():
    from tests.cases import random_case as rc
    case = rc(32, 8, 2, seed=9, dx=5e-9, dy=5e-9, dz=5e-9, ms=8e5, alpha=0.01, h_ext=None)
    wm = np.zeros((2, 8, 32), dtype=np.uint8)
    wm[:, :, :8] = 1
    case.wave_kind, case.wave_amp, case.wave_t0, case.wave_freq = 2, (5e-3 / (4e-7 * np.pi), 0.0, 0.0), 1e-12, 20e9
    case.wave_mask = wm.ravel()
    _steps_vs_oracle(case, "rk4", 8, 5e-14, 1e-11)


def test_uniform_maps_equal_scalar_plan():
    from fullmag_amd.backend import FdmBackend
    base = random_case(12, 10, 2, seed=2, dx=3e-9, dy=3e-9, dz=3e-9, axis=(0.3, 0.4, 0.866), ku1=2e5)
    mapped = random_case(12, 10, 2, seed=2, dx=3e-9, dy=3e-9, dz=3e-9, axis=(0.3, 0.4, 0.866), ku1=2e5)
    n = mapped.n
    mapped.ms_cell = np.full(n, mapped.ms)
    mapped.ku1_cell = np.full(n, 2e5)
    mapped.axis_cell = np.tile([0.3, 0.4, 0.866], (n, 1))
    with FdmBackend(base.plan()) as a, FdmBackend(mapped.plan()) as b:
        for nm in ("h_eff", "h_demag"):
            assert rel_err(b.field(nm), a.field(nm)) < 1e-13
        sa, sb = a.step(1e-13), b.step(1e-13)
        assert abs(sa["total_energy_joules"] - sb["total_energy_joules"]) <= 1e-12 * abs(sa["total_energy_joules"])


def test_mixed_tolerance_controller_matches_oracle():
    case = random_case(20, 16, 1, seed=4, alpha=0.5)
    case.rtol = 1e-6
    case.atol = 1e-6
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t, dt_o = 0.0, 1e-13
    with FdmBackend(case.plan()) as be:
        dt_e = 1e-13
        for _ in range(5):
            st = be.step(dt_e)
            t, rep = orc.step("rk45", m, t, dt_o)
            dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]
            assert abs(st["dt_seconds"] - rep["dt_used"]) <= 1e-9 * rep["dt_used"]
            assert np.abs(be.field("m") - m).max() < 1e-11
