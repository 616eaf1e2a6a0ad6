"""GPU parity of the region-table exchange (the plan's region_mask and
256 x 256 exchange_lut, native/backends/fdm/src/exchange_fp64.cu:80-114)
against the oracle's restatement, which tests/test_oracle_regions.py pins
by closed forms.  Non-uniform tables, symmetric and asymmetric (the lookup
is A[r_cell][r_neighbour]), with a hole mask, 2-D and 3-D, both precisions;
fields and DP45 steps.  Tolerances as in test_gpu_fields / test_gpu_steps:
FP64 1e-11 of the field scale and |dm| < 1e-11; FP32 2e-5 and 1e-5."""
import numpy as np
import pytest

from tests.cases import random_case, rel_err

pytestmark = pytest.mark.gpu

A0 = 1.3e-11


def _table(asym):
    rng = np.random.default_rng(12)
    t = rng.uniform(0.2, 2.0, (4, 4)) * A0
    if not asym:
        t = 0.5 * (t + t.T)
    t[1, 3] = t[3, 1] = 0.0  # two regions that do not couple
    lut = np.zeros((256, 256))
    lut[:4, :4] = t
    lut[200, 200] = 1.7 * A0
    return lut.ravel()


def _case(shape, asym):
    nx, ny, nz = shape
    n = nx * ny * nz
    rng = np.random.default_rng(3)
    regions = rng.integers(0, 4, n).astype(np.uint32)
    regions[::17] = 200
    return random_case(nx, ny, nz, seed=21, hole_radius_cells=max(2, nx // 6), alpha=0.1,
                       terms=dict(regions=regions, exchange_lut=_table(asym)))


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("shape,asym", [((48, 40, 1), False), ((48, 40, 1), True), ((20, 16, 6), False),
                                        ((20, 16, 6), True)])
def test_region_fields_match_oracle(precision, shape, asym):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = _case(shape, asym)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    ref = orc.fields(m)
    tol = 1e-11 if precision == "double" else 2e-5
    with FdmBackend(case.plan(precision=precision)) as be:
        for name in ("h_ex", "h_eff"):
            e = rel_err(be.field(name), ref[name])
            print(f"{precision} {shape} asym={asym} {name}: {e:.3e}", flush=True)
            assert e < tol, (name, e)
        ob = orc.observables(m)
        st = be.stats()
        assert abs(st["exchange_energy_joules"] - ob["e_ex"]) <= 50 * tol * abs(ob["e_total"])


@pytest.mark.parametrize("precision,shape", [("double", (48, 40, 1)), ("double", (20, 16, 6)),
                                             ("single", (48, 40, 1))])
def test_region_dp45_steps_match_oracle(precision, shape):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = _case(shape, True)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    dt_e = dt_o = 1e-13
    tol = 1e-11 if precision == "double" else 1e-5
    with FdmBackend(case.plan(precision=precision, integrator="dp45")) as be:
        for _ in range(5):
            st = be.step(dt_e if precision == "double" else dt_o)
            t, rep = orc.step("rk45", m, t, dt_o)
            dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]
            assert np.abs(be.field("m") - m).max() < tol
            if precision == "double":
                assert abs(st["time_seconds"] - t) <= 1e-9 * t
