"""GPU parity of feature combinations against the FP64 oracle (or, where the
oracle has no counterpart, against the engine's own reduction): thermal
field with periodic boundaries, with a multilayer stack and with a boundary
tier; batched multilayer stacks; a region table with periodic boundaries.
Tolerances as the other FP64 step tests."""
import numpy as np
import pytest

from tests.cases import Case, disk_boundary, random_case

pytestmark = pytest.mark.gpu


def _steps(case, integrator, nsteps, dt):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    with FdmBackend(case.plan(integrator=integrator)) as be:
        dt_e = dt_o = dt
        for i in range(nsteps):
            st = be.step(dt_e)
            t, rep = orc.step("rk45" if integrator == "dp45" else integrator, m, t, dt_o)
            assert np.abs(be.field("m") - m).max() < 1e-11, i
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-9 * abs(rep["e_total"]), i
            if integrator == "dp45":
                dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]


def test_thermal_with_periodic_boundaries():
    case = random_case(16, 12, 1, seed=6, alpha=0.1, terms=dict(temperature=100.0, thermal_seed=5,
                                                                periodic=(True, True, False),
                                                                periodic_images=(2, 2, 0)))
    _steps(case, "rk4", 5, 2e-14)
    _steps(case, "dp45", 5, 2e-14)


def test_thermal_with_a_multilayer_stack():
    case = random_case(16, 12, 3, seed=8, alpha=0.1, dz=3e-9,
                       terms=dict(temperature=50.0, layer_z=np.array([0.0, 7e-9, 15.5e-9])))
    _steps(case, "dp45", 5, 2e-14)
    _steps(case, "heun", 5, 2e-14)


def test_thermal_with_a_boundary_tier():
    from fullmag_amd.backend import random_unit_vectors
    nx, ny = 20, 18
    mask, phi, links, _ = disk_boundary(nx, ny, 1, 2e-9, 2e-9, 3e-9, 0.42 * nx * 2e-9)
    m0 = random_unit_vectors(2, nx * ny)
    m0[mask == 0] = 0.0
    case = Case(nx, ny, 1, 2e-9, 2e-9, 3e-9, 8e5, 1.3e-11, 0.1, m0, mask=mask,
                terms=dict(boundary_tier=1, volume_fraction=phi, face_links=links, phi_floor=0.1,
                           temperature=80.0))
    _steps(case, "rk4", 4, 2e-14)


def test_batched_multilayer_stacks_equal_individual_runs():
    from fullmag_amd.backend import FdmBackend, FdmBatch
    plans = [random_case(16, 12, 3, seed=s, alpha=0.1, dz=3e-9, terms=dict(layer_z=np.array(z))).plan()
             for s, z in ((1, [0.0, 6e-9, 14e-9]), (2, [0.0, 9e-9, 10e-8]), (3, [-5e-9, 0.0, 5e-9]))]
    with FdmBatch(plans) as b:
        _, steps = b.advance(1e-12, 100000, 5e-14)
        got = [s.field("m") for s in b.systems]
    for i, p in enumerate(plans):
        with FdmBackend(p) as be:
            r = be.advance(1e-12, 100000, 5e-14)
            assert r["steps_taken"] == steps[i], i
            assert np.array_equal(be.field("m"), got[i]), i


def test_region_table_with_periodic_boundaries():
    # two regions whose table holds the plain stiffness everywhere give the
    # periodic result without regions
    from fullmag_amd.backend import FdmBackend
    case = random_case(16, 12, 2, seed=9, alpha=0.1, terms=dict(periodic=(True, False, True),
                                                                periodic_images=(2, 0, 1)))
    regions = (np.arange(case.n) % 5 == 0).astype(np.uint32)
    lut = np.zeros(256 * 256)
    lut[[0, 1, 256, 257]] = case.a_ex
    with FdmBackend(case.plan()) as a, FdmBackend(case.plan(regions=regions, exchange_lut=lut)) as b:
        ha, hb = a.field("h_ex"), b.field("h_ex")
        assert np.abs(ha - hb).max() <= 1e-13 * np.abs(ha).max()
