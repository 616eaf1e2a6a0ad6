"""GPU parity of the boundary-correction tiers against the FP64 oracle,
through the C ABI: T0 (face-link weighted exchange), T1 (boundary-distance
stencil), the volume-fraction weighted demag source and energies, and the
sparse demag correction (the reference's native/backends/fdm/src/
exchange_t0_fp64.cu, exchange_t1_fp64.cu, demag_boundary_fp64.cu).

Geometry: a disk (and a stack of disks) with volume fractions, face links
and boundary distances sampled from the circle.  Tolerances as the other
step tests: FP64 fields and states within 1e-11 relative / absolute,
energies 1e-9 relative; FP32 within the Tier-B band (1e-5).
"""
import numpy as np
import pytest

from tests.cases import Case, disk_boundary

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _disk_case(tier, nx=20, ny=18, nz=1, corr=False, seed=4, phi_floor=0.0, **extra):
    from fullmag_amd.backend import random_unit_vectors
    dx = dy = 2e-9
    dz = 3e-9
    mask, phi, links, deltas = disk_boundary(nx, ny, nz, dx, dy, dz, 0.42 * nx * dx)
    n = nx * ny * nz
    m0 = random_unit_vectors(seed, n)
    m0[mask == 0] = 0.0
    terms = dict(boundary_tier=tier, volume_fraction=phi, phi_floor=phi_floor)
    if tier == 1:
        terms["face_links"] = links
    if tier == 2:
        terms["deltas"] = deltas
    if corr:
        terms["demag_corr"] = _correction(nx, ny, nz, phi, seed)
    terms.update(extra)
    return Case(nx, ny, nz, dx, dy, dz, 8.0e5, 1.3e-11, 0.1, m0, mask=mask, h_ext=(0.01 / MU0, 0.0, 0.0),
                terms=terms)


def _correction(nx, ny, nz, phi, seed):
    """Targets on the partially filled rim, each with its 3x3 in-plane
    neighbourhood as sources (outside the box: -1); one target repeated."""
    rng = np.random.default_rng(seed)
    rim = np.flatnonzero((phi > 0) & (phi < 1))
    tg = rng.choice(rim, size=min(12, rim.size), replace=False)
    tg = np.concatenate([tg, tg[:1]]).astype(np.int32)
    src = np.full((tg.size, 9), -1, np.int32)
    for t, c in enumerate(tg):
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        k = 0
        for dy_ in (-1, 0, 1):
            for dx_ in (-1, 0, 1):
                xx, yy = x + dx_, y + dy_
                if 0 <= xx < nx and 0 <= yy < ny:
                    src[t, k] = (z * ny + yy) * nx + xx
                k += 1
    ten = rng.normal(0.0, 0.02, (tg.size, 9, 6))
    return tg, src, ten


def _check_fields(case, precision="double", rel=1e-11):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    fo = orc.fields(m)
    with FdmBackend(case.plan(precision=precision)) as be:
        for k in ("h_ex", "h_demag", "h_eff"):
            got, want = be.field(k), fo[k]
            assert np.abs(got - want).max() <= rel * np.abs(want).max(), k


def _check_steps(case, integrator, nsteps, dt, tol_m=1e-11):
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
            assert np.abs(be.field("m") - m).max() <= tol_m, i
            assert abs(st["dt_seconds"] - rep["dt_used"]) <= 1e-9 * rep["dt_used"], i
            for k, o in (("total_energy_joules", "e_total"), ("exchange_energy_joules", "e_ex"),
                         ("demag_energy_joules", "e_demag"), ("external_energy_joules", "e_ext")):
                assert abs(st[k] - rep[o]) <= 1e-9 * abs(rep["e_total"]), (i, k)
            assert abs(st["max_demag_field_amplitude"] - rep["max_hdemag"]) <= 1e-9 * rep["max_hdemag"], i
            if integrator == "dp45":
                dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]


@pytest.mark.parametrize("tier,nz,corr", [(1, 1, False), (2, 1, False), (1, 3, True), (2, 3, True), (1, 1, True)])
def test_fp64_fields_match_oracle(tier, nz, corr):
    _check_fields(_disk_case(tier, nz=nz, corr=corr))


@pytest.mark.parametrize("tier,integrator", [(1, "dp45"), (2, "dp45"), (2, "rk4"), (1, "heun")])
def test_fp64_steps_match_oracle(tier, integrator):
    _check_steps(_disk_case(tier, corr=True, phi_floor=0.1), integrator, 5, 5e-14)


def test_volume_fraction_without_stencil_data():
    # a tier without face links / distances keeps the standard exchange and
    # weights the demag source and the energies by phi
    case = _disk_case(1, nz=2, corr=True)
    case.terms.pop("face_links")
    _check_fields(case)
    _check_steps(case, "rk4", 3, 5e-14)


def test_region_table_with_a_tier():
    # two regions whose table holds the plain stiffness everywhere give the
    # tier's result without regions
    from fullmag_amd.backend import FdmBackend
    case = _disk_case(1, corr=True)
    regions = (np.arange(case.n) % 3 == 0).astype(np.uint32)
    lut = np.zeros(256 * 256)
    lut[[0, 1, 256, 257]] = 1.3e-11
    with FdmBackend(case.plan()) as a, FdmBackend(case.plan(regions=regions, exchange_lut=lut)) as b:
        ha, hb = a.field("h_ex"), b.field("h_ex")
        assert np.abs(ha - hb).max() <= 1e-13 * np.abs(ha).max()
        for _ in range(3):
            a.step(5e-14)
            b.step(5e-14)
        assert np.abs(a.field("m") - b.field("m")).max() <= 1e-13


@pytest.mark.parametrize("tier", [1, 2])
def test_fp32_within_tier_b_band(tier):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = _disk_case(tier, nx=32, ny=32, corr=True)
    _check_fields(case, precision="single", rel=1e-5)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    with FdmBackend(case.plan(precision="single", integrator="rk4")) as be:
        for _ in range(3):
            st = be.step(5e-14)
            t, rep = orc.step("rk4", m, t, 5e-14)
            assert np.abs(be.field("m") - m).max() <= 1e-5
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-4 * abs(rep["e_total"])


def test_batched_tier_systems_equal_individual_runs():
    from fullmag_amd.backend import FdmBackend, FdmBatch
    plans = [_disk_case(t, seed=s, corr=c).plan(precision="double")
             for t, s, c in ((1, 1, True), (2, 2, False), (1, 3, False), (2, 5, True))]
    with FdmBatch(plans) as b:
        _, steps = b.advance(1e-12, 100000, 5e-14)
        got = [s.field("m") for s in b.systems]
    for i, p in enumerate(plans):
        with FdmBackend(p) as be:
            r = be.advance(1e-12, 100000, 5e-14)
            assert r["steps_taken"] == steps[i], i
            assert np.array_equal(be.field("m"), got[i]), i


def test_bad_boundary_data_is_rejected():
    from fullmag_amd.backend import FdmBackend, FdmError
    case = _disk_case(1, corr=True)
    phi = case.terms["volume_fraction"].copy()
    phi[5] = 1.5
    bad = [dict(volume_fraction=phi), dict(volume_fraction=phi[:-1])]
    tg, src, ten = case.terms["demag_corr"]
    src2 = src.copy()
    src2[0, 0] = case.n
    bad.append(dict(demag_corr=(tg, src2, ten)))
    tg2 = tg.copy()
    tg2[1] = case.n + 3
    bad.append(dict(demag_corr=(tg2, src, ten)))
    for b in bad:
        c = _disk_case(1, corr=True, **b)
        with pytest.raises(FdmError):
            FdmBackend(c.plan())
