"""GPU parity of whole integrator steps against the FP64 oracle, through the
C ABI: DP45 (adaptive, FSAL, device-side step control), RK23 (adaptive,
Bogacki-Shampine), ABM3 (multistep with Heun start-up and restart on dt
change), RK4, Heun, the
device-resident runner loop, interrupts, uploads and snapshots.

Tolerances: FP64 engine vs oracle, m within 1e-11 after a few steps (same
algorithm, different FFT/summation order), dt sequence within 1e-9 relative
(the controller's pow() sees roundoff-level differences of the error norm).
FP32 engine vs FP64 oracle: max |dm| <= 1e-5 (the reference's Tier-B band,
native/backends/fdm/tests/tier_b_compare.cu:453-456).
"""
import numpy as np
import pytest

from tests.cases import masked_3x3, random_case, rel_err, thin_film_8x6

pytestmark = pytest.mark.gpu

ADAPTIVE = ("dp45", "rk23")


def _orc_name(integrator):
    return "rk45" if integrator == "dp45" else integrator


def _run_pair(case, integrator, precision, nsteps, dt):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle

    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    out = []
    with FdmBackend(case.plan(precision=precision, integrator=integrator)) as be:
        dt_e = dt_o = dt
        for _ in range(nsteps):
            st = be.step(dt_e)
            t, rep = orc.step(_orc_name(integrator), m, t, dt_o)
            if integrator in ADAPTIVE:
                dt_e, dt_o = st["suggested_next_dt"], rep["dt_next"]
            out.append((st, rep, be.field("m"), m.copy()))
    return out


@pytest.mark.parametrize("integrator", ["dp45", "rk4", "heun", "rk23", "abm3"])
def test_steps_fp64_match_oracle_film(integrator):
    case = random_case(48, 40, 1, seed=21, hole_radius_cells=8)
    res = _run_pair(case, integrator, "double", 8 if integrator == "abm3" else 6, 1e-13)
    for st, rep, me, mo in res:
        assert abs(st["time_seconds"] - rep["time"]) <= 1e-9 * rep["time"]
        assert abs(st["dt_seconds"] - rep["dt_used"]) <= 1e-9 * rep["dt_used"]
        assert np.abs(me - mo).max() < 1e-11
        assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-9 * abs(rep["e_total"])
        assert abs(st["max_rhs_amplitude"] - rep["max_rhs"]) <= 1e-9 * rep["max_rhs"]
        if integrator in ADAPTIVE:
            assert abs(st["suggested_next_dt"] - rep["dt_next"]) <= 1e-9 * rep["dt_next"]


def test_dp45_fp64_3d_and_masked():
    for case in (random_case(10, 8, 4, seed=4, dx=3e-9, dy=3e-9, dz=3e-9, alpha=0.1), masked_3x3(True),
                 thin_film_8x6()):
        res = _run_pair(case, "dp45", "double", 5, 2e-13)
        for st, rep, me, mo in res:
            assert np.abs(me - mo).max() < 1e-11
            assert abs(st["dt_seconds"] - rep["dt_used"]) <= 1e-9 * rep["dt_used"]


@pytest.mark.parametrize("integrator", ADAPTIVE)
def test_rejection_path_matches_oracle(integrator):
    # a too-large first trial step forces rejections; both sides must shrink
    # dt identically and accept the same step
    case = random_case(32, 32, 1, seed=8, alpha=0.3)
    res = _run_pair(case, integrator, "double", 3, 1e-11)
    st, rep, me, mo = res[0]
    assert rep["rejected"] >= 1
    assert abs(st["dt_seconds"] - rep["dt_used"]) <= 1e-9 * rep["dt_used"]
    for st, rep, me, mo in res:
        assert np.abs(me - mo).max() < 1e-10


def test_tier_b_heun_exchange_8x8x4():
    """Mirror of the reference's Tier-B gate (native/backends/fdm/tests/
    tier_b_compare.cu:33-120): 8x8x4 cells of 2 nm, Ms 800e3, A 13e-12,
    alpha 0.5, exchange only, random m0, Heun dt 1e-14, 100 steps; FP64 vs
    FP32 engine: max |dm| < 1e-5 and exchange energy within 1e-4."""
    from fullmag_amd.backend import FdmBackend, random_unit_vectors
    from tests.cases import Case
    m0 = random_unit_vectors(12345, 8 * 8 * 4)
    case = Case(8, 8, 4, 2e-9, 2e-9, 2e-9, 800e3, 13e-12, 0.5, m0, demag=False)
    e64, e32 = [], []
    with FdmBackend(case.plan(precision="double", integrator="heun")) as b64, \
            FdmBackend(case.plan(precision="single", integrator="heun")) as b32:
        for _ in range(100):
            e64.append(b64.step(1e-14)["exchange_energy_joules"])
            e32.append(b32.step(1e-14)["exchange_energy_joules"])
        dm = np.abs(b64.field("m") - b32.field("m")).max()
    erel = max(abs(a - b) / abs(a) for a, b in zip(e64, e32))
    print(f"tier B: max|dm| {dm:.3e}, max rel dE {erel:.3e}")
    assert dm < 1e-5 and erel < 1e-4


def _fp32_vs_oracle(case, integrator, nsteps, dt0):
    """FP32 engine vs FP64 oracle on the engine's own accepted dt sequence."""
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t, dt, errs = 0.0, dt0, []
    with FdmBackend(case.plan(precision="single", integrator=integrator)) as be:
        for _ in range(nsteps):
            st = be.step(dt)
            # the oracle takes exactly the step the engine accepted
            t, rep = orc.step(_orc_name(integrator), m, t, st["dt_seconds"])
            if integrator in ADAPTIVE:
                assert rep["rejected"] == 0 and abs(rep["dt_used"] - st["dt_seconds"]) < 1e-30
                dt = st["suggested_next_dt"]
            errs.append(float(np.abs(be.field("m") - m).max()))
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-4 * abs(rep["e_total"])
    return errs


@pytest.mark.parametrize("integrator", ["dp45", "rk4", "rk23", "abm3"])
def test_fp32_random_state_vs_oracle(integrator):
    # random m0 with a hole: exchange fields ~1e7 A/m; a few steps stay
    # inside the Tier-B band before the dynamics amplify FP32 roundoff
    errs = _fp32_vs_oracle(random_case(64, 64, 1, seed=31, hole_radius_cells=12), integrator, 4, 1e-13)
    print(integrator, errs)
    assert max(errs) <= 1e-5


def test_fp32_smooth_state_dp45_vs_oracle():
    # smooth texture (reference thin-film texture, native_fdm.rs:1326-1338,
    # on a 64x48 film with a hole): 30 DP45 steps stay within 1e-5
    from tests.cases import Case
    nx, ny = 64, 48
    m0 = []
    for y in range(ny):
        for x in range(nx):
            th, ph = 0.11 * x, 0.07 * y
            v = np.array([np.cos(th) * np.cos(ph), np.sin(th) * np.cos(ph), 0.2 * np.sin(ph)])
            m0.append(v / np.linalg.norm(v))
    from fullmag_amd.backend import hole_mask
    mask = hole_mask(nx, ny, 1, 4e-9, 4e-9, 40e-9)
    m0 = np.array(m0)
    m0[mask == 0] = 0
    case = Case(nx, ny, 1, 4e-9, 4e-9, 10e-9, 800e3, 13e-12, 0.1, m0, mask=mask, h_ext=(2.0e3, -1.0e3, 5.0e2))
    errs = _fp32_vs_oracle(case, "dp45", 30, 1e-13)
    print(errs[-1])
    assert max(errs) <= 1e-5


def test_advance_equals_step_loop():
    from fullmag_amd.backend import FdmBackend
    case = random_case(40, 36, 1, seed=2, hole_radius_cells=6)
    with FdmBackend(case.plan()) as a, FdmBackend(case.plan()) as b:
        dt = 1e-13
        for _ in range(12):
            st = a.step(dt)
            dt = st["suggested_next_dt"]
        r = b.advance(1.0, 12, 1e-13)
        assert r["steps_taken"] == 12
        assert r["time_seconds"] == st["time_seconds"]
        assert np.array_equal(a.field("m"), b.field("m"))
        assert r["suggested_next_dt"] == st["suggested_next_dt"]


@pytest.mark.parametrize("integrator", ADAPTIVE)
def test_advance_matches_oracle_runner_loop_until(integrator):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = random_case(24, 20, 1, seed=13)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    until = 3e-12
    ro = orc.run(_orc_name(integrator), m, 0.0, until, 1e-13, 10000)
    with FdmBackend(case.plan(integrator=integrator)) as be:
        r = be.advance(until, 10000, 1e-13)
        assert r["steps_taken"] == ro["steps"]
        assert abs(r["time_seconds"] - ro["t"]) <= 1e-20
        assert np.abs(be.field("m") - m).max() < 1e-10


@pytest.mark.parametrize("integrator", ["rk4", "abm3"])
def test_fixed_step_advance_matches_oracle(integrator):
    # 20 steps of 1e-13 and a last one clamped to 5e-15 (for ABM3 a dt change
    # that restarts the history at the final push)
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = random_case(16, 16, 2, seed=17, dx=4e-9, dy=4e-9, dz=4e-9, alpha=0.05)
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    ro = orc.run(integrator, m, 0.0, 2.05e-12, 1e-13, 1000)
    with FdmBackend(case.plan(integrator=integrator)) as be:
        r = be.advance(2.05e-12, 1000, 1e-13)
        assert r["steps_taken"] == ro["steps"] == 21
        assert abs(r["dt_seconds"] - ro["last"]["dt_used"]) < 1e-25
        assert np.abs(be.field("m") - m).max() < 1e-11


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("integrator", ["dp45", "rk4", "heun", "rk23", "abm3"])
@pytest.mark.parametrize("shape", [(64, 48, 1), (50, 30, 1), (12, 10, 4), (8, 32, 16)])
def test_joined_stage_launches_equal_separate(precision, integrator, shape):
    # A trial step joins the inverse x pass of stage s with the forward x pass
    # of stage s+1 in one launch; with a poll callback the engine runs one
    # separate evaluation per stage instead.  Both orders do the same
    # arithmetic on the same values.
    from fullmag_amd.backend import FdmBackend
    nx, ny, nz = shape
    case = random_case(nx, ny, nz, seed=nx + ny + nz, hole_radius_cells=min(nx, ny) // 5 if nz == 1 else None,
                       ku1=2e4, axis=(0.0, 0.0, 1.0))
    with FdmBackend(case.plan(precision=precision, integrator=integrator)) as a, \
            FdmBackend(case.plan(precision=precision, integrator=integrator)) as b:
        b.set_interrupt_poll(lambda: False)
        dt_a = dt_b = 1e-13
        for _ in range(6):
            sa, sb = a.step(dt_a), b.step(dt_b)
            assert sa["time_seconds"] == sb["time_seconds"]
            if integrator in ADAPTIVE:
                dt_a, dt_b = sa["suggested_next_dt"], sb["suggested_next_dt"]
        ma, mb = a.field("m"), b.field("m")
        d = float(np.abs(ma - mb).max())
        print(f"{precision} {integrator} {shape}: max |dm| {d:.3e}", flush=True)
        assert d <= (1e-14 if precision == "double" else 1e-6)


def test_interrupt_leaves_state_unchanged():
    from fullmag_amd.backend import FdmBackend, Interrupted
    case = random_case(32, 24, 1, seed=3)
    with FdmBackend(case.plan()) as be:
        be.step(1e-13)
        m_before = be.field("m")
        calls = {"n": 0}

        def poll():
            calls["n"] += 1
            return calls["n"] >= 3
        be.set_interrupt_poll(poll)
        with pytest.raises(Interrupted) as ei:
            be.step(1e-13)
        assert ei.value.code == -4
        assert np.array_equal(be.field("m"), m_before)
        be.set_interrupt_poll(None)
        st = be.step(1e-13)
        assert st["step"] == 2


def test_upload_and_snapshots():
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = thin_film_8x6()
    orc = Oracle(case.problem())
    with FdmBackend(case.plan()) as be:
        be.step(1e-13)
        m_new = np.tile([0.0, 0.6, 0.8], (case.n, 1))
        be.upload_m(m_new * 3.0)  # normalised on upload
        assert np.abs(be.field("m") - m_new).max() < 1e-15
        h = be.field("h_eff")
        ref = orc.fields(orc.prepare(m_new))["h_eff"]
        assert rel_err(h, ref) < 1e-11
        snap = be.field_snapshot("h_eff")
        assert np.abs(snap.T - h).max() == 0.0
        pv = be.preview("m", 4, 3, 1)
        assert np.abs(pv - np.array([0.0, 0.6, 0.8])).max() < 1e-15
        st = be.stats()
        assert st["step"] == 1
        info = be.device_info()
        assert "gfx9" in info["name"]


def test_fp32_field_copy_and_snapshot_precision():
    from fullmag_amd.backend import FdmBackend
    case = thin_film_8x6()
    with FdmBackend(case.plan(precision="single")) as be:
        f32 = be.field("m", dtype=np.float32)
        f64 = be.field("m")
        assert f32.dtype == np.float32 and np.array_equal(f32.astype(np.float64), f64)
        snap = be.field_snapshot("m")
        assert snap.dtype == np.float32


def test_rejected_plans_report_errors():
    from fullmag_amd.backend import FdmBackend, FdmError
    case = thin_film_8x6()
    for extra, msg in (({"boundary_correction": 3}, "boundary_correction must be"),
                       ({"boundary_correction": 1}, "needs volume_fraction"),
                       ({"has_demag_boundary_corr": 1, "demag_corr_target_count": 2}, "needs a boundary_correction")):
        with pytest.raises(FdmError) as ei:
            FdmBackend(case.plan(extra=extra))
        assert msg in str(ei.value)
    bad = case.plan()
    bad.mask = np.ones(5, dtype=np.uint8)
    with pytest.raises(FdmError) as ei:
        FdmBackend(bad)
    assert "active_mask_len" in str(ei.value)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("shape", [(64, 48, 1), (50, 30, 1), (16, 32, 8)])
def test_autotune_keeps_state_and_results(precision, shape):
    # the tuner only changes the x tile (rows per work-group, FFT form): the
    # state is untouched and later steps agree with an untuned handle to
    # roundoff (the two-for-one packing pairs different rows)
    from fullmag_amd.backend import FdmBackend
    nx, ny, nz = shape
    case = random_case(nx, ny, nz, seed=nx + nz, hole_radius_cells=min(nx, ny) // 5 if nz == 1 else None)
    with FdmBackend(case.plan(precision=precision)) as a, FdmBackend(case.plan(precision=precision)) as b:
        m0 = a.field("m")
        choice = a.autotune(reps=3)
        assert choice["rows_per_group"] in (1, 2, 4)
        assert np.array_equal(a.field("m"), m0)
        dt_a = dt_b = 1e-13
        for _ in range(4):
            sa, sb = a.step(dt_a), b.step(dt_b)
            dt_a, dt_b = sa["suggested_next_dt"], sb["suggested_next_dt"]
        d = float(np.abs(a.field("m") - b.field("m")).max())
        print(shape, precision, choice, f"max |dm| {d:.2e}", flush=True)
        # FP32 from a random state: roundoff differences grow ~10x per step
        # (chaotic relaxation), so 4 steps end near 1e-5
        assert d <= (1e-12 if precision == "double" else 1e-4)


def test_abm3_restart_and_upload_follow_oracle():
    # dt sequence with a 5% change (history kept), a 30% change (full step on
    # the old history, then restart: two Heun start-up steps follow) and an
    # upload of the state (history dropped, as reset_abm_history)
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle
    case = random_case(40, 24, 1, seed=23, hole_radius_cells=5, ku1=2e4, axis=(0.0, 0.0, 1.0))
    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    t = 0.0
    dts = [1e-13] * 5 + [1.05e-13, 1.3e-13, 1.3e-13, 1.3e-13, 1.3e-13]
    counts = []
    with FdmBackend(case.plan(integrator="abm3")) as be:
        for dt in dts:
            st = be.step(dt)
            t, rep = orc.step("abm3", m, t, dt)
            counts.append(orc.abm_count())
            assert np.abs(be.field("m") - m).max() < 1e-11
            assert abs(st["total_energy_joules"] - rep["e_total"]) <= 1e-9 * abs(rep["e_total"])
        assert counts == [1, 2, 3, 3, 3, 3, 1, 2, 3, 3]
        be.upload_m(m)
        orc.reset_abm()
        for _ in range(4):
            be.step(1.3e-13)
            t, rep = orc.step("abm3", m, t, 1.3e-13)
            assert np.abs(be.field("m") - m).max() < 1e-11


@pytest.mark.parametrize("integrator,call", [("abm3", "time_kernels"), ("abm3", "autotune"), ("heun", "time_kernels"),
                                              ("dp45", "time_kernels"), ("dp45", "time_trial_kernels"),
                                              ("abm3", "time_trial_kernels"), ("rk4", "time_trial_kernels")])
def test_state_and_history_survive_tuning_calls(integrator, call):
    # the timing launches run trial stages and the controller on the live
    # buffers; the state, k1 and (ABM3) the RHS history must come out unchanged
    from fullmag_amd.backend import FdmBackend
    case = random_case(48, 40, 1, seed=3, hole_radius_cells=6)
    with FdmBackend(case.plan(integrator=integrator)) as a, FdmBackend(case.plan(integrator=integrator)) as b:
        for _ in range(5):
            a.step(1e-13)
            b.step(1e-13)
        getattr(b, call)()
        for _ in range(4):
            a.step(1e-13)
            b.step(1e-13)
        d = float(np.abs(a.field("m") - b.field("m")).max())
        # the tuner may pick another tile shape (other FFT rounding)
        assert d <= (1e-13 if call == "autotune" else 0.0), d


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("tile", ["128,0", "256,8", "128,4", "256,1", "64,2"])
def test_yz_tiles_give_identical_fields(precision, tile):
    # the 3-D y-line tile (threads per work-group) and z-pass tile (ky per
    # work-group) only regroup independent lines: the demag field and a step
    # are bit-identical to the default tiles (FULLMAG_FDM_YZ_TILE forces them)
    import os
    from fullmag_amd.backend import FdmBackend
    case = random_case(24, 32, 8, seed=3)
    plan = case.plan(precision=precision)
    out = []
    for t in (None, tile):
        old = os.environ.pop("FULLMAG_FDM_YZ_TILE", None)
        if t:
            os.environ["FULLMAG_FDM_YZ_TILE"] = t
        try:
            be = FdmBackend(plan)
        finally:
            os.environ.pop("FULLMAG_FDM_YZ_TILE", None)
            if old is not None:
                os.environ["FULLMAG_FDM_YZ_TILE"] = old
        with be:
            h = be.field("h_demag")
            be.step(1e-13)
            out.append((h, be.field("m")))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
