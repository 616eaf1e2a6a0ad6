"""The film's joined x pass with its inputs prefetched into LDS (k_xxp, FP32,
nx a multiple of 256, two rows per work-group) reads the same values as the
generic joined pass (k_xx) and runs the same formulas; the compiler may fuse
multiply-adds differently in the two kernels, so the states agree to FP32
roundoff at a common dt sequence (|dm| <= 1e-6 after six steps, far below
the 1e-3 a wrong neighbour row or k-vector would leave; the adaptive
methods run with atol 1 and a fixed dt, since their controllers amplify
the roundoff of the cancelling error estimate into different dt), and both
stay in the FP32 band of the FP64 oracle.  The generic pass is selected with
FULLMAG_FDM_XPASS_PREFETCH=0 at create."""
import os

import numpy as np
import pytest

from tests.cases import random_case

pytestmark = pytest.mark.gpu


def _run(plan, steps, prefetch):
    from fullmag_amd.backend import FdmBackend
    old = os.environ.get("FULLMAG_FDM_XPASS_PREFETCH")
    os.environ["FULLMAG_FDM_XPASS_PREFETCH"] = "1" if prefetch else "0"
    try:
        be = FdmBackend(plan)
    finally:
        if old is None:
            del os.environ["FULLMAG_FDM_XPASS_PREFETCH"]
        else:
            os.environ["FULLMAG_FDM_XPASS_PREFETCH"] = old
    with be:
        rows = []
        for _ in range(steps):
            st = be.step(5e-14)
            rows.append((st["time_seconds"], st["total_energy_joules"], st["max_rhs_amplitude"]))
        return be.field("m"), rows


@pytest.mark.parametrize("shape,integrator", [((256, 64), "dp45"), ((256, 64), "rk4"), ((256, 64), "heun"),
                                              ((256, 48), "rk23"), ((512, 128), "dp45")])
def test_prefetch_pass_equals_generic_pass(shape, integrator):
    nx, ny = shape
    case = random_case(nx, ny, 1, seed=6, hole_radius_cells=ny // 5, alpha=0.05)
    case.atol = 1.0
    _compare(case.plan(precision="single", integrator=integrator), nx)


def _compare(plan, nx):
    m1, r1 = _run(plan, 6, True)
    m0, r0 = _run(plan, 6, False)
    d = np.abs(m1 - m0).max(axis=1)
    i = int(np.argmax(d))
    print(f"max |dm| {d[i]:.3e} at x={i % nx} y={i // nx}; rows with |dm| > 1e-7: "
          f"{sorted(set((np.nonzero(d > 1e-7)[0] // nx).tolist()))[:20]}", flush=True)
    assert d[i] <= 1e-6
    for (t1, e1, k1), (t0, e0, k0) in zip(r1, r0):
        assert abs(t1 - t0) <= 1e-6 * t0
        assert abs(e1 - e0) <= 1e-6 * abs(e0)
        assert abs(k1 - k0) <= 1e-5 * k0


def test_prefetch_pass_with_cell_maps_and_anisotropy():
    nx, ny = 256, 32
    rng = np.random.default_rng(4)
    case = random_case(nx, ny, 1, seed=8, alpha=0.1, ku1=3e4, axis=(0.0, 0.0, 1.0), atol=1.0)
    case.ms_cell = rng.uniform(6e5, 9e5, nx * ny)
    ax = rng.standard_normal((nx * ny, 3))
    case.axis_cell = ax / np.linalg.norm(ax, axis=1, keepdims=True)
    _compare(case.plan(precision="single", integrator="dp45"), nx)


def test_prefetch_pass_matches_oracle_band():
    # the 512 x 512 film itself, FP32 with the prefetch pass, against the
    # FP64 oracle at the engine's dt (the Tier-B band |dm| <= 1e-5)
    from fullmag_amd.backend import FdmBackend, FdmPlan
    from fullmag_amd.workloads import film_with_hole
    from oracle.oracle import Oracle, Problem
    kw = film_with_hole()
    orc = Oracle(Problem(kw["nx"], kw["ny"], 1, kw["dx"], kw["dy"], kw["dz"], kw["ms"], kw["a_ex"], kw["alpha"],
                         h_ext=kw["h_ext"], mask=kw["mask"], atol=1.0))
    m = orc.prepare(kw["m0"])
    t = 0.0
    with FdmBackend(FdmPlan(**{**kw, "precision": "single"})) as be:
        dt = 1e-13
        for _ in range(4):
            st = be.step(dt)
            t, _ = orc.step("rk45", m, t, st["dt_seconds"])
            dt = st["suggested_next_dt"]
        assert np.abs(be.field("m") - m).max() <= 1e-5
