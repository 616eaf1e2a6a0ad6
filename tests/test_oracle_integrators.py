"""Pins the oracle's integrators and runner loop with closed forms and
invariants rather than by restating the tableaux:

- single-cell macrospin in a static field has the analytic LLG solution
  phi(t) = gb H t, tan(theta/2) = tan(theta0/2) exp(-alpha gb H t);
- observed convergence orders (RK4: 4, Heun: 2, DP45 5th-order solution: 5,
  RK23 3rd-order solution: 3, ABM3: 3) fail for any wrong tableau coefficient;
- ABM3 (fdm_integrators.rs:683-820) against a scheme built here from the
  definition of Adams methods: weights are integrals of the Lagrange
  polynomials through the history nodes (numpy.polynomial), three Heun
  start-up steps, predictor-evaluate-correct pushing f(pred); plus the
  history's restart rule (fdm_state.rs:247-257: a dt change above 10%);
- the DP45 error estimate scales like dt^5 (embedded 4th-order pair), the
  RK23 one like dt^3 (embedded 2nd-order pair), and each controller uses
  its own exponent 1/(order+1);
- norm preservation, monotone energy decay with damping, and the reference's
  guardrail that RK45 and Heun relax to energies within 5%
  (crates/fullmag-engine/tests/physics_guardrails.rs:122-152);
- the runner loop's dt carry and until-clamp
  (crates/fullmag-runner/src/cpu_reference.rs:622-639).
"""
import numpy as np
import pytest

from oracle import oracle as O

H = 2e5
ALPHA = 0.1
GB = 2.211e5 / (1 + ALPHA ** 2)


def _macrospin(**kw):
    return O.Oracle(O.Problem(1, 1, 1, 1e-9, 1e-9, 1e-9, 8e5, 1.3e-11, ALPHA, demag=False, exchange=False,
                              h_ext=(0.0, 0.0, H), **kw))


def _exact(theta0, t):
    th = 2 * np.arctan(np.tan(theta0 / 2) * np.exp(-ALPHA * GB * H * t))
    ph = GB * H * t
    return np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])


def _integrate(integ, dt, t_end):
    o = _macrospin(atol=1e9, dt_max=1.0)   # huge atol: DP45 accepts the given dt
    th0 = 1.0
    m = np.array([[np.sin(th0), 0.0, np.cos(th0)]])
    t = 0.0
    while t < t_end - 1e-25:
        t, rep = o.step(integ, m, t, dt)
    return m[0], _exact(th0, t)


@pytest.mark.parametrize("integ,order,n0", [("rk4", 4, 80), ("heun", 2, 80), ("rk45", 5, 80), ("rk23", 3, 80),
                                            ("abm3", 3, 160)])
def test_macrospin_convergence_order(integ, order, n0):
    # (the multistep method leaves its pre-asymptotic range at ~0.05 rad/step)
    t_end = 2e-10
    errs = []
    for nsteps in (n0, 2 * n0, 4 * n0):
        m, ex = _integrate(integ, t_end / nsteps, t_end)
        errs.append(np.abs(m - ex).max())
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(abs(r - order) < 0.4 for r in rates), (errs, rates)


def test_dp45_error_estimate_scales_as_dt5():
    o = _macrospin(atol=1e9, dt_max=1.0)
    ests = []
    for dt in (4e-12, 2e-12):
        o.invalidate_fsal()
        m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
        _, rep = o.step("rk45", m, 0.0, dt)
        # dt_next = headroom * dt * (thr/err)^(1/5) -> err = thr*(0.8 dt/dt_next)^5
        ests.append(1e9 * (0.8 * dt / rep["dt_next"]) ** 5)
    assert abs(np.log2(ests[0] / ests[1]) - 5.0) < 0.3


def test_rk23_error_estimate_scales_as_dt3():
    o = _macrospin(atol=1e9, dt_max=1.0)
    ests = []
    for dt in (4e-12, 2e-12):
        m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
        _, rep = o.step("rk23", m, 0.0, dt)
        ests.append(1e9 * (0.8 * dt / rep["dt_next"]) ** 3)
    assert abs(np.log2(ests[0] / ests[1]) - 3.0) < 0.2


@pytest.mark.parametrize("integ,expo", [("rk45", 0.2), ("rk23", 1.0 / 3.0)])
def test_controller_exponent(integ, expo):
    # same attempt at two tolerances -> same err; dt_next ratio = (atol ratio)^expo
    res = []
    for atol in (1e-3, 3.2e-2):
        o = _macrospin(atol=atol, dt_max=1.0)
        m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
        _, rep = o.step(integ, m, 0.0, 1e-13)
        assert rep["rejected"] == 0
        res.append(rep["dt_next"])
    assert abs(res[1] / res[0] - (3.2e-2 / 1e-3) ** expo) < 1e-12


def test_rk23_rejection_shrinks_and_retries():
    o = _macrospin(atol=1e-7, dt_max=1.0)
    m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
    _, rep = o.step("rk23", m, 0.0, 5e-11)
    assert rep["rejected"] >= 1 and rep["dt_used"] < 5e-11
    assert np.abs(m[0] - _exact(1.0, rep["dt_used"])).max() < 1e-6


def test_adaptive_controller_formula():
    # accept: dt_next = clamp(headroom*dt*(atol/err)^0.2) with err recovered
    # from a one-step run at two tolerances -> same err, ratio of dt_next is
    # (atol1/atol2)^0.2
    res = []
    for atol in (1e-3, 3.2e-2):
        o = _macrospin(atol=atol, dt_max=1.0)
        m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
        _, rep = o.step("rk45", m, 0.0, 1e-13)
        assert rep["rejected"] == 0
        res.append(rep["dt_next"])
    assert abs(res[1] / res[0] - (3.2e-2 / 1e-3) ** 0.2) < 1e-12


def test_rejection_shrinks_and_retries():
    o = _macrospin(atol=1e-9, dt_max=1.0)
    m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
    _, rep = o.step("rk45", m, 0.0, 5e-11)
    assert rep["rejected"] >= 1 and rep["dt_used"] < 5e-11
    assert np.abs(m[0] - _exact(1.0, rep["dt_used"])).max() < 1e-8


def test_dt_clamped_to_bounds():
    o = _macrospin(atol=1e9, dt_max=1e-13, dt_min=1e-15)
    m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
    _, rep = o.step("rk45", m, 0.0, 1e-11)
    assert rep["dt_used"] == 1e-13 and rep["dt_next"] == 1e-13


def test_runner_loop_until_and_dt_carry():
    o = _macrospin(atol=1e-6)
    m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
    r = o.run("rk45", m, 0.0, 1e-11, 1e-14, 100000, trace=True)
    tr = r["trace"]
    assert r["t"] >= 1e-11 and abs(r["t"] - 1e-11) < 1e-24
    # the last step is clamped to the remaining time
    assert tr[-1, 1] <= tr[-2, 1] * 1.0000001 + 1e-30
    assert np.abs(m[0] - _exact(1.0, r["t"])).max() < 1e-5


def _film(alpha=0.5):
    from tests.cases import random_case
    return random_case(12, 10, 1, seed=6, alpha=alpha, dx=4e-9, dy=4e-9, dz=4e-9, h_ext=None)


def test_norm_preserved_and_energy_decays():
    c = _film()
    o = O.Oracle(c.problem())
    m = o.prepare(c.m0)
    t, dt, e_prev = 0.0, 1e-13, None
    for _ in range(30):
        t, rep = o.step("rk45", m, t, dt)
        dt = rep["dt_next"]
        assert np.abs(np.linalg.norm(m, axis=1) - 1).max() < 1e-12
        if e_prev is not None:
            assert rep["e_total"] <= e_prev + 1e-12 * abs(e_prev)
        e_prev = rep["e_total"]


def test_rk45_and_heun_relax_to_same_energy():
    c = _film(alpha=1.0)
    o1, o2 = O.Oracle(c.problem()), O.Oracle(c.problem())
    m1, m2 = o1.prepare(c.m0), o2.prepare(c.m0)
    r1 = o1.run("rk45", m1, 0.0, 2e-10, 1e-13, 200000)
    r2 = o2.run("heun", m2, 0.0, 2e-10, 2e-14, 200000)
    e1, e2 = r1["last"]["e_total"], r2["last"]["e_total"]
    assert abs(e1 - e2) <= 0.05 * abs(e1)


def _adams_weights(nodes):
    """Weights w_j of int_0^1 p(s) ds where p interpolates f at the given
    nodes (in units of dt, 0 = start of the step)."""
    from numpy.polynomial import polynomial as P
    w = []
    for j, xj in enumerate(nodes):
        poly = np.array([1.0])
        for i, xi in enumerate(nodes):
            if i != j:
                poly = P.polymul(poly, np.array([-xi, 1.0]) / (xj - xi))
        integ = P.polyint(poly)
        w.append(P.polyval(1.0, integ) - P.polyval(0.0, integ))
    return np.array(w)


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_abm3_matches_adams_scheme_built_from_definition():
    ab = _adams_weights([0.0, -1.0, -2.0])     # f_n, f_{n-1}, f_{n-2}
    am = _adams_weights([1.0, 0.0, -1.0])      # f*, f_n, f_{n-1}
    o = _macrospin()
    o2 = _macrospin()
    m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
    x = m.copy()
    hist = []
    dt, t = 2e-12, 0.0
    for step in range(10):
        t, _ = o.step("abm3", m, t, dt)
        if len(hist) < 3:
            k1 = o2.rhs(x)
            pred = _unit(x + dt * k1)
            x = _unit(x + 0.5 * dt * (k1 + o2.rhs(pred)))
            hist.insert(0, o2.rhs(x))
        else:
            pred = _unit(x + dt * sum(w * f for w, f in zip(ab, hist)))
            fs = o2.rhs(pred)
            x = _unit(x + dt * (am[0] * fs + am[1] * hist[0] + am[2] * hist[1]))
            hist.insert(0, fs)
        hist = hist[:3]
        assert o.abm_count() == min(step + 1, 3)
        assert np.abs(m - x).max() < 1e-13, step


def test_abm3_restarts_on_dt_change_above_ten_percent():
    o = _macrospin()
    m = np.array([[np.sin(1.0), 0.0, np.cos(1.0)]])
    t = 0.0
    for _ in range(5):
        t, _ = o.step("abm3", m, t, 2e-12)
    t, _ = o.step("abm3", m, t, 2.1e-12)        # +5%: no restart
    assert o.abm_count() == 3
    t, _ = o.step("abm3", m, t, 2.5e-12)        # +19%: full step, then restart at the push
    assert o.abm_count() == 1
    # the next step is a Heun start-up step
    ref = _macrospin()
    x = m.copy()
    ref.step("heun", x, t, 2.5e-12)
    o.step("abm3", m, t, 2.5e-12)
    assert o.abm_count() == 2
    assert np.array_equal(m, x)
