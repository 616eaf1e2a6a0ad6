"""Adjoint sensitivities of fixed-step runs (fullmag_fdm_adjoint_*) against
central finite differences of the FP64 oracle's objective.

The objective and its gradient have no counterpart in the reference; they
are pinned by (a) the oracle's forward objective (same RK4 with per-cell Ms
and a sinc pulse, FP64 CPU) and (b) central differences of that objective
with h = 1e-4 relative on 16 seeded parameters (15 design-map cells and the
pulse amplitude).  FP64 adjoint vs FD: 1e-6 of the gradient scale (the FD
truncation error is O(h^2)); FP32 adjoint vs FP64 adjoint: 1e-4.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _problem(precision="double", integrator="rk4", demag=True, nz=2, steps=160, interval=0, budget=0):
    from fullmag_amd.adjoint import WaveguideProblem
    rng = np.random.default_rng(5)
    nx, ny = 24, 8
    n = nx * ny * nz
    m0 = np.array([1.0, 0.2, 0.1]) + 0.05 * rng.standard_normal((n, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    return WaveguideProblem(nx=nx, ny=ny, nz=nz, design_x=(8, 16), design_y=(0, 8), pulse_cols=3, probe_cols=3,
                            t0=4e-12, dt=50e-15, steps=steps, window=4e-12, precision=precision,
                            integrator=integrator, checkpoint_interval=interval, record_budget_bytes=budget, m0=m0,
                            extra={} if demag else {"demag": False})


def _design0(pr, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(6e5, 1e6, size=pr.design_shape())


def _oracle_objective(pr, ms_design, amp_t):
    from oracle.oracle import Oracle, Problem
    x = np.tile(np.arange(pr.nx), pr.ny * pr.nz)
    wmask = (x < pr.pulse_cols).astype(np.uint8)
    h1 = amp_t / MU0 * np.asarray(pr.direction)
    orc = Oracle(Problem(pr.nx, pr.ny, pr.nz, pr.cell, pr.cell, pr.cell, pr.ms_base, pr.a_ex, pr.alpha,
                         demag=pr.extra.get("demag", True), ms_cell=pr.ms_map(ms_design), wave_kind=2,
                         wave_amp=tuple(h1), wave_t0=pr.t0, wave_freq=pr.f_cut, wave_mask=wmask))
    m = orc.prepare(pr.initial_m())
    w = pr.weights()
    first = pr.first_step()
    t, acc = 0.0, 0.0
    for s in range(1, pr.steps + 1):
        t, _ = orc.step(pr.integrator, m, t, pr.dt)
        if s >= first:
            acc += float((w * m[:, pr.component] ** 2).sum())
    return acc / (pr.steps - first + 1)


def _params(pr, seed=11, count=15):
    rng = np.random.default_rng(seed)
    dy, dx = pr.design_shape()
    flat = rng.choice(dy * dx, size=count, replace=False)
    return [(int(f // dx), int(f % dx)) for f in flat]


def _fd_oracle(pr, ms_design, amp, params, h=1e-4):
    out = []
    for (iy, ix) in params:
        v = ms_design[iy, ix]
        p, m = ms_design.copy(), ms_design.copy()
        p[iy, ix] = v * (1 + h)
        m[iy, ix] = v * (1 - h)
        out.append((_oracle_objective(pr, p, amp) - _oracle_objective(pr, m, amp)) / (2 * h * v))
    out.append((_oracle_objective(pr, ms_design, amp * (1 + h)) - _oracle_objective(pr, ms_design, amp * (1 - h)))
               / (2 * h * amp))
    return np.array(out)


def _adjoint(pr, ms_design, amp, params):
    j, gd, ga = pr.value_and_grad(ms_design, amp)
    return j, np.array([gd[iy, ix] for (iy, ix) in params] + [ga]), gd


def test_forward_objective_matches_oracle():
    pr = _problem()
    d = _design0(pr)
    j_engine, be = pr.run(d, 5e-3)
    be.close()
    j_orc = _oracle_objective(pr, d, 5e-3)
    assert abs(j_engine - j_orc) <= 1e-10 * abs(j_orc)


@pytest.mark.parametrize("integrator,demag", [("rk4", True), ("heun", True), ("rk4", False)])
def test_fp64_adjoint_matches_oracle_central_differences(integrator, demag):
    pr = _problem(integrator=integrator, demag=demag)
    d = _design0(pr)
    params = _params(pr)
    j, g_adj, _ = _adjoint(pr, d, 5e-3, params)
    h = 1e-4
    g_fd = _fd_oracle(pr, d, 5e-3, params, h)
    scale = np.abs(g_fd).max()
    # central differences cannot resolve less than the roundoff of J over 2 h p
    p = np.array([d[iy, ix] for (iy, ix) in params] + [5e-3])
    floor = 10 * 2.2e-16 * abs(j) / (2 * h * p)
    err = np.abs(g_adj - g_fd) / (1e-6 * scale + floor)
    print(f"{integrator} demag={demag}: grad scale {scale:.3e}, worst |adjoint - fd| / tolerance {err.max():.3f}",
          flush=True)
    assert scale > 0
    assert err.max() <= 1.0


# record budgets: 1 byte keeps only the last segment (every other segment is
# recomputed), 3 segments of 7 steps (86 KB per FP64 step of this problem),
# and the default (the whole run)
BUDGETS = (1, 3 * 7 * 4 * 7 * 384 * 8, 0)


def test_checkpoint_interval_does_not_change_the_gradient():
    d = None
    ref = None
    for interval, budget in [(i, b) for i in (0, 1, 7, 160) for b in BUDGETS]:
        pr = _problem(interval=interval, budget=budget)
        if d is None:
            d = _design0(pr)
        j, gd, ga = pr.value_and_grad(d, 5e-3)
        if ref is None:
            ref = (j, gd, ga)
            continue
        assert j == ref[0]
        assert np.abs(gd - ref[1]).max() <= 1e-13 * np.abs(ref[1]).max()
        assert abs(ga - ref[2]) <= 1e-13 * abs(ref[2])


def test_fp32_adjoint_matches_fp64_adjoint():
    p64, p32 = _problem("double"), _problem("single")
    d = _design0(p64)
    j64, gd64, ga64 = p64.value_and_grad(d, 5e-3)
    j32, gd32, ga32 = p32.value_and_grad(d, 5e-3)
    assert abs(j32 - j64) <= 1e-5 * abs(j64)
    e_ms = np.abs(gd32 - gd64).max() / np.abs(gd64).max()
    e_amp = abs(ga32 - ga64) / abs(ga64)
    print(f"fp32 vs fp64 adjoint: ms {e_ms:.3e}, amplitude {e_amp:.3e}", flush=True)
    assert e_ms <= 1e-4 and e_amp <= 1e-4


def test_torch_autograd_function():
    import torch
    from fullmag_amd.adjoint import LLGObjective
    pr = _problem()
    d = _design0(pr)
    ms = torch.tensor(d, dtype=torch.float64, requires_grad=True)
    amp = torch.tensor(5e-3, dtype=torch.float64, requires_grad=True)
    j = LLGObjective.apply(ms, amp, pr)
    (2.0 * j).backward()
    j2, gd, ga = pr.value_and_grad(d, 5e-3)
    assert float(j.detach()) == j2
    assert np.allclose(ms.grad.numpy(), 2.0 * gd, rtol=1e-12, atol=0)
    assert abs(float(amp.grad) - 2.0 * ga) <= 1e-12 * abs(2.0 * ga)


def test_adjoint_rejects_adaptive_plans_and_backward_first():
    from fullmag_amd.backend import FdmBackend, FdmError
    pr = _problem()
    plan = pr.plan(_design0(pr), 5e-3)
    plan.integrator = "dp45"
    with FdmBackend(plan) as be:
        with pytest.raises(FdmError):
            be.adjoint_forward(10, 1e-13, pr.weights())
    with FdmBackend(pr.plan(_design0(pr), 5e-3)) as be:
        with pytest.raises(FdmError):
            be.adjoint_backward()
        with pytest.raises(FdmError):
            be.adjoint_forward(10, 1e-13, pr.weights(), first_step=11)


def test_fp64_field_cotangent_for_fp32_plans():
    # FP32 plans convolve the field cotangent in FP64 by default; the FP32
    # convolution (beta_state_precision) stays within the FP32 band of it
    p64 = _problem("double")
    d = _design0(p64)
    _, gd64, ga64 = p64.value_and_grad(d, 5e-3)
    out = {}
    for state in (False, True):
        pr = _problem("single")
        pr.beta_state_precision = state
        out[state] = pr.value_and_grad(d, 5e-3)
    for state, (_, gd, ga) in out.items():
        e_ms = np.abs(gd - gd64).max() / np.abs(gd64).max()
        e_amp = abs(ga - ga64) / abs(ga64)
        print(f"fp32 adjoint, beta in {'fp32' if state else 'fp64'}: ms {e_ms:.3e}, amplitude {e_amp:.3e}", flush=True)
        assert e_ms <= 1e-4 and e_amp <= 1e-4
    # the two forms differ (the FP64 context is used) but only in the FP32 band
    _, gd_a, ga_a = out[False]
    _, gd_b, ga_b = out[True]
    assert ga_a != ga_b
    assert abs(ga_a - ga_b) <= 1e-4 * abs(ga64)


@pytest.mark.parametrize("precision,interval,budget", [("double", 0, 1), ("double", 7, BUDGETS[1]), ("double", 160, 1),
                                                      ("double", 0, 0), ("single", 0, 1), ("single", 40, 1)])
def test_repeated_backward_and_recorded_last_segment(precision, interval, budget):
    # the forward records the trailing segments; a second backward after the
    # first (whose recomputes reuse those record slots) gives the same result
    pr = _problem(precision, interval=interval, budget=budget)
    d = _design0(pr)
    j, be = pr.run(d, 5e-3)
    try:
        g1, a1 = be.adjoint_backward()
        g2, a2 = be.adjoint_backward()
    finally:
        be.close()
    assert np.array_equal(g1, g2) and np.array_equal(a1, a2)
    j_ref, gd_ref, ga_ref = _problem(precision, interval=1, budget=1).value_and_grad(d, 5e-3)
    assert j == j_ref
    assert np.abs(pr.design_grad(g1) - gd_ref).max() <= 1e-12 * np.abs(gd_ref).max()
