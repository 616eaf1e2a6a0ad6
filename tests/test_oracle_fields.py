"""Pins the oracle's FFT, demag convolution, exchange, local terms and LLG
right-hand side.  Each check is independent of the oracle's own formulas:
numpy.fft (library DFT), brute-force real-space convolution, the reference's
closed-form stencil tests (crates/fullmag-engine/src/lib.rs:257-306), the
discrete Laplacian eigenvalue of a sinusoid, and analytic LLG identities."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

MU0 = 4e-7 * np.pi
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 14, 22, 100, 254, 400, 1024])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = O.fft(x)
    assert np.abs(f - np.fft.fft(x)).max() <= 1e-13 * max(1.0, np.abs(x).sum())
    assert np.abs(O.fft(f, inverse=True) / n - x).max() < 1e-13


def _problem(nx, ny, nz, **kw):
    base = dict(dx=2e-9, dy=3e-9, dz=4e-9, ms=8e5, a_ex=1.3e-11, alpha=0.1)
    base.update(kw)
    return O.Oracle(O.Problem(nx, ny, nz, **base))


@pytest.mark.parametrize("shape,masked", [((8, 6, 1), False), ((5, 4, 3), True), ((7, 3, 2), False)])
def test_demag_fft_equals_direct_sum(shape, masked):
    nx, ny, nz = shape
    n = nx * ny * nz
    mask = None
    if masked:
        mask = (np.arange(n) % 5 != 2).astype(np.uint8)
    o = _problem(nx, ny, nz, mask=mask, exchange=False)
    m = o.prepare(O.random_unit_vectors(17, n))
    hf = o.fields(m)["h_demag"]
    hd = o.demag_direct(m)
    assert np.abs(hf - hd).max() < 1e-12 * np.abs(hd).max()
    if masked:
        assert np.all(hf[mask == 0] == 0.0)


def test_demag_energy_positive_and_film_anisotropy():
    # E_demag >= 0 (crates/fullmag-engine/src/lib.rs:438-467) and a thin film
    # magnetised out of plane costs more than in plane
    # (crates/fullmag-engine/tests/physics_guardrails.rs:101-119)
    o = _problem(16, 16, 1, dx=5e-9, dy=5e-9, dz=1e-9, exchange=False)
    n = 256
    for seed in (1, 2, 3):
        m = o.prepare(O.random_unit_vectors(seed, n))
        assert o.observables(m)["e_demag"] > 0.0
    e_out = o.observables(np.tile([0.0, 0.0, 1.0], (n, 1)))["e_demag"]
    e_in = o.observables(np.tile([1.0, 0.0, 0.0], (n, 1)))["e_demag"]
    assert e_out > 2.0 * e_in


def test_uniform_film_demag_equals_tensor_sum():
    # a uniformly magnetised body: H(r_i) = -Ms * sum_j N(r_i - r_j) . m
    nx, ny = 6, 4
    o = _problem(nx, ny, 1, exchange=False)
    k = O.newell_kernels(nx, ny, 1, 2e-9, 3e-9, 4e-9)
    h = o.fields(np.tile([0.0, 0.0, 1.0], (nx * ny, 1)))["h_demag"]
    i, j = 2, 1
    s = sum(k["zz"][0, (j - b) % (2 * ny), (i - a) % (2 * nx)] for a in range(nx) for b in range(ny))
    assert abs(h[j * nx + i, 2] + 8e5 * s) < 1e-9 * 8e5


def _exchange_case(key):
    g = GOLD["exchange"][key]
    nx, ny, nz = g["grid"]
    mask = np.array(g["mask"], dtype=np.uint8) if "mask" in g else None
    o = O.Oracle(O.Problem(nx, ny, nz, *g["cell"], ms=g["ms"], a_ex=g["a_over_mu0"] * MU0, alpha=0.1,
                           demag=False, mask=mask))
    return g, o


def test_exchange_reference_stencils():
    g, o = _exchange_case("uniform_zero")
    m = o.prepare(np.array(g["m"], dtype=float))
    f = o.fields(m)
    assert np.abs(f["h_ex"]).max() <= g["tol"]
    assert np.abs(o.rhs(m)).max() <= g["tol"]
    assert abs(o.observables(m)["e_ex"]) <= g["tol"]
    g, o = _exchange_case("three_cell_stencil")
    h = o.fields(o.prepare(np.array(g["m"], dtype=float)))["h_ex"]
    assert np.abs(h[g["cell_index"]] - g["h"]).max() <= g["tol"]
    g, o = _exchange_case("masked_free_surface")
    m = o.prepare(np.array(g["m"], dtype=float))
    h = o.fields(m)["h_ex"]
    assert np.abs(h[1] - g["h1"]).max() <= g["tol"]
    assert np.abs(h[2] - g["h2"]).max() <= g["tol"]
    assert np.all(m[2] == 0.0)


def test_exchange_discrete_laplacian_eigenvalue():
    # m = (cos(q x), sin(q x), 0) on a periodic-free interior: the stencil gives
    # H = (2A/mu0 Ms) * 2 (cos(q dx) - 1)/dx^2 * m at interior cells
    nx, dx, q = 32, 2e-9, 0.15 / 2e-9
    x = np.arange(nx) * dx
    m = np.stack([np.cos(q * x), np.sin(q * x), np.zeros(nx)], axis=1)
    o = O.Oracle(O.Problem(nx, 1, 1, dx, 2e-9, 2e-9, 8e5, 1.3e-11, 0.1, demag=False))
    h = o.fields(m)["h_ex"]
    lam = 2 * 1.3e-11 / (MU0 * 8e5) * 2 * (np.cos(q * dx) - 1) / dx ** 2
    assert np.abs(h[1:-1] - lam * m[1:-1]).max() < 1e-12 * abs(lam)
    # Neumann edge: the missing neighbour is the cell itself
    assert np.abs(h[0] - 2 * 1.3e-11 / (MU0 * 8e5) * (m[1] - m[0]) / dx ** 2).max() < 1e-9 * abs(lam)


def test_exchange_energy_pair_form_equals_field_form():
    # for |m| = 1: A V sum_pairs |dm|^2/d^2 == -1/2 mu0 Ms V sum m.H_ex
    nx, ny, nz = 5, 4, 3
    d = (2e-9, 2.5e-9, 3e-9)
    o = O.Oracle(O.Problem(nx, ny, nz, *d, ms=8e5, a_ex=1.3e-11, alpha=0.1, demag=False))
    m = o.prepare(O.random_unit_vectors(4, nx * ny * nz)).reshape(nz, ny, nx, 3)
    pair = 0.0
    for ax, dd in zip((2, 1, 0), d):
        diff = np.diff(m, axis=ax)
        pair += (diff ** 2).sum() / dd ** 2
    pair *= 1.3e-11 * np.prod(d)
    assert abs(o.observables(m.reshape(-1, 3))["e_ex"] - pair) < 1e-12 * pair


def test_local_terms_and_llg_identities():
    n = 6
    o = O.Oracle(O.Problem(n, 1, 1, 2e-9, 2e-9, 2e-9, 8e5, 1.3e-11, 0.3, demag=False, exchange=False,
                           h_ext=(1e4, -2e4, 3e4), ku1=5e4, ku2=1e4, axis=(0.0, 0.6, 0.8)))
    m = o.prepare(O.random_unit_vectors(9, n))
    f = o.fields(m)
    u = np.array([0.0, 0.6, 0.8])
    mu = m @ u
    expect = (2 * 5e4 / (MU0 * 8e5) * mu + 4 * 1e4 / (MU0 * 8e5) * mu ** 3)[:, None] * u
    assert np.abs(f["h_ani"] - expect).max() < 1e-10 * np.abs(expect).max()
    assert np.abs(f["h_ext"] - [1e4, -2e4, 3e4]).max() == 0.0
    k = o.rhs(m)
    h = f["h_eff"]
    gb = 2.211e5 / (1 + 0.3 ** 2)
    ref = -gb * (np.cross(m, h) + 0.3 * np.cross(m, np.cross(m, h)))
    assert np.abs(k - ref).max() < 1e-10 * np.abs(ref).max()
    assert np.abs((k * m).sum(1)).max() < 1e-9 * np.abs(k).max()   # tangency m.k = 0


def test_pure_damping_has_no_precession():
    o = O.Oracle(O.Problem(1, 1, 1, 1e-9, 1e-9, 1e-9, 8e5, 1.3e-11, 0.5, demag=False, exchange=False,
                           h_ext=(0.0, 0.0, 1e5), precession=False))
    m = o.prepare(np.array([[1.0, 0.0, 1.0]]))
    k = o.rhs(m)[0]
    assert abs(k[1]) < 1e-20            # no azimuthal component
    assert k[2] > 0 and k[0] < 0        # relaxes towards +z


def test_extension_maps_reduce_to_scalar_problem():
    # per-cell Ms/Ku/axis maps filled with the uniform values and a waveform
    # that is identically 1 reproduce the scalar problem exactly
    n = 5 * 4 * 3
    kw = dict(dx=2e-9, dy=3e-9, dz=4e-9, ms=8e5, a_ex=1.3e-11, alpha=0.1, ku1=3e5, ku2=2e4, axis=(0.0, 0.6, 0.8))
    base = O.Oracle(O.Problem(5, 4, 3, h_ext=(1e3, 2e3, -5e3), **kw))
    mapped = O.Oracle(O.Problem(5, 4, 3, ms_cell=np.full(n, 8e5), ku1_cell=np.full(n, 3e5),
                                ku2_cell=np.full(n, 2e4), axis_cell=np.tile([0.0, 0.6, 0.8], (n, 1)),
                                wave_kind=1, wave_amp=(1e3, 2e3, -5e3), wave_t0=-2.0, wave_t1=-1.0, **kw))
    m = base.prepare(O.random_unit_vectors(8, n))
    fa, fb = base.fields(m), mapped.fields(m)
    for k in ("h_ex", "h_demag", "h_ext", "h_ani", "h_eff"):
        assert np.abs(fa[k] - fb[k]).max() <= 1e-12 * max(1.0, np.abs(fa[k]).max())
    assert abs(base.observables(m)["e_total"] - mapped.observables(m)["e_total"]) < 1e-12 * abs(
        base.observables(m)["e_total"])


def test_extension_waveforms_closed_forms():
    # single cell, no internal fields: H_ext = f(t) * amp with f from the
    # waveform definitions (ramp clamp, sinc(0) = 1, sine phase)
    amp = (0.0, 0.0, 1e4)
    cases = [(1, dict(wave_t0=1e-12, wave_t1=3e-12), [(0.0, 0.0), (2e-12, 0.5), (5e-12, 1.0)]),
             (2, dict(wave_t0=1e-12, wave_freq=2e10), [(1e-12, 1.0), (1e-12 + 0.25 / 2e10, np.sin(np.pi / 2) / (np.pi / 2))]),
             (3, dict(wave_freq=1e9, wave_phase=0.5), [(0.0, np.sin(0.5)), (2.5e-10, np.sin(np.pi / 2 + 0.5))])]
    for kind, kw, pts in cases:
        o = O.Oracle(O.Problem(1, 1, 1, 1e-9, 1e-9, 1e-9, 8e5, 1.3e-11, 0.1, exchange=False, demag=False,
                               wave_kind=kind, wave_amp=amp, **kw))
        for t, f in pts:
            o.set_time(t)
            assert abs(o.fields(np.array([[1.0, 0.0, 0.0]]))["h_ext"][0, 2] - f * 1e4) < 1e-9
