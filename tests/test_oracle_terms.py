"""Pins the oracle's further field and torque terms (cubic anisotropy, DMI,
magnetoelastic coupling, Oersted cylinder, Zhang-Li / Slonczewski / spin-orbit
torques) with checks that do not restate the oracle's formulas:

- every field term is the negative variational derivative of its stated
  energy, H = -(1/(mu0 Ms V)) dE/dm, checked by central differences of an
  energy written here from the definitions (cubic Kc1 S + Kc2 P + Kc3 S^2 in
  a rotated frame, fdm_fields.rs:648-705 and demag_fp64.cu:301-330;
  magnetoelastic magnetoelastic.rs:155-176; interfacial and bulk DMI energy
  densities D (m_z div m - (m.grad) m_z) and D m.curl m with central
  differences, fdm_fields.rs:707-816, on cells away from the clamped box
  faces);
- the Oersted field against Ampere's law outside and inside the conductor
  for an arbitrary axis (fdm_fields.rs:1049-1150) and the time envelopes;
- the torques are perpendicular to m, scale as stated with current, Ms and
  thickness, and their prefactors match hbar J / (2 e mu0 Ms d) and
  P mu_B J / (e Ms (1 + beta^2)) built from scipy.constants (CODATA 2018;
  the reference hard-codes CODATA 2014 values, hence the 1e-7 tolerance).
"""
import numpy as np
import pytest
from scipy import constants as C

from oracle import oracle as O

MU0 = 4e-7 * np.pi


def _problem(nx, ny, nz, **kw):
    base = dict(dx=2e-9, dy=3e-9, dz=4e-9, ms=8e5, a_ex=1.3e-11, alpha=0.1, exchange=False, demag=False)
    base.update(kw)
    return O.Oracle(O.Problem(nx, ny, nz, **base))


def _random_m(n, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def _grad_check(orc, energy, m, cells, ms, vol, h=1e-6, rtol=1e-6):
    """Field of the oracle against -(1/(mu0 Ms V)) dE/dm at the given cells."""
    f = orc.fields(m)["h_eff"]
    scale = np.abs(f).max()
    for s in cells:
        for c in range(3):
            mp, mm = m.copy(), m.copy()
            mp[s, c] += h
            mm[s, c] -= h
            g = (energy(mp) - energy(mm)) / (2 * h)
            expect = -g / (MU0 * ms[s] * vol)
            assert abs(f[s, c] - expect) <= rtol * scale, (s, c, f[s, c], expect)


def _rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    return q


def test_cubic_field_is_gradient_of_cubic_energy():
    R = _rotation(3)
    a1, a2 = R[0] * 2.0, R[1] * 0.5                 # unnormalised axes: the engine normalises them
    nx, ny, nz = 4, 3, 2
    n = nx * ny * nz
    rng = np.random.default_rng(4)
    kc1 = rng.uniform(-5e4, 5e4, n)
    ms = rng.uniform(5e5, 1e6, n)
    orc = _problem(nx, ny, nz, cubic=(0.0, 3e4, -2e4, tuple(a1), tuple(a2)), kc1_cell=kc1, ms_cell=ms)
    c1, c2 = R[0], R[1]
    c3 = np.cross(c1, c2)
    vol = 2e-9 * 3e-9 * 4e-9

    def energy(m):
        m1, m2, m3 = m @ c1, m @ c2, m @ c3
        S = m1 ** 2 * m2 ** 2 + m2 ** 2 * m3 ** 2 + m1 ** 2 * m3 ** 2
        P = m1 ** 2 * m2 ** 2 * m3 ** 2
        return vol * np.sum(kc1 * S + 3e4 * P - 2e4 * S ** 2)

    m = _random_m(n, 5)
    _grad_check(orc, energy, m, range(n), ms, vol)


def test_magnetoelastic_field_and_energy():
    strain = (1e-3, -2e-4, 5e-4, 3e-4, -6e-4, 2e-4)   # Voigt, engineering shears
    b1, b2 = -2.0e7, 3.5e7
    n = 6
    orc = _problem(3, 2, 1, mel=(b1, b2, strain))
    vol = 2e-9 * 3e-9 * 4e-9
    e11, e22, e33 = strain[:3]
    eps = np.array([[e11, strain[5] / 2, strain[4] / 2], [strain[5] / 2, e22, strain[3] / 2],
                    [strain[4] / 2, strain[3] / 2, e33]])

    def energy(m):
        # B1 sum_i m_i^2 eps_ii + B2 sum_{i != j} m_i m_j eps_ij (tensor strain)
        diag = np.einsum("ni,i->n", m ** 2, np.diag(eps))
        off = np.einsum("ni,ij,nj->n", m, eps - np.diag(np.diag(eps)), m)
        return vol * np.sum(b1 * diag + b2 * off)

    m = _random_m(n, 6)
    _grad_check(orc, energy, m, range(n), np.full(n, 8e5), vol)
    r = orc.observables(m)
    assert abs(r["e_mel"] - energy(m)) <= 1e-12 * abs(energy(m))
    assert abs(r["e_total"] - (r["e_ani"] + r["e_mel"])) <= 1e-12 * abs(r["e_total"])


def _dmi_energy(m, nx, ny, nz, d, kind, vol, dx, dy, dz):
    g = m.reshape(nz, ny, nx, 3)

    def diff(a, axis, h):
        out = np.zeros_like(a)
        sl_p = [slice(None)] * 3
        sl_m = [slice(None)] * 3
        sl_c = [slice(None)] * 3
        sl_p[axis], sl_m[axis], sl_c[axis] = slice(2, None), slice(0, -2), slice(1, -1)
        out[tuple(sl_c)] = (a[tuple(sl_p)] - a[tuple(sl_m)]) / (2 * h)
        return out

    # axes of g: 0 z, 1 y, 2 x
    dxm = diff(g, 2, dx)
    dym = diff(g, 1, dy)
    dzm = diff(g, 0, dz)
    if kind == "interfacial":
        e = g[..., 2] * (dxm[..., 0] + dym[..., 1]) - (g[..., 0] * dxm[..., 2] + g[..., 1] * dym[..., 2])
    else:
        curl = np.stack([dym[..., 2] - dzm[..., 1], dzm[..., 0] - dxm[..., 2], dxm[..., 1] - dym[..., 0]], -1)
        e = np.sum(g * curl, axis=-1)
    return d * vol * e.sum()


@pytest.mark.parametrize("kind", ["interfacial", "bulk"])
def test_dmi_field_is_gradient_of_dmi_energy_in_the_interior(kind):
    nx, ny, nz = 8, 7, 6
    n = nx * ny * nz
    dx, dy, dz = 2e-9, 3e-9, 4e-9
    vol = dx * dy * dz
    d = 2.5e-3
    orc = _problem(nx, ny, nz, **({"dmi_i": d} if kind == "interfacial" else {"dmi_b": d}))
    m = _random_m(n, 7)
    # cells whose stencil and whose neighbours' stencils stay inside the box
    cells = [((z * ny) + y) * nx + x for z in range(2, nz - 2) for y in range(2, ny - 2) for x in range(2, nx - 2)]
    _grad_check(orc, lambda mm: _dmi_energy(mm, nx, ny, nz, d, kind, vol, dx, dy, dz), m, cells,
                np.full(n, 8e5), vol, rtol=1e-6)


def test_dmi_helix_closed_form():
    # interfacial DMI on m = (cos kx, 0, sin kx): H = (2D/(mu0 Ms)) (d_x m_z, 0, -d_x m_x)
    # with central differences, sin(k dx)/dx k-scaling of the continuum form
    nx, dx, d, ms = 32, 2e-9, 3e-3, 8e5
    k = 2 * np.pi / (nx * dx) * 3
    orc = _problem(nx, 1, 1, dx=dx, dy=2e-9, dz=2e-9, ms=ms, dmi_i=d)
    x = (np.arange(nx) + 0.5) * dx
    m = np.stack([np.cos(k * x), np.zeros(nx), np.sin(k * x)], 1)
    h = orc.fields(m)["h_eff"]
    pf = 2 * d / (MU0 * ms) * np.sin(k * dx) / dx
    inner = slice(1, nx - 1)
    assert np.allclose(h[inner, 0], pf * np.cos(k * x[inner]), rtol=1e-10, atol=1e-9 * pf)
    assert np.allclose(h[inner, 2], pf * np.sin(k * x[inner]), rtol=1e-10, atol=1e-9 * pf)
    assert np.abs(h[:, 1]).max() == 0.0


def test_oersted_field_obeys_ampere_law():
    nx, ny, nz = 20, 18, 3
    dx = dy = dz = 5e-9
    axis = np.array([0.3, -0.2, 1.0])
    au = axis / np.linalg.norm(axis)
    center = (47e-9, 41e-9, 7e-9)
    current, radius = 2.5e-3, 18e-9
    orc = _problem(nx, ny, nz, dx=dx, dy=dy, dz=dz, oersted=(current, radius, center, tuple(axis)))
    h = orc.fields(np.tile([1.0, 0.0, 0.0], (nx * ny * nz, 1)))["h_eff"]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([(x.ravel() + 0.5) * dx, (y.ravel() + 0.5) * dy, (z.ravel() + 0.5) * dz], 1) - center
    r = pos - np.outer(pos @ au, au)
    rp = np.linalg.norm(r, axis=1)
    # azimuthal (h . r = h . axis = 0), right-handed, |H| from the enclosed current
    assert np.abs(np.sum(h * r, 1)).max() <= 1e-9 * np.abs(h).max() * rp.max()
    assert np.abs(h @ au).max() <= 1e-9 * np.abs(h).max()
    hm = np.linalg.norm(h, axis=1)
    enclosed = np.where(rp <= radius, current * (rp / radius) ** 2, current)
    assert np.allclose(hm * 2 * np.pi * rp, enclosed, rtol=1e-12)
    assert np.all(np.einsum("ni,ni->n", np.cross(au, r), h) >= 0)


def test_oersted_envelopes():
    base = (1e-3, 10e-9, (10e-9, 10e-9, 0.0), (0.0, 0.0, 1.0))
    m = np.tile([1.0, 0.0, 0.0], (16, 1))
    ref = _problem(4, 4, 1, oersted=base).fields(m)["h_eff"]
    f, ph, off = 3e9, 0.4, 0.25
    sine = _problem(4, 4, 1, oersted=base + (1, f, ph, off, 0.0, 0.0))
    pulse = _problem(4, 4, 1, oersted=base + (2, 0.0, 0.0, 0.0, 1e-10, 2e-10))
    for t in (0.0, 3.3e-11, 1.0e-10, 1.5e-10, 2.0e-10):
        sine.set_time(t)
        pulse.set_time(t)
        assert np.allclose(sine.fields(m)["h_eff"], ref * (np.sin(2 * np.pi * f * t + ph) + off), rtol=1e-13,
                           atol=1e-13 * np.abs(ref).max())
        on = 1.0 if 1e-10 <= t < 2e-10 else 0.0
        assert np.array_equal(pulse.fields(m)["h_eff"], ref * on)


def test_sot_torque_prefactor_and_geometry():
    n, ms, thick, je = 6, 7e5, 1.5e-9, 2e11
    sigma = np.array([0.0, 2.0, 0.0])
    kw = dict(alpha=0.0, precession=False)
    m = _random_m(n, 8)
    for xi_dl, xi_fl in ((0.3, 0.0), (0.0, -0.2), (0.1, 0.05)):
        orc = _problem(3, 2, 1, ms=ms, sot=(je, xi_dl, xi_fl, tuple(sigma), thick), **kw)
        k = orc.rhs(m)
        amp = C.hbar * je / (2 * C.e * C.mu_0 * ms * thick)
        s = sigma / np.linalg.norm(sigma)
        mxs = np.cross(m, s)
        expect = amp * (-xi_dl * np.cross(m, mxs) + xi_fl * mxs)
        assert np.allclose(k, expect, rtol=1e-7, atol=1e-7 * np.abs(expect).max())
        assert np.abs(np.einsum("ni,ni->n", k, m)).max() <= 1e-9 * np.abs(k).max()


def test_slonczewski_torque_properties():
    n, ms, dz = 8, 9e5, 3e-9
    j = (0.0, 0.0, 4e10)
    p = (0.0, 0.6, 0.8)
    kw = dict(alpha=0.0, precession=False, dz=dz)
    m = _random_m(n, 9)
    # lambda = 1: the efficiency is P/2 for every m.p
    orc = _problem(4, 2, 1, ms=ms, j=j, stt_degree=0.7, slon_p=p, slon_lambda=1.0, slon_eps=0.2, **kw)
    k = orc.rhs(m)
    pre = C.hbar * 4e10 / (2 * C.e * C.mu_0 * ms * dz)
    mp = np.cross(m, p)
    expect = pre * 0.35 * (np.cross(m, mp) + 0.2 * mp)
    assert np.allclose(k, expect, rtol=1e-7, atol=1e-7 * np.abs(expect).max())
    # lambda != 1: g(m.p = 1) / g(m.p = -1) = 1 / lambda^2
    lam = 2.5
    orc = _problem(2, 1, 1, ms=ms, j=j, stt_degree=0.7, slon_p=(0.0, 0.0, 1.0), slon_lambda=lam, **kw)
    tilt = 1e-3
    mu = np.array([[np.sin(tilt), 0.0, np.cos(tilt)], [np.sin(tilt), 0.0, -np.cos(tilt)]])
    kk = orc.rhs(mu)
    ratio = np.linalg.norm(kk[0]) / np.linalg.norm(kk[1])
    cu, cd = np.cos(tilt), -np.cos(tilt)
    g = lambda c: lam ** 2 / ((lam ** 2 + 1) + (lam ** 2 - 1) * c)
    assert abs(ratio - g(cu) / g(cd)) <= 1e-12 * ratio
    assert abs(g(1.0) / g(-1.0) - 1 / lam ** 2) < 1e-15


def test_zhang_li_torque():
    # uniform state: no torque; helix along x with u || x: upwind derivative,
    # perpendicular projection and beta term, u = P mu_B J / (e Ms (1 + beta^2))
    nx, dx, ms, P, beta, jx = 24, 2e-9, 8e5, 0.6, 0.04, 3e11
    kw = dict(alpha=0.0, precession=False, dx=dx, ms=ms, j=(jx, 0.0, 0.0), stt_degree=P, stt_beta=beta)
    orc = _problem(nx, 1, 1, **kw)
    assert np.abs(orc.rhs(np.tile([0.0, 0.6, 0.8], (nx, 1)))).max() == 0.0
    k = 2 * np.pi / (nx * dx) * 2
    x = (np.arange(nx) + 0.5) * dx
    m = np.stack([np.cos(k * x), np.sin(k * x), np.zeros(nx)], 1)
    tq = orc.rhs(m)
    u = P * C.physical_constants["Bohr magneton"][0] * jx / (C.e * ms * (1 + beta ** 2))
    du = np.zeros_like(m)
    du[1:] = u * (m[1:] - m[:-1]) / dx                 # u > 0: backward difference, none at x = 0
    mdu = np.cross(m, du)
    expect = -np.cross(m, mdu) - beta * mdu
    assert np.allclose(tq, expect, rtol=1e-7, atol=1e-7 * np.abs(expect).max())
    assert np.abs(tq[0]).max() == 0.0
    # reversing the current switches to forward differences
    orc2 = _problem(nx, 1, 1, **{**kw, "j": (-jx, 0.0, 0.0)})
    tq2 = orc2.rhs(m)
    du2 = np.zeros_like(m)
    du2[:-1] = -u * (m[1:] - m[:-1]) / dx
    md2 = np.cross(m, du2)
    assert np.allclose(tq2, -np.cross(m, md2) - beta * md2, rtol=1e-7, atol=1e-7 * np.abs(expect).max())


def test_energy_total_includes_every_term():
    nx, ny, nz = 6, 5, 2
    m = _random_m(nx * ny * nz, 10)
    orc = _problem(nx, ny, nz, exchange=True, demag=True, h_ext=(1e4, 0.0, 0.0), ku1=3e4, axis=(0.0, 0.0, 1.0),
                   cubic=(2e4, 1e4, 0.0, (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)), dmi_i=1e-3,
                   mel=(1e7, 2e7, (1e-3, 0.0, 0.0, 0.0, 0.0, 2e-4)))
    r = orc.observables(m)
    parts = r["e_ex"] + r["e_demag"] + r["e_ext"] + r["e_ani"] + r["e_cub"] + r["e_dmi"] + r["e_mel"]
    assert abs(r["e_total"] - parts) <= 1e-12 * abs(parts)
    assert all(r[k] != 0.0 for k in ("e_cub", "e_dmi", "e_mel"))


# ---------------------------------------------------------------- thermal --
# The thermal field (fdm_fields.rs:818-905) is pinned by: the SplitMix64
# finaliser against the published output sequence of its reference
# implementation (state 1234567, increments of 0x9E3779B97F4A7C15); the
# sample statistics of the Box-Muller normals (zero mean, the stated
# variance with scipy's Boltzmann constant, Gaussian kurtosis, no correlation
# between components); sigma ~ dt^-1/2; the counter semantics (one value per
# step call, shared by all stages, fdm_problem.rs:225-273).

def test_splitmix64_published_sequence():
    st = 1234567
    out = []
    for _ in range(5):
        st = (st + 0x9E3779B97F4A7C15) % 2 ** 64
        out.append(O.splitmix64(st))
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]


def _thermal(nx=64, ny=64, nz=4, **kw):
    base = dict(dx=3e-9, dy=3e-9, dz=3e-9, ms=8e5, alpha=0.1, temperature=300.0, thermal_dt=1e-13)
    base.update(kw)
    return _problem(nx, ny, nz, **base)


def test_thermal_field_statistics():
    orc = _thermal()
    h = orc.fields(np.tile([0.0, 0.0, 1.0], (64 * 64 * 4, 1)))["h_eff"]
    a, gamma, vol, dt = 0.1, 2.211e5, 27e-27, 1e-13
    sigma = np.sqrt(2 * a * C.k * 300.0 / (gamma * (1 + a * a) * MU0 * 8e5 * vol * dt))
    x = h / sigma
    n = x.size
    assert abs(x.mean()) < 4 / np.sqrt(n)
    assert abs(x.var() - 1.0) < 0.03
    kurt = np.mean(x ** 4) / np.mean(x ** 2) ** 2
    assert abs(kurt - 3.0) < 0.1
    c = np.corrcoef(x.T)
    assert np.abs(c - np.eye(3)).max() < 0.03


def test_thermal_counter_and_dt_scaling():
    m = np.tile([0.0, 0.0, 1.0], (64 * 64 * 4, 1))
    a = _thermal()
    b = _thermal(thermal_dt=4e-13)
    ha, hb = a.fields(m)["h_eff"], b.fields(m)["h_eff"]
    assert np.allclose(hb, 0.5 * ha, rtol=1e-13, atol=0)
    assert np.array_equal(a.fields(m)["h_eff"], ha)                 # same counter, same field
    c = _thermal(thermal_seed=43)
    assert np.abs(np.corrcoef(c.fields(m)["h_eff"].ravel(), ha.ravel())[0, 1]) < 0.03
    # one counter value per step call (also through rejections)
    small = _thermal(nx=4, ny=4, nz=1, exchange=False, atol=1e-12)
    mm = _random_m(16, 3)
    assert small.thermal_step() == 0
    t, rep = small.step("rk45", mm, 0.0, 1e-11)
    assert rep["rejected"] >= 1 and small.thermal_step() == 1
    small.step("heun", mm, t, 1e-13)
    assert small.thermal_step() == 2
    h2 = small.fields(mm)["h_eff"]
    small.step("rk4", mm, t, 1e-13)
    assert not np.array_equal(small.fields(mm)["h_eff"], h2)
