"""Pins of the oracle's boundary-correction tiers (the reference's CUDA
kernels native/backends/fdm/src/exchange_t0_fp64.cu, exchange_t1_fp64.cu and
demag_boundary_fp64.cu; the CPU engine has none).

Each check is fixed by the mathematics rather than by the oracle's own code:
T0 with unit face links on interior faces and T1 with distances equal to
the cell size are the standard 6-neighbour Laplacian; the field scales with
the face links and with 1/max(phi, floor); T1's boundary weight is the
ghost-point (Neumann) formula 2/(d (d + delta)) of Garcia-Cervera et al.
(2003); T0 with symmetric links is the gradient of its phi-weighted field-form
energy; a uniform volume fraction scales the demag source linearly; the
sparse correction is -dN.(phi Ms m) at its target only, and repeated entries
add up.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, Problem, hole_mask, random_unit_vectors

MU0 = 4e-7 * np.pi
A_EX, MS = 1.3e-11, 8.0e5


def _problem(nx, ny, nz, mask=None, **kw):
    base = dict(nx=nx, ny=ny, nz=nz, dx=2e-9, dy=3e-9, dz=4e-9, ms=MS, a_ex=A_EX, alpha=0.1, mask=mask,
                h_ext=None)
    base.update(kw)
    return Problem(**base)


def _state(n, seed, mask=None):
    m = random_unit_vectors(seed, n)
    if mask is not None:
        m[mask == 0] = 0.0
    return m


def _interior_links(nx, ny, nz, mask):
    """Face links 1 towards active neighbours inside the box, 0 otherwise."""
    act = np.ones(nx * ny * nz, bool) if mask is None else mask.astype(bool)
    a = act.reshape(nz, ny, nx)
    out = []
    for axis in (2, 1, 0):  # x, y, z in (z, y, x) storage
        for step in (+1, -1):
            nb = np.zeros_like(a)
            sl_dst = [slice(None)] * 3
            sl_src = [slice(None)] * 3
            if step > 0:
                sl_dst[axis], sl_src[axis] = slice(0, -1), slice(1, None)
            else:
                sl_dst[axis], sl_src[axis] = slice(1, None), slice(0, -1)
            nb[tuple(sl_dst)] = a[tuple(sl_src)]
            out.append((nb & a).astype(np.float64).ravel())
    return tuple(out)


def _fields(pb, m):
    o = Oracle(pb)
    mm = o.prepare(m)
    return o.fields(mm), o.observables(mm)


@pytest.mark.parametrize("nz", [1, 3])
def test_t0_with_interior_links_is_the_standard_exchange(nz):
    nx, ny = 10, 8
    mask = hole_mask(nx, ny, nz, 2e-9, 3e-9, 5e-9)
    m = _state(nx * ny * nz, 3, mask)
    ref, rep0 = _fields(_problem(nx, ny, nz, mask, demag=False), m)
    links = _interior_links(nx, ny, nz, mask)
    got, rep1 = _fields(_problem(nx, ny, nz, mask, demag=False, boundary_tier=1,
                                 volume_fraction=np.ones(nx * ny * nz), face_links=links), m)
    s = np.abs(ref["h_ex"]).max()
    assert np.abs(got["h_ex"] - ref["h_ex"]).max() <= 1e-13 * s
    assert abs(rep1["e_ex"] - rep0["e_ex"]) <= 1e-12 * abs(rep0["e_ex"])


@pytest.mark.parametrize("nz", [1, 3])
def test_t1_with_cell_size_distances_is_the_standard_exchange(nz):
    nx, ny = 9, 7
    mask = hole_mask(nx, ny, nz, 2e-9, 3e-9, 4e-9)
    m = _state(nx * ny * nz, 5, mask)
    n = nx * ny * nz
    ref, _ = _fields(_problem(nx, ny, nz, mask, demag=False), m)
    d = tuple(np.full(n, v) for v in (2e-9, 2e-9, 3e-9, 3e-9, 4e-9, 4e-9))
    got, _ = _fields(_problem(nx, ny, nz, mask, demag=False, boundary_tier=2, volume_fraction=np.ones(n),
                              deltas=d), m)
    assert np.abs(got["h_ex"] - ref["h_ex"]).max() <= 1e-13 * np.abs(ref["h_ex"]).max()


def test_t0_scales_with_links_and_inverse_clamped_phi():
    nx, ny, nz = 6, 5, 2
    n = nx * ny * nz
    m = _state(n, 8)
    links = _interior_links(nx, ny, nz, None)
    phi = np.random.default_rng(1).uniform(0.0, 1.0, n)
    phi[3] = 0.01        # below the floor
    phi[7] = 0.0         # empty: no exchange field
    ref, _ = _fields(_problem(nx, ny, nz, demag=False, boundary_tier=1, volume_fraction=np.ones(n),
                              face_links=links), m)
    got, _ = _fields(_problem(nx, ny, nz, demag=False, boundary_tier=1, volume_fraction=phi,
                              face_links=tuple(0.37 * f for f in links), phi_floor=0.08), m)
    scale = 0.37 / np.maximum(phi, 0.08)
    scale[7] = 0.0
    want = ref["h_ex"] * scale[:, None]
    assert np.abs(got["h_ex"] - want).max() <= 1e-13 * np.abs(want).max()


def test_t1_boundary_weight_is_the_ghost_point_formula():
    # two cells along x; the boundary lies delta beyond each cell's outer face
    # side; the neighbour's weight is 2 / (dx (dx + delta_opposite))
    dx, delta = 2e-9, 0.7e-9
    m = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    n = 2
    d = (np.array([dx, delta]), np.array([delta, dx]), np.full(n, 3e-9), np.full(n, 3e-9), np.full(n, 4e-9),
         np.full(n, 4e-9))
    got, _ = _fields(_problem(2, 1, 1, demag=False, boundary_tier=2, volume_fraction=np.array([1.0, 0.6]),
                              deltas=d), m)
    pre = 2 * A_EX / (MU0 * MS)
    h0 = pre * 2 / (dx * (dx + delta)) * (m[1] - m[0])
    h1 = pre / 0.6 * 2 / (dx * (dx + delta)) * (m[0] - m[1])
    assert np.allclose(got["h_ex"][0], h0, rtol=1e-13, atol=0)
    assert np.allclose(got["h_ex"][1], h1, rtol=1e-13, atol=0)
    # a distance below the minimum is clamped to it
    d2 = (np.array([dx, 1e-15]), np.array([1e-15, dx])) + d[2:]
    got2, _ = _fields(_problem(2, 1, 1, demag=False, boundary_tier=2, volume_fraction=np.ones(2), deltas=d2,
                               delta_min=0.25e-9), m)
    assert np.allclose(got2["h_ex"][0], pre * 2 / (dx * (dx + 0.25e-9)) * (m[1] - m[0]), rtol=1e-13, atol=0)


def test_t0_field_is_the_gradient_of_its_energy():
    # symmetric links (f_ij = f_ji) and phi above the floor: the phi-weighted
    # field-form energy is a symmetric quadratic form whose gradient is
    # -mu0 Ms phi V H
    nx, ny, nz = 5, 4, 3
    n = nx * ny * nz
    rng = np.random.default_rng(4)
    a = np.ones((nz, ny, nx))
    fx = rng.uniform(0.2, 1.0, (nz, ny, nx - 1))
    fy = rng.uniform(0.2, 1.0, (nz, ny - 1, nx))
    fz = rng.uniform(0.2, 1.0, (nz - 1, ny, nx))
    xp, xm, yp, ym, zp, zm = (np.zeros_like(a) for _ in range(6))
    xp[:, :, :-1], xm[:, :, 1:] = fx, fx
    yp[:, :-1, :], ym[:, 1:, :] = fy, fy
    zp[:-1], zm[1:] = fz, fz
    phi = rng.uniform(0.3, 1.0, n)
    pb = _problem(nx, ny, nz, demag=False, boundary_tier=1, volume_fraction=phi,
                  face_links=tuple(f.ravel() for f in (xp, xm, yp, ym, zp, zm)))
    o = Oracle(pb)
    m = _state(n, 6)
    h = o.fields(m)["h_ex"]
    vol = 2e-9 * 3e-9 * 4e-9
    for cell, c in ((0, 0), (17, 1), (n - 1, 2), (33, 0)):
        eps = 1e-6
        mp, mm = m.copy(), m.copy()
        mp[cell, c] += eps
        mm[cell, c] -= eps
        g = (o.observables(mp)["e_ex"] - o.observables(mm)["e_ex"]) / (2 * eps)
        want = -MU0 * MS * phi[cell] * vol * h[cell, c]
        assert abs(g - want) <= 1e-7 * abs(want)


def test_uniform_volume_fraction_scales_the_demag_source():
    # a tier without stencil data keeps the standard exchange; the demag
    # source is phi Ms m and every energy is weighted by phi
    nx, ny, nz = 8, 6, 2
    n = nx * ny * nz
    m = _state(n, 9)
    ref, r0 = _fields(_problem(nx, ny, nz), m)
    got, r1 = _fields(_problem(nx, ny, nz, boundary_tier=1, volume_fraction=np.full(n, 0.4)), m)
    assert np.abs(got["h_demag"] - 0.4 * ref["h_demag"]).max() <= 1e-12 * np.abs(ref["h_demag"]).max()
    assert np.abs(got["h_ex"] - ref["h_ex"]).max() <= 1e-13 * np.abs(ref["h_ex"]).max()
    assert abs(r1["e_demag"] - 0.16 * r0["e_demag"]) <= 1e-12 * abs(r0["e_demag"])
    assert abs(r1["e_ex"] - 0.4 * r0["e_ex"]) <= 1e-12 * abs(r0["e_ex"])


def test_sparse_demag_correction_closed_form():
    nx, ny, nz = 6, 5, 1
    n = nx * ny * nz
    m = _state(n, 12)
    phi = np.random.default_rng(2).uniform(0.5, 1.0, n)
    base = dict(boundary_tier=1, volume_fraction=phi)
    ref, _ = _fields(_problem(nx, ny, nz, **base), m)
    N1 = np.array([0.11, -0.02, 0.03, 0.2, 0.05, -0.07])
    N2 = np.array([-0.04, 0.01, 0.0, 0.09, -0.03, 0.12])
    # target 7 with sources 8 and 13 (and a skipped -1), target 20 with 20;
    # target 7 again with source 8 and tensor N2
    tg = np.array([7, 20, 7], np.int32)
    src = np.array([[8, 13, -1], [20, -1, -1], [8, -1, -1]], np.int32)
    ten = np.zeros((3, 3, 6))
    ten[0, 0], ten[0, 1], ten[1, 0], ten[2, 0] = N1, N2, N2, N2
    got, _ = _fields(_problem(nx, ny, nz, demag_corr=(tg, src, ten), **base), m)

    def sym(v):
        return np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]])

    def src_m(i):
        return phi[i] * MS * m[i]
    want = ref["h_demag"].copy()
    want[7] -= sym(N1) @ src_m(8) + sym(N2) @ src_m(13) + sym(N2) @ src_m(8)
    want[20] -= sym(N2) @ src_m(20)
    assert np.abs(got["h_demag"] - want).max() <= 1e-12 * np.abs(want).max()
    assert np.array_equal(np.delete(got["h_demag"], [7, 20], axis=0), np.delete(ref["h_demag"], [7, 20], axis=0))
