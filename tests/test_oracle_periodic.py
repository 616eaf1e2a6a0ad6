"""Pins of the oracle's periodic boundaries (the CPU engine's
FdmBoundaryPolicy, crates/fullmag-engine/src/fdm_types.rs:94-132, and its
truncated-image demag kernel, fdm_fft.rs:140-229,611-717).

Exchange and DMI of a spin spiral that fits the box an integer number of
times are the same in every cell with wrap-around neighbours (the discrete
Laplacian's eigenvalue 2(cos(2 pi k/n) - 1)/dx^2 and the central
difference sin(2 pi k/n)/dx); the FFT convolution of the unpadded periodic
grid equals a direct sum over source cells of
K(i) = sum_{m=-I..I} N_open(i + m n), i = (x_target - x_source) mod n,
built in the test from the open tensor; more images drive a uniformly
magnetised periodic film towards the infinite sheet (H = -Ms m_z z,
no in-plane field).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, Problem, newell_kernels, random_unit_vectors

MU0 = 4e-7 * np.pi
A_EX, MS = 1.3e-11, 8.0e5


def _pb(nx, ny, nz, periodic, images=(0, 0, 0), **kw):
    base = dict(nx=nx, ny=ny, nz=nz, dx=2e-9, dy=3e-9, dz=4e-9, ms=MS, a_ex=A_EX, alpha=0.1, h_ext=None,
                periodic=periodic, periodic_images=images)
    base.update(kw)
    return Problem(**base)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_spiral_exchange_and_dmi_are_uniform_with_wrap(axis):
    n = (12, 10, 8)
    k = 3
    L = n[axis]
    idx = np.indices((n[2], n[1], n[0]))[2 - axis].ravel()
    th = 2 * np.pi * k * idx / L
    m = np.zeros((idx.size, 3))
    m[:, 0], m[:, 1] = np.cos(th), np.sin(th)  # rotates in the x-y plane
    per = tuple(a == axis for a in range(3))
    d = (2e-9, 3e-9, 4e-9)[axis]
    o = Oracle(_pb(*n, per, demag=False))
    h = o.fields(m)["h_ex"]
    lam = 2 * (np.cos(2 * np.pi * k / L) - 1) / d ** 2
    want = 2 * A_EX / (MU0 * MS) * lam * m
    assert np.abs(h - want).max() <= 1e-10 * np.abs(want).max()
    # without wrap the end cells differ
    h_open = Oracle(_pb(*n, (False, False, False), demag=False)).fields(m)["h_ex"]
    assert np.abs(h_open - want).max() > 1e-3 * np.abs(want).max()
    # bulk DMI of the spiral: H = -2D/(mu0 Ms) curl m; with m in the x-y
    # plane varying along `axis` the curl is uniform in magnitude
    D = 2e-3
    hd = Oracle(_pb(*n, per, demag=False, exchange=False, dmi_b=D)).fields(m)["h_eff"]
    s = np.sin(2 * np.pi * k / L) / d
    dm = np.stack([-np.sin(th), np.cos(th), np.zeros_like(th)], axis=1) * s  # d m / d(axis) (central)
    curl = np.zeros_like(m)
    if axis == 0:
        curl[:, 1], curl[:, 2] = -dm[:, 2], dm[:, 1]
    elif axis == 1:
        curl[:, 0], curl[:, 2] = dm[:, 2], -dm[:, 0]
    else:
        curl[:, 0], curl[:, 1] = -dm[:, 1], dm[:, 0]
    want_d = -2 * D / (MU0 * MS) * curl
    assert np.abs(hd - want_d).max() <= 1e-10 * np.abs(want_d).max()


def _direct_periodic_field(nx, ny, nz, dims, per, images, m):
    """-sum_j K(x_i - x_j) Ms m_j with the truncated image kernel."""
    dx, dy, dz = dims
    n = (nx, ny, nz)
    I = [images[a] if per[a] else 0 for a in range(3)]
    en = [n[a] * (1 + 2 * I[a]) for a in range(3)]
    nk = newell_kernels(en[0], en[1], en[2], dx, dy, dz)  # (2enz, 2eny, 2enx) each
    comps = ("xx", "yy", "zz", "xy", "xz", "yz")

    def Kfold(off):
        # offset per axis: periodic -> base (o mod n) with images, open -> o
        tot = np.zeros(6)
        bases = []
        for a in range(3):
            o = off[a]
            if per[a]:
                b = o % n[a]
                bases.append([b + q * n[a] for q in range(-I[a], I[a] + 1)])
            else:
                bases.append([o])
        for oz in bases[2]:
            for oy in bases[1]:
                for ox in bases[0]:
                    ix, iy, iz = ox % (2 * en[0]), oy % (2 * en[1]), oz % (2 * en[2])
                    tot += np.array([nk[c][iz, iy, ix] for c in comps])
        return tot

    cells = [(x, y, z) for z in range(nz) for y in range(ny) for x in range(nx)]
    h = np.zeros((len(cells), 3))
    cache = {}
    for i, ci in enumerate(cells):
        for j, cj in enumerate(cells):
            off = tuple(ci[a] - cj[a] for a in range(3))
            if off not in cache:
                cache[off] = Kfold(off)
            xx, yy, zz, xy, xz, yz = cache[off]
            M = MS * m[j]
            h[i] -= np.array([xx * M[0] + xy * M[1] + xz * M[2], xy * M[0] + yy * M[1] + yz * M[2],
                              xz * M[0] + yz * M[1] + zz * M[2]])
    return h


@pytest.mark.parametrize("shape,per,images", [((6, 5, 1), (True, False, False), (2, 0, 0)),
                                              ((4, 5, 3), (True, True, False), (1, 2, 0)),
                                              ((4, 3, 3), (False, True, True), (0, 1, 1)),
                                              ((5, 4, 1), (True, True, True), (1, 1, 3))])
def test_periodic_demag_equals_direct_image_sum(shape, per, images):
    nx, ny, nz = shape
    m = random_unit_vectors(7, nx * ny * nz)
    o = Oracle(_pb(nx, ny, nz, per, images, exchange=False))
    h = o.fields(m)["h_demag"]
    want = _direct_periodic_field(nx, ny, nz, (2e-9, 3e-9, 4e-9), per, images, m)
    assert np.abs(h - want).max() <= 1e-10 * np.abs(want).max()


def test_many_images_approach_the_infinite_sheet():
    nx, ny = 8, 8
    n = nx * ny
    mz = np.tile([0.0, 0.0, 1.0], (n, 1))
    mx = np.tile([1.0, 0.0, 0.0], (n, 1))
    errs_z, errs_x = [], []
    for I in (1, 4, 16):
        o = Oracle(_pb(nx, ny, 1, (True, True, False), (I, I, 0), dz=2e-9, exchange=False))
        errs_z.append(np.abs(o.fields(mz)["h_demag"][:, 2] + MS).max() / MS)
        errs_x.append(np.abs(o.fields(mx)["h_demag"]).max() / MS)
    assert errs_z[0] > errs_z[1] > errs_z[2] and errs_z[2] < 0.02
    assert errs_x[0] > errs_x[1] > errs_x[2] and errs_x[2] < 0.02
