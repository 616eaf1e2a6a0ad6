"""Pins of the oracle's multilayer runtime (crates/fullmag-engine/src/
multilayer.rs:129-260 with the shifted pair kernels of
crates/fullmag-fdm-demag/src/newell.rs:528-660 and shifted_kernel.rs).

Layers stacked exactly one cell apart are the uniform 3-D grid, so their
demag field equals the 3-D convolution; at arbitrary heights the mutual
energy is reciprocal (M2 . H(M1) summed over layer 2 equals M1 . H(M2)
over layer 1, which fixes the parity of the shifted xz / yz components); a
layer's own field does not depend on the other layers' heights; exchange
stays inside each layer.
"""
import numpy as np

from oracle.oracle import Oracle, Problem, random_unit_vectors

MS = 8.0e5


def _pb(nx, ny, nz, layer_z=None, **kw):
    base = dict(nx=nx, ny=ny, nz=nz, dx=2e-9, dy=3e-9, dz=4e-9, ms=MS, a_ex=1.3e-11, alpha=0.1, h_ext=None,
                layer_z=layer_z)
    base.update(kw)
    return Problem(**base)


def test_layers_one_cell_apart_are_the_uniform_grid():
    nx, ny, L = 9, 7, 3
    m = random_unit_vectors(5, nx * ny * L)
    ref = Oracle(_pb(nx, ny, L, exchange=False)).fields(m)["h_demag"]
    got = Oracle(_pb(nx, ny, L, layer_z=np.arange(L) * 4e-9 + 1e-8, exchange=False)).fields(m)["h_demag"]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_mutual_energy_is_reciprocal_at_arbitrary_heights():
    nx, ny = 8, 6
    A = nx * ny
    m = random_unit_vectors(11, 2 * A)
    z = np.array([0.0, 10.7e-9])
    o = Oracle(_pb(nx, ny, 2, layer_z=z, exchange=False))
    m1 = m.copy()
    m1[A:] = 0.0
    m2 = m.copy()
    m2[:A] = 0.0
    h1 = o.fields(m1)["h_demag"]
    h2 = o.fields(m2)["h_demag"]
    e21 = np.sum(m[A:] * h1[A:])
    e12 = np.sum(m[:A] * h2[:A])
    assert abs(e21 - e12) <= 1e-12 * np.abs(h1[A:]).max() * A
    assert abs(e21) > 1e-3 * np.abs(h1[A:]).max()   # the layers do interact
    # swapping the heights swaps the roles (the kernel depends on z_dst - z_src)
    o2 = Oracle(_pb(nx, ny, 2, layer_z=z[::-1].copy(), exchange=False))
    hs = o2.fields(np.concatenate([m[A:], m[:A]]))["h_demag"]
    full = o.fields(m)["h_demag"]
    assert np.abs(np.concatenate([hs[A:], hs[:A]]) - full).max() <= 1e-12 * np.abs(full).max()


def test_self_field_and_exchange_stay_inside_each_layer():
    nx, ny, L = 8, 6, 3
    A = nx * ny
    m = random_unit_vectors(3, L * A)
    z = np.array([0.0, 9.1e-9, 31.0e-9])
    o = Oracle(_pb(nx, ny, L, layer_z=z))
    f = o.fields(m)
    for l in range(L):
        single = Oracle(_pb(nx, ny, 1)).fields(m[l * A:(l + 1) * A])
        sl = slice(l * A, (l + 1) * A)
        assert np.abs(f["h_ex"][sl] - single["h_ex"]).max() <= 1e-12 * np.abs(single["h_ex"]).max()
        # the self part of the demag field: only this layer magnetised
        ml = np.zeros_like(m)
        ml[sl] = m[sl]
        hd = o.fields(ml)["h_demag"][sl]
        assert np.abs(hd - single["h_demag"]).max() <= 1e-12 * np.abs(single["h_demag"]).max()
