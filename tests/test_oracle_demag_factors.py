"""The oracle's Newell convolution against the closed-form magnetometric
factors of a uniformly magnetised prism (tests/demag_factors.py): the
cell-averaged field summed over a fully occupied grid is -Ms N m exactly,
for films, 3-D blocks, anisotropic cells and one cell along an axis.  A
dropped or mis-scaled diagonal term or a swapped diagonal component moves
the average by far more than the tolerance (the off-diagonal sums vanish by
the prism's symmetry, so their signs are pinned by the Newell golden
vectors instead)."""
import numpy as np
import pytest

from oracle.oracle import MU0, Oracle, Problem
from tests.demag_factors import prism_factors

MS = 8.0e5
M_DIR = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)


def mean_field_ratio(h):
    """Mean of the demag field over the cells divided by -Ms m, per axis."""
    return -h.mean(axis=0) / (MS * M_DIR)


@pytest.mark.parametrize("nx,ny,nz,dx,dy,dz", [
    (24, 16, 1, 2e-9, 3e-9, 5e-9),
    (12, 10, 8, 3e-9, 2e-9, 4e-9),
    (8, 8, 8, 3e-9, 3e-9, 3e-9),
    (1, 9, 7, 4e-9, 2e-9, 3e-9),
    (20, 1, 1, 2e-9, 2e-9, 2e-9),
])
def test_uniform_prism_average_is_the_closed_form_factor(nx, ny, nz, dx, dy, dz):
    pb = Problem(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, ms=MS, a_ex=1.3e-11, alpha=0.1, h_ext=None,
                 exchange=False)
    o = Oracle(pb)
    m = np.tile(M_DIR, (nx * ny * nz, 1))
    mp = o.prepare(m)
    got = mean_field_ratio(o.fields(mp)["h_demag"])
    want = np.array(prism_factors(nx * dx, ny * dy, nz * dz))
    assert np.abs(got - want).max() <= 1e-11, (got, want)
    # energy (mu0/2) Ms^2 V m.N.m
    e = 0.5 * MU0 * MS * MS * (nx * dx * ny * dy * nz * dz) * float(np.sum(want * M_DIR * M_DIR))
    assert abs(o.observables(mp)["e_demag"] - e) <= 1e-10 * e


def test_the_pin_separates_a_transposed_prism():
    # the same grid read with x and y exchanged (a transposed operand) misses
    # the closed form by orders of magnitude more than the tolerance above
    nx, ny, dx, dy, dz = 24, 16, 2e-9, 3e-9, 5e-9
    o = Oracle(Problem(nx=nx, ny=ny, nz=1, dx=dx, dy=dy, dz=dz, ms=MS, a_ex=1.3e-11, alpha=0.1, h_ext=None,
                       exchange=False))
    got = mean_field_ratio(o.fields(o.prepare(np.tile(M_DIR, (nx * ny, 1))))["h_demag"])
    swapped = np.array(prism_factors(nx * dy, ny * dx, dz))
    assert np.abs(got - swapped).max() > 1e-3
    assert np.abs(got - np.array(prism_factors(nx * dx, ny * dy, dz))).max() <= 1e-11
