"""Pins the oracle's region-table exchange (the reference CUDA kernel's
per-neighbour form, native/backends/fdm/src/exchange_fp64.cu:80-114:
H_i = 2/(mu0 Ms) sum_j A[r_i][r_j] (m_j - m_i)/d^2 over the clamped
neighbours, inactive neighbours contributing nothing).  The checks do not
reuse the oracle's formula:
- a two-cell and a three-cell chain against the closed form written out by
  hand, with an asymmetric table, so that a transposed lookup
  (A[r_j][r_i]) or a neighbour taken from the wrong side fails;
- a table that is A everywhere equals the uniform-stiffness stencil (the
  reference runner's default table, native_fdm.rs:215-238, reduces to it);
- a table with A_ij = 0 between regions decouples them exactly as the
  inactive-neighbour rule does (each region alone, the other masked out);
- for a symmetric table the field-form energy -1/2 mu0 Ms V sum m.H equals
  the pair sum of A_ij |m_j - m_i|^2 / d^2 V over the grid's links
  (computed here with numpy), which also fixes the prefactor and the
  per-axis weights."""
import numpy as np
import pytest

from oracle import oracle as O

MU0 = 4e-7 * np.pi
A0 = 1.3e-11


def _oracle(nx, ny, nz, regions=None, lut=None, mask=None, a_ex=A0, dx=2e-9, dy=3e-9, dz=4e-9):
    return O.Oracle(O.Problem(nx, ny, nz, dx, dy, dz, 8e5, a_ex, 0.1, demag=False, mask=mask, regions=regions,
                              exchange_lut=lut))


def _table(pairs):
    lut = np.zeros((256, 256))
    for (i, j), v in pairs.items():
        lut[i, j] = v
    return lut.ravel()


def test_two_cell_closed_form_with_asymmetric_table():
    dx = 2e-9
    lut = _table({(0, 0): A0, (1, 1): 2 * A0, (0, 1): 0.7 * A0, (1, 0): 0.3 * A0})
    o = _oracle(2, 1, 1, regions=np.array([0, 1], dtype=np.uint32), lut=lut, dx=dx)
    m = np.array([[1.0, 0.0, 0.0], [0.6, 0.8, 0.0]])
    h = o.fields(m)["h_ex"]
    pre = 2.0 / (MU0 * 8e5)
    assert np.allclose(h[0], pre * 0.7 * A0 * (m[1] - m[0]) / dx ** 2, rtol=1e-14, atol=0)
    assert np.allclose(h[1], pre * 0.3 * A0 * (m[0] - m[1]) / dx ** 2, rtol=1e-14, atol=0)


def test_three_cell_chain_along_y():
    dy = 3e-9
    lut = _table({(0, 1): 1.0 * A0, (1, 0): 2.0 * A0, (1, 2): 3.0 * A0, (2, 1): 4.0 * A0, (0, 2): 9.0 * A0,
                  (2, 0): 9.0 * A0})
    o = _oracle(1, 3, 1, regions=np.array([0, 1, 2], dtype=np.uint32), lut=lut, dy=dy)
    m = O.random_unit_vectors(5, 3)
    h = o.fields(m)["h_ex"]
    pre = 2.0 / (MU0 * 8e5) / dy ** 2
    # cell 0 sees only cell 1, cell 2 only cell 1 (the 0-2 entries are not
    # neighbours and must not enter)
    assert np.allclose(h[0], pre * 1.0 * A0 * (m[1] - m[0]), rtol=1e-13, atol=0)
    assert np.allclose(h[1], pre * (2.0 * A0 * (m[0] - m[1]) + 3.0 * A0 * (m[2] - m[1])), rtol=1e-13, atol=0)
    assert np.allclose(h[2], pre * 4.0 * A0 * (m[1] - m[2]), rtol=1e-13, atol=0)


@pytest.mark.parametrize("shape", [(9, 7, 1), (6, 5, 4)])
def test_constant_table_equals_uniform_stencil(shape):
    nx, ny, nz = shape
    n = nx * ny * nz
    mask = (np.arange(n) % 7 != 3).astype(np.uint8)
    regions = (np.arange(n) * 7919 % 5).astype(np.uint32)
    lut = np.full(256 * 256, A0)
    m = O.random_unit_vectors(11, n)
    m[mask == 0] = 0.0
    h_lut = _oracle(nx, ny, nz, regions=regions, lut=lut, mask=mask).fields(m)["h_ex"]
    h_uni = _oracle(nx, ny, nz, mask=mask).fields(m)["h_ex"]
    assert np.abs(h_lut - h_uni).max() <= 1e-13 * np.abs(h_uni).max()


@pytest.mark.parametrize("shape", [(10, 8, 1), (6, 5, 3)])
def test_zero_coupling_decouples_like_masked_regions(shape):
    nx, ny, nz = shape
    n = nx * ny * nz
    x = np.arange(n) % nx
    regions = (x >= nx // 2).astype(np.uint32) + 2 * ((np.arange(n) // (nx * ny)) % 2 == 1)
    lut = _table({(r, r): A0 * (1 + 0.25 * r) for r in range(4)})
    m = O.random_unit_vectors(3, n)
    h = _oracle(nx, ny, nz, regions=regions.astype(np.uint32), lut=lut).fields(m)["h_ex"]
    for r in range(4):
        sel = regions == r
        if not sel.any():
            continue
        alone = _oracle(nx, ny, nz, mask=sel.astype(np.uint8), a_ex=A0 * (1 + 0.25 * r)).fields(
            np.where(sel[:, None], m, 0.0))["h_ex"]
        assert np.abs(h[sel] - alone[sel]).max() <= 1e-13 * np.abs(alone).max()


@pytest.mark.parametrize("shape", [(8, 6, 1), (5, 4, 3)])
def test_field_energy_equals_pair_sum_for_symmetric_table(shape):
    nx, ny, nz = shape
    n = nx * ny * nz
    dx, dy, dz = 2e-9, 3e-9, 4e-9
    rng = np.random.default_rng(1)
    regions = rng.integers(0, 3, n).astype(np.uint32)
    mask = (rng.random(n) > 0.15).astype(np.uint8)
    tab = np.zeros((256, 256))
    sym = A0 * np.array([[1.0, 0.4, 1.7], [0.4, 2.0, 0.0], [1.7, 0.0, 0.6]])
    tab[:3, :3] = sym
    o = _oracle(nx, ny, nz, regions=regions, lut=tab.ravel(), mask=mask)
    m = O.random_unit_vectors(8, n)
    m[mask == 0] = 0.0
    e_field = o.observables(m)["e_ex"]
    M = m.reshape(nz, ny, nx, 3)
    R = regions.reshape(nz, ny, nx)
    act = mask.reshape(nz, ny, nx).astype(bool)
    e_pair = 0.0
    for ax, d in ((2, dx), (1, dy), (0, dz)):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[ax] = slice(0, -1)
        sl_b[ax] = slice(1, None)
        a, b = tuple(sl_a), tuple(sl_b)
        both = act[a] & act[b]
        diff2 = ((M[b] - M[a]) ** 2).sum(-1)
        e_pair += (tab[R[a], R[b]] * diff2 * both).sum() / d ** 2
    e_pair *= dx * dy * dz
    assert abs(e_field - e_pair) <= 1e-12 * abs(e_pair)
