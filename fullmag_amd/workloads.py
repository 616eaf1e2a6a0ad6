"""Synthetic problem builders for the benchmark and test workloads.

Each builder returns the keyword arguments of `FdmPlan` (physical units,
SoA-free host arrays) for one named workload.  Random inputs are seeded so
every run sees the same problem.
"""
from __future__ import annotations

import numpy as np

MU0 = 4e-7 * np.pi


def film_with_hole(seed: int = 42, n: int = 512) -> dict:
    """n x n x 1 film of 2x2x5 nm cells with a centred circular hole of
    radius n/8 cells (geometry mask by cell centres), Ms 8.6e5 A/m,
    A 1.3e-11 J/m, alpha 0.02, B = (10 mT, 0, 0), seeded random m0 zeroed on
    the hole, DP45 atol 1e-5."""
    from .backend import hole_mask, random_unit_vectors
    dx = 2e-9
    mask = hole_mask(n, n, 1, dx, dx, (n // 8) * dx)
    m0 = random_unit_vectors(seed, n * n)
    m0[mask == 0] = 0.0
    return dict(nx=n, ny=n, nz=1, dx=dx, dy=dx, dz=5e-9, ms=8.6e5, a_ex=1.3e-11, alpha=0.02, m0=m0,
                integrator="dp45", h_ext=(0.01 / MU0, 0.0, 0.0), mask=mask, atol=1e-5, dt_min=1e-18,
                dt_max=1e-10, headroom=0.8)


def xorshift_uniform(seed: int, count: int) -> np.ndarray:
    """count values in [-1, 1] from the planner's xorshift64 stream
    (crates/fullmag-plan/src/util.rs:118-147: shifts 13/7/17, value
    state / u64::MAX * 2 - 1), without its rejection and normalisation."""
    if seed == 0:
        raise ValueError("xorshift64 seed 0 never leaves the zero state")
    mask = (1 << 64) - 1
    s = seed & mask
    out = np.empty(count)
    for i in range(count):
        s ^= (s << 13) & mask
        s ^= s >> 7
        s ^= (s << 17) & mask
        out[i] = (s / float(mask)) * 2.0 - 1.0
    return out


def texture_relax(seed: int = 1, nx: int = 200, ny: int = 50) -> dict:
    """Plain small relaxation: nx x ny x 1 cells of 2.5 x 2.5 x 3 nm, Ms
    8.0e5 A/m, A 1.3e-11 J/m, alpha 0.5, no anisotropy, exchange + Newell
    demag + Zeeman with B = 0; m0 = normalise(normalise(1, 0.25, 0.1) + n_i)
    with per-component noise n uniform in [-0.01, 0.01] from the seeded
    xorshift64 stream (cell-major, x y z per cell); DP45 with rtol 1e-6,
    atol 1e-5, FP64."""
    n = nx * ny
    base = np.array([1.0, 0.25, 0.1])
    base /= np.linalg.norm(base)
    m0 = base[None, :] + 0.01 * xorshift_uniform(seed, 3 * n).reshape(n, 3)
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    return dict(nx=nx, ny=ny, nz=1, dx=2.5e-9, dy=2.5e-9, dz=3e-9, ms=8.0e5, a_ex=1.3e-11, alpha=0.5, m0=m0,
                integrator="dp45", precision="double", h_ext=(0.0, 0.0, 0.0), atol=1e-5, rtol=1e-6, dt_min=1e-18,
                dt_max=1e-10, headroom=0.8)


def voronoi_axes(nx: int, ny: int, nz: int, cell: float, grains: int, seed: int):
    """Per-cell easy axes of a seeded Voronoi grain structure: `grains`
    sites uniform in the box, each cell takes the axis of the nearest site
    (cell centres, Euclidean, non-periodic), axes uniform on the sphere
    (normalised Gaussian triples).  Returns (axes (n,3), grain id (n,))."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    size = np.array([nx, ny, nz], dtype=float) * cell
    sites = rng.random((grains, 3)) * size
    ax = rng.standard_normal((grains, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    centres = np.stack([(x.ravel() + 0.5) * cell, (y.ravel() + 0.5) * cell, (z.ravel() + 0.5) * cell], axis=1)
    _, gid = cKDTree(sites).query(centres, k=1)
    return np.ascontiguousarray(ax[gid]), gid.astype(np.int64)


def granular_block(nx: int = 256, ny: int = 256, nz: int = 64, grains: int = 512, seed: int = 7) -> dict:
    """3-D granular block of 3 nm cubes: Ms 1.0e6 A/m, A 1.5e-11 J/m,
    uniaxial Ku 5e5 J/m^3 with per-grain easy axes (seeded Voronoi, axes
    uniform on the sphere), alpha 0.1, B ramp 0 -> 1 T along z over 1 ns,
    seeded random m0, DP45 atol 1e-5."""
    from .backend import random_unit_vectors
    from . import _abi as abi
    cell = 3e-9
    axes, _ = voronoi_axes(nx, ny, nz, cell, grains, seed)
    m0 = random_unit_vectors(seed, nx * ny * nz)
    return dict(nx=nx, ny=ny, nz=nz, dx=cell, dy=cell, dz=cell, ms=1.0e6, a_ex=1.5e-11, alpha=0.1, m0=m0,
                integrator="dp45", ku1=5e5, axis=(0.0, 0.0, 1.0), axis_field=axes, wave=abi.WAVE_RAMP,
                wave_amp=(0.0, 0.0, 1.0 / MU0), wave_t0=0.0, wave_t1=1e-9, atol=1e-5, dt_min=1e-18,
                dt_max=1e-10, headroom=0.8)


BATCH_SHAPES = ((64, 64, 1), (128, 64, 2), (256, 128, 1), (256, 256, 4))


def batch_sweep(count: int = 96, seed: int = 4, shapes=BATCH_SHAPES, precision: str = "single") -> list:
    """Independent systems of a parameter sweep: sizes drawn by seed from
    `shapes` (cells of 4x4x4 nm), Ms ~ U[6e5, 1.2e6] A/m, A ~ U[1e-11,
    2e-11] J/m, alpha ~ U[0.01, 0.5], seeded random m0 per system, exchange
    + demag, DP45 atol 1e-5 with its own adaptive dt.  Returns FdmPlan
    keyword dicts."""
    from .backend import random_unit_vectors
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx, ny, nz = shapes[int(rng.integers(len(shapes)))]
        out.append(dict(nx=nx, ny=ny, nz=nz, dx=4e-9, dy=4e-9, dz=4e-9, ms=float(rng.uniform(6e5, 1.2e6)),
                        a_ex=float(rng.uniform(1e-11, 2e-11)), alpha=float(rng.uniform(0.01, 0.5)),
                        m0=random_unit_vectors(seed * 1000 + i + 1, nx * ny * nz), precision=precision,
                        integrator="dp45", atol=1e-5, dt_min=1e-18, dt_max=1e-10, headroom=0.8))
    return out
