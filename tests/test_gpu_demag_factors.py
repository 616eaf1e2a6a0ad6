"""Size-independent pins of the device demag path: a uniformly magnetised,
fully occupied grid's volume-averaged demag field is -Ms N m.

Where every cell pair is nearer than 40 cells, the Newell tensor is exact
and N equals the closed-form prism factors (tests/demag_factors.py).
Beyond that the reference switches to the point-dipole far field
(crates/fullmag-fdm-demag/src/newell.rs:212-229, followed by the engine and
the oracle), whose tensors are traceless like the exact ones: at the
benchmarks' full sizes (512x512x1 film, 256x256x64 block) and a mixed-radix
shape the trace of N stays 1 to roundoff, a square film keeps Nx = Ny, the
factors stay within 1e-4 of the closed form (the far field moves them by
about 1e-5), and the demag energy statistic equals (mu0/2) Ms^2 V sum N_i
m_i^2 of the measured field.  FP64 to 1e-10, FP32 to 1e-5; the full-size
FP64 trace to 1e-9 (the f/g difference formula cancels about ten digits at
offsets near 40 cells even with compensated sums, and the 256x256x64 block
adds up some 10^5 such offsets: measured 1.1e-10)."""
import numpy as np
import pytest

from tests.cases import Case
from tests.demag_factors import prism_factors

pytestmark = pytest.mark.gpu

MS = 8.0e5
MU0 = 4.0e-7 * np.pi
M_DIR = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)
TOL = {"double": 1e-10, "single": 1e-5}
TRACE_TOL = {"double": 1e-9, "single": 1e-5}


def _factors(precision, nx, ny, nz, dx, dy, dz):
    from fullmag_amd.backend import FdmBackend
    n = nx * ny * nz
    case = Case(nx, ny, nz, dx, dy, dz, MS, 1.3e-11, 0.1, np.tile(M_DIR, (n, 1)), exchange=False)
    print(f"uniform prism {nx}x{ny}x{nz} {precision}", flush=True)
    with FdmBackend(case.plan(precision=precision)) as be:
        h = be.field("h_demag")
        e_got = be.stats()["demag_energy_joules"]
    got = -h.mean(axis=0) / (MS * M_DIR)
    want = np.array(prism_factors(nx * dx, ny * dy, nz * dz))
    print(f"  N got {got} closed form {want}", flush=True)
    return got, want, e_got, nx * dx * ny * dy * nz * dz


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("nx,ny,nz,dx,dy,dz", [
    (28, 26, 1, 2e-9, 2e-9, 5e-9),
    (28, 24, 8, 3e-9, 2e-9, 4e-9),
    (39, 1, 1, 2e-9, 3e-9, 3e-9),
])
def test_near_field_grid_gives_the_closed_form_factor(precision, nx, ny, nz, dx, dy, dz):
    got, want, e_got, vol = _factors(precision, nx, ny, nz, dx, dy, dz)
    assert np.abs(got - want).max() <= TOL[precision], (got, want)
    e = 0.5 * MU0 * MS * MS * vol * float(np.sum(want * M_DIR * M_DIR))
    assert abs(e_got - e) <= TOL[precision] * e, (e_got, e)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("nx,ny,nz,dx,dy,dz", [
    (512, 512, 1, 2e-9, 2e-9, 5e-9),
    (256, 256, 64, 3e-9, 3e-9, 3e-9),
    (200, 50, 3, 2e-9, 3e-9, 4e-9),
])
def test_full_size_factors_keep_the_trace(precision, nx, ny, nz, dx, dy, dz):
    got, want, e_got, vol = _factors(precision, nx, ny, nz, dx, dy, dz)
    tol = TOL[precision]
    assert abs(got.sum() - 1.0) <= TRACE_TOL[precision], got.sum()
    if nx * dx == ny * dy:
        assert abs(got[0] - got[1]) <= tol * got[2], got
    assert np.abs(got - want).max() <= 1e-4, (got, want)
    e = 0.5 * MU0 * MS * MS * vol * float(np.sum(got * M_DIR * M_DIR))
    assert abs(e_got - e) <= tol * e, (e_got, e)
