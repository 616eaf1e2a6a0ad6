"""The allocation-free preview entry point
(fullmag_fdm_backend_copy_field_preview_f64_noalloc) through the C ABI:
bit-identical to copy_field_preview_f64 on two calls in a row, the second
smaller than the first (the staging array keeps the first call's contents,
which must not leak into the second), for the state and an evaluated field,
in both state precisions; the size limit is enforced with last_error set."""
import ctypes

import numpy as np
import pytest

from tests.cases import random_case

pytestmark = pytest.mark.gpu


def _noalloc(be, name, pnx, pny, pnz, z0=0, zs=1):
    from fullmag_amd import _abi as abi
    from fullmag_amd.backend import _OBS
    out = np.full((pnx * pny * pnz, 3), np.nan)
    rc = be.lib.fullmag_fdm_backend_copy_field_preview_f64_noalloc(
        be.h, _OBS[name], pnx, pny, pnz, z0, zs, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size)
    return rc, out


@pytest.mark.parametrize("precision", ["double", "single"])
def test_noalloc_preview_equals_preview(precision):
    from fullmag_amd import _abi as abi
    from fullmag_amd.backend import FdmBackend
    case = random_case(96, 80, 3, seed=31, hole_radius_cells=12, alpha=0.1)
    with FdmBackend(case.plan(precision=precision)) as be:
        be.step(3e-14)
        for name in ("m", "h_demag", "h_eff", "h_ext"):
            for shape in ((48, 40, 3, 0, 1), (7, 5, 1, 1, 2)):
                rc, got = _noalloc(be, name, *shape)
                assert rc == abi.OK, be.lib.fullmag_fdm_backend_last_error(be.h)
                want = be.preview(name, *shape[:3], z_origin=shape[3], z_stride=shape[4])
                assert np.array_equal(got, want), (name, shape)


def test_noalloc_preview_rejects_oversize():
    from fullmag_amd import _abi as abi
    from fullmag_amd.backend import FdmBackend
    case = random_case(300, 300, 1, seed=2, alpha=0.1)
    with FdmBackend(case.plan(precision="single")) as be:
        rc, _ = _noalloc(be, "m", 300, 300, 1)
        assert rc == abi.ERR_INVALID
        assert b"exceeds the allocation-free staging array" in be.lib.fullmag_fdm_backend_last_error(be.h)
        rc, got = _noalloc(be, "m", 170, 170, 1)  # 28900 cells: fits
        assert rc == abi.OK
        assert np.array_equal(got, be.preview("m", 170, 170, 1))
