"""Readout paths of the C ABI (reference context.cu:254-317,1271-1745,
fullmag_fdm.h:380-473): device preview binning against a numpy restatement
of the bin rule on non-uniform fields (z origin and stride included),
asynchronous field and preview snapshots staged on the device and copied on
the readout stream, and the H_EXT observable with a field waveform."""
import numpy as np
import pytest

from tests.cases import MU0, random_case

pytestmark = pytest.mark.gpu


def _bin_rule(f, shape, pshape, z0, zs):
    """Preview bin rule of native/backends/fdm/src/context.cu:279-296 over an
    AoS field (n, 3): ranges floor(p*full/np) .. floor((p+1)*full/np), at
    least one cell and clipped; z from z0 + p*zs over zs planes."""
    nx, ny, nz = shape
    pnx, pny, pnz = pshape
    F = f.reshape(nz, ny, nx, 3)
    out = np.zeros((pnz, pny, pnx, 3))

    def rng(p, full, npv):
        b, e = (p * full) // npv, ((p + 1) * full) // npv
        e = max(e, b + 1)
        return b, min(e, full)
    for pz in range(pnz):
        zb = min(z0 + pz * zs, nz - 1)
        ze = min(max(z0 + (pz + 1) * zs, zb + 1), nz)
        for py in range(pny):
            yb, ye = rng(py, ny, pny)
            for px in range(pnx):
                xb, xe = rng(px, nx, pnx)
                out[pz, py, px] = F[zb:ze, yb:ye, xb:xe].reshape(-1, 3).mean(axis=0)
    return out.reshape(-1, 3)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("shape,pshape,z0,zs", [((40, 30, 1), (7, 5, 1), 0, 1), ((24, 18, 6), (5, 4, 2), 1, 2),
                                                 ((24, 18, 6), (24, 18, 6), 0, 1), ((16, 12, 5), (3, 3, 3), 4, 3)])
def test_preview_bins_match_rule(precision, shape, pshape, z0, zs):
    from fullmag_amd.backend import FdmBackend
    case = random_case(*shape, seed=5, hole_radius_cells=shape[0] // 5, alpha=0.1)
    with FdmBackend(case.plan(precision=precision)) as be:
        be.step(2e-14)
        for name in ("m", "h_ex", "h_demag", "h_eff"):
            full = be.field(name)
            want = _bin_rule(full, shape, pshape, z0, zs)
            scale = max(np.abs(want).max(), 1e-300)
            got64 = be.preview(name, *pshape, z_origin=z0, z_stride=zs)
            assert np.abs(got64 - want).max() <= 1e-14 * scale, name
            got32 = be.preview(name, *pshape, z_origin=z0, z_stride=zs, dtype=np.float32)
            assert np.array_equal(got32, got64.astype(np.float32)), name
            snap = be.preview_snapshot(name, *pshape, z_origin=z0, z_stride=zs)
            assert snap.dtype == (np.float64 if precision == "double" else np.float32)
            assert np.array_equal(snap, got64.astype(snap.dtype)), name


def test_snapshot_holds_state_of_its_call_and_outlives_backend():
    from fullmag_amd import _abi as abi
    from fullmag_amd.backend import FdmBackend, _OBS
    import ctypes
    case = random_case(64, 48, 1, seed=3, hole_radius_cells=6, alpha=0.1)
    be = FdmBackend(case.plan())
    m_before = be.field("m")
    s = be.lib.fullmag_fdm_backend_begin_field_snapshot(be.h, _OBS["m"])
    ps = be.lib.fullmag_fdm_backend_begin_preview_snapshot(be.h, _OBS["m"], 8, 6, 1, 0, 1)
    assert s and ps
    # the next steps are queued on the compute stream behind the staging copy
    for _ in range(3):
        be.step(5e-14)
    m_after = be.field("m")
    assert np.abs(m_after - m_before).max() > 1e-6
    pv_before = _bin_rule(m_before, (64, 48, 1), (8, 6, 1), 0, 1)
    be.close()
    try:
        for handle, wait, shape in ((s, be.lib.fullmag_fdm_field_snapshot_wait, (3, case.n)),
                                    (ps, be.lib.fullmag_fdm_preview_snapshot_wait, (48, 3))):
            p, ln, d = ctypes.c_void_p(), ctypes.c_uint64(), abi.SnapshotDesc()
            assert wait(handle, ctypes.byref(p), ctypes.byref(ln), ctypes.byref(d)) == abi.OK
            buf = (ctypes.c_char * ln.value).from_address(p.value)
            got = np.frombuffer(bytes(buf), dtype=np.float64).reshape(shape)
            if shape[0] == 3:
                assert np.array_equal(got.T, m_before)
            else:
                assert np.abs(got - pv_before).max() <= 1e-15
    finally:
        be.lib.fullmag_fdm_field_snapshot_destroy(s)
        be.lib.fullmag_fdm_preview_snapshot_destroy(ps)


def test_h_ext_includes_waveform():
    from fullmag_amd.backend import FdmBackend
    case = random_case(20, 16, 1, seed=4, hole_radius_cells=3, alpha=0.1, h_ext=(1e3, -2e3, 5e2))
    n = case.n
    wmask = (np.arange(n) % 3 == 0).astype(np.uint8)
    amp = np.array([0.0, 3e4, -1e4])
    case.wave_kind, case.wave_amp, case.wave_t0, case.wave_t1, case.wave_mask = 1, tuple(amp), 0.0, 1e-12, wmask
    with FdmBackend(case.plan()) as be:
        for _ in range(4):
            st = be.step(1e-13)
        t = st["time_seconds"]
        f = min(max(t / 1e-12, 0.0), 1.0)
        assert 0.0 < f < 1.0
        act = case.mask.astype(bool)
        want = np.zeros((n, 3))
        want[act] = np.array(case.h_ext)
        want[act & (wmask == 1)] += f * amp
        h = be.field("h_ext")
        assert np.abs(h - want).max() <= 1e-12 * np.abs(want).max()
        assert np.array_equal(be.field_snapshot("h_ext").T, h)
        # the Zeeman energy is the one of this field: -mu0 Ms V sum m.H_ext
        m = be.field("m")
        e = -MU0 * case.ms * case.dx * case.dy * case.dz * float((m * h).sum())
        be.refresh()
        assert abs(be.stats()["external_energy_joules"] - e) <= 1e-12 * abs(e)
        pv = be.preview("h_ext", 5, 4, 1)
        assert np.abs(pv - _bin_rule(h, (20, 16, 1), (5, 4, 1), 0, 1)).max() <= 1e-12 * np.abs(h).max()
