"""GPU parity of the effective-field terms and the LLG right-hand side
against the FP64 oracle, through the C ABI (copy_field after create).

Tolerances: FP64 engine vs FP64 oracle 1e-11 of the field scale (the two
differ only in FFT algorithm and summation order); FP32 engine 2e-5 of the
field scale (FP32 FFT roundoff on ~1e6 A/m fields, cf. the reference's
CPU<->CUDA thin-film band H_d 5e-4 rel, native_fdm.rs:1866-1936).
"""
import numpy as np
import pytest

from tests.cases import masked_3x3, random_case, rel_err, thin_film_8x6

pytestmark = pytest.mark.gpu

TOL = {"double": 1e-11, "single": 2e-5}


def _compare(case, precision):
    from fullmag_amd.backend import FdmBackend
    from oracle.oracle import Oracle

    orc = Oracle(case.problem())
    m = orc.prepare(case.m0)
    ref = orc.fields(m)
    with FdmBackend(case.plan(precision=precision)) as be:
        got_m = be.field("m")
        assert rel_err(got_m, m) < (1e-15 if precision == "double" else 1e-7)
        for name in ("h_ex", "h_demag", "h_eff", "h_ext"):
            got = be.field(name)
            r = ref[name]
            if not np.any(r):
                assert np.abs(got).max() == 0.0, name
                continue
            e = rel_err(got, r)
            print(f"{precision} {name}: rel err {e:.3e}", flush=True)
            assert e < TOL[precision], (name, e)
        # RHS through a stage-1 evaluation is the FSAL k1; its max norm is in the stats
        st = be.stats()
        ob = orc.observables(m)
        for k_eng, k_orc in (("exchange_energy_joules", "e_ex"), ("demag_energy_joules", "e_demag"),
                             ("external_energy_joules", "e_ext"), ("max_rhs_amplitude", "max_rhs"),
                             ("max_effective_field_amplitude", "max_heff")):
            ref_v = ob[k_orc]
            scale = max(abs(ref_v), abs(ob["e_total"]) if "energy" in k_eng else 0.0, 1e-300)
            assert abs(st[k_eng] - ref_v) <= 50 * TOL[precision] * scale, (k_eng, st[k_eng], ref_v)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_fields_masked_3x3_no_demag(precision):
    _compare(masked_3x3(False), precision)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_fields_masked_3x3_demag(precision):
    _compare(masked_3x3(True), precision)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_fields_thin_film_8x6(precision):
    _compare(thin_film_8x6(), precision)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_fields_film_with_hole_64x48(precision):
    _compare(random_case(64, 48, 1, seed=11, hole_radius_cells=10, axis=(0.3, 0.2, 1.0), ku1=4e4, ku2=1e3),
             precision)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_fields_3d_12x10x6(precision):
    _compare(random_case(12, 10, 6, seed=5, dx=3e-9, dy=3e-9, dz=3e-9), precision)


@pytest.mark.parametrize("precision", ["double"])
def test_fields_mixed_radix_200x50(precision):
    # padded 400 x 100: radix 5 and 2/4/8 passes
    _compare(random_case(200, 50, 1, seed=3, dx=2.5e-9, dy=2.5e-9, dz=3e-9, ms=8e5, alpha=0.5), precision)


@pytest.mark.parametrize("precision", ["double"])
def test_fields_odd_sizes_7x5x3(precision):
    # padded 14 x 10 x 6: radix 7, 5, 3 passes and a 3-D z pass
    _compare(random_case(7, 5, 3, seed=9, dx=2e-9, dy=3e-9, dz=4e-9), precision)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("nx,ny", [(40, 32), (24, 64), (16, 128), (20, 256), (12, 512), (8, 1024)])
def test_fields_register_four_step_y_shapes(precision, nx, ny):
    # padded y lengths 64..2048 take the register four-step y pass
    _compare(random_case(nx, ny, 1, seed=ny, hole_radius_cells=min(nx, ny) // 4), precision)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_fields_full_size_film_512(precision):
    from tests.cases import thin_film_512
    _compare(thin_film_512(), precision)


def test_fp32_film_y_pass_forms_agree(monkeypatch):
    # at py = 1024 FP32 plans take the three-step y pass (k_yf8); the 32 x 32
    # four-step pass (FULLMAG_FDM_YPASS8=0) gives the same demag field to
    # FP32 roundoff, and both the same DP45 steps within the FP32 band
    from fullmag_amd.backend import FdmBackend
    case = random_case(12, 512, 1, seed=7, hole_radius_cells=3)
    out = {}
    for form in ("1", "0"):
        monkeypatch.setenv("FULLMAG_FDM_YPASS8", form)
        with FdmBackend(case.plan(precision="single")) as be:
            h = be.field("h_demag")
            be.step(1e-13)
            out[form] = (h, be.field("m"))
    scale = np.abs(out["0"][0]).max()
    assert scale > 0
    assert np.abs(out["1"][0] - out["0"][0]).max() <= 2e-6 * scale
    assert np.abs(out["1"][1] - out["0"][1]).max() <= 1e-6


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("nx,ny", [(32, 20), (64, 12), (128, 8), (256, 6), (512, 5), (1024, 3)])
def test_fields_register_four_step_x_shapes(precision, nx, ny):
    # padded x lengths 64..2048 take the register four-step x passes
    _compare(random_case(nx, ny, 1, seed=nx + 1, hole_radius_cells=min(nx, ny) // 3), precision)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("nx,ny,nz", [(6, 32, 2), (5, 64, 4), (6, 32, 8), (5, 64, 16), (4, 128, 32), (3, 256, 64),
                                      (4, 32, 128)])
def test_fields_register_four_step_3d_shapes(precision, nx, ny, nz):
    # 3-D grids whose padded y (64..512) and z (16..256) lengths take the
    # register four-step y-line and z passes
    _compare(random_case(nx, ny, nz, seed=ny + nz, dx=3e-9, dy=3e-9, dz=3e-9), precision)
