"""ctypes mirror of include/fullmag_fdm.h and include/fullmag_fdm_ext.h.

The structure layouts are those of the reference ABI
(native/include/fullmag_fdm.h:84-306, mirrored in Rust by
crates/fullmag-fdm-sys/src/lib.rs:71-262).
"""
from __future__ import annotations

import ctypes
import os

c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_i32_p = ctypes.POINTER(ctypes.c_int32)

OK = 0
ERR_INVALID = -1
ERR_DEVICE = -2
ERR_INTERNAL = -3
ERR_INTERRUPTED = -4
MAX_EXCHANGE_REGIONS = 256

PRECISION_SINGLE = 1
PRECISION_DOUBLE = 2

INTEGRATOR_HEUN = 1
INTEGRATOR_DP45 = 2
INTEGRATOR_ABM3 = 3
INTEGRATOR_RK4 = 4
INTEGRATOR_RK23 = 5

OBSERVABLE_M = 1
OBSERVABLE_H_EX = 2
OBSERVABLE_H_DEMAG = 3
OBSERVABLE_H_EXT = 4
OBSERVABLE_H_EFF = 5

SNAPSHOT_F32 = 1
SNAPSHOT_F64 = 2

BOUNDARY_NONE = 0
BOUNDARY_VOLUME = 1
BOUNDARY_FULL = 2


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


class MaterialDesc(ctypes.Structure):
    _fields_ = [("saturation_magnetisation", ctypes.c_double), ("exchange_stiffness", ctypes.c_double),
                ("damping", ctypes.c_double), ("gyromagnetic_ratio", ctypes.c_double)]


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("grid", GridDesc),
        ("material", MaterialDesc),
        ("precision", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("disable_precession", ctypes.c_int),
        ("enable_exchange", ctypes.c_int),
        ("enable_demag", ctypes.c_int),
        ("has_external_field", ctypes.c_int),
        ("external_field_am", ctypes.c_double * 3),
        ("has_uniaxial_anisotropy", ctypes.c_int),
        ("uniaxial_anisotropy_constant", ctypes.c_double),
        ("uniaxial_anisotropy_k2", ctypes.c_double),
        ("anisotropy_axis", ctypes.c_double * 3),
        ("ku1_field", c_double_p),
        ("ku2_field", c_double_p),
        ("has_cubic_anisotropy", ctypes.c_int),
        ("cubic_Kc1", ctypes.c_double),
        ("cubic_Kc2", ctypes.c_double),
        ("cubic_Kc3", ctypes.c_double),
        ("cubic_axis1", ctypes.c_double * 3),
        ("cubic_axis2", ctypes.c_double * 3),
        ("kc1_field", c_double_p),
        ("kc2_field", c_double_p),
        ("kc3_field", c_double_p),
        ("has_interfacial_dmi", ctypes.c_int),
        ("dmi_D_interfacial", ctypes.c_double),
        ("has_bulk_dmi", ctypes.c_int),
        ("dmi_D_bulk", ctypes.c_double),
        ("has_magnetoelastic", ctypes.c_int),
        ("mel_b1", ctypes.c_double),
        ("mel_b2", ctypes.c_double),
        ("mel_strain", ctypes.c_double * 6),
        ("temperature", ctypes.c_double),
        ("current_density_x", ctypes.c_double),
        ("current_density_y", ctypes.c_double),
        ("current_density_z", ctypes.c_double),
        ("stt_degree", ctypes.c_double),
        ("stt_beta", ctypes.c_double),
        ("stt_p_x", ctypes.c_double),
        ("stt_p_y", ctypes.c_double),
        ("stt_p_z", ctypes.c_double),
        ("stt_lambda", ctypes.c_double),
        ("stt_epsilon_prime", ctypes.c_double),
        ("has_sot", ctypes.c_int),
        ("sot_je", ctypes.c_double),
        ("sot_xi_dl", ctypes.c_double),
        ("sot_xi_fl", ctypes.c_double),
        ("sot_sigma", ctypes.c_double * 3),
        ("sot_thickness", ctypes.c_double),
        ("has_oersted_cylinder", ctypes.c_int),
        ("oersted_current", ctypes.c_double),
        ("oersted_radius", ctypes.c_double),
        ("oersted_center", ctypes.c_double * 3),
        ("oersted_axis", ctypes.c_double * 3),
        ("oersted_time_dep_kind", ctypes.c_uint32),
        ("oersted_time_dep_freq", ctypes.c_double),
        ("oersted_time_dep_phase", ctypes.c_double),
        ("oersted_time_dep_offset", ctypes.c_double),
        ("oersted_time_dep_t_on", ctypes.c_double),
        ("oersted_time_dep_t_off", ctypes.c_double),
        ("demag_kernel_xx_spectrum", c_double_p),
        ("demag_kernel_yy_spectrum", c_double_p),
        ("demag_kernel_zz_spectrum", c_double_p),
        ("demag_kernel_xy_spectrum", c_double_p),
        ("demag_kernel_xz_spectrum", c_double_p),
        ("demag_kernel_yz_spectrum", c_double_p),
        ("demag_kernel_spectrum_len", ctypes.c_uint64),
        ("active_mask", c_u8_p),
        ("active_mask_len", ctypes.c_uint64),
        ("region_mask", c_u32_p),
        ("region_mask_len", ctypes.c_uint64),
        ("exchange_lut", c_double_p),
        ("exchange_lut_len", ctypes.c_uint64),
        ("boundary_correction", ctypes.c_int),
        ("boundary_phi_floor", ctypes.c_double),
        ("boundary_delta_min", ctypes.c_double),
        ("volume_fraction", c_double_p),
        ("volume_fraction_len", ctypes.c_uint64),
        ("face_link_xp", c_double_p),
        ("face_link_xm", c_double_p),
        ("face_link_yp", c_double_p),
        ("face_link_ym", c_double_p),
        ("face_link_zp", c_double_p),
        ("face_link_zm", c_double_p),
        ("delta_xp", c_double_p),
        ("delta_xm", c_double_p),
        ("delta_yp", c_double_p),
        ("delta_ym", c_double_p),
        ("delta_zp", c_double_p),
        ("delta_zm", c_double_p),
        ("has_demag_boundary_corr", ctypes.c_int),
        ("demag_corr_target_idx", c_i32_p),
        ("demag_corr_source_idx", c_i32_p),
        ("demag_corr_tensor", c_double_p),
        ("demag_corr_target_count", ctypes.c_uint32),
        ("demag_corr_stencil_size", ctypes.c_uint32),
        ("initial_magnetization_xyz", c_double_p),
        ("initial_magnetization_len", ctypes.c_uint64),
        ("adaptive_max_error", ctypes.c_double),
        ("adaptive_dt_min", ctypes.c_double),
        ("adaptive_dt_max", ctypes.c_double),
        ("adaptive_headroom", ctypes.c_double),
    ]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("step", ctypes.c_uint64),
        ("time_seconds", ctypes.c_double),
        ("dt_seconds", ctypes.c_double),
        ("exchange_energy_joules", ctypes.c_double),
        ("demag_energy_joules", ctypes.c_double),
        ("external_energy_joules", ctypes.c_double),
        ("anisotropy_energy_joules", ctypes.c_double),
        ("cubic_energy_joules", ctypes.c_double),
        ("dmi_energy_joules", ctypes.c_double),
        ("total_energy_joules", ctypes.c_double),
        ("max_effective_field_amplitude", ctypes.c_double),
        ("max_demag_field_amplitude", ctypes.c_double),
        ("max_rhs_amplitude", ctypes.c_double),
        ("suggested_next_dt", ctypes.c_double),
        ("wall_time_ns", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("compute_capability_major", ctypes.c_int),
                ("compute_capability_minor", ctypes.c_int), ("driver_version", ctypes.c_int),
                ("runtime_version", ctypes.c_int)]


class SnapshotDesc(ctypes.Structure):
    _fields_ = [("cell_count", ctypes.c_uint64), ("component_count", ctypes.c_uint32),
                ("scalar_bytes", ctypes.c_uint32), ("scalar_type", ctypes.c_int)]


class PlanExt(ctypes.Structure):
    """fullmag_fdm_plan_ext (include/fullmag_fdm_ext.h)."""
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("ms_field", c_double_p),
        ("anisotropy_axis_field", c_double_p),
        ("field_waveform", ctypes.c_uint32),
        ("field_amplitude_am", ctypes.c_double * 3),
        ("field_t0", ctypes.c_double),
        ("field_t1", ctypes.c_double),
        ("field_frequency", ctypes.c_double),
        ("field_phase", ctypes.c_double),
        ("field_mask", c_u8_p),
        ("adaptive_rtol", ctypes.c_double),
        ("adaptive_growth_limit", ctypes.c_double),
        ("adaptive_shrink_limit", ctypes.c_double),
        ("thermal_dt", ctypes.c_double),
        ("thermal_seed", ctypes.c_uint64),
        ("periodic_axes", ctypes.c_uint32),
        ("periodic_images", ctypes.c_uint32 * 3),
        ("layer_z", c_double_p),
    ]


WAVE_NONE, WAVE_RAMP, WAVE_SINC, WAVE_SINE = 0, 1, 2, 3

POLL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)

ADJOINT_BETA_STATE_PRECISION = 1


class AdjointDesc(ctypes.Structure):
    """fullmag_fdm_adjoint_desc (include/fullmag_fdm_ext.h)."""
    _fields_ = [
        ("struct_size", ctypes.c_uint64),
        ("steps", ctypes.c_uint64),
        ("dt", ctypes.c_double),
        ("objective_first_step", ctypes.c_uint64),
        ("objective_component", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("objective_weight", c_double_p),
        ("objective_weight_len", ctypes.c_uint64),
        ("checkpoint_interval", ctypes.c_uint64),
        ("record_budget_bytes", ctypes.c_uint64),
    ]


# name -> (restype, argtypes); every symbol of include/fullmag_fdm.h first,
# then include/fullmag_fdm_ext.h.
_H = ctypes.c_void_p
_B = ctypes.c_void_p  # fullmag_fdm_batch *
REFERENCE_SYMBOLS = {
    "fullmag_fdm_is_available": (ctypes.c_int, []),
    "fullmag_fdm_backend_create": (_H, [ctypes.POINTER(PlanDesc)]),
    "fullmag_fdm_backend_step": (ctypes.c_int, [_H, ctypes.c_double, ctypes.POINTER(StepStats)]),
    "fullmag_fdm_backend_set_interrupt_poll": (ctypes.c_int, [_H, POLL_FN, ctypes.c_void_p]),
    "fullmag_fdm_backend_copy_field_f64": (ctypes.c_int, [_H, ctypes.c_int, c_double_p, ctypes.c_uint64]),
    "fullmag_fdm_backend_copy_field_f32": (ctypes.c_int, [_H, ctypes.c_int, c_float_p, ctypes.c_uint64]),
    "fullmag_fdm_backend_copy_field_preview_f64": (
        ctypes.c_int, [_H, ctypes.c_int] + [ctypes.c_uint32] * 5 + [c_double_p, ctypes.c_uint64]),
    "fullmag_fdm_backend_copy_field_preview_f32": (
        ctypes.c_int, [_H, ctypes.c_int] + [ctypes.c_uint32] * 5 + [c_float_p, ctypes.c_uint64]),
    "fullmag_fdm_backend_begin_field_snapshot": (_H, [_H, ctypes.c_int]),
    "fullmag_fdm_backend_begin_preview_snapshot": (_H, [_H, ctypes.c_int] + [ctypes.c_uint32] * 5),
    "fullmag_fdm_field_snapshot_wait": (
        ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64),
                       ctypes.POINTER(SnapshotDesc)]),
    "fullmag_fdm_preview_snapshot_wait": (
        ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64),
                       ctypes.POINTER(SnapshotDesc)]),
    "fullmag_fdm_field_snapshot_destroy": (None, [_H]),
    "fullmag_fdm_preview_snapshot_destroy": (None, [_H]),
    "fullmag_fdm_backend_upload_magnetization_f64": (ctypes.c_int, [_H, c_double_p, ctypes.c_uint64]),
    "fullmag_fdm_backend_upload_magnetization_f32": (ctypes.c_int, [_H, c_float_p, ctypes.c_uint64]),
    "fullmag_fdm_backend_refresh_observables": (ctypes.c_int, [_H]),
    "fullmag_fdm_backend_refresh_demag_observable": (ctypes.c_int, [_H]),
    "fullmag_fdm_backend_snapshot_stats": (ctypes.c_int, [_H, ctypes.POINTER(StepStats)]),
    "fullmag_fdm_backend_get_device_info": (ctypes.c_int, [_H, ctypes.POINTER(DeviceInfo)]),
    "fullmag_fdm_backend_last_error": (ctypes.c_char_p, [_H]),
    "fullmag_fdm_backend_destroy": (None, [_H]),
}
EXT_SYMBOLS = {
    "fullmag_fdm_ext_version": (ctypes.c_int, []),
    "fullmag_fdm_backend_advance": (
        ctypes.c_int, [_H, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(StepStats),
                       ctypes.POINTER(ctypes.c_uint64)]),
    "fullmag_fdm_backend_set_adaptive_ext": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "fullmag_fdm_backend_get_counters": (
        ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                       ctypes.POINTER(ctypes.c_double)]),
    "fullmag_fdm_backend_time_kernels": (ctypes.c_int, [_H, ctypes.c_int, c_double_p, ctypes.c_int]),
    "fullmag_fdm_backend_time_trial_kernels": (ctypes.c_int, [_H, ctypes.c_int, c_double_p, ctypes.c_int]),
    "fullmag_fdm_random_unit_vectors": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, c_double_p]),
    "fullmag_fdm_backend_create_ext": (_H, [ctypes.POINTER(PlanDesc), ctypes.POINTER(PlanExt)]),
    "fullmag_fdm_adjoint_forward": (ctypes.c_int, [_H, ctypes.POINTER(AdjointDesc), c_double_p]),
    "fullmag_fdm_adjoint_backward": (ctypes.c_int, [_H, c_double_p, c_double_p]),
    "fullmag_fdm_backend_autotune": (ctypes.c_int, [_H, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                    ctypes.POINTER(ctypes.c_int)]),
    "fullmag_fdm_batch_create": (_B, [ctypes.POINTER(ctypes.POINTER(PlanDesc)),
                                      ctypes.POINTER(ctypes.POINTER(PlanExt)), ctypes.c_uint64]),
    "fullmag_fdm_batch_groups": (ctypes.c_int, [_B, ctypes.POINTER(ctypes.c_uint64)]),
    "fullmag_fdm_batch_advance": (ctypes.c_int, [_B, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, c_double_p,
                                                 ctypes.POINTER(ctypes.c_uint64)]),
    "fullmag_fdm_batch_system": (_H, [_B, ctypes.c_uint64]),
    "fullmag_fdm_batch_last_error": (ctypes.c_char_p, [_B]),
    "fullmag_fdm_batch_destroy": (None, [_B]),
    "fullmag_fdm_backend_copy_field_preview_f64_noalloc": (
        ctypes.c_int, [_H, ctypes.c_int] + [ctypes.c_uint32] * 5 + [c_double_p, ctypes.c_uint64]),
}

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libfullmag_fdm.so")
_lib = None


def load(path: str | None = None):
    """Load the engine library (fails loudly if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("FULLMAG_FDM_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libfullmag_fdm.so not built at {p}; run `python -m fullmag_amd.build`")
    lib = ctypes.CDLL(p)
    for table in (REFERENCE_SYMBOLS, EXT_SYMBOLS):
        for name, (res, args) in table.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
