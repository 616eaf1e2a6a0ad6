"""Python host mirror of the reference's `NativeFdmBackend`
(crates/fullmag-runner/src/native_fdm.rs:49-970): create from a plan, check
`last_error`, step, read fields, upload magnetisation, destroy.

All arithmetic happens in libfullmag_fdm.so through the C ABI; this module
only marshals numpy arrays.  There is no CPU fallback: without the built
library and a GPU every call raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _abi as abi

MU0 = 4.0 * np.pi * 1e-7


class FdmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class Interrupted(FdmError):
    pass


@dataclass
class FdmPlan:
    """Plan in physical units; the field meanings are those of
    `fullmag_fdm_plan_desc` (native/include/fullmag_fdm.h:100-269)."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    ms: float
    a_ex: float
    alpha: float
    m0: np.ndarray                      # (n,3) float64
    gamma: float = 2.211e5
    precision: str = "double"           # "single" | "double"
    integrator: str = "dp45"            # "heun" | "dp45" | "rk4" | "rk23" | "abm3"
    exchange: bool = True
    demag: bool = True
    precession: bool = True
    h_ext: tuple | None = None          # A/m
    ku1: float = 0.0
    ku2: float = 0.0
    axis: tuple | None = None           # enables uniaxial anisotropy
    mask: np.ndarray | None = None      # uint8 (n,)
    regions: np.ndarray | None = None   # uint32 (n,)
    exchange_lut: np.ndarray | None = None  # (256*256,) float64
    spectra: dict | None = None         # {"xx": interleaved complex f64, ...}
    spectra_len: int | None = None      # demag_kernel_spectrum_len (None: the arrays' length)
    atol: float = 0.0                   # 0 -> engine default 1e-5
    dt_min: float = 0.0
    dt_max: float = 0.0
    headroom: float = 0.0
    ku1_field: np.ndarray | None = None      # per-cell Ku1 (reference ABI ku1_field)
    ku2_field: np.ndarray | None = None
    kc1_field: np.ndarray | None = None      # per-cell cubic Kc1..Kc3 (reference ABI kc*_field)
    kc2_field: np.ndarray | None = None
    kc3_field: np.ndarray | None = None
    temperature: float = 0.0                 # K; thermal field (fdm_fields.rs:818-905)
    # boundary-correction tier (fullmag_fdm.h:222-258): 0 none, 1 volume (T0),
    # 2 full (T1); face_links / deltas: six per-cell arrays in the order
    # +x, -x, +y, -y, +z, -z; demag_corr: (targets, sources (count, stencil),
    # tensors (count, stencil, 6) as xx xy xz yy yz zz)
    boundary_tier: int = 0
    phi_floor: float = 0.0
    delta_min: float = 0.0
    volume_fraction: np.ndarray | None = None
    face_links: tuple | None = None
    deltas: tuple | None = None
    demag_corr: tuple | None = None
    # plan extension (include/fullmag_fdm_ext.h)
    ms_field: np.ndarray | None = None       # per-cell Ms
    axis_field: np.ndarray | None = None     # (n,3) per-cell uniaxial axis
    wave: int = 0                            # abi.WAVE_*
    wave_amp: tuple = (0.0, 0.0, 0.0)
    wave_t0: float = 0.0
    wave_t1: float = 0.0
    wave_freq: float = 0.0
    wave_phase: float = 0.0
    wave_mask: np.ndarray | None = None
    rtol: float = 0.0
    growth: float = 0.0
    shrink: float = 0.0
    thermal_dt: float = 0.0                  # noise dt; 0 = the dt of each step attempt
    thermal_seed: int = 0                    # counter-generator seed; 0 = 42
    periodic: tuple = (False, False, False)  # periodic x, y, z (fdm_types.rs:94-132)
    periodic_images: tuple = (0, 0, 0)       # images per periodic axis; 0 = 10
    layer_z: np.ndarray | None = None        # multilayer: z centre of each plane (m)
    extra: dict = field(default_factory=dict)  # raw PlanDesc fields (for error tests)

    def uses_ext(self) -> bool:
        return (self.ms_field is not None or self.axis_field is not None or self.wave != 0
                or self.rtol > 0 or self.growth > 0 or self.shrink > 0 or self.thermal_dt > 0
                or self.thermal_seed != 0 or any(self.periodic) or self.layer_z is not None)


_INTEGRATORS = {"heun": abi.INTEGRATOR_HEUN, "dp45": abi.INTEGRATOR_DP45, "rk45": abi.INTEGRATOR_DP45,
                "rk4": abi.INTEGRATOR_RK4, "abm3": abi.INTEGRATOR_ABM3, "rk23": abi.INTEGRATOR_RK23}
_OBS = {"m": abi.OBSERVABLE_M, "h_ex": abi.OBSERVABLE_H_EX, "h_demag": abi.OBSERVABLE_H_DEMAG,
        "h_ext": abi.OBSERVABLE_H_EXT, "h_eff": abi.OBSERVABLE_H_EFF}


def _dp(a):
    return a.ctypes.data_as(abi.c_double_p)


def is_available() -> bool:
    return bool(abi.load().fullmag_fdm_is_available())


def marshal_plan(plan: FdmPlan):
    """PlanDesc (and PlanExt when the plan uses extensions) for `plan`, plus
    the numpy arrays they point into (keep them alive while the library
    reads the plan)."""
    keep = []
    desc = abi.PlanDesc()
    desc.grid = abi.GridDesc(plan.nx, plan.ny, plan.nz, plan.dx, plan.dy, plan.dz)
    desc.material = abi.MaterialDesc(plan.ms, plan.a_ex, plan.alpha, plan.gamma)
    desc.precision = abi.PRECISION_DOUBLE if plan.precision == "double" else abi.PRECISION_SINGLE
    desc.integrator = _INTEGRATORS[plan.integrator]
    desc.disable_precession = 0 if plan.precession else 1
    desc.enable_exchange = int(plan.exchange)
    desc.enable_demag = int(plan.demag)
    if plan.h_ext is not None:
        desc.has_external_field = 1
        desc.external_field_am[:] = list(plan.h_ext)
    if plan.axis is not None:
        desc.has_uniaxial_anisotropy = 1
        desc.uniaxial_anisotropy_constant = plan.ku1
        desc.uniaxial_anisotropy_k2 = plan.ku2
        desc.anisotropy_axis[:] = list(plan.axis)
    m0 = np.ascontiguousarray(np.asarray(plan.m0, dtype=np.float64).reshape(-1))
    keep.append(m0)
    desc.initial_magnetization_xyz = _dp(m0)
    desc.initial_magnetization_len = m0.size
    if plan.mask is not None:
        mk = np.ascontiguousarray(np.asarray(plan.mask, dtype=np.uint8).reshape(-1))
        keep.append(mk)
        desc.active_mask = mk.ctypes.data_as(abi.c_u8_p)
        desc.active_mask_len = mk.size
    if plan.regions is not None:
        rg = np.ascontiguousarray(np.asarray(plan.regions, dtype=np.uint32).reshape(-1))
        keep.append(rg)
        desc.region_mask = rg.ctypes.data_as(abi.c_u32_p)
        desc.region_mask_len = rg.size
    if plan.exchange_lut is not None:
        lt = np.ascontiguousarray(np.asarray(plan.exchange_lut, dtype=np.float64).reshape(-1))
        keep.append(lt)
        desc.exchange_lut = _dp(lt)
        desc.exchange_lut_len = lt.size
    if plan.spectra is not None:
        ln = None
        for nm in ("xx", "yy", "zz", "xy", "xz", "yz"):
            if nm in plan.spectra:
                a = np.ascontiguousarray(np.asarray(plan.spectra[nm], dtype=np.float64))
                keep.append(a)
                setattr(desc, f"demag_kernel_{nm}_spectrum", _dp(a))
                ln = a.size
        desc.demag_kernel_spectrum_len = ln if plan.spectra_len is None else plan.spectra_len
    for nm in ("ku1_field", "ku2_field", "kc1_field", "kc2_field", "kc3_field"):
        a = getattr(plan, nm)
        if a is not None:
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
            keep.append(a)
            setattr(desc, nm, _dp(a))
    desc.adaptive_max_error = plan.atol
    desc.adaptive_dt_min = plan.dt_min
    desc.adaptive_dt_max = plan.dt_max
    desc.adaptive_headroom = plan.headroom
    desc.temperature = plan.temperature
    desc.boundary_correction = plan.boundary_tier
    desc.boundary_phi_floor, desc.boundary_delta_min = plan.phi_floor, plan.delta_min
    if plan.volume_fraction is not None:
        a = np.ascontiguousarray(np.asarray(plan.volume_fraction, dtype=np.float64).reshape(-1))
        keep.append(a)
        desc.volume_fraction, desc.volume_fraction_len = _dp(a), a.size
    for prefix, arrays in (("face_link_", plan.face_links), ("delta_", plan.deltas)):
        if arrays is not None:
            for suffix, arr in zip(("xp", "xm", "yp", "ym", "zp", "zm"), arrays):
                a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1))
                keep.append(a)
                setattr(desc, prefix + suffix, _dp(a))
    if plan.demag_corr is not None:
        tg, src, ten = plan.demag_corr
        tg = np.ascontiguousarray(np.asarray(tg, dtype=np.int32).reshape(-1))
        src = np.ascontiguousarray(np.asarray(src, dtype=np.int32).reshape(tg.size, -1))
        ten = np.ascontiguousarray(np.asarray(ten, dtype=np.float64).reshape(-1))
        keep += [tg, src, ten]
        desc.has_demag_boundary_corr = 1
        desc.demag_corr_target_idx = tg.ctypes.data_as(abi.c_i32_p)
        desc.demag_corr_source_idx = src.ctypes.data_as(abi.c_i32_p)
        desc.demag_corr_tensor = _dp(ten)
        desc.demag_corr_target_count, desc.demag_corr_stencil_size = tg.size, src.shape[1]
    for k, v in plan.extra.items():
        setattr(desc, k, v)
    ext = None
    if plan.uses_ext():
        ext = abi.PlanExt()
        ext.struct_size = ctypes.sizeof(abi.PlanExt)
        if plan.ms_field is not None:
            a = np.ascontiguousarray(np.asarray(plan.ms_field, dtype=np.float64).reshape(-1))
            keep.append(a)
            ext.ms_field = _dp(a)
        if plan.axis_field is not None:
            a = np.ascontiguousarray(np.asarray(plan.axis_field, dtype=np.float64).reshape(-1))
            keep.append(a)
            ext.anisotropy_axis_field = _dp(a)
        ext.field_waveform = plan.wave
        ext.field_amplitude_am[:] = list(plan.wave_amp)
        ext.field_t0, ext.field_t1 = plan.wave_t0, plan.wave_t1
        ext.field_frequency, ext.field_phase = plan.wave_freq, plan.wave_phase
        if plan.wave_mask is not None:
            wm = np.ascontiguousarray(np.asarray(plan.wave_mask, dtype=np.uint8).reshape(-1))
            keep.append(wm)
            ext.field_mask = wm.ctypes.data_as(abi.c_u8_p)
        ext.adaptive_rtol, ext.adaptive_growth_limit = plan.rtol, plan.growth
        ext.adaptive_shrink_limit = plan.shrink
        ext.thermal_dt, ext.thermal_seed = plan.thermal_dt, plan.thermal_seed
        ext.periodic_axes = sum(1 << a for a in range(3) if plan.periodic[a])
        ext.periodic_images[:] = [int(v) for v in plan.periodic_images]
        if plan.layer_z is not None:
            lz = np.ascontiguousarray(np.asarray(plan.layer_z, dtype=np.float64).reshape(-1))
            keep.append(lz)
            ext.layer_z = _dp(lz)
    return desc, ext, keep


class FdmBackend:
    """One engine handle (one host thread drives it)."""

    def __init__(self, plan: FdmPlan, lib_path: str | None = None):
        self.lib = abi.load(lib_path)
        self.plan = plan
        self.n = plan.nx * plan.ny * plan.nz
        desc, ext, self._keep = marshal_plan(plan)
        self._desc, self._ext = desc, ext
        if ext is not None:
            self.h = self.lib.fullmag_fdm_backend_create_ext(ctypes.byref(desc), ctypes.byref(ext))
        else:
            self.h = self.lib.fullmag_fdm_backend_create(ctypes.byref(desc))
        if not self.h:
            raise FdmError(abi.ERR_INTERNAL, "fullmag_fdm_backend_create returned NULL")
        err = self.lib.fullmag_fdm_backend_last_error(self.h)
        if err:
            msg = err.decode()
            self.lib.fullmag_fdm_backend_destroy(self.h)
            self.h = None
            raise FdmError(abi.ERR_INVALID, msg)
        self._poll = None

    # ---------------------------------------------------------------- core
    def _check(self, rc):
        if rc == abi.OK:
            return
        err = self.lib.fullmag_fdm_backend_last_error(self.h)
        msg = err.decode() if err else "error"
        if rc == abi.ERR_INTERRUPTED:
            raise Interrupted(rc, msg)
        raise FdmError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.lib.fullmag_fdm_backend_destroy(self.h)
            self.h = None

    @classmethod
    def _view(cls, lib, handle, plan: FdmPlan):
        """Non-owning wrapper of a handle that a batch owns."""
        obj = cls.__new__(cls)
        obj.lib, obj.h, obj.plan = lib, handle, plan
        obj.n = plan.nx * plan.ny * plan.nz
        obj._keep, obj._poll, obj._owned = [], None, False
        return obj

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, dt: float) -> dict:
        st = abi.StepStats()
        self._check(self.lib.fullmag_fdm_backend_step(self.h, dt, ctypes.byref(st)))
        return st.as_dict()

    def advance(self, t_end: float, max_steps: int, dt: float) -> dict:
        st = abi.StepStats()
        n = ctypes.c_uint64(0)
        self._check(self.lib.fullmag_fdm_backend_advance(self.h, t_end, max_steps, dt, ctypes.byref(st),
                                                         ctypes.byref(n)))
        d = st.as_dict()
        d["steps_taken"] = int(n.value)
        return d

    def adjoint_forward(self, steps: int, dt: float, weight: np.ndarray, first_step: int = 1, component: int = 2,
                        checkpoint_interval: int = 0, beta_state_precision: bool = False,
                        record_budget_bytes: int = 0) -> float:
        """Fixed-step run of `steps` steps with checkpoints; returns
        J = sum_{n>=first_step} sum_i w_i m_{c,i}(t_n)^2 / (steps - first_step + 1).
        beta_state_precision: an FP32 plan convolves the field cotangent in
        FP32 instead of FP64.  record_budget_bytes: device memory for the
        forward's record of the trailing checkpoint segments (0: half the
        free device memory; 1: the last segment only)."""
        w = np.ascontiguousarray(weight, dtype=np.float64).reshape(-1)
        d = abi.AdjointDesc()
        d.struct_size = ctypes.sizeof(abi.AdjointDesc)
        d.steps = steps
        d.dt = dt
        d.objective_first_step = first_step
        d.objective_component = component
        d.objective_weight = _dp(w)
        d.objective_weight_len = w.size
        d.checkpoint_interval = checkpoint_interval
        d.flags = abi.ADJOINT_BETA_STATE_PRECISION if beta_state_precision else 0
        d.record_budget_bytes = record_budget_bytes
        j = ctypes.c_double(0.0)
        self._check(self.lib.fullmag_fdm_adjoint_forward(self.h, ctypes.byref(d), ctypes.byref(j)))
        return j.value

    def adjoint_backward(self):
        """Reverse sweep after adjoint_forward: (dJ/dMs per cell (n,),
        dJ/d waveform amplitude in A/m (3,))."""
        g = np.zeros(self.n)
        ga = np.zeros(3)
        self._check(self.lib.fullmag_fdm_adjoint_backward(self.h, _dp(g), _dp(ga)))
        return g, ga

    def autotune(self, reps: int = 10) -> dict:
        """Times the x tiles and keeps the fastest (fullmag_fdm_backend_autotune)."""
        rows, fs = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.fullmag_fdm_backend_autotune(self.h, reps, ctypes.byref(rows), ctypes.byref(fs)))
        return {"rows_per_group": rows.value, "four_step_x": bool(fs.value)}

    def set_adaptive_ext(self, rtol: float, growth: float = 0.0, shrink: float = 0.0):
        self._check(self.lib.fullmag_fdm_backend_set_adaptive_ext(self.h, rtol, growth, shrink))

    def counters(self):
        a, r, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        self._check(self.lib.fullmag_fdm_backend_get_counters(self.h, ctypes.byref(a), ctypes.byref(r),
                                                              ctypes.byref(t)))
        return {"accepted": int(a.value), "rejected": int(r.value), "time": t.value}

    def set_interrupt_poll(self, fn):
        """fn() -> bool; called on the stepping thread between stages."""
        if fn is None:
            self._poll = None
            self._check(self.lib.fullmag_fdm_backend_set_interrupt_poll(self.h, abi.POLL_FN(), None))
            return
        self._poll = abi.POLL_FN(lambda _u: 1 if fn() else 0)
        self._check(self.lib.fullmag_fdm_backend_set_interrupt_poll(self.h, self._poll, None))

    def field(self, name: str = "m", dtype=np.float64) -> np.ndarray:
        if dtype == np.float32:
            out = np.zeros((self.n, 3), dtype=np.float32)
            self._check(self.lib.fullmag_fdm_backend_copy_field_f32(
                self.h, _OBS[name], out.ctypes.data_as(abi.c_float_p), out.size))
        else:
            out = np.zeros((self.n, 3))
            self._check(self.lib.fullmag_fdm_backend_copy_field_f64(self.h, _OBS[name], _dp(out), out.size))
        return out

    def preview(self, name, pnx, pny, pnz, z_origin=0, z_stride=1, dtype=np.float64):
        cnt = pnx * pny * pnz
        if dtype == np.float32:
            out = np.zeros((cnt, 3), dtype=np.float32)
            self._check(self.lib.fullmag_fdm_backend_copy_field_preview_f32(
                self.h, _OBS[name], pnx, pny, pnz, z_origin, z_stride, out.ctypes.data_as(abi.c_float_p), out.size))
        else:
            out = np.zeros((cnt, 3))
            self._check(self.lib.fullmag_fdm_backend_copy_field_preview_f64(
                self.h, _OBS[name], pnx, pny, pnz, z_origin, z_stride, _dp(out), out.size))
        return out

    def field_snapshot(self, name: str = "m") -> np.ndarray:
        """Asynchronous snapshot API; returns the SoA payload as (3, n)."""
        s = self.lib.fullmag_fdm_backend_begin_field_snapshot(self.h, _OBS[name])
        if not s:
            self._check(abi.ERR_INTERNAL)
        try:
            p, ln, d = ctypes.c_void_p(), ctypes.c_uint64(), abi.SnapshotDesc()
            rc = self.lib.fullmag_fdm_field_snapshot_wait(s, ctypes.byref(p), ctypes.byref(ln), ctypes.byref(d))
            if rc != abi.OK:
                raise FdmError(rc, "snapshot wait failed")
            dt = np.float64 if d.scalar_type == abi.SNAPSHOT_F64 else np.float32
            buf = (ctypes.c_char * ln.value).from_address(p.value)
            return np.frombuffer(bytes(buf), dtype=dt).reshape(3, d.cell_count).copy()
        finally:
            self.lib.fullmag_fdm_field_snapshot_destroy(s)

    def preview_snapshot(self, name, pnx, pny, pnz, z_origin=0, z_stride=1) -> np.ndarray:
        """Asynchronous preview API; returns the AoS payload as (cells, 3) in
        the state precision."""
        s = self.lib.fullmag_fdm_backend_begin_preview_snapshot(self.h, _OBS[name], pnx, pny, pnz, z_origin,
                                                                z_stride)
        if not s:
            self._check(abi.ERR_INTERNAL)
        try:
            p, ln, d = ctypes.c_void_p(), ctypes.c_uint64(), abi.SnapshotDesc()
            rc = self.lib.fullmag_fdm_preview_snapshot_wait(s, ctypes.byref(p), ctypes.byref(ln), ctypes.byref(d))
            if rc != abi.OK:
                raise FdmError(rc, "preview snapshot wait failed")
            dt = np.float64 if d.scalar_type == abi.SNAPSHOT_F64 else np.float32
            buf = (ctypes.c_char * ln.value).from_address(p.value)
            return np.frombuffer(bytes(buf), dtype=dt).reshape(d.cell_count, 3).copy()
        finally:
            self.lib.fullmag_fdm_preview_snapshot_destroy(s)

    def upload_m(self, m: np.ndarray):
        a = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(-1))
        self._check(self.lib.fullmag_fdm_backend_upload_magnetization_f64(self.h, _dp(a), a.size))

    def refresh(self):
        self._check(self.lib.fullmag_fdm_backend_refresh_observables(self.h))

    def stats(self) -> dict:
        st = abi.StepStats()
        self._check(self.lib.fullmag_fdm_backend_snapshot_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def device_info(self) -> dict:
        d = abi.DeviceInfo()
        self._check(self.lib.fullmag_fdm_backend_get_device_info(self.h, ctypes.byref(d)))
        return {"name": d.name.decode(), "major": d.compute_capability_major,
                "minor": d.compute_capability_minor, "driver": d.driver_version, "runtime": d.runtime_version}

    def time_kernels(self, reps: int = 20) -> list:
        out = np.zeros(8)
        self._check(self.lib.fullmag_fdm_backend_time_kernels(self.h, reps, _dp(out), 8))
        return out[:7].tolist()

    def time_trial_kernels(self, trials: int = 20) -> list:
        """Mean µs of the first launch of each kernel kind inside whole trial
        steps (each kernel reading what the one before it wrote; stage 2,
        the final x pass at the last stage), less the per-launch cost of the
        timing events; [7] whole trial step, [8] that event cost.  State
        unchanged."""
        out = np.zeros(9)
        self._check(self.lib.fullmag_fdm_backend_time_trial_kernels(self.h, trials, _dp(out), 9))
        return out.tolist()


# ------------------------------------------------------------ input builders
def random_unit_vectors(seed: int, count: int) -> np.ndarray:
    """xorshift64 random unit vectors of the planner,
    crates/fullmag-plan/src/util.rs:118-147 (seed 0 never terminates there
    and is refused here); generated by the library's host code."""
    if seed == 0:
        raise ValueError("seed 0 is not allowed")
    out = np.empty((count, 3))
    rc = abi.load().fullmag_fdm_random_unit_vectors(seed, count, _dp(out))
    if rc != abi.OK:
        raise FdmError(rc, "random_unit_vectors failed")
    return out


def hole_mask(nx: int, ny: int, nz: int, dx: float, dy: float, radius: float) -> np.ndarray:
    """Box minus a centred cylinder by cell centres,
    crates/fullmag-plan/src/geometry.rs:159-231."""
    x = (np.arange(nx) + 0.5) * dx - nx * dx * 0.5
    y = (np.arange(ny) + 0.5) * dy - ny * dy * 0.5
    r2 = x[None, :] ** 2 + y[:, None] ** 2
    plane = (r2 > radius * radius).astype(np.uint8)
    return np.ascontiguousarray(np.broadcast_to(plane, (nz, ny, nx)).reshape(-1))


class FdmBatch:
    """Batched sweep of independent systems (fullmag_fdm_batch_*): systems
    with the same launch geometry share every kernel launch."""

    def __init__(self, plans: list, lib_path: str | None = None):
        self.lib = abi.load(lib_path)
        self.plans = list(plans)
        marshalled = [marshal_plan(p) for p in self.plans]
        self._keep = marshalled
        n = len(self.plans)
        descs = (ctypes.POINTER(abi.PlanDesc) * n)(*[ctypes.pointer(d) for d, _, _ in marshalled])
        exts = (ctypes.POINTER(abi.PlanExt) * n)(*[ctypes.pointer(e) if e is not None else None
                                                    for _, e, _ in marshalled])
        self._arrays = (descs, exts)
        self.b = self.lib.fullmag_fdm_batch_create(descs, exts, n)
        if not self.b:
            raise FdmError(abi.ERR_INTERNAL, "fullmag_fdm_batch_create returned NULL")
        err = self.lib.fullmag_fdm_batch_last_error(self.b)
        if err:
            msg = err.decode()
            self.lib.fullmag_fdm_batch_destroy(self.b)
            self.b = None
            raise FdmError(abi.ERR_INVALID, msg)
        self.systems = [FdmBackend._view(self.lib, self.lib.fullmag_fdm_batch_system(self.b, i), p)
                        for i, p in enumerate(self.plans)]

    def groups(self) -> int:
        g = ctypes.c_uint64(0)
        if self.lib.fullmag_fdm_batch_groups(self.b, ctypes.byref(g)) != 0:
            raise FdmError(abi.ERR_INVALID, "batch_groups failed")
        return int(g.value)

    def advance(self, t_end: float, max_steps: int, dt0: float):
        """Every system to t_end with its own dt control; returns (wall
        seconds after which each system was seen finished, accepted steps)."""
        n = len(self.plans)
        wall = np.zeros(n)
        steps = np.zeros(n, dtype=np.uint64)
        rc = self.lib.fullmag_fdm_batch_advance(self.b, t_end, max_steps, dt0, _dp(wall),
                                                steps.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if rc != 0:
            err = self.lib.fullmag_fdm_batch_last_error(self.b)
            raise FdmError(rc, err.decode() if err else "batch advance failed")
        return wall, steps.astype(np.int64)

    def close(self):
        if getattr(self, "b", None):
            for s_ in self.systems:
                s_.h = None
            self.lib.fullmag_fdm_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
