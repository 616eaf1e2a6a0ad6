"""ctypes wrapper of the FP64 CPU oracle (oracle/fdm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
The oracle restates the reference CPU engine `fdm_cpu_reference`
(crates/fullmag-engine, crates/fullmag-fdm-demag, crates/fullmag-runner);
see the header of fdm_oracle.cpp for the file:line map.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfdm_oracle.so")
_lib = None

MU0 = 4.0 * np.pi * 1e-7


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class _Desc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("ms", ctypes.c_double), ("a_ex", ctypes.c_double), ("alpha", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("precession", ctypes.c_int), ("exchange", ctypes.c_int), ("demag", ctypes.c_int),
        ("has_ext", ctypes.c_int), ("h_ext", ctypes.c_double * 3),
        ("has_uni", ctypes.c_int), ("ku1", ctypes.c_double), ("ku2", ctypes.c_double),
        ("axis", ctypes.c_double * 3),
        ("mask", ctypes.POINTER(ctypes.c_uint8)),
        ("atol", ctypes.c_double), ("rtol", ctypes.c_double), ("dt_min", ctypes.c_double),
        ("dt_max", ctypes.c_double), ("headroom", ctypes.c_double), ("growth", ctypes.c_double),
        ("shrink", ctypes.c_double),
        ("ms_cell", ctypes.POINTER(ctypes.c_double)), ("ku1_cell", ctypes.POINTER(ctypes.c_double)),
        ("ku2_cell", ctypes.POINTER(ctypes.c_double)), ("axis_cell", ctypes.POINTER(ctypes.c_double)),
        ("wave_kind", ctypes.c_int), ("wave_amp", ctypes.c_double * 3),
        ("wave_t0", ctypes.c_double), ("wave_t1", ctypes.c_double), ("wave_freq", ctypes.c_double),
        ("wave_phase", ctypes.c_double), ("wave_mask", ctypes.POINTER(ctypes.c_uint8)),
        ("has_cub", ctypes.c_int), ("kc1", ctypes.c_double), ("kc2", ctypes.c_double), ("kc3", ctypes.c_double),
        ("cub_axis1", ctypes.c_double * 3), ("cub_axis2", ctypes.c_double * 3),
        ("kc1_cell", ctypes.POINTER(ctypes.c_double)), ("kc2_cell", ctypes.POINTER(ctypes.c_double)),
        ("kc3_cell", ctypes.POINTER(ctypes.c_double)),
        ("dmi_i", ctypes.c_double), ("dmi_b", ctypes.c_double),
        ("has_mel", ctypes.c_int), ("mel_b1", ctypes.c_double), ("mel_b2", ctypes.c_double),
        ("mel_strain", ctypes.c_double * 6),
        ("j", ctypes.c_double * 3), ("stt_degree", ctypes.c_double), ("stt_beta", ctypes.c_double),
        ("slon_p", ctypes.c_double * 3), ("slon_lambda", ctypes.c_double), ("slon_eps", ctypes.c_double),
        ("has_sot", ctypes.c_int), ("sot_je", ctypes.c_double), ("sot_xi_dl", ctypes.c_double),
        ("sot_xi_fl", ctypes.c_double), ("sot_sigma", ctypes.c_double * 3), ("sot_thickness", ctypes.c_double),
        ("has_oe", ctypes.c_int), ("oe_current", ctypes.c_double), ("oe_radius", ctypes.c_double),
        ("oe_center", ctypes.c_double * 3), ("oe_axis", ctypes.c_double * 3), ("oe_kind", ctypes.c_int),
        ("oe_freq", ctypes.c_double), ("oe_phase", ctypes.c_double), ("oe_offset", ctypes.c_double),
        ("oe_t_on", ctypes.c_double), ("oe_t_off", ctypes.c_double),
        ("temperature", ctypes.c_double), ("thermal_dt", ctypes.c_double), ("thermal_seed", ctypes.c_uint64),
        ("bc_tier", ctypes.c_int), ("bc_phi_floor", ctypes.c_double), ("bc_delta_min", ctypes.c_double),
        ("bc_phi", ctypes.POINTER(ctypes.c_double)), ("bc_face", ctypes.POINTER(ctypes.c_double) * 6),
        ("bc_delta", ctypes.POINTER(ctypes.c_double) * 6),
        ("dc_count", ctypes.c_uint32), ("dc_stencil", ctypes.c_uint32),
        ("dc_target", ctypes.POINTER(ctypes.c_int32)), ("dc_source", ctypes.POINTER(ctypes.c_int32)),
        ("dc_tensor", ctypes.POINTER(ctypes.c_double)),
        ("periodic", ctypes.c_uint32), ("images", ctypes.c_uint32 * 3),
        ("layer_z", ctypes.POINTER(ctypes.c_double)),
        ("region", ctypes.POINTER(ctypes.c_uint32)), ("exchange_lut", ctypes.POINTER(ctypes.c_double)),
    ]


class Report(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("dt_used", ctypes.c_double), ("dt_next", ctypes.c_double),
        ("e_ex", ctypes.c_double), ("e_demag", ctypes.c_double), ("e_ext", ctypes.c_double),
        ("e_ani", ctypes.c_double), ("e_total", ctypes.c_double),
        ("max_heff", ctypes.c_double), ("max_hdemag", ctypes.c_double), ("max_rhs", ctypes.c_double),
        ("rejected", ctypes.c_uint64),
        ("e_cub", ctypes.c_double), ("e_dmi", ctypes.c_double), ("e_mel", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_newell_f.restype = ctypes.c_double
        L.orc_newell_f.argtypes = [ctypes.c_double] * 3
        L.orc_newell_g.restype = ctypes.c_double
        L.orc_newell_g.argtypes = [ctypes.c_double] * 3
        L.orc_newell_kernels.argtypes = [ctypes.c_uint32] * 3 + [ctypes.c_double] * 3 + [_dp]
        L.orc_fft.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        L.orc_random_unit_vectors.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _dp]
        L.orc_xorshift64_states.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.orc_state_component.restype = ctypes.c_double
        L.orc_state_component.argtypes = [ctypes.c_uint64]
        L.orc_hole_mask.restype = ctypes.c_uint64
        L.orc_hole_mask.argtypes = [ctypes.c_uint32] * 3 + [ctypes.c_double] * 3 + [ctypes.POINTER(ctypes.c_uint8)]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(_Desc)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_kernel_spectrum.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_fields.argtypes = [ctypes.c_void_p] + [_dp] * 6
        L.orc_demag_direct.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_observables.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(Report)]
        L.orc_prepare_state.argtypes = [ctypes.c_void_p, _dp]
        L.orc_invalidate_fsal.argtypes = [ctypes.c_void_p]
        L.orc_reset_abm.argtypes = [ctypes.c_void_p]
        L.orc_abm_count.argtypes = [ctypes.c_void_p]
        L.orc_abm_count.restype = ctypes.c_int
        L.orc_thermal_step.argtypes = [ctypes.c_void_p]
        L.orc_thermal_step.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
        for nm in ("orc_heun_step", "orc_rk4_step", "orc_rk45_step", "orc_rk23_step", "orc_abm3_step"):
            getattr(L, nm).argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.POINTER(Report)]
        L.orc_run.restype = ctypes.c_uint64
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double,
                              ctypes.c_uint64, ctypes.POINTER(Report), _dp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def newell_f(x, y, z):
    return lib().orc_newell_f(x, y, z)


def newell_g(x, y, z):
    return lib().orc_newell_g(x, y, z)


def newell_kernels(nx, ny, nz, dx, dy, dz):
    """Real-space padded tensor; dict of 6 arrays shaped (2nz, 2ny, 2nx)."""
    P = 8 * nx * ny * nz
    out = np.zeros(6 * P)
    assert lib().orc_newell_kernels(nx, ny, nz, dx, dy, dz, _ptr(out)) == 0
    shp = (2 * nz, 2 * ny, 2 * nx)
    names = ("xx", "yy", "zz", "xy", "xz", "yz")
    return {nm: out[i * P:(i + 1) * P].reshape(shp) for i, nm in enumerate(names)}


def fft(x, inverse=False):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    assert lib().orc_fft(a.view(np.float64).ctypes.data_as(_dp), a.size, int(inverse)) == 0
    return a


def random_unit_vectors(seed, count):
    out = np.zeros((count, 3))
    assert lib().orc_random_unit_vectors(seed, count, _ptr(out)) == 0
    return out


def xorshift64_states(seed, n):
    """The first n states of the planner's xorshift64 after `seed`
    (util.rs:128-130)."""
    out = np.zeros(n, dtype=np.uint64)
    assert lib().orc_xorshift64_states(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))) == 0
    return [int(v) for v in out]


def state_component(state):
    """util.rs:133: the [-1, 1] component of one state."""
    return lib().orc_state_component(state)


def hole_mask(nx, ny, nz, dx, dy, radius):
    m = np.zeros(nx * ny * nz, dtype=np.uint8)
    lib().orc_hole_mask(nx, ny, nz, dx, dy, radius, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return m


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


@dataclass
class Problem:
    """Problem description; same physics fields as the ABI plan descriptor."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    ms: float
    a_ex: float
    alpha: float
    gamma: float = 2.211e5
    precession: bool = True
    exchange: bool = True
    demag: bool = True
    h_ext: tuple | None = None
    ku1: float = 0.0
    ku2: float = 0.0
    axis: tuple | None = None
    mask: np.ndarray | None = None
    atol: float = 1e-5
    rtol: float = 0.0
    dt_min: float = 1e-18
    dt_max: float = 1e-10
    headroom: float = 0.8
    growth: float = float("inf")
    shrink: float = 0.0
    # extensions (no reference counterpart): per-cell maps and a Zeeman waveform
    ms_cell: np.ndarray | None = None
    ku1_cell: np.ndarray | None = None
    ku2_cell: np.ndarray | None = None
    axis_cell: np.ndarray | None = None
    wave_kind: int = 0
    wave_amp: tuple = (0.0, 0.0, 0.0)
    wave_t0: float = 0.0
    wave_t1: float = 0.0
    wave_freq: float = 0.0
    wave_phase: float = 0.0
    wave_mask: np.ndarray | None = None
    # further terms of the plan descriptor (fullmag_fdm.h:117-178)
    cubic: tuple | None = None          # (Kc1, Kc2, Kc3, axis1, axis2)
    kc1_cell: np.ndarray | None = None
    kc2_cell: np.ndarray | None = None
    kc3_cell: np.ndarray | None = None
    dmi_i: float = 0.0
    dmi_b: float = 0.0
    mel: tuple | None = None            # (B1, B2, Voigt strain[6])
    j: tuple = (0.0, 0.0, 0.0)          # current density (A/m^2)
    stt_degree: float = 0.0
    stt_beta: float = 0.0
    slon_p: tuple = (0.0, 0.0, 0.0)
    slon_lambda: float = 0.0
    slon_eps: float = 0.0
    sot: tuple | None = None            # (je, xi_dl, xi_fl, sigma[3], thickness)
    oersted: tuple | None = None        # (current, radius, center[3], axis[3], kind, freq, phase, offset, t_on, t_off)
    temperature: float = 0.0            # K
    thermal_dt: float = 0.0             # s; 0 = the dt of each step
    thermal_seed: int = 0               # 0 = 42
    # boundary-correction tier (fullmag_fdm.h:222-258): 0 none, 1 T0, 2 T1;
    # volume fraction per cell, face links / distances as six arrays in the
    # order +x, -x, +y, -y, +z, -z, and the sparse demag correction as
    # (targets, sources (count, stencil), tensors (count, stencil, 6))
    boundary_tier: int = 0
    phi_floor: float = 0.0
    delta_min: float = 0.0
    volume_fraction: np.ndarray | None = None
    face_links: tuple | None = None
    deltas: tuple | None = None
    demag_corr: tuple | None = None
    # periodic boundaries (fdm_types.rs:94-132) and demag images per axis
    # (0 = the CPU engine's default 10)
    periodic: tuple = (False, False, False)
    periodic_images: tuple = (0, 0, 0)
    # multilayer: z centre of each plane (single-cell layers coupled by demag)
    layer_z: np.ndarray | None = None
    # exchange region table (exchange_fp64.cu:80-114): region id per cell and
    # the 256 x 256 stiffness table (both or neither)
    regions: np.ndarray | None = None
    exchange_lut: np.ndarray | None = None


class Oracle:
    def __init__(self, pb: Problem):
        L = lib()
        d = _Desc()
        d.nx, d.ny, d.nz = pb.nx, pb.ny, pb.nz
        d.dx, d.dy, d.dz = pb.dx, pb.dy, pb.dz
        d.ms, d.a_ex, d.alpha, d.gamma = pb.ms, pb.a_ex, pb.alpha, pb.gamma
        d.precession, d.exchange, d.demag = int(pb.precession), int(pb.exchange), int(pb.demag)
        d.has_ext = int(pb.h_ext is not None)
        if pb.h_ext is not None:
            d.h_ext[:] = list(pb.h_ext)
        d.has_uni = int(pb.axis is not None)
        if pb.axis is not None:
            d.ku1, d.ku2 = pb.ku1, pb.ku2
            d.axis[:] = list(pb.axis)
        self._mask = None
        if pb.mask is not None:
            self._mask = np.ascontiguousarray(pb.mask, dtype=np.uint8).ravel()
            d.mask = self._mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        d.atol, d.rtol, d.dt_min, d.dt_max = pb.atol, pb.rtol, pb.dt_min, pb.dt_max
        d.headroom, d.growth, d.shrink = pb.headroom, pb.growth, pb.shrink
        self._keep = []
        for nm in ("ms_cell", "ku1_cell", "ku2_cell", "axis_cell"):
            a = getattr(pb, nm)
            if a is not None:
                a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
                self._keep.append(a)
                setattr(d, nm, _ptr(a))
        d.wave_kind = pb.wave_kind
        d.wave_amp[:] = list(pb.wave_amp)
        d.wave_t0, d.wave_t1, d.wave_freq, d.wave_phase = pb.wave_t0, pb.wave_t1, pb.wave_freq, pb.wave_phase
        if pb.wave_mask is not None:
            wm = np.ascontiguousarray(pb.wave_mask, dtype=np.uint8).ravel()
            self._keep.append(wm)
            d.wave_mask = wm.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        if pb.cubic is not None:
            d.has_cub = 1
            d.kc1, d.kc2, d.kc3 = pb.cubic[0], pb.cubic[1], pb.cubic[2]
            d.cub_axis1[:] = list(pb.cubic[3])
            d.cub_axis2[:] = list(pb.cubic[4])
            for nm in ("kc1_cell", "kc2_cell", "kc3_cell"):
                a = getattr(pb, nm)
                if a is not None:
                    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
                    self._keep.append(a)
                    setattr(d, nm, _ptr(a))
        d.dmi_i, d.dmi_b = pb.dmi_i, pb.dmi_b
        if pb.mel is not None:
            d.has_mel = 1
            d.mel_b1, d.mel_b2 = pb.mel[0], pb.mel[1]
            d.mel_strain[:] = list(pb.mel[2])
        d.j[:] = list(pb.j)
        d.stt_degree, d.stt_beta = pb.stt_degree, pb.stt_beta
        d.slon_p[:] = list(pb.slon_p)
        d.slon_lambda, d.slon_eps = pb.slon_lambda, pb.slon_eps
        if pb.sot is not None:
            d.has_sot = 1
            d.sot_je, d.sot_xi_dl, d.sot_xi_fl = pb.sot[0], pb.sot[1], pb.sot[2]
            d.sot_sigma[:] = list(pb.sot[3])
            d.sot_thickness = pb.sot[4]
        if pb.oersted is not None:
            o = pb.oersted
            d.has_oe = 1
            d.oe_current, d.oe_radius = o[0], o[1]
            d.oe_center[:] = list(o[2])
            d.oe_axis[:] = list(o[3])
            d.oe_kind = int(o[4]) if len(o) > 4 else 0
            if len(o) > 5:
                d.oe_freq, d.oe_phase, d.oe_offset, d.oe_t_on, d.oe_t_off = o[5:10]
        d.temperature, d.thermal_dt, d.thermal_seed = pb.temperature, pb.thermal_dt, pb.thermal_seed
        d.bc_tier, d.bc_phi_floor, d.bc_delta_min = pb.boundary_tier, pb.phi_floor, pb.delta_min
        d.periodic = sum(1 << a for a in range(3) if pb.periodic[a])
        d.images[:] = [int(v) for v in pb.periodic_images]

        def f64(a):
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
            self._keep.append(a)
            return _ptr(a)
        if pb.layer_z is not None:
            d.layer_z = f64(pb.layer_z)
        if pb.regions is not None:
            rg = np.ascontiguousarray(np.asarray(pb.regions, dtype=np.uint32).ravel())
            self._keep.append(rg)
            d.region = rg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
            d.exchange_lut = f64(pb.exchange_lut)
        if pb.volume_fraction is not None:
            d.bc_phi = f64(pb.volume_fraction)
        for nm, arrays in (("bc_face", pb.face_links), ("bc_delta", pb.deltas)):
            if arrays is not None:
                dst = getattr(d, nm)
                for q in range(6):
                    dst[q] = f64(arrays[q])
        if pb.demag_corr is not None:
            tg, src, ten = pb.demag_corr
            tg = np.ascontiguousarray(np.asarray(tg, dtype=np.int32).ravel())
            src = np.ascontiguousarray(np.asarray(src, dtype=np.int32).reshape(len(tg), -1))
            self._keep += [tg, src]
            d.dc_count, d.dc_stencil = len(tg), src.shape[1]
            d.dc_target = tg.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            d.dc_source = src.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            d.dc_tensor = f64(ten)
        self.pb = pb
        self.n = pb.nx * pb.ny * pb.nz
        self.h = L.orc_create(ctypes.byref(d))
        if not self.h:
            raise ValueError("oracle create failed")

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.orc_destroy(h)
            self.h = None

    def _m(self, m):
        a = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(self.n, 3))
        return a

    def prepare(self, m):
        a = self._m(m).copy()
        lib().orc_prepare_state(self.h, _ptr(a))
        return a

    def fields(self, m):
        a = self._m(m)
        outs = {k: np.zeros((self.n, 3)) for k in ("h_ex", "h_demag", "h_ext", "h_ani", "h_eff")}
        lib().orc_fields(self.h, _ptr(a), *[_ptr(outs[k]) for k in ("h_ex", "h_demag", "h_ext", "h_ani", "h_eff")])
        return outs

    def demag_direct(self, m):
        a = self._m(m)
        out = np.zeros((self.n, 3))
        lib().orc_demag_direct(self.h, _ptr(a), _ptr(out))
        return out

    def rhs(self, m):
        a = self._m(m)
        out = np.zeros((self.n, 3))
        lib().orc_rhs(self.h, _ptr(a), _ptr(out))
        return out

    def observables(self, m):
        a = self._m(m)
        r = Report()
        lib().orc_observables(self.h, _ptr(a), ctypes.byref(r))
        return r.as_dict()

    def kernel_spectrum(self, c):
        P = 8 * self.n
        out = np.zeros(2 * P)
        assert lib().orc_kernel_spectrum(self.h, c, _ptr(out)) == 0
        return out

    def invalidate_fsal(self):
        lib().orc_invalidate_fsal(self.h)

    def reset_abm(self):
        """Drop the ABM3 history (the next three steps are Heun start-up steps)."""
        lib().orc_reset_abm(self.h)

    def thermal_step(self):
        """Step calls so far (the thermal field's counter)."""
        return lib().orc_thermal_step(self.h)

    def abm_count(self):
        """RHS values in the ABM3 history since the last restart (0..3)."""
        return lib().orc_abm_count(self.h)

    def set_time(self, t):
        """Time at which fields()/rhs()/observables() evaluate the waveform."""
        lib().orc_set_time(self.h, t)

    def step(self, integrator, m, t, dt):
        """One step in place on `m` (n,3 float64 C-contiguous). Returns (t, report)."""
        tt = ctypes.c_double(t)
        r = Report()
        fn = {"heun": lib().orc_heun_step, "rk4": lib().orc_rk4_step, "rk45": lib().orc_rk45_step,
              "rk23": lib().orc_rk23_step, "abm3": lib().orc_abm3_step}[integrator]
        fn(self.h, _ptr(m), ctypes.byref(tt), dt, ctypes.byref(r))
        return tt.value, r.as_dict()

    def run(self, integrator, m, t, until, dt0, max_steps, trace=False):
        code = {"heun": 1, "rk45": 2, "abm3": 3, "rk4": 4, "rk23": 5}[integrator]
        tt = ctypes.c_double(t)
        r = Report()
        tr = np.zeros(2 * max_steps) if trace else None
        n = lib().orc_run(self.h, code, _ptr(m), ctypes.byref(tt), until, dt0, max_steps, ctypes.byref(r),
                          _ptr(tr) if trace else None)
        res = {"t": tt.value, "steps": int(n), "last": r.as_dict()}
        if trace:
            res["trace"] = tr[:2 * n].reshape(n, 2)
        return res


def splitmix64(z):
    """SplitMix64 finaliser of the oracle's thermal generator."""
    return lib().orc_splitmix64(z)
