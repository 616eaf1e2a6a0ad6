"""Shared problem builders: one parameter set -> engine plan + oracle problem."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MU0 = 4.0 * np.pi * 1e-7


@dataclass
class Case:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    ms: float
    a_ex: float
    alpha: float
    m0: np.ndarray
    mask: np.ndarray | None = None
    h_ext: tuple | None = None
    ku1: float = 0.0
    ku2: float = 0.0
    axis: tuple | None = None
    demag: bool = True
    exchange: bool = True
    precession: bool = True
    gamma: float = 2.211e5
    atol: float = 1e-5
    rtol: float = 0.0
    dt_min: float = 1e-18
    dt_max: float = 1e-10
    headroom: float = 0.8
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
    # further terms in the oracle Problem's form (cubic, kc*_cell, dmi_i,
    # dmi_b, mel, j, stt_degree, stt_beta, slon_p, slon_lambda, slon_eps, sot,
    # oersted, temperature, thermal_dt, thermal_seed, boundary_tier,
    # phi_floor, delta_min, volume_fraction, face_links, deltas, demag_corr);
    # plan() maps them onto the plan descriptor's and the plan extension's
    # fields
    terms: dict | None = None

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    def _desc_terms(self):
        """Plan-descriptor fields (fullmag_fdm.h) and per-cell arrays of `terms`."""
        t = self.terms or {}
        d, cells = {}, {}
        if "cubic" in t:
            k1, k2, k3, a1, a2 = t["cubic"]
            d.update(has_cubic_anisotropy=1, cubic_Kc1=k1, cubic_Kc2=k2, cubic_Kc3=k3, cubic_axis1=tuple(a1),
                     cubic_axis2=tuple(a2))
        for q in ("kc1", "kc2", "kc3"):
            if t.get(q + "_cell") is not None:
                cells[q + "_field"] = t[q + "_cell"]
        if t.get("dmi_i"):
            d.update(has_interfacial_dmi=1, dmi_D_interfacial=t["dmi_i"])
        if t.get("dmi_b"):
            d.update(has_bulk_dmi=1, dmi_D_bulk=t["dmi_b"])
        if "mel" in t:
            d.update(has_magnetoelastic=1, mel_b1=t["mel"][0], mel_b2=t["mel"][1], mel_strain=tuple(t["mel"][2]))
        j = t.get("j", (0.0, 0.0, 0.0))
        d.update(current_density_x=j[0], current_density_y=j[1], current_density_z=j[2],
                 stt_degree=t.get("stt_degree", 0.0), stt_beta=t.get("stt_beta", 0.0))
        sp = t.get("slon_p", (0.0, 0.0, 0.0))
        d.update(stt_p_x=sp[0], stt_p_y=sp[1], stt_p_z=sp[2], stt_lambda=t.get("slon_lambda", 0.0),
                 stt_epsilon_prime=t.get("slon_eps", 0.0))
        if "sot" in t:
            je, dl, fl, sg, th = t["sot"]
            d.update(has_sot=1, sot_je=je, sot_xi_dl=dl, sot_xi_fl=fl, sot_sigma=tuple(sg), sot_thickness=th)
        if "oersted" in t:
            o = tuple(t["oersted"]) + (0, 0.0, 0.0, 0.0, 0.0, 0.0)[len(t["oersted"]) - 4:]
            d.update(has_oersted_cylinder=1, oersted_current=o[0], oersted_radius=o[1], oersted_center=tuple(o[2]),
                     oersted_axis=tuple(o[3]), oersted_time_dep_kind=int(o[4]), oersted_time_dep_freq=o[5],
                     oersted_time_dep_phase=o[6], oersted_time_dep_offset=o[7], oersted_time_dep_t_on=o[8],
                     oersted_time_dep_t_off=o[9])
        return d, cells

    def plan(self, precision="double", integrator="dp45", **kw):
        from fullmag_amd.backend import FdmPlan
        d, cells = self._desc_terms()
        if d:
            kw["extra"] = {**d, **kw.get("extra", {})}
        kw.update(cells)
        for q in ("temperature", "thermal_dt", "thermal_seed", "boundary_tier", "phi_floor", "delta_min",
                  "volume_fraction", "face_links", "deltas", "demag_corr", "periodic", "periodic_images", "layer_z",
                  "regions", "exchange_lut"):
            if q in (self.terms or {}):
                kw.setdefault(q, self.terms[q])
        return FdmPlan(nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dy, dz=self.dz, ms=self.ms,
                       a_ex=self.a_ex, alpha=self.alpha, m0=self.m0, gamma=self.gamma, precision=precision,
                       integrator=integrator, exchange=self.exchange, demag=self.demag,
                       precession=self.precession, h_ext=self.h_ext, ku1=self.ku1, ku2=self.ku2, axis=self.axis,
                       mask=self.mask, atol=self.atol, dt_min=self.dt_min, dt_max=self.dt_max,
                       headroom=self.headroom, ku1_field=self.ku1_cell, ku2_field=self.ku2_cell,
                       ms_field=self.ms_cell, axis_field=self.axis_cell, wave=self.wave_kind, wave_amp=self.wave_amp,
                       wave_t0=self.wave_t0, wave_t1=self.wave_t1, wave_freq=self.wave_freq,
                       wave_phase=self.wave_phase, wave_mask=self.wave_mask, rtol=self.rtol, **kw)

    def problem(self):
        from oracle.oracle import Problem
        return Problem(nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dy, dz=self.dz, ms=self.ms,
                       a_ex=self.a_ex, alpha=self.alpha, gamma=self.gamma, precession=self.precession,
                       exchange=self.exchange, demag=self.demag, h_ext=self.h_ext, ku1=self.ku1, ku2=self.ku2,
                       axis=self.axis, mask=self.mask, atol=self.atol, rtol=self.rtol, dt_min=self.dt_min,
                       dt_max=self.dt_max, headroom=self.headroom, ms_cell=self.ms_cell, ku1_cell=self.ku1_cell,
                       ku2_cell=self.ku2_cell, axis_cell=self.axis_cell, wave_kind=self.wave_kind,
                       wave_amp=self.wave_amp, wave_t0=self.wave_t0, wave_t1=self.wave_t1, wave_freq=self.wave_freq,
                       wave_phase=self.wave_phase, wave_mask=self.wave_mask, **(self.terms or {}))


def masked_3x3(demag: bool) -> Case:
    """crates/fullmag-runner/src/native_fdm.rs:1278-1319 (masked 3x3x1 plan)."""
    m0 = np.array([
        [1.0, 0.0, 0.0],
        [0.9950041652780258, 0.09983341664682815, 0.0],
        [0.9800665778412416, 0.19866933079506122, 0.0],
        [0.9992009587217894, 0.0, 0.03996803834887158],
        [0.9937606691655043, 0.09970865087213879, 0.04972948160146045],
        [0.9778332467629838, 0.19771314245924698, 0.06988589031642899],
        [0.9968017063026194, -0.039904089712529575, 0.06972124896577284],
        [0.9892364775387807, 0.05946310942269411, 0.1338082836649087],
        [0.9711213242426827, 0.15730105252897553, 0.17902957342582418],
    ])
    mask = np.array([1, 1, 1, 1, 0, 1, 1, 1, 0], dtype=np.uint8)
    return Case(3, 3, 1, 5e-9, 5e-9, 10e-9, 800e3, 13e-12, 0.1, m0, mask=mask, h_ext=(1.5e3, -2.0e3, 7.5e2),
                demag=demag)


def thin_film_8x6() -> Case:
    """crates/fullmag-runner/src/native_fdm.rs:1321-1365 (thin film with demag)."""
    nx, ny = 8, 6
    m0 = []
    for y in range(ny):
        for x in range(nx):
            th, ph = 0.11 * x, 0.07 * y
            v = np.array([np.cos(th) * np.cos(ph), np.sin(th) * np.cos(ph), 0.2 * np.sin(ph)])
            m0.append(v / np.linalg.norm(v))
    return Case(nx, ny, 1, 4e-9, 4e-9, 10e-9, 800e3, 13e-12, 0.1, np.array(m0), h_ext=(2.0e3, -1.0e3, 5.0e2))


def random_case(nx, ny, nz, seed=7, hole_radius_cells=None, dx=2e-9, dy=2e-9, dz=5e-9, ms=8.6e5, a_ex=1.3e-11,
                alpha=0.02, h_ext=(0.01 / MU0, 0.0, 0.0), **kw) -> Case:
    from fullmag_amd.backend import hole_mask, random_unit_vectors
    m0 = random_unit_vectors(seed, nx * ny * nz)
    mask = None
    if hole_radius_cells:
        mask = hole_mask(nx, ny, nz, dx, dy, hole_radius_cells * dx)
        m0[mask == 0] = 0.0
    return Case(nx, ny, nz, dx, dy, dz, ms, a_ex, alpha, m0, mask=mask, h_ext=h_ext, **kw)


def thin_film_512(seed=42) -> Case:
    """512x512x1 film of 2x2x5 nm cells with a 64-cell hole, Ms 8.6e5 A/m,
    A 1.3e-11 J/m, alpha 0.02, B = 10 mT along x, seeded random m0."""
    return random_case(512, 512, 1, seed=seed, hole_radius_cells=64)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


def disk_boundary(nx, ny, nz, dx, dy, dz, radius, sub=8):
    """Boundary-tier data of a disk of `radius` centred in an nx x ny grid
    (every layer alike): volume fraction from sub x sub samples per cell, face
    link fractions from sub samples along each x/y face (z faces 1 between
    occupied layers), distances from the cell centre to the circle along
    +-x / +-y (capped at one cell size; z: dz).  Returns (mask, phi,
    face_links, deltas), arrays in the order +x, -x, +y, -y, +z, -z."""
    cx, cy = nx * dx / 2, ny * dy / 2
    u = (np.arange(sub) + 0.5) / sub
    X = (np.arange(nx)[None, :, None, None] + u[None, None, None, :]) * dx - cx
    Y = (np.arange(ny)[:, None, None, None] + u[None, None, :, None]) * dy - cy
    phi2 = ((X ** 2 + Y ** 2) <= radius ** 2).mean(axis=(2, 3))
    xc = (np.arange(nx) + 0.5) * dx - cx
    yc = (np.arange(ny) + 0.5) * dy - cy
    xf = (np.arange(nx + 1)) * dx - cx            # x faces
    yf = (np.arange(ny + 1)) * dy - cy
    fx = ((xf[None, :, None] ** 2 + ((np.arange(ny)[:, None, None] + u[None, None, :]) * dy - cy) ** 2)
          <= radius ** 2).mean(axis=2)            # (ny, nx+1)
    fy = ((yf[:, None, None] ** 2 + ((np.arange(nx)[None, :, None] + u[None, None, :]) * dx - cx) ** 2)
          <= radius ** 2).mean(axis=2)            # (ny+1, nx)
    xp, xm = fx[:, 1:].copy(), fx[:, :-1].copy()
    xp[:, -1] = 0.0
    xm[:, 0] = 0.0
    yp, ym = fy[1:, :].copy(), fy[:-1, :].copy()
    yp[-1, :] = 0.0
    ym[0, :] = 0.0
    # distance to the circle along each axis from the centre (inside cells)
    Xc, Yc = np.meshgrid(xc, yc)
    r2 = radius ** 2
    sx = np.sqrt(np.maximum(r2 - Yc ** 2, 0.0))
    sy = np.sqrt(np.maximum(r2 - Xc ** 2, 0.0))
    dxp = np.clip(sx - Xc, 0.0, dx)
    dxm = np.clip(Xc + sx, 0.0, dx)
    dyp = np.clip(sy - Yc, 0.0, dy)
    dym = np.clip(Yc + sy, 0.0, dy)

    def stack(a):
        return np.repeat(a[None], nz, axis=0).ravel()
    phi = stack(phi2)
    mask = (phi > 0).astype(np.uint8)
    zp = np.ones((nz, ny, nx))
    zm = np.ones((nz, ny, nx))
    zp[-1] = 0.0
    zm[0] = 0.0
    zp, zm = zp.ravel() * (phi > 0), zm.ravel() * (phi > 0)
    links = (stack(xp), stack(xm), stack(yp), stack(ym), zp, zm)
    dzs = np.full(nx * ny * nz, dz)
    deltas = (stack(dxp), stack(dxm), stack(dyp), stack(dym), dzs, dzs)
    return mask, phi, links, deltas
