"""Differentiable fixed-step LLG runs on the GPU engine.

`WaveguideProblem` describes a waveguide with a trainable per-cell Ms map
over a rectangular design region and a sinc field pulse of trainable
amplitude on the first columns; the objective is the mean of m_c^2 over
the last columns during the final time window.  `LLGObjective` is the
torch.autograd.Function: forward runs the engine's checkpointed fixed-step
forward pass (`fullmag_fdm_adjoint_forward`), backward its reverse sweep
(`fullmag_fdm_adjoint_backward`), which reuses the symmetric demag
convolution for the field cotangent.  No arithmetic happens here; the
module only maps parameters to plan arrays and gradients back.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _abi as abi
from .backend import FdmBackend, FdmPlan

MU0 = 4e-7 * np.pi


@dataclass
class WaveguideProblem:
    """Defaults: 256x64x2 cells of 5 nm, Ms 8e5 A/m outside a 64x64 design
    region (x 96..159, all y, both layers), A 1.3e-11 J/m, alpha 0.01, sinc
    pulse B(t) = A sinc(2 pi f_c (t - t0)) along y on the first 8 columns
    (f_c 20 GHz, t0 0.1 ns), RK4 with dt 50 fs for 20000 steps, objective
    mean m_z^2 over the last 8 columns in the last 0.5 ns."""
    nx: int = 256
    ny: int = 64
    nz: int = 2
    cell: float = 5e-9
    ms_base: float = 8e5
    a_ex: float = 1.3e-11
    alpha: float = 0.01
    design_x: tuple = (96, 160)
    design_y: tuple = (0, 64)
    pulse_cols: int = 8
    probe_cols: int = 8
    f_cut: float = 20e9
    t0: float = 0.1e-9
    direction: tuple = (0.0, 1.0, 0.0)
    dt: float = 50e-15
    steps: int = 20000
    window: float = 0.5e-9
    component: int = 2
    precision: str = "single"
    integrator: str = "rk4"
    checkpoint_interval: int = 0
    beta_state_precision: bool = False
    record_budget_bytes: int = 0
    m0: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    def design_shape(self):
        return (self.design_y[1] - self.design_y[0], self.design_x[1] - self.design_x[0])

    def _xy(self):
        z, y, x = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        return x.ravel(), y.ravel(), z.ravel()

    def design_cells(self):
        """Cell indices of the design region, shape (nz, dy, dx)."""
        x, y, _ = self._xy()
        sel = ((x >= self.design_x[0]) & (x < self.design_x[1]) & (y >= self.design_y[0]) & (y < self.design_y[1]))
        return np.nonzero(sel)[0].reshape(self.nz, *self.design_shape())

    def ms_map(self, ms_design: np.ndarray) -> np.ndarray:
        """Per-cell Ms: ms_base outside the design region, the (dy, dx)
        design map on every layer inside it."""
        ms = np.full(self.n, self.ms_base)
        idx = self.design_cells()
        for k in range(self.nz):
            ms[idx[k]] = np.asarray(ms_design, dtype=np.float64)
        return ms

    def design_grad(self, grad_ms: np.ndarray) -> np.ndarray:
        """dJ/d(design map): the per-cell gradient summed over the layers."""
        idx = self.design_cells()
        return sum(grad_ms[idx[k]] for k in range(self.nz))

    def weights(self) -> np.ndarray:
        x, _, _ = self._xy()
        w = (x >= self.nx - self.probe_cols).astype(np.float64)
        return w / w.sum()

    def first_step(self) -> int:
        return max(1, self.steps - int(round(self.window / self.dt)) + 1)

    def initial_m(self) -> np.ndarray:
        if self.m0 is not None:
            return self.m0
        m = np.tile(np.array([1.0, 0.0, 0.0]), (self.n, 1))
        return m

    def plan(self, ms_design: np.ndarray, amplitude_t: float) -> FdmPlan:
        x, _, _ = self._xy()
        wmask = (x < self.pulse_cols).astype(np.uint8)
        h1 = amplitude_t / MU0 * np.asarray(self.direction, dtype=np.float64)
        kw = dict(nx=self.nx, ny=self.ny, nz=self.nz, dx=self.cell, dy=self.cell, dz=self.cell, ms=self.ms_base,
                  a_ex=self.a_ex, alpha=self.alpha, m0=self.initial_m(), precision=self.precision,
                  integrator=self.integrator, ms_field=self.ms_map(ms_design), wave=abi.WAVE_SINC,
                  wave_amp=tuple(h1), wave_t0=self.t0, wave_freq=self.f_cut, wave_mask=wmask)
        kw.update(self.extra)
        return FdmPlan(**kw)

    def run(self, ms_design, amplitude_t, backend: FdmBackend | None = None):
        """Forward pass only; returns (J, backend) with the backend ready for
        `adjoint_backward`."""
        be = backend or FdmBackend(self.plan(ms_design, amplitude_t))
        j = be.adjoint_forward(self.steps, self.dt, self.weights(), self.first_step(), self.component,
                               self.checkpoint_interval, self.beta_state_precision, self.record_budget_bytes)
        return j, be

    def value_and_grad(self, ms_design, amplitude_t):
        """(J, dJ/d design map, dJ/d amplitude in T)."""
        j, be = self.run(ms_design, amplitude_t)
        try:
            g_ms, g_amp = be.adjoint_backward()
        finally:
            be.close()
        d = np.asarray(self.direction, dtype=np.float64)
        return j, self.design_grad(g_ms), float(g_amp @ d) / MU0


def _torch():
    import torch
    return torch


class LLGObjective:
    """torch.autograd.Function: J(ms_design, amplitude) of a WaveguideProblem.

        J = LLGObjective.apply(ms_design, amplitude, problem)

    ms_design: tensor of the design-region shape (A/m); amplitude: scalar
    tensor (T).  Gradients flow to both."""

    _fn = None

    @classmethod
    def _build(cls):
        torch = _torch()

        class _Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, ms_design, amplitude, problem):
                j, be = problem.run(ms_design.detach().cpu().double().numpy(), float(amplitude))
                ctx.be = be
                ctx.problem = problem
                ctx.meta = (ms_design.dtype, ms_design.device, amplitude.dtype, amplitude.device)
                return torch.tensor(j, dtype=torch.float64)

            @staticmethod
            def backward(ctx, g):
                be, pr = ctx.be, ctx.problem
                try:
                    g_ms, g_amp = be.adjoint_backward()
                finally:
                    be.close()
                d = np.asarray(pr.direction, dtype=np.float64)
                ms_dt, ms_dev, a_dt, a_dev = ctx.meta
                gd = torch.as_tensor(pr.design_grad(g_ms), dtype=ms_dt, device=ms_dev)
                ga = torch.tensor(float(g_amp @ d) / MU0, dtype=a_dt, device=a_dev)
                gs = g.to(ms_dt)
                return gs * gd, g.to(a_dt) * ga, None

        cls._fn = _Fn

    @classmethod
    def apply(cls, ms_design, amplitude, problem):
        if cls._fn is None:
            cls._build()
        return cls._fn.apply(ms_design, amplitude, problem)
