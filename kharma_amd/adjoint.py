"""Differentiation form: gradients of a multi-step GRMHD evolution with respect
to its initial primitive field, through the hand-written adjoint kernels of
libkharma_amd (kh_adjoint_step).

    P_N = evolve(P0, cfg, nsteps)          # P0: (8, nx3, nx2, nx1) float64 CUDA tensor
    loss = ((P_N - P_target) ** 2).sum()
    loss.backward()                         # P0.grad via the adjoint sweep

The forward pass runs the FP64 driver step by step and keeps one checkpoint
(the starting primitives) per step; the backward pass walks the steps in
reverse, each adjoint step recomputing its forward stages from its checkpoint.
The timestep sequence is an input of the map: by default it is the CFL sequence
the driver picks from P0 (recorded in the forward pass) and it is then held
fixed, so the gradient is that of the map P0 -> P_N at those timesteps.

Supported blocks are those of kh_adjoint_create: FP64, WENO5 + LLF, every axis
periodic, with no zone floored or failing to invert during the run.
"""
import math

import numpy as np
import torch

import kharma_amd as kh
from kharma_amd import problems as pr


class Evolution:
    """FP64 driver + adjoint context for one block configuration."""

    def __init__(self, cfg):
        self.cfg = dict(cfg, precision="fp64")
        self.drv = kh.Driver(self.cfg)
        self.adj = kh.AdjointStep(self.cfg)
        self.ndim, self.shape = pr.block_shape(self.cfg)
        self.inner = (slice(None),) + pr.interior(self.cfg)

    def _embed(self, P_int):
        P = torch.zeros((pr.NV,) + self.shape, dtype=torch.float64, device="cuda")
        P[self.inner] = P_int
        return P

    def run(self, P_int, nsteps, dts=None, checkpoints=False):
        """Advance the interior field P_int by nsteps.  Returns (P_N interior,
        timesteps used, checkpoints: the nsteps + 1 states P_0 .. P_N, or None).
        With checkpoints the second stage of every step is checked for floored
        or failed zones (the adjoint step recomputes and checks the first)."""
        self.drv.init_prims(self._embed(P_int.detach()))
        ckpts = [] if checkpoints else None
        flagged = torch.zeros((), dtype=torch.int64, device="cuda")
        used = []
        time = 0.0
        for n in range(nsteps):
            if dts is not None:
                self.drv.set_clock(time, float(dts[n]), n)
            if checkpoints:
                P, _ = self.drv.state()
                ckpts.append(P)
            self.drv.step(1)
            if checkpoints:
                pf, ff = self.drv.flags()
                flagged += ((pf[self.inner[1:]] != 0) | (ff[self.inner[1:]] != 0)).sum()
            dt = dts[n] if dts is not None else self.drv.dt_history(n + 1)[n]
            used.append(float(dt))
            time += float(dt)
        P, _ = self.drv.state()
        if checkpoints:
            ckpts.append(P)
            if int(flagged) > 0:
                raise kh.KharmaError(f"{int(flagged)} zone-steps were floored or failed to invert; "
                                     "the adjoint covers unfloored, converged runs only")
        return P[self.inner].contiguous(), used, ckpts

    def backward(self, lam_int, ckpts, dts):
        """dL/dP0 (interior) from dL/dP_N (interior) by the adjoint sweep."""
        lam = self._embed(lam_int.detach())
        for n in reversed(range(len(dts))):
            self.adj.step(ckpts[n], dts[n], lam, ckpts[n + 1])
        return lam[self.inner].contiguous()


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, P0, evo, nsteps, dts):
        PN, used, ckpts = evo.run(P0, nsteps, dts, checkpoints=True)
        ctx.evo, ctx.ckpts, ctx.dts = evo, ckpts, used
        return PN

    @staticmethod
    def backward(ctx, gPN):
        g = ctx.evo.backward(gPN.to(torch.float64), ctx.ckpts, ctx.dts)
        ctx.ckpts = None
        return g, None, None, None


def evolve(P0, cfg_or_evo, nsteps, dts=None):
    """P_N (interior) after nsteps from the interior field P0, differentiable in P0."""
    evo = cfg_or_evo if isinstance(cfg_or_evo, Evolution) else Evolution(cfg_or_evo)
    return _Evolve.apply(P0, evo, nsteps, dts)


def smooth_direction(cfg, rng, kmax=3, modes=4):
    """A unit-norm direction for gradient checks: in each of the 8 variables a sum
    of `modes` random Fourier modes with integer wavenumbers |k_d| <= kmax on the
    periodic block (interior shape, numpy).  WENO5's weights stay at their
    smooth-data values along such directions, also where a field is uniform."""
    x, y, z = pr.cell_centers(cfg)
    ndim, _ = pr.block_shape(cfg)
    Z, Y, X = np.meshgrid(z if ndim == 3 else [0.], y if ndim > 1 else [0.], x, indexing="ij")
    L = [b[1] - b[0] for b in cfg["bounds"]]
    v = np.zeros((pr.NV,) + X.shape)
    for q in range(pr.NV):
        for _ in range(modes):
            k = rng.integers(-kmax, kmax + 1, size=3)
            if ndim < 3:
                k[2] = 0
            if ndim < 2:
                k[1] = 0
            ph = rng.uniform(0, 2 * math.pi)
            v[q] += rng.normal() * np.cos(2 * math.pi * (k[0] * X / L[0] + k[1] * Y / L[1] + k[2] * Z / L[2]) + ph)
    return v / np.linalg.norm(v)


def noise_direction(cfg, rng, fields):
    """A unit-norm white-noise direction confined to the listed variables (interior shape)."""
    inner = pr.interior(cfg)
    v = np.zeros((pr.NV,) + tuple(s.stop - s.start for s in inner))
    for q in fields:
        v[q] = rng.normal(size=v[q].shape)
    return v / np.linalg.norm(v)


def generic_state(cfg, rng, amp=1e-3, base=None):
    """An interior state (numpy) on which the step is smooth in every direction:
    `base` (default: the configured linear mode) plus, in each of the 8
    variables, one random Fourier mode with |k_d| in {1, 2} along every axis and
    relative amplitude `amp`.  Unlike the modes themselves, no field is uniform
    along any axis, so WENO5's smoothness indicators are nowhere zero and central
    differences of the step measure its derivative (a uniform field sits at
    beta = 0, where any finite perturbation moves the nonlinear weights by O(1))."""
    x, y, z = pr.cell_centers(cfg)
    ndim, _ = pr.block_shape(cfg)
    Z, Y, X = np.meshgrid(z if ndim == 3 else [0.], y if ndim > 1 else [0.], x, indexing="ij")
    L = [b[1] - b[0] for b in cfg["bounds"]]
    P = pr.mhdmode_state(cfg) if base is None else np.array(base, dtype=np.float64)
    for q in range(pr.NV):
        k = rng.integers(1, 3, size=3) * rng.choice([-1, 1], size=3)
        if ndim < 3:
            k[2] = 0
        if ndim < 2:
            k[1] = 0
        ph = rng.uniform(0, 2 * math.pi)
        scale = max(abs(P[q]).max(), 0.1)
        P[q] = P[q] + amp * scale * np.cos(2 * math.pi * (k[0] * X / L[0] + k[1] * Y / L[1] + k[2] * Z / L[2]) + ph)
    return P
