"""Differentiation form: gradients of a multi-step GRMHD evolution with respect
to its initial primitive field, through the hand-written adjoint kernels of
libkharma_amd (kh_adjoint_step).

    P_N = evolve(P0, cfg, nsteps)          # P0: (8, nx3, nx2, nx1) float64 CUDA tensor
    loss = ((P_N - P_target) ** 2).sum()
    loss.backward()                         # P0.grad via the adjoint sweep

    P_N, J = evolve_objective(P0, evo, nsteps)   # J: a face-flux objective of the run
    J.backward()

The forward pass runs the FP64 driver and keeps checkpoints of the primitives:
every step (every=1), or every `every` steps, in which case the backward pass
first re-runs each segment from its checkpoint to recover the per-step states
(two-level, revolve-style checkpointing: N / every + every states held instead
of N).  The backward pass walks the steps in reverse, each adjoint step
recomputing its RK stages from its starting state.  The timestep sequence is an
input of the map: by default the CFL sequence the driver picks from P0
(recorded in the forward pass), then held fixed, so the gradient is that of the
map P0 -> P_N at those timesteps.

Supported blocks are those of kh_adjoint_create: FP64, WENO5 + LLF, Minkowski
(no zone floored) or Kerr-Schild (floors included); zones failing to invert are
covered through the transpose of their fixup.
"""
import math

import numpy as np
import torch

import kharma_amd as kh
from kharma_amd import problems as pr


class Evolution:
    """FP64 driver + adjoint context for one block configuration.  objective:
    None, or dict(dir, var, face, weight=1) of the face-flux objective
    J = weight sum_steps sum_stages w_s dt sum F_dir[var](face) dA
    (kh_driver_set_face_flux_objective)."""

    def __init__(self, cfg, objective=None):
        self.cfg = dict(cfg, precision="fp64")
        self.drv = kh.Driver(self.cfg)
        self.adj = kh.AdjointStep(self.cfg)
        self.ndim, self.shape = pr.block_shape(self.cfg)
        self.inner = (slice(None),) + pr.interior(self.cfg)
        # the sweep checks the steps' coverage once, at its end (no per-step sync)
        self.deferred_checks = True
        self.objective = dict(objective) if objective else None
        if self.objective:
            o = self.objective
            self.drv.set_face_flux_objective(o["dir"], o["var"], o["face"], o.get("weight", 1.0))

    def _embed(self, P_int):
        P = torch.zeros((pr.NV,) + self.shape, dtype=torch.float64, device="cuda")
        P[self.inner] = P_int
        return P

    def _check_flags(self, flagged):
        if int(flagged) > 0:
            raise kh.KharmaError(f"{int(flagged)} zone-steps were floored on a flat-space block or failed to "
                                 "invert with the fixup off; the adjoint does not cover them")

    def _count(self, flagged, floors_ok, fixup_ok):
        pf, ff = self.drv.flags()
        bad = torch.zeros_like(pf[self.inner[1:]], dtype=torch.bool)
        if not fixup_ok:
            bad = bad | (pf[self.inner[1:]] != 0)
        if not floors_ok:
            bad = bad | (ff[self.inner[1:]] != 0)
        flagged += bad.sum()

    def run(self, P_int, nsteps, dts=None, checkpoints=False, every=1):
        """Advance the interior field P_int by nsteps.  Returns (P_N interior,
        timesteps used, checkpoints {step: full P} every `every` steps and at
        nsteps, or None, objective J (0 without one)).  With checkpoints the
        second stage of every step is checked for zones the adjoint does not
        cover (floored in flat space; failed to invert with the fixup off); the
        adjoint step recomputes and checks the first.  With every = 1 the
        driver also records each step's stage intermediates
        (checkpoints["records"]: {step: record}), which the adjoint step then
        uses instead of recomputing the stages."""
        self.drv.init_prims(self._embed(P_int.detach()))
        if self.objective:
            self.drv.objective(reset=True)
        ckpts = {} if checkpoints else None
        flagged = torch.zeros((), dtype=torch.int64, device="cuda")
        floors_ok = self.cfg.get("coords", "cartesian_minkowski") != "cartesian_minkowski"
        # the adjoint step recomputes stage 2 (and transposes its fixup) except on
        # periodic flat-space blocks, where it takes stage 2 from the next checkpoint
        recomputes = floors_ok or bool(self.cfg.get("boundaries"))
        fixup_ok = recomputes and (self.cfg.get("fix_average", True) or self.cfg.get("fix_atmo", True))
        record = checkpoints and every == 1
        if record:
            ckpts["records"] = {}
        used = []
        time = 0.0
        try:
            for n in range(nsteps):
                if dts is not None:
                    self.drv.set_clock(time, float(dts[n]), n)
                if checkpoints and n % every == 0:
                    ckpts[n] = self.drv.state()[0]
                if record:
                    ckpts["records"][n] = self.drv.new_stage_record()
                    self.drv.set_stage_record(ckpts["records"][n])
                self.drv.step(1)
                if checkpoints:
                    self._count(flagged, floors_ok, fixup_ok)
                dt = dts[n] if dts is not None else self.drv.dt_history(n + 1)[n]
                used.append(float(dt))
                time += float(dt)
        finally:
            if record:
                self.drv.set_stage_record()   # the driver must not keep a pointer to a freed record
        P, _ = self.drv.state()
        J = self.drv.objective() if self.objective else 0.0
        if checkpoints:
            ckpts[nsteps] = P
            self._check_flags(flagged)
        return P[self.inner].contiguous(), used, ckpts, J

    def _segment_states(self, ckpts, dts, a, b):
        """States P_a .. P_b (full) and the records of steps a .. b-1 (None where
        not held): re-run from the checkpoint at step a, recording, unless all
        states are held."""
        held = ckpts.get("records", {})
        if all(n in ckpts for n in range(a, b + 1)):
            return [ckpts[n] for n in range(a, b + 1)], [held.get(n) for n in range(a, b)]
        self.drv.init_prims(ckpts[a])
        states, recs = [ckpts[a]], []
        time = float(sum(dts[:a]))
        try:
            for n in range(a, b):
                self.drv.set_clock(time, float(dts[n]), n)
                recs.append(self.drv.new_stage_record())
                self.drv.set_stage_record(recs[-1])
                self.drv.step(1)
                time += float(dts[n])
                if n < b - 1:
                    states.append(self.drv.state()[0])
        finally:
            self.drv.set_stage_record()
        states.append(ckpts[b])
        return states, recs

    def backward(self, lam_int, ckpts, dts, gJ=0.0):
        """dL/dP0 (interior) from dL/dP_N (interior) and dL/dJ by the adjoint sweep."""
        lam = self._embed(lam_int.detach()) if lam_int is not None else \
            torch.zeros((pr.NV,) + self.shape, dtype=torch.float64, device="cuda")
        if self.objective:
            o = self.objective
            self.adj.set_face_flux_objective(o["dir"], o["var"], o["face"], o.get("weight", 1.0) * float(gJ))
        marks = sorted(k for k in ckpts if isinstance(k, int))
        self.adj.set_deferred_checks(self.deferred_checks)
        try:
            for q in reversed(range(len(marks) - 1)):
                a, b = marks[q], marks[q + 1]
                states, recs = self._segment_states(ckpts, dts, a, b)
                for n in reversed(range(a, b)):
                    if recs[n - a] is not None:
                        self.adj.step_recorded(states[n - a], recs[n - a], dts[n], lam, states[n - a + 1])
                    else:
                        self.adj.step(states[n - a], dts[n], lam, states[n - a + 1])
            if self.deferred_checks:
                self.adj.check()
        finally:
            self.adj.set_deferred_checks(False)
            if self.objective:
                self.adj.set_face_flux_objective(0, 0, 0, 0.0)
        return lam[self.inner].contiguous()


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, P0, evo, nsteps, dts, every):
        PN, used, ckpts, J = evo.run(P0, nsteps, dts, checkpoints=True, every=every)
        ctx.evo, ctx.ckpts, ctx.dts = evo, ckpts, used
        return PN, torch.tensor(J, dtype=torch.float64, device=P0.device)

    @staticmethod
    def backward(ctx, gPN, gJ):
        g = ctx.evo.backward(gPN.to(torch.float64) if gPN is not None else None, ctx.ckpts, ctx.dts,
                             float(gJ) if gJ is not None else 0.0)
        ctx.ckpts = None
        return g, None, None, None, None


def evolve(P0, cfg_or_evo, nsteps, dts=None, every=1):
    """P_N (interior) after nsteps from the interior field P0, differentiable in P0."""
    evo = cfg_or_evo if isinstance(cfg_or_evo, Evolution) else Evolution(cfg_or_evo)
    return _Evolve.apply(P0, evo, nsteps, dts, every)[0]


def evolve_objective(P0, evo, nsteps, dts=None, every=1):
    """(P_N interior, J) after nsteps, both differentiable in P0; evo carries the objective."""
    return _Evolve.apply(P0, evo, nsteps, dts, every)


def face_index(cfg, dir, x):
    """Entire-block index of the face along dir nearest the native coordinate x."""
    lo, hi = cfg["bounds"][dir - 1]
    n = cfg["nx"][dir - 1]
    ng = cfg.get("ng", 4)
    return ng + int(round((x - lo) / ((hi - lo) / n)))


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


def smooth_torus_state(cfg, rng):
    """A torus state (interior, numpy) on which a multi-step objective is
    differentiable in every direction, for gradient checks: the torus with its
    atmosphere at 1.5 times its floors (no zone floored for a few steps), small
    smooth modes in every variable (the bare atmosphere is uniform in theta,
    where WENO5's smoothness indicators vanish and finite differences measure
    the weights of the perturbation).  Returns (P, direction), where
    direction() draws a random smooth direction in all variables, each scaled by
    the variable's local magnitude.  Run it with the recovery converged tightly
    (err_tol 1e-14) so the inversion's own tolerance does not limit the check."""
    x, y, _ = pr.cell_centers(cfg)
    X2, X1 = np.meshgrid(y, x, indexing="ij")
    L1 = cfg["bounds"][0][1] - cfg["bounds"][0][0]
    inner = (slice(None),) + pr.interior(cfg)

    def modes():
        v = np.zeros((1,) + X1.shape)
        for _m in range(3):
            k1, k2 = rng.integers(1, 4, size=2)
            v[0] += rng.normal() * np.cos(2 * math.pi * k1 * X1 / L1 + rng.uniform(0, 2 * math.pi)) * \
                np.cos(math.pi * k2 * X2 + rng.uniform(0, 2 * math.pi))
        return v
    P = pr.torus_prims(cfg)[inner].copy()
    bmax = np.abs(P[pr.B1:]).max()
    P[pr.RHO] *= 1 + 1e-3 * modes()[0]
    P[pr.UU] *= 1 + 1e-3 * modes()[0]
    for q in (pr.U1, pr.U2, pr.U3):
        P[q] += 1e-3 * modes()[0]
    for q in (pr.B1, pr.B2, pr.B3):
        P[q] += 1e-6 * bmax * modes()[0]
    atm = P[pr.RHO] < 1e-2
    P[pr.RHO][atm] *= 1.5
    P[pr.UU][atm] *= 1.5
    scale = np.empty_like(P)
    scale[pr.RHO] = np.abs(P[pr.RHO])
    scale[pr.UU] = np.abs(P[pr.UU])
    scale[pr.U1:pr.B1] = np.abs(P[pr.U1:pr.B1]) + 0.1
    scale[pr.B1:] = np.abs(P[pr.B1:]) + 1e-6 * bmax

    def direction():
        return np.concatenate([scale[q:q + 1] * modes() for q in range(pr.NV)])
    return P, direction
