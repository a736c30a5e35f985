"""Adjoint (differentiation form) against central finite differences.

The adjoint step transposes the forward step stage by stage (WENO5 + LLF face
fluxes, flux-CT + divergence, RK update, onedw recovery by the
implicit-function theorem) with the timestep held fixed, so the finite
differences here run the same forward map: the driver with the recorded
timestep sequence imposed.

States and directions.  WENO5's nonlinear weights c_k / (EPS + beta_k)^2
(EPS = 1e-26) are smooth functions of the state, but where a field is uniform
along an axis (every linear mode is uniform along x3; rho, u and B1 of the
Alfven mode are uniform everywhere) beta_k = 0, and any perturbation with
beta >> EPS -- every central difference with h = 1e-7 -- moves the weights by
O(1): the difference then measures the weights of the perturbation, not the
derivative at the state.  So the general checks run on a generic smooth state
(adjoint.generic_state: the mode plus a small 3-D Fourier mode in every
variable, uniform nowhere) in random smooth and white-noise directions, and the
Alfven-mode objective is checked in the directions along which that state's map
is smooth: x3-independent white noise in the wave's own fields u3 and B3.
Bar: 1e-6 relative to the largest directional derivative of the set.
"""
import math

import numpy as np
import pytest

from kharma_amd import problems as pr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

H = 1e-7
RTOL = 1e-6


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import kharma_amd
    kharma_amd.lib()


def smooth_direction(cfg, rng):
    from kharma_amd.adjoint import smooth_direction as sd
    return sd(cfg, rng)


def noise_direction(cfg, rng, fields):
    from kharma_amd.adjoint import noise_direction as nd
    return nd(cfg, rng, fields)


def _cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _embed_np(shape, cfg, Pi):
    P = np.zeros((pr.NV,) + shape)
    P[(slice(None),) + pr.interior(cfg)] = Pi
    return P


def _interior(cfg, P):
    return P[(slice(None),) + pr.interior(cfg)]


def _check(fd, ad, what):
    fd, ad = np.asarray(fd), np.asarray(ad)
    scale = np.max(np.abs(fd))
    err = np.abs(fd - ad) / scale
    print(f"{what}: max rel err {err.max():.3e} over {len(fd)} directions (scale {scale:.3e})", flush=True)
    assert err.max() < RTOL, (what, fd, ad)


CASES = {
    "alfven16": lambda: pr.mhdmode_config(n=16, ndim=3, mode="alfven", amp=1e-4, dt_start=0.02),
    "fast16": lambda: pr.mhdmode_config(n=16, ndim=3, mode="fast", amp=0.05, dt_start=0.02),
    "fast2d": lambda: pr.mhdmode_config(n=32, ndim=2, mode="fast", amp=0.05, dt_start=0.01),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_step_vs_fd(case):
    """One step: <lam, dP1/dP0 v> by central differences vs <adjoint(lam), v>."""
    from kharma_amd.adjoint import Evolution
    from kharma_amd.adjoint import generic_state
    cfg = CASES[case]()
    evo = Evolution(cfg)
    rng = np.random.default_rng(7)
    P0 = _cuda(generic_state(cfg, rng))
    _, dts, ck, _ = evo.run(P0, 1, checkpoints=True)
    lam = _cuda(rng.normal(size=P0.shape))
    g = evo.backward(lam, ck, dts).cpu().numpy()
    lam_np = lam.cpu().numpy()
    dirs = [smooth_direction(cfg, rng) for _ in range(3)] + [noise_direction(cfg, rng, range(pr.NV)) for _ in range(2)]
    fd, ad = [], []
    for v in dirs:
        Pp = evo.run(P0 + H * _cuda(v), 1, dts)[0]
        Pm = evo.run(P0 - H * _cuda(v), 1, dts)[0]
        fd.append(float((lam_np * (Pp - Pm).cpu().numpy()).sum() / (2 * H)))
        ad.append(float((g * v).sum()))
    _check(fd, ad, f"{case} one step")


def _xy_noise(cfg, rng, fields):
    """Unit-norm white noise in `fields`, independent of x3 (keeps an x3-uniform state x3-uniform)."""
    v = noise_direction(cfg, rng, fields)
    v[:] = v[:, :1]
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("state", ["generic", "alfven"])
def test_adjoint_evolve_loss(state):
    """The C4 objective on a 16^3 block, 6 steps: L = |P_N - P_target|^2 with the
    target the Alfven mode at phase + pi/4, gradient through torch.autograd."""
    from kharma_amd.adjoint import Evolution, evolve, generic_state
    cfg = pr.mhdmode_config(n=16, ndim=3, mode="alfven", amp=1e-4, dt_start=0.02)
    evo = Evolution(cfg)
    nsteps = 6
    rng = np.random.default_rng(11)
    P0n = generic_state(cfg, rng) if state == "generic" else _interior(cfg, pr.mhdmode_prims(cfg))
    P0 = _cuda(P0n).requires_grad_(True)
    Pt = _cuda(pr.mhdmode_state(cfg, 0.0, phase=math.pi / 4))
    PN = evolve(P0, evo, nsteps)
    loss = ((PN - Pt) ** 2).sum()
    loss.backward()
    g = P0.grad.cpu().numpy()
    dts = evo.run(P0.detach(), nsteps)[1]
    if state == "generic":
        dirs = [smooth_direction(cfg, rng) for _ in range(3)] + [noise_direction(cfg, rng, range(pr.NV)) for _ in range(2)]
    else:
        dirs = [_xy_noise(cfg, rng, (pr.U3, pr.B3)) for _ in range(5)]
    fd, ad = [], []
    for v in dirs:
        Lp = float(((evo.run(P0.detach() + H * _cuda(v), nsteps, dts)[0] - Pt) ** 2).sum())
        Lm = float(((evo.run(P0.detach() - H * _cuda(v), nsteps, dts)[0] - Pt) ** 2).sum())
        fd.append((Lp - Lm) / (2 * H))
        ad.append(float((g * v).sum()))
    _check(fd, ad, f"{state} 16^3 x 6 steps, |P_N - P_target|^2")


def test_adjoint_rejects_unsupported():
    import kharma_amd as kh
    cfg = pr.mhdmode_config(n=16, ndim=3, mode="alfven")
    with pytest.raises(kh.KharmaError):
        kh.AdjointStep(dict(cfg, recon="ppm"))
    with pytest.raises(kh.KharmaError):
        kh.AdjointStep(dict(cfg, riemann="hlle"))
    with pytest.raises(kh.KharmaError, match="onedw"):
        kh.AdjointStep(dict(cfg, inverter="kastaun"))
    tcfg = pr.torus_config(nx1=32, nx2=16, nx3=1)
    with pytest.raises(kh.KharmaError):
        kh.AdjointStep(dict(tcfg, floor_frame="normal"))


@pytest.mark.parametrize("case", ["fast16", "fast2d"])
def test_adjoint_recorded_stage_matches_recompute(case):
    """kh_adjoint_step_recorded (the step's stage intermediates recorded by the
    forward driver) against kh_adjoint_step (stages recomputed from the
    checkpoint P_n with U_n = p_to_u(P_n)).  At step 0 both start from the same
    U_0 and agree to rounding.  Later the forward's U_n comes out of the fused
    update, whose P->U is compiled with other FMA contractions than the
    stand-alone P->U, so the two stage-1 recoveries start a last bit apart and
    the onedw iteration (tolerance 1e-8) can end on slightly different P_1; the
    adjoints then differ by far less than the finite-difference bar."""
    from kharma_amd.adjoint import Evolution, generic_state
    cfg = CASES[case]()
    evo = Evolution(cfg)
    rng = np.random.default_rng(11)
    P0 = _cuda(generic_state(cfg, rng))
    _, dts, ck, _ = evo.run(P0, 3, checkpoints=True)
    recs = ck["records"]
    assert sorted(recs) == [0, 1, 2]
    lam = evo._embed(_cuda(rng.normal(size=P0.shape)))
    for n in (0, 1, 2):
        la, lb = lam.clone(), lam.clone()
        evo.adj.step(ck[n], dts[n], la, ck[n + 1])
        evo.adj.step_recorded(ck[n], recs[n], dts[n], lb, ck[n + 1])
        a, b = _interior(cfg, la).cpu().numpy(), _interior(cfg, lb).cpu().numpy()
        rel = np.abs(a - b).max() / np.abs(a).max()
        print(f"{case} step {n}: recorded vs recomputed stages, max rel diff {rel:.2e}", flush=True)
        assert rel <= (1e-13 if n == 0 else 1e-9), (n, rel)


def test_adjoint_against_tangent_linear_oracle(oracle_lib):
    """<lam, dP_N/dP_0 v> from the FP64 tangent-linear oracle (the CPU
    restatement on forward-mode dual numbers) against <adjoint(lam), v> from the
    GPU sweep, 3 steps of a smooth generic state with the timesteps fixed and
    the recovery converged to 1e-14: both are exact derivatives of the same map
    (up to the GPU/oracle per-step difference of ~1e-12), so they agree to 1e-8."""
    from kharma_amd.adjoint import Evolution, generic_state, smooth_direction
    cfg = pr.mhdmode_config(n=24, ndim=2, mode="fast", amp=0.05, dt_start=0.01, err_tol=1e-14, iter_max=30)
    rng = np.random.default_rng(23)
    Pi = generic_state(cfg, rng)
    dts = [0.01, 0.01, 0.01]
    evo = Evolution(cfg)
    _, _, ck, _ = evo.run(_cuda(Pi), 3, dts, checkpoints=True)
    lam = rng.normal(size=Pi.shape)
    g = evo.backward(_cuda(lam), ck, dts).cpu().numpy()
    _, shape = pr.block_shape(cfg)
    full = lambda a: pr.periodic_fill(_embed_np(shape, cfg, a), cfg)
    tb = oracle_lib.OracleBlock(cfg, variant="tangent")
    tb.init_state(full(Pi))
    for q in range(3):
        v = smooth_direction(cfg, rng)
        tb.init_state(full(Pi))
        tb.set_tangent(full(v), np.zeros((pr.NV,) + shape))
        tb.ptou()
        for n in range(3):
            tb.dt = dts[n]
            tb.step(1)
        tl = float((lam * _interior(cfg, tb.tangent()[0])).sum())
        ad = float((g * v).sum())
        rel = abs(tl - ad) / max(abs(tl), abs(ad))
        print(f"[tangent-linear] dir {q}: TL {tl:.15e} adjoint {ad:.15e} rel {rel:.2e}", flush=True)
        assert rel < 1e-8


# ------------------------------------------------------- task-level transposes
def _task_setup(case="fast16"):
    import kharma_amd as kh
    from kharma_amd.adjoint import generic_state
    cfg = CASES[case]()
    blk = kh.Block(dict(cfg, precision="fp64"))
    adj = kh.AdjointStep(cfg)
    ndim, shape = pr.block_shape(cfg)
    P = np.zeros((pr.NV,) + shape)
    P[(slice(None),) + pr.interior(cfg)] = generic_state(cfg, np.random.default_rng(1))
    P0 = _cuda(pr.periodic_fill(P, cfg))
    return cfg, blk, adj, P0


@pytest.mark.parametrize("case", ["fast16", "alfven16", "fast2d"])
def test_adjoint_flux_vs_fd(case):
    """lam_P = sum_d (dF_d/dP)^T lam_F_d against central differences of kh_get_flux, per direction."""
    cfg, blk, adj, P0 = _task_setup(case)
    ndim, _ = pr.block_shape(cfg)
    rng = np.random.default_rng(3)

    def fluxes(P):
        F, cmax, cmin = blk.new_fluxes()
        for d in range(1, ndim + 1):
            blk.get_flux(d, P, F, cmax, cmin)
        return F

    for d in range(1, ndim + 1):
        lamF = torch.zeros((3, pr.NV) + P0.shape[1:], dtype=torch.float64, device="cuda")
        lamF[d - 1] = _cuda(rng.normal(size=(pr.NV,) + P0.shape[1:]))
        lamP = torch.zeros_like(P0)
        adj.flux(P0, lamF, lamP)
        g = lamP.cpu().numpy()
        fd, ad = [], []
        for _ in range(3):
            v = rng.normal(size=P0.shape)
            v /= np.linalg.norm(v)
            dF = (fluxes(P0 + H * _cuda(v)) - fluxes(P0 - H * _cuda(v))) / (2 * H)
            fd.append(float((lamF * dF).sum()))
            ad.append(float((g * v).sum()))
        _check(fd, ad, f"{case} flux transpose, dir {d}")


@pytest.mark.parametrize("case", ["fast16", "fast2d"])
def test_adjoint_flux_div_dot(case):
    """<lam_d, Div(CT(F))> == <(dDiv CT/dF)^T lam_d, F> for random F (linear map)."""
    cfg, blk, adj, P0 = _task_setup(case)
    rng = np.random.default_rng(5)
    shape = P0.shape[1:]
    F = _cuda(rng.normal(size=(3, pr.NV) + shape))
    Fc = F.clone()
    emf = torch.zeros((3,) + shape, dtype=torch.float64, device="cuda")
    blk.flux_ct(Fc, emf)
    dudt = torch.zeros_like(P0)
    blk.flux_divergence(Fc, dudt)
    lam = torch.zeros_like(P0)
    inner = (slice(None),) + pr.interior(cfg)
    lam[inner] = _cuda(rng.normal(size=lam[inner].shape))
    lamF = torch.zeros_like(F)
    adj.flux_div(lam, lamF)
    lhs = float((lam * dudt).sum())
    ndim, _ = pr.block_shape(cfg)
    for v in range(pr.NV):
        a = float((lam[v] * dudt[v]).sum())
        b = float((lamF[:ndim, v] * F[:ndim, v]).sum())
        print(f"{case} var {v}: <lam, dudt> {a:.12e}  <lamF, F> {b:.12e}", flush=True)
    rhs = float((lamF[:ndim] * F[:ndim]).sum())
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("case", ["fast16", "alfven16"])
def test_adjoint_utop_ptou_vs_fd(case):
    """The inverter transpose (dU/dP)^-T and the P->U transpose against central differences."""
    cfg, blk, adj, P0 = _task_setup(case)
    rng = np.random.default_rng(9)
    inner = (slice(None),) + pr.interior(cfg)
    U0 = torch.zeros_like(P0)
    blk.ptou(P0, U0)
    pflag = torch.zeros(P0.shape[1:], dtype=torch.int32, device="cuda")
    fflag = torch.zeros_like(pflag)
    # P -> U
    lamU = torch.zeros_like(P0)
    lamU[inner] = _cuda(rng.normal(size=lamU[inner].shape))
    lamP = torch.zeros_like(P0)
    adj.ptou(P0, lamU, lamP)
    fd, ad = [], []
    for _ in range(3):
        v = np.zeros(P0.shape)
        v[inner] = rng.normal(size=v[inner].shape)
        Up, Um = torch.zeros_like(P0), torch.zeros_like(P0)
        blk.ptou(P0 + H * _cuda(v), Up)
        blk.ptou(P0 - H * _cuda(v), Um)
        fd.append(float((lamU * (Up - Um)).sum()) / (2 * H))
        ad.append(float((lamP.cpu().numpy() * v).sum()))
    _check(fd, ad, f"{case} P->U transpose")
    # U -> P (implicit-function theorem)
    lamP = torch.zeros_like(P0)
    lamP[inner] = _cuda(rng.normal(size=lamP[inner].shape))
    lamU = torch.zeros_like(P0)
    adj.utop(P0, lamP, lamU)
    hu = 1e-5
    fd, ad = [], []
    for _ in range(3):
        w = np.zeros(P0.shape)
        w[inner] = rng.normal(size=w[inner].shape)
        w /= np.linalg.norm(w)
        Pp, Pm = P0.clone(), P0.clone()
        blk.utop(U0 + hu * _cuda(w), Pp, pflag, fflag)
        blk.utop(U0 - hu * _cuda(w), Pm, pflag, fflag)
        fd.append(float((lamP[inner] * (Pp - Pm)[inner]).sum()) / (2 * hu))
        ad.append(float((lamU.cpu().numpy() * w).sum()))
    _check(fd, ad, f"{case} U->P transpose")
