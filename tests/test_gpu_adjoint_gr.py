"""Adjoint of the Kerr-Schild torus step (FMKS, outflow / reflecting boundaries
with the inflow check and polar flux fixes, drift-frame floors): each new
transpose against central differences of the forward entry point it reverses,
then whole steps and a multi-step face-flux objective with two-level
checkpointing against central differences of the driver.

The component states are the torus of the parity tests with a small smooth
perturbation of every variable (so that no field is uniform: WENO5's weights,
though differentiable there, have a second derivative of order 1/EPS at a
uniform field, which central differences cannot resolve), evolved a few steps so
that the atmosphere is floored and flowing.  Whole steps are checked on a
smooth state, the same torus with its atmosphere raised 50 % above the floors:
every zone that a stage floors leaves it at exactly the floor of its radius,
uniform in theta, which puts the X2 face states on the reconstruction floor's
kink (central differences there return the mean of the one-sided slopes); the
floors themselves are checked on their own.  The whole-step checks also run the
onedw recovery to tolerance 1e-14 (30 iterations): at its default 1e-8 the
recovered primitives of zones far from the recovery's guess carry errors up to
~1e-7 that change with the iteration count, which central differences of the
driver see as noise of order 1e-7 / h.  Directions are relative to each
variable's local magnitude.
"""
import math

import numpy as np
import pytest

from kharma_amd import problems as pr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import kharma_amd
    kharma_amd.lib()


def _cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=torch.float64)


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _modes(cfg, rng, n=3, kmax=3):
    """Sum of n random smooth modes over the block's interior cell centres (numpy)."""
    x, y, z = pr.cell_centers(cfg)
    X2, X1 = np.meshgrid(y, x, indexing="ij")
    L1 = cfg["bounds"][0][1] - cfg["bounds"][0][0]
    out = np.zeros((1,) + X1.shape)
    for _ in range(n):
        k1, k2 = rng.integers(1, kmax + 1, size=2)
        out[0] += rng.normal() * np.cos(2 * math.pi * k1 * X1 / L1 + rng.uniform(0, 2 * math.pi)) * \
            np.cos(math.pi * k2 * X2 + rng.uniform(0, 2 * math.pi))
    return out


def _scales(P):
    """Per-variable local magnitudes of an interior state (for relative directions)."""
    bmax = np.abs(P[pr.B1:]).max()
    s = np.empty_like(P)
    s[pr.RHO] = np.abs(P[pr.RHO])
    s[pr.UU] = np.abs(P[pr.UU])
    s[pr.U1:pr.B1] = np.abs(P[pr.U1:pr.B1]) + 0.1
    s[pr.B1:] = np.abs(P[pr.B1:]) + 1e-6 * bmax
    return s


def _direction(cfg, rng, P, vars_=range(pr.NV)):
    s = _scales(P)
    v = np.zeros_like(P)
    for q in vars_:
        v[q] = s[q] * _modes(cfg, rng)
    return v


CFG = lambda: pr.torus_config(64, 48, 1)


@pytest.fixture(scope="module")
def torus_state():
    """(cfg, Evolution, evolved interior state) shared by the tests."""
    from kharma_amd.adjoint import Evolution
    cfg = CFG()
    rng = np.random.default_rng(7)
    P = pr.torus_prims(cfg)[(slice(None),) + pr.interior(cfg)].copy()
    bmax = np.abs(P[pr.B1:]).max()
    P[pr.RHO] *= 1 + 1e-3 * _modes(cfg, rng)[0]
    P[pr.UU] *= 1 + 1e-3 * _modes(cfg, rng)[0]
    for q in (pr.U1, pr.U2, pr.U3):
        P[q] += 1e-3 * _modes(cfg, rng)[0]
    for q in (pr.B1, pr.B2, pr.B3):
        P[q] += 1e-6 * bmax * _modes(cfg, rng)[0]
    evo = Evolution(cfg)
    Ps = evo.run(_cuda(P), 4)[0]
    return cfg, evo, Ps


def _filled(blk, evo, Pint):
    """The full block of an interior state, ghosts filled by the physical boundaries."""
    Pf = evo._embed(Pint)
    U = torch.zeros_like(Pf)
    blk.ptou(Pf, U)
    blk.apply_bounds(U, Pf)
    return Pf


@pytest.fixture(scope="module")
def smooth_state():
    """(cfg with a converged recovery, Evolution, state): the perturbed torus with its
    atmosphere at 1.5 times its floors, so that no zone is floored in the steps."""
    from kharma_amd.adjoint import Evolution
    cfg = dict(CFG(), err_tol=1e-14, iter_max=30)
    rng = np.random.default_rng(8)
    P = pr.torus_prims(cfg)[(slice(None),) + pr.interior(cfg)].copy()
    bmax = np.abs(P[pr.B1:]).max()
    P[pr.RHO] *= 1 + 1e-3 * _modes(cfg, rng)[0]
    P[pr.UU] *= 1 + 1e-3 * _modes(cfg, rng)[0]
    for q in (pr.U1, pr.U2, pr.U3):
        P[q] += 1e-3 * _modes(cfg, rng)[0]
    for q in (pr.B1, pr.B2, pr.B3):
        P[q] += 1e-6 * bmax * _modes(cfg, rng)[0]
    atm = P[pr.RHO] < 1e-2
    P[pr.RHO][atm] *= 1.5
    P[pr.UU][atm] *= 1.5
    return cfg, Evolution(cfg), _cuda(P)


def _rel(fd, ad):
    return abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)


def test_geo_source_transpose(torus_state):
    import kharma_amd as kh
    cfg, evo, Ps = torus_state
    blk = kh.Block(cfg)
    P = evo._embed(Ps)
    rng = np.random.default_rng(11)
    inner = (slice(None),) + pr.interior(cfg)

    def S(Pf):
        d = torch.zeros_like(Pf)
        blk.add_geo_source(Pf, d)
        return d

    lam = torch.zeros_like(P)
    lam[inner] = _cuda(rng.normal(size=Ps.shape))
    out = torch.zeros_like(P)
    evo.adj.geo_source(P, lam, out)
    g = _np(out)
    for q in range(3):
        v = np.zeros(P.shape)
        v[inner] = _direction(cfg, rng, _np(Ps))
        h = 1e-6
        fd = float((lam * (S(P + h * _cuda(v)) - S(P - h * _cuda(v)))).sum()) / (2 * h)
        ad = float((g * v).sum())
        print(f"[gr geo source] dir {q}: fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
        assert _rel(fd, ad) < 1e-8


def test_floors_transpose(torus_state):
    """Drift-frame floors on a state whose atmosphere was pushed below the floors."""
    import kharma_amd as kh
    cfg, evo, Ps = torus_state
    blk = kh.Block(cfg)
    P0 = _np(evo._embed(Ps))
    atm = P0[pr.RHO] < 1e-2
    P0[pr.RHO][atm] *= 0.6
    P0[pr.UU][atm] *= 0.6
    P = _cuda(P0)
    inner = (slice(None),) + pr.interior(cfg)

    def F(Pf):
        Pf = Pf.clone()
        U = torch.zeros_like(Pf)
        blk.ptou(Pf, U)
        pf = torch.zeros(Pf.shape[1:], dtype=torch.int32, device="cuda")
        ff = torch.zeros_like(pf)
        blk.apply_floors(U, Pf, pf, ff)
        return Pf, ff

    _, ff = F(P)
    nfl = int((ff[inner[1:]] != 0).sum())
    print(f"[gr floors] {nfl} of {Ps[0].numel()} zones floored", flush=True)
    assert nfl > Ps[0].numel() // 4
    rng = np.random.default_rng(12)
    lam = torch.zeros_like(P)
    lam[inner] = _cuda(rng.normal(size=Ps.shape))
    out = torch.zeros_like(P)
    evo.adj.floors(P, lam, out)
    g = _np(out)
    for q in range(3):
        v = np.zeros(P.shape)
        v[inner] = _direction(cfg, rng, P0[inner])
        h = 1e-7
        fd = float((lam[inner] * (F(P + h * _cuda(v))[0] - F(P - h * _cuda(v))[0])[inner]).sum()) / (2 * h)
        ad = float((g * v).sum())
        print(f"[gr floors] dir {q}: fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
        assert _rel(fd, ad) < 1e-7


def test_boundary_fill_transpose(torus_state):
    """Outflow x1 faces with the inflow check, reflecting x2 faces, the inner corner fix."""
    import kharma_amd as kh
    cfg, evo, Ps = torus_state
    blk = kh.Block(cfg)
    inner = (slice(None),) + pr.interior(cfg)

    fill = lambda Pint: _filled(blk, evo, Pint)
    P = fill(Ps)
    rng = np.random.default_rng(13)
    lam = _cuda(rng.normal(size=P.shape))
    lam_in = lam.clone()
    evo.adj.bounds(P, lam_in)
    g = _np(lam_in)
    assert np.abs(g[:, :, :4, :]).max() == 0 and np.abs(g[:, :, :, :4]).max() == 0   # ghosts cleared
    for q in range(3):
        v = _direction(cfg, rng, _np(Ps))
        h = 1e-7
        fd = float((lam * (fill(Ps + h * _cuda(v)) - fill(Ps - h * _cuda(v)))).sum()) / (2 * h)
        ad = float((g[inner] * v).sum())
        print(f"[gr bounds] dir {q}: fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
        assert _rel(fd, ad) < 1e-8


def test_flux_fix_ct_divergence_transpose(torus_state):
    """div(CT(fix(F))) is positively homogeneous in F: <lam, D(F)> = <D^T lam, F> exactly."""
    import kharma_amd as kh
    cfg, evo, Ps = torus_state
    blk = kh.Block(cfg)
    P = _filled(blk, evo, Ps)
    F, cmax, cmin = blk.new_fluxes()
    for d in (1, 2):
        blk.get_flux(d, P, F, cmax, cmin)

    def D(Fr):
        Ff = Fr.clone()
        emf = torch.zeros((3,) + tuple(P.shape[1:]), dtype=torch.float64, device="cuda")
        blk.flux_ct(Ff, emf)
        dU = torch.zeros_like(P)
        blk.flux_divergence(Ff, dU)
        return dU, Ff

    dU, Ff = D(F)
    rng = np.random.default_rng(14)
    inner = (slice(None),) + pr.interior(cfg)
    lam = torch.zeros_like(P)
    lam[inner] = _cuda(rng.normal(size=Ps.shape))
    lF = torch.zeros_like(F)
    evo.adj.flux_div(lam, lF)
    evo.adj.fix_flux(Ff, lF)
    lhs = float((lam * dU).sum())
    rhs = float((lF * F).sum())
    print(f"[gr fix+ct+div] <lam, D F> {lhs:.15e}  <D^T lam, F> {rhs:.15e} rel {_rel(lhs, rhs):.2e}", flush=True)
    assert _rel(lhs, rhs) < 1e-12
    # and along a perturbation of the fluxes
    G = _cuda(rng.normal(size=F.shape)) * (F.abs() + 1e-12)
    h = 1e-8
    fd = float((lam * (D(F + h * G)[0] - D(F - h * G)[0])).sum()) / (2 * h)
    ad = float((lF * G).sum())
    print(f"[gr fix+ct+div] fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
    assert _rel(fd, ad) < 1e-8


def test_fixup_transpose(smooth_state):
    """Inversion fixup (kh_fix_utop): zones flagged as failed take the weighted
    average of their non-failed neighbours, then the geometric floors."""
    import kharma_amd as kh
    cfg, evo, Ps = smooth_state
    blk = kh.Block(cfg)
    P = _filled(blk, evo, Ps)
    rng = np.random.default_rng(18)
    inner = (slice(None),) + pr.interior(cfg)
    pf = np.zeros(P.shape[1:], dtype=np.int32)
    pf[inner[1:]] = (rng.uniform(size=Ps.shape[1:]) < 0.05).astype(np.int32)
    pflag = torch.as_tensor(pf, device="cuda")

    def F(Pq):
        Pq = Pq.clone()
        U = torch.zeros_like(Pq)
        blk.ptou(Pq, U)
        blk.fix_utop(U, Pq, pflag)
        return Pq

    Pfx = F(P)
    lam = torch.zeros_like(P)
    lam[inner] = _cuda(rng.normal(size=Ps.shape))
    lam_in = lam.clone()
    evo.adj.fixup(Pfx, pflag, lam_in)
    g = _np(lam_in)
    for q in range(3):
        v = np.zeros(P.shape)
        v[inner] = _direction(cfg, rng, _np(Ps))
        h = 1e-7
        fd = float((lam[inner] * (F(P + h * _cuda(v)) - F(P - h * _cuda(v)))[inner]).sum()) / (2 * h)
        ad = float((g[inner] * v[inner]).sum())
        print(f"[gr fixup] dir {q}: fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
        assert _rel(fd, ad) < 1e-8


def test_flux_transpose(smooth_state):
    """WENO5 + LLF face fluxes with the face metrics and reconstruction floors, all
    faces of both directions, ghost zones filled by the physical boundaries."""
    import kharma_amd as kh
    cfg, evo, Ps = smooth_state
    blk = kh.Block(cfg)
    P = _filled(blk, evo, Ps)

    def Fl(Pq):
        F, cmax, cmin = blk.new_fluxes()
        for d in (1, 2):
            blk.get_flux(d, Pq, F, cmax, cmin)
        return F

    F0 = Fl(P)
    rng = np.random.default_rng(17)
    lamF = _cuda(rng.normal(size=F0.shape)) * (F0 != 0) / (F0.abs() + 1e-30)
    lamP = torch.zeros_like(P)
    evo.adj.flux(P, lamF, lamP)
    g = _np(lamP)
    inner = (slice(None),) + pr.interior(cfg)
    for q in range(3):
        v = np.zeros(P.shape)
        v[inner] = _direction(cfg, rng, _np(Ps))
        h = 1e-8
        fd = float((lamF * (Fl(P + h * _cuda(v)) - Fl(P - h * _cuda(v)))).sum()) / (2 * h)
        ad = float((g * v).sum())
        print(f"[gr flux] dir {q}: fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
        assert _rel(fd, ad) < 1e-6


def test_step_against_finite_differences(smooth_state):
    cfg, evo, Ps = smooth_state
    rng = np.random.default_rng(15)
    _, dts, ck, _ = evo.run(Ps, 1, checkpoints=True)
    w = _cuda(rng.normal(size=Ps.shape) * _scales(_np(Ps)) ** -1)
    lam = evo.backward(w, ck, dts)
    g = _np(lam)
    L = lambda P: float((w * evo.run(P, 1, dts)[0]).sum())
    pf, ff = evo.drv.flags()
    assert int((ff != 0).sum()) == 0
    for q in range(3):
        v = _direction(cfg, rng, _np(Ps))
        h = 1e-6
        fd = (L(Ps + h * _cuda(v)) - L(Ps - h * _cuda(v))) / (2 * h)
        ad = float((g * v).sum())
        print(f"[gr step] dir {q}: fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
        assert _rel(fd, ad) < 1e-6


def test_floored_step_velocity_and_field_directions():
    """A step whose stages floor most of the atmosphere (drift frame), along
    directions in the velocities and the field, which reach the floored zones
    through the drift-frame floors.  Central differences of floored steps carry
    evaluation noise that grows as 1/h (the floored rows sit on ties: uniform in
    theta, on the reconstruction floor's kink, at zero WENO5 smoothness
    indicators; the drift frame is ill-conditioned where B crosses zero), so the
    check uses h = 1e-5; smaller steps are printed for the record.  The noise
    measured at h = 1e-6 and 1e-7 (up to 6e-5 and 2e-4 in a field direction)
    extrapolates as 1/h to about 6e-6 at h = 1e-5, and two builds whose forward
    steps differ only in rounding gave 2.6e-6 and 5.2e-6 there for the same
    direction, so the bound is 1.5e-5 (a wrong transpose term shows as O(1))."""
    from kharma_amd.adjoint import Evolution
    cfg = dict(CFG(), err_tol=1e-14, iter_max=30)
    rng = np.random.default_rng(9)
    P = pr.torus_prims(cfg)[(slice(None),) + pr.interior(cfg)].copy()
    bmax = np.abs(P[pr.B1:]).max()
    P[pr.RHO] *= 1 + 1e-3 * _modes(cfg, rng)[0]
    P[pr.UU] *= 1 + 1e-3 * _modes(cfg, rng)[0]
    for q in (pr.U1, pr.U2, pr.U3):
        P[q] += 1e-3 * _modes(cfg, rng)[0]
    for q in (pr.B1, pr.B2, pr.B3):
        P[q] += 1e-6 * bmax * _modes(cfg, rng)[0]
    evo = Evolution(cfg)
    Ps = evo.run(_cuda(P), 4)[0]
    _, dts, ck, _ = evo.run(Ps, 1, checkpoints=True)
    pf, ff = evo.drv.flags()
    nfl = int((ff != 0).sum())
    print(f"[gr floored step] {nfl} zones floored in stage 2", flush=True)
    assert nfl > Ps[0].numel() // 4
    w = _cuda(rng.normal(size=Ps.shape) * _scales(_np(Ps)) ** -1)
    lam = evo.backward(w, ck, dts)
    g = _np(lam)
    assert np.isfinite(g).all()
    L = lambda Pq: float((w * evo.run(Pq, 1, dts)[0]).sum())
    for name, vars_ in (("velocity", (pr.U1, pr.U2, pr.U3)), ("field", (pr.B1, pr.B2, pr.B3))):
        for q in range(2):
            v = _direction(cfg, rng, _np(Ps), vars_)
            ad = float((g * v).sum())
            errs = {}
            for h in (1e-5, 1e-6, 1e-7):
                fd = (L(Ps + h * _cuda(v)) - L(Ps - h * _cuda(v))) / (2 * h)
                errs[h] = _rel(fd, ad)
            print(f"[gr floored step] {name} dir {q}: adjoint {ad:.12e} rel err " +
                  " ".join(f"h {h:.0e}: {e:.2e}" for h, e in errs.items()), flush=True)
            assert errs[1e-5] < 1.5e-5


def test_recorded_steps_match_recompute(smooth_state):
    """kh_adjoint_step_recorded (each stage's U*, flags and clamped X1 fluxes
    recorded by the forward driver) against kh_adjoint_step (both stages
    recomputed from the checkpoint) on the torus.  The recorded stages are the
    forward trajectory's own; the recompute starts from U_n = p_to_u(P_n), a
    last bit away from the forward's U_n after step 0, and its recoveries
    converge to 1e-14, so the adjoints agree far inside the step bar."""
    cfg, evo, Ps = smooth_state
    _, dts, ck, _ = evo.run(Ps, 3, checkpoints=True)
    recs = ck["records"]
    rng = np.random.default_rng(21)
    lam = evo._embed(_cuda(rng.normal(size=tuple(Ps.shape))))
    for n in (0, 1, 2):
        la, lb = lam.clone(), lam.clone()
        evo.adj.step(ck[n], dts[n], la, ck[n + 1])
        evo.adj.step_recorded(ck[n], recs[n], dts[n], lb, ck[n + 1])
        a, b = _np(la[evo.inner]), _np(lb[evo.inner])
        rel = np.abs(a - b).max() / np.abs(a).max()
        print(f"[gr recorded] step {n}: max rel diff {rel:.2e}", flush=True)
        assert rel < 1e-9, (n, rel)


def test_step_against_tangent_linear_oracle(smooth_state, oracle_lib):
    """<lam, dP_2/dP_0 v> from the FP64 tangent-linear oracle (on the library's
    geometry cache) against <adjoint(lam), v> from the GPU sweep on the torus:
    2 steps of the smooth state (no zone floored, recovery to 1e-14), timesteps
    fixed, directions smooth in every variable and scaled per variable; the
    ghosts follow the interior through the boundary fill in both."""
    from tests.helpers import install_library_geometry
    cfg, evo, Ps = smooth_state
    nsteps = 2
    dts = evo.run(Ps, nsteps)[1]
    _, _, ck, _ = evo.run(Ps, nsteps, dts, checkpoints=True)
    rng = np.random.default_rng(29)
    lam = rng.normal(size=tuple(Ps.shape))
    g = _np(evo.backward(_cuda(lam), ck, dts))
    P0 = _np(evo._embed(Ps))
    tb = oracle_lib.OracleBlock(cfg, variant="tangent")
    install_library_geometry(tb, cfg)
    for q in range(3):
        v = _direction(cfg, rng, _np(Ps))
        tb.init_state(P0)
        tb.set_tangent(_np(evo._embed(_cuda(v))), np.zeros_like(P0))
        tb.ptou()
        tb.apply_bcs()
        for n in range(nsteps):
            tb.dt = dts[n]
            tb.step(1)
        tl = float((lam * tb.tangent()[0][evo.inner]).sum())
        ad = float((g * v).sum())
        print(f"[gr tangent-linear] dir {q}: TL {tl:.15e} adjoint {ad:.15e} rel {_rel(tl, ad):.2e}", flush=True)
        assert _rel(tl, ad) < 1e-8


def test_mass_flux_objective_two_level_checkpoints(smooth_state):
    """J = time-integrated mass flux through the face nearest r = 2.5 over 6 steps,
    gradient with checkpoints every 3 steps (segments re-run) and every step."""
    from kharma_amd.adjoint import Evolution, evolve_objective, face_index
    cfg, _, Ps = smooth_state
    face = face_index(cfg, 1, math.log(2.5))
    evo = Evolution(cfg, objective=dict(dir=1, var=pr.RHO, face=face, weight=1.0))
    nsteps = 6
    dts = evo.run(Ps, nsteps)[1]
    grads = {}
    for every in (3, 1):
        P0 = Ps.clone().requires_grad_(True)
        _, J = evolve_objective(P0, evo, nsteps, dts, every=every)
        J.backward()
        grads[every] = _np(P0.grad)
        print(f"[gr objective] every {every}: J = {float(J.detach()):.15e}", flush=True)
    dif = np.abs(grads[3] - grads[1]).max() / np.abs(grads[1]).max()
    print(f"[gr objective] max |g_3 - g_1| / max |g_1| = {dif:.2e}", flush=True)
    assert dif < 1e-10
    rng = np.random.default_rng(16)
    Jf = lambda P: evo.run(P, nsteps, dts)[3]
    for q in range(3):
        v = _direction(cfg, rng, _np(Ps))
        h = 1e-7
        fd = (Jf(Ps + h * _cuda(v)) - Jf(Ps - h * _cuda(v))) / (2 * h)
        ad = float((grads[3] * v).sum())
        print(f"[gr objective] dir {q}: fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
        assert _rel(fd, ad) < 1e-6


def test_step_3d_torus_periodic_x3():
    """3-D torus block (x3 periodic with ghost images, x1 / x2 physical): the
    physical-boundary transpose composed with the periodic fold, on a smooth
    state (atmosphere above its floors, recovery to 1e-12, which every zone of
    this block reaches in 30 iterations).  The directions are in rho and the
    velocities and leave out the three zones next to each x1 face.  Next to
    the outflow faces the ghosts copy the edge zone, so WENO5 sub-stencils
    reaching into them are exactly uniform, where central differences cannot
    resolve the derivative (as for any uniform field).  Along B^1 and B^2 the
    stage's right-hand side is so strongly curved at the edge of the field loop
    that central differences of it approach the transpose only for h <~ 1e-9
    (a relative gap of 3e-1, 3e-2, 4e-4 at h = 1e-7, 1e-8, 1e-9 in the worst
    slab), below the noise floor of a whole step."""
    from kharma_amd.adjoint import Evolution
    cfg = dict(pr.torus_config(32, 24, 8), err_tol=1e-12, iter_max=30)
    rng = np.random.default_rng(19)
    P = pr.torus_prims(cfg)[(slice(None),) + pr.interior(cfg)].copy()
    x, y, z = pr.cell_centers(cfg)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    L1 = cfg["bounds"][0][1] - cfg["bounds"][0][0]

    def mode3():
        k1, k2, k3 = rng.integers(1, 3, size=3)
        return np.cos(2 * math.pi * k1 * X / L1 + rng.uniform(0, 6.3)) * np.cos(math.pi * k2 * Y + rng.uniform(0, 6.3)) * \
            np.cos(k3 * Z + rng.uniform(0, 6.3))

    bmax = np.abs(P[pr.B1:]).max()
    P[pr.RHO] *= 1 + 1e-3 * mode3()
    P[pr.UU] *= 1 + 1e-3 * mode3()
    for q in (pr.U1, pr.U2, pr.U3):
        P[q] += 1e-3 * mode3()
    for q in (pr.B1, pr.B2, pr.B3):
        P[q] += 1e-6 * bmax * mode3()
    atm = P[pr.RHO] < 1e-2
    P[pr.RHO][atm] *= 1.5
    P[pr.UU][atm] *= 1.5
    evo = Evolution(cfg)
    Ps = _cuda(P)
    _, dts, ck, _ = evo.run(Ps, 1, checkpoints=True)
    pf, ff = evo.drv.flags()
    assert int((ff != 0).sum()) == 0
    w = _cuda(rng.normal(size=P.shape) * _scales(P) ** -1)
    g = _np(evo.backward(w, ck, dts))
    L = lambda Pq: float((w * evo.run(Pq, 1, dts)[0]).sum())
    s_ = _scales(P)
    away = np.zeros(P.shape[1:], bool)
    away[:, :, 3:-3] = True
    for q in range(2):
        v = np.zeros_like(P)
        for var in (pr.RHO, pr.U1, pr.U2, pr.U3):
            v[var] = s_[var] * mode3() * away
        h = 1e-6
        fd = (L(Ps + h * _cuda(v)) - L(Ps - h * _cuda(v))) / (2 * h)
        ad = float((g * v).sum())
        print(f"[gr 3d step] dir {q}: fd {fd:.12e} adjoint {ad:.12e} rel {_rel(fd, ad):.2e}", flush=True)
        assert _rel(fd, ad) < 1e-5
