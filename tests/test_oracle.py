"""Pins the CPU oracle (oracle/kharma_oracle.cpp) with checks that do not
restate its formulas: closed forms, invariants and the reference's own
known-answer data (the mhdmodes eigenmodes and convergence bounds of
tests/mhdmodes/check.py:48-147 in the reference)."""
import math

import numpy as np
import pytest

from kharma_amd import problems as pr
from tests.helpers import interior_of

RHO, UU, U1, U2, U3, B1, B2, B3 = range(8)


# ------------------------------------------------------------ reconstruction
@pytest.mark.parametrize("coef", [(1.0, 0.0, 0.0), (0.3, -1.2, 0.0), (2.0, 0.7, -0.4), (-1.0, 3.0, 5.0)])
def test_weno5_exact_on_quadratics(oracle_lib, coef):
    # Every 3-point sub-stencil is the quadratic Lagrange interpolant of point
    # values at cell centres, so any convex combination reproduces q(+-1/2).
    a, b, c = coef
    q = lambda x: a + b * x + c * x * x
    x = [q(i) for i in (-2, -1, 0, 1, 2)]
    l, r = oracle_lib.weno5(x)
    assert abs(l - q(-0.5)) < 1e-12 * (1 + abs(q(-0.5)))
    assert abs(r - q(0.5)) < 1e-12 * (1 + abs(q(0.5)))


def test_weno5_mirror_symmetry(oracle_lib):
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = rng.normal(size=5)
        l, r = oracle_lib.weno5(x)
        l2, r2 = oracle_lib.weno5(x[::-1])
        assert abs(l - r2) <= 1e-14 * (1 + abs(l))
        assert abs(r - l2) <= 1e-14 * (1 + abs(r))


def test_weno5_linear_weights_smooth_limit(oracle_lib):
    # For a smooth function the nonlinear weights tend to (1/16, 5/8, 5/16),
    # the 5th-order combination: interpolation error must fall ~h^5 (S11 eq. 9).
    errs = []
    hs = [0.1, 0.05, 0.025]
    for h in hs:
        x = [math.sin(1.0 + i * h) for i in (-2, -1, 0, 1, 2)]
        _, r = oracle_lib.weno5(x)
        errs.append(abs(r - math.sin(1.0 + 0.5 * h)))
    slope = np.polyfit(np.log(hs), np.log(errs), 1)[0]
    assert slope > 4.0


def test_linear_mc_and_ppm_exact_on_linear(oracle_lib):
    x = [1.0 + 0.25 * i for i in (-2, -1, 0, 1, 2)]
    for name in ("linear_mc", "ppm"):
        l, r = oracle_lib.reconstruct(name, x)
        assert abs(l - 0.875) < 1e-14 and abs(r - 1.125) < 1e-14, name
    # MC limiter clips to zero slope at an extremum
    l, r = oracle_lib.reconstruct("linear_mc", [0.0, 1.0, 2.0, 1.0, 0.0])
    assert l == 2.0 and r == 2.0


# ------------------------------------------------------------------ physics
def _embed(cfg, Pi):
    """A full block array with the interior Pi (ghosts zero)."""
    _, shape = pr.block_shape(cfg)
    P = np.zeros((pr.NV,) + shape)
    P[(slice(None),) + pr.interior(cfg)] = Pi
    return P


def _block(oracle_lib, cfg):
    return oracle_lib.OracleBlock(cfg)


def _random_prims(cfg, rng, lo=0.5, hi=1.5):
    ndim, shape = pr.block_shape(cfg)
    P = np.empty((8,) + shape)
    P[RHO] = rng.uniform(lo, hi, size=shape)
    P[UU] = rng.uniform(0.2, 1.0, size=shape)
    P[U1:U3 + 1] = rng.uniform(-0.5, 0.5, size=(3,) + shape)
    P[B1:B3 + 1] = rng.uniform(-0.7, 0.7, size=(3,) + shape)
    return P


def _srmhd_conserved(P, gam):
    """Textbook SR-MHD conserved variables (e.g. Mignone & McKinney 2007, eqs. 1-5),
    in KHARMA's form: U_rho = D, U_u = D - E, U_i = S_i, U_B = B."""
    rho, u = P[RHO], P[UU]
    ut = P[U1:U3 + 1]
    B = P[B1:B3 + 1]
    W = np.sqrt(1.0 + (ut * ut).sum(0))       # Lorentz factor
    v = ut / W
    p = (gam - 1.0) * u
    w = rho + u + p
    vB = (v * B).sum(0)
    B2 = (B * B).sum(0)
    v2 = (v * v).sum(0)
    D = rho * W
    S = (w * W * W + B2) * v - vB * B
    E = w * W * W - p + 0.5 * B2 + 0.5 * (v2 * B2 - vB * vB)
    return np.concatenate([D[None], (D - E)[None], S, B])


def test_ptou_matches_textbook_srmhd(oracle_lib):
    cfg = pr.mhdmode_config(n=8, ndim=3)
    rng = np.random.default_rng(11)
    P = _random_prims(cfg, rng)
    b = _block(oracle_lib, cfg)
    b.set_state(P, np.zeros_like(P))
    b.ptou()
    _, U = b.state()
    ref = _srmhd_conserved(P, cfg["gamma"])
    assert np.max(np.abs(U - ref) / (1 + np.abs(ref))) < 1e-13


def test_onedw_round_trip(oracle_lib):
    # u_to_p(p_to_u(P)) = P to the inverter tolerance, from a perturbed guess
    cfg = pr.mhdmode_config(n=8, ndim=3)
    rng = np.random.default_rng(5)
    P = _random_prims(cfg, rng)
    b = _block(oracle_lib, cfg)
    b.set_state(P, np.zeros_like(P))
    b.ptou()
    _, U = b.state()
    guess = P * (1.0 + 1e-3 * rng.standard_normal(P.shape))
    b.set_state(guess, U)
    b.utop()
    Pr, _ = b.state()
    assert np.all(b.array(4) == 0)
    assert np.max(np.abs(Pr[:5] - P[:5]) / np.abs(P[:5]).clip(1e-3)) < 1e-7
    assert np.array_equal(Pr[5:], U[5:])  # gdet = 1


@pytest.mark.parametrize("metric", ["minkowski", "ks"])
def test_kastaun_round_trip(oracle_lib, metric):
    """Inverter::u_to_p<kastaun> (inverter/kastaun.hpp:83-400): u_to_p(p_to_u(P)) = P
    to the solver's tolerance (1e-12 on mu) from no guess at all, in Minkowski and on
    the Kerr-Schild metric of the torus block (the normal-frame transform and the
    spatial inverse metric g^ij - g^0i g^0j / g^00 enter there), magnetised states."""
    if metric == "minkowski":
        cfg = pr.mhdmode_config(n=8, ndim=3, inverter="kastaun")
    else:
        cfg = dict(pr.torus_config(16, 12, 1), inverter="kastaun", err_tol=1e-12, iter_max=25)
    rng = np.random.default_rng(9)
    if metric == "ks":
        # the torus's own state: random native-coordinate values are unphysical at
        # large radius in log-r coordinates, where the inverter's ceilings would act
        P = pr.torus_prims(cfg)
    else:
        P = _random_prims(cfg, rng)
    b = _block(oracle_lib, cfg)
    b.set_state(P, np.zeros_like(P))
    b.ptou()
    _, U = b.state()
    b.set_state(np.zeros_like(P), U)
    b.utop()
    Pr, _ = b.state()
    sl = (slice(None),) + pr.interior(cfg)
    assert np.all(b.array(4)[sl[1:]] == 0) and np.all(b.array(5)[sl[1:]] == 0)
    # rho and u relative, velocities absolute: in the torus's magnetised, nearly
    # cold atmosphere P -> U -> P itself loses ~1e-9 of u to roundoff (onedw from
    # the exact answer loses the same), whatever the solver
    rel = np.abs(Pr[:2][sl] - P[:2][sl]) / np.abs(P[:2][sl])
    vel = np.abs(Pr[U1:U3 + 1][sl] - P[U1:U3 + 1][sl])
    assert rel.max() < (1e-12 if metric == "minkowski" else 1e-8), rel.max()
    assert vel.max() < (1e-12 if metric == "minkowski" else 1e-9), vel.max()
    # the same solution as onedw converged to 1e-14 from the exact guess
    o = _block(oracle_lib, dict(cfg, inverter="onedw", err_tol=1e-14, iter_max=30))
    o.set_state(P, U)
    o.utop()
    Po, _ = o.state()
    assert np.abs(Pr[:2][sl] - Po[:2][sl]).max() <= 1e-12 * np.abs(Po[:2][sl]).max()
    assert np.abs(Pr[U1:U3 + 1][sl] - Po[U1:U3 + 1][sl]).max() < 1e-11


def test_kastaun_inverter_density_floor(oracle_lib):
    """The density floor inside Kastaun's master function (kastaun.hpp:126-136):
    a zone whose D / W falls below rho_min_const is recovered at the floor and
    flagged INVERTER_RHO (16384) in fflag, with pflag 0."""
    cfg = pr.mhdmode_config(n=8, ndim=3, inverter="kastaun", floor_values=dict(rho_min_const=1e-3))
    rng = np.random.default_rng(4)
    P = _random_prims(cfg, rng)
    sl = pr.interior(cfg)
    z = tuple(s_.start for s_ in sl)
    P[(RHO,) + z] = 1e-5
    b = _block(oracle_lib, cfg)
    b.set_state(P, np.zeros_like(P))
    b.ptou()
    _, U = b.state()
    b.set_state(np.zeros_like(P), U)
    b.utop()
    Pr, _ = b.state()
    assert b.array(4)[z] == 0 and int(b.array(5)[z]) & 16384
    assert Pr[(RHO,) + z] == 1e-3


def test_onedw_failure_codes(oracle_lib):
    cfg = pr.mhdmode_config(n=8, ndim=3)
    P = pr.mhdmode_prims(cfg)
    b = _block(oracle_lib, cfg)
    b.set_state(P, np.zeros_like(P))
    b.ptou()
    Pg, U = b.state()
    U[RHO, 5, 5, 5] = -1.0            # neg_input (onedw.hpp:99-101)
    b.set_state(P, U)
    b.utop()
    pf = b.array(4)
    assert pf[5, 5, 5] == 1 and np.count_nonzero(pf) == 1
    cfg2 = dict(cfg, iter_max=0)      # max_iter unless the Halley step already converged
    b2 = _block(oracle_lib, cfg2)
    U2 = U.copy(); U2[RHO, 5, 5, 5] = 1.0
    b2.set_state(P * 1.05, U2)
    b2.utop()
    assert np.all(b2.array(4)[pr.interior(cfg)] == 2)


def test_static_fast_speed(oracle_lib):
    # Static uniform plasma: the signal speeds along each axis equal the fast
    # magnetosonic speed c_ms^2 = cs^2 + va^2 - cs^2 va^2, with
    # cs^2 = gam (gam-1) u / (rho + gam u), va^2 = B^2 / (B^2 + rho + gam u).
    cfg = pr.mhdmode_config(n=8, ndim=3)
    gam = cfg["gamma"]
    ndim, shape = pr.block_shape(cfg)
    P = np.zeros((8,) + shape)
    rho, u, B = 1.3, 0.7, (0.4, -0.2, 0.1)
    P[RHO], P[UU] = rho, u
    P[B1], P[B2], P[B3] = B
    b = _block(oracle_lib, cfg)
    b.set_state(P, np.zeros_like(P))
    for d in (1, 2, 3):
        b.get_flux(d)
    cs2 = gam * (gam - 1) * u / (rho + gam * u)
    Bsq = sum(x * x for x in B)
    va2 = Bsq / (Bsq + rho + gam * u)
    cms = math.sqrt(cs2 + va2 - cs2 * va2)
    sl = (slice(None),) + pr.interior(cfg)
    assert np.allclose(b.array(1)[sl], cms, rtol=1e-14)
    assert np.allclose(b.array(2)[sl], cms, rtol=1e-14)
    # LLF of a uniform state is the physical flux: mass flux 0, momentum flux
    # p + B^2/2 - B_d^2 on the diagonal (static fluid).
    F = b.array(0)
    p = (gam - 1) * u
    for d in (1, 2, 3):
        assert np.allclose(F[d - 1, RHO][pr.interior(cfg)], 0.0, atol=1e-15)
        assert np.allclose(F[d - 1, U1 + d - 1][pr.interior(cfg)], p + 0.5 * Bsq - B[d - 1] ** 2, rtol=1e-13)


@pytest.mark.parametrize("ndim", [2, 3])
def test_uniform_state_preserved(oracle_lib, ndim):
    cfg = pr.mhdmode_config(n=16 if ndim == 2 else 8, ndim=ndim)
    _, shape = pr.block_shape(cfg)
    P = np.zeros((8,) + shape)
    P[RHO], P[UU] = 1.0, 0.5
    P[U1], P[U2], P[U3] = 0.1, -0.2, 0.3
    P[B1], P[B2], P[B3] = 0.3, 0.1, -0.2
    b = _block(oracle_lib, cfg)
    b.init_state(P)
    b.step(3)
    Pn, _ = b.state()
    assert np.max(np.abs(Pn - P)) < 1e-13


def test_divb_roundoff_orszag_tang(oracle_lib):
    cfg = pr.orszag_tang_config(n=64)
    b = _block(oracle_lib, cfg)
    b.init_state(pr.orszag_tang_prims(cfg))
    assert b.max_divb() < 1e-14
    b.step(25)
    assert b.max_divb() < 1e-12      # reference torus_sanity bound is 1e-10


def test_orszag_tang_reference_generator(oracle_lib):
    """KHARMA's own Orszag-Tang set-up (prob/orszag_tang.hpp:25-62, B the
    averaged curl of seed_a<orszag_tang_a>): A3 = f(x) + g(y) is separable, so
    the averaged curl is exact differences of f and g across the zone,
    B1 = -ts sin(y + phase) sin(dy/2)/(dy/2), B2 = -ts sin(2x + phase) sin(dx)/dx,
    and the corner divB vanishes to roundoff through steps."""
    cfg = pr.orszag_tang_config(n=48, init="reference")
    P = pr.orszag_tang_prims(cfg)
    ts, ph, gam = cfg["ot_tscale"], cfg["ot_phase"], cfg["gamma"]
    x, y, _ = pr.cell_centers(cfg)
    X, Y = np.meshgrid(x, y, indexing="xy")
    dx = dy = 2 * math.pi / 48
    Pi = interior_of(P, cfg)[:, 0]
    assert np.allclose(Pi[UU], 5.0 / (3.0 * (gam - 1.0)) * ts * ts, rtol=1e-15, atol=0)
    assert np.allclose(Pi[U1], -ts * np.sin(Y + ph), rtol=0, atol=1e-16)
    assert np.allclose(Pi[U2], ts * np.sin(X + ph), rtol=0, atol=1e-16)
    assert np.abs(Pi[B1] - (-ts * np.sin(Y + ph) * np.sin(dy / 2) / (dy / 2))).max() < 1e-15
    assert np.abs(Pi[B2] - (-ts * np.sin(2 * X + ph) * np.sin(dx) / dx)).max() < 1e-15
    b = _block(oracle_lib, cfg)
    b.init_state(P)
    assert b.max_divb() < 1e-15
    b.step(10)
    assert b.max_divb() < 1e-13


def test_dt_ramp(oracle_lib):
    # First step takes dt_start = dt_min = 1e-5; dt then at most doubles per step
    # (grmhd.cpp:88-95,335-339; kharma_driver.cpp:442-448).
    cfg = pr.orszag_tang_config(n=32)
    b = _block(oracle_lib, cfg)
    b.init_state(pr.orszag_tang_prims(cfg))
    dts = b.step(20)
    assert dts[0] == 1e-5
    for a, c in zip(dts[:-1], dts[1:]):
        assert c <= 2 * a * (1 + 1e-15)
    assert np.allclose(dts[:10], 1e-5 * 2.0 ** np.arange(10), rtol=0)
    assert dts[-1] < 2 * dts[-2]     # ramp has saturated at the CFL limit


def test_fixup_weighted_average(oracle_lib):
    # A failed zone takes the 1/(|l|+|m|+|n|+1)-weighted mean of its 26
    # non-failed neighbours (inverter/fixup.cpp:87-104).
    cfg = pr.mhdmode_config(n=8, ndim=3, mode="fast", amp=0.05)
    P = pr.mhdmode_prims(cfg)
    b = _block(oracle_lib, cfg)
    b.set_state(P, np.zeros_like(P))
    b.ptou()
    _, U = b.state()
    U[RHO, 6, 5, 7] = -1.0
    b.set_state(P, U)
    b.utop()
    assert b.array(4)[6, 5, 7] == 1
    b.fix_utop()
    Pf, _ = b.state()
    num = np.zeros(5)
    den = 0.0
    for n in (-1, 0, 1):
        for m in (-1, 0, 1):
            for l in (-1, 0, 1):
                if n == m == l == 0:
                    continue
                w = 1.0 / (abs(l) + abs(m) + abs(n) + 1)
                num += w * P[:5, 6 + n, 5 + m, 7 + l]
                den += w
    assert np.allclose(Pf[:5, 6, 5, 7], num / den, rtol=1e-14)


def test_drift_floor_keeps_momentum_along_field(oracle_lib):
    # Drift-frame injection (R17 App. B3) raises rho to the floor and keeps the
    # normal-observer momentum projected on B fixed.
    cfg = pr.mhdmode_config(n=8, ndim=3)
    gam = cfg["gamma"]
    rng = np.random.default_rng(2)
    P = _random_prims(cfg, rng)
    P[RHO, 6, 6, 6] = 1e-11
    P[UU, 6, 6, 6] = 1e-3
    b = _block(oracle_lib, cfg)
    b.set_state(P, np.zeros_like(P))
    b.ptou()
    _, U = b.state()
    b.set_state(P, U)
    b.apply_floors()
    Pn, _ = b.state()
    assert Pn[RHO, 6, 6, 6] == 1e-8
    assert int(b.array(5)[6, 6, 6]) & 32

    def qdotb(Pz):
        ut = Pz[U1:U3 + 1]
        W = math.sqrt(1 + (ut * ut).sum())
        w = Pz[RHO] + gam * Pz[UU]
        return w * W * float((ut * Pz[B1:B3 + 1]).sum())   # w u^0 (u_i B^i), flat space

    assert abs(qdotb(Pn[:, 6, 6, 6]) - qdotb(P[:, 6, 6, 6])) < 1e-12 * (1 + abs(qdotb(P[:, 6, 6, 6])))
    assert np.array_equal(Pn[:, 5, 5, 5], P[:, 5, 5, 5])   # unfloored zones untouched


# ------------------------------------------------------ whole-step behaviour
def _run_mode(oracle_lib, n, mode, ndim=2):
    cfg = pr.mhdmode_config(n=n, mode=mode, ndim=ndim)
    b = _block(oracle_lib, cfg)
    b.init_state(pr.mhdmode_prims(cfg))
    steps = 0
    while b.time < cfg["tlim"] * (1 - 1e-12):
        b.step(1)
        steps += 1
        assert steps < 10000
    P, _ = b.state()
    sol = pr.mhdmode_state(cfg, 0.0)
    return np.abs(interior_of(P, cfg) - sol).mean(axis=tuple(range(1, 1 + 3)))


@pytest.mark.parametrize("mode", ["fast", "alfven", "slow"])
def test_mhdmode_convergence_bounds(oracle_lib, mode):
    """Slope bounds of the reference's mhdmodes check (tests/mhdmodes/check.py:130-147)."""
    res = [16, 24, 32]
    L1 = np.array([_run_mode(oracle_lib, n, mode) for n in res])
    ev = pr.MODES_DIR3[mode]
    keys = ["drho", "du", "du1", "du2", "du3", "dB1", "dB2", "dB3"]
    for v, key in enumerate(keys):
        if ev.get(key, 0.0) == 0.0:
            continue
        slope = np.polyfit(np.log(res), np.log(L1[:, v]), 1)[0]
        lo = -2.3 if v in (U2, U3) else -2.15
        assert lo <= slope <= -1.94, (mode, key, slope, L1[:, v])


# ------------------------------------------------------ driver/type = simple
def _both_orderings(oracle_lib, cfg, P0, nsteps):
    out = {}
    for drv in ("kharma", "simple"):
        b = _block(oracle_lib, dict(cfg, driver=drv))
        b.init_state(P0)
        b.step(nsteps)
        out[drv] = b.state()
    return out


def test_simple_ordering_periodic_block_bitwise(oracle_lib):
    """driver/type simple (driver/simple_step.cpp:39-168: inversion and floors on
    the interior, exchange of the primitives, then the fix-up) against the default
    ordering (driver/kharma_step.cpp:77-319) on a periodic block with failed
    inversions and floored zones.  Both invert, floor and fix the same zones from
    the same values, and every ghost is an exact periodic image, so the whole
    state agrees bit for bit."""
    cfg = pr.mhdmode_config(n=16, mode="fast", ndim=3, amp=1e-1)
    P0 = pr.mhdmode_prims(cfg)
    rng = np.random.default_rng(3)
    Pi = P0[(slice(None),) + pr.interior(cfg)]
    for k, j, i in rng.integers(0, 16, size=(12, 3)):
        Pi[U1:U3 + 1, k, j, i] = 30.0 * rng.normal(size=3)   # near-luminal, light zones
        Pi[RHO, k, j, i] = 1e-3
    P0[(slice(None),) + pr.interior(cfg)] = Pi
    P0 = pr.periodic_fill(P0, cfg)
    b = _block(oracle_lib, cfg)
    b.init_state(P0)
    fails = floored = 0
    for _ in range(6):
        b.step(1)
        fails += int((interior_of(b.array(4)[None], cfg) > 0).sum())
        floored += int((interior_of(b.array(5)[None], cfg) > 0).sum())
    assert fails > 0 and floored > 0                     # fix-ups and floors act
    out = _both_orderings(oracle_lib, cfg, P0, 6)
    for a, s in zip(out["kharma"], out["simple"]):
        assert np.array_equal(a, s)


def test_simple_ordering_torus_physical_boundaries(oracle_lib):
    """The same comparison on a torus block (outflow x1 with the inflow check,
    reflecting x2 with the polar fixes, floored atmosphere).  The interiors agree;
    in the ghost zones the simple fill recomputes U.B from P.B = U.B / sqrt(-g)
    (boundaries.cpp:561-571), which moves it by at most an ulp or two."""
    cfg = pr.torus_config(48, 32, 1)
    P0 = pr.torus_prims(cfg)
    out = _both_orderings(oracle_lib, cfg, P0, 10)
    (Pk, Uk), (Ps, Us) = out["kharma"], out["simple"]
    sl = (slice(None),) + pr.interior(cfg)
    assert np.array_equal(Pk[sl], Ps[sl]) and np.array_equal(Uk[sl], Us[sl])
    dB = np.abs(Uk[B1:] - Us[B1:])
    assert dB.max() > 0                                   # the fills do differ ...
    assert np.all(dB <= 4.5e-16 * np.abs(Uk[B1:]))        # ... by rounding only
    assert np.all(np.abs(Uk[:B1] - Us[:B1]) <= 4.5e-16 * np.abs(Uk[:B1]))
    assert np.all(np.abs(Pk - Ps) <= 4.5e-16 * np.abs(Pk))


# ------------------------------------------------------ tangent-linear build
def test_tangent_linear_matches_central_differences(oracle_lib):
    """The oracle compiled on forward-mode dual numbers (liboracle_tangent.so)
    gives the exact directional derivative of three FP64 steps (timesteps held
    fixed): central differences of the plain oracle converge to it at O(h^2)
    (error ratio 1e-2 per decade of h) down to 1e-7 relative.  A smooth generic
    state (no field uniform along an axis) and a smooth direction; recovery to
    1e-14.  At the initial state the tangent of P->U matches to roundoff."""
    from kharma_amd.adjoint import generic_state, smooth_direction
    cfg = pr.mhdmode_config(n=24, ndim=2, mode="fast", amp=0.05, dt_start=0.01, err_tol=1e-14, iter_max=30)
    rng = np.random.default_rng(5)
    full = lambda Pi: pr.periodic_fill(_embed(cfg, Pi), cfg)
    P0 = full(generic_state(cfg, rng))
    v = full(smooth_direction(cfg, rng))
    dts = [0.01, 0.01, 0.01]

    def tangent(nsteps):
        tb = oracle_lib.OracleBlock(cfg, variant="tangent")
        tb.init_state(P0)
        tb.set_tangent(v, np.zeros_like(v))   # the direction in P, carried into U = p_to_u(P)
        tb.ptou()
        for n in range(nsteps):
            tb.dt = dts[n]
            tb.step(1)
        return tb.tangent()

    def central(nsteps, h):
        out = []
        for sgn in (1, -1):
            q = _block(oracle_lib, cfg)
            q.init_state(P0 + sgn * h * v)
            for n in range(nsteps):
                q.dt = dts[n]
                q.step(1)
            out.append(q.state())
        return [(a - b) / (2 * h) for a, b in zip(out[0], out[1])]

    rel = lambda a, b: np.abs(interior_of(a - b, cfg)).max() / np.abs(interior_of(b, cfg)).max()
    dP0, dU0 = tangent(0)
    fP0, fU0 = central(0, 1e-4)
    assert rel(fP0, dP0) < 1e-10 and rel(fU0, dU0) < 1e-9
    dP, _ = tangent(3)
    errs = [rel(central(3, h)[0], dP) for h in (1e-4, 1e-5, 1e-6)]
    assert errs[1] < 0.05 * errs[0] and errs[2] < 0.1 * errs[1], errs
    assert errs[2] < 1e-6, errs


def test_fofc_every_zone_marked_is_the_donor_cell_scheme(oracle_lib):
    """First-order flux correction (flux/fofc.cpp:113-226) with every zone marked
    (an eh_buffer beyond the outer radius) replaces every consumed face flux by the
    LLF flux of the two zones' own primitives: the step is then the donor-cell
    scheme with LLF and no reconstruction floors, bit for bit, whatever the
    reconstruction the fluxes were first computed with."""
    cfg = pr.torus_config(48, 32, 1)
    P0 = pr.torus_prims(cfg)
    out = {}
    for name, kw in (("fofc", dict(fofc=True, fofc_eh_buffer=1e6)),
                     ("donor", dict(recon="donor_cell_c", recon_floors=False, riemann="llf"))):
        b = _block(oracle_lib, dict(cfg, **kw))
        b.init_state(P0)
        b.step(3)
        out[name] = b.state()
    for a, d in zip(out["fofc"], out["donor"]):
        assert np.array_equal(a, d)


def test_fofc_without_marks_changes_nothing(oracle_lib):
    """A smooth periodic flow far above the floors marks no zone, so the
    correction leaves the step bit for bit as it is."""
    cfg = pr.mhdmode_config(n=16, mode="fast", ndim=3, amp=1e-2)
    P0 = pr.mhdmode_prims(cfg)
    out = []
    for on in (False, True):
        b = _block(oracle_lib, dict(cfg, fofc=on))
        b.init_state(P0)
        b.step(4)
        out.append(b.state())
    for a, c in zip(*out):
        assert np.array_equal(a, c)


def test_fofc_marks_act_locally(oracle_lib):
    """Marks are OR-ed per zone, so adding the two polar rows at each pole
    (fofc/polar_cells) changes the step only near the poles, and widening the
    horizon buffer (fofc/eh_buffer 0.1 -> 0.6) only near the horizon: one step
    moves a difference at most the flux stencil (three zones for WENO5, one more
    for the corner EMFs, one for the fix-up's neighbours) from a marked face.
    Floors are off for the one step, so that the atmosphere's floor hits (which
    mark nearly every zone outside the torus) leave these marks visible."""
    cfg = dict(pr.torus_config(48, 32, 1), floors=False)
    P0 = pr.torus_prims(cfg)
    ng = cfg["ng"]

    def run(**kw):
        b = _block(oracle_lib, dict(cfg, fofc=True, **kw))
        b.init_state(P0)
        b.step(1)
        return b.state()[0][:, 0, ng:-ng, ng:-ng]                     # (v, j, i)

    base = run(fofc_eh_buffer=0.1)
    j, _ = np.nonzero(np.abs(run(fofc_eh_buffer=0.1, fofc_polar_cells=2) - base).max(axis=0))
    assert j.size > 0 and np.all((j < 2 + 5) | (j >= 32 - 2 - 5))
    _, i = np.nonzero(np.abs(run(fofc_eh_buffer=0.6) - base).max(axis=0))
    x1min, x1max = cfg["bounds"][0]
    r = np.exp(x1min + (i + 0.5) * (x1max - x1min) / 48)
    r_eh = 1 + math.sqrt(1 - cfg["a"] ** 2)
    assert i.size > 0 and np.all(r < np.exp(np.log(r_eh + 0.6) + 5 * (x1max - x1min) / 48))
