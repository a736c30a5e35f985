"""CPU checks of the Kerr-Schild geometry cache and the torus initial state.

The metric pieces are pinned against closed forms (sqrt(-g) of Kerr-Schild,
g^tt, g g^-1 = 1), the finite-difference connection against the metric
compatibility identity d_l g_mn = G^k_lm g_kn + G^k_ln g_mk evaluated with an
independent difference of the cached metric, and the library's host geometry
against the oracle's own restatement.  The torus state is pinned by its
invariants: corner divB at roundoff, beta = max P_gas / max P_B = 100, the
density maximum normalised to 1 and located at r_max in the midplane.
"""
import math

import numpy as np
import pytest

import kharma_amd as kh
from kharma_amd import problems as pr

SYM = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


def metric(cfg, loc, which):
    """[n2+1, n1+1, 4, 4] gcov (which=0) or gcon (which=10) from the host cache."""
    comps = [kh.geometry_host(cfg, loc, which + c) for c in range(10)]
    g = np.zeros(comps[0].shape + (4, 4))
    for c, (a, b) in enumerate(SYM):
        g[..., a, b] = comps[c]
        g[..., b, a] = comps[c]
    return g


def small_cfg(transform="fmks", nx=(32, 24, 1), x1=None):
    cfg = pr.torus_config(*nx, transform=transform)
    if x1 is not None:
        b = list(cfg["bounds"])
        b[0] = x1
        cfg["bounds"] = tuple(b)
    return cfg


def native_centres(cfg, loc="center"):
    ng = cfg["ng"]
    nx1, nx2, _ = cfg["nx"]
    (x1a, x1b), (x2a, x2b), _ = cfg["bounds"]
    d1, d2 = (x1b - x1a) / nx1, (x2b - x2a) / nx2
    i = np.arange(nx1 + 2 * ng + 1) - ng
    j = np.arange(nx2 + 2 * ng + 1) - ng
    o1 = 0.0 if loc in ("face1", "corner") else 0.5
    o2 = 0.0 if loc in ("face2", "corner") else 0.5
    return np.meshgrid(x1a + (i + o1) * d1, x2a + (j + o2) * d2, indexing="xy")


def theta_of(cfg, X1, X2):
    """FMKS / MKS polar angle (the transform's definition, coordinate_systems.hpp:586-620)."""
    h = cfg["hslope"]
    thG = math.pi * X2 + 0.5 * (1 - h) * np.sin(2 * math.pi * X2)
    if cfg["transform"] == "mks":
        return thG
    xt, al, sm, x0 = cfg["poly_xt"], cfg["poly_alpha"], cfg["mks_smooth"], cfg["fmks_zero_point"]
    norm = 0.5 * math.pi / (1 + 1 / (al + 1) / xt ** al)
    y = 2 * X2 - 1
    thJ = norm * y * (1 + (y / xt) ** al / (al + 1)) + 0.5 * math.pi
    return thG + np.exp(sm * (x0 - X1)) * (thJ - thG)


@pytest.mark.parametrize("transform", ["mks", "fmks"])
def test_ks_gdet_closed_form(transform):
    cfg = small_cfg(transform)
    a = cfg["a"]
    X1, X2 = native_centres(cfg)
    r = np.exp(X1)
    th = theta_of(cfg, X1, X2)
    h = 1e-6
    dth = (theta_of(cfg, X1, X2 + h) - theta_of(cfg, X1, X2 - h)) / (2 * h)
    # sqrt(-g)_KS = (r^2 + a^2 cos^2 th) sin th, times the Jacobian r dth/dX2 of (r, th) <- (X1, X2)
    expect = (r * r + a * a * np.cos(th) ** 2) * np.abs(np.sin(th)) * r * np.abs(dth)
    got = kh.geometry_host(cfg, "center", 20)
    assert np.allclose(got, expect, rtol=1e-8, atol=0)


def test_ks_inverse_and_lapse():
    cfg = small_cfg()
    # (X2 faces include the poles, where g_33 ~ sin^2 theta -> 1e-40 and the
    # inverse is only as good as that conditioning allows; cell centres are not)
    g = metric(cfg, "center", 0)
    gi = metric(cfg, "center", 10)
    eye = np.einsum("...ab,...bc->...ac", g, gi)
    # in the basis scaled to unit diagonal (the metric spans ~1 to ~r^2 sin^2 theta)
    d = 1 / np.sqrt(np.abs(np.diagonal(g, axis1=-2, axis2=-1)))
    err = (eye - np.eye(4)) * d[..., :, None] / d[..., None, :]
    # cofactor inversion of the Kerr-Schild metric over 0.6 < r < 1000 keeps
    # ~1e-10 in this norm (a wrong cofactor would give O(1))
    assert np.abs(err).max() < 1e-8
    # g^tt = -(1 + 2 r / rho^2) for Kerr-Schild, unchanged by the spatial transform
    X1, X2 = native_centres(cfg, "center")
    r = np.exp(X1)
    th = theta_of(cfg, X1, X2)
    rho2 = r * r + cfg["a"] ** 2 * np.cos(th) ** 2
    assert np.allclose(gi[..., 0, 0], -(1 + 2 * r / rho2), rtol=1e-10)
    alpha = kh.geometry_host(cfg, "center", 21)
    assert np.allclose(alpha, 1 / np.sqrt(1 + 2 * r / rho2), rtol=1e-10)


def test_connection_table_symmetric_in_lower_indices():
    """The cached sqrt(-g) Gamma^a_{b c} is exactly symmetric in (b, c): the build
    stores the mean of the two finite-difference orders, which differ by rounding
    only (the source kernel reads each (b, c) pair once)."""
    cfg = small_cfg(nx=(24, 16, 1))
    conn = np.stack([kh.geometry_host(cfg, "center", 22 + n) for n in range(64)], axis=-1)
    conn = conn.reshape(conn.shape[:2] + (4, 4, 4))
    assert np.array_equal(conn, conn.transpose(0, 1, 2, 4, 3))
    assert np.abs(conn).max() > 0


def test_connection_metric_compatibility():
    # a block 1e-3 wide in x1 so that differences of the cached face metrics
    # give d_1 g to ~1e-9 relative
    cfg = small_cfg(nx=(16, 16, 1), x1=(math.log(4.0), math.log(4.0) + 1e-3))
    cfg["bounds"] = (cfg["bounds"][0], (0.3, 0.3 + 1e-3), cfg["bounds"][2])
    conn = np.stack([kh.geometry_host(cfg, "center", 22 + n) for n in range(64)], axis=-1)
    conn = conn.reshape(conn.shape[:2] + (4, 4, 4))          # sqrt(-g) Gamma^a_{b c}
    gdet = kh.geometry_host(cfg, "center", 20)
    G = conn / gdet[..., None, None, None]
    g = metric(cfg, "center", 0)
    g1 = metric(cfg, "face1", 0)
    g2 = metric(cfg, "face2", 0)
    ng = cfg["ng"]
    sl = (slice(ng, -ng - 1), slice(ng, -ng - 1))
    (x1a, x1b), (x2a, x2b), _ = cfg["bounds"]
    d1, d2 = (x1b - x1a) / cfg["nx"][0], (x2b - x2a) / cfg["nx"][1]
    dg = np.zeros(g.shape[:2] + (4, 4, 4))                   # dg[..., l, m, n] = d_l g_mn
    dg[:-1, :-1, 1] = (g1[:-1, 1:] - g1[:-1, :-1]) / d1
    dg[:-1, :-1, 2] = (g2[1:, :-1] - g2[:-1, :-1]) / d2
    rhs = np.einsum("...klm,...kn->...lmn", G, g) + np.einsum("...kln,...mk->...lmn", G, g)
    err = np.abs(dg - rhs)[sl].max() / np.abs(dg[sl]).max()
    assert err < 1e-5, err
    # and a transposed index (Gamma^k_{ml} for a metric with a non-symmetric
    # FD stencil) would break it: the check is sensitive to the index order
    bad = np.einsum("...kml->...klm", G)
    rhs_bad = np.einsum("...klm,...kn->...lmn", bad, g) + np.einsum("...kln,...mk->...lmn", bad, g)
    assert np.abs(dg - rhs_bad)[sl].max() / np.abs(dg[sl]).max() < 1e-5  # Gamma is symmetric in its lower pair
    # contraction Gamma^a_{a l} = d_l ln sqrt(-g)
    tr = np.einsum("...aal->...l", G)
    gd1 = kh.geometry_host(cfg, "face1", 20)
    gd2 = kh.geometry_host(cfg, "face2", 20)
    dl1 = (np.log(gd1[:-1, 1:]) - np.log(gd1[:-1, :-1])) / d1
    dl2 = (np.log(gd2[1:, :-1]) - np.log(gd2[:-1, :-1])) / d2
    assert np.allclose(tr[:-1, :-1, 1][sl], dl1[sl], rtol=1e-5)
    assert np.allclose(tr[:-1, :-1, 2][sl], dl2[sl], rtol=1e-5)
    assert np.abs(tr[..., 0]).max() < 1e-12 and np.abs(tr[..., 3]).max() < 1e-12


def test_library_geometry_matches_oracle(oracle_lib):
    from oracle import oracle
    cfg = small_cfg(nx=(32, 24, 8))
    ob = oracle.OracleBlock(cfg)
    og = ob.geometry(0)          # [loc, j, i, 4, 4], loci face1 face2 face3 centre corner
    oi = ob.geometry(1)
    od = ob.geometry(2)
    oc = ob.geometry(3)          # [j, i, a, b, c]
    # Both follow the reference's arithmetic (adjugate of 3x3 minors); they differ
    # in the last bit where the oracle's full 4x4 inverse is not exactly symmetric
    # (the library stores its upper triangle).
    loc_map = {"center": 3, "face1": 0, "face2": 1, "corner": 4}
    for loc, ol in loc_map.items():
        g = metric(cfg, loc, 0)
        gi = metric(cfg, loc, 10)
        row = np.abs(og[ol]).max(axis=-1, keepdims=True)
        assert (np.abs(g - og[ol]) / row).max() < 1e-15
        # (the asymmetry of the adjugate inverse reaches ~1e-12 of the largest
        # component where the metric is ill-conditioned, inside the horizon)
        big = np.abs(oi[ol]).max(axis=(-2, -1), keepdims=True)
        assert (np.abs(gi - oi[ol]) / big).max() < 1e-11
        assert np.allclose(kh.geometry_host(cfg, loc, 20), od[ol], rtol=1e-15, atol=0)
    # The connection is a centred difference with delta = 1e-8 (gr_coordinates.cpp:51):
    # last-bit differences of the two metric evaluations reach it amplified by 1e8.
    conn = np.stack([kh.geometry_host(cfg, "center", 22 + n) for n in range(64)], axis=-1).reshape(oc.shape)
    inner = (slice(0, -1), slice(0, -1))
    scale = np.abs(oc[inner]).max(axis=(-3, -2, -1), keepdims=True)
    assert (np.abs(conn[inner] - oc[inner]) / scale).max() < 1e-7


def _corner_divb(UB, cfg):
    ng = cfg["ng"]
    (x1a, x1b), (x2a, x2b), (x3a, x3b) = cfg["bounds"]
    nx1, nx2, nx3 = cfg["nx"]
    d1, d2, d3 = (x1b - x1a) / nx1, (x2b - x2a) / nx2, (x3b - x3a) / nx3
    B1, B2, B3 = UB
    s = np.s_
    if nx3 > 1:
        c = s[1:, 1:, 1:]
        def sh(A, dk, dj, di):
            return A[1 - dk:A.shape[0] - dk, 1 - dj:A.shape[1] - dj, 1 - di:A.shape[2] - di]
        t1 = sh(B1, 0, 0, 0) - sh(B1, 0, 0, 1) + sh(B1, 0, 1, 0) - sh(B1, 0, 1, 1) + \
            sh(B1, 1, 0, 0) + sh(B1, 1, 1, 0) - sh(B1, 1, 0, 1) - sh(B1, 1, 1, 1)
        t2 = sh(B2, 0, 0, 0) - sh(B2, 0, 1, 0) + sh(B2, 0, 0, 1) - sh(B2, 0, 1, 1) + \
            sh(B2, 1, 0, 0) + sh(B2, 1, 0, 1) - sh(B2, 1, 1, 0) - sh(B2, 1, 1, 1)
        t3 = sh(B3, 0, 0, 0) + sh(B3, 0, 1, 0) + sh(B3, 0, 0, 1) + sh(B3, 0, 1, 1) - \
            sh(B3, 1, 0, 0) - sh(B3, 1, 1, 0) - sh(B3, 1, 0, 1) - sh(B3, 1, 1, 1)
        del c
        return 0.25 * t1 / d1 + 0.25 * t2 / d2 + 0.25 * t3 / d3
    B1, B2 = B1[0], B2[0]
    t1 = B1[1:, 1:] - B1[1:, :-1] + B1[:-1, 1:] - B1[:-1, :-1]
    t2 = B2[1:, 1:] - B2[:-1, 1:] + B2[1:, :-1] - B2[:-1, :-1]
    return 0.5 * t1 / d1 + 0.5 * t2 / d2


@pytest.mark.parametrize("nx", [(64, 48, 1), (64, 32, 8)])
def test_torus_initial_state_invariants(nx):
    cfg = pr.torus_config(*nx)
    P = pr.torus_prims(cfg)
    sl = pr.interior(cfg)
    ng = cfg["ng"]
    Pi = P[(slice(None),) + sl]
    assert np.isfinite(Pi).all()
    # density normalised by the midplane maximum of the analytic solution
    assert 0.9 < Pi[0].max() <= 1.0 + 1e-12
    # U.B = sqrt(-g) B from the corner-averaged curl: divB vanishes at interior corners
    gdet = kh.geometry_host(cfg, "center", 20)[:-1, :-1][ng:-ng, ng:-ng]
    UB = Pi[5:8] * gdet[None, None]
    divb = _corner_divb(UB, cfg)
    scale = np.abs(UB).max() / min((b[1] - b[0]) / n for b, n in zip(cfg["bounds"], cfg["nx"]) if n > 1)
    assert np.abs(divb).max() < 1e-12 * scale
    # legacy beta normalisation: max P_gas / max P_B = beta_min
    g = metric(cfg, "center", 0)[:-1, :-1][ng:-ng, ng:-ng]
    gi = metric(cfg, "center", 10)[:-1, :-1][ng:-ng, ng:-ng]
    alpha = 1 / np.sqrt(-gi[..., 0, 0])
    u = Pi[2:5].transpose(1, 2, 3, 0)
    qsq = np.einsum("...a,...ab,...b->...", u, g[None, :, :, 1:, 1:], u)
    gamma = np.sqrt(1 + qsq)
    ucon = np.empty(u.shape[:-1] + (4,))
    ucon[..., 0] = gamma / alpha
    ucon[..., 1:] = u - (gamma * alpha)[..., None] * gi[None, ..., 0, 1:]
    ucov = np.einsum("...ab,...b->...a", g[None], ucon)
    B = Pi[5:8].transpose(1, 2, 3, 0)
    bcon = np.empty_like(ucon)
    bcon[..., 0] = np.einsum("...a,...a->...", B, ucov[..., 1:])
    bcon[..., 1:] = (B + bcon[..., :1] * ucon[..., 1:]) / ucon[..., :1]
    bsq = np.einsum("...a,...ab,...b->...", bcon, g[None], bcon)
    beta = ((cfg["gamma"] - 1) * Pi[1]).max() / (0.5 * bsq.max())
    assert abs(beta - 100.0) < 1e-9 * 100


def test_torus_pressure_maximum_at_rmax():
    # 2-D, odd nx2 so a row of cell centres sits on the midplane; fine radial grid
    cfg = pr.torus_config(512, 33, 1, u_jitter=0.0)
    P = pr.torus_prims(cfg)
    ng = cfg["ng"]
    jmid = ng + 16
    X1, X2 = native_centres(cfg)
    assert abs(X2[jmid, 0] - 0.5) < 1e-12
    rho = P[0, 0, jmid, ng:-ng]
    r = np.exp(X1[jmid, ng:-ng - 1])
    imax = int(np.argmax(rho))
    dr = r[imax + 1] - r[imax]
    assert abs(r[imax] - cfg["torus"]["rmax"]) <= dr
