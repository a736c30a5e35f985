// kh_adjoint.hpp -- vector-Jacobian products (reverse mode) of the pointwise
// functions of the GRMHD step, FP64.
//
// Each *_vjp function recomputes the forward intermediates it needs from its
// inputs and returns lambda^T J for the Jacobian J of the forward function at
// that point.  The forward functions are those of kh_device.hpp:
//   weno5_cell (flux/reconstruction.hpp:128-204), calc_4vecs
//   (grmhd/grmhd_functions.hpp:225-246), prim_to_u_and_flux
//   (flux/flux_functions.hpp:116-178), vchar (flux_functions.hpp:368-439), the
//   LLF flux (flux/flux.hpp:110-121), apply_geo_floors (floors_functions.hpp:300-332)
// and p_to_u (flux_functions.hpp:343-352).  Derivatives are those of the
// mathematical functions the forward code evaluates (its reciprocal estimates
// and reassociated sums differ from them by rounding only); quotients here are
// products with rcp() (~1 ulp).  At the kinks of
// max/min the branch the forward code takes carries the whole derivative.
#pragma once
#include "kh_device.hpp"

namespace kh {

// ------------------------------------------------------------------ WENO5
// Reverse of weno5_cell(x1..x5) -> (lout, rout): g[k] = lam_l dl/dx_k + lam_r dr/dx_k.
// With t_k = EPS + beta_k and normalised weights w_k = (c_k / t_k^2) / sum_j (c_j / t_j^2):
//   d(out)/d(q_k) = w_k,   d(out)/d(t_k) = -2 w_k (q_k - out) / t_k.
__device__ __forceinline__ void weno5_cell_vjp(double x1, double x2, double x3, double x4, double x5,
                                               double lam_l, double lam_r, double g[5]) {
    constexpr double EPS = 1.e-26;
    const double a1 = x1 - 2. * x2 + x3, a2 = x1 - 4. * x2 + 3. * x3;
    const double b1 = x2 - 2. * x3 + x4, b2 = x4 - x2;
    const double c1 = x3 - 2. * x4 + x5, c2 = x5 - 4. * x4 + 3. * x3;
    const double t0 = EPS + ((13. / 12.) * a1 * a1 + 0.25 * a2 * a2);
    const double t1 = EPS + ((13. / 12.) * b1 * b1 + 0.25 * b2 * b2);
    const double t2 = EPS + ((13. / 12.) * c1 * c1 + 0.25 * c2 * c2);
    // weights in the product form of the forward code (no under/overflow for t >= EPS)
    const double s0 = t0 * t0, s1 = t1 * t1, s2 = t2 * t2;
    const double p12 = s1 * s2, p02 = s0 * s2, p01 = s0 * s1;
    const double ar0 = (1. / 16.) * p12, ar1 = (5. / 8.) * p02, ar2 = (5. / 16.) * p01;
    const double al0 = (1. / 16.) * p01, al1 = ar1, al2 = (5. / 16.) * p12;
    const double iWr = rcp(ar0 + ar1 + ar2), iWl = rcp(al0 + al1 + al2);
    const double wr0 = ar0 * iWr, wr1 = ar1 * iWr, wr2 = ar2 * iWr;
    const double wl0 = al0 * iWl, wl1 = al1 * iWl, wl2 = al2 * iWl;
    const double qr0 = (3. / 8.) * x1 - (5. / 4.) * x2 + (15. / 8.) * x3;
    const double qr1 = (-1. / 8.) * x2 + (3. / 4.) * x3 + (3. / 8.) * x4;
    const double qr2 = (3. / 8.) * x3 + (3. / 4.) * x4 - (1. / 8.) * x5;
    const double ql0 = (3. / 8.) * x5 - (5. / 4.) * x4 + (15. / 8.) * x3;   // weight of t2
    const double ql1 = (-1. / 8.) * x4 + (3. / 4.) * x3 + (3. / 8.) * x2;   // weight of t1
    const double ql2 = (3. / 8.) * x3 + (3. / 4.) * x2 - (1. / 8.) * x1;    // weight of t0
    const double rout = qr0 * wr0 + qr1 * wr1 + qr2 * wr2;
    const double lout = ql0 * wl0 + ql1 * wl1 + ql2 * wl2;
    // adjoints of the candidate values
    const double gr0 = lam_r * wr0, gr1 = lam_r * wr1, gr2 = lam_r * wr2;
    const double gl0 = lam_l * wl0, gl1 = lam_l * wl1, gl2 = lam_l * wl2;
    // adjoints of t_k
    const double gt0 = -2. * (gr0 * (qr0 - rout) + gl2 * (ql2 - lout)) * rcp(t0);
    const double gt1 = -2. * (gr1 * (qr1 - rout) + gl1 * (ql1 - lout)) * rcp(t1);
    const double gt2 = -2. * (gr2 * (qr2 - rout) + gl0 * (ql0 - lout)) * rcp(t2);
    // through the candidate values
    g[0] = (3. / 8.) * gr0 - (1. / 8.) * gl2;
    g[1] = (-5. / 4.) * gr0 - (1. / 8.) * gr1 + (3. / 8.) * gl1 + (3. / 4.) * gl2;
    g[2] = (15. / 8.) * gr0 + (3. / 4.) * gr1 + (3. / 8.) * gr2 + (15. / 8.) * gl0 + (3. / 4.) * gl1 + (3. / 8.) * gl2;
    g[3] = (3. / 8.) * gr1 + (3. / 4.) * gr2 - (5. / 4.) * gl0 - (1. / 8.) * gl1;
    g[4] = (-1. / 8.) * gr2 + (3. / 8.) * gl0;
    // through the smoothness indicators: dt/dx = (13/6) c1 dc1/dx + (1/2) c2 dc2/dx
    const double e0a = (13. / 6.) * a1 * gt0, e0b = 0.5 * a2 * gt0;
    g[0] += e0a + e0b;
    g[1] += -2. * e0a - 4. * e0b;
    g[2] += e0a + 3. * e0b;
    const double e1a = (13. / 6.) * b1 * gt1, e1b = 0.5 * b2 * gt1;
    g[1] += e1a - e1b;
    g[2] += -2. * e1a;
    g[3] += e1a + e1b;
    const double e2a = (13. / 6.) * c1 * gt2, e2b = 0.5 * c2 * gt2;
    g[2] += e2a + 3. * e2b;
    g[3] += -2. * e2a - 4. * e2b;
    g[4] += e2a + e2b;
}

// ------------------------------------------------------------ one face side
// Forward intermediates of one point shared by the reverse sweeps below
// (calc_4vecs, grmhd_functions.hpp:225-246, and the tensor invariants).
struct SideState {
    FourVec<double> D;
    double gamma, alpha, iu0, gdet, bsq, eta;
};
template <class G>
__device__ __forceinline__ void side_forward(const G& g, const double P[NV], double gam, SideState& S) {
    calc_4vecs<double>(g, P, S.D);
    S.gamma = g.lorentz(P[U1], P[U2], P[U3]);
    if constexpr (G::flat) S.alpha = 1.;
    else S.alpha = g.alpha();
    S.iu0 = rcp(S.D.ucon[0]);
    S.gdet = g.gdet();
    S.bsq = dot4(S.D.bcon, S.D.bcov);
    S.eta = (gam - 1.) * P[UU] + P[RHO] + P[UU] + S.bsq;
}

// Reverse of P -> (U = prim_to_u(P), F^DIR = prim_to_flux(P), cmax, cmin = vchar(P))
// at one point of geometry g, from the forward intermediates S of side_forward.
// lU, lF, lcmax, lcmin are the adjoints of the outputs; lP receives the
// adjoint of P.  DIR = 0 skips F and the speeds (the conserved variables
// alone, for the inverter and P->U adjoints).
// Adjoint accumulators of one point's forward intermediates (SideState).
struct SideAdj {
    double ucon[4] = {0., 0., 0., 0.}, ucov[4] = {0., 0., 0., 0.};
    double bcon[4] = {0., 0., 0., 0.}, bcov[4] = {0., 0., 0., 0.};
    double rho = 0., u = 0., eta = 0., ptot = 0., bsq = 0.;
    double B[3] = {0., 0., 0.};
};

// Shared tail of the reverse sweeps of one point: from the adjoints of the
// four-vectors, of rho, u, B^i and of the tensor invariants eta, ptot, bsq (a)
// to the adjoint of P, through calc_4vecs (grmhd_functions.hpp:225-246).
template <class G>
__device__ __forceinline__ void side_tail(const G& g, const double P[NV], double gam, const SideState& S, SideAdj& a,
                                          double lP[NV]) {
    const double ut1 = P[U1], ut2 = P[U2], ut3 = P[U3];
    const double* ucon = S.D.ucon;
    const double* ucov = S.D.ucov;
    const double* bcon = S.D.bcon;
    const double* bcov = S.D.bcov;
    const double gamma = S.gamma, alpha = S.alpha, iu0 = S.iu0;
    // eta = pgas + rho + u + bsq, ptot = pgas + bsq / 2, pgas = (gam - 1) u
    a.rho += a.eta;
    a.u += a.eta + (gam - 1.) * (a.eta + a.ptot);
    a.bsq += a.eta + 0.5 * a.ptot;
    // bsq = b^mu b_mu
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        a.bcon[mu] += a.bsq * bcov[mu];
        a.bcov[mu] += a.bsq * bcon[mu];
    }
    // b_mu = g_{mu nu} b^nu (g symmetric)
    {
        double t[4];
        g.lower(a.bcov, t);
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) a.bcon[mu] += t[mu];
    }
    // b^i = (B^i + b^0 u^i) / u^0
    double l_iu0 = 0.;
#pragma unroll
    for (int i = 1; i <= 3; ++i) {
        a.B[i - 1] += a.bcon[i] * iu0;
        a.bcon[0] += a.bcon[i] * ucon[i] * iu0;
        a.ucon[i] += a.bcon[i] * bcon[0] * iu0;
        l_iu0 += a.bcon[i] * (P[B1 + i - 1] + bcon[0] * ucon[i]);
    }
    a.ucon[0] -= l_iu0 * iu0 * iu0;
    // b^0 = B^i u_i
#pragma unroll
    for (int i = 1; i <= 3; ++i) {
        a.B[i - 1] += a.bcon[0] * ucov[i];
        a.ucov[i] += a.bcon[0] * P[B1 + i - 1];
    }
    // u_mu = g_{mu nu} u^nu
    {
        double t[4];
        g.lower(a.ucov, t);
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) a.ucon[mu] += t[mu];
    }
    // u^0 = gamma / alpha, u^i = ut^i - gamma alpha g^{0i}
    double l_gamma;
    if constexpr (G::flat) l_gamma = a.ucon[0];
    else l_gamma = a.ucon[0] * rcp(alpha) - alpha * (a.ucon[1] * g.gcon0(1) + a.ucon[2] * g.gcon0(2) + a.ucon[3] * g.gcon0(3));
    // gamma = sqrt(1 + g_ij ut^i ut^j)
    const double l_q = l_gamma * rcp(gamma);   // 2 * d(gamma)/d(qsq) = 1 / gamma
    double lut[3];
#pragma unroll
    for (int i = 1; i <= 3; ++i)
        lut[i - 1] = a.ucon[i] + l_q * (g.gcov(i, 1) * ut1 + g.gcov(i, 2) * ut2 + g.gcov(i, 3) * ut3);
    lP[RHO] = a.rho;
    lP[UU] = a.u;
    lP[U1] = lut[0]; lP[U2] = lut[1]; lP[U3] = lut[2];
    lP[B1] = a.B[0]; lP[B2] = a.B[1]; lP[B3] = a.B[2];
}

template <int DIR, class G>
__device__ __forceinline__ void side_reverse(const G& g, const double P[NV], double gam, const SideState& S,
                                             const double lU[NV], const double lF[NV], double lcmax, double lcmin,
                                             double lP[NV]) {
    const double* ucon = S.D.ucon;
    const double* ucov = S.D.ucov;
    const double* bcon = S.D.bcon;
    const double* bcov = S.D.bcov;
    const double gdet = S.gdet, bsq = S.bsq, eta = S.eta;
    // ---- reverse
    SideAdj acc;

    if constexpr (DIR > 0) {
        // vchar (flux_functions.hpp:368-439)
        const double ef = P[RHO] + gam * P[UU];
        const double ief = rcp(ef);
        const double cs2 = gam * (gam - 1.) * P[UU] * ief;
        const double ibef = rcp(bsq + ef);
        const double va2 = bsq * ibef;
        const double cms2 = cs2 + va2 - cs2 * va2;
        double Asq, Bsq, AB;
        if constexpr (G::flat) { Asq = 1.; Bsq = -1.; AB = 0.; }
        else { Asq = g.gcon(DIR, DIR); Bsq = g.gcon(0, 0); AB = g.gcon(0, DIR); }
        const double Au = ucon[DIR], Bu = ucon[0];
        const double A = Bu * Bu - (Bsq + Bu * Bu) * cms2;
        const double B = 2. * (Au * Bu - (AB + Au * Bu) * cms2);
        const double C = Au * Au - (Asq + Au * Au) * cms2;
        const double Dd = B * B - 4. * A * C;
        const double discr = sqrt(smax(Dd, 0.));
        const double i2A = rcp(2. * A);
        const double vp = (B - discr) * i2A, vm = (B + discr) * i2A;
        // cmax = smax(vp, vm), cmin = smin(vp, vm): ties take vp for both
        double l_vp, l_vm;
        if (vp < vm) { l_vp = lcmin; l_vm = lcmax; }
        else if (vm < vp) { l_vp = lcmax; l_vm = lcmin; }
        else { l_vp = lcmax + lcmin; l_vm = 0.; }
        double l_B2 = (l_vp + l_vm) * i2A;
        const double l_discr = (l_vm - l_vp) * i2A;
        double l_A = -2. * (l_vp * vp + l_vm * vm) * i2A;
        const double l_Dd = (Dd > 0.) ? 0.5 * l_discr * rcp(discr) : 0.;
        l_B2 += 2. * B * l_Dd;
        l_A += -4. * C * l_Dd;
        const double l_C = -4. * A * l_Dd;
        const double om = 1. - cms2;
        acc.ucon[DIR] += 2. * om * (l_B2 * Bu + l_C * Au);
        acc.ucon[0] += 2. * om * (l_A * Bu + l_B2 * Au);
        const double l_cms2 = -l_A * (Bsq + Bu * Bu) - 2. * l_B2 * (AB + Au * Bu) - l_C * (Asq + Au * Au);
        const double l_cs2 = l_cms2 * (1. - va2), l_va2 = l_cms2 * (1. - cs2);
        acc.bsq += l_va2 * ef * ibef * ibef;
        double l_ef = -l_va2 * bsq * ibef * ibef;
        acc.u += l_cs2 * gam * (gam - 1.) * ief;
        l_ef += -l_cs2 * cs2 * ief;
        acc.rho += l_ef;
        acc.u += gam * l_ef;
    }

    // prim_to_u / prim_to_flux (flux_functions.hpp:116-178): U[UU] and F[UU] carry
    // + U[RHO] / + F[RHO].
    {
        const double lU0 = lU[RHO] + lU[UU];
        acc.rho += lU0 * ucon[0] * gdet;
        acc.ucon[0] += lU0 * P[RHO] * gdet;
        if constexpr (DIR > 0) {
            const double lF0 = lF[RHO] + lF[UU];
            acc.rho += lF0 * ucon[DIR] * gdet;
            acc.ucon[DIR] += lF0 * P[RHO] * gdet;
        }
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) {
            // U[UU+mu] = (eta u^0 u_mu + delta^0_mu ptot - b^0 b_mu) gdet
            const double a = lU[UU + mu] * gdet;
            acc.eta += a * ucon[0] * ucov[mu];
            acc.ucon[0] += a * eta * ucov[mu];
            acc.ucov[mu] += a * eta * ucon[0];
            if (mu == 0) acc.ptot += a;
            acc.bcon[0] -= a * bcov[mu];
            acc.bcov[mu] -= a * bcon[0];
            if constexpr (DIR > 0) {
                // F[UU+mu] = (eta u^d u_mu + delta^d_mu ptot - b^d b_mu) gdet
                const double f = lF[UU + mu] * gdet;
                acc.eta += f * ucon[DIR] * ucov[mu];
                acc.ucon[DIR] += f * eta * ucov[mu];
                acc.ucov[mu] += f * eta * ucon[DIR];
                if (mu == DIR) acc.ptot += f;
                acc.bcon[DIR] -= f * bcov[mu];
                acc.bcov[mu] -= f * bcon[DIR];
            }
        }
#pragma unroll
        for (int i = 1; i <= 3; ++i) {
            acc.B[i - 1] += lU[B1 + i - 1] * gdet;
            if constexpr (DIR > 0) {
                // F[Bi] = (b^i u^d - b^d u^i) gdet
                const double f = lF[B1 + i - 1] * gdet;
                acc.bcon[i] += f * ucon[DIR];
                acc.ucon[DIR] += f * bcon[i];
                acc.bcon[DIR] -= f * ucon[i];
                acc.ucon[i] -= f * bcon[DIR];
            }
        }
    }
    side_tail(g, P, gam, S, acc, lP);
}

template <int DIR, class G>
__device__ __forceinline__ void side_vjp(const G& g, const double P[NV], double gam, const double lU[NV],
                                         const double lF[NV], double lcmax, double lcmin, double lP[NV]) {
    SideState S;
    side_forward(g, P, gam, S);
    side_reverse<DIR>(g, P, gam, S, lU, lF, lcmax, lcmin, lP);
}

// -------------------------------------------------------------- face (LLF)
// Reverse of the LLF face flux of face_flux_g (kh_flux.hip; get_flux.hpp:237-367)
// from the two reconstructed states, including the reconstruction floors.
template <int DIR, class G>
__device__ __forceinline__ void face_vjp_llf(const G& g, double gam, bool recon_floors, double rflr, double uflr,
                                             const double Pl_in[NV], const double Pr_in[NV], const double lF[NV],
                                             double lPl[NV], double lPr[NV]) {
    double Pl[NV], Pr[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { Pl[v] = Pl_in[v]; Pr[v] = Pr_in[v]; }
    // P += max(0, floor - P): the state moves only where it is below the floor
    bool kl_rho = true, kl_u = true, kr_rho = true, kr_u = true;
    if (recon_floors) {
        kl_rho = !(rflr - Pl[RHO] > 0.); kl_u = !(uflr - Pl[UU] > 0.);
        kr_rho = !(rflr - Pr[RHO] > 0.); kr_u = !(uflr - Pr[UU] > 0.);
        apply_geo_floors<double>(Pl, rflr, uflr);
        apply_geo_floors<double>(Pr, rflr, uflr);
    }
    double Ul[NV], Fl[NV], Ur[NV], Fr[NV], cmaxL, cminL, cmaxR, cminR;
    SideState SL, SR;
    side_forward(g, Pl, gam, SL);
    prim_to_u_and_flux<double, G, DIR>(g, Pl, SL.D, gam, Ul, Fl);
    vchar<double, G, DIR>(g, Pl, SL.D, gam, cmaxL, cminL);
    side_forward(g, Pr, gam, SR);
    prim_to_u_and_flux<double, G, DIR>(g, Pr, SR.D, gam, Ur, Fr);
    vchar<double, G, DIR>(g, Pr, SR.D, gam, cmaxR, cminR);
    // cmx = smax(smax(0, cmaxL), cmaxR); cmn = -smin(smin(0, cminL), cminR); ctop = smax(cmx, cmn)
    const double t1 = smax(0., cmaxL);
    const bool mx_R = t1 < cmaxR, mx_L = !mx_R && (0. < cmaxL);
    const double cmx = mx_R ? cmaxR : t1;
    const double s1 = smin(0., cminL);
    const bool mn_R = cminR < s1, mn_L = !mn_R && (cminL < 0.);
    const double cmn = -(mn_R ? cminR : s1);
    const bool use_cmn = cmx < cmn;
    const double ctop = use_cmn ? cmn : cmx;
    double lUl[NV], lUr[NV], lFh[NV], l_ctop = 0.;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        lFh[v] = 0.5 * lF[v];
        lUl[v] = 0.5 * ctop * lF[v];
        lUr[v] = -0.5 * ctop * lF[v];
        l_ctop -= 0.5 * lF[v] * (Ur[v] - Ul[v]);
    }
    double lcmaxL = 0., lcmaxR = 0., lcminL = 0., lcminR = 0.;
    if (use_cmn) {
        if (mn_R) lcminR = -l_ctop;
        else if (mn_L) lcminL = -l_ctop;
    } else {
        if (mx_R) lcmaxR = l_ctop;
        else if (mx_L) lcmaxL = l_ctop;
    }
    side_reverse<DIR>(g, Pl, gam, SL, lUl, lFh, lcmaxL, lcminL, lPl);
    side_reverse<DIR>(g, Pr, gam, SR, lUr, lFh, lcmaxR, lcminR, lPr);
    if (!kl_rho) lPl[RHO] = 0.;
    if (!kl_u) lPl[UU] = 0.;
    if (!kr_rho) lPr[RHO] = 0.;
    if (!kr_u) lPr[UU] = 0.;
}

// ------------------------------------------------------- inverter (IFT)
// Adjoint of the converged U->P recovery at P (Inverter::u_to_p<onedw>,
// inverter/onedw.hpp:89-233, with P.B = U.B / gdet, b_flux_ct.cpp:172-217): by the
// implicit-function theorem dP/dU = (dU/dP)^-1 for U = p_to_u(P), so
// lU = (dU/dP)^-T lP.  dU/dP has the block form [[J_hh, J_hB], [0, gdet I]]
// (hydro rows rho, u, u^1..3; B rows); J_hh^T x = lP_h is solved by Gaussian
// elimination with partial pivoting, then lU_B = (lP_B - J_hB^T x) / gdet.
// Returns false if J_hh is singular to working precision.
template <class G>
__device__ __forceinline__ bool inverter_vjp(const G& g, const double P[NV], double gam, const double lP[NV],
                                             double lU[NV]) {
    double M[5][6];   // [row][col] of J_hh^T with lP_h appended
    double N[3][5];   // J_hB^T
    const double zero[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
    SideState S;
    side_forward(g, P, gam, S);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double e[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
        e[k] = 1.;
        double row[NV];   // (dU_k/dP)^T: row k of dU/dP
        side_reverse<0>(g, P, gam, S, e, zero, 0., 0., row);
#pragma unroll
        for (int j = 0; j < 5; ++j) M[j][k] = row[j];
#pragma unroll
        for (int j = 0; j < 3; ++j) N[j][k] = row[5 + j];
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) M[j][5] = lP[j];
    // Partial pivoting by conditional row swaps with static indices only (a
    // runtime-indexed row would send M to scratch memory).
#pragma unroll
    for (int col = 0; col < 5; ++col) {
#pragma unroll
        for (int r = col + 1; r < 5; ++r) {
            const bool sw = fabs(M[r][col]) > fabs(M[col][col]);
#pragma unroll
            for (int q = col; q < 6; ++q) {
                const double a = M[col][q], c = M[r][q];
                M[col][q] = sw ? c : a;
                M[r][q] = sw ? a : c;
            }
        }
        if (!(fabs(M[col][col]) > 0.)) return false;
        const double ip = rcp(M[col][col]);
#pragma unroll
        for (int r = col + 1; r < 5; ++r) {
            const double f = M[r][col] * ip;
#pragma unroll
            for (int q = col; q < 6; ++q) M[r][q] -= f * M[col][q];
        }
    }
    double x[5];
#pragma unroll
    for (int r = 4; r >= 0; --r) {
        double s = M[r][5];
#pragma unroll
        for (int q = r + 1; q < 5; ++q) s -= M[r][q] * x[q];
        x[r] = s * rcp(M[r][r]);
    }
    const double igd = rcp(g.gdet());
#pragma unroll
    for (int j = 0; j < 5; ++j) lU[j] = x[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) lU[5 + j] = (lP[5 + j] - (N[j][0] * x[0] + N[j][1] * x[1] + N[j][2] * x[2] +
                                                           N[j][3] * x[3] + N[j][4] * x[4])) * igd;
    return true;
}

// ------------------------------------------------------ dual numbers
// Forward-mode dual numbers v + d e (e^2 = 0): one directional derivative
// carried through the templated floor and inflow functions of kh_device.hpp,
// whose Jacobians (a few branches, no fixed structure) are taken column by
// column this way.  Values follow the forward arithmetic operation for
// operation (IEEE quotients where the forward code divides), so the branches
// taken are the forward ones.  They live in their own namespace and are found
// by argument-dependent lookup, so double-typed calls never see them.
namespace ad {
struct Dual {
    double v, d;
    __host__ __device__ constexpr Dual(double x = 0., double dx = 0.) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
__device__ __forceinline__ Dual sqrt(const Dual& a) {
    const double r = ::sqrt(a.v);
    return Dual(r, 0.5 * a.d / r);
}
__device__ __forceinline__ Dual rcp(const Dual& a) {
    const double r = kh::rcp(a.v);
    return Dual(r, -a.d * r * r);
}
__device__ __forceinline__ Dual pow(const Dual& a, double e) {
    const double r = ::pow(a.v, e);
    return Dual(r, e * r / a.v * a.d);
}
__device__ __forceinline__ double val(const Dual& a) { return a.v; }
}  // namespace ad

// -------------------------------------------------- geometric source
// Reverse of add_geo_source (Flux::AddGeoSource, flux/flux.cpp:395-444) at one
// zone: S_lam = sum_{mu,nu} T^mu_nu (gdet Gamma^nu_{lam mu}) for lam = 0..3
// (dudt[UU..U3]), T^mu_nu = eta u^mu u_nu + ptot delta^mu_nu - b^mu b_nu.
// lP += (dS/dP)^T lS.
template <class G>
__device__ __forceinline__ void geo_source_vjp(const G& g, const GeomDev& gd, long q, const double P[NV], double gam,
                                               const double lS[4], double lP[NV]) {
    SideState S;
    side_forward(g, P, gam, S);
    SideAdj a;
    const double* ucon = S.D.ucon;
    const double* ucov = S.D.ucov;
    const double* bcon = S.D.bcon;
    const double* bcov = S.D.bcov;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        double cm[16];
#pragma unroll
        for (int nl = 0; nl < 16; ++nl) cm[nl] = gd.conn[(long)(nl * 4 + mu) * gd.np + q];
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) {
            const double lT = lS[0] * cm[nu * 4] + lS[1] * cm[nu * 4 + 1] + lS[2] * cm[nu * 4 + 2] + lS[3] * cm[nu * 4 + 3];
            a.eta += lT * ucon[mu] * ucov[nu];
            a.ucon[mu] += lT * S.eta * ucov[nu];
            a.ucov[nu] += lT * S.eta * ucon[mu];
            if (mu == nu) a.ptot += lT;
            a.bcon[mu] -= lT * bcov[nu];
            a.bcov[nu] -= lT * bcon[mu];
        }
    }
    double out[NV];
    side_tail(g, P, gam, S, a, out);
#pragma unroll
    for (int v = 0; v < NV; ++v) lP[v] += out[v];
}

// ------------------------------------------------------------- floors
// Reverse of the floors of one zone (floors_zone: determine_floors, the frame's
// floors, apply_ceilings; floors/floors_impl.hpp:45-128) at the inverted state
// Pin: lPin = (dP_out / dPin)^T lPout.  The B components pass through.
// FRAME 0 (fluid) or 2 (drift).  One dual-number pass per input component.
// Component k of lPin (one column of the Jacobian, one dual-number pass).
template <int FRAME, class G>
__device__ __forceinline__ double floors_vjp_col(const G& g, const double Pin[NV], double gam, const FloorParams& f,
                                                 double rflr, double uflr, const double lPout[NV], int k) {
    using ad::Dual;
    Dual P[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) P[v] = Dual(Pin[v], v == k ? 1. : 0.);
    FourVec<Dual> D;
    calc_4vecs<Dual>(g, P, D);
    Dual rf, uf;
    const int ff = determine_floors(g, P, dot4(D.bcon, D.bcov), gam, f, rflr, uflr, rf, uf);
    if (ff) {
        if constexpr (FRAME == 0) {
            P[RHO] = P[RHO] + smax(Dual(0.), rf - P[RHO]);
            P[UU] = P[UU] + smax(Dual(0.), uf - P[UU]);
        } else {
            apply_floors_drift(g, P, gam, rf, uf);
        }
        apply_ceilings(g, P, gam, f);
    }
    double dk = 0.;
#pragma unroll
    for (int m = 0; m < 5; ++m) dk += lPout[m] * P[m].d;
#pragma unroll
    for (int v = B1; v < NV; ++v)
        if (v == k) dk += lPout[v];
    return dk;
}
template <int FRAME, class G>
__device__ __forceinline__ void floors_vjp(const G& g, const double Pin[NV], double gam, const FloorParams& f,
                                           double rflr, double uflr, const double lPout[NV], double lPin[NV]) {
#pragma unroll 1
    for (int k = 0; k < NV; ++k) {
        const double dk = floors_vjp_col<FRAME>(g, Pin, gam, f, rflr, uflr, lPout, k);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            if (v == k) lPin[v] = dk;
    }
}

// ------------------------------------------------------- inflow check
// Reverse of check_inflow (boundaries/boundaries.hpp:101-136) at the state Pin:
// l[U1..U3] <- (dP_out / dPin)^T l[U1..U3] (the other components pass through).
template <class G>
__device__ __forceinline__ void inflow_vjp(const G& g, const double Pin[NV], bool inner, double l[NV]) {
    using ad::Dual;
    double out[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        Dual P[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) P[v] = Dual(Pin[v], v == U1 + k ? 1. : 0.);
        check_inflow(g, P, inner);
        out[k] = l[U1] * P[U1].d + l[U2] * P[U2].d + l[U3] * P[U3].d;
    }
    l[U1] = out[0]; l[U2] = out[1]; l[U3] = out[2];
}

}  // namespace kh
