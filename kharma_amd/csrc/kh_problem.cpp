// kh_problem.cpp -- host-side problem set-up for the torus benchmark and host
// access to the geometry cache.  Initial conditions are generated once, on the
// host, in FP64; the step never calls into this file.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "kh_internal.hpp"

using namespace kh;

namespace {

int pfail(int code, const std::string& msg) { return kh::api_fail(code, msg); }

// Fishbone & Moncrief (1976) torus in Boyer-Lindquist coordinates,
// prob/fm_torus.hpp:18-105.
double lnh_calc(double a, double l, double rin, double r, double th) {
    if (r < rin) return 1.;
    auto piece = [&](double rr, double tt) {
        const double sth = std::sin(tt), cth = std::cos(tt);
        const double r2 = rr * rr, a2 = a * a;
        const double DD = r2 - 2. * rr + a2;
        const double AA = (r2 + a2) * (r2 + a2) - DD * a2 * sth * sth;
        const double SS = r2 + a2 * cth * cth;
        const double q = std::sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth));
        return 0.5 * std::log((1. + q) / (SS * DD / AA)) - 0.5 * q - 2. * a * rr * l / AA;
    };
    return piece(r, th) - piece(rin, M_PI / 2.);
}

// Specific angular momentum of the torus from its pressure maximum rmax
double lfish_calc(double a, double r) {
    const double sqtr = std::sqrt(r);
    return ((a * a - 2. * a * sqtr + r * r) *
            ((-2. * a * r * (a * a - 2. * a * sqtr + r * r)) / std::sqrt(2. * a * sqtr + (-3. + r) * r) +
             ((a + (-2. + r) * sqtr) * (r * r * r + a * a * (2. + r))) /
                 std::sqrt(1 + (2. * a) / std::pow(r, 1.5) - 3. / r))) /
           (r * r * r * std::sqrt(2. * a * sqtr + (-3. + r) * r) * (a * a + (-2. + r) * r));
}

double fm_torus_rho(double a, double rin, double rmax, double gam, double kappa, double r, double th) {
    const double l = lfish_calc(a, rmax);
    const double lnh = lnh_calc(a, l, rin, r, th);
    if (lnh >= 0. && r >= rin) {
        const double hm1 = std::exp(lnh) - 1.;
        return std::pow(hm1 * (gam - 1.) / (kappa * gam), 1. / (gam - 1.));
    }
    return 0.;
}

// Boyer-Lindquist metric (coordinate_systems.hpp:268-300)
void gcov_bl(double a, double r, double th_in, double g[4][4]) {
    const double th = (std::fabs(th_in) > 1e-20) ? ((std::fabs(th_in - M_PI) > 1e-20) ? th_in : M_PI - 1e-20) : 1e-20;
    const double cth = std::cos(th), sth = std::sin(th), s2 = sth * sth;
    const double a2 = a * a, r2 = r * r;
    const double mmu = 1. + a2 * cth * cth / r2;
    std::memset(g, 0, sizeof(double) * 16);
    g[0][0] = -(1. - 2. / (r * mmu));
    g[0][3] = g[3][0] = -2. * a * s2 / (r * mmu);
    g[1][1] = mmu / (1. - 2. / r + a2 / r2);
    g[2][2] = r2 * mmu;
    g[3][3] = s2 * (r2 + a2 + 2. * a2 * s2 / (r * mmu));
}

// u^t from the normalisation u.u = -1 (coordinate_utils.hpp:170-187)
void set_ut(const double g[4][4], double u[4]) {
    const double AA = g[0][0];
    const double BB = 2. * (g[0][1] * u[1] + g[0][2] * u[2] + g[0][3] * u[3]);
    const double CC = 1. + g[1][1] * u[1] * u[1] + g[2][2] * u[2] * u[2] + g[3][3] * u[3] * u[3] +
                      2. * (g[1][2] * u[1] * u[2] + g[1][3] * u[1] * u[3] + g[2][3] * u[2] * u[3]);
    u[0] = (-BB - std::sqrt(BB * BB - 4. * AA * CC)) / (2. * AA);
}


}  // namespace

extern "C" {

int kh_geometry_host(const kh_params* p, int loc, int comp, double* out) {
    if (!p || !out) return pfail(KH_ERR_INVALID, "null argument");
    if (check_params(p) != KH_OK) return pfail(KH_ERR_INVALID, "invalid parameters");
    if (loc < 0 || loc >= GLOC_N || comp < 0 || comp > KH_GEOM_RADIUS) return pfail(KH_ERR_INVALID, "bad locus/component");
    BlockDims b;
    make_dims(p, b);
    GeomHost G;
    build_geometry(*p, b, G);
    const size_t np = (size_t)G.np;
    const double* src;
    if (comp < GEO_NCOMP) src = G.g[loc].data() + (size_t)comp * np;
    else if (comp < KH_GEOM_FLOOR_RHO) src = G.conn.data() + (size_t)(comp - GEO_NCOMP) * np;
    else if (comp < KH_GEOM_RADIUS) src = G.flr.data() + (size_t)(comp - KH_GEOM_FLOOR_RHO) * np;
    else src = G.r.data();
    std::memcpy(out, src, sizeof(double) * np);
    return KH_OK;
}

int kh_torus_prims(const kh_params* p, const kh_torus_params* t, double* P) {
    if (!p || !t || !P) return pfail(KH_ERR_INVALID, "null argument");
    if (check_params(p) != KH_OK) return pfail(KH_ERR_INVALID, "invalid parameters");
    if (p->coords != KH_COORDS_KS) return pfail(KH_ERR_INVALID, "the torus needs Kerr-Schild coordinates");
    BlockDims b;
    make_dims(p, b);
    GeomHost G;
    build_geometry(*p, b, G);
    const long nc = b.nc, np = G.np;
    const double a = p->a, gam = p->gam;
    std::fill(P, P + (size_t)NV * nc, 0.);
    auto id = [&](int k, int j, int i) { return ((size_t)k * b.n2 + j) * b.n1 + i; };
    auto gq = [&](int j, int i) { return (size_t)j * G.gn1 + i; };
    auto gcomp = [&](int loc, int c, int j, int i) { return G.g[loc][(size_t)c * np + gq(j, i)]; };
    auto gcon_c = [&](int j, int i, int m, int n) {
        static const int IDX[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
        return gcomp(GLOC_CENTER, GEO_GCON + IDX[m][n], j, i);
    };

    // InitializeFMTorus, prob/fm_torus.cpp:41-197 (no tilt): rho, u, u-tilde in the torus
    const double l = lfish_calc(a, t->rmax);
    for (int k = b.ks; k <= b.ke; ++k)
        for (int j = b.js; j <= b.je; ++j)
            for (int i = b.is; i <= b.ie; ++i) {
                double X[4], Xe[4];
                locus_coords(*p, b, GLOC_CENTER, j, i, X);
                embed_coords(*p, X, Xe);
                const double r = Xe[1], th = Xe[2];
                const double lnh = lnh_calc(a, l, t->rin, r, th);
                if (!(lnh >= 0. && r >= t->rin)) continue;
                const double sth = std::sin(th), cth = std::cos(th);
                const double r2 = r * r, a2 = a * a;
                const double DD = r2 - 2. * r + a2;
                const double AA = (r2 + a2) * (r2 + a2) - DD * a2 * sth * sth;
                const double SS = r2 + a2 * cth * cth;
                const double hm1 = std::exp(lnh) - 1.;
                const double rho_l = std::pow(hm1 * (gam - 1.) / (t->kappa * gam), 1. / (gam - 1.));
                const double u_l = t->kappa * std::pow(rho_l, gam) / (gam - 1.);
                const double expm2chi = SS * SS * DD / (AA * AA * sth * sth);
                const double up1 = std::sqrt((-1. + std::sqrt(1. + 4. * l * l * expm2chi)) / 2.);
                const double up = 2. * a * r * std::sqrt(1. + up1 * up1) / std::sqrt(AA * SS * DD) +
                                  std::sqrt(SS / AA) * up1 / sth;
                // BL four-velocity -> KS -> native (coordinate_embedding.hpp:552-583)
                double ubl[4] = {0., 0., 0., up}, gbl[4][4];
                gcov_bl(a, r, th, gbl);
                set_ut(gbl, ubl);
                double uks[4];
                const double den = r * r - 2. * r + a * a;
                uks[0] = ubl[0] + 2. * r / den * ubl[1];
                uks[1] = ubl[1];
                uks[2] = ubl[2];
                uks[3] = ubl[3] + a / den * ubl[1];
                double J[4][4], Ji[4][4], un[4];
                native_dxdX(*p, X, J);
                invert_metric(J, Ji);
                for (int m = 0; m < 4; ++m) {
                    un[m] = 0.;
                    for (int n = 0; n < 4; ++n) un[m] += Ji[m][n] * uks[n];
                }
                // fourvel_to_prim, coordinate_utils.hpp:195-202
                const double alpha2 = -1.0 / gcon_c(j, i, 0, 0);
                const size_t c = id(k, j, i);
                P[RHO * nc + c] = rho_l;
                P[UU * nc + c] = u_l;
                P[U1 * nc + c] = un[1] + un[0] * alpha2 * gcon_c(j, i, 0, 1);
                P[U2 * nc + c] = un[2] + un[0] * alpha2 * gcon_c(j, i, 0, 2);
                P[U3 * nc + c] = un[3] + un[0] * alpha2 * gcon_c(j, i, 0, 3);
            }

    // Maximum density along the midplane, sampled every 0.001 in x1 (fm_torus.cpp:139-181)
    const double ddx = 0.001;
    const int nxs = (int)((p->x1max - p->x1min) / ddx);
    double rho_max = 0.;
    for (int i = 0; i <= nxs; ++i) {
        const double X[4] = {0., p->x1min + i * ddx, 0., 0.};
        double Xe[4];
        embed_coords(*p, X, Xe);
        const double rho = fm_torus_rho(a, t->rin, t->rmax, gam, t->kappa, Xe[1], M_PI / 2.);
        if (rho > rho_max) rho_max = rho;
    }
    if (!(rho_max > 0.)) return pfail(KH_ERR_INVALID, "torus does not intersect the grid");
    for (int k = b.ks; k <= b.ke; ++k)
        for (int j = b.js; j <= b.je; ++j)
            for (int i = b.is; i <= b.ie; ++i) {
                P[RHO * nc + id(k, j, i)] /= rho_max;
                P[UU * nc + id(k, j, i)] /= rho_max;
            }

    // Floors::ApplyInitialFloors (floors.cpp:151-195): fluid-frame floors with B = 0, then ceilings
    for (int k = b.ks; k <= b.ke; ++k)
        for (int j = b.js; j <= b.je; ++j)
            for (int i = b.is; i <= b.ie; ++i) {
                const size_t c = id(k, j, i);
                double rho = P[RHO * nc + c], u = P[UU * nc + c];
                const double rhoflr_geom = G.flr[gq(j, i)], uflr_geom = G.flr[np + gq(j, i)];
                const double uflr_max = std::max(uflr_geom, 0.);
                double rhoflr_max;
                int fflag = 0;
                if (!p->temp_adjust_u) {
                    const double rhoflr_temp = std::max(u, uflr_max) / p->u_over_rho_max;
                    fflag |= (rhoflr_temp > rho) * F_TEMP;
                    rhoflr_max = std::max(std::max(rhoflr_geom, 0.), rhoflr_temp);
                } else {
                    rhoflr_max = std::max(rhoflr_geom, 0.);
                }
                if (rhoflr_max > rho || uflr_max > u) fflag |= F_GEOM_RHO;
                const double u1 = P[U1 * nc + c], u2 = P[U2 * nc + c], u3 = P[U3 * nc + c];
                auto gcv = [&](int m, int n) {
                    static const int IDX[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
                    return gcomp(GLOC_CENTER, GEO_GCOV + IDX[m][n], j, i);
                };
                const double qsq = gcv(1, 1) * u1 * u1 + gcv(2, 2) * u2 * u2 + gcv(3, 3) * u3 * u3 +
                                   2. * (gcv(1, 2) * u1 * u2 + gcv(1, 3) * u1 * u3 + gcv(2, 3) * u2 * u3);
                const double gamma = std::sqrt(1. + qsq);
                if (gamma > p->gamma_max) fflag |= F_GAMMA;
                if ((gam - 1.) * u / std::pow(rho, gam) > p->ktot_max) fflag |= F_KTOT;
                if (!fflag) continue;
                rho += std::max(0., rhoflr_max - rho);
                u += std::max(0., uflr_max - u);
                P[RHO * nc + c] = rho;
                P[UU * nc + c] = u;
                if (gamma > p->gamma_max) {
                    const double f = std::sqrt((p->gamma_max * p->gamma_max - 1.) / (gamma * gamma - 1.));
                    P[U1 * nc + c] *= f; P[U2 * nc + c] *= f; P[U3 * nc + c] *= f;
                }
                const double ktot = (gam - 1.) * u / std::pow(rho, gam);
                if (ktot > p->ktot_max) P[UU * nc + c] = p->ktot_max / ktot * u;
                if (p->temp_adjust_u && P[UU * nc + c] / rho > p->u_over_rho_max) P[UU * nc + c] = p->u_over_rho_max * rho;
            }

    // PerturbU, serial path (prob/utils/perturbation.hpp:81-92)
    if (t->u_jitter > 0.) {
        std::mt19937 gen((unsigned)t->rng_seed);
        std::uniform_real_distribution<double> dis(-t->u_jitter / 2, t->u_jitter / 2);
        for (int k = b.ks; k <= b.ke; ++k)
            for (int j = b.js; j <= b.je; ++j)
                for (int i = b.is; i <= b.ie; ++i) P[UU * nc + id(k, j, i)] *= 1. + dis(gen);
    }

    // SeedBField<sane> (prob/seed_B.cpp:215-400): A_phi = max(rho/rho_norm - min_A, 0) at
    // cell corners from the analytic torus density, then B^i sqrt(-g) from the
    // corner-averaged curl (b_flux_ct_functions.hpp:109-150); A has no k dependence.
    if (b.ndim < 2) return pfail(KH_ERR_INVALID, "the torus needs at least 2 dimensions");
    std::vector<double> A((size_t)G.gn2 * G.gn1, 0.);
    for (int j = 0; j < G.gn2; ++j)
        for (int i = 0; i < G.gn1; ++i) {
            double X[4], Xe[4];
            locus_coords(*p, b, GLOC_CORNER, j, i, X);
            embed_coords(*p, X, Xe);
            const double rho_av = fm_torus_rho(a, t->rin, t->rmax, gam, t->kappa, Xe[1], Xe[2]) / rho_max;
            A[gq(j, i)] = std::max(rho_av - t->min_A, 0.);
        }
    const int k3d = (b.ndim > 2);
    double bsq_max = 0., p_max = 0.;
    for (int k = b.ks; k <= b.ke; ++k)
        for (int j = b.js; j <= b.je; ++j)
            for (int i = b.is; i <= b.ie; ++i) {
                double bu1, bu2;
                if (k3d) {
                    const double A3c2f = (A[gq(j + 1, i)] + A[gq(j + 1, i + 1)] + A[gq(j + 1, i)] + A[gq(j + 1, i + 1)]) / 4;
                    const double A3c2b = (A[gq(j, i)] + A[gq(j, i + 1)] + A[gq(j, i)] + A[gq(j, i + 1)]) / 4;
                    const double A3c1f = (A[gq(j, i + 1)] + A[gq(j, i + 1)] + A[gq(j + 1, i + 1)] + A[gq(j + 1, i + 1)]) / 4;
                    const double A3c1b = (A[gq(j, i)] + A[gq(j, i)] + A[gq(j + 1, i)] + A[gq(j + 1, i)]) / 4;
                    bu1 = (A3c2f - A3c2b) / b.dx2 - (0. - 0.) / b.dx3;
                    bu2 = (0. - 0.) / b.dx3 - (A3c1f - A3c1b) / b.dx1;
                } else {
                    const double A3c2f = (A[gq(j + 1, i)] + A[gq(j + 1, i + 1)]) / 2;
                    const double A3c2b = (A[gq(j, i)] + A[gq(j, i + 1)]) / 2;
                    const double A3c1f = (A[gq(j, i + 1)] + A[gq(j + 1, i + 1)]) / 2;
                    const double A3c1b = (A[gq(j, i)] + A[gq(j + 1, i)]) / 2;
                    bu1 = (A3c2f - A3c2b) / b.dx2;
                    bu2 = -(A3c1f - A3c1b) / b.dx1;
                }
                const double gdet = gcomp(GLOC_CENTER, GEO_GDET, j, i);
                const size_t c = id(k, j, i);
                P[B1 * nc + c] = bu1 / gdet;
                P[B2 * nc + c] = bu2 / gdet;
                P[B3 * nc + c] = 0. / gdet;
                // b^2 and gas pressure for the normalisation (reductions_variables.hpp:76-86)
                double Pc[NV];
                for (int v = 0; v < NV; ++v) Pc[v] = P[(size_t)v * nc + c];
                double gc[4][4], gi[4][4];
                static const int IDX[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
                for (int m = 0; m < 4; ++m)
                    for (int n = 0; n < 4; ++n) {
                        gc[m][n] = gcomp(GLOC_CENTER, GEO_GCOV + IDX[m][n], j, i);
                        gi[m][n] = gcomp(GLOC_CENTER, GEO_GCON + IDX[m][n], j, i);
                    }
                const double qsq = gc[1][1] * Pc[U1] * Pc[U1] + gc[2][2] * Pc[U2] * Pc[U2] + gc[3][3] * Pc[U3] * Pc[U3] +
                                   2. * (gc[1][2] * Pc[U1] * Pc[U2] + gc[1][3] * Pc[U1] * Pc[U3] + gc[2][3] * Pc[U2] * Pc[U3]);
                const double gamma = std::sqrt(1. + qsq);
                const double alpha = gcomp(GLOC_CENTER, GEO_ALPHA, j, i);
                double ucon[4], ucov[4], bcon[4], bcov[4];
                ucon[0] = gamma / alpha;
                for (int v = 0; v < 3; ++v) ucon[v + 1] = Pc[U1 + v] - gamma * alpha * gi[0][v + 1];
                for (int m = 0; m < 4; ++m) { ucov[m] = 0.; for (int n = 0; n < 4; ++n) ucov[m] += gc[m][n] * ucon[n]; }
                bcon[0] = Pc[B1] * ucov[1] + Pc[B2] * ucov[2] + Pc[B3] * ucov[3];
                for (int v = 0; v < 3; ++v) bcon[v + 1] = (Pc[B1 + v] + bcon[0] * ucon[v + 1]) / ucon[0];
                for (int m = 0; m < 4; ++m) { bcov[m] = 0.; for (int n = 0; n < 4; ++n) bcov[m] += gc[m][n] * bcon[n]; }
                const double bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
                bsq_max = std::max(bsq_max, bsq);
                p_max = std::max(p_max, (gam - 1.) * Pc[UU]);
            }
    // NormalizeBField, legacy beta = max(P_gas) / max(P_B) (prob/seed_B.cpp:453-489)
    const double beta_now = p_max / (0.5 * bsq_max);
    if (beta_now > 0.) {
        const double norm = std::sqrt(beta_now / t->beta_min);
        for (int k = b.ks; k <= b.ke; ++k)
            for (int j = b.js; j <= b.je; ++j)
                for (int i = b.is; i <= b.ie; ++i)
                    for (int v = B1; v <= B3; ++v) P[(size_t)v * nc + id(k, j, i)] *= norm;
    }
    return KH_OK;
}

}  // extern "C"
