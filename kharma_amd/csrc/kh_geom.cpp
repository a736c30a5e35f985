// kh_geom.cpp -- host-side geometry of one meshblock: the metric cache the
// kernels read (GRCoordinates' cache, coordinates/gr_coordinates.hpp:86-90,
// filled as init_GRCoordinates does, coordinates/gr_coordinates.cpp:96-248).
//
// The metric is axisymmetric and stationary, so every cached quantity depends
// on (j, i) only and is stored once per 2-D point, shared by all k.  It is
// evaluated once, on the host, in FP64 with the C library's transcendental
// functions; the kernels never call sin/cos/exp/pow on geometry.
#include <cmath>
#include <cstring>

#include "kh_internal.hpp"

namespace kh {

namespace {

constexpr double SMALL = 1.e-20;   // decs.hpp:82
constexpr double DELTA = 1.e-8;    // finite-difference step of the connection, gr_coordinates.cpp:51

// kharma_utils.hpp:119-123: move n out of (center - range, center + range)
double excise(double n, double center, double range) {
    return (std::fabs(n - center) > range) ? n : ((n > center) ? center + range : center - range);
}

struct Embedding {
    int spherical = 0;
    double a = 0.;
    int transform = KH_TRANSFORM_NULL;
    double hslope = 0.3, mks_smooth = 0.5, poly_xt = 0.82, poly_alpha = 14., startx1 = 0.;
    double poly_norm = 0.;

    // Native -> embedding (r, th, phi) coordinates, coordinate_systems.hpp:351-672
    void to_embed(const double X[4], double Xe[4]) const {
        Xe[0] = X[0];
        Xe[3] = X[3];
        if (!spherical || transform == KH_TRANSFORM_NULL) { Xe[1] = X[1]; Xe[2] = X[2]; return; }
        Xe[1] = std::exp(X[1]);
        const double thG = M_PI * X[2] + ((1. - hslope) / 2.) * std::sin(2. * M_PI * X[2]);
        if (transform == KH_TRANSFORM_MKS) { Xe[2] = thG; return; }
        const double y = 2. * X[2] - 1.;
        const double thJ = poly_norm * y * (1. + std::pow(y / poly_xt, poly_alpha) / (poly_alpha + 1.)) + 0.5 * M_PI;
        Xe[2] = thG + std::exp(mks_smooth * (startx1 - X[1])) * (thJ - thG);
    }

    // d(embedding)/d(native), the contravariant transformation matrix
    void dxdX(const double X[4], double M[4][4]) const {
        std::memset(M, 0, sizeof(double) * 16);
        M[0][0] = 1.; M[1][1] = 1.; M[2][2] = 1.; M[3][3] = 1.;
        if (!spherical || transform == KH_TRANSFORM_NULL) return;
        M[1][1] = std::exp(X[1]);
        if (transform == KH_TRANSFORM_MKS) { M[2][2] = M_PI - (hslope - 1.) * M_PI * std::cos(2. * M_PI * X[2]); return; }
        // d theta / d X1 = -s e (thJ - thG), d theta / d X2 = thG' + e (thJ' - thG'),
        // e = exp(s (x0 - X1)), grouped as coordinate_systems.hpp:630-655 groups them
        const double x2 = X[2];
        const double e = std::exp(mks_smooth * (startx1 - X[1]));
        M[2][1] = -e * mks_smooth *
                  (M_PI / 2. - M_PI * x2 +
                   poly_norm * (2. * x2 - 1.) * (1 + std::pow((-1. + 2 * x2) / poly_xt, poly_alpha) / (1 + poly_alpha)) -
                   1. / 2. * (1. - hslope) * std::sin(2. * M_PI * x2));
        M[2][2] = M_PI + (1. - hslope) * M_PI * std::cos(2. * M_PI * x2) +
                  e * (-M_PI + 2. * poly_norm * (1. + std::pow((2. * x2 - 1.) / poly_xt, poly_alpha) / (poly_alpha + 1.)) +
                       (2. * poly_alpha * poly_norm * (2. * x2 - 1.) * std::pow((2. * x2 - 1.) / poly_xt, poly_alpha - 1.)) /
                           ((1. + poly_alpha) * poly_xt) -
                       (1. - hslope) * M_PI * std::cos(2. * M_PI * x2));
    }

    // Metric of the embedding system at Xe (coordinate_systems.hpp:77-151)
    void gcov_embed(const double Xe[4], double g[4][4]) const {
        std::memset(g, 0, sizeof(double) * 16);
        if (!spherical) { g[0][0] = -1.; g[1][1] = 1.; g[2][2] = 1.; g[3][3] = 1.; return; }
        const double r = Xe[1];
        const double th = excise(excise(Xe[2], 0.0, SMALL), M_PI, SMALL);
        const double cth = std::cos(th), sth = std::sin(th);
        const double s2 = sth * sth;
        const double rho2 = r * r + a * a * cth * cth;
        // Kerr-Schild: g = eta + f l l with f = 2 r / rho^2, l = (1, 1, 0, -a sin^2 th)
        g[0][0] = -1. + 2. * r / rho2;
        g[0][1] = g[1][0] = 2. * r / rho2;
        g[0][3] = g[3][0] = -2. * a * r * s2 / rho2;
        g[1][1] = 1. + 2. * r / rho2;
        g[1][3] = g[3][1] = -a * s2 * (1. + 2. * r / rho2);
        g[2][2] = rho2;
        g[3][3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2));
    }

    // Metric in native coordinates: g_native = J^T g_embed J, J = dx/dX
    void gcov_native(const double X[4], double g[4][4]) const {
        double Xe[4], ge[4][4], J[4][4];
        to_embed(X, Xe);
        gcov_embed(Xe, ge);
        dxdX(X, J);
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = 0; nu < 4; ++nu) {
                double s = 0.;
                for (int lam = 0; lam < 4; ++lam)
                    for (int kap = 0; kap < 4; ++kap) s += ge[lam][kap] * J[lam][mu] * J[kap][nu];
                g[mu][nu] = s;
            }
    }
};

// 4x4 inverse as adjugate / determinant, each cofactor a 3x3 minor expanded
// along its first row (the arithmetic of coordinates/matrix.hpp:57-108, so the
// cache matches the reference's to the last bit where the metric is
// ill-conditioned, e.g. inside the horizon).  Returns the determinant.
double minor3(const double m[16], int r0, int r1, int r2, int c0, int c1, int c2) {
    return m[4 * r0 + c0] * (m[4 * r1 + c1] * m[4 * r2 + c2] - m[4 * r2 + c1] * m[4 * r1 + c2]) -
           m[4 * r0 + c1] * (m[4 * r1 + c0] * m[4 * r2 + c2] - m[4 * r2 + c0] * m[4 * r1 + c2]) +
           m[4 * r0 + c2] * (m[4 * r1 + c0] * m[4 * r2 + c1] - m[4 * r2 + c0] * m[4 * r1 + c1]);
}
double invert4(const double mm[4][4], double inv[4][4]) {
    const double* m = &mm[0][0];
    static const int K[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            // (r, c) of the adjugate: signed minor without row c and column r
            const double mn = minor3(m, K[c][0], K[c][1], K[c][2], K[r][0], K[r][1], K[r][2]);
            inv[r][c] = ((r + c) & 1) ? -mn : mn;
        }
    const double det = m[0] * minor3(m, 1, 2, 3, 1, 2, 3) - m[1] * minor3(m, 1, 2, 3, 0, 2, 3) +
                       m[2] * minor3(m, 1, 2, 3, 0, 1, 3) - m[3] * minor3(m, 1, 2, 3, 0, 1, 2);
    const double id = 1. / det;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[r][c] *= id;
    return det;
}

// Christoffel symbols Gamma^lam_{nu mu} from centred differences of the native
// metric (CoordinateEmbedding::conn_native, coordinate_embedding.hpp:500-545)
void conn_native(const Embedding& E, const double X[4], double conn[4][4][4]) {
    double dg[4][4][4];   // dg[lam][kap][nu] = d_nu g_{lam kap}
    for (int nu = 0; nu < 4; ++nu) {
        double Xh[4], Xl[4], gh[4][4], gl[4][4];
        for (int mu = 0; mu < 4; ++mu) { Xl[mu] = X[mu] - DELTA * (mu == nu); Xh[mu] = X[mu] + DELTA * (mu == nu); }
        E.gcov_native(Xh, gh);
        E.gcov_native(Xl, gl);
        for (int lam = 0; lam < 4; ++lam)
            for (int kap = 0; kap < 4; ++kap) dg[lam][kap][nu] = (gh[lam][kap] - gl[lam][kap]) / (Xh[nu] - Xl[nu]);
    }
    double low[4][4][4];  // Gamma_{lam nu mu}
    for (int lam = 0; lam < 4; ++lam)
        for (int nu = 0; nu < 4; ++nu)
            for (int mu = 0; mu < 4; ++mu)
                low[lam][nu][mu] = 0.5 * (dg[nu][lam][mu] + dg[mu][lam][nu] - dg[mu][nu][lam]);
    double g[4][4], gi[4][4];
    E.gcov_native(X, g);
    invert4(g, gi);
    for (int lam = 0; lam < 4; ++lam)
        for (int nu = 0; nu < 4; ++nu)
            for (int mu = 0; mu < 4; ++mu) {
                double s = 0.;
                for (int kap = 0; kap < 4; ++kap) s += gi[lam][kap] * low[kap][nu][mu];
                conn[lam][nu][mu] = s;
            }
}

Embedding make_embedding(const kh_params& p) {
    Embedding E;
    E.spherical = (p.coords == KH_COORDS_KS);
    E.a = p.a;
    E.transform = p.transform;
    E.hslope = p.hslope; E.mks_smooth = p.mks_smooth; E.poly_xt = p.poly_xt; E.poly_alpha = p.poly_alpha;
    E.startx1 = p.fmks_zero_point;
    E.poly_norm = 0.5 * M_PI * 1. / (1. + 1. / (p.poly_alpha + 1.) * 1. / std::pow(p.poly_xt, p.poly_alpha));
    return E;
}

}  // namespace

// Native coordinates of locus `loc` at entire-block indices (j, i): cell
// centres at x_min + (idx - start + 1/2) dx, faces at x_min + (idx - start) dx
// (Parthenon UniformCartesian Xc/Xf), X3 at the first cell centre.
void locus_coords(const kh_params& p, const BlockDims& b, int loc, int j, int i, double X[4]) {
    const bool f1 = (loc == GLOC_FACE1 || loc == GLOC_CORNER);
    const bool f2 = (loc == GLOC_FACE2 || loc == GLOC_CORNER);
    X[0] = 0.;
    X[1] = p.x1min + ((i - b.is + b.goff[0]) + (f1 ? 0. : 0.5)) * b.dx1;
    X[2] = (b.ndim > 1) ? p.x2min + ((j - b.js + b.goff[1]) + (f2 ? 0. : 0.5)) * b.dx2 : 0.5 * (p.x2min + p.x2max);
    X[3] = p.x3min + 0.5 * b.dx3;
}

double embed_r(const kh_params& p, const double X[4]) {
    Embedding E = make_embedding(p);
    double Xe[4];
    E.to_embed(X, Xe);
    return E.spherical ? Xe[1] : std::sqrt(Xe[1] * Xe[1] + Xe[2] * Xe[2] + Xe[3] * Xe[3]);
}

void embed_coords(const kh_params& p, const double X[4], double Xe[4]) {
    make_embedding(p).to_embed(X, Xe);
}

void native_dxdX(const kh_params& p, const double X[4], double M[4][4]) {
    make_embedding(p).dxdX(X, M);
}

void native_gcov(const kh_params& p, const double X[4], double g[4][4]) {
    make_embedding(p).gcov_native(X, g);
}

double invert_metric(const double g[4][4], double gi[4][4]) { return invert4(g, gi); }

static const int SYM[10][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};

void build_geometry(const kh_params& p, const BlockDims& b, GeomHost& G) {
    const Embedding E = make_embedding(p);
    G.gn1 = b.n1 + 1;
    G.gn2 = b.n2 + 1;
    const long np = (long)G.gn1 * G.gn2;
    G.np = np;
    for (int loc = 0; loc < GLOC_N; ++loc) G.g[loc].assign((size_t)GEO_NCOMP * np, 0.);
    G.conn.assign((size_t)64 * np, 0.);
    G.flr.assign((size_t)2 * np, 0.);
    G.r.assign((size_t)np, 0.);
    for (int j = 0; j < G.gn2; ++j)
        for (int i = 0; i < G.gn1; ++i) {
            const long q = (long)j * G.gn1 + i;
            for (int loc = 0; loc < GLOC_N; ++loc) {
                double X[4], g[4][4], gi[4][4];
                locus_coords(p, b, loc, j, i, X);
                E.gcov_native(X, g);
                const double det = invert4(g, gi);
                double* o = G.g[loc].data();
                for (int c = 0; c < 10; ++c) {
                    o[(size_t)c * np + q] = g[SYM[c][0]][SYM[c][1]];
                    o[(size_t)(10 + c) * np + q] = gi[SYM[c][0]][SYM[c][1]];
                }
                const double gdet = std::sqrt(std::fabs(det));
                o[(size_t)GEO_GDET * np + q] = gdet;
                o[(size_t)GEO_ALPHA * np + q] = 1. / std::sqrt(-gi[0][0]);
                if (loc == GLOC_CENTER) {
                    if (E.spherical && j < b.n2 && i < b.n1) {
                        double cn[4][4][4];
                        conn_native(E, X, cn);
                        // Gamma^mu_{nu lam} is symmetric in (nu, lam); the centred
                        // differences leave the two orders apart by rounding (up to
                        // ~2e-11 relative), and the table holds their mean, exactly
                        // symmetric, so the source kernel reads each pair once.
                        for (int mu = 0; mu < 4; ++mu)
                            for (int nu = 0; nu < 4; ++nu)
                                for (int lam = 0; lam < 4; ++lam)
                                    G.conn[(size_t)(mu * 16 + nu * 4 + lam) * np + q] =
                                        gdet * (0.5 * (cn[mu][nu][lam] + cn[mu][lam][nu]));
                    }
                    // Geometric floors, floors_functions.hpp:97-108 / 312-323 (radius of the cell centre)
                    const double r = embed_r(p, X);
                    G.r[q] = r;
                    double rhoflr, uflr;
                    if (E.spherical) {
                        const double rhoscal = p.use_r_char ? 1. / ((r * r) * (1 + r / p.r_char)) : 1. / std::sqrt(r * r * r);
                        rhoflr = std::max(p.rho_min_geom * rhoscal, p.rho_min_const);
                        uflr = std::max(p.u_min_geom * std::pow(rhoscal, p.gam), p.u_min_const);
                    } else {
                        rhoflr = p.rho_min_const;
                        uflr = p.u_min_const;
                    }
                    G.flr[q] = rhoflr;
                    G.flr[(size_t)np + q] = uflr;
                }
            }
        }
}

}  // namespace kh
