/*
 * kharma_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++17 FP64 CPU restatement of KHARMA's explicit GRMHD step (the
 * "kharma" driver, RK2, flux-CT, onedw inverter, floors, fixup), written to be
 * the parity oracle for the HIP implementation in kharma_amd/.  Nothing in the
 * product links or calls this file: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker.
 *
 * Every function names the reference file:line it restates.  The closed-form
 * and invariant checks in tests/test_oracle.py pin reconstruction, P->U, the
 * onedw round trip, signal speeds, LLF, flux-CT (divB invariant), the dt ramp,
 * the fix-up weights, the drift-floor momentum invariant and the whole step
 * (linear-mode convergence).  Parity unpinned (restated, checked only end to
 * end): the onedw iterate sequence and exit tests, the drift-frame floor algebra
 * beyond its invariant, and the Parthenon pieces the reference tree does not
 * hold (RK2/VL2 weights, FluxDivHelper summation order, periodic fill).
 *
 * Layout: every field is SoA, [var][k][j][i] over the *entire* block
 * (interior + ng ghosts per side in each active dimension), i fastest,
 * exactly like Parthenon's VariablePack(v,k,j,i).  Fluxes are [dir-1][var][k][j][i]
 * with face index i meaning the left face of cell i (reference convention,
 * flux/get_flux.hpp, reconstruction.hpp:503-509).
 *
 * Build: oracle/Makefile -> oracle/liboracle.so (g++ -O2 -fopenmp -ffp-contract=off).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

#ifdef KH_ORACLE_TANGENT
// Forward-mode dual number (value, tangent): the tangent-linear build of this
// file (liboracle_tangent.so) carries one directional derivative through the
// step with every branch taken on the values.
struct Dual {
    double v = 0., d = 0.;
    constexpr Dual() = default;
    constexpr Dual(double x) : v(x), d(0.) {}
    constexpr Dual(double x, double dx) : v(x), d(dx) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
    Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
    Dual& operator*=(const Dual& o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    Dual& operator/=(const Dual& o) { const double q = v / o.v; d = (d - q * o.d) / o.v; v = q; return *this; }
};
inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
using Real = Dual;
inline double val(const Dual& x) { return x.v; }
#else
using Real = double;
#endif
inline double val(double x) { return x; }
constexpr int NV = 8;   // rho, u, u1, u2, u3, B1, B2, B3
enum { RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4, B1 = 5, B2 = 6, B3 = 7 };
// Loci, decs.hpp:103-104 (outer/inner_half are not used on this path)
enum { FACE1 = 0, FACE2 = 1, FACE3 = 2, CENTER = 3, CORNER = 4, NLOC = 5 };

// FFlag bits, floors/floors.hpp:47-68
constexpr int F_GEOM_RHO = 32, F_GEOM_U = 64, F_B_RHO = 128, F_B_U = 256, F_TEMP = 512,
              F_GAMMA = 1024, F_KTOT = 2048, F_GEOM_RHO_FLUX = 4096, F_GEOM_U_FLUX = 8192;
constexpr int FFLAG_MINIMUM = F_GEOM_RHO;
constexpr int F_INVERTER_RHO = 16384, F_INVERTER_U = 32768, F_INVERTER_GAMMA = 65536, F_INVERTER_U_MAX = 131072;
// Inverter::Status, inverter/invert_template.hpp:53
enum { S_SUCCESS = 0, S_NEG_INPUT, S_MAX_ITER, S_BAD_UT, S_BAD_GAMMA, S_NEG_RHO, S_NEG_U, S_NEG_RHOU };

constexpr Real SMALL = 1e-20;  // decs.hpp:90

inline Real SQR(Real x) { return x * x; }
// kharma_utils.hpp:100-110
inline Real clip(Real n, Real lower, Real upper) { return std::min(std::max(lower, n), upper); }
// coordinates/matrix.hpp:51-54
inline Real dot(const Real a[4], const Real b[4]) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2] + a[3]*b[3]; }

struct FourVectors { Real ucon[4], ucov[4], bcon[4], bcov[4]; };

}  // namespace

#ifdef KH_ORACLE_TANGENT
// Elementary functions of the dual number, found through the std:: calls of the
// restatement.
namespace std {
inline Dual sqrt(const Dual& a) { const double r = std::sqrt(a.v); return Dual(r, a.d / (2. * r)); }
inline Dual exp(const Dual& a) { const double e = std::exp(a.v); return Dual(e, e * a.d); }
inline Dual log(const Dual& a) { return Dual(std::log(a.v), a.d / a.v); }
inline Dual sin(const Dual& a) { return Dual(std::sin(a.v), std::cos(a.v) * a.d); }
inline Dual cos(const Dual& a) { return Dual(std::cos(a.v), -std::sin(a.v) * a.d); }
inline Dual fabs(const Dual& a) { return a.v < 0. ? Dual(-a.v, -a.d) : a; }
inline Dual abs(const Dual& a) { return fabs(a); }
inline Dual pow(const Dual& a, double p) {
    const double r = std::pow(a.v, p);
    return Dual(r, a.d == 0. ? 0. : p * std::pow(a.v, p - 1.) * a.d);
}
inline Dual pow(const Dual& a, const Dual& p) {
    const double r = std::pow(a.v, p.v);
    double d = 0.;
    if (a.d != 0.) d += p.v * std::pow(a.v, p.v - 1.) * a.d;
    if (p.d != 0.) d += r * std::log(a.v) * p.d;
    return Dual(r, d);
}
inline Dual pow(double a, const Dual& p) { return pow(Dual(a), p); }
inline const Dual& max(const Dual& a, const Dual& b) { return (a < b) ? b : a; }
inline const Dual& min(const Dual& a, const Dual& b) { return (b < a) ? b : a; }
inline Dual max(double a, const Dual& b) { return (Dual(a) < b) ? b : Dual(a); }
inline Dual max(const Dual& a, double b) { return (a < Dual(b)) ? Dual(b) : a; }
inline Dual min(double a, const Dual& b) { return (b < Dual(a)) ? b : Dual(a); }
inline Dual min(const Dual& a, double b) { return (Dual(b) < a) ? Dual(b) : a; }
inline bool isnan(const Dual& a) { return std::isnan(a.v); }
}  // namespace std
#endif

extern "C" {
// Parameters of one block / one run.  Mirrors the .par options the path reads.
struct or_params {
    int nx1, nx2, nx3, ng;            // interior sizes (nx3 == 1 for 2D), ghost depth
    double x1min, x1max, x2min, x2max, x3min, x3max;
    double gam;                        // GRMHD/gamma
    double cfl, dt_min, dt_start, max_dt_increase, tlim;
    int recon;                         // 0 donor_cell_c, 1 linear_mc, 2 weno5, 3 ppm
    int use_hlle;                      // flux/type
    int floors_on;                     // Floors package loaded
    int recon_floors;                  // flux/reconstruction_floors
    int floor_frame;                   // 0 fluid, 1 normal, 2 drift
    double rho_min_geom, u_min_geom, rho_min_const, u_min_const, r_char;
    double bsq_over_rho_max, bsq_over_u_max, u_over_rho_max, ktot_max, gamma_max;
    int use_r_char, temp_adjust_u;
    int iter_max;                      // inverter/iter_max
    double err_tol;                    // inverter/err_tol
    int fix_average, fix_atmo;         // inverter/fix_average_neighbors, fix_atmosphere
    int integrator;                    // 0 rk2, 1 vl2
    int spherical;                     // coordinates/base: 0 cartesian_minkowski, 1 spherical_ks
    double a;                          // coordinates/a
    int transform;                     // coordinates/transform: 0 null, 1 mks, 2 fmks
    double hslope, mks_smooth, poly_xt, poly_alpha, fmks_zero_point;
    int bc_inner_x1, bc_outer_x1, bc_inner_x2, bc_outer_x2;   // 1 outflow, 2 reflecting
    int check_inflow_inner_x1, check_inflow_outer_x1;
    int zero_polar_flux, fix_polar_flux, fix_corner_inner;
    int periodic1, periodic2, periodic3;
    int driver;                        // driver/type: 0 kharma, 1 simple
    int inverter;                      // inverter/type: 0 onedw, 1 kastaun
    int fofc, fofc_polar_cells;        // fofc/on, fofc/polar_cells
    double fofc_eh_buffer;             // fofc/eh_buffer
    int face_ct;                       // b_field/solver: 0 flux_ct, 1 face_ct (the B_CT package)
    int ct_scheme;                     // b_field/ct_scheme: 0 bs99, 1 gs05_0, 2 gs05_c, 3 sg07
};
}

namespace {

struct Block {
    or_params p;
    int ndim;
    int n1, n2, n3;          // entire sizes
    int is, ie, js, je, ks, ke;
    size_t nc;               // cells per var
    Real dx[3];
    // Geometry cache, gr_coordinates.hpp:86-90: (loc, j, i, mu, nu) over (n2+1) x (n1+1)
    int gn1, gn2;
    std::vector<Real> gcov, gcon, gdet, gdet_conn;
    std::vector<Real> rcen;   // embedding radius of each (j, i) cell centre
    bool cart_minkowski = true;
    // Containers: base (full step init / final), s1 (stage 1)
    std::vector<Real> U0, P0, U1c, P1c;
    std::vector<Real> F;      // [3][NV][n3][n2][n1] fluxes (OneCopy-style: shared by containers)
    std::vector<Real> cmax, cmin;  // [3][...]
    std::vector<Real> Pl, Pr, Ul, Ur, Fl, Fr;  // Flux.* temporaries [NV][...]
    std::vector<Real> emf;   // [3][...]
    std::vector<Real> dUdt;  // [NV][...]
    std::vector<Real> pflag, fflag, rho_floor, u_floor;
    // First-order flux correction, last call of fofc(): interior marks, the trial
    // primitives and each trial state's relative distance to its nearest floor threshold
    std::vector<Real> fofc_mark, fofc_trial, fofc_margin;
    // B_CT: face fields cons.fB of the base and stage-1 containers, edge EMFs
    // B_CT.emf, their rate dB_Uf/dt and the cell-centre EMFs B_CT.cemf.  Face and
    // edge arrays are [3][n3+1][n2+1][n1+1] (component F1/F2/F3 or E1/E2/E3).
    int fn1 = 0, fn2 = 0, fn3 = 0;
    size_t nf = 0;
    std::vector<Real> fB0, fB1c, femf, dfB, cemf;
    double time = 0., dt = 0., dt_last = 0.;
    long ncycle = 0;
    bool in_loop = false;

    size_t id(int k, int j, int i) const { return ((size_t)k * n2 + j) * n1 + i; }
    size_t fid(int k, int j, int i) const { return ((size_t)k * fn2 + j) * fn1 + i; }
    size_t gidx(int loc, int j, int i) const { return ((size_t)loc * gn2 + j) * gn1 + i; }
    Real gcov_(int loc, int j, int i, int mu, int nu) const { return gcov[gidx(loc, j, i) * 16 + mu * 4 + nu]; }
    Real gcon_(int loc, int j, int i, int mu, int nu) const { return gcon[gidx(loc, j, i) * 16 + mu * 4 + nu]; }
    Real gdet_(int loc, int j, int i) const { return gdet[gidx(loc, j, i)]; }
    Real gdet_conn_(int j, int i, int mu, int nu, int lam) const {
        return gdet_conn[(((size_t)j * gn1 + i) * 4 + mu) * 16 + nu * 4 + lam];
    }

    // GRCoordinates::lower/raise, gr_coordinates.hpp:227-238
    void lower(const Real vcon[4], Real vcov[4], int j, int i, int loc) const {
        for (int mu = 0; mu < 4; ++mu) vcov[mu] = 0.;
        for (int mu = 0; mu < 4; ++mu) for (int nu = 0; nu < 4; ++nu) vcov[mu] += gcov_(loc, j, i, mu, nu) * vcon[nu];
    }
    void raise(const Real vcov[4], Real vcon[4], int j, int i, int loc) const {
        for (int mu = 0; mu < 4; ++mu) vcon[mu] = 0.;
        for (int mu = 0; mu < 4; ++mu) for (int nu = 0; nu < 4; ++nu) vcon[mu] += gcon_(loc, j, i, mu, nu) * vcov[nu];
    }
};

// ------------------------------------------------------------------ geometry
// Cartesian Minkowski: coordinate_systems.hpp:77-86 (gcov = diag(-1,1,1,1)); the
// cache is filled the way init_GRCoordinates does (gr_coordinates.cpp:120-205):
// gcon = inverse, gdet = sqrt|det g| = 1, connection = 0.
void init_geometry_minkowski(Block& b) {
    b.gn1 = b.n1 + 1; b.gn2 = b.n2 + 1;
    const size_t ng = (size_t)b.gn1 * b.gn2;
    b.gcov.assign(NLOC * ng * 16, 0.); b.gcon.assign(NLOC * ng * 16, 0.); b.gdet.assign(NLOC * ng, 1.);
    b.gdet_conn.assign(ng * 64, 0.);
    for (size_t c = 0; c < NLOC * ng; ++c) {
        for (int mu = 0; mu < 4; ++mu) {
            b.gcov[c * 16 + mu * 5] = (mu == 0) ? -1. : 1.;
            b.gcon[c * 16 + mu * 5] = (mu == 0) ? -1. : 1.;
        }
    }
    b.cart_minkowski = true;
}

// --------------------------------------------------------- reconstruction
// reconstruction.hpp:46 (EPS) and 128-204 (weno5 both sides / left / right)
constexpr Real WENO_EPS = 1.e-26;
void weno5(Real x1, Real x2, Real x3, Real x4, Real x5, Real& lout, Real& rout) {
    Real beta[3], c1, c2;
    c1 = x1 - 2.*x2 + x3; c2 = x1 - 4.*x2 + 3.*x3;
    beta[0] = (13./12.)*c1*c1 + (1./4.)*c2*c2;
    c1 = x2 - 2.*x3 + x4; c2 = x4 - x2;
    beta[1] = (13./12.)*c1*c1 + (1./4.)*c2*c2;
    c1 = x3 - 2.*x4 + x5; c2 = x5 - 4.*x4 + 3.*x3;
    beta[2] = (13./12.)*c1*c1 + (1./4.)*c2*c2;
    Real den[3] = {WENO_EPS + beta[0], WENO_EPS + beta[1], WENO_EPS + beta[2]};
    den[0] *= den[0]; den[1] *= den[1]; den[2] *= den[2];
    Real wtr[3] = {(1./16.)/den[0], (5./8.)/den[1], (5./16.)/den[2]};
    Real Wr = wtr[0] + wtr[1] + wtr[2];
    Real wtl[3] = {(1./16.)/den[2], (5./8.)/den[1], (5./16.)/den[0]};
    Real Wl = wtl[0] + wtl[1] + wtl[2];
    lout = ((3./8.)*x5 - (5./4.)*x4 + (15./8.)*x3)*(wtl[0] / Wl) +
           ((-1./8.)*x4 + (3./4.)*x3 + (3./8.)*x2)*(wtl[1] / Wl) +
           ((3./8.)*x3 + (3./4.)*x2 - (1./8.)*x1)*(wtl[2] / Wl);
    rout = ((3./8.)*x1 - (5./4.)*x2 + (15./8.)*x3)*(wtr[0] / Wr) +
           ((-1./8.)*x2 + (3./4.)*x3 + (3./8.)*x4)*(wtr[1] / Wr) +
           ((3./8.)*x3 + (3./4.)*x4 - (1./8.)*x5)*(wtr[2] / Wr);
}
// reconstruction.hpp:52-56
Real mc_lim(Real dm, Real dp) {
    const Real r = (std::fabs(dp) > 0. ? dm/dp : 2.0);
    return std::max(0.0, std::min(2.0, std::min(2*r, 0.5*(1+r))));
}
// reconstruction.hpp:107-112
void linear_mc(Real, Real x2, Real x3, Real x4, Real, Real& lout, Real& rout) {
    const Real dq = mc_lim(x3 - x2, x4 - x3)*(x4 - x3);
    rout = x3 + 0.5*dq;
    lout = x3 - 0.5*dq;
}
// reconstruction.hpp:351-382 (PPM, CW84 limiter)
void ppm(Real q_im2, Real q_im1, Real q_i, Real q_ip1, Real q_ip2, Real& qlv, Real& qrv) {
    qlv = (7.*(q_i + q_im1) - (q_im2 + q_ip1))/12.0;
    qrv = (7.*(q_i + q_ip1) - (q_im1 + q_ip2))/12.0;
    qlv = std::max(qlv, std::min(q_i, q_im1));
    qlv = std::min(qlv, std::max(q_i, q_im1));
    qrv = std::max(qrv, std::min(q_i, q_ip1));
    qrv = std::min(qrv, std::max(q_i, q_ip1));
    Real qc = qrv - q_i;
    Real qd = qlv - q_i;
    if ((qc*qd) >= 0.0) {
        qlv = q_i; qrv = q_i;
    } else {
        if (std::fabs(qc) >= 2.0*std::fabs(qd)) qrv = q_i - 2.0*qd;
        if (std::fabs(qd) >= 2.0*std::fabs(qc)) qlv = q_i - 2.0*qc;
    }
}
// dispatch: "lout" is the value at the left face of the centre cell, "rout" at its right face
void reconstruct(int recon, Real x1, Real x2, Real x3, Real x4, Real x5, Real& lout, Real& rout) {
    switch (recon) {
    case 0: lout = x3; rout = x3; break;            // donor_cell_c, :92-96
    case 1: linear_mc(x1, x2, x3, x4, x5, lout, rout); break;
    case 3: ppm(x1, x2, x3, x4, x5, lout, rout); break;
    default: weno5(x1, x2, x3, x4, x5, lout, rout); break;
    }
}

// Spherical Kerr-Schild base with null / "modified" / "funky" transforms:
// coordinate_systems.hpp:113-151 (gcov_embed), :351-672 (transforms),
// coordinate_embedding.hpp:461-545 (gcov_native, conn_native), matrix.hpp:57-108
// (adjoint/determinant inverse), cached as init_GRCoordinates does
// (gr_coordinates.cpp:96-205, connection_average_points = 1).
struct KSEmbed {
    Real a, hslope, smooth, xt, alpha, x0, norm;
    int transform;
    void embed(const Real X[4], Real& r, Real& th) const {
        if (transform == 0) { r = X[1]; th = X[2]; return; }
        r = std::exp(X[1]);
        const Real thG = M_PI*X[2] + ((1. - hslope)/2.)*std::sin(2.*M_PI*X[2]);
        if (transform == 1) { th = thG; return; }
        const Real y = 2*X[2] - 1.;
        const Real thJ = norm * y * (1. + std::pow(y/xt, alpha) / (alpha + 1.)) + 0.5 * M_PI;
        th = thG + std::exp(smooth * (x0 - X[1])) * (thJ - thG);
    }
    void dxdX(const Real X[4], Real d[4][4]) const {
        for (int m = 0; m < 4; ++m) for (int n = 0; n < 4; ++n) d[m][n] = (m == n);
        if (transform == 0) return;
        d[1][1] = std::exp(X[1]);
        if (transform == 1) { d[2][2] = M_PI - (hslope - 1.)*M_PI*std::cos(2.*M_PI*X[2]); return; }
        const Real e = std::exp(smooth * (x0 - X[1]));
        d[2][1] = -e * smooth * (M_PI/2. - M_PI*X[2] + norm*(2.*X[2] - 1.)*(1 + std::pow((-1. + 2*X[2])/xt, alpha)/(1 + alpha))
                                 - 1./2.*(1. - hslope)*std::sin(2.*M_PI*X[2]));
        d[2][2] = M_PI + (1. - hslope)*M_PI*std::cos(2.*M_PI*X[2])
                + e * (-M_PI + 2.*norm*(1. + std::pow((2.*X[2] - 1.)/xt, alpha)/(alpha + 1.))
                       + (2.*alpha*norm*(2.*X[2] - 1.)*std::pow((2.*X[2] - 1.)/xt, alpha - 1.))/((1. + alpha)*xt)
                       - (1. - hslope)*M_PI*std::cos(2.*M_PI*X[2]));
    }
    void gcov_ks(Real r, Real th_in, Real g[4][4]) const {
        auto excise = [](Real n, Real c, Real w) { return (std::fabs(n - c) > w) ? n : ((n > c) ? c + w : c - w); };
        const Real th = excise(excise(th_in, 0.0, 1e-20), M_PI, 1e-20);
        const Real cth = std::cos(th), sth = std::sin(th), sin2 = sth*sth, rho2 = r*r + a*a*cth*cth;
        g[0][0] = -1. + 2.*r/rho2;  g[0][1] = 2.*r/rho2;  g[0][2] = 0.;  g[0][3] = -2.*a*r*sin2/rho2;
        g[1][0] = 2.*r/rho2;  g[1][1] = 1. + 2.*r/rho2;  g[1][2] = 0.;  g[1][3] = -a*sin2*(1. + 2.*r/rho2);
        g[2][0] = 0.;  g[2][1] = 0.;  g[2][2] = rho2;  g[2][3] = 0.;
        g[3][0] = -2.*a*r*sin2/rho2;  g[3][1] = -a*sin2*(1. + 2.*r/rho2);  g[3][2] = 0.;
        g[3][3] = sin2*(rho2 + a*a*sin2*(1. + 2.*r/rho2));
    }
    void gcov_native(const Real X[4], Real g[4][4]) const {
        Real r, th, ge[4][4], d[4][4];
        embed(X, r, th);
        gcov_ks(r, th, ge);
        dxdX(X, d);
        for (int mu = 0; mu < 4; ++mu) for (int nu = 0; nu < 4; ++nu) {
            g[mu][nu] = 0;
            for (int lam = 0; lam < 4; ++lam) for (int kap = 0; kap < 4; ++kap)
                g[mu][nu] += ge[lam][kap]*d[lam][mu]*d[kap][nu];
        }
    }
};
Real minor3(const Real m[16], int r0, int r1, int r2, int c0, int c1, int c2) {
    return m[4*r0+c0]*(m[4*r1+c1]*m[4*r2+c2] - m[4*r2+c1]*m[4*r1+c2]) -
           m[4*r0+c1]*(m[4*r1+c0]*m[4*r2+c2] - m[4*r2+c0]*m[4*r1+c2]) +
           m[4*r0+c2]*(m[4*r1+c0]*m[4*r2+c1] - m[4*r2+c0]*m[4*r1+c1]);
}
// inverse = adjugate / determinant; returns the determinant
Real invert4(const Real* m, Real* inv) {
    static const int R[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) {
        // adj[i][j] = (-1)^(i+j) * minor deleting row j, column i
        const Real mn = minor3(m, R[j][0], R[j][1], R[j][2], R[i][0], R[i][1], R[i][2]);
        inv[4*i + j] = ((i + j) % 2) ? -mn : mn;
    }
    const Real det = m[0]*minor3(m,1,2,3,1,2,3) - m[1]*minor3(m,1,2,3,0,2,3) +
                     m[2]*minor3(m,1,2,3,0,1,3) - m[3]*minor3(m,1,2,3,0,1,2);
    const Real id = 1. / det;
    for (int c = 0; c < 16; ++c) inv[c] *= id;
    return det;
}
void init_geometry_ks(Block& b) {
    const or_params& p = b.p;
    KSEmbed E{p.a, p.hslope, p.mks_smooth, p.poly_xt, p.poly_alpha, p.fmks_zero_point, 0., p.transform};
    E.norm = 0.5 * M_PI * 1./(1. + 1./(p.poly_alpha + 1.) * 1./std::pow(p.poly_xt, p.poly_alpha));
    b.gn1 = b.n1 + 1; b.gn2 = b.n2 + 1;
    const size_t ng = (size_t)b.gn1 * b.gn2;
    b.gcov.assign(NLOC * ng * 16, 0.); b.gcon.assign(NLOC * ng * 16, 0.); b.gdet.assign(NLOC * ng, 0.);
    b.gdet_conn.assign(ng * 64, 0.);
    b.rcen.assign(ng, 0.);
    const Real d1 = b.dx[0], d2 = b.dx[1];
    for (int j = 0; j < b.gn2; ++j)
    for (int i = 0; i < b.gn1; ++i) {
        for (int loc = 0; loc < NLOC; ++loc) {
            // GRCoordinates::coord (gr_coordinates.hpp:184-225) at k = 0
            const bool f1 = (loc == FACE1 || loc == CORNER), f2 = (loc == FACE2 || loc == CORNER);
            Real X[4] = {0., p.x1min + (i - b.is + (f1 ? 0. : 0.5)) * d1,
                         b.ndim > 1 ? p.x2min + (j - b.js + (f2 ? 0. : 0.5)) * d2 : 0.5*(p.x2min + p.x2max),
                         p.x3min + 0.5 * b.dx[2]};
            Real g[4][4], gi[4][4];
            E.gcov_native(X, g);
            const Real det = invert4(&g[0][0], &gi[0][0]);
            const size_t q = b.gidx(loc, j, i);
            for (int mu = 0; mu < 4; ++mu) for (int nu = 0; nu < 4; ++nu) {
                b.gcov[q*16 + mu*4 + nu] = g[mu][nu];
                b.gcon[q*16 + mu*4 + nu] = gi[mu][nu];
            }
            const Real gdet = std::sqrt(std::fabs(det));
            b.gdet[q] = gdet;
            if (loc == CENTER) {
                Real r, th; E.embed(X, r, th);
                b.rcen[(size_t)j * b.gn1 + i] = r;
                if (j < b.n2 && i < b.n1) {
                    // conn_native with delta = 1e-8 (gr_coordinates.cpp:51)
                    Real conn[4][4][4], tmp[4][4][4];
                    for (int nu = 0; nu < 4; ++nu) {
                        Real Xl[4], Xh[4], gh[4][4], gl[4][4];
                        for (int mu = 0; mu < 4; ++mu) { Xl[mu] = X[mu] - 1e-8*(mu == nu); Xh[mu] = X[mu] + 1e-8*(mu == nu); }
                        E.gcov_native(Xh, gh); E.gcov_native(Xl, gl);
                        for (int lam = 0; lam < 4; ++lam) for (int kap = 0; kap < 4; ++kap)
                            conn[lam][kap][nu] = (gh[lam][kap] - gl[lam][kap]) / (Xh[nu] - Xl[nu]);
                    }
                    for (int lam = 0; lam < 4; ++lam) for (int nu = 0; nu < 4; ++nu) for (int mu = 0; mu < 4; ++mu)
                        tmp[lam][nu][mu] = 0.5 * (conn[nu][lam][mu] + conn[mu][lam][nu] - conn[mu][nu][lam]);
                    for (int lam = 0; lam < 4; ++lam) for (int nu = 0; nu < 4; ++nu) for (int mu = 0; mu < 4; ++mu) {
                        Real c = 0.;
                        for (int kap = 0; kap < 4; ++kap) c += gi[lam][kap] * tmp[kap][nu][mu];
                        b.gdet_conn[(((size_t)j * b.gn1 + i) * 4 + lam) * 16 + nu * 4 + mu] = gdet * c;
                    }
                }
            }
        }
    }
    b.cart_minkowski = false;
}

// Geometric floors at cell centre (j, i): floors_functions.hpp:97-108 / 312-323
void geom_floors(const Block& b, int j, int i, Real& rhoflr, Real& uflr) {
    if (b.p.spherical) {
        const Real r = b.rcen[(size_t)j * b.gn1 + i];
        const Real rhoscal = b.p.use_r_char ? 1. / ((r*r) * (1 + r / b.p.r_char)) : 1. / std::sqrt(r*r*r);
        rhoflr = std::max(b.p.rho_min_geom * rhoscal, b.p.rho_min_const);
        uflr = std::max(b.p.u_min_geom * std::pow(rhoscal, b.p.gam), b.p.u_min_const);
    } else {
        rhoflr = b.p.rho_min_const;
        uflr = b.p.u_min_const;
    }
}

// ----------------------------------------------------------- GRMHD physics
// grmhd_functions.hpp:116-129 (Local lorentz_calc)
Real lorentz_calc(const Block& b, const Real* P, int j, int i, int loc) {
    const Real qsq = b.gcov_(loc, j, i, 1, 1) * P[U1] * P[U1] +
                     b.gcov_(loc, j, i, 2, 2) * P[U2] * P[U2] +
                     b.gcov_(loc, j, i, 3, 3) * P[U3] * P[U3] +
                     2. * (b.gcov_(loc, j, i, 1, 2) * P[U1] * P[U2] +
                           b.gcov_(loc, j, i, 1, 3) * P[U1] * P[U3] +
                           b.gcov_(loc, j, i, 2, 3) * P[U2] * P[U3]);
    return std::sqrt(1. + qsq);
}
// grmhd_functions.hpp:225-246 (Local calc_4vecs, B present)
void calc_4vecs(const Block& b, const Real* P, int j, int i, int loc, FourVectors& D) {
    const Real gamma = lorentz_calc(b, P, j, i, loc);
    const Real alpha = 1. / std::sqrt(-b.gcon_(loc, j, i, 0, 0));
    D.ucon[0] = gamma / alpha;
    for (int v = 0; v < 3; ++v) D.ucon[v+1] = P[U1 + v] - gamma * alpha * b.gcon_(loc, j, i, 0, v+1);
    b.lower(D.ucon, D.ucov, j, i, loc);
    D.bcon[0] = 0;
    for (int v = 0; v < 3; ++v) D.bcon[0] += P[B1 + v] * D.ucov[v+1];
    for (int v = 0; v < 3; ++v) D.bcon[v+1] = (P[B1 + v] + D.bcon[0] * D.ucon[v+1]) / D.ucon[0];
    b.lower(D.bcon, D.bcov, j, i, loc);
}
// grmhd_functions.hpp:142-155
void calc_tensor_mhd(Real rho, Real u, Real pgas, const FourVectors& D, int dir, Real mhd[4]) {
    const Real bsq = dot(D.bcon, D.bcov);
    const Real eta = pgas + rho + u + bsq;
    const Real ptot = pgas + 0.5 * bsq;
    for (int mu = 0; mu < 4; ++mu)
        mhd[mu] = eta * D.ucon[dir] * D.ucov[mu] + ptot * (dir == mu) - D.bcon[dir] * D.bcov[mu];
}
// flux/flux_functions.hpp:116-178 (Local prim_to_flux, GRMHD + B_FluxCT variables only)
void prim_to_flux(const Block& b, const Real* P, const FourVectors& D, int j, int i, int dir,
                  Real* flux, int loc, bool with_B = true) {
    const Real gam = b.p.gam;
    const Real gdet = b.gdet_(loc, j, i);
    flux[RHO] = P[RHO] * D.ucon[dir] * gdet;
    Real T[4];
    calc_tensor_mhd(P[RHO], P[UU], (gam - 1) * P[UU], D, dir, T);
    flux[UU] = T[0] * gdet + flux[RHO];
    flux[U1] = T[1] * gdet;
    flux[U2] = T[2] * gdet;
    flux[U3] = T[3] * gdet;
    if (!with_B) return;
    if (dir == 0) {
        for (int v = 0; v < 3; ++v) flux[B1 + v] = P[B1 + v] * gdet;
    } else {
        for (int v = 0; v < 3; ++v) flux[B1 + v] = (D.bcon[v+1] * D.ucon[dir] - D.bcon[dir] * D.ucon[v+1]) * gdet;
    }
}
// flux/flux_functions.hpp:368-439.  Note both clip() calls are no-ops in the
// reference (their results are discarded, :378 and :409), so cs2/cms2 are NOT clamped.
void vchar(const Block& b, const Real* P, const FourVectors& D, int j, int i, int loc, int dir,
           Real& cmax, Real& cmin) {
    const Real gam = b.p.gam;
    const Real ef = P[RHO] + gam * P[UU];
    const Real cs2 = gam * (gam - 1) * P[UU] / ef;
    const Real bsq = dot(D.bcon, D.bcov);
    const Real va2 = bsq / (bsq + ef);
    const Real cms2 = cs2 + va2 - cs2 * va2;
    Real A, B, C;
    {
        Real Bcov[4] = {1., 0., 0., 0.};
        Real Acov[4] = {0., 0., 0., 0.}; Acov[dir] = 1.;
        Real Acon[4], Bcon[4];
        b.raise(Acov, Acon, j, i, loc);
        b.raise(Bcov, Bcon, j, i, loc);
        const Real Asq = dot(Acon, Acov);
        const Real Bsq = dot(Bcon, Bcov);
        const Real Au = dot(Acov, D.ucon);
        const Real Bu = dot(Bcov, D.ucon);
        const Real AB = dot(Acon, Bcov);
        A = Bu*Bu - (Bsq + Bu*Bu) * cms2;
        B = 2. * (Au*Bu - (AB + Au*Bu) * cms2);
        C = Au*Au - (Asq + Au*Au) * cms2;
    }
    Real discr = std::sqrt(std::max(B * B - 4. * A * C, 0.));
    Real vp = -(-B + discr) / (2. * A);
    Real vm = -(-B - discr) / (2. * A);
    cmax = std::max(vp, vm);
    cmin = std::min(vp, vm);
}
// flux/flux.hpp:110-121
inline Real llf(Real fL, Real fR, Real cmax, Real cmin, Real Ul, Real Ur) {
    Real ctop = std::max(cmax, cmin);
    return 0.5 * (fL + fR - ctop * (Ur - Ul));
}
inline Real hlle(Real fL, Real fR, Real cmax, Real cmin, Real Ul, Real Ur) {
    return (cmax*fL + cmin*fR - cmax*cmin*(Ur - Ul)) / (cmax + cmin);
}

// -------------------------------------------------------------- floors
// floors/floors_functions.hpp:300-332 (Local apply_geo_floors): floors of the
// cell centre (j, i) whatever the face locus (G.r(0, j, i) at :311).
int apply_geo_floors(const Block& b, Real* P, int j, int i) {
    Real rhoflr_geom, uflr_geom;
    geom_floors(b, j, i, rhoflr_geom, uflr_geom);
    int fflag = 0;
    fflag |= (rhoflr_geom > P[RHO]) * F_GEOM_RHO_FLUX;
    fflag |= (uflr_geom > P[UU]) * F_GEOM_U_FLUX;
    P[RHO] += std::max(0., rhoflr_geom - P[RHO]);
    P[UU]  += std::max(0., uflr_geom - P[UU]);
    return fflag;
}

// ---------------------------------------------------------------- helpers
inline void load_prims(const Block& b, const std::vector<Real>& P, size_t c, Real out[NV]) {
    for (int v = 0; v < NV; ++v) out[v] = P[v * b.nc + c];
}

// =================================================================== tasks

// Flux::GetFlux<Recon,dir>, flux/get_flux.hpp:60-373.  Range: GetRange(interior,
// FaceOf(dir), -1, +1) (:124 + domain.hpp:102-124).  The X1 left state of the
// first face of that range is left unset by the reference's row
// reconstruction (ReconstructX1 writes ql(i+1), reconstruction.hpp:511-527)
// and nothing downstream reads that face; here it is reconstructed from cell is-2.
void get_flux(Block& b, const std::vector<Real>& P, int dir) {
    if (b.ndim < 3 && dir == 3) return;
    if (b.ndim < 2 && dir == 2) return;
    const int loc = dir - 1;  // face1/2/3
    int is = b.is - 1, ie = b.ie + 1, js = b.js, je = b.je, ks = b.ks, ke = b.ke;
    if (b.ndim > 1) { js = b.js - 1; je = b.je + 1; }
    if (b.ndim > 2) { ks = b.ks - 1; ke = b.ke + 1; }
    if (dir == 1) ie += 1;  // F1 bounds are [is, ie+1]
    if (dir == 2) je += 1;
    if (dir == 3) ke += 1;
    const int s1 = (dir == 1) ? 1 : 0, s2 = (dir == 2) ? b.n1 : 0, s3 = (dir == 3) ? b.n1 * b.n2 : 0;
    const long st = s1 + s2 + (long)s3;
    const Real gam = b.p.gam;
    (void)gam;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
        for (int i = is; i <= ie; ++i) {
            const size_t c = b.id(k, j, i);
            Real Pl[NV], Pr[NV];
            for (int v = 0; v < NV; ++v) {
                const Real* q = &P[v * b.nc];
                Real l_unused, r_unused;
                // left state at this face = right edge of the cell below (c - st)
                reconstruct(b.p.recon, q[c - 3*st], q[c - 2*st], q[c - st], q[c], q[c + st], l_unused, Pl[v]);
                // right state = left edge of this cell
                reconstruct(b.p.recon, q[c - 2*st], q[c - st], q[c], q[c + st], q[c + 2*st], Pr[v], r_unused);
            }
            // get_flux.hpp:170-182
            if (b.p.recon_floors) {
                apply_geo_floors(b, Pl, j, i);
                apply_geo_floors(b, Pr, j, i);
            }
            for (int v = 0; v < NV; ++v) { b.Pl[v * b.nc + c] = Pl[v]; b.Pr[v * b.nc + c] = Pr[v]; }
            Real Ul[NV], Ur[NV], Fl[NV], Fr[NV];
            FourVectors D;
            // calc_flux_left, get_flux.hpp:256-277
            calc_4vecs(b, Pl, j, i, loc, D);
            prim_to_flux(b, Pl, D, j, i, 0, Ul, loc);
            prim_to_flux(b, Pl, D, j, i, dir, Fl, loc);
            Real cmaxL, cminL;
            vchar(b, Pl, D, j, i, loc, dir, cmaxL, cminL);
            Real cmx = std::max(0., cmaxL);
            Real cmn = std::min(0., cminL);
            // calc_flux_right, get_flux.hpp:313-333
            calc_4vecs(b, Pr, j, i, loc, D);
            prim_to_flux(b, Pr, D, j, i, 0, Ur, loc);
            prim_to_flux(b, Pr, D, j, i, dir, Fr, loc);
            Real cmaxR, cminR;
            vchar(b, Pr, D, j, i, loc, dir, cmaxR, cminR);
            cmx = std::max(cmx, cmaxR);
            cmn = -std::min(cmn, cminR);
            b.cmax[(dir-1) * b.nc + c] = cmx;
            b.cmin[(dir-1) * b.nc + c] = cmn;
            // flux_llf / flux_hlle, get_flux.hpp:353-367
            for (int v = 0; v < NV; ++v) {
                b.Ul[v * b.nc + c] = Ul[v]; b.Ur[v * b.nc + c] = Ur[v];
                b.Fl[v * b.nc + c] = Fl[v]; b.Fr[v * b.nc + c] = Fr[v];
                b.F[((size_t)(dir-1) * NV + v) * b.nc + c] = b.p.use_hlle
                    ? hlle(Fl[v], Fr[v], cmx, cmn, Ul[v], Ur[v])
                    : llf(Fl[v], Fr[v], cmx, cmn, Ul[v], Ur[v]);
            }
        }
    }
}

// B_FluxCT::FluxCT, b_flux_ct/b_flux_ct.cpp:307-370
void flux_ct(Block& b) {
    if (b.ndim < 2) return;
    const int nd = b.ndim;
    auto F = [&](int dir, int v, int k, int j, int i) -> Real& {
        return b.F[((size_t)(dir-1) * NV + v) * b.nc + b.id(k, j, i)];
    };
    auto E = [&](int v, int k, int j, int i) -> Real& { return b.emf[(size_t)v * b.nc + b.id(k, j, i)]; };
    const int ils = b.is, ile = b.ie + 1, jls = b.js, jle = b.je + 1;
    const int kls = b.ks, kle = (nd > 2) ? b.ke + 1 : b.ke;
#pragma omp parallel for collapse(2)
    for (int k = kls; k <= kle; ++k)
    for (int j = jls; j <= jle; ++j)
    for (int i = ils; i <= ile; ++i) {
        if (nd > 2) {
            E(0, k, j, i) = 0.25 * (F(2, B3, k, j, i) + F(2, B3, k-1, j, i) -
                                    F(3, B2, k, j, i) - F(3, B2, k, j-1, i));
            E(1, k, j, i) = 0.25 * (F(3, B1, k, j, i) + F(3, B1, k, j, i-1) -
                                    F(1, B3, k, j, i) - F(1, B3, k-1, j, i));
        }
        E(2, k, j, i) = 0.25 * (F(1, B2, k, j, i) + F(1, B2, k, j-1, i) -
                                F(2, B1, k, j, i) - F(2, B1, k, j, i-1));
    }
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = b.js; j <= b.je; ++j)
    for (int i = ils; i <= ile; ++i) {
        F(1, B1, k, j, i) = 0.0;
        F(1, B2, k, j, i) = 0.5 * (E(2, k, j, i) + E(2, k, j+1, i));
        if (nd > 2) F(1, B3, k, j, i) = -0.5 * (E(1, k, j, i) + E(1, k+1, j, i));
    }
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = jls; j <= jle; ++j)
    for (int i = b.is; i <= b.ie; ++i) {
        F(2, B1, k, j, i) = -0.5 * (E(2, k, j, i) + E(2, k, j, i+1));
        F(2, B2, k, j, i) = 0.0;
        if (nd > 2) F(2, B3, k, j, i) = 0.5 * (E(0, k, j, i) + E(0, k+1, j, i));
    }
    if (nd > 2) {
#pragma omp parallel for collapse(2)
        for (int k = kls; k <= kle; ++k)
        for (int j = b.js; j <= b.je; ++j)
        for (int i = b.is; i <= b.ie; ++i) {
            F(3, B1, k, j, i) = 0.5 * (E(1, k, j, i) + E(1, k, j, i+1));
            F(3, B2, k, j, i) = -0.5 * (E(0, k, j, i) + E(0, k, j+1, i));
            F(3, B3, k, j, i) = 0.0;
        }
    }
}

// A deliberate mistake planted in the restatement by a test (or_plant), to show
// that the test's pin fails on it.  0: none.  1: the geometric source contracts
// T^mu_nu with Gamma^lam_{nu mu} (upper and first lower index exchanged) instead
// of Gamma^nu_{lam mu}.  2: the drift-frame floors solve for the parallel
// velocity with the pre-floor enthalpy.  3: B_CT's AddSource takes the F1 face's
// upper E3 edge one cell along x1 instead of along x2.  4: B_CT's gs05_c
// upwinding takes the downwind side.
static int g_plant = 0;

// ============================================================== B_CT
// Face-centred constrained transport (b_ct/b_ct.cpp).  Parthenon's
// UniformCartesian element sizes (absent from the tree): an edge along x_d has
// length dx_d, a face normal to x_d area dx_e dx_f, whatever its position.
inline Real& fface(Block& b, std::vector<Real>& A, int el, int k, int j, int i) {
    return A[(size_t)el * b.nf + b.fid(k, j, i)];
}

// B_CT::CalculateEMF, b_ct.cpp:263-416: the BS99 average of the B fluxes onto
// each edge (the whole of bs99), then the GS05 corrections from the cell-centre
// EMF -(v x B) sqrt(-g): gs05_0 subtracts its 4-cell average from twice the
// average; gs05_c / sg07 add the upwinded differences, the upwind side chosen
// by the sign of the mass flux through the neighbouring faces.  Edges over
// GetRange(interior, 0, 1): i in [is, ie+1], j in [js, je+1], k in [ks, ke+1]
// (collapsed in inactive directions).
void ct_calculate_emf(Block& b, const std::vector<Real>& P) {
    const int nd = b.ndim;
    auto F = [&](int dir, int v, int k, int j, int i) -> Real {
        return b.F[((size_t)(dir-1) * NV + v) * b.nc + b.id(k, j, i)];
    };
    auto E = [&](int e, int k, int j, int i) -> Real& { return fface(b, b.femf, e, k, j, i); };
    auto C = [&](int v, int k, int j, int i) -> Real { return b.cemf[(size_t)v * b.nc + b.id(k, j, i)]; };
    const int ile = b.ie + 1, jle = (nd > 1) ? b.je + 1 : b.je, kle = (nd > 2) ? b.ke + 1 : b.ke;
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= kle; ++k)
    for (int j = b.js; j <= jle; ++j)
    for (int i = b.is; i <= ile; ++i) {
        if (nd > 2) {
            E(0, k, j, i) = 0.25 * (-F(2, B3, k - 1, j, i) - F(2, B3, k, j, i)
                                    + F(3, B2, k, j - 1, i) + F(3, B2, k, j, i));
            E(1, k, j, i) = 0.25 * (-F(3, B1, k, j, i - 1) - F(3, B1, k, j, i)
                                    + F(1, B3, k - 1, j, i) + F(1, B3, k, j, i));
            E(2, k, j, i) = 0.25 * (-F(1, B2, k, j - 1, i) - F(1, B2, k, j, i)
                                    + F(2, B1, k, j, i - 1) + F(2, B1, k, j, i));
        } else if (nd > 1) {
            E(0, k, j, i) = -F(2, B3, k, j, i);
            E(1, k, j, i) = F(1, B3, k, j, i);
            E(2, k, j, i) = 0.25 * (-F(1, B2, k, j - 1, i) - F(1, B2, k, j, i)
                                    + F(2, B1, k, j, i - 1) + F(2, B1, k, j, i));
        } else {
            E(0, k, j, i) = 0;
            E(1, k, j, i) = F(1, B3, k, j, i);
            E(2, k, j, i) = -F(1, B2, k, j, i);
        }
    }
    if (nd < 2 || b.p.ct_scheme == 0) return;
    // cell-centre EMF over the entire block (b_ct.cpp:318-335)
#pragma omp parallel for collapse(2)
    for (int k = 0; k < b.n3; ++k)
    for (int j = 0; j < b.n2; ++j)
    for (int i = 0; i < b.n1; ++i) {
        const size_t c = b.id(k, j, i);
        Real Pc[NV]; load_prims(b, P, c, Pc);
        FourVectors D;
        calc_4vecs(b, Pc, j, i, CENTER, D);
        const Real gdet = b.gdet_(CENTER, j, i);
        b.cemf[0 * b.nc + c] = (D.bcon[2] * D.ucon[3] - D.bcon[3] * D.ucon[2]) * gdet;
        b.cemf[1 * b.nc + c] = (D.bcon[3] * D.ucon[1] - D.bcon[1] * D.ucon[3]) * gdet;
        b.cemf[2 * b.nc + c] = (D.bcon[1] * D.ucon[2] - D.bcon[2] * D.ucon[1]) * gdet;
    }
    if (b.p.ct_scheme == 1) {
        // gs05_0, b_ct.cpp:337-356.  The E1 average repeats (k, j-1) and omits
        // (k-1, j), as the reference writes it.
#pragma omp parallel for collapse(2)
        for (int k = b.ks; k <= kle; ++k)
        for (int j = b.js; j <= jle; ++j)
        for (int i = b.is; i <= ile; ++i) {
            if (nd > 2) {
                E(0, k, j, i) = 2 * E(0, k, j, i)
                    - 0.25 * (C(0, k, j, i) + C(0, k, j - 1, i) + C(0, k, j - 1, i) + C(0, k - 1, j - 1, i));
                E(1, k, j, i) = 2 * E(1, k, j, i)
                    - 0.25 * (C(1, k, j, i) + C(1, k, j, i - 1) + C(1, k - 1, j, i) + C(1, k - 1, j, i - 1));
            }
            E(2, k, j, i) = 2 * E(2, k, j, i)
                - 0.25 * (C(2, k, j, i) + C(2, k, j, i - 1) + C(2, k, j - 1, i) + C(2, k, j - 1, i - 1));
        }
        return;
    }
    // gs05_c / sg07, b_ct.cpp:357-412
    auto up = [&](Real mass_flux) { return g_plant == 4 ? (mass_flux < 0.0) : (mass_flux >= 0.0); };
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= kle; ++k)
    for (int j = b.js; j <= jle; ++j)
    for (int i = b.is; i <= ile; ++i) {
        if (nd > 2) {
            const Real e1_l3 = up(F(2, RHO, k - 1, j, i)) ? F(3, B2, k, j - 1, i) - C(0, k - 1, j - 1, i)
                                                               : F(3, B2, k, j, i) - C(0, k - 1, j, i);
            const Real e1_r3 = up(F(2, RHO, k, j, i)) ? F(3, B2, k, j - 1, i) - C(0, k, j - 1, i)
                                                           : F(3, B2, k, j, i) - C(0, k, j, i);
            const Real e1_l2 = up(F(3, RHO, k, j - 1, i)) ? -F(2, B3, k - 1, j, i) - C(0, k - 1, j - 1, i)
                                                               : -F(2, B3, k, j, i) - C(0, k, j - 1, i);
            const Real e1_r2 = up(F(3, RHO, k, j, i)) ? -F(2, B3, k - 1, j, i) - C(0, k - 1, j, i)
                                                           : -F(2, B3, k, j, i) - C(0, k, j, i);
            E(0, k, j, i) += 0.25 * (e1_l3 + e1_r3 + e1_l2 + e1_r2);
            const Real e2_l3 = up(F(1, RHO, k - 1, j, i)) ? -F(3, B1, k, j, i - 1) - C(1, k - 1, j, i - 1)
                                                               : -F(3, B1, k, j, i) - C(1, k - 1, j, i);
            const Real e2_r3 = up(F(1, RHO, k, j, i)) ? -F(3, B1, k, j, i - 1) - C(1, k, j, i - 1)
                                                           : -F(3, B1, k, j, i) - C(1, k, j, i);
            const Real e2_l1 = up(F(3, RHO, k, j, i - 1)) ? F(1, B3, k - 1, j, i) - C(1, k - 1, j, i - 1)
                                                               : F(1, B3, k, j, i) - C(1, k, j, i - 1);
            const Real e2_r1 = up(F(3, RHO, k, j, i)) ? F(1, B3, k - 1, j, i) - C(1, k - 1, j, i)
                                                           : F(1, B3, k, j, i) - C(1, k, j, i);
            E(1, k, j, i) += 0.25 * (e2_l3 + e2_r3 + e2_l1 + e2_r1);
        }
        const Real e3_l2 = up(F(1, RHO, k, j - 1, i)) ? F(2, B1, k, j, i - 1) - C(2, k, j - 1, i - 1)
                                                           : F(2, B1, k, j, i) - C(2, k, j - 1, i);
        const Real e3_r2 = up(F(1, RHO, k, j, i)) ? F(2, B1, k, j, i - 1) - C(2, k, j, i - 1)
                                                       : F(2, B1, k, j, i) - C(2, k, j, i);
        const Real e3_l1 = up(F(2, RHO, k, j, i - 1)) ? -F(1, B2, k, j - 1, i) - C(2, k, j - 1, i - 1)
                                                           : -F(1, B2, k, j, i) - C(2, k, j, i - 1);
        const Real e3_r1 = up(F(2, RHO, k, j, i)) ? -F(1, B2, k, j - 1, i) - C(2, k, j - 1, i)
                                                       : -F(1, B2, k, j, i) - C(2, k, j, i);
        E(2, k, j, i) += 0.25 * (e3_l2 + e3_r2 + e3_l1 + e3_r1);
    }
}

// B_CT::AddSource, b_ct.cpp:418-470: dB_Uf/dt on each face from the circulation
// of the edge EMFs, per unit face area.  Faces over GetRange(interior, F_d): the
// face direction's index runs to e+1 (in 2-D F3 has its one layer k = ks).
void ct_add_source(Block& b) {
    const int nd = b.ndim;
    const Real dx1 = b.dx[0], dx2 = b.dx[1], dx3 = b.dx[2];
    auto E = [&](int e, int k, int j, int i) -> Real { return fface(b, b.femf, e, k, j, i); };
    auto D = [&](int f, int k, int j, int i) -> Real& { return fface(b, b.dfB, f, k, j, i); };
    const int jle = (nd > 1) ? b.je + 1 : b.je, kle = (nd > 2) ? b.ke + 1 : b.ke;
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = b.js; j <= b.je; ++j)
    for (int i = b.is; i <= b.ie + 1; ++i) {
        D(0, k, j, i) = (-dx3 * (g_plant == 3 ? E(2, k, j, i + 1) : E(2, k, j + 1, i)) + dx3 * E(2, k, j, i));
        if (nd > 2) D(0, k, j, i) += (dx2 * E(1, k + 1, j, i) - dx2 * E(1, k, j, i));
        D(0, k, j, i) /= (dx2 * dx3);
    }
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = b.js; j <= jle; ++j)
    for (int i = b.is; i <= b.ie; ++i) {
        D(1, k, j, i) = (dx3 * E(2, k, j, i + 1) - dx3 * E(2, k, j, i));
        if (nd > 2) D(1, k, j, i) += (-dx1 * E(0, k + 1, j, i) + dx1 * E(0, k, j, i));
        D(1, k, j, i) /= (dx1 * dx3);
    }
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= kle; ++k)
    for (int j = b.js; j <= b.je; ++j)
    for (int i = b.is; i <= b.ie; ++i) {
        D(2, k, j, i) = (-dx2 * E(1, k, j, i + 1) + dx2 * E(1, k, j, i)
                         + dx1 * E(0, k, j + 1, i) - dx1 * E(0, k, j, i)) / (dx1 * dx2);
    }
}

// KHARMADriver::WeightedSumDataFace, driver/kharma_driver.hpp:167-190: z = w1 x + w2 y
// over the whole face arrays.
void ct_weighted_sum_face(Block& b, const std::vector<Real>& x, const std::vector<Real>& y, Real w1, Real w2,
                          std::vector<Real>& z) {
    const size_t n = (size_t)3 * b.nf;
#pragma omp parallel for
    for (size_t q = 0; q < n; ++q) z[q] = w1 * x[q] + w2 * y[q];
}

// Periodic exchange of the face fields of one block with itself (Parthenon's
// face-field boundary communication, not in the reference tree): along a face's
// own direction its faces s..e+1 are the block's and the others take the value
// N faces in; across, its cells wrap as a cell field's do.
void ct_sync_faces(Block& b, std::vector<Real>& A) {
    const int N[3] = {b.p.nx1, b.p.nx2, b.p.nx3};
    const int per[3] = {b.p.periodic1, b.ndim > 1 && b.p.periodic2, b.ndim > 2 && b.p.periodic3};
    const int lo[3] = {b.is, b.js, b.ks}, hi[3] = {b.ie, b.je, b.ke}, n[3] = {b.n1, b.n2, b.n3};
    for (int f = 0; f < 3; ++f) {
        // face index bounds per axis: along f one more face than cells
        int top[3];
        for (int d = 0; d < 3; ++d) top[d] = (d == f && d < b.ndim) ? n[d] : n[d] - 1;
        for (int d = 0; d < 3; ++d) {
            if (!per[d]) continue;
            const int ehi = (d == f) ? hi[d] + 1 : hi[d];
#pragma omp parallel for collapse(2)
            for (int k = 0; k <= top[2]; ++k)
            for (int j = 0; j <= top[1]; ++j)
            for (int i = 0; i <= top[0]; ++i) {
                int x[3] = {i, j, k};
                if (x[d] >= lo[d] && x[d] <= ehi) continue;
                int y[3] = {i, j, k};
                y[d] = (x[d] < lo[d]) ? x[d] + N[d] : x[d] - N[d];
                fface(b, A, f, k, j, i) = fface(b, A, f, y[2], y[1], y[0]);
            }
        }
    }
}

// B_CT::BlockUtoP, b_ct.cpp:173-205, over the entire block: cell-centred
// primitive B from the faces (each divided by its sqrt(-g)), conserved B.
void ct_block_utop(Block& b, std::vector<Real>& fB, std::vector<Real>& U, std::vector<Real>& P) {
    const int nd = b.ndim;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < b.n3; ++k)
    for (int j = 0; j < b.n2; ++j)
    for (int i = 0; i < b.n1; ++i) {
        const size_t c = b.id(k, j, i);
        const Real p1 = (fface(b, fB, 0, k, j, i) / b.gdet_(FACE1, j, i)
                         + fface(b, fB, 0, k, j, i + 1) / b.gdet_(FACE1, j, i + 1)) / 2;
        const Real p2 = (nd > 1) ? (fface(b, fB, 1, k, j, i) / b.gdet_(FACE2, j, i)
                                    + fface(b, fB, 1, k, j + 1, i) / b.gdet_(FACE2, j + 1, i)) / 2
                                 : fface(b, fB, 1, k, j, i) / b.gdet_(FACE2, j, i);
        const Real p3 = (nd > 2) ? (fface(b, fB, 2, k, j, i) / b.gdet_(FACE3, j, i)
                                    + fface(b, fB, 2, k + 1, j, i) / b.gdet_(FACE3, j, i)) / 2
                                 : fface(b, fB, 2, k, j, i) / b.gdet_(FACE3, j, i);
        P[(size_t)B1 * b.nc + c] = p1;
        P[(size_t)B2 * b.nc + c] = p2;
        P[(size_t)B3 * b.nc + c] = p3;
        for (int v = 0; v < 3; ++v)
            U[(size_t)(B1 + v) * b.nc + c] = P[(size_t)(B1 + v) * b.nc + c] * b.gdet_(CENTER, j, i);
    }
}

// B_CT::MaxDivB, b_ct.cpp:472-497, with face_div (b_ct_functions.hpp:47-56) on
// the interior cells.
double ct_max_divb(const Block& b, const std::vector<Real>& fB) {
    const int nd = b.ndim;
    Block& bb = const_cast<Block&>(b);
    std::vector<Real>& A = const_cast<std::vector<Real>&>(fB);
    double m = 0.;
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = b.js; j <= b.je; ++j)
    for (int i = b.is; i <= b.ie; ++i) {
        Real du = (fface(bb, A, 0, k, j, i + 1) - fface(bb, A, 0, k, j, i)) / b.dx[0];
        if (nd > 1) du += (fface(bb, A, 1, k, j + 1, i) - fface(bb, A, 1, k, j, i)) / b.dx[1];
        if (nd > 2) du += (fface(bb, A, 2, k + 1, j, i) - fface(bb, A, 2, k, j, i)) / b.dx[2];
        m = std::max(m, std::fabs(val(du)));
    }
    return m;
}

// KHARMADriver::FluxDivergence, driver/kharma_driver.hpp:208-231 ->
// parthenon::Update::FluxDivHelper (not in the reference tree; restated from
// UniformCartesian face areas / cell volume -- summation order parity unpinned).
void flux_divergence(Block& b) {
    const Real dx1 = b.dx[0], dx2 = b.dx[1], dx3 = b.dx[2];
    const Real A1 = dx2 * dx3, A2 = dx1 * dx3, A3 = dx1 * dx2, V = dx1 * dx2 * dx3;
    const int nd = b.ndim;
    const long s2 = b.n1, s3 = (long)b.n1 * b.n2;
#pragma omp parallel for collapse(3)
    for (int v = 0; v < NV; ++v)
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = b.js; j <= b.je; ++j)
    for (int i = b.is; i <= b.ie; ++i) {
        const size_t c = b.id(k, j, i);
        const Real* f1 = &b.F[((size_t)0 * NV + v) * b.nc];
        const Real* f2 = &b.F[((size_t)1 * NV + v) * b.nc];
        const Real* f3 = &b.F[((size_t)2 * NV + v) * b.nc];
        Real du = (A1 * f1[c + 1] - A1 * f1[c]);
        if (nd >= 2) du += (A2 * f2[c + s2] - A2 * f2[c]);
        if (nd == 3) du += (A3 * f3[c + s3] - A3 * f3[c]);
        b.dUdt[(size_t)v * b.nc + c] = -du / V;
    }
}


// Flux::AddGeoSource, flux/flux.cpp:395-444.  Skipped in Cartesian Minkowski (:407).
void add_geo_source(Block& b, const std::vector<Real>& P) {
    if (b.cart_minkowski) return;
    const Real gam = b.p.gam;
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = b.js; j <= b.je; ++j)
    for (int i = b.is; i <= b.ie; ++i) {
        const size_t c = b.id(k, j, i);
        Real Pc[NV]; load_prims(b, P, c, Pc);
        FourVectors D;
        calc_4vecs(b, Pc, j, i, CENTER, D);
        Real Tmu[4] = {0}, new_du[4] = {0};
        for (int mu = 0; mu < 4; ++mu) {
            calc_tensor_mhd(Pc[RHO], Pc[UU], (gam - 1) * Pc[UU], D, mu, Tmu);
            for (int nu = 0; nu < 4; ++nu)
                for (int lam = 0; lam < 4; ++lam)
                    new_du[lam] += Tmu[nu] * (g_plant == 1 ? b.gdet_conn_(j, i, lam, nu, mu)
                                                           : b.gdet_conn_(j, i, nu, lam, mu));
        }
        b.dUdt[(size_t)UU * b.nc + c] += new_du[0];
        for (int v = 0; v < 3; ++v) b.dUdt[(size_t)(U1 + v) * b.nc + c] += new_du[1 + v];
    }
}

// KHARMADriver::AddStateUpdate, driver/kharma_driver.cpp:339-387, via
// parthenon::Update::WeightedSumData (entire domain; not in the reference tree).
void state_update(Block& b, const std::vector<Real>& Usub, const std::vector<Real>& Ufull,
                  const std::vector<Real>& Psub, std::vector<Real>& Uout, std::vector<Real>& Pout,
                  Real gam0, Real gam1, Real beta_dt) {
    const size_t n = (size_t)NV * b.nc;
#pragma omp parallel for
    for (size_t c = 0; c < n; ++c) Uout[c] = gam0 * Usub[c] + gam1 * Ufull[c];
#pragma omp parallel for
    for (size_t c = 0; c < n; ++c) Uout[c] = 1.0 * Uout[c] + beta_dt * b.dUdt[c];
    // Copy of the HD primitives (rho, u, uvec) sub -> out as the U->P guess (:379-384)
    if (&Psub != &Pout) {
#pragma omp parallel for
        for (size_t c = 0; c < (size_t)5 * b.nc; ++c) Pout[c] = Psub[c];
    }
}

// Periodic boundary exchange of one block with itself (the Parthenon ghost
// exchange, AddBoundaryExchangeTasks at kharma_driver.cpp:151, is not in the
// reference tree).  Synced variables: those flagged FillGhost = cons (8) and
// the 5 GRMHD primitives (grmhd.cpp:155-165, kharma_step.cpp:129-138).
void sync_periodic_var(Block& b, std::vector<Real>& A, int v) {
    Real* a = &A[(size_t)v * b.nc];
    const int N1 = b.p.nx1, N2 = b.p.nx2, N3 = b.p.nx3;
    auto wrap = [](int x, int s, int N) { int r = (x - s) % N; if (r < 0) r += N; return s + r; };
    // x1 then x2 then x3 sweep over the full extents fills edges and corners
    if (b.p.periodic1)
    for (int k = 0; k < b.n3; ++k) for (int j = 0; j < b.n2; ++j) for (int i = 0; i < b.n1; ++i) {
        if (i >= b.is && i <= b.ie) continue;
        a[b.id(k, j, i)] = a[b.id(k, j, wrap(i, b.is, N1))];
    }
    if (b.ndim > 1 && b.p.periodic2)
        for (int k = 0; k < b.n3; ++k) for (int j = 0; j < b.n2; ++j) {
            if (j >= b.js && j <= b.je) continue;
            for (int i = 0; i < b.n1; ++i) a[b.id(k, j, i)] = a[b.id(k, wrap(j, b.js, N2), i)];
        }
    if (b.ndim > 2 && b.p.periodic3)
        for (int k = 0; k < b.n3; ++k) {
            if (k >= b.ks && k <= b.ke) continue;
            for (int j = 0; j < b.n2; ++j) for (int i = 0; i < b.n1; ++i)
                a[b.id(k, j, i)] = a[b.id(wrap(k, b.ks, N3), j, i)];
        }
}
void sync_bounds(Block& b, std::vector<Real>& U, std::vector<Real>& P) {
#pragma omp parallel for
    for (int v = 0; v < NV + 5; ++v) {
        if (v < NV) sync_periodic_var(b, U, v);
        else sync_periodic_var(b, P, v - NV);
    }
}

// The exchange of driver "simple" (sync_prims, kharma_driver.cpp:109-121): the
// eight primitives, the conserved B and pflag (FillGhost when the fix-up
// averages neighbours, inverter.cpp:97-104), then B_FluxCT::BlockUtoP over each
// periodic boundary's ghost slab (kharma_driver.cpp:155-180).  Fluid conserved
// variables are not exchanged.
void sync_bounds_prims(Block& b, std::vector<Real>& U, std::vector<Real>& P) {
#pragma omp parallel for
    for (int v = 0; v < NV + 3; ++v) {
        if (v < NV) sync_periodic_var(b, P, v);
        else sync_periodic_var(b, U, B1 + v - NV);
    }
    if (b.p.fix_average) sync_periodic_var(b, b.pflag, 0);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < b.n3; ++k)
    for (int j = 0; j < b.n2; ++j)
    for (int i = 0; i < b.n1; ++i) {
        const bool g1 = b.p.periodic1 && (i < b.is || i > b.ie);
        const bool g2 = b.ndim > 1 && b.p.periodic2 && (j < b.js || j > b.je);
        const bool g3 = b.ndim > 2 && b.p.periodic3 && (k < b.ks || k > b.ke);
        if (!(g1 || g2 || g3)) continue;
        const size_t c = b.id(k, j, i);
        for (int v = 0; v < 3; ++v) P[(B1 + v) * b.nc + c] = U[(B1 + v) * b.nc + c] / b.gdet_(CENTER, j, i);
    }
}

// Inverter::u_to_p<onedw>, inverter/onedw.hpp:49-233
Real lorentz_calc_w(Real Bsq, Real D, Real QdB, Real Qtsq, Real Wp) {
    const Real QdBsq = QdB * QdB;
    const Real W = Wp + D;
    const Real W2 = W * W;
    const Real WB = W + Bsq;
    const Real utsq = -((W + WB) * QdBsq + W2 * Qtsq) / (QdBsq * (W + WB) + W2 * (Qtsq - WB * WB));
    if (utsq < -1.e-15 || utsq > 1.e7) return -1.;
    return std::sqrt(1. + std::fabs(utsq));
}
Real err_eqn(Real gam, Real Bsq, Real D, Real Ep, Real QdB, Real Qtsq, Real Wp, int& eflag) {
    const Real W = Wp + D;
    const Real gamma = lorentz_calc_w(Bsq, D, QdB, Qtsq, Wp);
    if (gamma < 1) eflag = S_BAD_UT;
    const Real w = W / (gamma*gamma);
    const Real rho = D / gamma;
    const Real p = (w - rho) * (gam - 1) / gam;
    return -Ep + Wp - p + 0.5 * Bsq + 0.5 * (Bsq * Qtsq - QdB * QdB) / SQR(Bsq + W);
}
int u_to_p_onedw(const Block& b, const Real* U, Real* P, int j, int i, int loc, int max_iterations, Real tol) {
    const Real gam = b.p.gam;
    if (U[RHO] <= 0.) return S_NEG_INPUT;
    const Real alpha = 1./std::sqrt(-b.gcon_(loc, j, i, 0, 0));
    const Real gdet = b.gdet_(loc, j, i);
    const Real a_over_g = alpha / gdet;
    const Real D = U[RHO] * a_over_g;
    Real Bcon[4] = {0, 0, 0, 0};
    Bcon[1] = U[B1] * a_over_g;
    Bcon[2] = U[B2] * a_over_g;
    Bcon[3] = U[B3] * a_over_g;
    const Real Qcov[4] = {(U[UU] - U[RHO]) * a_over_g, U[U1] * a_over_g, U[U2] * a_over_g, U[U3] * a_over_g};
    const Real ncov[4] = {-alpha, 0., 0., 0.};
    Real Bcov[4], Qcon[4], ncon[4];
    b.lower(Bcon, Bcov, j, i, loc);
    b.raise(Qcov, Qcon, j, i, loc);
    b.raise(ncov, ncon, j, i, loc);
    const Real Bsq = dot(Bcon, Bcov);
    const Real QdB = dot(Bcon, Qcov);
    const Real Qdotn = dot(Qcon, ncov);
    Real Qtcon[4];
    for (int mu = 0; mu < 4; ++mu) Qtcon[mu] = Qcon[mu] + ncon[mu] * Qdotn;
    const Real Qtsq = dot(Qcon, Qcov) + Qdotn*Qdotn;
    const Real Ep = -Qdotn - D;
    int eflag = S_SUCCESS;
    Real Wp, err;
    {
        const Real gamma = lorentz_calc(b, P, j, i, loc);
        if (gamma < 1) return S_BAD_UT;
        const Real rho = P[RHO], u = P[UU];
        Wp = (rho + u + (gam - 1) * u) * gamma * gamma - rho * gamma;
        err = err_eqn(gam, Bsq, D, Ep, QdB, Qtsq, Wp, eflag);
    }
    Real dW;
    {
        const Real Wpm = (1. - 1.e-5) * Wp;
        const Real h = Wp - Wpm;
        const Real Wpp = Wp + h;
        const Real errm = err_eqn(gam, Bsq, D, Ep, QdB, Qtsq, Wpm, eflag);
        const Real errp = err_eqn(gam, Bsq, D, Ep, QdB, Qtsq, Wpp, eflag);
        const Real dedW = (errp - errm) / (Wpp - Wpm);
        const Real dedW2 = (errp - 2. * err + errm) / (h*h);
        const Real f = clip(0.5 * err * dedW2 / (dedW*dedW), -0.3, 0.3);
        dW = clip(-err / dedW / (1. - f), -0.5*Wp, 2.0*Wp);
    }
    Real Wp1 = Wp;
    Real err1 = err;
    Wp += dW;
    err = err_eqn(gam, Bsq, D, Ep, QdB, Qtsq, Wp, eflag);
    int iter = 0;
    for (iter = 0; iter < max_iterations; iter++) {
        dW = clip((Wp1 - Wp) * err / (err - err1), -0.5*Wp, 2.0*Wp);
        Wp1 = Wp;
        err1 = err;
        Wp += dW;
        if (std::fabs(dW / Wp) < tol) break;
        err = err_eqn(gam, Bsq, D, Ep, QdB, Qtsq, Wp, eflag);
        if (std::fabs(err / Wp) < tol) break;
    }
    if (iter == max_iterations) return S_MAX_ITER;
    const Real gamma = lorentz_calc_w(Bsq, D, QdB, Qtsq, Wp);
    if (gamma < 1) return S_BAD_UT;
    const Real rho = D / gamma;
    const Real W = Wp + D;
    const Real w = W / (gamma*gamma);
    const Real p = (w - rho) * (gam - 1) / gam;
    const Real u = w - (rho + p);
    if (rho < 0 && u < 0) return S_NEG_RHOU;
    else if (rho < 0) return S_NEG_RHO;
    else if (u < 0) return S_NEG_U;
    P[RHO] = rho;
    P[UU] = u;
    const Real pre = (gamma / (W + Bsq));
    P[U1] = pre * (Qtcon[1] + QdB * Bcon[1] / W);
    P[U2] = pre * (Qtcon[2] + QdB * Bcon[2] / W);
    P[U3] = pre * (Qtcon[3] + QdB * Bcon[3] / W);
    return S_SUCCESS;
}

// KDomain::GetPhysicalRange (domain.hpp:161-187): the interior along
// non-periodic (physical-boundary) dimensions, the entire block along periodic ones.
struct Range3 { int is, ie, js, je, ks, ke; };
Range3 physical_range(const Block& b) {
    Range3 r{0, b.n1 - 1, 0, b.n2 - 1, 0, b.n3 - 1};
    if (!b.p.periodic1) { r.is = b.is; r.ie = b.ie; }
    if (b.ndim > 1 && !b.p.periodic2) { r.js = b.js; r.je = b.je; }
    if (b.ndim > 2 && !b.p.periodic3) { r.ks = b.ks; r.ke = b.ke; }
    return r;
}

// Inverter::u_to_p<kastaun>, inverter/kastaun.hpp:83-400 (Kastaun, Kalinani &
// Ciolfi 2021).  The residual of the master function (their eq. 44) with the
// inverter's floors (KastaunResidual, kastaun.hpp:83-206), the bracketing
// auxiliary function (eq. 49), and the two false-position solves with the
// Illinois halving (kastaun.hpp:290-364); returns the INVERTER_* floor bits, or
// S_MAX_ITER when the master solve does not converge.
struct KastaunResidual {
    Real D, q, bsq, bsq_rpsq, rsq, rbsq, v0sq, gam, rho_floor, e_floor, e_max;
    bool rho_fl = false, e_fl = false, e_mx = false, g_mx = false;
    Real x_mu(Real mu) const { return 1.0 / (1.0 + mu * bsq); }
    Real rbarsq_mu(Real mu, Real x) const { return x * (x * rsq + mu * (1.0 + x) * rbsq); }
    Real qbar_mu(Real mu, Real x) const { const Real mux = mu * x; return q - 0.5 * (bsq + mux * mux * bsq_rpsq); }
    Real vhatsq_mu(Real mu, Real rbarsq) {
        const Real v = mu * mu * rbarsq;
        g_mx = v > v0sq;
        return g_mx ? v0sq : v;
    }
    Real rhohat_mu(Real iW) {
        const Real r = D * iW;
        rho_fl = r <= rho_floor;
        return rho_fl ? rho_floor : r;
    }
    Real ehat_mu(Real mu, Real qbar, Real rbarsq, Real vhatsq, Real W) {
        const Real e = W * (qbar - mu * rbarsq) + vhatsq * W * W / (1.0 + W);
        e_fl = false; e_mx = false;
        if (e <= e_floor) { e_fl = true; return e_floor; }
        if (e > e_max) { e_mx = true; return e_max; }
        return e;
    }
    Real master(Real mu) {
        const Real x = x_mu(mu);
        const Real rbarsq = rbarsq_mu(mu, x);
        const Real qbar = qbar_mu(mu, x);
        const Real vhatsq = vhatsq_mu(mu, rbarsq);
        const Real iW = std::sqrt(1.0 - vhatsq);
        const Real W = 1.0 / iW;
        const Real rhohat = rhohat_mu(iW);
        const Real ehat = ehat_mu(mu, qbar, rbarsq, vhatsq, W);
        const Real Phat = ehat * rhohat * (gam - 1.0);
        const Real hhat = rhohat * (1.0 + ehat) + Phat;
        const Real ahat = Phat / (hhat - Phat);
        const Real nua = (1.0 + ahat) * (1.0 + ehat) * iW;
        const Real nub = (1.0 + ahat) * (1.0 + qbar - mu * rbarsq);
        const Real nuhat = std::max(nua, nub);
        return mu - 1.0 / (nuhat + mu * rbarsq);
    }
    Real aux(Real mu) const {
        const Real x = 1.0 / (1.0 + mu * bsq);
        const Real rbarsq = x * (rsq * x + mu * (1.0 + x) * rbsq);
        return mu * std::sqrt(1.0 + rbarsq) - 1.0;
    }
};
template <class F>
Real false_position_illinois(F&& f, Real zm, Real zp, int max_iterations, Real tol, int& iter) {
    Real fm = f(zm), fp = f(zp);
    int iterations = max_iterations;
    if (std::fabs(zm - zp) < tol || (std::fabs(fm) + std::fabs(fp)) < 2.0 * tol) iterations = -1;
    Real z = 0.5 * (zm + zp);
    for (iter = 0; iter < iterations; ++iter) {
        z = (zm * fp - zp * fm) / (fp - fm);
        const Real fz = f(z);
        if (std::fabs(zm - zp) < tol || std::fabs(fz) < tol) break;
        if (fz * fp < 0.0) { zm = zp; fm = fp; zp = z; fp = fz; }
        else { fm = 0.5 * fm; zp = z; fp = fz; }
    }
    if (iterations < 0) iter = 0;
    return z;
}
int u_to_p_kastaun(const Block& b, const Real* U, Real* P, int j, int i, int loc, int max_iterations, Real tol) {
    const auto& f = b.p;
    const Real alpha = 1. / std::sqrt(-b.gcon_(loc, j, i, 0, 0));
    const Real a_over_g = alpha / b.gdet_(loc, j, i);
    const Real D = U[RHO] * a_over_g;
    const Real D_fl = std::max(D, Real(f.rho_min_const));
    const Real Qcov[4] = {(U[UU] - U[RHO]) * a_over_g, U[U1] * a_over_g, U[U2] * a_over_g, U[U3] * a_over_g};
    const Real ncov[4] = {-alpha, 0., 0., 0.};
    Real ncon[4];
    b.raise(ncov, ncon, j, i, loc);
    const Real q = (-dot(Qcov, ncon) - D) / D;
    const Real rcov[3] = {U[U1] / U[RHO], U[U2] / U[RHO], U[U3] / U[RHO]};
    Real rcon[3];
    // spatial inverse metric gamma^ij = g^ij - g^0i g^0j / g^00 (kastaun.hpp:241-257)
    const Real g00 = b.gcon_(loc, j, i, 0, 0);
    for (int a = 0; a < 3; ++a) {
        rcon[a] = 0.;
        for (int c = 0; c < 3; ++c)
            rcon[a] += (b.gcon_(loc, j, i, a + 1, c + 1) - b.gcon_(loc, j, i, 0, a + 1) * b.gcon_(loc, j, i, 0, c + 1) / g00) * rcov[c];
    }
    const Real rsq = rcon[0] * rcov[0] + rcon[1] * rcov[1] + rcon[2] * rcov[2];
    const Real sD = 1.0 / std::sqrt(D_fl);
    Real bu[3], bdotr = 0.;
    for (int a = 0; a < 3; ++a) { bu[a] = (U[B1 + a] * a_over_g) * sD; bdotr += bu[a] * rcov[a]; }
    Real bsq = 0.;
    for (int a = 0; a < 3; ++a) for (int c = 0; c < 3; ++c) bsq += b.gcov_(loc, j, i, a + 1, c + 1) * bu[a] * bu[c];
    bsq = std::max(Real(0.), bsq);
    KastaunResidual res;
    res.D = D; res.q = q; res.bsq = bsq; res.rsq = rsq; res.rbsq = bdotr * bdotr; res.bsq_rpsq = bsq * rsq - res.rbsq;
    res.v0sq = std::min(rsq / (1.0 + rsq), Real(1.0 - 1.0 / SQR(f.gamma_max)));
    res.gam = f.gam; res.rho_floor = f.rho_min_const; res.e_floor = f.u_min_const / D_fl; res.e_max = f.u_over_rho_max;
    int iter;
    const Real zb = false_position_illinois([&](Real mu) { return res.aux(mu); }, Real(0.), Real(1.), max_iterations, tol, iter);
    const Real mu = false_position_illinois([&](Real z) { return res.master(z); }, Real(0.), zb, max_iterations, tol, iter);
    if (iter == max_iterations) return S_MAX_ITER;
    const Real x = res.x_mu(mu);
    const Real rbarsq = res.rbarsq_mu(mu, x);
    const Real vsq = res.vhatsq_mu(mu, rbarsq);
    const Real iW = std::sqrt(1.0 - vsq);
    const Real W = 1. / iW;
    P[RHO] = res.rhohat_mu(iW);
    const Real qbar = res.qbar_mu(mu, x);
    P[UU] = res.ehat_mu(mu, qbar, rbarsq, vsq, W) * P[RHO];
    for (int a = 0; a < 3; ++a) P[U1 + a] = W * mu * x * (rcon[a] + mu * bdotr * bu[a]);
    int fl = 0;
    if (res.rho_fl) fl |= F_INVERTER_RHO;
    if (res.e_fl) fl |= F_INVERTER_U;
    if (res.g_mx) fl |= F_INVERTER_GAMMA;
    if (res.e_mx) fl |= F_INVERTER_U_MAX;
    return fl;
}

// Packages::MeshUtoP -> Inverter::BlockUtoP (inverter.cpp:131-200, over the
// physical range whatever the domain) then B_FluxCT::BlockUtoP
// (b_flux_ct.cpp:195-217, over the domain: entire, or the interior when
// b_interior is set).
void utop(Block& b, std::vector<Real>& U, std::vector<Real>& P, bool b_interior = false) {
    const Range3 pr = physical_range(b);
#pragma omp parallel for collapse(2)
    for (int k = pr.ks; k <= pr.ke; ++k)
    for (int j = pr.js; j <= pr.je; ++j)
    for (int i = pr.is; i <= pr.ie; ++i) {
        const size_t c = b.id(k, j, i);
        Real Uc[NV], Pc[NV];
        load_prims(b, U, c, Uc); load_prims(b, P, c, Pc);
        int pflagl = (b.p.inverter == 1) ? u_to_p_kastaun(b, Uc, Pc, j, i, CENTER, b.p.iter_max, b.p.err_tol)
                                         : u_to_p_onedw(b, Uc, Pc, j, i, CENTER, b.p.iter_max, b.p.err_tol);
        for (int v = 0; v < 5; ++v) P[v * b.nc + c] = Pc[v];
        b.pflag[c] = pflagl % FFLAG_MINIMUM;
        b.fflag[c] = (pflagl / FFLAG_MINIMUM) * FFLAG_MINIMUM;
    }
    // B_FluxCT::BlockUtoP (b_flux_ct.cpp:172-217); with B_CT the primitive B came
    // from the faces before the inversion (Packages::BlockUtoP runs B_CT first,
    // kharma_package.cpp:58-90)
    if (b.p.face_ct) return;
    const Range3 br = b_interior ? Range3{b.is, b.ie, b.js, b.je, b.ks, b.ke}
                                 : Range3{0, b.n1 - 1, 0, b.n2 - 1, 0, b.n3 - 1};
#pragma omp parallel for collapse(2)
    for (int v = 0; v < 3; ++v)
    for (int k = br.ks; k <= br.ke; ++k)
    for (int j = br.js; j <= br.je; ++j)
    for (int i = br.is; i <= br.ie; ++i) {
        const size_t c = b.id(k, j, i);
        P[(B1 + v) * b.nc + c] = U[(B1 + v) * b.nc + c] / b.gdet_(CENTER, j, i);
    }
}

// floors/floors_functions.hpp:84-160 (Cartesian or spherical geometric floors)
int determine_floors(const Block& b, const Real* P, int j, int i, Real& rhoflr_max, Real& uflr_max) {
    const auto& f = b.p;
    const Real gam = f.gam;
    Real rhoflr_geom, uflr_geom;
    geom_floors(b, j, i, rhoflr_geom, uflr_geom);
    FourVectors D;
    calc_4vecs(b, P, j, i, CENTER, D);
    Real bsq = dot(D.bcon, D.bcov);
    Real rhoflr_b = bsq / f.bsq_over_rho_max;
    Real uflr_b = bsq / f.bsq_over_u_max;
    uflr_max = std::max(uflr_geom, uflr_b);
    const Real rho = P[RHO];
    const Real u = P[UU];
    int fflag = 0;
    if (!f.temp_adjust_u) {
        const Real rhoflr_temp = std::max(u, uflr_max) / f.u_over_rho_max;
        fflag |= (rhoflr_temp > rho) * F_TEMP;
        rhoflr_max = std::max(std::max(rhoflr_geom, rhoflr_b), rhoflr_temp);
    } else {
        rhoflr_max = std::max(rhoflr_geom, rhoflr_b);
    }
    if (rhoflr_max > rho || uflr_max > u) {
        fflag |= (rhoflr_geom > rho) * F_GEOM_RHO;
        fflag |= (uflr_geom > u) * F_GEOM_U;
        fflag |= (rhoflr_b > rho) * F_B_RHO;
        fflag |= (uflr_b > u) * F_B_U;
    }
    if (lorentz_calc(b, P, j, i, CENTER) > f.gamma_max) fflag |= F_GAMMA;
    if ((gam - 1.) * P[UU] / std::pow(P[RHO], gam) > f.ktot_max) fflag |= F_KTOT;
    if (f.temp_adjust_u && (P[UU] / P[RHO] > f.u_over_rho_max)) fflag |= F_TEMP;
    return fflag;
}
// floors_functions.hpp:186-257 (drift frame)
int apply_floors_drift(const Block& b, Real* P, int j, int i, Real rhoflr_max, Real uflr_max) {
    const Real gam = b.p.gam;
    const Real lapse2 = 1. / (-b.gcon_(CENTER, j, i, 0, 0));
    Real beta[4] = {0., 0., 0., 0.};
    beta[1] = lapse2 * b.gcon_(CENTER, j, i, 0, 1);
    beta[2] = lapse2 * b.gcon_(CENTER, j, i, 0, 2);
    beta[3] = lapse2 * b.gcon_(CENTER, j, i, 0, 3);
    const Real rho = P[RHO];
    const Real uu = P[UU];
    const Real pg = (gam - 1.) * uu;
    const Real w_old = std::max(rho + uu + pg, SMALL);
    Real Bcon[4] = {0}, Bcov[4] = {0};
    Bcon[0] = 0; Bcon[1] = P[B1]; Bcon[2] = P[B2]; Bcon[3] = P[B3];
    for (int mu = 0; mu < 4; ++mu) for (int nu = 0; nu < 4; ++nu) Bcov[mu] += b.gcov_(CENTER, j, i, mu, nu) * Bcon[nu];
    const Real Bsq = std::max(dot(Bcon, Bcov), SMALL);
    const Real B_mag = std::sqrt(Bsq);
    FourVectors D;
    calc_4vecs(b, P, j, i, CENTER, D);
    Real Qcov[4] = {0};
    Qcov[0] = w_old * D.ucon[0] * D.ucov[0] + pg;
    Qcov[1] = w_old * D.ucon[0] * D.ucov[1];
    Qcov[2] = w_old * D.ucon[0] * D.ucov[2];
    Qcov[3] = w_old * D.ucon[0] * D.ucov[3];
    Real QdotB = dot(Bcon, Qcov);
    Real vpar = QdotB / (B_mag * w_old * D.ucon[0]*D.ucon[0]);
    Real ucon_dr[4] = {0};
    ucon_dr[0] = 1. / std::sqrt(1. / (D.ucon[0]*D.ucon[0]) + vpar*vpar);
    for (int mu = 0; mu < 4; ++mu) ucon_dr[mu] = D.ucon[mu] * (ucon_dr[0] / D.ucon[0]) - (vpar * Bcon[mu] * ucon_dr[0] / B_mag);
    P[RHO] = std::max(rho, rhoflr_max);
    P[UU] = std::max(uu, uflr_max);
    const Real pg_new = (gam - 1.) * P[UU];
    const Real w_new = P[RHO] + P[UU] + pg_new;
    // plant 2: the pre-floor enthalpy in the parallel-momentum solve
    const Real x = (2. * QdotB) / (B_mag * (g_plant == 2 ? w_old : w_new) * ucon_dr[0]);
    vpar = x / (1 + std::sqrt(1 + x*x)) * (1. / ucon_dr[0]);
    D.ucon[0] = 1. / std::sqrt(1/(ucon_dr[0]*ucon_dr[0]) - vpar*vpar);
    for (int mu = 0; mu < 4; ++mu) D.ucon[mu] = ucon_dr[mu] * (D.ucon[0] / ucon_dr[0]) + (vpar * Bcon[mu] * D.ucon[0] / B_mag);
    b.lower(D.ucon, D.ucov, j, i, CENTER);
    P[U1] = D.ucon[1] + (beta[1] * D.ucon[0]);
    P[U2] = D.ucon[2] + (beta[2] * D.ucon[0]);
    P[U3] = D.ucon[3] + (beta[3] * D.ucon[0]);
    return 0;
}
// floors_functions.hpp:178-183 (fluid frame)
int apply_floors_fluid(Real* P, Real rhoflr_max, Real uflr_max) {
    P[RHO] += std::max(0., rhoflr_max - P[RHO]);
    P[UU]  += std::max(0., uflr_max - P[UU]);
    return 0;
}
// floors_functions.hpp:50-82
void apply_ceilings(const Block& b, Real* P, int j, int i) {
    const auto& f = b.p;
    const Real gam = f.gam;
    Real gamma = lorentz_calc(b, P, j, i, CENTER);
    Real ktot = (gam - 1.) * P[UU] / std::pow(P[RHO], gam);
    if (gamma > f.gamma_max) {
        Real fr = std::sqrt((SQR(f.gamma_max) - 1.) / (SQR(gamma) - 1.));
        for (int v = 0; v < 3; ++v) P[U1 + v] *= fr;
    }
    if (ktot > f.ktot_max) P[UU] = f.ktot_max / ktot * P[UU];
    if (f.temp_adjust_u && P[UU] / P[RHO] > f.u_over_rho_max) P[UU] = f.u_over_rho_max * P[RHO];
}
// normal observer frame, floors_functions.hpp:260-289 (onedw recovery, 8 iters, 1e-8)
int apply_floors_normal(const Block& b, Real* P, Real* U, int j, int i, Real rhoflr_max, Real uflr_max) {
    const Real gam = b.p.gam;
    const Real rho_add = std::max(0., rhoflr_max - P[RHO]);
    const Real u_add = std::max(0., uflr_max - P[UU]);
    Real Padd[NV] = {rho_add, u_add, 0, 0, 0, 0, 0, 0};
    // GRMHD::p_to_u_mhd, grmhd_functions.hpp:356-375
    FourVectors D;
    calc_4vecs(b, Padd, j, i, CENTER, D);
    const Real gdet = b.gdet_(CENTER, j, i);
    Real rho_ut = rho_add * D.ucon[0] * gdet;
    Real mhd[4];
    calc_tensor_mhd(rho_add, u_add, (gam - 1) * u_add, D, 0, mhd);
    Real T[4];
    T[0] = mhd[0] * gdet + rho_ut;
    for (int v = 0; v < 3; ++v) T[1 + v] = mhd[1 + v] * gdet;
    P[RHO] += rho_add;
    P[UU] += u_add;
    U[RHO] += rho_ut;
    U[UU] += T[0];
    U[U1] += T[1];
    U[U2] += T[2];
    U[U3] += T[3];
    return u_to_p_onedw(b, U, P, j, i, CENTER, 8, 1e-8);
}
// Flux::p_to_u_mhd (flux_functions.hpp:321-328): calc_4vecs at centre + 5 conserved
void p_to_u_mhd(const Block& b, const Real* P, Real* U, int j, int i) {
    FourVectors D;
    calc_4vecs(b, P, j, i, CENTER, D);
    prim_to_flux(b, P, D, j, i, 0, U, CENTER, false);
}
// Floors::ApplyGRMHDFloors -> ApplyFloorsInFrame, floors/floors_impl.hpp:45-128 and
// DetermineGRMHDFloors, floors.cpp:197-231, over the entire domain (or the
// interior, when `interior` is set).
void apply_floors(Block& b, std::vector<Real>& U, std::vector<Real>& P, bool interior = false) {
    if (!b.p.floors_on) return;
    const Range3 r = interior ? Range3{b.is, b.ie, b.js, b.je, b.ks, b.ke}
                              : Range3{0, b.n1 - 1, 0, b.n2 - 1, 0, b.n3 - 1};
#pragma omp parallel for collapse(2)
    for (int k = r.ks; k <= r.ke; ++k)
    for (int j = r.js; j <= r.je; ++j)
    for (int i = r.is; i <= r.ie; ++i) {
        const size_t c = b.id(k, j, i);
        Real Pc[NV]; load_prims(b, P, c, Pc);
        Real rf, uf;
        int ff = determine_floors(b, Pc, j, i, rf, uf);
        b.fflag[c] = (Real)((int)b.fflag[c] | ff);
        b.rho_floor[c] = rf; b.u_floor[c] = uf;
    }
#pragma omp parallel for collapse(2)
    for (int k = r.ks; k <= r.ke; ++k)
    for (int j = r.js; j <= r.je; ++j)
    for (int i = r.is; i <= r.ie; ++i) {
        const size_t c = b.id(k, j, i);
        if (!(int)b.fflag[c]) continue;
        Real Pc[NV], Uc[NV];
        load_prims(b, P, c, Pc); load_prims(b, U, c, Uc);
        int pflag_l = 0;
        if (b.p.floor_frame == 0) pflag_l = apply_floors_fluid(Pc, b.rho_floor[c], b.u_floor[c]);
        else if (b.p.floor_frame == 1) pflag_l = apply_floors_normal(b, Pc, Uc, j, i, b.rho_floor[c], b.u_floor[c]);
        else pflag_l = apply_floors_drift(b, Pc, j, i, b.rho_floor[c], b.u_floor[c]);
        if (pflag_l) b.pflag[c] = pflag_l;
        apply_ceilings(b, Pc, j, i);
        p_to_u_mhd(b, Pc, Uc, j, i);
        for (int v = 0; v < 5; ++v) { P[v * b.nc + c] = Pc[v]; U[v * b.nc + c] = Uc[v]; }
    }
}

// Inverter::FixUtoP, inverter/fixup.cpp:47-156 over the physical range
void fix_utop(Block& b, std::vector<Real>& U, std::vector<Real>& P) {
    if (!b.p.fix_average && !b.p.fix_atmo) return;
    const Range3 pr = physical_range(b);
    // First pass reads only non-failed zones and writes only failed ones, so it is
    // order independent.
#pragma omp parallel for collapse(2)
    for (int k = pr.ks; k <= pr.ke; ++k)
    for (int j = pr.js; j <= pr.je; ++j)
    for (int i = pr.is; i <= pr.ie; ++i) {
        const size_t c = b.id(k, j, i);
        if (!(b.pflag[c] > 0)) continue;
        Real wsum = 0., sum[5] = {0., 0., 0., 0., 0.};
        if (b.p.fix_average) {
            for (int n = -1; n <= 1; n++) for (int m = -1; m <= 1; m++) for (int l = -1; l <= 1; l++) {
                int ii = i + l, jj = j + m, kk = k + n;
                if (ii < pr.is || ii > pr.ie || jj < pr.js || jj > pr.je || kk < pr.ks || kk > pr.ke) continue;
                const size_t cc = b.id(kk, jj, ii);
                if (!(b.pflag[cc] > 0)) {
                    double w = 1./(std::abs(l) + std::abs(m) + std::abs(n) + 1);
                    wsum += w;
                    for (int p = 0; p < 5; ++p) sum[p] += w * P[p * b.nc + cc];
                }
            }
        }
        if (wsum < 1.e-10) { for (int p = 0; p < 5; ++p) P[p * b.nc + c] = 0.; }
        else { for (int p = 0; p < 5; ++p) P[p * b.nc + c] = sum[p]/wsum; }
    }
#pragma omp parallel for collapse(2)
    for (int k = pr.ks; k <= pr.ke; ++k)
    for (int j = pr.js; j <= pr.je; ++j)
    for (int i = pr.is; i <= pr.ie; ++i) {
        const size_t c = b.id(k, j, i);
        if (!(b.pflag[c] > 0)) continue;
        Real Pc[NV], Uc[NV];
        load_prims(b, P, c, Pc); load_prims(b, U, c, Uc);
        // Floors::apply_geo_floors (Global), floors_functions.hpp:334-367
        Real rf, uf;
        geom_floors(b, j, i, rf, uf);
        Pc[RHO] += std::max(0., rf - Pc[RHO]);
        Pc[UU]  += std::max(0., uf - Pc[UU]);
        // GRMHD::p_to_u, grmhd_functions.hpp:270-298 (5 conserved)
        FourVectors D;
        calc_4vecs(b, Pc, j, i, CENTER, D);
        prim_to_flux(b, Pc, D, j, i, 0, Uc, CENTER, false);
        for (int v = 0; v < 5; ++v) { P[v * b.nc + c] = Pc[v]; U[v * b.nc + c] = Uc[v]; }
    }
}

// Flux::MeshPtoU(entire) -> Flux::p_to_u, flux/flux.cpp:276-322, flux_functions.hpp:343-352
void ptou(Block& b, const std::vector<Real>& P, std::vector<Real>& U) {
#pragma omp parallel for collapse(2)
    for (int k = 0; k < b.n3; ++k)
    for (int j = 0; j < b.n2; ++j)
    for (int i = 0; i < b.n1; ++i) {
        const size_t c = b.id(k, j, i);
        Real Pc[NV], Uc[NV];
        load_prims(b, P, c, Pc);
        FourVectors D;
        calc_4vecs(b, Pc, j, i, CENTER, D);
        prim_to_flux(b, Pc, D, j, i, 0, Uc, CENTER, true);
        for (int v = 0; v < NV; ++v) U[v * b.nc + c] = Uc[v];
    }
}

// GRMHD::EstimateTimestep in-loop branch, grmhd/grmhd.cpp:310-339
double estimate_timestep(const Block& b) {
    double min_ndt = std::numeric_limits<double>::max();
    const Real dx1 = b.dx[0], dx2 = b.dx[1], dx3 = b.dx[2];
#pragma omp parallel for collapse(2) reduction(min:min_ndt)
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = b.js; j <= b.je; ++j)
    for (int i = b.is; i <= b.ie; ++i) {
        const size_t c = b.id(k, j, i);
        double ndt_zone = val(1.0 / (1 / (dx1 / std::max(b.cmax[c], b.cmin[c])) +
                                     1 / (dx2 / std::max(b.cmax[b.nc + c], b.cmin[b.nc + c])) +
                                     1 / (dx3 / std::max(b.cmax[2 * b.nc + c], b.cmin[2 * b.nc + c]))));
        if (!std::isnan(ndt_zone) && ndt_zone < min_ndt) min_ndt = ndt_zone;
    }
    const double dt_max = b.p.max_dt_increase * b.dt_last;
    return val(clip(min_ndt * b.p.cfl, b.p.dt_min, dt_max));
}

// B_FluxCT::MaxDivB with corner_div, b_flux_ct.cpp:566-607, b_flux_ct_functions.hpp:52-75
double max_divb(const Block& b, const std::vector<Real>& U) {
    if (b.ndim < 2) return 0.;
    const bool do3 = b.ndim > 2;
    const int jsb = b.js, jeb = b.je + 1, ksb = b.ks, keb = b.ke + (do3 ? 1 : 0);
    // ValidDivBX1 (b_flux_ct.cpp:554-564): skip the corners on physical X1 faces
    const int isb = b.is + (b.p.periodic1 ? 0 : 1), ieb = b.ie + (b.p.periodic1 ? 1 : 0);
    const Real* B1v = &U[(size_t)B1 * b.nc];
    const Real* B2v = &U[(size_t)B2 * b.nc];
    const Real* B3v = &U[(size_t)B3 * b.nc];
    double mx = 0.;
#pragma omp parallel for collapse(2) reduction(max:mx)
    for (int k = ksb; k <= keb; ++k)
    for (int j = jsb; j <= jeb; ++j)
    for (int i = isb; i <= ieb; ++i) {
        auto id = [&](int kk, int jj, int ii) { return b.id(kk, jj, ii); };
        const double norm = do3 ? 0.25 : 0.5;
        Real term1 = B1v[id(k, j, i)] - B1v[id(k, j, i-1)] + B1v[id(k, j-1, i)] - B1v[id(k, j-1, i-1)];
        Real term2 = B2v[id(k, j, i)] - B2v[id(k, j-1, i)] + B2v[id(k, j, i-1)] - B2v[id(k, j-1, i-1)];
        Real term3 = 0.;
        if (do3) {
            term1 += B1v[id(k-1, j, i)] + B1v[id(k-1, j-1, i)] - B1v[id(k-1, j, i-1)] - B1v[id(k-1, j-1, i-1)];
            term2 += B2v[id(k-1, j, i)] + B2v[id(k-1, j, i-1)] - B2v[id(k-1, j-1, i)] - B2v[id(k-1, j-1, i-1)];
            term3 = B3v[id(k, j, i)] + B3v[id(k, j-1, i)] + B3v[id(k, j, i-1)] + B3v[id(k, j-1, i-1)]
                  - B3v[id(k-1, j, i)] - B3v[id(k-1, j-1, i)] - B3v[id(k-1, j, i-1)] - B3v[id(k-1, j-1, i-1)];
        }
        double d = val(std::fabs(norm*term1/b.dx[0] + norm*term2/b.dx[1] + norm*term3/b.dx[2]));
        if (d > mx) mx = d;
    }
    return mx;
}

// ------------------------------------------------------------ boundaries
// KBoundaries::check_inflow, boundaries.hpp:101-136: remove velocity into the domain.
// The clip() of vsq there discards its result.
void check_inflow(const Block& b, Real* Pc, int j, int i, bool inner) {
    Real uvec[3] = {Pc[U1], Pc[U2], Pc[U3]};
    const Real gamma0 = lorentz_calc(b, Pc, j, i, CENTER);
    const Real alpha0 = 1. / std::sqrt(-b.gcon_(CENTER, j, i, 0, 0));
    const Real ucon1 = uvec[0] - gamma0 * alpha0 * b.gcon_(CENTER, j, i, 0, 1);
    if ((ucon1 > 0. && inner) || (ucon1 < 0. && !inner)) {
        for (int v = 0; v < 3; ++v) uvec[v] /= gamma0;
        const Real alpha = 1. / std::sqrt(-b.gcon_(CENTER, j, i, 0, 0));
        const Real beta1 = b.gcon_(CENTER, j, i, 0, 1) * alpha * alpha;
        uvec[0] = beta1 / alpha;
        const Real vsq = b.gcov_(CENTER, j, i, 1, 1) * uvec[0] * uvec[0] +
                         b.gcov_(CENTER, j, i, 2, 2) * uvec[1] * uvec[1] +
                         b.gcov_(CENTER, j, i, 3, 3) * uvec[2] * uvec[2] +
                         2. * (b.gcov_(CENTER, j, i, 1, 2) * uvec[0] * uvec[1] +
                               b.gcov_(CENTER, j, i, 1, 3) * uvec[0] * uvec[2] +
                               b.gcov_(CENTER, j, i, 2, 3) * uvec[1] * uvec[2]);
        const Real gamma = 1. / std::sqrt(1. - vsq);
        for (int v = 0; v < 3; ++v) Pc[U1 + v] = uvec[v] * gamma;
    }
}

// One physical face, KBoundaries::ApplyBoundary (boundaries.cpp:369-589):
// Parthenon's outflow (copy of the first/last interior zone) or reflecting
// (mirror copy, x2 vector components negated) BoundaryFunction over the whole
// ghost slab -- restated, the functions are not in the reference tree -- on the
// synced variables (8 conserved + rho, u, uvec), check_inflow on X1 faces, then
// BoundaryPtoUElseUtoP (kharma_package.cpp:125-156): Flux::BlockPtoUMHD and
// B_FluxCT::BlockUtoP over the slab.
enum { INNER_X1 = 0, OUTER_X1 = 1, INNER_X2 = 2, OUTER_X2 = 3 };
//
// With `prims` set (driver "simple": the primitives, the conserved B and pflag
// are the FillGhost fields, kharma_driver.cpp:109-121, inverter.cpp:97-104) the
// copy is of those fields, and the slab then gets B_FluxCT::BlockUtoP and
// Flux::BlockPtoU of all eight variables (boundaries.cpp:561-571).
void apply_boundary(Block& b, std::vector<Real>& U, std::vector<Real>& P, int face, bool prims = false) {
    const bool x1 = (face == INNER_X1 || face == OUTER_X1);
    const bool inner = (face == INNER_X1 || face == INNER_X2);
    const int type = (face == INNER_X1) ? b.p.bc_inner_x1 : (face == OUTER_X1) ? b.p.bc_outer_x1 :
                     (face == INNER_X2) ? b.p.bc_inner_x2 : b.p.bc_outer_x2;
    const int i0 = x1 ? (inner ? 0 : b.ie + 1) : 0, i1 = x1 ? (inner ? b.is - 1 : b.n1 - 1) : b.n1 - 1;
    const int j0 = x1 ? 0 : (inner ? 0 : b.je + 1), j1 = x1 ? b.n2 - 1 : (inner ? b.js - 1 : b.n2 - 1);
    const bool inflow = x1 && (inner ? b.p.check_inflow_inner_x1 : b.p.check_inflow_outer_x1);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < b.n3; ++k)
    for (int j = j0; j <= j1; ++j)
    for (int i = i0; i <= i1; ++i) {
        int si = i, sj = j;
        Real sign2 = 1.;
        if (x1) {
            si = (type == 1) ? (inner ? b.is : b.ie) : (inner ? 2 * b.is - 1 - i : 2 * b.ie + 1 - i);
        } else {
            sj = (type == 1) ? (inner ? b.js : b.je) : (inner ? 2 * b.js - 1 - j : 2 * b.je + 1 - j);
        }
        if (type == 2) sign2 = -1.;
        const size_t c = b.id(k, j, i), sc = b.id(k, sj, si);
        auto flip = [&](int v) {
            return (type == 2 && ((x1 && (v == U1 || v == B1)) || (!x1 && (v == U2 || v == B2)))) ? sign2 : 1.;
        };
        for (int v = prims ? B1 : 0; v < NV; ++v) U[v * b.nc + c] = flip(v) * U[v * b.nc + sc];
        for (int v = 0; v < (prims ? NV : 5); ++v) P[v * b.nc + c] = flip(v) * P[v * b.nc + sc];
        if (prims) b.pflag[c] = b.pflag[sc];
    }
    if (inflow) {
#pragma omp parallel for collapse(2)
        for (int k = 0; k < b.n3; ++k)
        for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
            const size_t c = b.id(k, j, i);
            Real Pc[NV]; load_prims(b, P, c, Pc);
            check_inflow(b, Pc, j, i, inner);
            for (int v = U1; v <= U3; ++v) P[v * b.nc + c] = Pc[v];
        }
    }
#pragma omp parallel for collapse(2)
    for (int k = 0; k < b.n3; ++k)
    for (int j = j0; j <= j1; ++j)
    for (int i = i0; i <= i1; ++i) {
        const size_t c = b.id(k, j, i);
        Real Pc[NV], Uc[NV];
        if (prims) {
            for (int v = 0; v < 3; ++v) P[(B1 + v) * b.nc + c] = U[(B1 + v) * b.nc + c] / b.gdet_(CENTER, j, i);
            load_prims(b, P, c, Pc);
            FourVectors D;
            calc_4vecs(b, Pc, j, i, CENTER, D);
            prim_to_flux(b, Pc, D, j, i, 0, Uc, CENTER, true);
            for (int v = 0; v < NV; ++v) U[v * b.nc + c] = Uc[v];
            continue;
        }
        load_prims(b, P, c, Pc); load_prims(b, U, c, Uc);
        p_to_u_mhd(b, Pc, Uc, j, i);
        for (int v = 0; v < 5; ++v) U[v * b.nc + c] = Uc[v];
        for (int v = 0; v < 3; ++v) P[(B1 + v) * b.nc + c] = U[(B1 + v) * b.nc + c] / b.gdet_(CENTER, j, i);
    }
}
// ApplyBoundaryConditions over the faces in Parthenon's order, plus KHARMA's
// inner-X1 corner fix after the X2 faces (boundaries.cpp:556-570).
void apply_physical_bcs(Block& b, std::vector<Real>& U, std::vector<Real>& P, bool prims = false) {
    if (!b.p.periodic1) { apply_boundary(b, U, P, INNER_X1, prims); apply_boundary(b, U, P, OUTER_X1, prims); }
    if (b.ndim > 1 && !b.p.periodic2) {
        apply_boundary(b, U, P, INNER_X2, prims); apply_boundary(b, U, P, OUTER_X2, prims);
        if (b.p.fix_corner_inner && !b.p.periodic1) apply_boundary(b, U, P, INNER_X1, prims);
    }
}

// B_FluxCT::ZeroBoundaryFlux on both X2 faces (b_flux_ct.cpp:372-420): no B
// flux through the pole, and the mirrored B2 fluxes that make the polar EMFs vanish.
void zero_polar_b_flux(Block& b) {
    if (b.ndim < 2 || b.p.periodic2 || !b.p.fix_polar_flux) return;
    auto F = [&](int dir, int v, int k, int j, int i) -> Real& {
        return b.F[((size_t)(dir-1) * NV + v) * b.nc + b.id(k, j, i)];
    };
    const int ks = b.ks - (b.ndim > 2), ke = b.ke + (b.ndim > 2);
    for (int k = ks; k <= ke; ++k)
    for (int i = b.is - 1; i <= b.ie + 1; ++i) {
        int j = b.js;
        F(2, B1, k, j, i) = 0.; F(2, B3, k, j, i) = 0.;
        F(1, B2, k, j-1, i) = -F(1, B2, k, j, i);
        if (b.ndim > 2) F(3, B2, k, j-1, i) = -F(3, B2, k, j, i);
        j = b.je + 1;
        F(2, B1, k, j, i) = 0.; F(2, B3, k, j, i) = 0.;
        F(1, B2, k, j, i) = -F(1, B2, k, j-1, i);
        if (b.ndim > 2) F(3, B2, k, j, i) = -F(3, B2, k, j-1, i);
    }
}
// KBoundaries::FixFlux (boundaries.cpp:611-690): no mass inflow through
// physical X1 faces, and zero fluxes through the X2 (polar) faces.
void kboundaries_fix_flux(Block& b) {
    auto F = [&](int dir, int v, int k, int j, int i) -> Real& {
        return b.F[((size_t)(dir-1) * NV + v) * b.nc + b.id(k, j, i)];
    };
    const int is = b.is - 1, ie = b.ie + 1;
    const int js = b.js - (b.ndim > 1), je = b.je + (b.ndim > 1);
    const int ks = b.ks - (b.ndim > 2), ke = b.ke + (b.ndim > 2);
    if (!b.p.periodic1) {
        for (int side = 0; side < 2; ++side) {
            const bool on = side == 0 ? b.p.check_inflow_inner_x1 : b.p.check_inflow_outer_x1;
            if (!on) continue;
            const int i = side == 0 ? b.is : b.ie + 1;
            for (int k = ks; k <= ke; ++k) for (int j = js; j <= je; ++j) {
                Real& f = F(1, RHO, k, j, i);
                f = side == 0 ? std::min(f, 0.) : std::max(f, 0.);
            }
        }
    }
    if (b.ndim > 1 && !b.p.periodic2 && b.p.zero_polar_flux) {
        for (int side = 0; side < 2; ++side) {
            const int j = side == 0 ? b.js : b.je + 1;
            for (int v = 0; v < NV; ++v) for (int k = ks; k <= ke; ++k) for (int i = is; i <= ie; ++i)
                F(2, v, k, j, i) = 0.;
        }
    }
}

// Integrator weights for parthenon's low-storage "rk2" / "vl2" (not in the
// reference tree; values restated, parity unpinned).  (gam0, gam1, beta)
void integrator_weights(int integrator, int stage, Real& gam0, Real& gam1, Real& beta) {
    if (integrator == 1) {  // vl2
        gam0 = 0.; gam1 = 1.; beta = (stage == 1) ? 0.5 : 1.0;
    } else {                // rk2
        gam0 = (stage == 1) ? 0. : 0.5; gam1 = (stage == 1) ? 1. : 0.5; beta = (stage == 1) ? 1. : 0.5;
    }
}

// First-order flux correction (KHARMADriver::AddFOFC, driver/kharma_driver.cpp:292-337;
// Flux::MarkFOFC and Flux::FOFC, flux/fofc.cpp:45-226) with its defaults (the
// floors' own prescription, LLF speeds from both states, no source term in the
// trial).  A trial update of the stage from the raw fluxes (FluxDivergence and
// AddStateUpdate into the guess, B included), its U->P with the stage's
// primitives as guess and DetermineGRMHDFloors; zones whose floor or inverter
// flags are set, zones within eh_buffer of the horizon and the first
// polar_cells rows at physical X2 faces are marked; the marks are synced (ghost
// zones take those of their periodic images or, at physical faces, of the
// nearest interior zone, as the outflow/reflecting fill of a scalar gives); every
// face next to a marked zone gets the LLF flux of the two zones' own primitives.
// Relative distance of the primitives P from the nearest threshold that
// determine_floors compares them with (test instrumentation: a zone within
// rounding of a threshold takes its floor flag from the last bits of P).
Real floor_margin(const Block& b, const Real* P, int j, int i) {
    const auto& f = b.p;
    Real rg, ug;
    geom_floors(b, j, i, rg, ug);
    FourVectors D;
    calc_4vecs(b, P, j, i, CENTER, D);
    const Real bsq = dot(D.bcon, D.bcov);
    auto rel = [](Real x, Real y) {
        const Real s = std::max(std::fabs(val(x)), std::fabs(val(y)));
        return s > 0 ? Real(std::fabs(val(x) - val(y)) / s) : Real(1.);
    };
    Real m = std::min(rel(rg, P[RHO]), rel(ug, P[UU]));
    m = std::min(m, std::min(rel(bsq / f.bsq_over_rho_max, P[RHO]), rel(bsq / f.bsq_over_u_max, P[UU])));
    m = std::min(m, rel(std::max(P[UU], std::max(ug, bsq / f.bsq_over_u_max)) / f.u_over_rho_max, P[RHO]));
    m = std::min(m, rel(lorentz_calc(b, P, j, i, CENTER), f.gamma_max));
    return m;
}

void fofc(Block& b, int stage, const std::vector<Real>& Usub, const std::vector<Real>& Psub) {
    Real gam0, gam1, beta;
    integrator_weights(b.p.integrator, stage, gam0, gam1, beta);
    flux_divergence(b);
    std::vector<int> mark(b.nc, 0);
    b.fofc_mark.assign(b.nc, 0.);
    b.fofc_trial.assign((size_t)NV * b.nc, 0.);
    b.fofc_margin.assign(b.nc, 1.);
    const Real r_eh = 1. + std::sqrt(1. - b.p.a * b.p.a);
#pragma omp parallel for collapse(2)
    for (int k = b.ks; k <= b.ke; ++k)
    for (int j = b.js; j <= b.je; ++j)
    for (int i = b.is; i <= b.ie; ++i) {
        const size_t c = b.id(k, j, i);
        Real Ug[NV], Pg[NV];
        for (int v = 0; v < NV; ++v) {
            const Real u = gam0 * Usub[v * b.nc + c] + gam1 * b.U0[v * b.nc + c];
            Ug[v] = 1.0 * u + (beta * b.dt) * b.dUdt[v * b.nc + c];
            Pg[v] = Psub[v * b.nc + c];
        }
        const int ret = (b.p.inverter == 1) ? u_to_p_kastaun(b, Ug, Pg, j, i, CENTER, b.p.iter_max, b.p.err_tol)
                                            : u_to_p_onedw(b, Ug, Pg, j, i, CENTER, b.p.iter_max, b.p.err_tol);
        for (int v = 0; v < 3; ++v) Pg[B1 + v] = Ug[B1 + v] / b.gdet_(CENTER, j, i);
        int ff = (ret / FFLAG_MINIMUM) * FFLAG_MINIMUM;
        if (b.p.floors_on) {
            Real rf, uf;
            ff |= determine_floors(b, Pg, j, i, rf, uf);
        }
        bool m = ff != 0;
        if (b.p.spherical && b.rcen[(size_t)j * b.gn1 + i] < r_eh + b.p.fofc_eh_buffer) m = true;
        if (b.p.spherical && !b.p.periodic2 && b.p.fofc_polar_cells > 0 &&
            (j < b.js + b.p.fofc_polar_cells || j > b.je - b.p.fofc_polar_cells)) m = true;
        mark[c] = m;
        b.fofc_mark[c] = m;
        for (int v = 0; v < NV; ++v) b.fofc_trial[v * b.nc + c] = Pg[v];
        if (b.p.floors_on) b.fofc_margin[c] = floor_margin(b, Pg, j, i);
    }
    auto src = [&](int v, int s0, int e0, int n, bool periodic) {
        if (v >= s0 && v <= e0) return v;
        if (periodic) { int r = (v - s0) % n; if (r < 0) r += n; return s0 + r; }
        return v < s0 ? s0 : e0;
    };
    auto marked = [&](int k, int j, int i) {
        const int ii = src(i, b.is, b.ie, b.p.nx1, b.p.periodic1);
        const int jj = b.ndim > 1 ? src(j, b.js, b.je, b.p.nx2, b.p.periodic2) : j;
        const int kk = b.ndim > 2 ? src(k, b.ks, b.ke, b.p.nx3, b.p.periodic3) : k;
        return mark[b.id(kk, jj, ii)] != 0;
    };
    for (int dir = 1; dir <= b.ndim; ++dir) {
        const int loc = dir - 1;
        // the faces the step consumes (X1: i in [is, ie+1], transverse cells +-1)
        int is = b.is, ie = b.ie, js = b.js, je = b.je, ks = b.ks, ke = b.ke;
        if (dir == 1) ie += 1; else { is -= 1; ie += 1; }
        if (b.ndim > 1) { if (dir == 2) je += 1; else { js -= 1; je += 1; } }
        if (b.ndim > 2) { if (dir == 3) ke += 1; else { ks -= 1; ke += 1; } }
        const long st = (dir == 1) ? 1 : (dir == 2) ? (long)b.n1 : (long)b.n1 * b.n2;
#pragma omp parallel for collapse(2)
        for (int k = ks; k <= ke; ++k)
        for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie; ++i) {
            const int km = dir == 3 ? k - 1 : k, jm = dir == 2 ? j - 1 : j, im = dir == 1 ? i - 1 : i;
            if (!marked(k, j, i) && !marked(km, jm, im)) continue;
            const size_t c = b.id(k, j, i);
            Real Pl[NV], Pr[NV], Ul[NV], Ur[NV], Fl[NV], Fr[NV];
            for (int v = 0; v < NV; ++v) { Pl[v] = Psub[v * b.nc + c - st]; Pr[v] = Psub[v * b.nc + c]; }
            FourVectors D;
            calc_4vecs(b, Pl, j, i, loc, D);
            prim_to_flux(b, Pl, D, j, i, 0, Ul, loc);
            prim_to_flux(b, Pl, D, j, i, dir, Fl, loc);
            Real cmaxL, cminL;
            vchar(b, Pl, D, j, i, loc, dir, cmaxL, cminL);
            calc_4vecs(b, Pr, j, i, loc, D);
            prim_to_flux(b, Pr, D, j, i, 0, Ur, loc);
            prim_to_flux(b, Pr, D, j, i, dir, Fr, loc);
            Real cmaxR, cminR;
            vchar(b, Pr, D, j, i, loc, dir, cmaxR, cminR);
            const Real cmx = std::max(std::max(Real(0.), cmaxL), cmaxR);
            const Real cmn = -std::min(std::min(Real(0.), cminL), cminR);
            b.cmax[(dir - 1) * b.nc + c] = cmx;
            b.cmin[(dir - 1) * b.nc + c] = cmn;
            for (int v = 0; v < NV; ++v)
                b.F[((size_t)(dir - 1) * NV + v) * b.nc + c] = llf(Fl[v], Fr[v], cmx, cmn, Ul[v], Ur[v]);
        }
    }
}

// One stage of KHARMADriver::MakeDefaultTaskCollection, driver/kharma_step.cpp:77-319.
// Stage containers follow parthenon's MultiStageDriver naming: stage 1 is
// base -> "1", stage 2 is "1" -> base.
void run_stage(Block& b, int stage) {
    std::vector<Real>& Usub = (stage == 1) ? b.U0 : b.U1c;
    std::vector<Real>& Psub = (stage == 1) ? b.P0 : b.P1c;
    std::vector<Real>& Uout = (stage == 1) ? b.U1c : b.U0;
    std::vector<Real>& Pout = (stage == 1) ? b.P1c : b.P0;
    // Flux region (kharma_step.cpp:141-210)
    for (int dir = 1; dir <= 3; ++dir) get_flux(b, Psub, dir);
    // AddFOFC on the raw fluxes (kharma_step.cpp:166-172), before the flux fixes
    if (b.p.fofc) fofc(b, stage, Usub, Psub);
    Real gam0, gam1, beta;
    integrator_weights(b.p.integrator, stage, gam0, gam1, beta);
    std::vector<Real>& fBsub = (stage == 1) ? b.fB0 : b.fB1c;
    std::vector<Real>& fBout = (stage == 1) ? b.fB1c : b.fB0;
    if (b.p.face_ct) {
        // Packages::FixFlux (KBoundaries only), then B_CT::CalculateEMF
        // (kharma_step.cpp:175-190); FluxDivergence does not update the derived
        // cell-centred B, B_CT::AddSource sets dB_Uf/dt (kharma_step.cpp:196-202)
        kboundaries_fix_flux(b);
        ct_calculate_emf(b, Psub);
        flux_divergence(b);
        add_geo_source(b, Psub);
        ct_add_source(b);
    } else {
        // Packages::FixFlux: B_FluxCT (ZeroBoundaryFlux, FluxCT), then KBoundaries
        zero_polar_b_flux(b);
        flux_ct(b);
        kboundaries_fix_flux(b);
        flux_divergence(b);
        add_geo_source(b, Psub);
    }
    state_update(b, Usub, b.U0, Psub, Uout, Pout, gam0, gam1, beta * b.dt);
    if (b.p.face_ct) {
        // AddStateUpdate's face half (kharma_driver.cpp:353-373)
        ct_weighted_sum_face(b, fBsub, b.fB0, gam0, gam1, fBout);
        ct_weighted_sum_face(b, fBout, b.dfB, 1.0, beta * b.dt, fBout);
    }
    // boundary exchange, which applies the physical boundaries too (kharma_driver.cpp:148-150)
    sync_bounds(b, Uout, Pout);
    if (b.p.face_ct) ct_sync_faces(b, fBout);
    apply_physical_bcs(b, Uout, Pout);
    // Fix region (kharma_step.cpp:216-288); Packages::MeshUtoP runs B_CT's
    // BlockUtoP over the entire block before the inversion
    if (b.p.face_ct) ct_block_utop(b, fBout, Uout, Pout);
    utop(b, Uout, Pout);
    apply_floors(b, Uout, Pout);
    fix_utop(b, Uout, Pout);
    apply_physical_bcs(b, Uout, Pout);
    ptou(b, Pout, Uout);
    // Second sync (two_sync, kharma_step.cpp:307-314)
    sync_bounds(b, Uout, Pout);
    apply_physical_bcs(b, Uout, Pout);
}

// One stage of KHARMADriver::MakeSimpleTaskCollection, driver/simple_step.cpp:39-168
// (driver/type = simple): inversion and floors on the interior before the
// exchange (:122-128), the fix-up after it (:141), then the boundary fill and
// P->U over the entire block; the exchanged fields are the primitives.
void run_stage_simple(Block& b, int stage) {
    std::vector<Real>& Usub = (stage == 1) ? b.U0 : b.U1c;
    std::vector<Real>& Psub = (stage == 1) ? b.P0 : b.P1c;
    std::vector<Real>& Uout = (stage == 1) ? b.U1c : b.U0;
    std::vector<Real>& Pout = (stage == 1) ? b.P1c : b.P0;
    for (int dir = 1; dir <= 3; ++dir) get_flux(b, Psub, dir);          // :86
    zero_polar_b_flux(b);                                                // :93 Packages::FixFlux
    flux_ct(b);
    kboundaries_fix_flux(b);
    flux_divergence(b);                                                  // :96
    add_geo_source(b, Psub);                                             // :100
    Real gam0, gam1, beta;
    integrator_weights(b.p.integrator, stage, gam0, gam1, beta);
    state_update(b, Usub, b.U0, Psub, Uout, Pout, gam0, gam1, beta * b.dt);  // :104-119
    utop(b, Uout, Pout, true);                                           // :123 (interior)
    apply_floors(b, Uout, Pout, true);                                   // :126 (interior)
    sync_bounds_prims(b, Uout, Pout);                                    // :129
    apply_physical_bcs(b, Uout, Pout, true);
    fix_utop(b, Uout, Pout);                                             // :141
    apply_physical_bcs(b, Uout, Pout, true);                             // :143
    ptou(b, Pout, Uout);                                                 // :146 (entire)
    sync_bounds_prims(b, Uout, Pout);                                    // :160-165 two_sync
    apply_physical_bcs(b, Uout, Pout, true);
}

// EvolutionDriver step + KHARMA dt logic: PreStepWork (kharma.cpp:121-129),
// stages, EstimateTimestep (grmhd.cpp:252-343), SetGlobalTimeStep
// (kharma_driver.cpp:439-460).  Returns the dt that was taken.
double step(Block& b) {
    b.in_loop = true;
    b.dt_last = b.dt;
    for (int stage = 1; stage <= 2; ++stage) {
        if (b.p.driver == 1) run_stage_simple(b, stage);
        else run_stage(b, stage);
    }
    double new_dt = estimate_timestep(b);
    const double taken = b.dt;
    b.time += b.dt;
    b.ncycle++;
    double dt = b.dt;
    if (dt < 0.1 * std::numeric_limits<Real>::max()) dt *= 2.0;
    dt = std::min(dt, new_dt);
    if (b.time < b.p.tlim && (b.p.tlim - b.time) < dt) dt = b.p.tlim - b.time;
    b.dt = dt;
    return taken;
}

}  // namespace

// ================================================================== C ABI
namespace {
// Values between the ABI's double arrays and the block's Real arrays.
void load_values(std::vector<Real>& dst, const double* src, size_t n) {
    for (size_t q = 0; q < n; ++q) dst[q] = Real(src[q]);
}
void store_values(const std::vector<Real>& src, double* dst, size_t n) {
    for (size_t q = 0; q < n; ++q) dst[q] = val(src[q]);
}
}  // namespace

extern "C" {

void* or_create(const or_params* p) {
    Block* b = new Block();
    b->p = *p;
    b->ndim = (p->nx3 > 1) ? 3 : ((p->nx2 > 1) ? 2 : 1);
    const int ng = p->ng;
    b->n1 = p->nx1 + 2 * ng;
    b->n2 = (b->ndim > 1) ? p->nx2 + 2 * ng : 1;
    b->n3 = (b->ndim > 2) ? p->nx3 + 2 * ng : 1;
    b->is = ng; b->ie = ng + p->nx1 - 1;
    b->js = (b->ndim > 1) ? ng : 0; b->je = (b->ndim > 1) ? ng + p->nx2 - 1 : 0;
    b->ks = (b->ndim > 2) ? ng : 0; b->ke = (b->ndim > 2) ? ng + p->nx3 - 1 : 0;
    b->nc = (size_t)b->n1 * b->n2 * b->n3;
    b->dx[0] = (p->x1max - p->x1min) / p->nx1;
    b->dx[1] = (p->x2max - p->x2min) / p->nx2;
    b->dx[2] = (p->x3max - p->x3min) / p->nx3;
    if (p->spherical) init_geometry_ks(*b);
    else init_geometry_minkowski(*b);
    const size_t nvc = (size_t)NV * b->nc;
    b->U0.assign(nvc, 0.); b->P0.assign(nvc, 0.); b->U1c.assign(nvc, 0.); b->P1c.assign(nvc, 0.);
    b->F.assign(3 * nvc, 0.);
    b->cmax.assign(3 * b->nc, 0.); b->cmin.assign(3 * b->nc, 0.);
    b->Pl.assign(nvc, 0.); b->Pr.assign(nvc, 0.); b->Ul.assign(nvc, 0.); b->Ur.assign(nvc, 0.);
    b->Fl.assign(nvc, 0.); b->Fr.assign(nvc, 0.);
    b->emf.assign(3 * b->nc, 0.); b->dUdt.assign(nvc, 0.);
    b->pflag.assign(b->nc, 0.); b->fflag.assign(b->nc, 0.);
    b->rho_floor.assign(b->nc, 0.); b->u_floor.assign(b->nc, 0.);
    b->fn1 = b->n1 + 1; b->fn2 = b->n2 + 1; b->fn3 = b->n3 + 1;
    b->nf = (size_t)b->fn1 * b->fn2 * b->fn3;
    if (p->face_ct) {
        b->fB0.assign(3 * b->nf, 0.); b->fB1c.assign(3 * b->nf, 0.);
        b->femf.assign(3 * b->nf, 0.); b->dfB.assign(3 * b->nf, 0.);
        b->cemf.assign(3 * b->nc, 0.);
    }
    return b;
}

void or_destroy(void* h) { delete static_cast<Block*>(h); }

size_t or_ncells(void* h) { return static_cast<Block*>(h)->nc; }

// Set the initial primitives (all 8, entire block) and do what
// KHARMA::PostInitialize does for the fields on this path
// (post_initialize.cpp:55-193): sync, P->U over the entire domain, sync.
// dt is dt_start for the first step (grmhd.cpp:283-289).
void or_init_state(void* h, const double* P) {
    Block& b = *static_cast<Block*>(h);
    load_values(b.P0, P, (size_t)NV * b.nc);
    // KHARMA::PostInitialize (prob/post_initialize.cpp:55-193): P->U over the
    // entire block, then the boundary sync with physical boundaries.  The input
    // may hold the interior only; its periodic ghosts (all 8 primitives, B
    // included, as a field seeded over the entire block would have) come first.
    for (int v = 0; v < NV; ++v) sync_periodic_var(b, b.P0, v);
    ptou(b, b.P0, b.U0);
    sync_bounds(b, b.U0, b.P0);
    apply_physical_bcs(b, b.U0, b.P0);
    b.time = 0.; b.ncycle = 0; b.in_loop = false;
    b.dt = b.p.dt_start;
    b.dt_last = b.dt;
}

// Initial state with face-centred B (B_CT): the primitives P (B ignored) and the
// face fields fB [3][n3+1][n2+1][n1+1] as seeded (prob/seed_B.cpp:137-191,
// 336-353): face sync, B_CT::BlockUtoP over the entire block, then what
// or_init_state does.
void or_init_state_faces(void* h, const double* P, const double* fB) {
    Block& b = *static_cast<Block*>(h);
    load_values(b.P0, P, (size_t)NV * b.nc);
    load_values(b.fB0, fB, 3 * b.nf);
    ct_sync_faces(b, b.fB0);
    for (int v = 0; v < NV; ++v) sync_periodic_var(b, b.P0, v);
    ct_block_utop(b, b.fB0, b.U0, b.P0);
    ptou(b, b.P0, b.U0);
    sync_bounds(b, b.U0, b.P0);
    apply_physical_bcs(b, b.U0, b.P0);
    b.time = 0.; b.ncycle = 0; b.in_loop = false;
    b.dt = b.p.dt_start;
    b.dt_last = b.dt;
}
size_t or_nfaces(void* h) { return static_cast<Block*>(h)->nf; }
// B_CT task entry points on the block's arrays: CalculateEMF from the stored
// fluxes and P0; AddSource into dfB; BlockUtoP of fB0 into U0, P0; the face sync
// of fB0; MaxDivB of fB0; WeightedSumDataFace on caller arrays [3][nf].
void or_ct_calculate_emf(void* h) { Block& b = *static_cast<Block*>(h); ct_calculate_emf(b, b.P0); }
void or_ct_add_source(void* h) { ct_add_source(*static_cast<Block*>(h)); }
void or_ct_block_utop(void* h) { Block& b = *static_cast<Block*>(h); ct_block_utop(b, b.fB0, b.U0, b.P0); }
void or_ct_sync_faces(void* h) { Block& b = *static_cast<Block*>(h); ct_sync_faces(b, b.fB0); }
double or_ct_max_divb(void* h) { Block& b = *static_cast<Block*>(h); return ct_max_divb(b, b.fB0); }
void or_ct_weighted_sum_face(void* h, const double* x, const double* y, double w1, double w2, double* z) {
    Block& b = *static_cast<Block*>(h);
    const size_t n = 3 * b.nf;
    std::vector<Real> X(n), Y(n), Z(n);
    load_values(X, x, n); load_values(Y, y, n);
    ct_weighted_sum_face(b, X, Y, w1, w2, Z);
    store_values(Z, z, n);
}

// Overwrite base-container U and P directly (no P->U), for task-level tests.
void or_set_state(void* h, const double* P, const double* U) {
    Block& b = *static_cast<Block*>(h);
    load_values(b.P0, P, (size_t)NV * b.nc);
    load_values(b.U0, U, (size_t)NV * b.nc);
}

void or_get_state(void* h, double* P, double* U) {
    Block& b = *static_cast<Block*>(h);
    if (P) store_values(b.P0, P, (size_t)NV * b.nc);
    if (U) store_values(b.U0, U, (size_t)NV * b.nc);
}

// Geometry cache: which = 0 gcov [NLOC][gn2][gn1][4][4], 1 gcon, 2 gdet [NLOC][gn2][gn1],
// 3 gdet*Gamma [gn2][gn1][4][4][4], 4 centre radius [gn2][gn1].  Returns the element count.
long or_get_geometry(void* h, int which, double* out) {
    Block& b = *static_cast<Block*>(h);
    const std::vector<Real>* v = which == 0 ? &b.gcov : which == 1 ? &b.gcon : which == 2 ? &b.gdet :
                                 which == 3 ? &b.gdet_conn : &b.rcen;
    if (out) store_values(*v, out, v->size());
    return (long)v->size();
}
// Replace the geometry cache (same layouts as or_get_geometry), e.g. to run the
// step on exactly the metric another implementation uses.
void or_set_geometry(void* h, int which, const double* in) {
    Block& b = *static_cast<Block*>(h);
    std::vector<Real>* v = which == 0 ? &b.gcov : which == 1 ? &b.gcon : which == 2 ? &b.gdet :
                           which == 3 ? &b.gdet_conn : &b.rcen;
    load_values(*v, in, v->size());
}
// Physical boundaries on the base state (KBoundaries::ApplyBoundary for each face).
// The physical boundary fill of the block's driver type (the synced-primitive
// fill for driver "simple").
void or_apply_bcs(void* h) { Block& b = *static_cast<Block*>(h); apply_physical_bcs(b, b.U0, b.P0, b.p.driver == 1); }
// Flux fixes on the stored fluxes (ZeroBoundaryFlux, FluxCT, KBoundaries::FixFlux).
void or_fix_flux(void* h) { Block& b = *static_cast<Block*>(h); zero_polar_b_flux(b); flux_ct(b); kboundaries_fix_flux(b); }

void or_set_dt(void* h, double dt) { static_cast<Block*>(h)->dt = dt; static_cast<Block*>(h)->dt_last = dt; }
double or_get_dt(void* h) { return static_cast<Block*>(h)->dt; }
double or_get_time(void* h) { return static_cast<Block*>(h)->time; }

// Run nsteps full RK2 steps; dts[n] receives the dt of each step (may be null).
int or_step(void* h, int nsteps, double* dts) {
    Block& b = *static_cast<Block*>(h);
    for (int n = 0; n < nsteps; ++n) {
        double taken = step(b);
        if (dts) dts[n] = taken;
    }
    return 0;
}

// Task-level entry points on the base container.
void or_get_flux(void* h, int dir) { Block& b = *static_cast<Block*>(h); get_flux(b, b.P0, dir); }
void or_flux_ct(void* h) { flux_ct(*static_cast<Block*>(h)); }
// The first-order flux correction of stage 1 from the block's state (U_sub = U_full = U0, P_sub = P0)
// on the fluxes of the preceding or_get_flux calls, which it rewrites with cmax/cmin.
void or_fofc_stage1(void* h) { Block& b = *static_cast<Block*>(h); fofc(b, 1, b.U0, b.P0); }
void or_flux_divergence(void* h) { flux_divergence(*static_cast<Block*>(h)); }
void or_add_geo_source(void* h) { Block& b = *static_cast<Block*>(h); add_geo_source(b, b.P0); }
void or_utop(void* h) { Block& b = *static_cast<Block*>(h); utop(b, b.U0, b.P0); }
void or_apply_floors(void* h) { Block& b = *static_cast<Block*>(h); apply_floors(b, b.U0, b.P0); }
void or_fix_utop(void* h) { Block& b = *static_cast<Block*>(h); fix_utop(b, b.U0, b.P0); }
void or_ptou(void* h) { Block& b = *static_cast<Block*>(h); ptou(b, b.P0, b.U0); }
double or_estimate_timestep(void* h) { return estimate_timestep(*static_cast<Block*>(h)); }
double or_max_divb(void* h) { Block& b = *static_cast<Block*>(h); return max_divb(b, b.U0); }

// Copy out internal arrays: which = 0 F (3*NV*nc), 1 cmax (3*nc), 2 cmin (3*nc),
// 3 dUdt (NV*nc), 4 pflag (nc), 5 fflag (nc), 6 emf (3*nc), 7 Pl, 8 Pr (NV*nc),
// after fofc(): 9 marks (nc), 10 trial primitives (NV*nc), 11 floor margins (nc);
// B_CT: 12 face fields fB0, 13 edge EMFs, 14 dB_Uf/dt (3*nf each), 15 cell-centre EMFs (3*nc)
void or_get_array(void* h, int which, double* out) {
    Block& b = *static_cast<Block*>(h);
    const std::vector<Real>* src = nullptr;
    switch (which) {
    case 0: src = &b.F; break; case 1: src = &b.cmax; break; case 2: src = &b.cmin; break;
    case 3: src = &b.dUdt; break; case 4: src = &b.pflag; break; case 5: src = &b.fflag; break;
    case 6: src = &b.emf; break; case 7: src = &b.Pl; break; case 8: src = &b.Pr; break;
    case 9: src = &b.fofc_mark; break; case 10: src = &b.fofc_trial; break; case 11: src = &b.fofc_margin; break;
    case 12: src = &b.fB0; break; case 13: src = &b.femf; break; case 14: src = &b.dfB; break;
    case 15: src = &b.cemf; break;
    default: return;
    }
    store_values(*src, out, src->size());
}
void or_set_array(void* h, int which, const double* in) {
    Block& b = *static_cast<Block*>(h);
    std::vector<Real>* dst = nullptr;
    switch (which) {
    case 0: dst = &b.F; break; case 1: dst = &b.cmax; break; case 2: dst = &b.cmin; break;
    case 3: dst = &b.dUdt; break; case 4: dst = &b.pflag; break; case 5: dst = &b.fflag; break;
    case 12: dst = &b.fB0; break; case 13: dst = &b.femf; break; case 14: dst = &b.dfB; break;
    default: return;
    }
    load_values(*dst, in, dst->size());
}

// Scalar kernels exposed for closed-form tests of the oracle itself.
void or_weno5(double x1, double x2, double x3, double x4, double x5, double* l, double* r) {
    Real lr, rr;
    weno5(x1, x2, x3, x4, x5, lr, rr);
    *l = val(lr); *r = val(rr);
}
void or_reconstruct(int recon, double x1, double x2, double x3, double x4, double x5, double* l, double* r) {
    Real lr, rr;
    reconstruct(recon, x1, x2, x3, x4, x5, lr, rr);
    *l = val(lr); *r = val(rr);
}
#ifdef KH_ORACLE_TANGENT
// Tangent-linear build: seed the directions of the base state P, U (entire
// block, [8][nc]); read the tangents of the base state after steps.
void or_set_tangent(void* h, const double* dP, const double* dU) {
    Block& b = *static_cast<Block*>(h);
    for (size_t q = 0; q < (size_t)NV * b.nc; ++q) {
        b.P0[q].d = dP ? dP[q] : 0.;
        b.U0[q].d = dU ? dU[q] : 0.;
    }
}
void or_get_tangent(void* h, double* dP, double* dU) {
    Block& b = *static_cast<Block*>(h);
    for (size_t q = 0; q < (size_t)NV * b.nc; ++q) {
        if (dP) dP[q] = b.P0[q].d;
        if (dU) dU[q] = b.U0[q].d;
    }
}
#endif
// WeightedSumData pair of AddStateUpdate (state_update's arithmetic) on
// caller arrays [8][nc], with the block's dUdt; Uout may alias Usub or Ufull.
void or_weighted_sum(void* h, const double* Usub, const double* Ufull, double gam0, double gam1, double beta_dt,
                     double* Uout) {
    Block& b = *static_cast<Block*>(h);
    const size_t n = (size_t)NV * b.nc;
    std::vector<Real> s(n), f(n), o(n);
    load_values(s, Usub, n);
    load_values(f, Ufull, n);
    for (size_t c = 0; c < n; ++c) o[c] = gam0 * s[c] + gam1 * f[c];
    for (size_t c = 0; c < n; ++c) o[c] = 1.0 * o[c] + beta_dt * b.dUdt[c];
    store_values(o, Uout, n);
}
// The periodic exchange of one block with itself on nvar caller arrays [nvar][nc].
void or_sync_vars(void* h, double* A, int nvar) {
    Block& b = *static_cast<Block*>(h);
    const size_t n = (size_t)nvar * b.nc;
    std::vector<Real> a(n);
    load_values(a, A, n);
    for (int v = 0; v < nvar; ++v) sync_periodic_var(b, a, v);
    store_values(a, A, n);
}

// Plant a known mistake in the restatement (see g_plant); 0 removes it.
void or_plant(int code) { g_plant = code; }

int or_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
