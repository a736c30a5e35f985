// kh_device.hpp -- device-side GRMHD physics for the gfx950 kernels.
//
// Every function here is templated on the arithmetic type R (double for the
// plain FP64 path, float for the split-precision flux path) and on a geometry
// policy G that supplies the metric at one point.  MinkowskiGeom specialises
// lower/raise/lorentz/alpha to the flat metric so that no 4x4 metric is read or
// multiplied; its results are identical to contracting with diag(-1,1,1,1).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace kh {

enum { RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4, B1 = 5, B2 = 6, B3 = 7, NV = 8 };

// Floor flag bits (floors/floors.hpp:47-68) and inversion status codes
// (inverter/invert_template.hpp:53): the flag arrays carry the same values.
constexpr int F_GEOM_RHO = 32, F_GEOM_U = 64, F_B_RHO = 128, F_B_U = 256, F_TEMP = 512,
              F_GAMMA = 1024, F_KTOT = 2048, F_GEOM_RHO_FLUX = 4096, F_GEOM_U_FLUX = 8192;
enum { S_SUCCESS = 0, S_NEG_INPUT, S_MAX_ITER, S_BAD_UT, S_BAD_GAMMA, S_NEG_RHO, S_NEG_U, S_NEG_RHOU };

template <typename R> struct FourVec { R ucon[4], ucov[4], bcon[4], bcov[4]; };

template <typename R> __device__ __forceinline__ R dot4(const R a[4], const R b[4]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}
template <typename R> __device__ __forceinline__ R dmax(R a, R b) { return a > b ? a : (b != b ? b : a); }
// std::max / std::min semantics (return the first argument unless the second compares greater/less)
template <typename R> __device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }
template <typename R> __device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }
// kharma_utils.hpp:100-110: min(max(lower, n), upper)
template <typename R> __device__ __forceinline__ R clip(R n, R lo, R hi) { return smin(smax(lo, n), hi); }

// Enables the geometry overloads below for non-arithmetic scalar types only, so
// that calls with float or double arguments resolve exactly as without them.
template <class T> using NonArith = std::enable_if_t<!std::is_arithmetic<T>::value>;
// Value of a scalar (the identity for double; the adjoint's dual numbers
// overload it).
__device__ __forceinline__ double val(double x) { return x; }

// Flat space, Cartesian coordinates.
template <typename R> struct MinkowskiGeom {
    static constexpr bool flat = true;
    __device__ __forceinline__ R gdet() const { return R(1); }
    __device__ __forceinline__ R alpha() const { return R(1); }
    __device__ __forceinline__ R gcov(int a, int b) const { return a != b ? R(0) : (a == 0 ? R(-1) : R(1)); }
    __device__ __forceinline__ R gcon(int a, int b) const { return a != b ? R(0) : (a == 0 ? R(-1) : R(1)); }
    __device__ __forceinline__ R gcon0(int v) const { return gcon(0, v); }
    __device__ __forceinline__ void lower(const R v[4], R o[4]) const { o[0] = -v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3]; }
    __device__ __forceinline__ void raise(const R v[4], R o[4]) const { o[0] = -v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3]; }
    __device__ __forceinline__ R lorentz(R u1, R u2, R u3) const {
        const R qsq = u1 * u1 + u2 * u2 + u3 * u3;
        return sqrt(R(1) + qsq);
    }
    // the same for non-arithmetic scalars (the adjoint's dual numbers)
    template <class T, class = NonArith<T>>
    __device__ __forceinline__ void lower(const T v[4], T o[4]) const { o[0] = -v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3]; }
    template <class T, class = NonArith<T>>
    __device__ __forceinline__ void raise(const T v[4], T o[4]) const { o[0] = -v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3]; }
    template <class T, class = NonArith<T>>
    __device__ __forceinline__ T lorentz(T u1, T u2, T u3) const {
        const T qsq = u1 * u1 + u2 * u2 + u3 * u3;
        return sqrt(T(1.) + qsq);
    }
};

// Element e of the array at p through a 32-bit byte offset: a wave-uniform base
// plus an unsigned offset is one base-register + offset memory instruction,
// where a 64-bit index costs 64-bit address arithmetic per access.  The arrays
// addressed this way (one variable of a block, one geometry table) are kept
// below 4 GiB by the block-size check of the C ABI.
template <class T> __device__ __forceinline__ T& at32(T* p, unsigned e) {
    return *(T*)((char*)p + (size_t)(e * (unsigned)sizeof(T)));
}
template <class T> __device__ __forceinline__ const T& at32(const T* p, unsigned e) {
    return *(const T*)((const char*)p + (size_t)(e * (unsigned)sizeof(T)));
}

// Metric of one point read from the geometry cache: 10 + 10 symmetric
// components, sqrt(-g) and the lapse, stored component-major ([22][np]) so a
// wavefront's loads of one component are contiguous.  Components are read
// where they are used; the pointer is const and restrict, so the compiler is
// free to keep or reload them.
// A32: components addressed through 32-bit offsets (at32), which the VALU-bound
// flux kernels use; the update kernels keep 64-bit indices, with which their
// register allocation measured better.
template <typename R, bool A32 = false> struct CachedGeom {
    static constexpr bool flat = false;
    const R* __restrict__ p;
    long np, q;
    __device__ __forceinline__ CachedGeom(const R* __restrict__ base, long np_, long q_) : p(base), np(np_), q(q_) {}
    static __device__ __forceinline__ constexpr int sym(int a, int b) {
        return (a <= b) ? (a == 0 ? b : a == 1 ? 3 + b : a == 2 ? 5 + b : 9) : sym(b, a);
    }
    __device__ __forceinline__ R comp(int n) const {
        if constexpr (A32) return at32(p, (unsigned)(n * np + q));
        else return p[(long)n * np + q];
    }
    __device__ __forceinline__ R gcov(int a, int b) const { return comp(sym(a, b)); }
    __device__ __forceinline__ R gcon(int a, int b) const { return comp(10 + sym(a, b)); }
    __device__ __forceinline__ R gcon0(int v) const { return gcon(0, v); }
    __device__ __forceinline__ R gdet() const { return comp(20); }
    __device__ __forceinline__ R alpha() const { return comp(21); }
    __device__ __forceinline__ void lower(const R v[4], R o[4]) const {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            o[m] = gcov(m, 0) * v[0] + gcov(m, 1) * v[1] + gcov(m, 2) * v[2] + gcov(m, 3) * v[3];
    }
    __device__ __forceinline__ void raise(const R v[4], R o[4]) const {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            o[m] = gcon(m, 0) * v[0] + gcon(m, 1) * v[1] + gcon(m, 2) * v[2] + gcon(m, 3) * v[3];
    }
    // GRMHD::lorentz_calc, grmhd_functions.hpp:116-129
    __device__ __forceinline__ R lorentz(R u1, R u2, R u3) const {
        const R qsq = gcov(1, 1) * u1 * u1 + gcov(2, 2) * u2 * u2 + gcov(3, 3) * u3 * u3 +
                      R(2) * (gcov(1, 2) * u1 * u2 + gcov(1, 3) * u1 * u3 + gcov(2, 3) * u2 * u3);
        return sqrt(R(1) + qsq);
    }
    // the same for non-arithmetic scalars (the adjoint's dual numbers)
    template <class T, class = NonArith<T>>
    __device__ __forceinline__ void lower(const T v[4], T o[4]) const {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            o[m] = gcov(m, 0) * v[0] + gcov(m, 1) * v[1] + gcov(m, 2) * v[2] + gcov(m, 3) * v[3];
    }
    template <class T, class = NonArith<T>>
    __device__ __forceinline__ void raise(const T v[4], T o[4]) const {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            o[m] = gcon(m, 0) * v[0] + gcon(m, 1) * v[1] + gcon(m, 2) * v[2] + gcon(m, 3) * v[3];
    }
    template <class T, class = NonArith<T>>
    __device__ __forceinline__ T lorentz(T u1, T u2, T u3) const {
        const T qsq = gcov(1, 1) * u1 * u1 + gcov(2, 2) * u2 * u2 + gcov(3, 3) * u3 * u3 +
                      R(2) * (gcov(1, 2) * u1 * u2 + gcov(1, 3) * u1 * u3 + gcov(2, 3) * u2 * u3);
        return sqrt(T(1.) + qsq);
    }
};

// Reciprocal from the hardware estimate plus Newton steps (~1 ulp; not the
// correctly rounded IEEE quotient, which costs ~12 instructions in FP64).
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
// FP32 (split-form fluxes): the hardware estimate alone, which is within 1 ulp.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Both face values of one cell from its 5-point stencil (x3 is the cell):
// lout at the left face, rout at the right face, WENO5 weights of
// flux/reconstruction.hpp:128-204.  The normalised weights
// (c_k / s_k) / sum_j (c_j / s_j), s_k = (EPS + beta_k)^2, are formed with one
// reciprocal per side: in FP64 by multiplying through by s0 s1 s2 (s >= 1e-52,
// so the products stay normal); in FP32 with r_k = min(t)/t_k in (0, 1],
// which neither overflows nor lets the dominant weight underflow.
__device__ __forceinline__ void weno5_cell(double x1, double x2, double x3, double x4, double x5,
                                           double& lout, double& rout) {
    constexpr double EPS = 1.e-26;
    double c1 = x1 - 2. * x2 + x3, c2 = x1 - 4. * x2 + 3. * x3;
    const double t0 = EPS + ((13. / 12.) * c1 * c1 + 0.25 * c2 * c2);
    c1 = x2 - 2. * x3 + x4; c2 = x4 - x2;
    const double t1 = EPS + ((13. / 12.) * c1 * c1 + 0.25 * c2 * c2);
    c1 = x3 - 2. * x4 + x5; c2 = x5 - 4. * x4 + 3. * x3;
    const double t2 = EPS + ((13. / 12.) * c1 * c1 + 0.25 * c2 * c2);
    const double s0 = t0 * t0, s1 = t1 * t1, s2 = t2 * t2;
    const double p12 = s1 * s2, p02 = s0 * s2, p01 = s0 * s1;
    const double ar0 = (1. / 16.) * p12, ar1 = (5. / 8.) * p02, ar2 = (5. / 16.) * p01;
    const double al0 = (1. / 16.) * p01, al1 = ar1, al2 = (5. / 16.) * p12;
    const double iWr = rcp(ar0 + ar1 + ar2), iWl = rcp(al0 + al1 + al2);
    rout = ((3. / 8.) * x1 - (5. / 4.) * x2 + (15. / 8.) * x3) * (ar0 * iWr) +
           ((-1. / 8.) * x2 + (3. / 4.) * x3 + (3. / 8.) * x4) * (ar1 * iWr) +
           ((3. / 8.) * x3 + (3. / 4.) * x4 - (1. / 8.) * x5) * (ar2 * iWr);
    lout = ((3. / 8.) * x5 - (5. / 4.) * x4 + (15. / 8.) * x3) * (al0 * iWl) +
           ((-1. / 8.) * x4 + (3. / 4.) * x3 + (3. / 8.) * x2) * (al1 * iWl) +
           ((3. / 8.) * x3 + (3. / 4.) * x2 - (1. / 8.) * x1) * (al2 * iWl);
}
__device__ __forceinline__ void weno5_cell(float x1, float x2, float x3, float x4, float x5,
                                           float& lout, float& rout) {
    constexpr float EPS = 1.e-26f;
    float c1 = x1 - 2.f * x2 + x3, c2 = x1 - 4.f * x2 + 3.f * x3;
    const float t0 = EPS + ((13.f / 12.f) * c1 * c1 + 0.25f * c2 * c2);
    c1 = x2 - 2.f * x3 + x4; c2 = x4 - x2;
    const float t1 = EPS + ((13.f / 12.f) * c1 * c1 + 0.25f * c2 * c2);
    c1 = x3 - 2.f * x4 + x5; c2 = x5 - 4.f * x4 + 3.f * x3;
    const float t2 = EPS + ((13.f / 12.f) * c1 * c1 + 0.25f * c2 * c2);
    const float tm = fminf(fminf(t0, t1), t2);
    const float r0 = tm * rcp(t0), r1 = tm * rcp(t1), r2 = tm * rcp(t2);
    const float q0 = r0 * r0, q1 = r1 * r1, q2 = r2 * r2;
    const float ar0 = (1.f / 16.f) * q0, ar1 = (5.f / 8.f) * q1, ar2 = (5.f / 16.f) * q2;
    const float al0 = (1.f / 16.f) * q2, al2 = (5.f / 16.f) * q0;
    const float iWr = rcp(ar0 + ar1 + ar2), iWl = rcp(al0 + ar1 + al2);
    rout = ((3.f / 8.f) * x1 - (5.f / 4.f) * x2 + (15.f / 8.f) * x3) * (ar0 * iWr) +
           ((-1.f / 8.f) * x2 + (3.f / 4.f) * x3 + (3.f / 8.f) * x4) * (ar1 * iWr) +
           ((3.f / 8.f) * x3 + (3.f / 4.f) * x4 - (1.f / 8.f) * x5) * (ar2 * iWr);
    lout = ((3.f / 8.f) * x5 - (5.f / 4.f) * x4 + (15.f / 8.f) * x3) * (al0 * iWl) +
           ((-1.f / 8.f) * x4 + (3.f / 4.f) * x3 + (3.f / 8.f) * x2) * (ar1 * iWl) +
           ((3.f / 8.f) * x3 + (3.f / 4.f) * x2 - (1.f / 8.f) * x1) * (al2 * iWl);
}

// ------------------------------------------------------------ reconstruction
// WENO5 of Tchekhovskoy+07 / Shu 2011 as in flux/reconstruction.hpp:128-204.
// The nonlinear weights (c_i / (EPS+beta_i)^2) / sum are evaluated in the
// equivalent scaled form c_i r_i^2 / sum(c r^2), r_i = min_j(EPS+beta_j)/(EPS+beta_i),
// which cannot under/overflow in FP32 (EPS^2 = 1e-52 underflows a float).
template <typename R>
__device__ __forceinline__ void weno5_beta(R x1, R x2, R x3, R x4, R x5, R t[3]) {
    constexpr R EPS = R(1.e-26);
    R c1 = x1 - R(2) * x2 + x3, c2 = x1 - R(4) * x2 + R(3) * x3;
    t[0] = EPS + (R(13. / 12.) * c1 * c1 + R(0.25) * c2 * c2);
    c1 = x2 - R(2) * x3 + x4; c2 = x4 - x2;
    t[1] = EPS + (R(13. / 12.) * c1 * c1 + R(0.25) * c2 * c2);
    c1 = x3 - R(2) * x4 + x5; c2 = x5 - R(4) * x4 + R(3) * x3;
    t[2] = EPS + (R(13. / 12.) * c1 * c1 + R(0.25) * c2 * c2);
}
// value at the right face of the centre cell (x3)
template <typename R>
__device__ __forceinline__ R weno5_right(R x1, R x2, R x3, R x4, R x5) {
    R t[3];
    weno5_beta(x1, x2, x3, x4, x5, t);
    const R tm = smin(smin(t[0], t[1]), t[2]);
    const R r0 = tm / t[0], r1 = tm / t[1], r2 = tm / t[2];
    const R w0 = R(1. / 16.) * r0 * r0, w1 = R(5. / 8.) * r1 * r1, w2 = R(5. / 16.) * r2 * r2;
    const R iW = R(1) / (w0 + w1 + w2);
    return ((R(3. / 8.) * x1 - R(5. / 4.) * x2 + R(15. / 8.) * x3) * (w0 * iW) +
            (R(-1. / 8.) * x2 + R(3. / 4.) * x3 + R(3. / 8.) * x4) * (w1 * iW) +
            (R(3. / 8.) * x3 + R(3. / 4.) * x4 - R(1. / 8.) * x5) * (w2 * iW));
}
// value at the left face of the centre cell (x3)
template <typename R>
__device__ __forceinline__ R weno5_left(R x1, R x2, R x3, R x4, R x5) {
    R t[3];
    weno5_beta(x1, x2, x3, x4, x5, t);
    const R tm = smin(smin(t[0], t[1]), t[2]);
    const R r0 = tm / t[0], r1 = tm / t[1], r2 = tm / t[2];
    const R w0 = R(1. / 16.) * r2 * r2, w1 = R(5. / 8.) * r1 * r1, w2 = R(5. / 16.) * r0 * r0;
    const R iW = R(1) / (w0 + w1 + w2);
    return ((R(3. / 8.) * x5 - R(5. / 4.) * x4 + R(15. / 8.) * x3) * (w0 * iW) +
            (R(-1. / 8.) * x4 + R(3. / 4.) * x3 + R(3. / 8.) * x2) * (w1 * iW) +
            (R(3. / 8.) * x3 + R(3. / 4.) * x2 - R(1. / 8.) * x1) * (w2 * iW));
}
// MC-limited linear (reconstruction.hpp:52-56,107-123)
template <typename R> __device__ __forceinline__ R mc_lim(R dm, R dp) {
    const R r = (fabs(dp) > R(0) ? dm / dp : R(2));
    return smax(R(0), smin(R(2), smin(R(2) * r, R(0.5) * (R(1) + r))));
}
// PPM with the CW84 limiter (reconstruction.hpp:351-382)
template <typename R>
__device__ __forceinline__ void ppm(R q_im2, R q_im1, R q_i, R q_ip1, R q_ip2, R& qlv, R& qrv) {
    qlv = (R(7) * (q_i + q_im1) - (q_im2 + q_ip1)) / R(12);
    qrv = (R(7) * (q_i + q_ip1) - (q_im1 + q_ip2)) / R(12);
    qlv = smax(qlv, smin(q_i, q_im1));
    qlv = smin(qlv, smax(q_i, q_im1));
    qrv = smax(qrv, smin(q_i, q_ip1));
    qrv = smin(qrv, smax(q_i, q_ip1));
    const R qc = qrv - q_i, qd = qlv - q_i;
    if ((qc * qd) >= R(0)) { qlv = q_i; qrv = q_i; }
    else {
        if (fabs(qc) >= R(2) * fabs(qd)) qrv = q_i - R(2) * qd;
        if (fabs(qd) >= R(2) * fabs(qc)) qlv = q_i - R(2) * qc;
    }
}

enum Recon { RECON_DONOR = 0, RECON_LINEAR_MC = 1, RECON_WENO5 = 2, RECON_PPM = 3 };

// Face states from a 6-point stencil x[-3..2] around face "0" (left face of cell 0):
// ql = right edge of cell -1, qr = left edge of cell 0.
template <int RECON, typename R>
__device__ __forceinline__ void face_states(const R x[6], R& ql, R& qr) {
    if constexpr (RECON == RECON_WENO5) {
        ql = weno5_right(x[0], x[1], x[2], x[3], x[4]);
        qr = weno5_left(x[1], x[2], x[3], x[4], x[5]);
    } else if constexpr (RECON == RECON_LINEAR_MC) {
        ql = x[2] + R(0.5) * (mc_lim(x[2] - x[1], x[3] - x[2]) * (x[3] - x[2]));
        qr = x[3] - R(0.5) * (mc_lim(x[3] - x[2], x[4] - x[3]) * (x[4] - x[3]));
    } else if constexpr (RECON == RECON_PPM) {
        R a, b;
        ppm(x[0], x[1], x[2], x[3], x[4], a, ql);
        ppm(x[1], x[2], x[3], x[4], x[5], qr, b);
    } else {
        ql = x[2];
        qr = x[3];
    }
}

// ----------------------------------------------------------------- GRMHD
// GRMHD::calc_4vecs, grmhd/grmhd_functions.hpp:225-246
template <typename R, class G>
__device__ __forceinline__ void calc_4vecs(const G& g, const R P[NV], FourVec<R>& D) {
    const R gamma = g.lorentz(P[U1], P[U2], P[U3]);
    if constexpr (G::flat) {
        D.ucon[0] = gamma;
        D.ucon[1] = P[U1]; D.ucon[2] = P[U2]; D.ucon[3] = P[U3];
    } else {
        const R alpha = g.alpha();
        D.ucon[0] = gamma / alpha;
        for (int v = 0; v < 3; ++v) D.ucon[v + 1] = P[U1 + v] - gamma * alpha * g.gcon0(v + 1);
    }
    g.lower(D.ucon, D.ucov);
    D.bcon[0] = P[B1] * D.ucov[1] + P[B2] * D.ucov[2] + P[B3] * D.ucov[3];
    const R iu0 = rcp(D.ucon[0]);
    for (int v = 0; v < 3; ++v) D.bcon[v + 1] = (P[B1 + v] + D.bcon[0] * D.ucon[v + 1]) * iu0;
    g.lower(D.bcon, D.bcov);
}

// Flux::prim_to_flux for dir = 0 (conserved) and dir = d together, sharing the
// tensor invariants.  flux_functions.hpp:116-178 + grmhd_functions.hpp:142-155.
template <typename R, class G, int DIR>
__device__ __forceinline__ void prim_to_u_and_flux(const G& g, const R P[NV], const FourVec<R>& D, R gam,
                                                   R U[NV], R F[NV]) {
    const R gdet = g.gdet();
    const R bsq = dot4(D.bcon, D.bcov);
    const R pgas = (gam - R(1)) * P[UU];
    const R eta = pgas + P[RHO] + P[UU] + bsq;
    const R ptot = pgas + R(0.5) * bsq;
    // dir = 0
    U[RHO] = P[RHO] * D.ucon[0] * gdet;
    {
        const R eu = eta * D.ucon[0];
        U[UU] = (eu * D.ucov[0] + ptot - D.bcon[0] * D.bcov[0]) * gdet + U[RHO];
        U[U1] = (eu * D.ucov[1] - D.bcon[0] * D.bcov[1]) * gdet;
        U[U2] = (eu * D.ucov[2] - D.bcon[0] * D.bcov[2]) * gdet;
        U[U3] = (eu * D.ucov[3] - D.bcon[0] * D.bcov[3]) * gdet;
        U[B1] = P[B1] * gdet; U[B2] = P[B2] * gdet; U[B3] = P[B3] * gdet;
    }
    // dir = DIR
    F[RHO] = P[RHO] * D.ucon[DIR] * gdet;
    {
        const R eu = eta * D.ucon[DIR];
        const R bd = D.bcon[DIR];
        F[UU] = (eu * D.ucov[0] - bd * D.bcov[0]) * gdet + F[RHO];
        F[U1] = (eu * D.ucov[1] + (DIR == 1 ? ptot : R(0)) - bd * D.bcov[1]) * gdet;
        F[U2] = (eu * D.ucov[2] + (DIR == 2 ? ptot : R(0)) - bd * D.bcov[2]) * gdet;
        F[U3] = (eu * D.ucov[3] + (DIR == 3 ? ptot : R(0)) - bd * D.bcov[3]) * gdet;
        const R ud = D.ucon[DIR];
        F[B1] = (D.bcon[1] * ud - bd * D.ucon[1]) * gdet;
        F[B2] = (D.bcon[2] * ud - bd * D.ucon[2]) * gdet;
        F[B3] = (D.bcon[3] * ud - bd * D.ucon[3]) * gdet;
    }
}

// Flux::vchar, flux_functions.hpp:368-439 (the two clip() calls there are no-ops,
// so cs2 and cms2 are deliberately left unclamped here too).
template <typename R, class G, int DIR>
__device__ __forceinline__ void vchar(const G& g, const R P[NV], const FourVec<R>& D, R gam, R& cmax, R& cmin) {
    const R ef = P[RHO] + gam * P[UU];
    const R cs2 = gam * (gam - R(1)) * P[UU] * rcp(ef);
    const R bsq = dot4(D.bcon, D.bcov);
    const R va2 = bsq * rcp(bsq + ef);
    const R cms2 = cs2 + va2 - cs2 * va2;
    R A, B, C;
    if constexpr (G::flat) {
        const R Au = D.ucon[DIR], Bu = D.ucon[0];
        A = Bu * Bu - (R(-1) + Bu * Bu) * cms2;
        B = R(2) * (Au * Bu - (Au * Bu) * cms2);
        C = Au * Au - (R(1) + Au * Au) * cms2;
    } else {
        const R Asq = g.gcon(DIR, DIR), Bsq = g.gcon(0, 0), AB = g.gcon(0, DIR);
        const R Au = D.ucon[DIR], Bu = D.ucon[0];
        A = Bu * Bu - (Bsq + Bu * Bu) * cms2;
        B = R(2) * (Au * Bu - (AB + Au * Bu) * cms2);
        C = Au * Au - (Asq + Au * Au) * cms2;
    }
    const R discr = sqrt(smax(B * B - R(4) * A * C, R(0)));
    const R i2A = rcp(R(2) * A);
    const R vp = -(-B + discr) * i2A;
    const R vm = -(-B - discr) * i2A;
    cmax = smax(vp, vm);
    cmin = smin(vp, vm);
}

// ----------------------------------------------------------- inversion (FP64)
// Inverter::u_to_p<onedw>, inverter/onedw.hpp:49-233.  Always FP64.  Quotients
// are formed as products with reciprocals (~1 ulp), and the relative
// convergence tests |x / Wp| < tol as |x| < tol |Wp|.
__device__ __forceinline__ double lorentz_calc_w(double Bsq, double D, double QdB, double Qtsq, double Wp) {
    const double QdBsq = QdB * QdB;
    const double W = Wp + D;
    const double W2 = W * W;
    const double WB = W + Bsq;
    const double utsq = -((W + WB) * QdBsq + W2 * Qtsq) * rcp(QdBsq * (W + WB) + W2 * (Qtsq - WB * WB));
    if (utsq < -1.e-15 || utsq > 1.e7) return -1.;
    return sqrt(1. + fabs(utsq));
}
// gm1g = (gam - 1) / gam
__device__ __forceinline__ double err_eqn(double gm1g, double Bsq, double D, double Ep, double QdB, double Qtsq,
                                          double Wp, int& eflag) {
    const double W = Wp + D;
    const double gamma = lorentz_calc_w(Bsq, D, QdB, Qtsq, Wp);
    if (gamma < 1) eflag = S_BAD_UT;
    const double ig = rcp(gamma);
    const double w = W * (ig * ig);
    const double rho = D * ig;
    const double p = (w - rho) * gm1g;
    const double bw = Bsq + W;
    return -Ep + Wp - p + 0.5 * Bsq + 0.5 * (Bsq * Qtsq - QdB * QdB) * rcp(bw * bw);
}
// U: 8 conserved at the point; P: in = guess (rho,u,u1..3), out = recovered (only on success)
template <class G>
__device__ __forceinline__ int u_to_p_onedw(const G& g, const double U[NV], double P[NV], double gam,
                                            int max_iterations, double tol) {
    if (U[RHO] <= 0.) return S_NEG_INPUT;
    double alpha, a_over_g;
    if constexpr (G::flat) { alpha = 1.; a_over_g = 1.; }
    else { alpha = g.alpha(); a_over_g = alpha * rcp(g.gdet()); }
    const double gm1g = (gam - 1) / gam;
    const double D = U[RHO] * a_over_g;
    double Bcon[4] = {0., U[B1] * a_over_g, U[B2] * a_over_g, U[B3] * a_over_g};
    const double Qcov[4] = {(U[UU] - U[RHO]) * a_over_g, U[U1] * a_over_g, U[U2] * a_over_g, U[U3] * a_over_g};
    const double ncov[4] = {-alpha, 0., 0., 0.};
    double Bcov[4], Qcon[4], ncon[4];
    g.lower(Bcon, Bcov);
    g.raise(Qcov, Qcon);
    if constexpr (G::flat) { ncon[0] = alpha; ncon[1] = 0.; ncon[2] = 0.; ncon[3] = 0.; }
    else g.raise(ncov, ncon);
    const double Bsq = dot4(Bcon, Bcov);
    const double QdB = dot4(Bcon, Qcov);
    const double Qdotn = dot4(Qcon, ncov);
    double Qtcon[4];
    for (int mu = 0; mu < 4; ++mu) Qtcon[mu] = Qcon[mu] + ncon[mu] * Qdotn;
    const double Qtsq = dot4(Qcon, Qcov) + Qdotn * Qdotn;
    const double Ep = -Qdotn - D;
    int eflag = S_SUCCESS;
    double Wp, err;
    {
        const double gamma = g.lorentz(P[U1], P[U2], P[U3]);
        if (gamma < 1) return S_BAD_UT;
        const double rho = P[RHO], u = P[UU];
        Wp = (rho + u + (gam - 1) * u) * gamma * gamma - rho * gamma;
        err = err_eqn(gm1g, Bsq, D, Ep, QdB, Qtsq, Wp, eflag);
    }
    double dW;
    {
        const double Wpm = (1. - 1.e-5) * Wp;
        const double h = Wp - Wpm;
        const double Wpp = Wp + h;
        const double errm = err_eqn(gm1g, Bsq, D, Ep, QdB, Qtsq, Wpm, eflag);
        const double errp = err_eqn(gm1g, Bsq, D, Ep, QdB, Qtsq, Wpp, eflag);
        const double dedW = (errp - errm) * rcp(Wpp - Wpm);
        const double dedW2 = (errp - 2. * err + errm) * rcp(h * h);
        const double f = clip(0.5 * err * dedW2 * rcp(dedW * dedW), -0.3, 0.3);
        dW = clip(-err * rcp(dedW * (1. - f)), -0.5 * Wp, 2.0 * Wp);
    }
    double Wp1 = Wp;
    double err1 = err;
    Wp += dW;
    err = err_eqn(gm1g, Bsq, D, Ep, QdB, Qtsq, Wp, eflag);
    int iter = 0;
    for (iter = 0; iter < max_iterations; iter++) {
        dW = clip((Wp1 - Wp) * err * rcp(err - err1), -0.5 * Wp, 2.0 * Wp);
        Wp1 = Wp;
        err1 = err;
        Wp += dW;
        if (fabs(dW) < tol * fabs(Wp)) break;
        err = err_eqn(gm1g, Bsq, D, Ep, QdB, Qtsq, Wp, eflag);
        if (fabs(err) < tol * fabs(Wp)) break;
    }
    if (iter == max_iterations) return S_MAX_ITER;
    const double gamma = lorentz_calc_w(Bsq, D, QdB, Qtsq, Wp);
    if (gamma < 1) return S_BAD_UT;
    const double ig = rcp(gamma);
    const double rho = D * ig;
    const double W = Wp + D;
    const double w = W * (ig * ig);
    const double p = (w - rho) * gm1g;
    const double u = w - (rho + p);
    if (rho < 0 && u < 0) return S_NEG_RHOU;
    else if (rho < 0) return S_NEG_RHO;
    else if (u < 0) return S_NEG_U;
    P[RHO] = rho;
    P[UU] = u;
    const double pre = gamma * rcp(W + Bsq);
    const double qw = QdB * rcp(W);
    P[U1] = pre * (Qtcon[1] + qw * Bcon[1]);
    P[U2] = pre * (Qtcon[2] + qw * Bcon[2]);
    P[U3] = pre * (Qtcon[3] + qw * Bcon[3]);
    return S_SUCCESS;
}

// Conserved variables of all 8 fields from P and its four-vectors D
// (Flux::p_to_u, flux_functions.hpp:343-352).
template <class G>
__device__ __forceinline__ void p_to_u_D(const G& g, const double P[NV], const FourVec<double>& D, double gam,
                                         double U[NV]) {
    const double gdet = g.gdet();
    const double bsq = dot4(D.bcon, D.bcov);
    const double pgas = (gam - 1.) * P[UU];
    const double eta = pgas + P[RHO] + P[UU] + bsq;
    const double ptot = pgas + 0.5 * bsq;
    const double eu = eta * D.ucon[0];
    U[RHO] = P[RHO] * D.ucon[0] * gdet;
    U[UU] = (eu * D.ucov[0] + ptot - D.bcon[0] * D.bcov[0]) * gdet + U[RHO];
    U[U1] = (eu * D.ucov[1] - D.bcon[0] * D.bcov[1]) * gdet;
    U[U2] = (eu * D.ucov[2] - D.bcon[0] * D.bcov[2]) * gdet;
    U[U3] = (eu * D.ucov[3] - D.bcon[0] * D.bcov[3]) * gdet;
    U[B1] = P[B1] * gdet; U[B2] = P[B2] * gdet; U[B3] = P[B3] * gdet;
}

// (gam - 1) u / rho^gam > ktot_max (the entropy ceiling test of
// floors_functions.hpp:147-148), decided in log2 space from hardware log2
// estimates and falling back to the exact pow() only within 1e-3 of the
// threshold, so the outcome equals the exact comparison.
__device__ __forceinline__ double log2_est(double x) {
    int e;
    const double m = frexp(x, &e);   // x = m 2^e, m in [0.5, 1)
    return (double)e + (double)__builtin_amdgcn_logf((float)m);
}
__device__ __forceinline__ bool ktot_exceeds(double u, double rho, double gam, double ktot_max) {
    const double a = (gam - 1.) * u;
    if (a > 0. && rho > 0. && ktot_max > 0. && a < 1e300 && rho < 1e300 && ktot_max < 1e300 &&
        a > 1e-300 && rho > 1e-300) {
        const double d = log2_est(a) - gam * log2_est(rho) - log2_est(ktot_max);
        if (d > 1e-3) return true;
        if (d < -1e-3) return false;
    }
    return a / pow(rho, gam) > ktot_max;
}

// Conserved variables of all 8 fields (Flux::p_to_u, flux_functions.hpp:343-352)
template <class G>
__device__ __forceinline__ void p_to_u(const G& g, const double P[NV], double gam, double U[NV]) {
    FourVec<double> D;
    calc_4vecs<double>(g, P, D);
    p_to_u_D(g, P, D, gam, U);
}

// ------------------------------------------------------------------- floors
struct FloorParams {
    double rho_min_geom, u_min_geom, rho_min_const, u_min_const, r_char;
    double bsq_over_rho_max, bsq_over_u_max, u_over_rho_max, ktot_max, gamma_max;
    int use_r_char, temp_adjust_u, frame, on;
};

// Floors::determine_floors, floors/floors_functions.hpp:84-160.  bsq = b^mu b_mu
// of the zone's current state; rhoflr_geom / uflr_geom are the geometric floors
// of the zone (constants in Cartesian coordinates, r-dependent in spherical ones).
// T is double, or the adjoint's dual numbers (its floor Jacobians).
template <class G, class T>
__device__ __forceinline__ int determine_floors(const G& g, const T P[NV], T bsq, double gam,
                                                const FloorParams& f, double rhoflr_geom, double uflr_geom,
                                                T& rhoflr_max, T& uflr_max) {
    const T rhoflr_b = bsq / f.bsq_over_rho_max;
    const T uflr_b = bsq / f.bsq_over_u_max;
    uflr_max = smax(T(uflr_geom), uflr_b);
    const T rho = P[RHO], u = P[UU];
    int fflag = 0;
    if (!f.temp_adjust_u) {
        const T rhoflr_temp = smax(u, uflr_max) / f.u_over_rho_max;
        fflag |= (rhoflr_temp > rho) * F_TEMP;
        rhoflr_max = smax(smax(T(rhoflr_geom), rhoflr_b), rhoflr_temp);
    } else {
        rhoflr_max = smax(T(rhoflr_geom), rhoflr_b);
    }
    if (rhoflr_max > rho || uflr_max > u) {
        fflag |= (rhoflr_geom > rho) * F_GEOM_RHO;
        fflag |= (uflr_geom > u) * F_GEOM_U;
        fflag |= (rhoflr_b > rho) * F_B_RHO;
        fflag |= (uflr_b > u) * F_B_U;
    }
    if (g.lorentz(P[U1], P[U2], P[U3]) > f.gamma_max) fflag |= F_GAMMA;
    if (ktot_exceeds(val(P[UU]), val(P[RHO]), gam, f.ktot_max)) fflag |= F_KTOT;
    if (f.temp_adjust_u && (P[UU] / P[RHO] > f.u_over_rho_max)) fflag |= F_TEMP;
    return fflag;
}

// Drift-frame floors, floors_functions.hpp:186-257
template <class G, class T>
__device__ __forceinline__ void apply_floors_drift(const G& g, T P[NV], double gam, T rhoflr_max, T uflr_max) {
    double beta[4] = {0., 0., 0., 0.};
    if constexpr (!G::flat) {
        const double lapse2 = rcp(-g.gcon(0, 0));
        beta[1] = lapse2 * g.gcon(0, 1);
        beta[2] = lapse2 * g.gcon(0, 2);
        beta[3] = lapse2 * g.gcon(0, 3);
    }
    // Quotients are products with rcp() (~1 ulp) rather than IEEE divisions.
    const T rho = P[RHO], uu = P[UU];
    const T pg = (gam - 1.) * uu;
    const T w_old = smax(rho + uu + pg, T(1e-20));
    T Bcon[4] = {T(0.), P[B1], P[B2], P[B3]};
    T Bcov[4];
    g.lower(Bcon, Bcov);
    const T Bsq = smax(dot4(Bcon, Bcov), T(1e-20));
    const T B_mag = sqrt(Bsq);
    const T iBm = rcp(B_mag);
    FourVec<T> D;
    calc_4vecs<T>(g, P, D);
    const T u0 = D.ucon[0], iu0 = rcp(u0);
    T Qcov[4];
    Qcov[0] = w_old * u0 * D.ucov[0] + pg;
    Qcov[1] = w_old * u0 * D.ucov[1];
    Qcov[2] = w_old * u0 * D.ucov[2];
    Qcov[3] = w_old * u0 * D.ucov[3];
    const T QdotB = dot4(Bcon, Qcov);
    T vpar = QdotB * rcp(B_mag * w_old * u0 * u0);
    T ucon_dr[4];
    const T udr0 = rcp(sqrt(iu0 * iu0 + vpar * vpar));
    {
        const T a = udr0 * iu0, c = vpar * udr0 * iBm;
        for (int mu = 0; mu < 4; ++mu) ucon_dr[mu] = D.ucon[mu] * a - Bcon[mu] * c;
    }
    P[RHO] = smax(rho, rhoflr_max);
    P[UU] = smax(uu, uflr_max);
    const T pg_new = (gam - 1.) * P[UU];
    const T w_new = P[RHO] + P[UU] + pg_new;
    const T iudr0 = rcp(ucon_dr[0]);
    const T x = (2. * QdotB) * rcp(B_mag * w_new * ucon_dr[0]);
    vpar = x * rcp(1. + sqrt(1. + x * x)) * iudr0;
    T ucon[4];
    ucon[0] = rcp(sqrt(iudr0 * iudr0 - vpar * vpar));
    {
        const T a = ucon[0] * iudr0, c = vpar * ucon[0] * iBm;
        for (int mu = 1; mu < 4; ++mu) ucon[mu] = ucon_dr[mu] * a + Bcon[mu] * c;
    }
    P[U1] = ucon[1] + (beta[1] * ucon[0]);
    P[U2] = ucon[2] + (beta[2] * ucon[0]);
    P[U3] = ucon[3] + (beta[3] * ucon[0]);
}

// Inverter::u_to_p<kastaun>, inverter/kastaun.hpp:83-400 (Kastaun, Kalinani &
// Ciolfi 2021): the conserved state is moved to the normal observer's frame,
// the auxiliary function mu sqrt(1 + rbar^2(mu)) - 1 (their eq. 49) is
// bracketed on [0, 1] and the master function f(mu) = mu - muhat(mu) (eq. 44)
// solved on [0, root], both by false position with the Illinois halving.  The
// inverter's own floors (density and specific-energy floors, the velocity
// ceiling from gamma_max, the specific-energy ceiling) act inside the master
// function and are reported as the INVERTER_* floor bits; a master solve that
// does not converge in max_iterations returns S_MAX_ITER.  Always FP64, with
// IEEE quotients.
constexpr int F_INVERTER_RHO = 16384, F_INVERTER_U = 32768, F_INVERTER_GAMMA = 65536, F_INVERTER_U_MAX = 131072;
struct KastaunMaster {
    double D, q, bsq, bsq_rpsq, rsq, rbsq, v0sq, gam, rho_floor, e_floor, e_max;
    __device__ __forceinline__ double x_mu(double mu) const { return 1.0 / (1.0 + mu * bsq); }
    __device__ __forceinline__ double rbarsq_mu(double mu, double x) const { return x * (x * rsq + mu * (1.0 + x) * rbsq); }
    __device__ __forceinline__ double qbar_mu(double mu, double x) const {
        const double mux = mu * x;
        return q - 0.5 * (bsq + mux * mux * bsq_rpsq);
    }
    __device__ __forceinline__ double vhatsq_mu(double mu, double rbarsq, int& fl) const {
        const double v = mu * mu * rbarsq;
        if (v > v0sq) { fl |= F_INVERTER_GAMMA; return v0sq; }
        return v;
    }
    __device__ __forceinline__ double rhohat_mu(double iW, int& fl) const {
        const double r = D * iW;
        if (r <= rho_floor) { fl |= F_INVERTER_RHO; return rho_floor; }
        return r;
    }
    __device__ __forceinline__ double ehat_mu(double mu, double qbar, double rbarsq, double vhatsq, double W,
                                              int& fl) const {
        const double e = W * (qbar - mu * rbarsq) + vhatsq * W * W / (1.0 + W);
        if (e <= e_floor) { fl |= F_INVERTER_U; return e_floor; }
        if (e > e_max) { fl |= F_INVERTER_U_MAX; return e_max; }
        return e;
    }
    __device__ __forceinline__ double aux(double mu) const {
        const double x = 1.0 / (1.0 + mu * bsq);
        const double rbarsq = x * (rsq * x + mu * (1.0 + x) * rbsq);
        return mu * sqrt(1.0 + rbarsq) - 1.0;
    }
    __device__ __forceinline__ double master(double mu) const {
        int fl = 0;
        const double x = x_mu(mu);
        const double rbarsq = rbarsq_mu(mu, x);
        const double qbar = qbar_mu(mu, x);
        const double vhatsq = vhatsq_mu(mu, rbarsq, fl);
        const double iW = sqrt(1.0 - vhatsq);
        const double W = 1.0 / iW;
        const double rhohat = rhohat_mu(iW, fl);
        const double ehat = ehat_mu(mu, qbar, rbarsq, vhatsq, W, fl);
        const double Phat = ehat * rhohat * (gam - 1.0);
        const double hhat = rhohat * (1.0 + ehat) + Phat;
        const double ahat = Phat / (hhat - Phat);
        const double nua = (1.0 + ahat) * (1.0 + ehat) * iW;
        const double nub = (1.0 + ahat) * (1.0 + qbar - mu * rbarsq);
        const double nuhat = nua > nub ? nua : nub;
        return mu - 1.0 / (nuhat + mu * rbarsq);
    }
};
// False position with the Illinois halving on [zm, zp] (the reference's loop,
// kastaun.hpp:298-324 and 338-364); returns the last iterate, iter the count.
template <class F>
__device__ __forceinline__ double illinois(F&& f, double zm, double zp, int max_iterations, double tol, int& iter) {
    double fm = f(zm), fp = f(zp);
    int iterations = max_iterations;
    if (fabs(zm - zp) < tol || (fabs(fm) + fabs(fp)) < 2.0 * tol) iterations = -1;
    double z = 0.5 * (zm + zp);
    for (iter = 0; iter < iterations; ++iter) {
        z = (zm * fp - zp * fm) / (fp - fm);
        const double fz = f(z);
        if (fabs(zm - zp) < tol || fabs(fz) < tol) break;
        if (fz * fp < 0.0) { zm = zp; fm = fp; zp = z; fp = fz; }
        else { fm = 0.5 * fm; zp = z; fp = fz; }
    }
    if (iterations < 0) iter = 0;
    return z;
}
template <class G>
__device__ __forceinline__ int u_to_p_kastaun(const G& g, const double U[NV], double P[NV], double gam,
                                              const FloorParams& fp, int max_iterations, double tol) {
    double alpha, a_over_g;
    if constexpr (G::flat) { alpha = 1.; a_over_g = 1.; }
    else { alpha = 1. / sqrt(-g.gcon(0, 0)); a_over_g = alpha / g.gdet(); }
    const double D = U[RHO] * a_over_g;
    const double D_fl = D > fp.rho_min_const ? D : fp.rho_min_const;
    const double Qcov[4] = {(U[UU] - U[RHO]) * a_over_g, U[U1] * a_over_g, U[U2] * a_over_g, U[U3] * a_over_g};
    const double ncov[4] = {-alpha, 0., 0., 0.};
    double ncon[4];
    g.raise(ncov, ncon);
    const double q = (-dot4(Qcov, ncon) - D) / D;
    const double rcov[3] = {U[U1] / U[RHO], U[U2] / U[RHO], U[U3] / U[RHO]};
    double rcon[3];
    const double g00 = g.gcon(0, 0);
    for (int a = 0; a < 3; ++a) {
        rcon[a] = 0.;
        for (int b = 0; b < 3; ++b) rcon[a] += (g.gcon(a + 1, b + 1) - g.gcon(0, a + 1) * g.gcon(0, b + 1) / g00) * rcov[b];
    }
    const double rsq = rcon[0] * rcov[0] + rcon[1] * rcov[1] + rcon[2] * rcov[2];
    const double sD = 1.0 / sqrt(D_fl);
    double bu[3], bdotr = 0.;
    for (int a = 0; a < 3; ++a) {
        bu[a] = (U[B1 + a] * a_over_g) * sD;
        bdotr += bu[a] * rcov[a];
    }
    double bsq = 0.;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) bsq += g.gcov(a + 1, b + 1) * bu[a] * bu[b];
    bsq = bsq > 0. ? bsq : 0.;
    const double rbsq = bdotr * bdotr;
    KastaunMaster m;
    m.D = D; m.q = q; m.bsq = bsq; m.bsq_rpsq = bsq * rsq - rbsq; m.rsq = rsq; m.rbsq = rbsq;
    const double vmax = 1.0 - 1.0 / (fp.gamma_max * fp.gamma_max);
    m.v0sq = rsq / (1.0 + rsq) < vmax ? rsq / (1.0 + rsq) : vmax;
    m.gam = gam; m.rho_floor = fp.rho_min_const; m.e_floor = fp.u_min_const / D_fl; m.e_max = fp.u_over_rho_max;
    int iter;
    const double zb = illinois([&](double mu) { return m.aux(mu); }, 0., 1., max_iterations, tol, iter);
    const double mu = illinois([&](double z) { return m.master(z); }, 0., zb, max_iterations, tol, iter);
    if (iter == max_iterations) return S_MAX_ITER;
    int fl = 0;
    const double x = m.x_mu(mu);
    const double rbarsq = m.rbarsq_mu(mu, x);
    const double vsq = m.vhatsq_mu(mu, rbarsq, fl);
    const double iW = sqrt(1.0 - vsq);
    const double W = 1.0 / iW;
    const double rho = m.rhohat_mu(iW, fl);
    const double qbar = m.qbar_mu(mu, x);
    P[RHO] = rho;
    P[UU] = m.ehat_mu(mu, qbar, rbarsq, vsq, W, fl) * rho;
    for (int a = 0; a < 3; ++a) P[U1 + a] = W * mu * x * (rcon[a] + mu * bdotr * bu[a]);
    return fl;
}

// Normal-observer-frame floors, floors_functions.hpp:260-289 (U is modified in registers)
template <class G>
__device__ __forceinline__ int apply_floors_normal(const G& g, double P[NV], double U[NV], double gam,
                                                   double rhoflr_max, double uflr_max) {
    const double rho_add = smax(0., rhoflr_max - P[RHO]);
    const double u_add = smax(0., uflr_max - P[UU]);
    double Padd[NV] = {rho_add, u_add, 0., 0., 0., 0., 0., 0.};
    double Uadd[NV];
    p_to_u(g, Padd, gam, Uadd);
    P[RHO] += rho_add;
    P[UU] += u_add;
    for (int v = 0; v < 5; ++v) U[v] += Uadd[v];
    return u_to_p_onedw(g, U, P, gam, 8, 1e-8);
}

// Floors::apply_ceilings, floors_functions.hpp:50-82
// ktot = (gam - 1) u / rho^gam is formed (with its pow) only where the entropy
// ceiling applies; ktot_exceeds decides that exactly from log2 estimates.
template <class G, class T>
__device__ __forceinline__ void apply_ceilings(const G& g, T P[NV], double gam, const FloorParams& f) {
    const T gamma = g.lorentz(P[U1], P[U2], P[U3]);
    const T u0 = P[UU], rho0 = P[RHO];
    if (gamma > f.gamma_max) {
        const T fr = sqrt((f.gamma_max * f.gamma_max - 1.) / (gamma * gamma - 1.));
        P[U1] = P[U1] * fr; P[U2] = P[U2] * fr; P[U3] = P[U3] * fr;
    }
    if (ktot_exceeds(val(u0), val(rho0), gam, f.ktot_max)) {
        const T ktot = (gam - 1.) * u0 / pow(rho0, gam);
        P[UU] = f.ktot_max / ktot * P[UU];
    }
    if (f.temp_adjust_u && P[UU] / P[RHO] > f.u_over_rho_max) P[UU] = f.u_over_rho_max * P[RHO];
}

// One zone of Floors::ApplyFloorsInFrame (floors_impl.hpp:45-128): determine,
// apply in the configured frame, ceilings.  Returns fflag; may overwrite pflag.
// D holds the four-vectors of P on entry and on return (recomputed if P changed).
// FRAME (0 fluid, 1 normal, 2 drift) fixes the frame at compile time, which
// keeps the other frames' code and registers out of the kernel; FRAME_RUNTIME
// reads it from f.frame.
constexpr int FRAME_RUNTIME = -1;
template <int FRAME = FRAME_RUNTIME, class G>
__device__ __forceinline__ int floors_zone(const G& g, double P[NV], double U[NV], FourVec<double>& D, double gam,
                                           const FloorParams& f, double rhoflr_geom, double uflr_geom,
                                           int fflag_in, int& pflag) {
    double rf, uf;
    const int fflag = fflag_in | determine_floors(g, P, dot4(D.bcon, D.bcov), gam, f, rhoflr_geom, uflr_geom, rf, uf);
    if (fflag) {
        int pflag_l = 0;
        const int frame = (FRAME == FRAME_RUNTIME) ? f.frame : FRAME;
        if (frame == 0) {
            P[RHO] += smax(0., rf - P[RHO]);
            P[UU] += smax(0., uf - P[UU]);
        } else if (frame == 1) {
            pflag_l = apply_floors_normal(g, P, U, gam, rf, uf);
        } else {
            apply_floors_drift(g, P, gam, rf, uf);
        }
        if (pflag_l) pflag = pflag_l;
        apply_ceilings(g, P, gam, f);
        calc_4vecs<double>(g, P, D);
    }
    return fflag;
}

// KBoundaries::check_inflow, boundaries/boundaries.hpp:101-136: velocity into
// the domain through a physical X1 face is replaced by zero radial 4-velocity.
// (The clip() of vsq there discards its result, so vsq is used unclamped.)
template <class G, class T>
__device__ __forceinline__ void check_inflow(const G& g, T P[NV], bool inner) {
    T uvec[3] = {P[U1], P[U2], P[U3]};
    const T gamma0 = g.lorentz(uvec[0], uvec[1], uvec[2]);
    const T ucon1 = uvec[0] - gamma0 * g.alpha() * g.gcon0(1);
    if ((ucon1 > 0. && inner) || (ucon1 < 0. && !inner)) {
        for (int v = 0; v < 3; ++v) uvec[v] = uvec[v] / gamma0;
        const double alpha = g.alpha();
        const double beta1 = g.gcon(0, 1) * alpha * alpha;
        uvec[0] = beta1 / alpha;
        const T vsq = g.gcov(1, 1) * uvec[0] * uvec[0] + g.gcov(2, 2) * uvec[1] * uvec[1] +
                           g.gcov(3, 3) * uvec[2] * uvec[2] +
                           2. * (g.gcov(1, 2) * uvec[0] * uvec[1] + g.gcov(1, 3) * uvec[0] * uvec[2] +
                                 g.gcov(2, 3) * uvec[1] * uvec[2]);
        const T gamma = 1. / sqrt(1. - vsq);
        for (int v = 0; v < 3; ++v) P[U1 + v] = uvec[v] * gamma;
    }
}

// Geometric floors on one state (Floors::apply_geo_floors, floors_functions.hpp:300-367)
template <typename R>
__device__ __forceinline__ void apply_geo_floors(R P[NV], R rho_min, R u_min) {
    P[RHO] += smax(R(0), rho_min - P[RHO]);
    P[UU] += smax(R(0), u_min - P[UU]);
}

// ------------------------------------------------------------ reductions
// Wavefront reduction with cross-lane shuffles (wave64), then LDS across waves.
__device__ __forceinline__ double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
// Non-negative doubles order like their bit patterns: atomic min/max on u64.
__device__ __forceinline__ void atomic_min_pos(unsigned long long* addr, double v) {
    atomicMin(addr, (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ void atomic_max_pos(unsigned long long* addr, double v) {
    atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

}  // namespace kh
