// kh_step.hip -- flux-CT, the fused stage update, inversion fix-ups, reductions
// and the task-level kernels that mirror the reference's individual tasks.
//
// The fused update kernel performs, for one interior zone and in registers:
//   flux-CT B fluxes from the corner EMFs (B_FluxCT::FluxCT, b_flux_ct.cpp:347-369)
//   + flux divergence (KHARMADriver::FluxDivergence, kharma_driver.hpp:208-231)
//   + RK state update (KHARMADriver::AddStateUpdate, kharma_driver.cpp:339-387)
//   + onedw U->P (Inverter::BlockUtoP, inverter.cpp:131-182) and B U->P (b_flux_ct.cpp:195-217)
//   + floors (Floors::ApplyGRMHDFloors, floors_impl.hpp:45-128)
//   + P->U (Flux::MeshPtoU, flux.cpp:276-322)
//   + (last stage) the CFL estimate (GRMHD::EstimateTimestep, grmhd.cpp:313-328)
// and writes U, P and the flags once, also into the periodic ghost images of
// the zone (the boundary exchange), so no separate sync pass reads the state.
#include <algorithm>
#include <mutex>

#include "kh_internal.hpp"

#include <cfloat>

namespace kh {

static constexpr int BS = 256;
// Work-groups per CU the update kernel is compiled for: 4 (128 VGPRs) with the
// flat metric; the cached-metric variant needs ~240 VGPRs and would spill at 4.
template <int GEO> struct UpdateOcc { static constexpr int min_blocks = GEO ? 2 : 4; };

static inline dim3 grid_for(long n, int bs = BS) { return dim3((unsigned)((n + bs - 1) / bs)); }
static inline long ninterior_host(const BlockDims& b) {
    return (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
}

// FP32 shadow of the primitives read by the split-precision flux kernels.
__global__ void __launch_bounds__(256) convert_f32_kernel(const double* __restrict__ A, float* __restrict__ B, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) B[t] = (float)A[t];
}

// ------------------------------------------------------------ ghost images
// Images of interior zone (k,j,i) in ghost regions: per axis, the zone's
// offset in the image block's layout (0 = no image) and the index offset of
// that block (0: this block, a periodic axis of a one-block mesh; otherwise
// the neighbour across the face the zone lies within ng of).  With nx >= 2 ng
// on every image axis a zone has at most one image per axis, so up to 7 image
// copies (edges and corners land in the diagonal neighbour, whose block offset
// is the sum of the axes' offsets).
struct ImageSet {
    long oi, oj, ok;
    int bi, bj, bk;
};
__device__ __forceinline__ void image_set(const BlockDims& b, int k, int j, int i, ImageSet& s) {
    s.oi = s.oj = s.ok = 0;
    s.bi = s.bj = s.bk = 0;
    if (b.img[0][0] && i - b.is < b.ng) { s.oi = b.nx1; s.bi = b.nbd[0][0]; }
    else if (b.img[0][1] && b.ie - i < b.ng) { s.oi = -b.nx1; s.bi = b.nbd[0][1]; }
    if (b.ndim > 1) {
        if (b.img[1][0] && j - b.js < b.ng) { s.oj = (long)b.nx2 * b.s2; s.bj = b.nbd[1][0]; }
        else if (b.img[1][1] && b.je - j < b.ng) { s.oj = -(long)b.nx2 * b.s2; s.bj = b.nbd[1][1]; }
    }
    if (b.ndim > 2) {
        if (b.img[2][0] && k - b.ks < b.ng) { s.ok = (long)b.nx3 * b.s3; s.bk = b.nbd[2][0]; }
        else if (b.img[2][1] && b.ke - k < b.ng) { s.ok = -(long)b.nx3 * b.s3; s.bk = b.nbd[2][1]; }
    }
}
// base[c] = val and the same at every image; bstride: elements between the
// same array of consecutive blocks (NV nc for a state, nc for a flag array).
template <typename T>
__device__ __forceinline__ void store_with_images(const BlockDims&, const ImageSet& s, T* base, long c, T val,
                                                  long bstride) {
    base[c] = val;
    if ((s.oi | s.oj | s.ok) == 0) return;
    const long di = s.oi + s.bi * bstride, dj = s.oj + s.bj * bstride, dk = s.ok + s.bk * bstride;
    if (s.oi) base[c + di] = val;
    if (s.oj) base[c + dj] = val;
    if (s.ok) base[c + dk] = val;
    if (s.oi && s.oj) base[c + di + dj] = val;
    if (s.oi && s.ok) base[c + di + dk] = val;
    if (s.oj && s.ok) base[c + dj + dk] = val;
    if (s.oi && s.oj && s.ok) base[c + di + dj + dk] = val;
}

__device__ __forceinline__ void interior_index(const BlockDims& b, unsigned t, int& k, int& j, int& i) {
    const unsigned nij = (unsigned)b.nx1 * (unsigned)(b.ndim > 1 ? b.nx2 : 1);
    const unsigned kk = t / nij;
    const unsigned rem = t - kk * nij;
    const unsigned jj = rem / (unsigned)b.nx1;
    i = b.is + (int)(rem - jj * (unsigned)b.nx1);
    j = b.js + (int)jj;
    k = b.ks + (int)kk;
}
__device__ __forceinline__ long ninterior(const BlockDims& b) {
    return (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
}

// Metric of cell centre (j, i): flat, or read from the geometry cache.
template <int GEO, class Fn>
__device__ __forceinline__ void with_center_geom(const GeomDev& gd, int j, int i, Fn&& fn) {
    if constexpr (GEO) {
        CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, (long)j * gd.gn1 + i);
        fn(g);
    } else {
        MinkowskiGeom<double> g;
        fn(g);
    }
}
// Geometric floors of cell centre (j, i) (floors_functions.hpp:97-108)
template <int GEO>
__device__ __forceinline__ void geom_floors(const StepConsts& c, const GeomDev& gd, int j, int i, double& rf, double& uf) {
    if constexpr (GEO) {
        const long q = (long)j * gd.gn1 + i;
        rf = gd.flr[q];
        uf = gd.flr[gd.np + q];
    } else {
        rf = c.floors.rho_min_const;
        uf = c.floors.u_min_const;
    }
}
// KDomain::GetPhysicalRange (domain.hpp:161-187): the entire block, less the
// ghosts across physical-boundary faces (the ghosts of periodic faces and of
// faces shared with another block are in range).
__device__ __forceinline__ bool in_physical(const BlockDims& b, int k, int j, int i) {
    if (i < 0 || i >= b.n1 || j < 0 || j >= b.n2 || k < 0 || k >= b.n3) return false;
    if ((i < b.is && !b.img[0][0]) || (i > b.ie && !b.img[0][1])) return false;
    if (b.ndim > 1 && ((j < b.js && !b.img[1][0]) || (j > b.je && !b.img[1][1]))) return false;
    if (b.ndim > 2 && ((k < b.ks && !b.img[2][0]) || (k > b.ke && !b.img[2][1]))) return false;
    return true;
}

// Flux::AddGeoSource (flux/flux.cpp:395-444): the T^mu_nu Gamma contraction of
// zone (j, i) added to the energy and momentum rates,
// S_lam = sum_{mu, nu} T^mu_nu sqrt(-g) Gamma^nu_{lam mu}.  conn holds
// sqrt(-g) Gamma^a_{b c} at index 16 a + 4 b + c, exactly symmetric in (b, c)
// (the geometry build stores the mean of the two finite-difference orders), so
// each of the 40 distinct components is loaded once and serves both (lam, mu)
// and (mu, lam): 40 loads instead of 64 (the reference sums all 64; the sum
// here differs from it by rounding only).  No test for zero components: a
// branch per component would make every load wait for the one before it.
// conn(n) returns component n = 16 a + 4 b + c of this zone's table.
template <class G, class C>
__device__ __forceinline__ void add_geo_source_c(const G& g, const C& conn, const double Psub[NV], double gam,
                                                 double dudt[NV]) {
    FourVec<double> D;
    calc_4vecs<double>(g, Psub, D);
    const double bsq = dot4(D.bcon, D.bcov);
    const double pgas = (gam - 1.) * Psub[UU];
    const double eta = pgas + Psub[RHO] + Psub[UU] + bsq;
    const double ptot = pgas + 0.5 * bsq;
    double nd[4] = {0., 0., 0., 0.};
    double T[4][4];   // T^mu_nu
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) T[mu][nu] = eta * D.ucon[mu] * D.ucov[nu] + (mu == nu ? ptot : 0.) - D.bcon[mu] * D.bcov[nu];
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) {
        double cs[10];
#pragma unroll
        for (int lam = 0, n = 0; lam < 4; ++lam)
#pragma unroll
            for (int mu = lam; mu < 4; ++mu, ++n) cs[n] = conn(nu * 16 + lam * 4 + mu);
#pragma unroll
        for (int lam = 0, n = 0; lam < 4; ++lam)
#pragma unroll
            for (int mu = lam; mu < 4; ++mu, ++n) {
                nd[lam] += T[mu][nu] * cs[n];
                if (mu != lam) nd[mu] += T[lam][nu] * cs[n];
            }
    }
    dudt[UU] += nd[0];
    dudt[U1] += nd[1];
    dudt[U2] += nd[2];
    dudt[U3] += nd[3];
}
template <class G>
__device__ __forceinline__ void add_geo_source(const G& g, const GeomDev& gd, long q, const double Psub[NV], double gam,
                                               double dudt[NV]) {
    add_geo_source_c(g, [&](int n) { return gd.conn[(long)n * gd.np + q]; }, Psub, gam, dudt);
}

// ----------------------------------------------------------------- flux-CT
// Corner EMFs, b_flux_ct.cpp:331-342: emf_d at corners i in [is, ie+1],
// j in [js, je+1], k in [ks, ke(+1 in 3D)].  F = [3][8][nc], emf = [3][nc].
template <typename R>
__global__ void __launch_bounds__(BS) emf_kernel(BlockDims b, FluxVars<const R> F, R* __restrict__ emf) {
    const int ni = b.nx1 + 1, nj = b.nx2 + 1, nk = (b.ndim > 2) ? b.nx3 + 1 : 1;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)(ni * nj * nk)) return;
    const int kk = t / (ni * nj), rem = t - kk * (ni * nj), jj = rem / ni;
    const int i = b.is + rem - jj * ni, j = b.js + jj, k = b.ks + kk;
    const long c = (long)k * b.s3 + (long)j * b.s2 + i;
    const Vars<const R>& F1 = F[0];
    const Vars<const R>& F2 = F[1];
    const Vars<const R>& F3 = F[2];
    if (b.ndim > 2) {
        emf[c] = R(0.25) * (F2[B3][c] + F2[B3][c - b.s3] - F3[B2][c] - F3[B2][c - b.s2]);
        emf[b.nc + c] = R(0.25) * (F3[B1][c] + F3[B1][c - 1] - F1[B3][c] - F1[B3][c - b.s3]);
    }
    emf[2 * b.nc + c] = R(0.25) * (F1[B2][c] + F1[B2][c - b.s2] - F2[B1][c] - F2[B1][c - 1]);
}

// Rewrite the B fluxes from the EMFs (b_flux_ct.cpp:347-369), for the task-level API.
template <typename R>
__global__ void __launch_bounds__(BS) ct_rewrite_kernel(BlockDims b, FluxVars<R> F, const R* __restrict__ emf) {
    const int ni = b.nx1 + 1, nj = b.nx2 + 1, nk = (b.ndim > 2) ? b.nx3 + 1 : 1;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)(ni * nj * nk)) return;
    const int kk = t / (ni * nj), rem = t - kk * (ni * nj), jj = rem / ni;
    const int i = b.is + rem - jj * ni, j = b.js + jj, k = b.ks + kk;
    const long c = (long)k * b.s3 + (long)j * b.s2 + i;
    const Vars<R>& F1 = F[0];
    const Vars<R>& F2 = F[1];
    const Vars<R>& F3 = F[2];
    const R* e1 = emf;
    const R* e2 = emf + b.nc;
    const R* e3 = emf + 2 * b.nc;
    const bool d3 = b.ndim > 2;
    const bool kin = k <= b.ke, jin = j <= b.je, iin = i <= b.ie;
    if (kin && jin) {  // flux_ct_1: i in [is, ie+1]
        F1[B1][c] = R(0);
        F1[B2][c] = R(0.5) * (e3[c] + e3[c + b.s2]);
        if (d3) F1[B3][c] = R(-0.5) * (e2[c] + e2[c + b.s3]);
    }
    if (kin && iin) {  // flux_ct_2: j in [js, je+1]
        F2[B1][c] = R(-0.5) * (e3[c] + e3[c + 1]);
        F2[B2][c] = R(0);
        if (d3) F2[B3][c] = R(0.5) * (e1[c] + e1[c + b.s3]);
    }
    if (d3 && jin && iin) {  // flux_ct_3: k in [ks, ke+1]
        F3[B1][c] = R(0.5) * (e2[c] + e2[c + 1]);
        F3[B2][c] = R(-0.5) * (e1[c] + e1[c + b.s2]);
        F3[B3][c] = R(0);
    }
}

// Flux divergence of all 8 variables of zone c (FluxDivHelper form:
// -(sum_d A_d (F_d(hi) - F_d(lo))) / V), with the flux-CT B fluxes
// evaluated from the EMFs when `ct` is set.
template <typename R>
__device__ __forceinline__ void zone_divergence(const BlockDims& b, const FluxVars<const R>& F, const R* __restrict__ emf,
                                                long c, bool ct, double dudt[NV], bool skip_b = false) {
    const double A1 = b.dx2 * b.dx3, A2 = b.dx1 * b.dx3, A3 = b.dx1 * b.dx2, V = b.dx1 * b.dx2 * b.dx3;
    const double iV = 1. / V;   // one division per zone instead of eight
    const Vars<const R>& F1 = F[0];
    const Vars<const R>& F2 = F[1];
    const Vars<const R>& F3 = F[2];
    const long s2 = b.s2, s3 = b.s3;
    const int nd = b.ndim;
    double f1lo[NV], f1hi[NV], f2lo[NV], f2hi[NV], f3lo[NV], f3hi[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        f1lo[v] = f1hi[v] = f2lo[v] = f2hi[v] = f3lo[v] = f3hi[v] = 0.;
        if (skip_b && v >= B1) continue;   // face-centred B: the cell B comes from the faces
        f1lo[v] = (double)F1[v][c];
        f1hi[v] = (double)F1[v][c + 1];
        if (nd >= 2 && !(ct && v >= B1)) { f2lo[v] = (double)F2[v][c]; f2hi[v] = (double)F2[v][c + s2]; }
        if (nd == 3 && !(ct && v >= B1)) { f3lo[v] = (double)F3[v][c]; f3hi[v] = (double)F3[v][c + s3]; }
    }
    if (ct) {
        const R* e1 = emf;
        const R* e2 = emf + b.nc;
        const R* e3 = emf + 2 * b.nc;
        f1lo[B1] = 0.; f1hi[B1] = 0.;
        f1lo[B2] = (double)(R(0.5) * (e3[c] + e3[c + s2]));
        f1hi[B2] = (double)(R(0.5) * (e3[c + 1] + e3[c + 1 + s2]));
        f2lo[B1] = (double)(R(-0.5) * (e3[c] + e3[c + 1]));
        f2hi[B1] = (double)(R(-0.5) * (e3[c + s2] + e3[c + s2 + 1]));
        f2lo[B2] = 0.; f2hi[B2] = 0.;
        if (nd == 3) {
            f1lo[B3] = (double)(R(-0.5) * (e2[c] + e2[c + s3]));
            f1hi[B3] = (double)(R(-0.5) * (e2[c + 1] + e2[c + 1 + s3]));
            f2lo[B3] = (double)(R(0.5) * (e1[c] + e1[c + s3]));
            f2hi[B3] = (double)(R(0.5) * (e1[c + s2] + e1[c + s2 + s3]));
            f3lo[B1] = (double)(R(0.5) * (e2[c] + e2[c + 1]));
            f3hi[B1] = (double)(R(0.5) * (e2[c + s3] + e2[c + s3 + 1]));
            f3lo[B2] = (double)(R(-0.5) * (e1[c] + e1[c + s2]));
            f3hi[B2] = (double)(R(-0.5) * (e1[c + s3] + e1[c + s3 + s2]));
            f3lo[B3] = 0.; f3hi[B3] = 0.;
        } else {
            // 2D: B3 keeps its Riemann fluxes
            f2lo[B3] = (double)F2[B3][c];
            f2hi[B3] = (double)F2[B3][c + s2];
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double du = (A1 * f1hi[v] - A1 * f1lo[v]);
        if (nd >= 2) du += (A2 * f2hi[v] - A2 * f2lo[v]);
        if (nd == 3) du += (A3 * f3hi[v] - A3 * f3lo[v]);
        dudt[v] = -du * iV;
    }
}

__device__ __forceinline__ double zone_ndt(const BlockDims& b, double c1, double c2, double c3) {
    // GRMHD::EstimateTimestep, grmhd.cpp:320-322, with c_d = max(cmax_d, cmin_d)
    return 1.0 / (1 / (b.dx1 / c1) + 1 / (b.dx2 / c2) + 1 / (b.dx3 / c3));
}

// Block-wide min / max (wavefront shuffles, then LDS across the block's waves).
__device__ __forceinline__ double block_min(double v) {
    __shared__ double red[1024 / 64];
    v = wave_min(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmin(m, red[q]);
    return m;
}
__device__ __forceinline__ double block_max(double v) {
    __shared__ double red[1024 / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmax(m, red[q]);
    return m;
}

// Corner EMFs of flux-CT (b_flux_ct.cpp:331-342) evaluated from the Riemann
// B fluxes at corner c.
template <typename R>
__device__ __forceinline__ R emf3_at(const R* F1B2, const R* F2B1, long c, long s2) {
    return R(0.25) * (F1B2[c] + F1B2[c - s2] - F2B1[c] - F2B1[c - 1]);
}
template <typename R>
__device__ __forceinline__ R emf1_at(const R* F2B3, const R* F3B2, long c, long s2, long s3) {
    return R(0.25) * (F2B3[c] + F2B3[c - s3] - F3B2[c] - F3B2[c - s2]);
}
template <typename R>
__device__ __forceinline__ R emf2_at(const R* F3B1, const R* F1B3, long c, long s3) {
    return R(0.25) * (F3B1[c] + F3B1[c - 1] - F1B3[c] - F1B3[c - s3]);
}

// Divergence of zone c with the flux-CT B fluxes (b_flux_ct.cpp:347-369) built
// from EMFs computed here from the Riemann fluxes, so neither the EMFs nor the
// rewritten B fluxes make a trip through HBM.
// ND (2 or 3) fixes the dimensionality at compile time; ND = 0 reads b.ndim.
template <typename R, int ND = 0>
__device__ __forceinline__ void zone_divergence_ct(const BlockDims& b, const R* __restrict__ F, long c, double dudt[NV]) {
    const double A1 = b.dx2 * b.dx3, A2 = b.dx1 * b.dx3, A3 = b.dx1 * b.dx2, V = b.dx1 * b.dx2 * b.dx3;
    const double iV = 1. / V;   // one division per zone instead of eight
    const R* F1 = F;
    const R* F2 = F + 8 * b.nc;
    const R* F3 = F + 16 * b.nc;
    const long s2 = b.s2, s3 = b.s3, nc = b.nc;
    const int nd = ND ? ND : b.ndim;
    double f1lo[NV], f1hi[NV], f2lo[NV], f2hi[NV], f3lo[NV], f3hi[NV];
#pragma unroll
    for (int v = 0; v < B1; ++v) {
        f1lo[v] = (double)F1[v * nc + c];
        f1hi[v] = (double)F1[v * nc + c + 1];
        f2lo[v] = (double)F2[v * nc + c];
        f2hi[v] = (double)F2[v * nc + c + s2];
        if (nd == 3) { f3lo[v] = (double)F3[v * nc + c]; f3hi[v] = (double)F3[v * nc + c + s3]; }
        else { f3lo[v] = 0.; f3hi[v] = 0.; }
    }
    const R* F1B2 = F1 + B2 * nc;
    const R* F2B1 = F2 + B1 * nc;
    const R e3_00 = emf3_at(F1B2, F2B1, c, s2), e3_10 = emf3_at(F1B2, F2B1, c + 1, s2);
    const R e3_01 = emf3_at(F1B2, F2B1, c + s2, s2), e3_11 = emf3_at(F1B2, F2B1, c + s2 + 1, s2);
    f1lo[B1] = 0.; f1hi[B1] = 0.;
    f1lo[B2] = (double)(R(0.5) * (e3_00 + e3_01));
    f1hi[B2] = (double)(R(0.5) * (e3_10 + e3_11));
    f2lo[B1] = (double)(R(-0.5) * (e3_00 + e3_10));
    f2hi[B1] = (double)(R(-0.5) * (e3_01 + e3_11));
    f2lo[B2] = 0.; f2hi[B2] = 0.;
    f3lo[B1] = f3hi[B1] = f3lo[B2] = f3hi[B2] = f3lo[B3] = f3hi[B3] = 0.;
    if (nd == 3) {
        const R* F2B3 = F2 + B3 * nc;
        const R* F3B2 = F3 + B2 * nc;
        const R* F3B1 = F3 + B1 * nc;
        const R* F1B3 = F1 + B3 * nc;
        const R e1_000 = emf1_at(F2B3, F3B2, c, s2, s3), e1_010 = emf1_at(F2B3, F3B2, c + s2, s2, s3);
        const R e1_001 = emf1_at(F2B3, F3B2, c + s3, s2, s3), e1_011 = emf1_at(F2B3, F3B2, c + s2 + s3, s2, s3);
        const R e2_000 = emf2_at(F3B1, F1B3, c, s3), e2_100 = emf2_at(F3B1, F1B3, c + 1, s3);
        const R e2_001 = emf2_at(F3B1, F1B3, c + s3, s3), e2_101 = emf2_at(F3B1, F1B3, c + s3 + 1, s3);
        f1lo[B3] = (double)(R(-0.5) * (e2_000 + e2_001));
        f1hi[B3] = (double)(R(-0.5) * (e2_100 + e2_101));
        f2lo[B3] = (double)(R(0.5) * (e1_000 + e1_001));
        f2hi[B3] = (double)(R(0.5) * (e1_010 + e1_011));
        f3lo[B1] = (double)(R(0.5) * (e2_000 + e2_100));
        f3hi[B1] = (double)(R(0.5) * (e2_001 + e2_101));
        f3lo[B2] = (double)(R(-0.5) * (e1_000 + e1_010));
        f3hi[B2] = (double)(R(-0.5) * (e1_001 + e1_011));
    } else {
        // 2D: B3 keeps its Riemann fluxes
        f1lo[B3] = (double)F1[B3 * nc + c];
        f1hi[B3] = (double)F1[B3 * nc + c + 1];
        f2lo[B3] = (double)F2[B3 * nc + c];
        f2hi[B3] = (double)F2[B3 * nc + c + s2];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double du = (A1 * f1hi[v] - A1 * f1lo[v]);
        du += (A2 * f2hi[v] - A2 * f2lo[v]);
        if (nd == 3) du += (A3 * f3hi[v] - A3 * f3lo[v]);
        dudt[v] = -du * iV;
    }
}

// The U->P solver of inverter/type (inverter.cpp:53-70): INV 0 onedw, 1 Kastaun.
template <int INV, class G>
__device__ __forceinline__ int invert_zone(const G& g, const double U[NV], double P[NV], const StepConsts& cst) {
    if constexpr (INV == 1) return u_to_p_kastaun(g, U, P, cst.gam, cst.floors, cst.iter_max, cst.err_tol);
    else return u_to_p_onedw(g, U, P, cst.gam, cst.iter_max, cst.err_tol);
}

// ------------------------------------------------------- fused stage update
// Stage 1 (gam0 = 0 and U_sub == U_full) reads the state once: 0*U + 1*U == U.
// BM: 0 no B transport (1-D), 1 flux-CT EMFs from the face fluxes, 2 face-centred
// B (B_CT): the cell B is the average of the stage's new faces fB (B_CT::BlockUtoP,
// b_ct.cpp:173-205, which the reference runs before the inversion).
template <typename R, bool DO_DT, int BM, int GEO, int INV = 0, bool PK = false>
__global__ void __launch_bounds__(BS, UpdateOcc<GEO>::min_blocks)
update_kernel(BlockDims b, StepConsts cst, GeomDev gd, int stage, const R* __restrict__ F, const R* __restrict__ ctop,
              const double* __restrict__ Usub, const double* Ufull, const double* __restrict__ Psub,
              double* Uout, double* __restrict__ Pout, float* __restrict__ Pout32,
              int* __restrict__ pflag, int* __restrict__ fflag, int* __restrict__ fail_list,
              DeviceState* __restrict__ ds, double* __restrict__ dt_part, const double* __restrict__ fB,
              const BlockEntry* __restrict__ pk) {
    if constexpr (PK) {
        const long blk = pack_select(pk, b, gd);
        const long sv = blk * NV * b.nc;
        F = pack_at(F, 3 * sv);
        ctop = pack_at(ctop, blk * 3 * b.nc);
        Usub = pack_at(Usub, sv); Ufull = pack_at(Ufull, sv); Psub = pack_at(Psub, sv);
        Uout = pack_at(Uout, sv); Pout = pack_at(Pout, sv); Pout32 = pack_at(Pout32, sv);
        pflag = pack_at(pflag, blk * b.nc); fflag = pack_at(fflag, blk * b.nc); fail_list = pack_at(fail_list, blk * b.nc);
        ds += blk;
        dt_part = pack_at(dt_part, blk * gridDim.x);
    }
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = t < (unsigned)ninterior(b);
    double ndt = DBL_MAX;
    if (active) {
        int k, j, i;
        interior_index(b, t, k, j, i);
        const long c = (long)k * b.s3 + (long)j * b.s2 + i;
        const double gam = cst.gam;
        double dudt[NV];
        if (BM == 1) zone_divergence_ct<R>(b, F, c, dudt);
        else zone_divergence<R>(b, flux_vars_of(F, b.nc), nullptr, c, false, dudt, BM == 2);
        if constexpr (GEO) {
            double Ps[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) Ps[v] = Psub[v * b.nc + c];
            const long q = (long)j * gd.gn1 + i;
            CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, q);
            add_geo_source(g, gd, q, Ps, gam, dudt);
        }
        const double g0 = cst.gam0[stage - 1], g1 = cst.gam1[stage - 1];
        const double beta_dt = cst.beta[stage - 1] * ds->dt;
        const bool same = (g0 == 0.) && (Usub == Ufull);
        double U[NV], P[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double uf = Ufull[v * b.nc + c];
            const double u = same ? uf : g0 * Usub[v * b.nc + c] + g1 * uf;
            U[v] = 1.0 * u + beta_dt * dudt[v];
        }
        if constexpr (BM == 2) {
            const long f = (long)k * b.fs3 + (long)j * b.fs2 + i;
            U[B1] = (fB[f] + fB[f + 1]) / 2;
            U[B2] = (fB[b.nf + f] + fB[b.nf + f + b.fs2]) / 2;
            U[B3] = (b.ndim > 2) ? (fB[2 * b.nf + f] + fB[2 * b.nf + f + b.fs3]) / 2 : fB[2 * b.nf + f];
        }
        // U->P guess: the stage's starting primitives (kharma_driver.cpp:379-384)
#pragma unroll
        for (int v = 0; v < 5; ++v) P[v] = Psub[v * b.nc + c];
        int pf = 0, ff = 0;
        double rf, uf;
        geom_floors<GEO>(cst, gd, j, i, rf, uf);
        with_center_geom<GEO>(gd, j, i, [&](const auto& g) {
            pf = invert_zone<INV>(g, U, P, cst);
            const int ff0 = (pf / F_GEOM_RHO) * F_GEOM_RHO;
            pf = pf % F_GEOM_RHO;
            const double gd_ = g.gdet();
            const double igd = rcp(gd_);
            P[B1] = U[B1] * igd; P[B2] = U[B2] * igd; P[B3] = U[B3] * igd;
            FourVec<double> D;
            calc_4vecs<double>(g, P, D);
            ff = ff0;
            if (cst.floors.on) ff = floors_zone(g, P, U, D, gam, cst.floors, rf, uf, ff0, pf);
            p_to_u_D(g, P, D, gam, U);
        });
        ImageSet s;
        image_set(b, k, j, i, s);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            store_with_images(b, s, Uout + v * b.nc, c, U[v], NV * b.nc);
            store_with_images(b, s, Pout + v * b.nc, c, P[v], NV * b.nc);
            if (Pout32) store_with_images(b, s, Pout32 + v * b.nc, c, (float)P[v], NV * b.nc);
        }
        store_with_images(b, s, pflag, c, pf, b.nc);
        fflag[c] = ff;
        if (pf > 0) {
            const int slot = atomicAdd(&ds->fail_count[stage - 1], 1);
            fail_list[slot] = (int)c;
        }
        if (DO_DT) {
            const double n = zone_ndt(b, (double)ctop[c], b.ndim > 1 ? (double)ctop[b.nc + c] : 0.,
                                      b.ndim > 2 ? (double)ctop[2 * b.nc + c] : 0.);
            if (!(n != n)) ndt = n;
        }
    }
    if (DO_DT) {
        const double m = block_min(ndt);
        if (threadIdx.x == 0) dt_part[blockIdx.x] = m;
    }
}

// ------------------------------------------- curved-space stage update, split
// In Kerr-Schild geometry the fused update needs ~240 VGPRs (two waves per
// SIMD) and spends most of its time waiting.  It runs as two kernels instead:
// combine_kernel (flux divergence, geometric source, RK combination -> U) and
// recover_kernel (U->P, floors, P->U, flags).  The arithmetic is the same as
// the fused kernel's, so the results are bit-identical; the cost is one extra
// write and read of U (128 bytes per zone).
//
// Both walk the interior in (j, k, i) order, i fastest, with the workgroups
// renumbered so each XCD (workgroup w runs on XCD w % 8) owns one contiguous
// range of j: the axisymmetric geometry tables ([comp][j][i]) a workgroup
// reads then stay within one eighth of the table per XCD L2.
// Waves per SIMD the two halves are built for (the second launch bound).  The
// staged 3-D recovery on 64-wide tiles runs 4 (128 VGPRs with a 12-byte spill
// instead of 140: 262 -> 234 us on the 192x128x128 torus); on 32-wide tiles
// (28-byte spill) 4 waves measured 1 % slower, and the recovery in the
// normal-observer frame (182-241 VGPRs) and the unstaged ones would spill
// hundreds of bytes, so those keep their natural allocation.  The combine keeps
// its 2 waves: a 3-wave bound (168 VGPRs, 20-90 byte spill) measured 1-7 % slower.
template <int FRAME, bool STG, int TI> struct RecOcc {
    static constexpr int w = (STG && FRAME != 1 && TI == 64) ? 4 : 1;
};
constexpr unsigned N_XCD = 8;
__device__ __forceinline__ unsigned xcd_logical_block(unsigned b, unsigned nb) {
    const unsigned x = b % N_XCD, r = b / N_XCD, q = nb / N_XCD, m = nb % N_XCD;
    return x * q + (x < m ? x : m) + r;
}
__device__ __forceinline__ void interior_index_jki(const BlockDims& b, unsigned t, int& k, int& j, int& i) {
    const unsigned n3 = (unsigned)(b.ndim > 2 ? b.nx3 : 1);
    const unsigned row = t / (unsigned)b.nx1;
    const unsigned jj = row / n3;
    i = b.is + (int)(t - row * (unsigned)b.nx1);
    k = b.ks + (int)(row - jj * n3);
    j = b.js + (int)jj;
}

// In 3-D the recover kernel runs on tiles of TI = 64 x1 cells by 4 x3 rows at
// one x2 row, one x3 row per wavefront: the four wavefronts of a work-group
// share the geometry of their 64 (j, i) columns (the 22 metric components and
// the 2 geometric floors), which the work-group stages once in LDS as
// [component][64] instead of each zone loading its own.  A block whose x1 width
// is an odd multiple of 32 (the 96-wide blocks of a 2x2x4 mesh of the
// 192x128x128 torus) runs TI = 32 by 8 rows instead, two rows per wavefront, so
// that no tile is half empty.  Tiles are numbered i fastest, then k, then j, and
// renumbered per XCD as above.  Each zone's U and P loads are issued before the
// staging so that they overlap it (272 -> 262 us on the 192x128x128 torus).
// (Staging the connection as well and adding the geometric source here instead
// of in the combine shortens the combine -- 75-102 VGPRs instead of 190-254 --
// but slows the recovery by more: measured 223 + 469 against 371 + 262 us for
// the split stage-2 update.)
constexpr int STG_RECOVER = GEO_NCOMP + 2;
template <int TI> __host__ __device__ __forceinline__ int tiles_i(const BlockDims& b) { return (b.nx1 + TI - 1) / TI; }
template <int TI> __host__ __device__ __forceinline__ int tiles_k(const BlockDims& b) {
    return (b.nx3 + BS / TI - 1) / (BS / TI);
}
// The recovery's tile width for block b: 32 where it leaves fewer idle lanes.
static inline int recover_tile_i(const BlockDims& b) {
    return (tiles_i<32>(b) * 32 - b.nx1 < tiles_i<64>(b) * 64 - b.nx1) ? 32 : 64;
}
static inline long update_tiles_host(const BlockDims& b) {
    return recover_tile_i(b) == 32 ? (long)tiles_i<32>(b) * tiles_k<32>(b) * b.nx2
                                   : (long)tiles_i<64>(b) * tiles_k<64>(b) * b.nx2;
}
// Zone (k, j, i) of this thread; i0 is the tile's first x1 cell.
template <int TI>
__device__ __forceinline__ bool tile_zone(const BlockDims& b, int& k, int& j, int& i, int& i0) {
    const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const unsigned nti = (unsigned)tiles_i<TI>(b), ntk = (unsigned)tiles_k<TI>(b);
    const unsigned ti = lb % nti, rest = lb / nti;
    const unsigned tk = rest % ntk, jj = rest / ntk;
    const int row = (TI == 64) ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x / TI)) : (int)(threadIdx.x / TI);
    i0 = b.is + (int)ti * TI;
    i = i0 + (int)(threadIdx.x % TI);
    k = b.ks + (int)tk * (BS / TI) + row;
    j = b.js + (int)jj;
    return i <= b.ie && k <= b.ke;
}
// Stage the tile's metric and floors (x1 clamped to the table for a ragged
// last tile) into sg; every thread of the work-group calls it.
template <int TI>
__device__ __forceinline__ void stage_recover_geometry(const GeomDev& gd, int j, int i0, double* sg) {
    const int lane = (int)(threadIdx.x % TI);
    const int w = (TI == 64) ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x / TI)) : (int)(threadIdx.x / TI);
    const long q = (long)j * gd.gn1 + min(i0 + lane, gd.gn1 - 1);
#pragma unroll
    for (int c0 = 0; c0 < STG_RECOVER; c0 += BS / TI) {
        const int comp = c0 + w;
        if (comp < STG_RECOVER)
            sg[comp * TI + lane] = comp < GEO_NCOMP ? gd.g[GLOC_CENTER][(long)comp * gd.np + q]
                                                    : gd.flr[(long)(comp - GEO_NCOMP) * gd.np + q];
    }
    __syncthreads();
}

// ND: dimensionality (flux-CT for 2 and 3); SAME: stage with gam0 = 0 and
// U_sub == U_full, which reads the state once (0*U + 1*U == U).
template <typename R, bool DO_DT, int ND, bool SAME, bool PK>
__global__ void __launch_bounds__(BS)
combine_kernel(BlockDims b, StepConsts cst, GeomDev gd, int stage, const R* __restrict__ F, const R* __restrict__ ctop,
               const double* __restrict__ Usub, const double* Ufull, const double* __restrict__ Psub, double* Uout,
               const DeviceState* __restrict__ ds, double* __restrict__ dt_part, const BlockEntry* __restrict__ pk) {
    if constexpr (PK) {
        const long blk = pack_select(pk, b, gd);
        const long sv = blk * NV * b.nc;
        F = pack_at(F, 3 * sv);
        ctop = pack_at(ctop, blk * 3 * b.nc);
        Usub = pack_at(Usub, sv); Ufull = pack_at(Ufull, sv); Psub = pack_at(Psub, sv); Uout = pack_at(Uout, sv);
        ds += blk;
        dt_part = pack_at(dt_part, blk * gridDim.x);
    }
    const unsigned t = xcd_logical_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool active = t < (unsigned)ninterior(b);
    double ndt = DBL_MAX;
    if (active) {
        int k, j, i;
        interior_index_jki(b, t, k, j, i);
        const long c = (long)k * b.s3 + (long)j * b.s2 + i;
        const double g0 = cst.gam0[stage - 1], g1 = cst.gam1[stage - 1];
        // The geometric source (flux.cpp:395-444) of the stage's primitives.  The
        // FP32-flux form computes it first, before the state and the fluxes are
        // loaded (100 instead of 174-190 VGPRs, 1-2 % faster); the FP64 form keeps
        // it after the divergence (first measured 2 % slower there).  Both add
        // it to the divergence in the same order, so the results are the same.
        constexpr bool SRC_FIRST = sizeof(R) == 4;
        auto source = [&](double acc[NV]) {
            double Ps[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) Ps[v] = Psub[v * b.nc + c];
            const long q = (long)j * gd.gn1 + i;
            CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, q);
            add_geo_source(g, gd, q, Ps, cst.gam, acc);
        };
        double dS[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
        if constexpr (SRC_FIRST) source(dS);
        double Ub[NV];   // the stage's base state, loaded early so its latency overlaps the rest
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double uf = Ufull[v * b.nc + c];
            Ub[v] = SAME ? uf : g0 * Usub[v * b.nc + c] + g1 * uf;
        }
        double dudt[NV];
        if constexpr (ND > 1) zone_divergence_ct<R, ND>(b, F, c, dudt);
        else zone_divergence<R>(b, flux_vars_of(F, b.nc), nullptr, c, false, dudt);
        if constexpr (SRC_FIRST) {
#pragma unroll
            for (int v = UU; v <= U3; ++v) dudt[v] += dS[v];
        } else {
            source(dudt);
        }
        const double beta_dt = cst.beta[stage - 1] * ds->dt;
#pragma unroll
        for (int v = 0; v < NV; ++v) Uout[v * b.nc + c] = 1.0 * Ub[v] + beta_dt * dudt[v];
        if (DO_DT) {
            const double n = zone_ndt(b, (double)ctop[c], ND > 1 ? (double)ctop[b.nc + c] : 0.,
                                      ND > 2 ? (double)ctop[2 * b.nc + c] : 0.);
            if (!(n != n)) ndt = n;
        }
    }
    if (DO_DT) {
        const double m = block_min(ndt);
        if (threadIdx.x == 0) dt_part[blockIdx.x] = m;
    }
}

// U->P with the stage's starting primitives as guess (kharma_driver.cpp:379-384),
// B from U, floors in frame FRAME, P->U; U is read from and written back to Uout.
// One zone (k, j, i) at c, whose U and guess P the caller has loaded.
template <int FRAME, int INV, bool STG, int TI>
__device__ __forceinline__ void recover_zone(const BlockDims& b, const StepConsts& cst, const GeomDev& gd, int stage,
                                             double* Uout, double* __restrict__ Pout,
                                             float* __restrict__ Pout32, int* __restrict__ pflag,
                                             int* __restrict__ fflag, int* __restrict__ fail_list,
                                             DeviceState* __restrict__ ds, double* __restrict__ Ustar,
                                             const double* sg, int k, int j, int i, long c, double (&U)[NV],
                                             double (&P)[NV]) {
    const double gam = cst.gam;
    int pf = 0, ff = 0;
    double rf, uf;
    const int lane = (int)(threadIdx.x % TI);
    if (Ustar) {   // the stage's U* for the differentiation form's record
#pragma unroll
        for (int v = 0; v < NV; ++v) Ustar[v * b.nc + c] = U[v];
    }
    if constexpr (STG) {
        rf = sg[GEO_NCOMP * TI + lane];
        uf = sg[(GEO_NCOMP + 1) * TI + lane];
    } else {
        geom_floors<1>(cst, gd, j, i, rf, uf);
    }
    {
        CachedGeom<double> g(STG ? sg : gd.g[GLOC_CENTER], STG ? TI : gd.np, STG ? lane : (long)j * gd.gn1 + i);
        pf = invert_zone<INV>(g, U, P, cst);
        const int ff0 = (pf / F_GEOM_RHO) * F_GEOM_RHO;
        pf = pf % F_GEOM_RHO;
        const double gd_ = g.gdet();
        const double igd = rcp(gd_);
        P[B1] = U[B1] * igd; P[B2] = U[B2] * igd; P[B3] = U[B3] * igd;
        FourVec<double> D;
        calc_4vecs<double>(g, P, D);
        ff = ff0;
        if (cst.floors.on) ff = floors_zone<FRAME>(g, P, U, D, gam, cst.floors, rf, uf, ff0, pf);
        p_to_u_D(g, P, D, gam, U);
    }
    ImageSet s;
    image_set(b, k, j, i, s);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        store_with_images(b, s, Uout + v * b.nc, c, U[v], NV * b.nc);
        store_with_images(b, s, Pout + v * b.nc, c, P[v], NV * b.nc);
        if (Pout32) store_with_images(b, s, Pout32 + v * b.nc, c, (float)P[v], NV * b.nc);
    }
    store_with_images(b, s, pflag, c, pf, b.nc);
    fflag[c] = ff;
    if (pf > 0) {
        const int slot = atomicAdd(&ds->fail_count[stage - 1], 1);
        fail_list[slot] = (int)c;
    }
}

template <int FRAME, int INV, bool STG, bool PK, int TI = 64>
__global__ void __launch_bounds__(BS, (RecOcc<FRAME, STG, TI>::w))
recover_kernel(BlockDims b, StepConsts cst, GeomDev gd, int stage, const double* __restrict__ Psub, double* Uout,
               double* __restrict__ Pout, float* __restrict__ Pout32, int* __restrict__ pflag,
               int* __restrict__ fflag, int* __restrict__ fail_list, DeviceState* __restrict__ ds,
               double* __restrict__ Ustar, const BlockEntry* __restrict__ pk) {
    __shared__ double sg[STG ? STG_RECOVER * TI : 1];
    if constexpr (PK) {
        const long blk = pack_select(pk, b, gd);
        const long sv = blk * NV * b.nc;
        Psub = pack_at(Psub, sv); Uout = pack_at(Uout, sv); Pout = pack_at(Pout, sv); Pout32 = pack_at(Pout32, sv);
        Ustar = pack_at(Ustar, sv);
        pflag = pack_at(pflag, blk * b.nc); fflag = pack_at(fflag, blk * b.nc); fail_list = pack_at(fail_list, blk * b.nc);
        ds += blk;
    }
    constexpr int NP = 5;   // the guess: the fluid primitives (B follows from U)
    int k, j, i;
    long c;
    double U[NV], P[NV];
    if constexpr (STG) {
        // the zone's U and P are requested before the tile's geometry is staged,
        // so that their latency overlaps the staging and its barrier
        int i0;
        const bool active = tile_zone<TI>(b, k, j, i, i0);
        c = (long)k * b.s3 + (long)j * b.s2 + i;
        if (active) {
#pragma unroll
            for (int v = 0; v < NV; ++v) U[v] = Uout[v * b.nc + c];
#pragma unroll
            for (int v = 0; v < NP; ++v) P[v] = Psub[v * b.nc + c];
        }
        stage_recover_geometry<TI>(gd, j, i0, sg);
        if (!active) return;
    } else {
        const unsigned t = xcd_logical_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
        if (t >= (unsigned)ninterior(b)) return;
        interior_index_jki(b, t, k, j, i);
        c = (long)k * b.s3 + (long)j * b.s2 + i;
#pragma unroll
        for (int v = 0; v < NV; ++v) U[v] = Uout[v * b.nc + c];
#pragma unroll
        for (int v = 0; v < NP; ++v) P[v] = Psub[v * b.nc + c];
    }
    recover_zone<FRAME, INV, STG, TI>(b, cst, gd, stage, Uout, Pout, Pout32, pflag, fflag, fail_list, ds, Ustar, sg, k, j,
                                  i, c, U, P);
}

// Inverter::FixUtoP (inverter/fixup.cpp:47-156) on the zones listed by the
// update kernel: distance-weighted average of non-failed neighbours, atmosphere
// fallback, geometric floors, then the conserved state of the final P->U.
template <int GEO, bool PK>
__global__ void __launch_bounds__(BS)
fixup_list_kernel(BlockDims b, StepConsts cst, GeomDev gd, int stage, double* __restrict__ U, double* __restrict__ P,
                  float* __restrict__ P32, const int* __restrict__ pflag, const int* __restrict__ fail_list,
                  const DeviceState* __restrict__ ds, const BlockEntry* __restrict__ pk) {
    if constexpr (PK) {
        const long blk = pack_select(pk, b, gd);
        const long sv = blk * NV * b.nc;
        U = pack_at(U, sv); P = pack_at(P, sv); P32 = pack_at(P32, sv);
        pflag = pack_at(pflag, blk * b.nc); fail_list = pack_at(fail_list, blk * b.nc);
        ds += blk;
    }
    const int n = ds->fail_count[stage - 1];
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const long c = fail_list[q];
        const int k = (int)(c / b.s3), j = (int)((c - (long)k * b.s3) / b.s2), i = (int)(c - (long)k * b.s3 - (long)j * b.s2);
        double Pz[NV];
        for (int v = 0; v < NV; ++v) Pz[v] = P[v * b.nc + c];
        double wsum = 0., sum[5] = {0., 0., 0., 0., 0.};
        if (cst.fix_average) {
            for (int dn = -1; dn <= 1; dn++) for (int dm = -1; dm <= 1; dm++) for (int dl = -1; dl <= 1; dl++) {
                const int ii = i + dl, jj = j + dm, kk = k + dn;
                if (!in_physical(b, kk, jj, ii)) continue;
                const long cc = (long)kk * b.s3 + (long)jj * b.s2 + ii;
                if (!(pflag[cc] > 0)) {
                    const double w = 1. / (abs(dl) + abs(dm) + abs(dn) + 1);
                    wsum += w;
                    for (int p = 0; p < 5; ++p) sum[p] += w * P[p * b.nc + cc];
                }
            }
        }
        if (wsum < 1.e-10) { for (int p = 0; p < 5; ++p) Pz[p] = 0.; }
        else { for (int p = 0; p < 5; ++p) Pz[p] = sum[p] / wsum; }
        double rf, uf;
        geom_floors<GEO>(cst, gd, j, i, rf, uf);
        apply_geo_floors<double>(Pz, rf, uf);
        double Uz[NV];
        with_center_geom<GEO>(gd, j, i, [&](const auto& g) { p_to_u(g, Pz, cst.gam, Uz); });
        ImageSet s;
        image_set(b, k, j, i, s);
        for (int v = 0; v < 5; ++v) {
            store_with_images(b, s, P + v * b.nc, c, Pz[v], NV * b.nc);
            if (P32) store_with_images(b, s, P32 + v * b.nc, c, (float)Pz[v], NV * b.nc);
        }
        for (int v = 0; v < NV; ++v) store_with_images(b, s, U + v * b.nc, c, Uz[v], NV * b.nc);
    }
}

// End-of-step scalar work on the device: the CFL minimum and divB maximum of
// the per-block partials, then SetGlobalTimeStep (kharma_driver.cpp:439-460)
// with GRMHD::EstimateTimestep's clip (grmhd.cpp:335-339), time advance, resets.
// ds holds nds states, one per block of a mesh: the first carries the mesh's
// clock and totals, which are then copied to the others; every block's
// failed-inversion counts are added up and cleared.
__global__ void __launch_bounds__(1024)
finalize_kernel(StepConsts cst, DeviceState* ds, int nds, const double* __restrict__ dt_part, int n_dt,
                const double* __restrict__ divb_part, int n_divb, double* dt_hist, int hist_cap) {
    double mn0 = DBL_MAX, mn1 = DBL_MAX, mx0 = 0., mx1 = 0.;
    const int st = blockDim.x;
#pragma unroll 4
    for (int q = threadIdx.x; q < n_dt; q += 2 * st) {
        mn0 = fmin(mn0, dt_part[q]);
        if (q + st < n_dt) mn1 = fmin(mn1, dt_part[q + st]);
    }
#pragma unroll 4
    for (int q = threadIdx.x; q < n_divb; q += 2 * st) {
        mx0 = fmax(mx0, divb_part[q]);
        if (q + st < n_divb) mx1 = fmax(mx1, divb_part[q + st]);
    }
    const double mn = block_min(fmin(mn0, mn1));
    const double mx = block_max(fmax(mx0, mx1));
    if (threadIdx.x != 0) return;
    const double min_ndt = mn;
    const double dt = ds->dt;
    const double dt_max = cst.max_dt_increase * dt;
    const double new_dt = clip(min_ndt * cst.cfl, cst.dt_min, dt_max);
    if (dt_hist && ds->ncycle < hist_cap) dt_hist[ds->ncycle] = dt;
    ds->time += dt;
    double nd = dt;
    if (nd < 0.1 * DBL_MAX) nd *= 2.0;
    nd = smin(nd, new_dt);
    if (ds->time < cst.tlim && (cst.tlim - ds->time) < nd) nd = cst.tlim - ds->time;
    ds->dt_last = dt;
    ds->dt = nd;
    ds->last_ndt = min_ndt;
    ds->ncycle += 1;
    for (int q = 0; q < nds; ++q) {
        ds->total_fails += ds[q].fail_count[0] + ds[q].fail_count[1];
        ds[q].fail_count[0] = 0;
        ds[q].fail_count[1] = 0;
    }
    ds->last_divb = mx;
    if (cst.kill_divb && mx > cst.divb_over && ds->divb_over_cycle == 0) {
        ds->divb_over_cycle = ds->ncycle;
        ds->divb_over_value = mx;
    }
    for (int q = 1; q < nds; ++q) {
        ds[q].dt = ds->dt; ds[q].dt_last = ds->dt_last; ds[q].time = ds->time;
        ds[q].last_divb = ds->last_divb; ds[q].last_ndt = ds->last_ndt;
        ds[q].ncycle = ds->ncycle; ds[q].total_fails = ds->total_fails;
        ds[q].divb_over_cycle = ds->divb_over_cycle; ds[q].divb_over_value = ds->divb_over_value;
    }
}

// One-block reduction of per-block partials into *out (min or max).
__global__ void __launch_bounds__(BS) reduce_partials_kernel(const double* __restrict__ part, int n, int is_max, double* out) {
    double v = is_max ? 0. : DBL_MAX;
    for (int q = threadIdx.x; q < n; q += blockDim.x) v = is_max ? fmax(v, part[q]) : fmin(v, part[q]);
    v = is_max ? block_max(v) : block_min(v);
    if (threadIdx.x == 0) *out = v;
}

__global__ void reset_state_kernel(DeviceState* ds, double dt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ds->dt = dt; ds->dt_last = dt; ds->time = 0.;
    ds->last_divb = 0.; ds->last_ndt = 0.;
    ds->fail_count[0] = ds->fail_count[1] = 0;
    ds->ncycle = 0; ds->total_fails = 0;
    ds->divb_over_cycle = 0; ds->divb_over_value = 0.;
}

__global__ void set_clock_kernel(DeviceState* ds, double time, double dt, int ncycle) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ds->time = time; ds->dt = dt; ds->dt_last = dt; ds->ncycle = ncycle;
}

// B_FluxCT::MaxDivB with corner_div (b_flux_ct.cpp:566-607, b_flux_ct_functions.hpp:52-75)
// Corners i in [is, ie+1]; ValidDivBX1 (b_flux_ct.cpp:554-564) skips the
// corners on a physical X1 face (i = is inner, i = ie + 1 outer).
template <bool PK>
__global__ void __launch_bounds__(BS) max_divb_kernel(BlockDims b, Vars<const double> U, double* __restrict__ part,
                                                      const BlockEntry* __restrict__ pk) {
    if constexpr (PK) {
        const long blk = pack_select(pk, b);
        U = pack_at(U, blk * NV * b.nc);
        part = pack_at(part, blk * gridDim.x);
    }
    const int i0 = b.img[0][0] ? b.is : b.is + 1;
    const int ni = (b.img[0][1] ? b.ie + 1 : b.ie) - i0 + 1, nj = b.nx2 + 1, nk = (b.ndim > 2) ? b.nx3 + 1 : 1;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.;
    if (t < (unsigned)(ni * nj * nk)) {
        const int kk = t / (ni * nj), rem = t - kk * (ni * nj), jj = rem / ni;
        const int i = i0 + rem - jj * ni, j = b.js + jj, k = b.ks + kk;
        const long c = (long)k * b.s3 + (long)j * b.s2 + i;
        const double* __restrict__ Bx = U[B1];
        const double* __restrict__ By = U[B2];
        const double* __restrict__ Bz = U[B3];
        const long s2 = b.s2, s3 = b.s3;
        const bool do3 = b.ndim > 2;
        const double norm = do3 ? 0.25 : 0.5;
        double term1 = Bx[c] - Bx[c - 1] + Bx[c - s2] - Bx[c - s2 - 1];
        double term2 = By[c] - By[c - s2] + By[c - 1] - By[c - s2 - 1];
        double term3 = 0.;
        if (do3) {
            term1 += Bx[c - s3] + Bx[c - s3 - s2] - Bx[c - s3 - 1] - Bx[c - s3 - s2 - 1];
            term2 += By[c - s3] + By[c - s3 - 1] - By[c - s3 - s2] - By[c - s3 - s2 - 1];
            term3 = Bz[c] + Bz[c - s2] + Bz[c - 1] + Bz[c - s2 - 1]
                  - Bz[c - s3] - Bz[c - s3 - s2] - Bz[c - s3 - 1] - Bz[c - s3 - s2 - 1];
        }
        d = fabs(norm * term1 / b.dx1 + norm * term2 / b.dx2 + norm * term3 / b.dx3);
    }
    const double m = block_max(d);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// ------------------------------------------------- physical boundaries
// Flux fixes of physical faces on the Riemann fluxes (B_FluxCT::ZeroBoundaryFlux,
// b_flux_ct.cpp:372-420; KBoundaries::FixFlux, boundaries.cpp:611-690).  One
// thread per face of the fixed rows; the two kinds of rows are disjoint.
template <typename R>
__device__ __forceinline__ void fix_flux_zone(const BlockDims& b, const FluxVars<R>& F, long off) {
    const int nd = b.ndim;
    const int ni = b.nx1 + 2, nk = (nd > 2) ? b.nx3 + 2 : 1, nj = (nd > 1) ? b.nx2 + 2 : 1;
    const int i0 = b.is - 1, k0 = (nd > 2) ? b.ks - 1 : b.ks, j0 = (nd > 1) ? b.js - 1 : b.js;
    const Vars<R>& F1 = F[0];
    const Vars<R>& F2 = F[1];
    const Vars<R>& F3 = F[2];
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool pol[2] = {nd > 1 && !b.img[1][0], nd > 1 && !b.img[1][1]};
    const unsigned npolar = (pol[0] || pol[1]) ? 2u * ni * nk : 0u;
    if (t < npolar) {
        // polar faces j = js (side 0) and j = je + 1 (side 1), i in [is-1, ie+1], k in [ks-1, ke+1]
        const int side = t / (ni * nk), rem = t - side * (ni * nk);
        if (!pol[side]) return;
        const int k = k0 + rem / ni, i = i0 + rem % ni;
        const int j = side == 0 ? b.js : b.je + 1;
        const long c = (long)k * b.s3 + (long)j * b.s2 + i + off;
        const long cg = side == 0 ? c - b.s2 : c;        // the face row that is mirrored into
        const long cs = side == 0 ? c : c - b.s2;        // and its source
        if (b.fix_polar_flux) {
            F2[B1][c] = R(0);
            F2[B3][c] = R(0);
            F1[B2][cg] = -F1[B2][cs];
            if (nd > 2) F3[B2][cg] = -F3[B2][cs];
        }
        if (b.zero_polar_flux)
            for (int v = 0; v < NV; ++v) F2[v][c] = R(0);
        return;
    }
    const unsigned t2 = t - npolar;
    if ((!b.img[0][0] || !b.img[0][1]) && t2 < 2u * nj * nk) {
        // X1 faces i = is (side 0) and i = ie + 1 (side 1), j in [js-1, je+1], k in [ks-1, ke+1]
        const int side = t2 / (nj * nk), rem = t2 - side * (nj * nk);
        const int k = k0 + rem / nj, j = j0 + rem % nj;
        if (!b.check_inflow[side]) return;
        const int i = side == 0 ? b.is : b.ie + 1;
        const long c = (long)k * b.s3 + (long)j * b.s2 + i + off;
        const R f = F1[RHO][c];
        F1[RHO][c] = side == 0 ? smin(f, R(0)) : smax(f, R(0));
    }
}
// off: this block's part of the pack's fluxes, as an index offset; the block's
// dims are read in place from the pack's table.
template <typename R, bool PK>
__global__ void __launch_bounds__(BS) fix_flux_kernel(BlockDims b, FluxVars<R> F, const BlockEntry* __restrict__ pk) {
    if constexpr (PK) {
        const BlockDims& bb = pk[blockIdx.z].dims;
        fix_flux_zone<R>(bb, F, (long)blockIdx.z * 3 * NV * bb.nc);
    } else {
        fix_flux_zone<R>(b, F, 0);
    }
}

// One ghost zone of a physical X1 or X2 face, written from the final interior
// state the way KBoundaries::ApplyBoundary leaves it (boundaries.cpp:369-589):
// Parthenon's outflow (copy of the first / last interior zone) or reflecting
// (mirror copy with the face-normal vector components negated) fill of the 8
// conserved and 5 fluid primitive variables, check_inflow on X1 faces, then
// BoundaryPtoUElseUtoP (kharma_package.cpp:125-156): B^i = U.B / sqrt(-g) and the
// 5 fluid conserved variables from the primitives.  Faces apply in the order
// inner/outer x1, inner/outer x2, then inner x1 again when fix_corner_inner
// (boundaries.cpp:556-570); corner zones compose the two fills accordingly.
// With b.prim_fill (driver "simple", boundaries.cpp:561-571) the fill copies
// the primitives and U.B, then sets B^i = U.B / sqrt(-g) and all 8 conserved
// variables from the primitives, so every filled zone -- a corner's
// intermediate zone included -- holds U.B = (U.B_src / sqrt(-g)) sqrt(-g).
template <int GEO>
__device__ __forceinline__ void bc_fill_x1(const BlockDims& b, const GeomDev& gd, const Vars<double>& Usrc,
                                           const Vars<double>& Psrc, int k, int j, int i, double P[NV], double Bu[3]) {
    const bool inner = i < b.is;
    const int type = b.bc[inner ? 0 : 1];
    const int si = (type == KH_BC_OUTFLOW) ? (inner ? b.is : b.ie) : (inner ? 2 * b.is - 1 - i : 2 * b.ie + 1 - i);
    const long sc = (long)k * b.s3 + (long)j * b.s2 + si;
    for (int v = 0; v < 5; ++v) P[v] = Psrc[v][sc];
    for (int v = 0; v < 3; ++v) Bu[v] = Usrc[B1 + v][sc];
    if (type == KH_BC_REFLECT) { P[U1] = -P[U1]; Bu[0] = -Bu[0]; }
    if (b.check_inflow[inner ? 0 : 1]) {
        with_center_geom<GEO>(gd, j, i, [&](const auto& g) { check_inflow(g, P, inner); });
    }
}
template <int GEO>
__device__ __forceinline__ void bc_fill_x2(const BlockDims& b, const Vars<double>& Usrc, const Vars<double>& Psrc,
                                           int k, int j, int i, double P[NV], double Bu[3]) {
    const bool inner = j < b.js;
    const int type = b.bc[inner ? 2 : 3];
    const int sj = (type == KH_BC_OUTFLOW) ? (inner ? b.js : b.je) : (inner ? 2 * b.js - 1 - j : 2 * b.je + 1 - j);
    const long sc = (long)k * b.s3 + (long)sj * b.s2 + i;
    for (int v = 0; v < 5; ++v) P[v] = Psrc[v][sc];
    for (int v = 0; v < 3; ++v) Bu[v] = Usrc[B1 + v][sc];
    if (type == KH_BC_REFLECT) { P[U2] = -P[U2]; Bu[1] = -Bu[1]; }
}
// The zones the physical boundaries fill, enumerated without the interior:
// slab 1 is every zone of the ng X1 ghost columns (all j, k) of each physical
// X1 face; slab 2 is the ng X2 ghost rows of each physical X2 face over the X1
// interior, widened to the ghost columns of X1 faces that are not physical.
__host__ __device__ inline int nphys(const BlockDims& b, int d) {
    return (d < b.ndim) ? (!b.img[d][0]) + (!b.img[d][1]) : 0;
}
__host__ __device__ inline long bounds_slab1(const BlockDims& b) {
    return (long)b.n3 * b.n2 * b.ng * nphys(b, 0);
}
__host__ __device__ inline int bounds_slab2_width(const BlockDims& b) {
    return (b.img[0][1] ? b.n1 - 1 : b.ie) - (b.img[0][0] ? 0 : b.is) + 1;
}
__host__ __device__ inline long bounds_slab2(const BlockDims& b) {
    return (long)b.n3 * b.ng * nphys(b, 1) * bounds_slab2_width(b);
}
template <int GEO, bool PK>
__global__ void __launch_bounds__(BS) bounds_kernel(BlockDims b, StepConsts cst, GeomDev gd, Vars<double> U,
                                                    Vars<double> P, float* __restrict__ P32,
                                                    const BlockEntry* __restrict__ pk) {
    if constexpr (PK) {
        const long blk = pack_select(pk, b, gd);
        if (!has_physical_face(b)) return;
        const long sv = blk * NV * b.nc;
        U = pack_at(U, sv); P = pack_at(P, sv); P32 = pack_at(P32, sv);
    }
    // slab sizes stay below 2^31 (the C ABI bounds a block's size): 32-bit division
    unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned n1s = (unsigned)bounds_slab1(b);
    int k, j, i;
    if (r < n1s) {
        const unsigned m = (unsigned)(b.ng * nphys(b, 0));
        const int g = (int)(r % m);
        r /= m;
        j = (int)(r % (unsigned)b.n2);
        k = (int)(r / (unsigned)b.n2);
        i = (!b.img[0][0] && g < b.ng) ? g : b.ie + 1 + (g - (b.img[0][0] ? 0 : b.ng));
    } else {
        r -= n1s;
        if (r >= (unsigned)bounds_slab2(b)) return;
        const unsigned w = (unsigned)bounds_slab2_width(b);
        const int ii = (int)(r % w);
        r /= w;
        const unsigned m = (unsigned)(b.ng * nphys(b, 1));
        const int g = (int)(r % m);
        k = (int)(r / m);
        j = (!b.img[1][0] && g < b.ng) ? g : b.je + 1 + (g - (b.img[1][0] ? 0 : b.ng));
        i = (b.img[0][0] ? 0 : b.is) + ii;
    }
    const long t = (long)k * b.s3 + (long)j * b.s2 + i;
    const bool x1g = (i < b.is && !b.img[0][0]) || (i > b.ie && !b.img[0][1]);
    const bool x2g = b.ndim > 1 && ((j < b.js && !b.img[1][0]) || (j > b.je && !b.img[1][1]));
    double Pz[NV], Bu[3];
    // U.B that a filled zone (jj, ii) ends with, under b.prim_fill
    auto settle = [&](double B[3], int jj, int ii) {
        if (!b.prim_fill) return;
        with_center_geom<GEO>(gd, jj, ii, [&](const auto& g) {
            const double gdet = g.gdet();
            for (int v = 0; v < 3; ++v) B[v] = (B[v] / gdet) * gdet;
        });
    };
    if (x1g && !x2g) {
        bc_fill_x1<GEO>(b, gd, U, P, k, j, i, Pz, Bu);
    } else if (x2g && !x1g) {
        bc_fill_x2<GEO>(b, U, P, k, j, i, Pz, Bu);
    } else if (i < b.is && b.fix_corner_inner) {
        // X2 fill of column is, then the inner-X1 fill over it
        double Pm[NV], Bm[3];
        bc_fill_x2<GEO>(b, U, P, k, j, b.is, Pm, Bm);
        settle(Bm, j, b.is);
        const int type = b.bc[0];
        for (int v = 0; v < 5; ++v) Pz[v] = Pm[v];
        for (int v = 0; v < 3; ++v) Bu[v] = Bm[v];
        if (type == KH_BC_REFLECT) {
            // a reflecting X1 face would read its mirror column instead
            const int si = 2 * b.is - 1 - i;
            bc_fill_x2<GEO>(b, U, P, k, j, si, Pz, Bu);
            settle(Bu, j, si);
            Pz[U1] = -Pz[U1]; Bu[0] = -Bu[0];
        }
        if (b.check_inflow[0]) with_center_geom<GEO>(gd, j, i, [&](const auto& g) { check_inflow(g, Pz, true); });
    } else {
        // X1 fill of the mirror / edge row, then the X2 fill over it
        const bool inner = j < b.js;
        const int type = b.bc[inner ? 2 : 3];
        const int sj = (type == KH_BC_OUTFLOW) ? (inner ? b.js : b.je) : (inner ? 2 * b.js - 1 - j : 2 * b.je + 1 - j);
        bc_fill_x1<GEO>(b, gd, U, P, k, sj, i, Pz, Bu);
        settle(Bu, sj, i);
        if (type == KH_BC_REFLECT) { Pz[U2] = -Pz[U2]; Bu[1] = -Bu[1]; }
    }
    double Uz[NV];
    with_center_geom<GEO>(gd, j, i, [&](const auto& g) {
        const double gdet = g.gdet();
        Pz[B1] = Bu[0] / gdet; Pz[B2] = Bu[1] / gdet; Pz[B3] = Bu[2] / gdet;
        p_to_u(g, Pz, cst.gam, Uz);
    });
    for (int v = 0; v < 5; ++v) U[v][t] = Uz[v];
    for (int v = 0; v < 3; ++v) U[B1 + v][t] = b.prim_fill ? Uz[B1 + v] : Bu[v];
    for (int v = 0; v < NV; ++v) P[v][t] = Pz[v];
    if (P32) for (int v = 0; v < NV; ++v) P32[v * b.nc + t] = (float)Pz[v];
}

// ----------------------------------------------------- task-level kernels
template <typename R>
__global__ void __launch_bounds__(BS) divergence_kernel(BlockDims b, FluxVars<const R> F, Vars<double> dUdt) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)ninterior(b)) return;
    int k, j, i;
    interior_index(b, t, k, j, i);
    const long c = (long)k * b.s3 + (long)j * b.s2 + i;
    double d[NV];
    zone_divergence<R>(b, F, nullptr, c, false, d);
    for (int v = 0; v < NV; ++v) dUdt[v][c] = d[v];
}

// Flux::AddGeoSource (flux/flux.cpp:395-444) over the interior, from the primitives P
__global__ void __launch_bounds__(BS) geo_source_kernel(BlockDims b, StepConsts cst, GeomDev gd, Vars<const double> P,
                                                        Vars<double> dUdt) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)ninterior(b)) return;
    int k, j, i;
    interior_index(b, t, k, j, i);
    const long c = (long)k * b.s3 + (long)j * b.s2 + i;
    double Pz[NV], d[NV];
    for (int v = 0; v < NV; ++v) { Pz[v] = P[v][c]; d[v] = dUdt[v][c]; }
    const long q = (long)j * gd.gn1 + i;
    CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, q);
    add_geo_source(g, gd, q, Pz, cst.gam, d);
    for (int v = UU; v <= U3; ++v) dUdt[v][c] = d[v];
}

// parthenon::Update::WeightedSumData pair used by AddStateUpdate, entire block;
// variable blockIdx.y.  Uout may be Usub or Ufull: AddStateUpdate updates in
// place (kharma_driver.cpp:366-370), and each element is read before it is written.
__global__ void __launch_bounds__(BS) weighted_sum_kernel(long n, Vars<const double> Usub, Vars<const double> Ufull,
                                                          Vars<const double> dUdt, double g0, double g1, double bdt,
                                                          Vars<double> Uout) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (t >= n) return;
    const double u = g0 * Usub[v][t] + g1 * Ufull[v][t];
    Uout[v][t] = 1.0 * u + bdt * dUdt[v][t];
}

__device__ __forceinline__ void zone_ijk(const BlockDims& b, long c, int& k, int& j, int& i) {
    k = (int)(c / b.s3);
    j = (int)((c - (long)k * b.s3) / b.s2);
    i = (int)(c - (long)k * b.s3 - (long)j * b.s2);
}

// Inverter::BlockUtoP over the physical range + B_FluxCT::BlockUtoP over the entire block
template <int GEO, int INV = 0>
__global__ void __launch_bounds__(BS) utop_kernel(BlockDims b, StepConsts cst, GeomDev gd, Vars<const double> U,
                                                  Vars<double> P, int* __restrict__ pflag, int* __restrict__ fflag) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc) return;
    int k, j, i;
    zone_ijk(b, c, k, j, i);
    with_center_geom<GEO>(gd, j, i, [&](const auto& g) {
        if (in_physical(b, k, j, i)) {
            double Uz[NV], Pz[NV];
            for (int v = 0; v < NV; ++v) { Uz[v] = U[v][c]; Pz[v] = P[v][c]; }
            const int r = invert_zone<INV>(g, Uz, Pz, cst);
            for (int v = 0; v < 5; ++v) P[v][c] = Pz[v];
            pflag[c] = r % F_GEOM_RHO;
            fflag[c] = (r / F_GEOM_RHO) * F_GEOM_RHO;
        }
        const double gdet = g.gdet();
        for (int v = 0; v < 3; ++v) P[B1 + v][c] = U[B1 + v][c] / gdet;
    });
}

// Floors::ApplyGRMHDFloors over the entire block: floors, ceilings and the
// 5 hydrodynamic conserved variables of modified zones (p_to_u_mhd).
template <int GEO>
__global__ void __launch_bounds__(BS) floors_kernel(BlockDims b, StepConsts cst, GeomDev gd, Vars<double> U,
                                                    Vars<double> P, int* __restrict__ pflag, int* __restrict__ fflag) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc) return;
    int k, j, i;
    zone_ijk(b, c, k, j, i);
    double Uz[NV], Pz[NV];
    for (int v = 0; v < NV; ++v) { Uz[v] = U[v][c]; Pz[v] = P[v][c]; }
    double rf, uf;
    geom_floors<GEO>(cst, gd, j, i, rf, uf);
    with_center_geom<GEO>(gd, j, i, [&](const auto& g) {
        int pf = pflag[c];
        FourVec<double> D;
        calc_4vecs<double>(g, Pz, D);
        const int ff = floors_zone(g, Pz, Uz, D, cst.gam, cst.floors, rf, uf, fflag[c], pf);
        fflag[c] = ff;
        pflag[c] = pf;
        if (ff) {
            double Un[NV];
            p_to_u_D(g, Pz, D, cst.gam, Un);
            for (int v = 0; v < 5; ++v) { P[v][c] = Pz[v]; U[v][c] = Un[v]; }
        }
    });
}

// Inverter::FixUtoP over the physical range; two passes so averages only read
// non-failed zones.
__global__ void __launch_bounds__(BS) fixup_avg_kernel(BlockDims b, StepConsts cst, Vars<double> P,
                                                       const int* __restrict__ pflag) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc || !(pflag[c] > 0)) return;
    int k, j, i;
    zone_ijk(b, c, k, j, i);
    if (!in_physical(b, k, j, i)) return;
    double wsum = 0., sum[5] = {0., 0., 0., 0., 0.};
    if (cst.fix_average) {
        for (int dn = -1; dn <= 1; dn++) for (int dm = -1; dm <= 1; dm++) for (int dl = -1; dl <= 1; dl++) {
            const int ii = i + dl, jj = j + dm, kk = k + dn;
            if (!in_physical(b, kk, jj, ii)) continue;
            const long cc = (long)kk * b.s3 + (long)jj * b.s2 + ii;
            if (!(pflag[cc] > 0)) {
                const double w = 1. / (abs(dl) + abs(dm) + abs(dn) + 1);
                wsum += w;
                for (int p = 0; p < 5; ++p) sum[p] += w * P[p][cc];
            }
        }
    }
    for (int p = 0; p < 5; ++p) P[p][c] = (wsum < 1.e-10) ? 0. : sum[p] / wsum;
}
template <int GEO>
__global__ void __launch_bounds__(BS) fixup_floor_kernel(BlockDims b, StepConsts cst, GeomDev gd, Vars<double> U,
                                                         Vars<double> P, const int* __restrict__ pflag) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc || !(pflag[c] > 0)) return;
    int k, j, i;
    zone_ijk(b, c, k, j, i);
    if (!in_physical(b, k, j, i)) return;
    double Pz[NV], Uz[NV];
    for (int v = 0; v < NV; ++v) Pz[v] = P[v][c];
    double rf, uf;
    geom_floors<GEO>(cst, gd, j, i, rf, uf);
    apply_geo_floors<double>(Pz, rf, uf);
    with_center_geom<GEO>(gd, j, i, [&](const auto& g) { p_to_u(g, Pz, cst.gam, Uz); });
    for (int v = 0; v < 5; ++v) { P[v][c] = Pz[v]; U[v][c] = Uz[v]; }
}

// Flux::MeshPtoU over the entire block
template <int GEO>
__global__ void __launch_bounds__(BS) ptou_kernel(BlockDims b, StepConsts cst, GeomDev gd, Vars<const double> P,
                                                  Vars<double> U) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc) return;
    int k, j, i;
    zone_ijk(b, c, k, j, i);
    double Pz[NV], Uz[NV];
    for (int v = 0; v < NV; ++v) Pz[v] = P[v][c];
    with_center_geom<GEO>(gd, j, i, [&](const auto& g) { p_to_u(g, Pz, cst.gam, Uz); });
    for (int v = 0; v < NV; ++v) U[v][c] = Uz[v];
}

// Periodic ghost fill of up to 8 variables (variable blockIdx.y): every ghost
// zone copies its interior image.
template <typename T>
__global__ void __launch_bounds__(BS) sync_kernel(BlockDims b, Vars<T> A) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (c >= b.nc) return;
    const int k = (int)(c / b.s3), j = (int)((c - (long)k * b.s3) / b.s2), i = (int)(c - (long)k * b.s3 - (long)j * b.s2);
    const bool in_i = i >= b.is && i <= b.ie, in_j = j >= b.js && j <= b.je, in_k = k >= b.ks && k <= b.ke;
    if (in_i && in_j && in_k) return;
    auto wrap = [](int x, int s, int n) { int r = (x - s) % n; if (r < 0) r += n; return s + r; };
    const int si = (b.periodic[0] && !in_i) ? wrap(i, b.is, b.nx1) : i;
    const int sj = (b.ndim > 1 && b.periodic[1] && !in_j) ? wrap(j, b.js, b.nx2) : j;
    const int sk = (b.ndim > 2 && b.periodic[2] && !in_k) ? wrap(k, b.ks, b.nx3) : k;
    const long src = (long)sk * b.s3 + (long)sj * b.s2 + si;
    if (src != c) A[v][c] = A[v][src];
}

template <typename R>
__global__ void __launch_bounds__(BS) estimate_dt_kernel(BlockDims b, const R* __restrict__ cmax, const R* __restrict__ cmin,
                                                         double* __restrict__ part) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    double ndt = DBL_MAX;
    if (t < (unsigned)ninterior(b)) {
        int k, j, i;
        interior_index(b, t, k, j, i);
        const long c = (long)k * b.s3 + (long)j * b.s2 + i;
        const double n = zone_ndt(b, smax((double)cmax[c], (double)cmin[c]),
                                  smax((double)cmax[b.nc + c], (double)cmin[b.nc + c]),
                                  smax((double)cmax[2 * b.nc + c], (double)cmin[2 * b.nc + c]));
        if (!(n != n)) ndt = n;
    }
    const double m = block_min(ndt);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// Face-flux objective of one stage: acc += w dt sum F_dir[var] dA over the
// faces of index `face` along dir and the interior transverse range (dA: the
// product of the two transverse zone widths).  One work-group, fixed order.
template <typename R>
__global__ void __launch_bounds__(BS) face_flux_objective_kernel(BlockDims b, const R* __restrict__ F, int dir, int var,
                                                                 int face, double w, const DeviceState* __restrict__ ds,
                                                                 double* __restrict__ acc) {
    const int na = (dir == 1) ? (b.ndim > 1 ? b.nx2 : 1) : b.nx1;
    const int nb = (dir == 3) ? (b.ndim > 1 ? b.nx2 : 1) : (b.ndim > 2 ? b.nx3 : 1);
    const R* Fd = F + ((long)(dir - 1) * NV + var) * b.nc;
    double sum = 0.;
    for (int t = threadIdx.x; t < na * nb; t += blockDim.x) {
        const int ta = t % na, tb = t / na;
        int i, j, k;
        if (dir == 1) { i = face; j = b.js + ta; k = b.ks + tb; }
        else if (dir == 2) { i = b.is + ta; j = face; k = b.ks + tb; }
        else { i = b.is + ta; k = face; j = b.js + tb; }
        sum += (double)Fd[(long)k * b.s3 + (long)j * b.s2 + i];
    }
    __shared__ double red[BS];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int h = BS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dA = (dir == 1) ? b.dx2 * b.dx3 : (dir == 2) ? b.dx1 * b.dx3 : b.dx1 * b.dx2;
        acc[0] += w * ds->dt * red[0] * dA;
    }
}

// Floors::DetermineGRMHDFloors (floors/floors.cpp:197-231) over the entire
// block: fflag |= the floor and ceiling bits the zone's primitives would take
// (determine_floors, floors/floors_functions.hpp), the primitives unchanged.
template <int GEO>
__global__ void __launch_bounds__(BS) determine_floors_kernel(BlockDims b, StepConsts cst, GeomDev gd, Vars<const double> P,
                                                              int* __restrict__ fflag) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc) return;
    int k, j, i;
    zone_ijk(b, c, k, j, i);
    double Pz[NV];
    for (int v = 0; v < NV; ++v) Pz[v] = P[v][c];
    double rf, uf;
    geom_floors<GEO>(cst, gd, j, i, rf, uf);
    int ff = 0;
    with_center_geom<GEO>(gd, j, i, [&](const auto& g) {
        FourVec<double> D;
        calc_4vecs<double>(g, Pz, D);
        double rfm, ufm;
        ff = determine_floors(g, Pz, dot4(D.bcon, D.bcov), cst.gam, cst.floors, rf, uf, rfm, ufm);
    });
    fflag[c] |= ff;
}

// Flux::MarkFOFC (flux/fofc.cpp:45-111) over the entire block: a zone is marked
// when the floor flag of its trial state is set, or when `fixed` ([n2+1][n1+1],
// or null) marks it: cell centres within eh_buffer of the horizon and the first
// polar_cells rows at physical X2 faces.
__global__ void __launch_bounds__(BS) mark_fofc_kernel(BlockDims b, int gn1, const int* __restrict__ fflag,
                                                       const int* __restrict__ fixed, int* __restrict__ mark) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc) return;
    int k, j, i;
    zone_ijk(b, c, k, j, i);
    mark[c] = (fflag[c] != 0 || (fixed && fixed[(long)j * gn1 + i])) ? 1 : 0;
}

// pflag / fflag between the int32 arrays of this library and the reference's
// Real-valued fields, read back with static_cast<int> as the reference does.
__global__ void __launch_bounds__(BS) flags_to_real_kernel(long n, const int* __restrict__ f, double* __restrict__ r) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) r[t] = (double)f[t];
}
__global__ void __launch_bounds__(BS) flags_from_real_kernel(long n, const double* __restrict__ r, int* __restrict__ f) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f[t] = (int)r[t];
}

// ---------------------------------------------------------------- launchers
hipError_t launch_determine_floors(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& P,
                                   int* fflag, hipStream_t s) {
    if (!c.floors.on) return hipSuccess;
    if (b.geo) hipLaunchKernelGGL(determine_floors_kernel<1>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, P, fflag);
    else hipLaunchKernelGGL(determine_floors_kernel<0>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, P, fflag);
    return hipGetLastError();
}
hipError_t launch_mark_fofc(const BlockDims& b, const GeomDev& gd, const int* fflag, const int* fixed, int* mark,
                            hipStream_t s) {
    hipLaunchKernelGGL(mark_fofc_kernel, grid_for(b.nc), dim3(BS), 0, s, b, gd.gn1, fflag, fixed, mark);
    return hipGetLastError();
}
hipError_t launch_flags_to_real(const int* f, double* r, long n, hipStream_t s) {
    hipLaunchKernelGGL(flags_to_real_kernel, grid_for(n), dim3(BS), 0, s, n, f, r);
    return hipGetLastError();
}
hipError_t launch_flags_from_real(const double* r, int* f, long n, hipStream_t s) {
    hipLaunchKernelGGL(flags_from_real_kernel, grid_for(n), dim3(BS), 0, s, n, r, f);
    return hipGetLastError();
}
hipError_t launch_face_flux_objective(const BlockDims& b, int prec, const void* F, int dir, int var, int face, double w,
                                      const DeviceState* ds, double* acc, hipStream_t s) {
    if (prec == PREC_SPLIT)
        hipLaunchKernelGGL(face_flux_objective_kernel<float>, dim3(1), dim3(BS), 0, s, b, (const float*)F, dir, var, face, w, ds, acc);
    else
        hipLaunchKernelGGL(face_flux_objective_kernel<double>, dim3(1), dim3(BS), 0, s, b, (const double*)F, dir, var, face, w, ds, acc);
    return hipGetLastError();
}
hipError_t launch_emf(const BlockDims& b, int prec, const FluxVars<void>& F, void* emf, hipStream_t s) {
    if (b.ndim < 2) return hipSuccess;
    const long n = (long)(b.nx1 + 1) * (b.nx2 + 1) * ((b.ndim > 2) ? b.nx3 + 1 : 1);
    if (prec == PREC_SPLIT)
        hipLaunchKernelGGL(emf_kernel<float>, grid_for(n), dim3(BS), 0, s, b, flux_vars_cast<const float>(F), (float*)emf);
    else
        hipLaunchKernelGGL(emf_kernel<double>, grid_for(n), dim3(BS), 0, s, b, flux_vars_cast<const double>(F), (double*)emf);
    return hipGetLastError();
}
hipError_t launch_flux_ct_rewrite(const BlockDims& b, int prec, const FluxVars<void>& F, const void* emf, hipStream_t s) {
    if (b.ndim < 2) return hipSuccess;
    const long n = (long)(b.nx1 + 1) * (b.nx2 + 1) * ((b.ndim > 2) ? b.nx3 + 1 : 1);
    if (prec == PREC_SPLIT)
        hipLaunchKernelGGL(ct_rewrite_kernel<float>, grid_for(n), dim3(BS), 0, s, b, flux_vars_cast<float>(F), (const float*)emf);
    else
        hipLaunchKernelGGL(ct_rewrite_kernel<double>, grid_for(n), dim3(BS), 0, s, b, flux_vars_cast<double>(F), (const double*)emf);
    return hipGetLastError();
}
hipError_t launch_geo_source(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& P,
                             const Vars<double>& dUdt, hipStream_t s) {
    if (!b.geo) return hipSuccess;   // Cartesian Minkowski: no source (flux.cpp:407)
    hipLaunchKernelGGL(geo_source_kernel, grid_for(ninterior_host(b)), dim3(BS), 0, s, b, c, gd, P, dUdt);
    return hipGetLastError();
}
hipError_t launch_flux_divergence(const BlockDims& b, int prec, const FluxVars<void>& F, const Vars<double>& dUdt,
                                  hipStream_t s) {
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (prec == PREC_SPLIT)
        hipLaunchKernelGGL(divergence_kernel<float>, grid_for(n), dim3(BS), 0, s, b, flux_vars_cast<const float>(F), dUdt);
    else
        hipLaunchKernelGGL(divergence_kernel<double>, grid_for(n), dim3(BS), 0, s, b, flux_vars_cast<const double>(F), dUdt);
    return hipGetLastError();
}
// ---------------------------------------------- first-order flux correction
// Flux::MarkFOFC (flux/fofc.cpp:45-111) on the trial state of
// KHARMADriver::AddFOFC (driver/kharma_driver.cpp:292-337): U_trial = gam0 U_sub
// + gam1 U_full + beta dt (-div F) from the stage's raw fluxes (no source term),
// its U->P from the stage's primitives with B = U.B / sqrt(-g), and
// DetermineGRMHDFloors.  A zone is marked when the inverter's floor bits or the
// floors' flags are set, or when `fixed` marks it (within the horizon buffer,
// polar rows).  Interior zones only.
// STG (3-D curved space): the zones run on the recovery's tiles, with the tile's
// metric and floors staged in LDS once, as the recovery does: 132 instead of 188
// VGPRs (3 waves per SIMD instead of 2) and 344 -> 319 us per launch on the FP64
// 192x128x128 torus, with the same marks.  (A 4-wave bound spills 36 B per lane and
// measured the same.)
template <int GEO, int INV, bool STG, int TI = 64>
__global__ void __launch_bounds__(BS)
fofc_mark_kernel(BlockDims b, StepConsts cst, GeomDev gd, int stage, const double* __restrict__ F,
                 const double* __restrict__ Usub, const double* Ufull, const double* __restrict__ Psub,
                 const int* __restrict__ fixed, const DeviceState* __restrict__ ds, int* __restrict__ mark,
                 bool images) {
    __shared__ double sg[STG ? STG_RECOVER * TI : 1];
    int k, j, i, i0 = 0;
    bool active;
    if constexpr (STG) {
        active = tile_zone<TI>(b, k, j, i, i0);
    } else {
        const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
        active = t < (unsigned)ninterior(b);
        if (!active) return;
        interior_index(b, t, k, j, i);
    }
    const long c = (long)k * b.s3 + (long)j * b.s2 + i;
    double U[NV], P[NV];
    if (active) {
        double dudt[NV];
        zone_divergence<double>(b, flux_vars_of(F, b.nc), nullptr, c, false, dudt);
        const double g0 = cst.gam0[stage - 1], g1 = cst.gam1[stage - 1];
        const double beta_dt = cst.beta[stage - 1] * ds->dt;
        const bool same = (g0 == 0.) && (Usub == Ufull);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double uf = Ufull[v * b.nc + c];
            const double u = same ? uf : g0 * Usub[v * b.nc + c] + g1 * uf;
            U[v] = 1.0 * u + beta_dt * dudt[v];
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) P[v] = Psub[v * b.nc + c];
    }
    if constexpr (STG) {
        stage_recover_geometry<TI>(gd, j, i0, sg);
        if (!active) return;
    }
    double rf, uf;
    const int lane = (int)(threadIdx.x % TI);
    if constexpr (STG) {
        rf = sg[GEO_NCOMP * TI + lane];
        uf = sg[(GEO_NCOMP + 1) * TI + lane];
    } else {
        geom_floors<GEO>(cst, gd, j, i, rf, uf);
    }
    int ff = 0;
    auto trial = [&](const auto& g) {
        const int pf = invert_zone<INV>(g, U, P, cst);
        ff = (pf / F_GEOM_RHO) * F_GEOM_RHO;
        const double igd = rcp(g.gdet());
        P[B1] = U[B1] * igd; P[B2] = U[B2] * igd; P[B3] = U[B3] * igd;
        if (cst.floors.on) {
            FourVec<double> D;
            calc_4vecs<double>(g, P, D);
            double rfm, ufm;
            ff |= determine_floors(g, P, dot4(D.bcon, D.bcov), cst.gam, cst.floors, rf, uf, rfm, ufm);
        }
    };
    if constexpr (STG) {
        CachedGeom<double> g(sg, TI, lane);
        trial(g);
    } else {
        with_center_geom<GEO>(gd, j, i, trial);
    }
    const int mv = (ff != 0 || (fixed && fixed[(long)j * gd.gn1 + i])) ? 1 : 0;
    if (images) {   // the boundary sync of the flag (kharma_driver.cpp:328-331)
        ImageSet s;
        image_set(b, k, j, i, s);
        store_with_images(b, s, mark, c, mv, b.nc);
    } else {
        mark[c] = mv;
    }
}

hipError_t launch_fofc_mark(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, const double* F,
                            const double* Usub, const double* Ufull, const double* Psub, const int* fixed,
                            const DeviceState* ds, int* mark, hipStream_t s, bool images) {
    const dim3 grid = grid_for(ninterior_host(b));
#define KH_MARK(GEO, INV, STG, TI, G) \
    hipLaunchKernelGGL((fofc_mark_kernel<GEO, INV, STG, TI>), G, dim3(BS), 0, s, b, c, gd, stage, F, Usub, Ufull, Psub, \
                       fixed, ds, mark, images)
    if (b.geo && b.ndim == 3) {
        const dim3 gt((unsigned)update_tiles_host(b));
        const bool ti32 = recover_tile_i(b) == 32;
        if (c.inverter == 1) { if (ti32) KH_MARK(1, 1, true, 32, gt); else KH_MARK(1, 1, true, 64, gt); }
        else { if (ti32) KH_MARK(1, 0, true, 32, gt); else KH_MARK(1, 0, true, 64, gt); }
    } else if (b.geo) {
        if (c.inverter == 1) KH_MARK(1, 1, false, 64, grid); else KH_MARK(1, 0, false, 64, grid);
    } else {
        if (c.inverter == 1) KH_MARK(0, 1, false, 64, grid); else KH_MARK(0, 0, false, 64, grid);
    }
#undef KH_MARK
    return hipGetLastError();
}

template <typename R, bool PK>
static void launch_update_geo_k(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, bool do_dt,
                              const void* F, const void* ctop, const double* Usub, const double* Ufull,
                              const double* Psub, double* Uout, double* Pout, float* Pout32, int* pflag, int* fflag,
                              int* fail_list, DeviceState* ds, double* dt_part, hipStream_t s, int parts,
                              double* Ustar, const Pack* pk) {
    const dim3 g = pack_grid(grid_for(ninterior_host(b)), pk);
    const bool rstg = b.ndim == 3;   // recover on LDS-staged tiles
    const bool ti32 = recover_tile_i(b) == 32;
    const dim3 gr = rstg ? pack_grid(dim3((unsigned)update_tiles_host(b)), pk) : g;
    const BlockEntry* pd = pack_dev(pk);
    const bool same = (c.gam0[stage - 1] == 0.) && (Usub == Ufull);
    if (parts & 1) {
#define KH_CMB(DT, ND, SAME)                                                                                          \
    hipLaunchKernelGGL((combine_kernel<R, DT, ND, SAME, PK>), g, dim3(BS), 0, s, b, c, gd, stage, (const R*)F, (const R*)ctop, \
                       Usub, Ufull, Psub, Uout, ds, dt_part, pd)
#define KH_CMB_ND(DT, SAME) \
    { if (b.ndim == 3) KH_CMB(DT, 3, SAME); else if (b.ndim == 2) KH_CMB(DT, 2, SAME); else KH_CMB(DT, 1, SAME); }
    if (do_dt) { if (same) KH_CMB_ND(true, true) else KH_CMB_ND(true, false) }
    else { if (same) KH_CMB_ND(false, true) else KH_CMB_ND(false, false) }
#undef KH_CMB_ND
#undef KH_CMB
    }
    if (!(parts & 2)) return;
#define KH_REC(FR, INV)                                                                                             \
    do {                                                                                                            \
        if (rstg && ti32) hipLaunchKernelGGL((recover_kernel<FR, INV, true, PK, 32>), gr, dim3(BS), 0, s, b, c, gd,     \
                                             stage, Psub, Uout, Pout, Pout32, pflag, fflag, fail_list, ds, Ustar, pd);  \
        else if (rstg) hipLaunchKernelGGL((recover_kernel<FR, INV, true, PK>), gr, dim3(BS), 0, s, b, c, gd, stage, Psub,   \
                                          Uout, Pout, Pout32, pflag, fflag, fail_list, ds, Ustar, pd);                  \
        else hipLaunchKernelGGL((recover_kernel<FR, INV, false, PK>), gr, dim3(BS), 0, s, b, c, gd, stage, Psub, Uout, Pout, \
                                Pout32, pflag, fflag, fail_list, ds, Ustar, pd);                                        \
    } while (0)
    if (c.inverter == 1) {
        if (c.floors.frame == 0) KH_REC(0, 1);
        else if (c.floors.frame == 1) KH_REC(1, 1);
        else KH_REC(2, 1);
    } else {
        if (c.floors.frame == 0) KH_REC(0, 0);
        else if (c.floors.frame == 1) KH_REC(1, 0);
        else KH_REC(2, 0);
    }
#undef KH_REC
}
template <typename R>
static void launch_update_geo(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, bool do_dt,
                              const void* F, const void* ctop, const double* Usub, const double* Ufull,
                              const double* Psub, double* Uout, double* Pout, float* Pout32, int* pflag, int* fflag,
                              int* fail_list, DeviceState* ds, double* dt_part, hipStream_t s, int parts = 3,
                              double* Ustar = nullptr, const Pack* pk = nullptr) {
    if (pk) launch_update_geo_k<R, true>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32, pflag,
                                         fflag, fail_list, ds, dt_part, s, parts, Ustar, pk);
    else launch_update_geo_k<R, false>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32, pflag,
                                       fflag, fail_list, ds, dt_part, s, parts, Ustar, pk);
}
template <typename R, int GEO, bool PK>
static void launch_update_t_k(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, bool do_dt,
                            const void* F, const void* ctop, const double* Usub, const double* Ufull, const double* Psub,
                            double* Uout, double* Pout, float* Pout32, int* pflag, int* fflag,
                            int* fail_list, DeviceState* ds, double* dt_part, hipStream_t s, const double* fB,
                            const Pack* pk) {
    const dim3 g = pack_grid(grid_for(ninterior_host(b)), pk);
    const int bm = fB ? 2 : (b.ndim > 1 ? 1 : 0);
    const BlockEntry* pd = pack_dev(pk);
#define KH_UPD(DT, BM, INV)                                                                                         \
    hipLaunchKernelGGL((update_kernel<R, DT, BM, GEO, INV, PK>), g, dim3(BS), 0, s, b, c, gd, stage, (const R*)F, (const R*)ctop, \
                       Usub, Ufull, Psub, Uout, Pout, Pout32, pflag, fflag, fail_list, ds, dt_part, fB, pd)
#define KH_UPD_BM(DT, INV) \
    { if (bm == 2) KH_UPD(DT, 2, INV); else if (bm == 1) KH_UPD(DT, 1, INV); else KH_UPD(DT, 0, INV); }
    if (c.inverter == 1) {
        if (do_dt) KH_UPD_BM(true, 1) else KH_UPD_BM(false, 1)
    } else {
        if (do_dt) KH_UPD_BM(true, 0) else KH_UPD_BM(false, 0)
    }
#undef KH_UPD_BM
#undef KH_UPD
}
template <typename R, int GEO>
static void launch_update_t(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, bool do_dt,
                            const void* F, const void* ctop, const double* Usub, const double* Ufull, const double* Psub,
                            double* Uout, double* Pout, float* Pout32, int* pflag, int* fflag,
                            int* fail_list, DeviceState* ds, double* dt_part, hipStream_t s, const double* fB = nullptr,
                            const Pack* pk = nullptr) {
    if (pk) launch_update_t_k<R, GEO, true>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32,
                                            pflag, fflag, fail_list, ds, dt_part, s, fB, pk);
    else launch_update_t_k<R, GEO, false>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32,
                                          pflag, fflag, fail_list, ds, dt_part, s, fB, pk);
}
hipError_t launch_update_faces(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int prec, int stage,
                               bool do_dt, const void* F, const void* ctop, const double* Usub, const double* Ufull,
                               const double* Psub, double* Uout, double* Pout, float* Pout32, int* pflag, int* fflag,
                               int* fail_list, DeviceState* ds, double* dt_part, const double* fB, hipStream_t s) {
    if (b.geo || !fB) return hipErrorInvalidValue;
    if (prec == PREC_SPLIT)
        launch_update_t<float, 0>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32, pflag, fflag,
                                  fail_list, ds, dt_part, s, fB);
    else
        launch_update_t<double, 0>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32, pflag, fflag,
                                   fail_list, ds, dt_part, s, fB);
    return hipGetLastError();
}
hipError_t launch_update(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int prec, int stage, bool do_dt,
                         const void* F, const void* ctop, const double* Usub, const double* Ufull, const double* Psub,
                         double* Uout, double* Pout, float* Pout32, int* pflag, int* fflag,
                         int* fail_list, DeviceState* ds, double* dt_part, hipStream_t s, const Pack* pk) {
    if (prec == PREC_SPLIT) {
        if (b.geo) launch_update_geo<float>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32, pflag, fflag, fail_list, ds, dt_part, s, 3, nullptr, pk);
        else launch_update_t<float, 0>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32, pflag, fflag, fail_list, ds, dt_part, s, nullptr, pk);
    } else {
        if (b.geo) launch_update_geo<double>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32, pflag, fflag, fail_list, ds, dt_part, s, 3, nullptr, pk);
        else launch_update_t<double, 0>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, Pout32, pflag, fflag, fail_list, ds, dt_part, s, nullptr, pk);
    }
    return hipGetLastError();
}
hipError_t launch_update_parts(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, int parts,
                               const double* F, const double* Usub, const double* Ufull, const double* Psub, double* Uout,
                               double* Pout, int* pflag, int* fflag, int* fail_list, DeviceState* ds, hipStream_t s,
                               double* Ustar) {
    if (!b.geo) return hipErrorInvalidValue;
    launch_update_geo<double>(b, c, gd, stage, false, F, nullptr, Usub, Ufull, Psub, Uout, Pout, nullptr, pflag, fflag,
                              fail_list, ds, nullptr, s, parts, Ustar);
    return hipGetLastError();
}
hipError_t launch_update_recorded(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, bool do_dt,
                                  const void* F, const void* ctop, const double* Usub, const double* Ufull,
                                  const double* Psub, double* Uout, double* Pout, int* pflag, int* fflag,
                                  int* fail_list, DeviceState* ds, double* dt_part, double* Ustar, hipStream_t s) {
    if (!b.geo)
        return launch_update(b, c, gd, PREC_FP64, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, nullptr, pflag,
                             fflag, fail_list, ds, dt_part, s);
    launch_update_geo<double>(b, c, gd, stage, do_dt, F, ctop, Usub, Ufull, Psub, Uout, Pout, nullptr, pflag, fflag,
                              fail_list, ds, dt_part, s, 3, Ustar);
    return hipGetLastError();
}
long update_blocks(const BlockDims& b) { return (ninterior_host(b) + BS - 1) / BS; }
long divb_blocks(const BlockDims& b) {
    return ((long)(b.nx1 + 1) * (b.nx2 + 1) * ((b.ndim > 2) ? b.nx3 + 1 : 1) + BS - 1) / BS;
}
hipError_t launch_fixup(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, double* U, double* P,
                        float* P32, const int* pflag, const int* fail_list, DeviceState* ds, hipStream_t s,
                        const Pack* pk) {
    const dim3 g = pack_grid(dim3(64), pk);
    const BlockEntry* pd = pack_dev(pk);
#define KH_FIX(GEO, PK) \
    hipLaunchKernelGGL((fixup_list_kernel<GEO, PK>), g, dim3(BS), 0, s, b, c, gd, stage, U, P, P32, pflag, fail_list, ds, pd)
    if (pk) { if (b.geo) KH_FIX(1, true); else KH_FIX(0, true); }
    else { if (b.geo) KH_FIX(1, false); else KH_FIX(0, false); }
#undef KH_FIX
    return hipGetLastError();
}
static long fix_flux_threads(const BlockDims& b) {
    const long npolar = nphys(b, 1) ? 2L * (b.nx1 + 2) * ((b.ndim > 2) ? b.nx3 + 2 : 1) : 0;
    const long nx1f = nphys(b, 0) ? 2L * ((b.ndim > 1) ? b.nx2 + 2 : 1) * ((b.ndim > 2) ? b.nx3 + 2 : 1) : 0;
    return npolar + nx1f;
}
static hipError_t launch_fix_flux_pk(const BlockDims& b, int prec, const FluxVars<void>& F, hipStream_t s,
                                     const Pack* pk) {
    const long n = pk ? pk->fix_flux_n : fix_flux_threads(b);
    if (n == 0) return hipSuccess;
    const dim3 g = pack_grid(grid_for(n), pk);
    const BlockEntry* pd = pack_dev(pk);
#define KH_FF(R, PK) hipLaunchKernelGGL((fix_flux_kernel<R, PK>), g, dim3(BS), 0, s, b, flux_vars_cast<R>(F), pd)
    if (pk) { if (prec == PREC_SPLIT) KH_FF(float, true); else KH_FF(double, true); }
    else { if (prec == PREC_SPLIT) KH_FF(float, false); else KH_FF(double, false); }
#undef KH_FF
    return hipGetLastError();
}
hipError_t launch_fix_flux(const BlockDims& b, int prec, const FluxVars<void>& F, hipStream_t s) {
    return launch_fix_flux_pk(b, prec, F, s, nullptr);
}
hipError_t launch_fix_flux(const BlockDims& b, int prec, void* F, hipStream_t s, const Pack* pk) {
    return launch_fix_flux_pk(b, prec, flux_vars_prec(F, prec, b.nc), s, pk);
}
static hipError_t launch_bounds_pk(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<double>& U,
                                   const Vars<double>& P, float* P32, hipStream_t s, const Pack* pk) {
    if (!pk && !has_physical_face(b)) return hipSuccess;
    const long n = pk ? pk->bounds_n : bounds_slab1(b) + bounds_slab2(b);
    if (n == 0) return hipSuccess;
    const dim3 g = pack_grid(grid_for(n), pk);
    const BlockEntry* pd = pack_dev(pk);
#define KH_BD(GEO, PK) hipLaunchKernelGGL((bounds_kernel<GEO, PK>), g, dim3(BS), 0, s, b, c, gd, U, P, P32, pd)
    if (pk) { if (b.geo) KH_BD(1, true); else KH_BD(0, true); }
    else { if (b.geo) KH_BD(1, false); else KH_BD(0, false); }
#undef KH_BD
    return hipGetLastError();
}
hipError_t launch_bounds(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<double>& U,
                         const Vars<double>& P, float* P32, hipStream_t s) {
    return launch_bounds_pk(b, c, gd, U, P, P32, s, nullptr);
}
hipError_t launch_bounds(const BlockDims& b, const StepConsts& c, const GeomDev& gd, double* U, double* P, float* P32,
                         hipStream_t s, const Pack* pk) {
    return launch_bounds_pk(b, c, gd, vars_of(U, b.nc), vars_of(P, b.nc), P32, s, pk);
}
void pack_set_sizes(Pack& pk, const BlockDims* dims, int nblk) {
    pk.fix_flux_n = pk.bounds_n = 0;
    for (int n = 0; n < nblk; ++n) {
        pk.fix_flux_n = std::max(pk.fix_flux_n, fix_flux_threads(dims[n]));
        if (has_physical_face(dims[n])) pk.bounds_n = std::max(pk.bounds_n, bounds_slab1(dims[n]) + bounds_slab2(dims[n]));
    }
}
hipError_t launch_finalize(const StepConsts& c, DeviceState* ds, const double* dt_part, int n_dt,
                           const double* divb_part, int n_divb, double* dt_hist, int hist_cap, hipStream_t s, int nds) {
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1024), 0, s, c, ds, nds, dt_part, n_dt, divb_part, n_divb, dt_hist,
                       hist_cap);
    return hipGetLastError();
}
static hipError_t launch_max_divb_pk(const BlockDims& b, const Vars<const double>& U, double* part, hipStream_t s,
                                     const Pack* pk) {
    if (b.ndim < 2) return hipSuccess;
    const dim3 g = pack_grid(dim3((unsigned)divb_blocks(b)), pk);
    if (pk) hipLaunchKernelGGL(max_divb_kernel<true>, g, dim3(BS), 0, s, b, U, part, pack_dev(pk));
    else hipLaunchKernelGGL(max_divb_kernel<false>, g, dim3(BS), 0, s, b, U, part, nullptr);
    return hipGetLastError();
}
hipError_t launch_max_divb(const BlockDims& b, const Vars<const double>& U, double* part, hipStream_t s) {
    return launch_max_divb_pk(b, U, part, s, nullptr);
}
hipError_t launch_max_divb(const BlockDims& b, const double* U, double* part, hipStream_t s, const Pack* pk) {
    return launch_max_divb_pk(b, vars_of(U, b.nc), part, s, pk);
}
hipError_t launch_reduce(const double* part, int n, int is_max, double* out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(BS), 0, s, part, n, is_max, out);
    return hipGetLastError();
}
hipError_t launch_reset_state(DeviceState* ds, double dt, hipStream_t s) {
    hipLaunchKernelGGL(reset_state_kernel, dim3(1), dim3(64), 0, s, ds, dt);
    return hipGetLastError();
}
hipError_t launch_set_clock(DeviceState* ds, double time, double dt, int ncycle, hipStream_t s) {
    hipLaunchKernelGGL(set_clock_kernel, dim3(1), dim3(64), 0, s, ds, time, dt, ncycle);
    return hipGetLastError();
}
hipError_t launch_weighted_sum(const BlockDims& b, const Vars<const double>& Usub, const Vars<const double>& Ufull,
                               const Vars<const double>& dUdt, double gam0, double gam1, double beta_dt,
                               const Vars<double>& Uout, hipStream_t s) {
    const dim3 grid((unsigned)((b.nc + BS - 1) / BS), NV);
    hipLaunchKernelGGL(weighted_sum_kernel, grid, dim3(BS), 0, s, b.nc, Usub, Ufull, dUdt, gam0, gam1, beta_dt, Uout);
    return hipGetLastError();
}
hipError_t launch_utop(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& U,
                       const Vars<double>& P, int* pflag, int* fflag, hipStream_t s) {
    if (c.inverter == 1) {
        if (b.geo) hipLaunchKernelGGL((utop_kernel<1, 1>), grid_for(b.nc), dim3(BS), 0, s, b, c, gd, U, P, pflag, fflag);
        else hipLaunchKernelGGL((utop_kernel<0, 1>), grid_for(b.nc), dim3(BS), 0, s, b, c, gd, U, P, pflag, fflag);
    } else {
        if (b.geo) hipLaunchKernelGGL(utop_kernel<1>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, U, P, pflag, fflag);
        else hipLaunchKernelGGL(utop_kernel<0>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, U, P, pflag, fflag);
    }
    return hipGetLastError();
}
hipError_t launch_floors(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<double>& U,
                         const Vars<double>& P, int* pflag, int* fflag, hipStream_t s) {
    if (!c.floors.on) return hipSuccess;
    if (b.geo) hipLaunchKernelGGL(floors_kernel<1>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, U, P, pflag, fflag);
    else hipLaunchKernelGGL(floors_kernel<0>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, U, P, pflag, fflag);
    return hipGetLastError();
}
hipError_t launch_fixup_entire(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<double>& U,
                               const Vars<double>& P, const int* pflag, hipStream_t s) {
    if (!c.fix_average && !c.fix_atmo) return hipSuccess;
    hipLaunchKernelGGL(fixup_avg_kernel, grid_for(b.nc), dim3(BS), 0, s, b, c, P, pflag);
    if (b.geo) hipLaunchKernelGGL(fixup_floor_kernel<1>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, U, P, pflag);
    else hipLaunchKernelGGL(fixup_floor_kernel<0>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, U, P, pflag);
    return hipGetLastError();
}
hipError_t launch_ptou(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& P,
                       const Vars<double>& U, hipStream_t s) {
    if (b.geo) hipLaunchKernelGGL(ptou_kernel<1>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, P, U);
    else hipLaunchKernelGGL(ptou_kernel<0>, grid_for(b.nc), dim3(BS), 0, s, b, c, gd, P, U);
    return hipGetLastError();
}
hipError_t launch_ptou(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P, double* U,
                       hipStream_t s) {
    return launch_ptou(b, c, gd, vars_of(P, b.nc), vars_of(U, b.nc), s);
}
hipError_t launch_sync(const BlockDims& b, const Vars<double>& A, int nvar, hipStream_t s) {
    if (nvar < 1 || nvar > NV) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((b.nc + BS - 1) / BS), (unsigned)nvar);
    hipLaunchKernelGGL(sync_kernel<double>, grid, dim3(BS), 0, s, b, A);
    return hipGetLastError();
}
hipError_t launch_sync(const BlockDims& b, double* A, int nvar, hipStream_t s) {
    for (int v0 = 0; v0 < nvar; v0 += NV) {
        const hipError_t e = launch_sync(b, vars_of(A + (long)v0 * b.nc, b.nc), nvar - v0 < NV ? nvar - v0 : NV, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t launch_sync_int(const BlockDims& b, int* A, hipStream_t s) {
    hipLaunchKernelGGL(sync_kernel<int>, grid_for(b.nc), dim3(BS), 0, s, b, vars_of(A, b.nc));
    return hipGetLastError();
}
hipError_t launch_estimate_dt(const BlockDims& b, const StepConsts& c, int prec, const void* cmax, const void* cmin,
                              double* part, hipStream_t s) {
    const dim3 g = grid_for(ninterior_host(b));
    if (prec == PREC_SPLIT) hipLaunchKernelGGL(estimate_dt_kernel<float>, g, dim3(BS), 0, s, b, (const float*)cmax, (const float*)cmin, part);
    else hipLaunchKernelGGL(estimate_dt_kernel<double>, g, dim3(BS), 0, s, b, (const double*)cmax, (const double*)cmin, part);
    return hipGetLastError();
}
// ---------------------------------------------- allocation-free timestep estimate
// Scratch of kh_estimate_timestep_noalloc: one CFL minimum per work-group of
// estimate_dt_kernel.  Each GPU holds its own copy for the life of the process;
// calls on one GPU share it, so calls that overlap on one GPU race.  Every
// partial a call reads is written by the same call, so nothing needs clearing.
static __device__ __align__(256) double scratch[KH_NOALLOC_PARTIALS];

// The reduced minimum clipped as kh_estimate_timestep does on the host:
// n = m cfl, then min(max(dt_min, n), max_dt_increase dt_last).
__global__ void __launch_bounds__(BS) reduce_dt_clip_kernel(const double* __restrict__ part, int n, double cfl,
                                                            double dt_min, double hi, double* __restrict__ out) {
    double v = DBL_MAX;
    for (int q = threadIdx.x; q < n; q += blockDim.x) v = fmin(v, part[q]);
    v = block_min(v);
    if (threadIdx.x == 0) {
        double w = v * cfl;
        w = (w < dt_min) ? dt_min : w;
        w = (hi < w) ? hi : w;
        *out = w;
    }
}

namespace {
std::mutex g_scratch_mu;
double* g_scratch_dev[KH_MAX_DEVICES] = {};
}  // namespace

// Device address of `scratch` on the current device, resolved once per device.
hipError_t noalloc_scratch(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= KH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = (double*)p;
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

hipError_t launch_estimate_dt_noalloc(const BlockDims& b, const StepConsts& c, int prec, const void* cmax,
                                      const void* cmin, double dt_last, double* ndt, hipStream_t s) {
    double* part = nullptr;
    hipError_t e = noalloc_scratch(&part);
    if (e != hipSuccess) return e;
    const long n = update_blocks(b);
    if (n > KH_NOALLOC_PARTIALS) return hipErrorInvalidValue;
    e = launch_estimate_dt(b, c, prec, cmax, cmin, part, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(reduce_dt_clip_kernel, dim3(1), dim3(BS), 0, s, part, (int)n, c.cfl, c.dt_min,
                       c.max_dt_increase * dt_last, ndt);
    return hipGetLastError();
}

hipError_t launch_convert_f32(const double* A, float* B, long n, hipStream_t s) {
    hipLaunchKernelGGL(convert_f32_kernel, grid_for(n), dim3(BS), 0, s, A, B, n);
    return hipGetLastError();
}

}  // namespace kh
