// kh_adjoint.hip -- reverse-mode (adjoint) kernels of one RK stage of the
// GRMHD step: WENO5 reconstruction + LLF face fluxes, flux-CT + divergence,
// the conserved update and the U->P recovery (implicit-function theorem).
//
// The forward stage (kh_flux.hip, kh_step.hip) is
//   F_d = LLF(WENO5(P))            per direction d (get_flux.hpp:60-373)
//   F_ct = CT(F), dudt = -div F_ct  (b_flux_ct.cpp:307-370, kharma_driver.hpp:208-231)
//   U* = g0 U_sub + g1 U_full + beta dt dudt          (kharma_driver.cpp:339-387)
//   P_out = u_to_p(U*), U_out = p_to_u(P_out)         (inverter/onedw.hpp:89-233)
// and these kernels apply its transpose, stage by stage, with the timestep
// held fixed.  Each kernel writes every output element it owns exactly once
// (gathers, no atomics), so the adjoint is deterministic.
//
// Face adjoints:  adj_face_* recompute the face states of the forward tiles and
// write, per cell m, lamL[v][m] (adjoint of the value at m's left face, from
// face m) and lamR[v][m] (at m's right face, from face m + 1).  adj_recon then
// applies the WENO5 transpose per cell and gathers the 5-point contributions
// into lamP through LDS; it reads lamL / lamR only where those faces exist, so
// the scratch needs no clearing.
#include "kh_internal.hpp"
#include "kh_adjoint.hpp"

namespace kh {

namespace {

constexpr int ABS = 256;

inline dim3 agrid(long n, int bs = ABS) { return dim3((unsigned)((n + bs - 1) / bs)); }

__global__ void __launch_bounds__(ABS) fill_kernel(double* __restrict__ a, long n, double v) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// Geometry of a face at entire-block indices (j, i), as face_flux() uses it.
template <int DIR, int GEO, class Fn>
__device__ __forceinline__ void with_face_geom(const StepConsts& c, const GeomDev& gd, int j, int i, Fn&& fn) {
    if constexpr (GEO) {
        const long q = (long)j * gd.gn1 + i;
        const int loc = (DIR == 1) ? GLOC_FACE1 : (DIR == 2) ? GLOC_FACE2 : GLOC_CENTER;
        CachedGeom<double> g(gd.g[loc], gd.np, q);
        fn(g, gd.flr[q], gd.flr[gd.np + q]);
    } else {
        MinkowskiGeom<double> g;
        fn(g, c.floors.rho_min_const, c.floors.u_min_const);
    }
}

template <int DIR, int GEO>
__device__ __forceinline__ void face_adjoint(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int j, int i,
                                             long ci, long st, const double Pl[NV], const double Pr[NV],
                                             const double* __restrict__ lF, double* __restrict__ lamL,
                                             double* __restrict__ lamR) {
    double lf[NV], lPl[NV], lPr[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) lf[v] = lF[v * b.nc + ci];
    with_face_geom<DIR, GEO>(c, gd, j, i, [&](const auto& g, double rflr, double uflr) {
        face_vjp_llf<DIR>(g, c.gam, c.recon_floors != 0, rflr, uflr, Pl, Pr, lf, lPl, lPr);
    });
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        lamL[v * b.nc + ci] = lPr[v];
        lamR[v * b.nc + ci - st] = lPl[v];
    }
}

// X1 faces: the tile scheme of flux_x1_kernel (flattened rows, 256 cells per
// block, consecutive blocks overlapping by one cell).
template <int GEO>
__global__ void __launch_bounds__(256)
adj_face_x1_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ P, const double* __restrict__ lF,
                   double* __restrict__ lamL, double* __restrict__ lamR, int j0, int nj, int k0, int nk) {
    __shared__ double sr[NV][256];
    const int L = b.nx1 + 2;
    const int tid = threadIdx.x;
    const long g = (long)blockIdx.x * 255 + tid;
    const bool valid = g < (long)L * nj * nk;
    int ii = 0, jf = 0;
    long ci = 0;
    double lo[NV];
    if (valid) {
        const int row = (int)(g / L);
        ii = (int)(g - (long)row * L);
        const int jj = row % nj, kk = row / nj;
        jf = j0 + jj;
        ci = (long)(k0 + kk) * b.s3 + (long)(j0 + jj) * b.s2 + (b.is - 1 + ii);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double* q = P + v * b.nc + ci;
            double ro;
            weno5_cell(q[-2], q[-1], q[0], q[1], q[2], lo[v], ro);
            sr[v][tid] = ro;
        }
    }
    __syncthreads();
    if (!valid || tid == 0 || ii == 0) return;
    double Pl[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Pl[v] = sr[v][tid - 1];
    face_adjoint<1, GEO>(b, c, gd, jf, b.is - 1 + ii, ci, 1, Pl, lo, lF, lamL, lamR);
}

// X2 / X3 faces: the 64 x 8 tiles of flux_xt_kernel.
template <int DIR, int GEO>
__global__ void __launch_bounds__(512)
adj_face_xt_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ P, const double* __restrict__ lF,
                   double* __restrict__ lamL, double* __restrict__ lamR, int o0, int no, int s0, int s1) {
    __shared__ double sr[NV][8][64];
    const int ni = b.nx1 + 2;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int col = blockIdx.x * 64 + tx;
    const int s = s0 - 1 + (int)blockIdx.y * 7 + ty;
    const bool valid = col < ni * no && s <= s1;
    const long st = (DIR == 2) ? b.s2 : b.s3;
    long ci = 0;
    int jf = 0, if_ = 0;
    double lo[NV];
    if (valid) {
        const int outer = o0 + col / ni;
        const int i = b.is - 1 + col % ni;
        const int j = (DIR == 2) ? s : outer;
        const int k = (DIR == 2) ? outer : s;
        jf = j; if_ = i;
        ci = (long)k * b.s3 + (long)j * b.s2 + i;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double* q = P + v * b.nc + ci;
            double ro;
            weno5_cell(q[-2 * st], q[-st], q[0], q[st], q[2 * st], lo[v], ro);
            sr[v][ty][tx] = ro;
        }
    }
    __syncthreads();
    if (!valid || ty == 0) return;
    double Pl[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Pl[v] = sr[v][ty - 1][tx];
    face_adjoint<DIR, GEO>(b, c, gd, jf, if_, ci, st, Pl, lo, lF, lamL, lamR);
}

// WENO5 transpose and 5-point gather along DIR, accumulated into lamP.
// Cells m along DIR in [s-1, e+1] were reconstructed by the forward pass; their
// stencils, and so the gather, cover cells [s-3, e+3].  Tiles hold TN cells along DIR and emit the
// TN - 4 cells whose whole +-2 neighbourhood lies inside the tile.
// DIR = 1: block (64, 4), x along i, y over transverse rows (j, k).
// DIR = 2, 3: block (64, 16), x over transverse columns (outer, i), y along DIR.
template <int DIR>
__global__ void __launch_bounds__(1024)
adj_recon_kernel(BlockDims b, const double* __restrict__ P, const double* __restrict__ lamL,
                 const double* __restrict__ lamR, double* __restrict__ lamP, int j0, int nj, int k0, int nk) {
    constexpr int TN = (DIR == 1) ? 64 : 16;
    constexpr int TT = (DIR == 1) ? 4 : 64;
    __shared__ double G[5][TN][TT];
    int tn, tt, n, i = 0, j = 0, k = 0;
    bool tvalid;
    long st;
    int s, e;
    if (DIR == 1) {
        tn = threadIdx.x; tt = threadIdx.y;
        s = b.is; e = b.ie; st = 1;
        n = s - 3 - 2 + (int)blockIdx.x * (TN - 4) + tn;
        const int row = (int)blockIdx.y * TT + tt;
        tvalid = row < nj * nk;
        j = j0 + (tvalid ? row % nj : 0);
        k = k0 + (tvalid ? row / nj : 0);
        i = n;
    } else {
        tt = threadIdx.x; tn = threadIdx.y;
        const int ni = b.nx1 + 2;
        const int col = (int)blockIdx.x * TT + tt;
        const int no = (DIR == 2) ? nk : nj;
        tvalid = col < ni * no;
        const int outer = (DIR == 2 ? k0 : j0) + (tvalid ? col / ni : 0);
        i = b.is - 1 + (tvalid ? col % ni : 0);
        if (DIR == 2) { s = b.js; e = b.je; st = b.s2; k = outer; }
        else { s = b.ks; e = b.ke; st = b.s3; j = outer; }
        n = s - 3 - 2 + (int)blockIdx.y * (TN - 4) + tn;
        if (DIR == 2) j = n; else k = n;
    }
    const bool recon = tvalid && n >= s - 1 && n <= e + 1;
    const bool out = tvalid && tn >= 2 && tn < TN - 2 && n >= s - 3 && n <= e + 3;
    const long c = recon || out ? (long)k * b.s3 + (long)j * b.s2 + i : 0;
#pragma unroll 1
    for (int v = 0; v < NV; ++v) {
        double g5[5] = {0., 0., 0., 0., 0.};
        if (recon) {
            const double* q = P + v * b.nc + c;
            // the left face of cell n is face n (exists for n >= s), its right face is n + 1 (n <= e)
            const double ll = (n >= s) ? lamL[v * b.nc + c] : 0.;
            const double lr = (n <= e) ? lamR[v * b.nc + c] : 0.;
            weno5_cell_vjp(q[-2 * st], q[-st], q[0], q[st], q[2 * st], ll, lr, g5);
        }
#pragma unroll
        for (int o = 0; o < 5; ++o) G[o][tn][tt] = g5[o];
        __syncthreads();
        if (out) {
            const double a = G[4][tn - 2][tt] + G[3][tn - 1][tt] + G[2][tn][tt] + G[1][tn + 1][tt] + G[0][tn + 2][tt];
            lamP[v * b.nc + c] += a;
        }
        __syncthreads();
    }
}

// Transpose of the divergence and flux-CT of zone_divergence_ct (ND = 2, 3)
// or of the plain divergence (ND = 1), at every point x of the block:
// non-B face adjoints lamF_d[v][x] = (A_d / V)(ld[v][x] - ld[v][x - s_d]) and the
// corner EMF adjoints; ld (= lam dudt) is zero outside the interior.
template <int ND>
__global__ void __launch_bounds__(ABS)
adj_divct_kernel(BlockDims b, const double* __restrict__ ld, double* __restrict__ lF, double* __restrict__ lemf) {
    const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= b.nc) return;
    const double V = b.dx1 * b.dx2 * b.dx3;
    const double a1 = b.dx2 * b.dx3 / V, a2 = b.dx1 * b.dx3 / V, a3 = b.dx1 * b.dx2 / V;
    const long nc = b.nc, s2 = b.s2, s3 = b.s3;
    auto L = [&](int v, long y) -> double { return (y >= 0 && y < nc) ? ld[v * nc + y] : 0.; };
    const long st[3] = {1, s2, s3};
    const double ad[3] = {a1, a2, a3};
    // variables whose face fluxes enter the divergence directly
#pragma unroll
    for (int d = 0; d < ND; ++d) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double val = 0.;
            const bool direct = (v < B1) || (ND == 1) || (ND == 2 && v == B3);
            if (direct) val = ad[d] * (L(v, x) - L(v, x - st[d]));
            lF[(d * NV + v) * nc + x] = val;
        }
    }
    if constexpr (ND >= 2) {
        // e3 (b_flux_ct.cpp:331-342) enters dudt[B1] through F2 and dudt[B2] through F1
        const double e3 = -a2 * (-0.5 * (L(B1, x - s2) + L(B1, x - s2 - 1)) + 0.5 * (L(B1, x) + L(B1, x - 1)))
                          - a1 * (0.5 * (L(B2, x - 1) + L(B2, x - 1 - s2)) - 0.5 * (L(B2, x) + L(B2, x - s2)));
        lemf[2 * nc + x] = e3;
    }
    if constexpr (ND == 3) {
        const double e1 = -a3 * (-0.5 * (L(B2, x - s3) + L(B2, x - s3 - s2)) + 0.5 * (L(B2, x) + L(B2, x - s2)))
                          - a2 * (0.5 * (L(B3, x - s2) + L(B3, x - s2 - s3)) - 0.5 * (L(B3, x) + L(B3, x - s3)));
        const double e2 = -a3 * (0.5 * (L(B1, x - s3) + L(B1, x - s3 - 1)) - 0.5 * (L(B1, x) + L(B1, x - 1)))
                          - a1 * (-0.5 * (L(B3, x - 1) + L(B3, x - 1 - s3)) + 0.5 * (L(B3, x) + L(B3, x - s3)));
        lemf[x] = e1;
        lemf[nc + x] = e2;
    }
}

// Riemann B-flux adjoints from the corner EMF adjoints (transpose of
// emf{1,2,3}_at): e3(x) = (F1B2[x] + F1B2[x-s2] - F2B1[x] - F2B1[x-1]) / 4, etc.
template <int ND>
__global__ void __launch_bounds__(ABS)
adj_emf_kernel(BlockDims b, const double* __restrict__ lemf, double* __restrict__ lF) {
    const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= b.nc) return;
    const long nc = b.nc, s2 = b.s2, s3 = b.s3;
    auto E = [&](int q, long y) -> double { return (y >= 0 && y < nc) ? lemf[q * nc + y] : 0.; };
    auto F = [&](int d, int v) -> double& { return lF[((d - 1) * NV + v) * nc + x]; };
    F(1, B1) = 0.;
    F(2, B2) = 0.;
    F(1, B2) = 0.25 * (E(2, x) + E(2, x + s2));
    F(2, B1) = -0.25 * (E(2, x) + E(2, x + 1));
    if constexpr (ND == 3) {
        F(3, B3) = 0.;
        F(2, B3) = 0.25 * (E(0, x) + E(0, x + s3));
        F(3, B2) = -0.25 * (E(0, x) + E(0, x + s2));
        F(3, B1) = 0.25 * (E(1, x) + E(1, x + 1));
        F(1, B3) = -0.25 * (E(1, x) + E(1, x + s3));
    }
}

__device__ __forceinline__ void interior_zone(const BlockDims& b, long t, int& k, int& j, int& i) {
    const long nij = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1);
    const long kk = t / nij;
    const long rem = t - kk * nij;
    const long jj = rem / b.nx1;
    i = b.is + (int)(rem - jj * b.nx1);
    j = b.js + (int)jj;
    k = b.ks + (int)kk;
}

// Stage transpose at each interior zone: lU* = (dU/dP)^-T lP_out + lU_out, then
// lam dudt = beta dt lU* and the base-state adjoints (same: U* = U_full).
template <int GEO>
__global__ void __launch_bounds__(ABS)
adj_stage_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ Pout,
                 const double* __restrict__ lPout, const double* __restrict__ lUout, double g0, double g1,
                 double beta_dt, int same, int accumulate, double* lUsub, double* lUfull,
                 double* __restrict__ ld, int* __restrict__ err) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n) return;
    int k, j, i;
    interior_zone(b, t, k, j, i);
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    double P[NV], lP[NV], lU[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { P[v] = Pout[v * b.nc + x]; lP[v] = lPout[v * b.nc + x]; }
    bool ok;
    if constexpr (GEO) {
        CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, (long)j * gd.gn1 + i);
        ok = inverter_vjp(g, P, c.gam, lP, lU);
    } else {
        MinkowskiGeom<double> g;
        ok = inverter_vjp(g, P, c.gam, lP, lU);
    }
    if (!ok) {
        atomicAdd(&err[1], 1);
#pragma unroll
        for (int v = 0; v < NV; ++v) lU[v] = 0.;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double l = lU[v] + (lUout ? lUout[v * b.nc + x] : 0.);
        ld[v * b.nc + x] = beta_dt * l;
        if (same) {
            lUfull[v * b.nc + x] = (accumulate ? lUfull[v * b.nc + x] : 0.) + l;
        } else {
            lUsub[v * b.nc + x] = (accumulate ? lUsub[v * b.nc + x] : 0.) + g0 * l;
            lUfull[v * b.nc + x] = (accumulate ? lUfull[v * b.nc + x] : 0.) + g1 * l;
        }
    }
}

// lamP += (dU/dP)^T lamU at each interior zone (U = p_to_u(P)).
template <int GEO>
__global__ void __launch_bounds__(ABS)
adj_ptou_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ P, const double* __restrict__ lU,
                double* __restrict__ lP) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n) return;
    int k, j, i;
    interior_zone(b, t, k, j, i);
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    double Pz[NV], lu[NV], out[NV];
    const double zero[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
#pragma unroll
    for (int v = 0; v < NV; ++v) { Pz[v] = P[v * b.nc + x]; lu[v] = lU[v * b.nc + x]; }
    if constexpr (GEO) {
        CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, (long)j * gd.gn1 + i);
        side_vjp<0>(g, Pz, c.gam, lu, zero, 0., 0., out);
    } else {
        MinkowskiGeom<double> g;
        side_vjp<0>(g, Pz, c.gam, lu, zero, 0., 0., out);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) lP[v * b.nc + x] += out[v];
}

// Transpose of the periodic ghost fill: each interior zone collects the
// adjoints held by its periodic images (up to 3 per axis offset, corners
// included); ghosts are cleared by adj_clear_ghosts afterwards.
__global__ void __launch_bounds__(ABS) adj_fold_kernel(BlockDims b, double* __restrict__ A, int nvar) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n) return;
    int k, j, i;
    interior_zone(b, t, k, j, i);
    int ni = 1, nj = 1, nk = 1;
    int oi[3] = {0, 0, 0}, oj[3] = {0, 0, 0}, ok[3] = {0, 0, 0};
    if (b.periodic[0]) {
        if (i - b.nx1 >= 0) oi[ni++] = -b.nx1;
        if (i + b.nx1 < b.n1) oi[ni++] = b.nx1;
    }
    if (b.ndim > 1 && b.periodic[1]) {
        if (j - b.nx2 >= 0) oj[nj++] = -b.nx2;
        if (j + b.nx2 < b.n2) oj[nj++] = b.nx2;
    }
    if (b.ndim > 2 && b.periodic[2]) {
        if (k - b.nx3 >= 0) ok[nk++] = -b.nx3;
        if (k + b.nx3 < b.n3) ok[nk++] = b.nx3;
    }
    if (ni == 1 && nj == 1 && nk == 1) return;   // no periodic image
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    for (int v = 0; v < nvar; ++v) {
        double s = 0.;
        for (int a = 0; a < nk; ++a)
            for (int q = 0; q < nj; ++q)
                for (int r = 0; r < ni; ++r) {
                    if (a == 0 && q == 0 && r == 0) continue;
                    s += A[v * b.nc + x + (long)ok[a] * b.s3 + (long)oj[q] * b.s2 + oi[r]];
                }
        A[v * b.nc + x] += s;
    }
}

__global__ void __launch_bounds__(ABS) adj_clear_ghosts_kernel(BlockDims b, double* __restrict__ A, int nvar) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.nc) return;
    const int k = (int)(t / b.s3), j = (int)((t - (long)k * b.s3) / b.s2), i = (int)(t - (long)k * b.s3 - (long)j * b.s2);
    const bool inside = i >= b.is && i <= b.ie && j >= b.js && j <= b.je && k >= b.ks && k <= b.ke;
    if (inside) return;
    for (int v = 0; v < nvar; ++v) A[v * b.nc + t] = 0.;
}

// Count of zones with a non-zero inversion or floor flag (the adjoint assumes
// an unfloored, converged recovery everywhere).
__global__ void __launch_bounds__(ABS) adj_count_flags_kernel(BlockDims b, const int* __restrict__ pflag,
                                                              const int* __restrict__ fflag, int* __restrict__ err) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    int bad = 0;
    if (t < n) {
        int k, j, i;
        interior_zone(b, t, k, j, i);
        const long x = (long)k * b.s3 + (long)j * b.s2 + i;
        bad = (pflag[x] != 0) || (fflag[x] != 0);
    }
    if (bad) atomicAdd(&err[0], 1);
}

long ninterior_l(const BlockDims& b) {
    return (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
}

template <int GEO>
hipError_t launch_adj_flux_t(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P,
                             const double* lF, double* lamL, double* lamR, double* lamP, hipStream_t s) {
    const int jlo = (b.ndim > 1) ? b.js - 1 : b.js, jhi = (b.ndim > 1) ? b.je + 1 : b.je;
    const int klo = (b.ndim > 2) ? b.ks - 1 : b.ks, khi = (b.ndim > 2) ? b.ke + 1 : b.ke;
    const int nj = jhi - jlo + 1, nk = khi - klo + 1;
    for (int dir = 1; dir <= b.ndim; ++dir) {
        const double* lFd = lF + (long)(dir - 1) * NV * b.nc;
        if (dir == 1) {
            const long cells = (long)(b.nx1 + 2) * nj * nk;
            hipLaunchKernelGGL((adj_face_x1_kernel<GEO>), dim3((unsigned)((cells - 1 + 254) / 255)), dim3(256), 0, s,
                               b, c, gd, P, lFd, lamL, lamR, jlo, nj, klo, nk);
            const int nout = b.nx1 + 6;
            hipLaunchKernelGGL((adj_recon_kernel<1>), dim3((unsigned)((nout + 59) / 60), (unsigned)((nj * nk + 3) / 4)),
                               dim3(64, 4), 0, s, b, P, lamL, lamR, lamP, jlo, nj, klo, nk);
        } else if (dir == 2) {
            const int ncol = (b.nx1 + 2) * nk;
            hipLaunchKernelGGL((adj_face_xt_kernel<2, GEO>), dim3((unsigned)((ncol + 63) / 64), (unsigned)((b.nx2 + 1 + 6) / 7)),
                               dim3(64, 8), 0, s, b, c, gd, P, lFd, lamL, lamR, klo, nk, b.js, b.je + 1);
            const int nout = b.nx2 + 6;
            hipLaunchKernelGGL((adj_recon_kernel<2>), dim3((unsigned)((ncol + 63) / 64), (unsigned)((nout + 11) / 12)),
                               dim3(64, 16), 0, s, b, P, lamL, lamR, lamP, jlo, nj, klo, nk);
        } else {
            const int ncol = (b.nx1 + 2) * nj;
            hipLaunchKernelGGL((adj_face_xt_kernel<3, GEO>), dim3((unsigned)((ncol + 63) / 64), (unsigned)((b.nx3 + 1 + 6) / 7)),
                               dim3(64, 8), 0, s, b, c, gd, P, lFd, lamL, lamR, jlo, nj, b.ks, b.ke + 1);
            const int nout = b.nx3 + 6;
            hipLaunchKernelGGL((adj_recon_kernel<3>), dim3((unsigned)((ncol + 63) / 64), (unsigned)((nout + 11) / 12)),
                               dim3(64, 16), 0, s, b, P, lamL, lamR, lamP, jlo, nj, klo, nk);
        }
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_adj_fill(double* a, long n, double v, hipStream_t s) {
    hipLaunchKernelGGL(fill_kernel, agrid(n), dim3(ABS), 0, s, a, n, v);
    return hipGetLastError();
}

hipError_t launch_adj_flux(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P,
                           const double* lF, double* lamLR, double* lamP, hipStream_t s) {
    double* lamL = lamLR;
    double* lamR = lamLR + (long)NV * b.nc;
    if (b.geo) return launch_adj_flux_t<1>(b, c, gd, P, lF, lamL, lamR, lamP, s);
    return launch_adj_flux_t<0>(b, c, gd, P, lF, lamL, lamR, lamP, s);
}

hipError_t launch_adj_divct(const BlockDims& b, const double* ld, double* lF, double* lemf, hipStream_t s) {
    const dim3 g = agrid(b.nc);
    if (b.ndim == 3) {
        hipLaunchKernelGGL(adj_divct_kernel<3>, g, dim3(ABS), 0, s, b, ld, lF, lemf);
        hipLaunchKernelGGL(adj_emf_kernel<3>, g, dim3(ABS), 0, s, b, lemf, lF);
    } else if (b.ndim == 2) {
        hipLaunchKernelGGL(adj_divct_kernel<2>, g, dim3(ABS), 0, s, b, ld, lF, lemf);
        hipLaunchKernelGGL(adj_emf_kernel<2>, g, dim3(ABS), 0, s, b, lemf, lF);
    } else {
        hipLaunchKernelGGL(adj_divct_kernel<1>, g, dim3(ABS), 0, s, b, ld, lF, lemf);
    }
    return hipGetLastError();
}

hipError_t launch_adj_stage(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* Pout,
                            const double* lPout, const double* lUout, double g0, double g1, double beta_dt, bool same,
                            bool accumulate, double* lUsub, double* lUfull, double* ld, int* err, hipStream_t s) {
    const dim3 g = agrid(ninterior_l(b));
    if (b.geo)
        hipLaunchKernelGGL(adj_stage_kernel<1>, g, dim3(ABS), 0, s, b, c, gd, Pout, lPout, lUout, g0, g1, beta_dt,
                           (int)same, (int)accumulate, lUsub, lUfull, ld, err);
    else
        hipLaunchKernelGGL(adj_stage_kernel<0>, g, dim3(ABS), 0, s, b, c, gd, Pout, lPout, lUout, g0, g1, beta_dt,
                           (int)same, (int)accumulate, lUsub, lUfull, ld, err);
    return hipGetLastError();
}

hipError_t launch_adj_ptou(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P,
                           const double* lU, double* lP, hipStream_t s) {
    const dim3 g = agrid(ninterior_l(b));
    if (b.geo) hipLaunchKernelGGL(adj_ptou_kernel<1>, g, dim3(ABS), 0, s, b, c, gd, P, lU, lP);
    else hipLaunchKernelGGL(adj_ptou_kernel<0>, g, dim3(ABS), 0, s, b, c, gd, P, lU, lP);
    return hipGetLastError();
}

hipError_t launch_adj_fold(const BlockDims& b, double* A, int nvar, hipStream_t s) {
    hipLaunchKernelGGL(adj_fold_kernel, agrid(ninterior_l(b)), dim3(ABS), 0, s, b, A, nvar);
    hipLaunchKernelGGL(adj_clear_ghosts_kernel, agrid(b.nc), dim3(ABS), 0, s, b, A, nvar);
    return hipGetLastError();
}

hipError_t launch_adj_count_flags(const BlockDims& b, const int* pflag, const int* fflag, int* err, hipStream_t s) {
    hipLaunchKernelGGL(adj_count_flags_kernel, agrid(ninterior_l(b)), dim3(ABS), 0, s, b, pflag, fflag, err);
    return hipGetLastError();
}

}  // namespace kh
