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
// Face adjoints:  adj_flux_*_fused recompute the face states of a tile of
// cells along one direction, take the adjoint of each face, keep the two face
// adjoints of every cell in LDS and apply the WENO5 transpose, gathering the
// 5-point contributions into lamP; no face-state adjoints go through HBM.
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

// Face adjoints and WENO5 transpose fused (no face-state adjoint arrays), X1.
// Rows (j, k) are laid end to end as padded segments of cells [is-5, ie+5]
// (length nx1 + 10); a workgroup of 256 threads covers 256 consecutive slots
// (from 3 before its 250 output slots), one cell each.  Per slot:
// reconstruct the cell (cells [is-1, ie+1]), take the adjoint of the face at its
// left (faces [is, ie+1]; the left state comes from the neighbour slot through
// LDS), apply the WENO5 transpose with the adjoints of both its faces, and
// gather the 5-point contributions into lamP for the output cells [is-3, ie+3]
// whose whole neighbourhood lies in the workgroup (slots 3 .. 252).
template <int GEO>
__global__ void __launch_bounds__(256)
adj_flux_x1_fused_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ P,
                         const double* __restrict__ lF, double* __restrict__ lamP, int j0, int nj, int k0, int nk) {
    __shared__ double sro[NV][256];   // right face value of each slot's cell, then the adjoint of it
    __shared__ double G[5][256];
    const int Lp = b.nx1 + 10;
    const int tid = threadIdx.x;
    const long g = (long)blockIdx.x * 250 + tid - 3;   // workgroup b outputs slots [250 b, 250 b + 250)
    const long nslots = (long)Lp * nj * nk;
    const bool valid = g >= 0 && g < nslots;
    int m = 0, jf = 0;
    long ci = 0;
    if (valid) {
        const int row = (int)(g / Lp);
        m = b.is - 5 + (int)(g - (long)row * Lp);
        const int jj = row % nj, kk = row / nj;
        jf = j0 + jj;
        ci = (long)(k0 + kk) * b.s3 + (long)jf * b.s2 + m;
    }
    const bool recon = valid && m >= b.is - 1 && m <= b.ie + 1;
    double lo[NV], ll[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double ro = 0.;
        lo[v] = 0.;
        if (recon) {
            const double* q = P + v * b.nc + ci;
            weno5_cell(q[-2], q[-1], q[0], q[1], q[2], lo[v], ro);
        }
        sro[v][tid] = ro;
        ll[v] = 0.;
    }
    __syncthreads();
    // the face at this cell's left: its adjoint splits into lamL (this cell) and
    // lamR (the left neighbour, written back into sro after every read of it)
    const bool face = valid && tid > 0 && m >= b.is && m <= b.ie + 1;
    double lPl[NV];
    if (face) {
        double Pl[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) Pl[v] = sro[v][tid - 1];
        double lf[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) lf[v] = lF[v * b.nc + ci];
        with_face_geom<1, GEO>(c, gd, jf, m, [&](const auto& gg, double rflr, double uflr) {
            face_vjp_llf<1>(gg, c.gam, c.recon_floors != 0, rflr, uflr, Pl, lo, lf, lPl, ll);
        });
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < NV; ++v) sro[v][tid] = 0.;
    __syncthreads();
    if (face) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sro[v][tid - 1] = lPl[v];
    }
    __syncthreads();
    const bool out = valid && tid >= 3 && tid < 253 && m >= b.is - 3 && m <= b.ie + 3;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double g5[5] = {0., 0., 0., 0., 0.};
        if (recon) {
            const double* q = P + v * b.nc + ci;
            weno5_cell_vjp(q[-2], q[-1], q[0], q[1], q[2], ll[v], sro[v][tid], g5);
        }
#pragma unroll
        for (int o = 0; o < 5; ++o) G[o][tid] = g5[o];
        __syncthreads();
        if (out) {
            const double a = G[4][tid - 2] + G[3][tid - 1] + G[2][tid] + G[1][tid + 1] + G[0][tid + 2];
            lamP[v * b.nc + ci] += a;
        }
        __syncthreads();
    }
}

// The same for X2 / X3: tiles of TX transverse columns (outer, i) x TY slots
// along DIR.  Tile by covers the cells s - 6 + by (TY - 6) + [0, TY) and
// outputs its slots 3 .. TY - 4; the outputs of a column are the cells
// [s-3, e+3], its valid slots [s-5, e+5].  Flat space: 6 x 41 (35 outputs a
// tile: two tiles for 64 cells, four for 128, 246 threads); curved space,
// whose face adjoint needs more registers: 8 x 32.
template <int GEO> constexpr int adj_xt_tx() { return GEO ? 8 : 6; }
template <int GEO> constexpr int adj_xt_ty() { return GEO ? 32 : 41; }
template <int DIR, int GEO>
__global__ void __launch_bounds__(256)
adj_flux_xt_fused_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ P,
                         const double* __restrict__ lF, double* __restrict__ lamP, int o0, int no) {
    constexpr int TX = adj_xt_tx<GEO>(), TY = adj_xt_ty<GEO>();
    __shared__ double sro[NV][TY][TX];
    __shared__ double G[5][TY][TX];
    const int ni = b.nx1 + 2;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int col = blockIdx.x * TX + tx;
    const int s0 = (DIR == 2) ? b.js : b.ks, e0 = (DIR == 2) ? b.je : b.ke;
    const int n = s0 - 6 + (int)blockIdx.y * (TY - 6) + ty;
    const bool valid = col < ni * no && n >= s0 - 5 && n <= e0 + 5;
    const long st = (DIR == 2) ? b.s2 : b.s3;
    int jf = 0, if_ = 0;
    long ci = 0;
    if (valid) {
        const int outer = o0 + col / ni;
        const int i = b.is - 1 + col % ni;
        const int j = (DIR == 2) ? n : outer;
        const int k = (DIR == 2) ? outer : n;
        jf = j; if_ = i;
        ci = (long)k * b.s3 + (long)j * b.s2 + i;
    }
    const bool recon = valid && n >= s0 - 1 && n <= e0 + 1;
    double lo[NV], ll[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double ro = 0.;
        lo[v] = 0.;
        if (recon) {
            const double* q = P + v * b.nc + ci;
            weno5_cell(q[-2 * st], q[-st], q[0], q[st], q[2 * st], lo[v], ro);
        }
        sro[v][ty][tx] = ro;
        ll[v] = 0.;
    }
    __syncthreads();
    const bool face = valid && ty > 0 && n >= s0 && n <= e0 + 1;
    double lPl[NV];
    if (face) {
        double Pl[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) Pl[v] = sro[v][ty - 1][tx];
        double lf[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) lf[v] = lF[v * b.nc + ci];
        with_face_geom<DIR, GEO>(c, gd, jf, if_, [&](const auto& gg, double rflr, double uflr) {
            face_vjp_llf<DIR>(gg, c.gam, c.recon_floors != 0, rflr, uflr, Pl, lo, lf, lPl, ll);
        });
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < NV; ++v) sro[v][ty][tx] = 0.;
    __syncthreads();
    if (face) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sro[v][ty - 1][tx] = lPl[v];
    }
    __syncthreads();
    const bool out = valid && ty >= 3 && ty < TY - 3 && n >= s0 - 3 && n <= e0 + 3;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double g5[5] = {0., 0., 0., 0., 0.};
        if (recon) {
            const double* q = P + v * b.nc + ci;
            weno5_cell_vjp(q[-2 * st], q[-st], q[0], q[st], q[2 * st], ll[v], sro[v][ty][tx], g5);
        }
#pragma unroll
        for (int o = 0; o < 5; ++o) G[o][ty][tx] = g5[o];
        __syncthreads();
        if (out) {
            const double a = G[4][ty - 2][tx] + G[3][ty - 1][tx] + G[2][ty][tx] + G[1][ty + 1][tx] + G[0][ty + 2][tx];
            lamP[v * b.nc + ci] += a;
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
adj_divct_kernel(BlockDims b, const double* __restrict__ ld, double* __restrict__ lF, double* __restrict__ lemf,
                 FluxSeed seed) {
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
    if (seed.dir) {
        // a face-flux objective's seed (adj_seed_kernel's faces, fluid variables)
        const int k = (int)(x / s3), j = (int)((x - (long)k * s3) / s2), i = (int)(x - (long)k * s3 - (long)j * s2);
        const int along = (seed.dir == 1) ? i : (seed.dir == 2) ? j : k;
        const bool t1 = (seed.dir == 1) || (i >= b.is && i <= b.ie);
        const bool t2 = (seed.dir == 2) || (b.ndim < 2) || (j >= b.js && j <= b.je);
        const bool t3 = (seed.dir == 3) || (b.ndim < 3) || (k >= b.ks && k <= b.ke);
        if (along == seed.face && t1 && t2 && t3) lF[((seed.dir - 1) * NV + seed.var) * nc + x] += seed.w;
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

// KDomain::GetPhysicalRange membership (domain.hpp:161-187), as the forward fixup uses it.
__device__ __forceinline__ bool phys_zone(const BlockDims& b, int k, int j, int i) {
    if (i < 0 || i >= b.n1 || j < 0 || j >= b.n2 || k < 0 || k >= b.n3) return false;
    if (!b.periodic[0] && (i < b.is || i > b.ie)) return false;
    if (b.ndim > 1 && !b.periodic[1] && (j < b.js || j > b.je)) return false;
    if (b.ndim > 2 && !b.periodic[2] && (k < b.ks || k > b.ke)) return false;
    return true;
}

// Transpose of fixup_list_kernel (Inverter::FixUtoP, inverter/fixup.cpp:47-156)
// for the fluid primitives, gathered at each zone y of the physical range: a
// zone f that failed to invert took P_f = geo_floors(sum_m w_fm P_m / sum_m w_fm)
// over its non-failed neighbours m (w = 1 / (|l| + |m| + |n| + 1)), and
// U_f = p_to_u(P_f).  y receives (w_fy / W_f) lam_f for each failed interior
// neighbour f, with lam_f = lP_out(f) + (dU/dP at P_f)^T lU_out(f) (fluid part,
// zero where the geometric floor set the value).  lPfx is written on every
// zone (periodic images are folded afterwards).
template <int GEO>
__global__ void __launch_bounds__(ABS)
adj_fixup_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ Pout,
                 const double* __restrict__ lPout, const double* __restrict__ lUout, const int* __restrict__ pf,
                 const int* __restrict__ gate, double* __restrict__ lPfx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.nc || (gate && *gate == 0)) return;   // no failed inversion in the stage
    const int k = (int)(t / b.s3), j = (int)((t - (long)k * b.s3) / b.s2), i = (int)(t - (long)k * b.s3 - (long)j * b.s2);
    double acc[5] = {0., 0., 0., 0., 0.};
    const int nd = b.ndim;
    if (c.fix_average && phys_zone(b, k, j, i) && !(pf[t] > 0)) {
        for (int dn = (nd > 2 ? -1 : 0); dn <= (nd > 2 ? 1 : 0); ++dn)
            for (int dm = (nd > 1 ? -1 : 0); dm <= (nd > 1 ? 1 : 0); ++dm)
                for (int dl = -1; dl <= 1; ++dl) {
                    const int fk = k + dn, fj = j + dm, fi = i + dl;
                    if (fi < b.is || fi > b.ie || fj < b.js || fj > b.je || fk < b.ks || fk > b.ke) continue;
                    const long f = (long)fk * b.s3 + (long)fj * b.s2 + fi;
                    if (!(pf[f] > 0)) continue;
                    // the failed zone's average over its non-failed neighbours
                    double wsum = 0., sum[2] = {0., 0.};
                    for (int en = -1; en <= 1; ++en)
                        for (int em = -1; em <= 1; ++em)
                            for (int el = -1; el <= 1; ++el) {
                                const int mk = fk + en, mj = fj + em, mi = fi + el;
                                if (!phys_zone(b, mk, mj, mi)) continue;
                                const long m = (long)mk * b.s3 + (long)mj * b.s2 + mi;
                                if (pf[m] > 0) continue;
                                const double w = 1. / (abs(el) + abs(em) + abs(en) + 1);
                                wsum += w;
                                sum[0] += w * Pout[RHO * b.nc + m];
                                sum[1] += w * Pout[UU * b.nc + m];
                            }
                    if (wsum < 1.e-10) continue;   // atmosphere fallback: no dependence
                    double rf, uf;
                    if constexpr (GEO) {
                        const long q = (long)fj * gd.gn1 + fi;
                        rf = gd.flr[q]; uf = gd.flr[gd.np + q];
                    } else {
                        rf = c.floors.rho_min_const; uf = c.floors.u_min_const;
                    }
                    double Pf[NV], lu[NV], lt[NV];
                    const double zero[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
#pragma unroll
                    for (int v = 0; v < NV; ++v) { Pf[v] = Pout[v * b.nc + f]; lu[v] = lUout ? lUout[v * b.nc + f] : 0.; }
                    if constexpr (GEO) {
                        CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, (long)fj * gd.gn1 + fi);
                        side_vjp<0>(g, Pf, c.gam, lu, zero, 0., 0., lt);
                    } else {
                        MinkowskiGeom<double> g;
                        side_vjp<0>(g, Pf, c.gam, lu, zero, 0., 0., lt);
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) lt[v] += lPout[v * b.nc + f];
                    if (rf - sum[0] / wsum > 0.) lt[RHO] = 0.;
                    if (uf - sum[1] / wsum > 0.) lt[UU] = 0.;
                    const double cw = 1. / ((abs(dl) + abs(dm) + abs(dn) + 1) * wsum);
#pragma unroll
                    for (int v = 0; v < 5; ++v) acc[v] += cw * lt[v];
                }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) lPfx[v * b.nc + t] = (v < 5) ? acc[v] : 0.;
}

// Floors of the stage transpose, first half (one thread per floored zone):
// P_inv = u_to_p(U*) recovered again from the stage's U* and guess Psub as the
// forward update does, and lPtot = lP_out + (dU/dP at P_out)^T lU_out (+ the
// fixup transpose lPfx when the stage had failed inversions), into Pinv, lPtot.
template <int GEO>
__global__ void __launch_bounds__(ABS)
adj_floors_prep_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ Pout,
                       const double* __restrict__ lPout, const double* __restrict__ lUout,
                       const double* __restrict__ Ustar, const double* __restrict__ Psub, const int* __restrict__ ffs,
                       const int* __restrict__ pfs, const int* __restrict__ gate, const double* __restrict__ lPfx,
                       double* __restrict__ Pinv, double* __restrict__ lPtot, int* __restrict__ err) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n) return;
    int k, j, i;
    interior_zone(b, t, k, j, i);
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    if (ffs[x] == 0 || (pfs && pfs[x] > 0)) return;
    const bool fx = lPfx && gate && *gate > 0;
    double P[NV], lP[NV], lUo[NV], Us[NV], Pi[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        P[v] = Pout[v * b.nc + x];
        lP[v] = lPout[v * b.nc + x] + (fx ? lPfx[v * b.nc + x] : 0.);
        lUo[v] = lUout ? lUout[v * b.nc + x] : 0.;
        Us[v] = Ustar[v * b.nc + x];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) Pi[v] = Psub[v * b.nc + x];
    auto body = [&](const auto& g) {
        double t8[NV];
        const double zero[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
        side_vjp<0>(g, P, c.gam, lUo, zero, 0., 0., t8);
        const int st = u_to_p_onedw(g, Us, Pi, c.gam, c.iter_max, c.err_tol);
        const double igd = rcp(g.gdet());
        Pi[B1] = Us[B1] * igd; Pi[B2] = Us[B2] * igd; Pi[B3] = Us[B3] * igd;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Pinv[v * b.nc + x] = Pi[v];
            lPtot[v * b.nc + x] = lP[v] + t8[v];
        }
        if (st % F_GEOM_RHO != 0) atomicAdd(&err[1], 1);
    };
    if constexpr (GEO) {
        CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, (long)j * gd.gn1 + i);
        body(g);
    } else {
        MinkowskiGeom<double> g;
        body(g);
    }
}

// Second half, eight threads per floored zone (one per Jacobian column):
// lPinv = (dfloors/dP at P_inv)^T lPtot.
template <int GEO, int FRAME>
__global__ void __launch_bounds__(ABS)
adj_floors_cols_kernel(BlockDims b, StepConsts c, GeomDev gd, const int* __restrict__ ffs, const int* __restrict__ pfs,
                       const double* __restrict__ Pinv, const double* __restrict__ lPtot, double* __restrict__ lPinv) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    const long z = t >> 3;
    const int col = (int)(t & 7);
    if (z >= n) return;
    int k, j, i;
    interior_zone(b, z, k, j, i);
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    if (ffs[x] == 0 || (pfs && pfs[x] > 0)) return;
    double Pi[NV], lP[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { Pi[v] = Pinv[v * b.nc + x]; lP[v] = lPtot[v * b.nc + x]; }
    if constexpr (GEO) {
        const long q = (long)j * gd.gn1 + i;
        CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, q);
        lPinv[col * b.nc + x] = floors_vjp_col<FRAME>(g, Pi, c.gam, c.floors, gd.flr[q], gd.flr[gd.np + q], lP, col);
    } else {
        MinkowskiGeom<double> g;
        lPinv[col * b.nc + x] =
            floors_vjp_col<FRAME>(g, Pi, c.gam, c.floors, c.floors.rho_min_const, c.floors.u_min_const, lP, col);
    }
}

// Stage transpose at each interior zone.  Unfloored zones: lU* = (dU/dP)^-T
// lP_out + lU_out at P_out (U_out = p_to_u(P_out) and P_out = u_to_p(U*)).
// Zones the stage floored (ffs != 0; FRAME >= 0): lU* = (dU/dP at P_inv)^-T
// lP_inv from adj_floors_prep_kernel's P_inv and lP_inv.  Zones that failed to
// invert (pfs > 0) pass only their B adjoint to U*.  With failed inversions in
// the stage (*gate > 0) every zone first adds the fixup transpose lPfx.  Then
// lam dudt = beta dt lU* and the
// base-state adjoints (same: U* = U_full).  With lPgeo (curved geometry), the
// geometric source's transpose at the stage input Psub is written (not added)
// to lPgeo at the zone.
template <int GEO, int FRAME>
__global__ void __launch_bounds__(ABS)
adj_stage_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ Pout,
                 const double* __restrict__ lPout, const double* __restrict__ lUout, double g0, double g1,
                 double beta_dt, int same, int accumulate, double* lUsub, double* lUfull,
                 double* __restrict__ ld, int* __restrict__ err, const double* __restrict__ Pinv,
                 const double* __restrict__ lPinv, const double* __restrict__ Psub, const int* __restrict__ ffs,
                 double* __restrict__ lPgeo, const int* __restrict__ pfs, const int* __restrict__ gate,
                 const double* __restrict__ lPfx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n) return;
    int k, j, i;
    interior_zone(b, t, k, j, i);
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    double P[NV], lP[NV], lU[NV], lUo[NV];
    const bool fx = lPfx && gate && *gate > 0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        P[v] = Pout[v * b.nc + x];
        lP[v] = lPout[v * b.nc + x] + (fx ? lPfx[v * b.nc + x] : 0.);
        lUo[v] = lUout ? lUout[v * b.nc + x] : 0.;
    }
    const bool failed = pfs && pfs[x] > 0;
    bool ok = true;
    auto body = [&](const auto& g) {
        if (failed) {
            // fixed up from its neighbours (adj_fixup_kernel carries the fluid part);
            // only P.B = U*.B / gdet depends on this zone's U*
            double t8[NV];
            const double zero[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
            side_vjp<0>(g, P, c.gam, lUo, zero, 0., 0., t8);
            const double igd = rcp(g.gdet());
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                lU[v] = (v >= B1) ? (lP[v] + t8[v]) * igd : 0.;
                lUo[v] = 0.;
            }
            return;
        }
        if constexpr (FRAME >= 0) {
            if (ffs[x] != 0) {
                // floors carried through by adj_floors_prep_kernel: lP_inv at P_inv
                double Pi[NV], lPi[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) { Pi[v] = Pinv[v * b.nc + x]; lPi[v] = lPinv[v * b.nc + x]; }
                ok = inverter_vjp(g, Pi, c.gam, lPi, lU);
#pragma unroll
                for (int v = 0; v < NV; ++v) lUo[v] = 0.;   // already carried through P_out
                return;
            }
        }
        ok = inverter_vjp(g, P, c.gam, lP, lU);
    };
    if constexpr (GEO) {
        CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, (long)j * gd.gn1 + i);
        body(g);
    } else {
        MinkowskiGeom<double> g;
        body(g);
    }
    if (!ok) {
        atomicAdd(&err[1], 1);
#pragma unroll
        for (int v = 0; v < NV; ++v) lU[v] = 0.;
    }
    double lS[4];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double l = lU[v] + lUo[v];
        ld[v * b.nc + x] = beta_dt * l;
        if (v >= UU && v <= U3) lS[v - UU] = beta_dt * l;
        if (same) {
            lUfull[v * b.nc + x] = (accumulate ? lUfull[v * b.nc + x] : 0.) + l;
        } else {
            lUsub[v * b.nc + x] = (accumulate ? lUsub[v * b.nc + x] : 0.) + g0 * l;
            lUfull[v * b.nc + x] = (accumulate ? lUfull[v * b.nc + x] : 0.) + g1 * l;
        }
    }
    if constexpr (GEO) {
        if (lPgeo) {
            double Pz[NV], out[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
#pragma unroll
            for (int v = 0; v < NV; ++v) Pz[v] = Psub[v * b.nc + x];
            const long q = (long)j * gd.gn1 + i;
            CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, q);
            geo_source_vjp(g, gd, q, Pz, c.gam, lS, out);
#pragma unroll
            for (int v = 0; v < NV; ++v) lPgeo[v * b.nc + x] = out[v];
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

// ------------------------------------------------ physical boundaries
// The interior zone a physical ghost zone (j, i) is filled from, and how
// (bounds_kernel, kh_step.hip; KBoundaries::ApplyBoundary, boundaries.cpp:369-589):
// P_ghost[0..4] = post(inflow(pre(P_src[0..4]))) and U_ghost.B = pre.post(U_src.B),
// P_ghost.B = U_ghost.B / gdet(ghost), with pre / post the reflections (bit 0:
// U1, B1 negated; bit 1: U2, B2) before and after the inflow check, which runs
// (inflow = 1 inner, 2 outer side) with the geometry of zone (gj, gi).
struct GhostSrc { int j, i, pre, post, inflow, gj, gi; };
__device__ __forceinline__ bool phys_ghost(const BlockDims& b, int j, int i) {
    const bool x1g = !b.periodic[0] && (i < b.is || i > b.ie);
    const bool x2g = b.ndim > 1 && !b.periodic[1] && (j < b.js || j > b.je);
    return x1g || x2g;
}
__device__ __forceinline__ GhostSrc ghost_source(const BlockDims& b, int j, int i) {
    const bool x1g = !b.periodic[0] && (i < b.is || i > b.ie);
    const bool x2g = b.ndim > 1 && !b.periodic[1] && (j < b.js || j > b.je);
    const bool in1 = i < b.is, in2 = j < b.js;
    const bool refl1 = b.bc[in1 ? 0 : 1] == KH_BC_REFLECT, refl2 = b.bc[in2 ? 2 : 3] == KH_BC_REFLECT;
    const int si = refl1 ? (in1 ? 2 * b.is - 1 - i : 2 * b.ie + 1 - i) : (in1 ? b.is : b.ie);
    const int sj = refl2 ? (in2 ? 2 * b.js - 1 - j : 2 * b.je + 1 - j) : (in2 ? b.js : b.je);
    const int infl = b.check_inflow[in1 ? 0 : 1] ? (in1 ? 1 : 2) : 0;
    GhostSrc r;
    if (x1g && !x2g) {
        r = {j, si, refl1 ? 1 : 0, 0, infl, j, i};
    } else if (x2g && !x1g) {
        r = {sj, i, refl2 ? 2 : 0, 0, 0, j, i};
    } else if (in1 && b.fix_corner_inner) {
        // the X2 fill of column is (or of the mirror column), then the inner X1 face over it
        const bool r1 = b.bc[0] == KH_BC_REFLECT;
        r = {sj, r1 ? 2 * b.is - 1 - i : b.is, (refl2 ? 2 : 0) | (r1 ? 1 : 0), 0, b.check_inflow[0] ? 1 : 0, j, i};
    } else {
        // the X1 fill of the mirror / edge row (inflow check with that row's geometry), then the X2 reflection
        r = {sj, si, refl1 ? 1 : 0, refl2 ? 2 : 0, infl, sj, i};
    }
    return r;
}

// Transpose of the physical ghost fill, in two passes.  adj_bounds_map_kernel,
// one thread per physical ghost zone (the slabs of bounds_kernel), maps the
// ghost's adjoint back through its fill to an adjoint of its source zone and
// stores it at the ghost's own position of C; adj_bounds_kernel then gathers,
// at each zone within ng of a physical face, the C entries of the ghosts filled
// from it.  Ghosts are cleared afterwards (adj_fold).
__host__ __device__ inline long ghost_slab1(const BlockDims& b) {
    return b.periodic[0] ? 0 : (long)b.n3 * b.n2 * 2 * b.ng;
}
__host__ __device__ inline long ghost_slab2(const BlockDims& b) {
    if (b.ndim < 2 || b.periodic[1]) return 0;
    return (long)b.n3 * 2 * b.ng * (b.periodic[0] ? b.n1 : b.nx1);
}
template <int GEO>
__global__ void __launch_bounds__(ABS)
adj_bounds_map_kernel(BlockDims b, GeomDev gd, const double* __restrict__ P, const double* __restrict__ A,
                      double* __restrict__ C) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n1s = ghost_slab1(b);
    int k, j, i;
    if (r < n1s) {
        const int g = (int)(r % (2 * b.ng));
        r /= 2 * b.ng;
        j = (int)(r % b.n2);
        k = (int)(r / b.n2);
        i = (g < b.ng) ? g : b.ie + 1 + (g - b.ng);
    } else {
        r -= n1s;
        if (r >= ghost_slab2(b)) return;
        const int w = b.periodic[0] ? b.n1 : b.nx1;
        const int ii = (int)(r % w);
        r /= w;
        const int g = (int)(r % (2 * b.ng));
        k = (int)(r / (2 * b.ng));
        j = (g < b.ng) ? g : b.je + 1 + (g - b.ng);
        i = (b.periodic[0] ? 0 : b.is) + ii;
    }
    const long gx = (long)k * b.s3 + (long)j * b.s2 + i;
    const GhostSrc m = ghost_source(b, j, i);
    const long sx = (long)k * b.s3 + (long)m.j * b.s2 + m.i;
    double l[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) l[v] = A[v * b.nc + gx];
    // B: P_g.B = flip(U_src.B) / gdet_g = flip(P_src.B) gdet_src / gdet_g
    double ratio = 1.;
    if constexpr (GEO) ratio = gd.g[GLOC_CENTER][20 * gd.np + (long)m.j * gd.gn1 + m.i] /
                               gd.g[GLOC_CENTER][20 * gd.np + (long)j * gd.gn1 + i];
    const int fl = m.pre ^ m.post;
    double c[NV];
    c[B1] = ((fl & 1) ? -l[B1] : l[B1]) * ratio;
    c[B2] = ((fl & 2) ? -l[B2] : l[B2]) * ratio;
    c[B3] = l[B3] * ratio;
    // fluid: post flips, inflow transpose at the pre-flipped source state, pre flips
    if (m.post & 1) l[U1] = -l[U1];
    if (m.post & 2) l[U2] = -l[U2];
    if (m.inflow) {
        double s1[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) s1[v] = P[v * b.nc + sx];
        if (m.pre & 1) s1[U1] = -s1[U1];
        if (m.pre & 2) s1[U2] = -s1[U2];
        if constexpr (GEO) {
            CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, (long)m.gj * gd.gn1 + m.gi);
            inflow_vjp(g, s1, m.inflow == 1, l);
        } else {
            MinkowskiGeom<double> g;
            inflow_vjp(g, s1, m.inflow == 1, l);
        }
    }
    if (m.pre & 1) l[U1] = -l[U1];
    if (m.pre & 2) l[U2] = -l[U2];
#pragma unroll
    for (int v = 0; v < 5; ++v) c[v] = l[v];
#pragma unroll
    for (int v = 0; v < NV; ++v) C[v * b.nc + gx] = c[v];
}

// Zones within ng of a physical face, enumerated: band 1 is the 2 ng interior
// columns next to the X1 faces (every j, k) when X1 is physical; band 2 the
// 2 ng interior rows next to the X2 faces over the other columns when X2 is.
__host__ __device__ inline int x1_band_cols(const BlockDims& b) { return b.periodic[0] ? 0 : 2 * b.ng; }
__host__ __device__ inline long bounds_band1(const BlockDims& b) { return (long)b.n3 * b.n2 * x1_band_cols(b); }
__host__ __device__ inline long bounds_band2(const BlockDims& b) {
    if (b.ndim < 2 || b.periodic[1]) return 0;
    return (long)b.n3 * 2 * b.ng * (b.n1 - x1_band_cols(b));
}
__global__ void __launch_bounds__(ABS)
adj_bounds_kernel(BlockDims b, const double* __restrict__ C, double* __restrict__ A) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nb1 = bounds_band1(b);
    int k, j, i;
    if (r < nb1) {
        const int c = (int)(r % x1_band_cols(b));
        r /= x1_band_cols(b);
        j = (int)(r % b.n2);
        k = (int)(r / b.n2);
        i = (c < b.ng) ? b.is + c : b.ie - (c - b.ng);
    } else {
        r -= nb1;
        if (r >= bounds_band2(b)) return;
        const int w = b.n1 - x1_band_cols(b);
        int c = (int)(r % w);
        r /= w;
        const int g = (int)(r % (2 * b.ng));
        k = (int)(r / (2 * b.ng));
        j = (g < b.ng) ? b.js + g : b.je - (g - b.ng);
        // the columns outside band 1, in order
        if (!b.periodic[0] && c >= b.is) c += b.ng;                 // skip is .. is+ng-1
        if (!b.periodic[0] && c > b.ie - b.ng) c += b.ng;           // skip ie-ng+1 .. ie
        i = c;
    }
    const long t = (long)k * b.s3 + (long)j * b.s2 + i;
    if (phys_ghost(b, j, i)) return;
    const bool p1 = !b.periodic[0], p2 = b.ndim > 1 && !b.periodic[1];
    const bool near1 = p1 && (i < b.is + b.ng || i > b.ie - b.ng);
    const bool near2 = p2 && (j < b.js + b.ng || j > b.je - b.ng);
    if (!near1 && !near2) return;
    // candidate ghosts: this row / column and the ghost columns / rows whose fill
    // reads this column / row (every composite fill reads its source column from
    // the X1 rule and its source row from the X2 rule, or column is for the
    // inner corner fix; ghost_source confirms each candidate)
    int ic[9], jc[9], ni = 0, nj = 0;
    ic[ni++] = i;
    jc[nj++] = j;
    if (near1)
        for (int g = 0; g < 2 * b.ng; ++g) {
            const int gi = (g < b.ng) ? g : b.ie + 1 + (g - b.ng);
            const bool in1 = gi < b.is;
            const bool refl1 = b.bc[in1 ? 0 : 1] == KH_BC_REFLECT;
            const int si = refl1 ? (in1 ? 2 * b.is - 1 - gi : 2 * b.ie + 1 - gi) : (in1 ? b.is : b.ie);
            if (si == i) ic[ni++] = gi;
        }
    if (near2)
        for (int g = 0; g < 2 * b.ng; ++g) {
            const int gj = (g < b.ng) ? g : b.je + 1 + (g - b.ng);
            const bool in2 = gj < b.js;
            const bool refl2 = b.bc[in2 ? 2 : 3] == KH_BC_REFLECT;
            const int sj = refl2 ? (in2 ? 2 * b.js - 1 - gj : 2 * b.je + 1 - gj) : (in2 ? b.js : b.je);
            if (sj == j) jc[nj++] = gj;
        }
    if (ni == 1 && nj == 1) return;
    double acc[NV] = {0., 0., 0., 0., 0., 0., 0., 0.};
    bool any = false;
    for (int qj = 0; qj < nj; ++qj)
        for (int qi = 0; qi < ni; ++qi) {
            const int gj = jc[qj], gi = ic[qi];
            if (!phys_ghost(b, gj, gi)) continue;
            const GhostSrc m = ghost_source(b, gj, gi);
            if (m.j != j || m.i != i) continue;
            any = true;
            const long gx = (long)k * b.s3 + (long)gj * b.s2 + gi;
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += C[v * b.nc + gx];
        }
    if (!any) return;
#pragma unroll
    for (int v = 0; v < NV; ++v) A[v * b.nc + t] += acc[v];
}

// Transpose of fix_flux_kernel (kh_step.hip) on the flux adjoints lF, in
// place: zeroed polar fluxes get zero adjoint, the mirrored polar B2 fluxes pass
// theirs (negated) to the mirrored source row, and the X1 mass-flux clamp
// passes the adjoint where the clamped flux is the Riemann flux.  x1f holds the
// clamped X1 mass fluxes of the forward stage (save_x1_flux_kernel).
__global__ void __launch_bounds__(ABS) adj_fix_flux_kernel(BlockDims b, double* __restrict__ lF,
                                                           const double* __restrict__ x1f) {
    const int nd = b.ndim;
    const int ni = b.nx1 + 2, nk = (nd > 2) ? b.nx3 + 2 : 1, nj = (nd > 1) ? b.nx2 + 2 : 1;
    const int i0 = b.is - 1, k0 = (nd > 2) ? b.ks - 1 : b.ks, j0 = (nd > 1) ? b.js - 1 : b.js;
    const long nc = b.nc;
    double* F1 = lF;
    double* F2 = lF + 8 * nc;
    double* F3 = lF + 16 * nc;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned npolar = (nd > 1 && !b.periodic[1]) ? 2u * ni * nk : 0u;
    if (t < npolar) {
        const int side = t / (ni * nk), rem = t - side * (ni * nk);
        const int k = k0 + rem / ni, i = i0 + rem % ni;
        const int j = side == 0 ? b.js : b.je + 1;
        const long c = (long)k * b.s3 + (long)j * b.s2 + i;
        const long cg = side == 0 ? c - b.s2 : c;
        const long cs = side == 0 ? c : c - b.s2;
        if (b.zero_polar_flux)
            for (int v = 0; v < NV; ++v) F2[v * nc + c] = 0.;
        if (b.fix_polar_flux) {
            F1[B2 * nc + cs] -= F1[B2 * nc + cg];
            F1[B2 * nc + cg] = 0.;
            if (nd > 2) {
                F3[B2 * nc + cs] -= F3[B2 * nc + cg];
                F3[B2 * nc + cg] = 0.;
            }
            F2[B1 * nc + c] = 0.;
            F2[B3 * nc + c] = 0.;
        }
        return;
    }
    const unsigned t2 = t - npolar;
    if (!b.periodic[0] && t2 < 2u * nj * nk) {
        const int side = t2 / (nj * nk), rem = t2 - side * (nj * nk);
        const int k = k0 + rem / nj, j = j0 + rem % nj;
        if (!b.check_inflow[side]) return;
        const int i = side == 0 ? b.is : b.ie + 1;
        const long c = (long)k * b.s3 + (long)j * b.s2 + i;
        const double f = x1f[t2];
        const bool pass = side == 0 ? (f < 0.) : (f > 0.);
        if (!pass) F1[RHO * nc + c] = 0.;
    }
}

// The clamped X1 mass fluxes of fix_flux_kernel's X1 rows (same enumeration).
__global__ void __launch_bounds__(ABS) save_x1_flux_kernel(BlockDims b, const double* __restrict__ F,
                                                           double* __restrict__ x1f) {
    const int nd = b.ndim;
    const int nk = (nd > 2) ? b.nx3 + 2 : 1, nj = (nd > 1) ? b.nx2 + 2 : 1;
    const int k0 = (nd > 2) ? b.ks - 1 : b.ks, j0 = (nd > 1) ? b.js - 1 : b.js;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2u * nj * nk) return;
    const int side = t / (nj * nk), rem = t - side * (nj * nk);
    const int k = k0 + rem / nj, j = j0 + rem % nj;
    const int i = side == 0 ? b.is : b.ie + 1;
    x1f[t] = F[RHO * b.nc + (long)k * b.s3 + (long)j * b.s2 + i];
}

// Adjoint seed of a face-flux objective: lF_dir[var] += w on the faces of
// index `face` along dir, over the interior transverse range.
__global__ void __launch_bounds__(ABS) adj_seed_kernel(BlockDims b, int dir, int var, int face, double w,
                                                       double* __restrict__ lF) {
    const int na = (dir == 1) ? (b.ndim > 1 ? b.nx2 : 1) : b.nx1;
    const int nb = (dir == 3) ? (b.ndim > 1 ? b.nx2 : 1) : (b.ndim > 2 ? b.nx3 : 1);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int ta = t % na, tb = t / na;
    int i, j, k;
    if (dir == 1) { i = face; j = b.js + ta; k = b.ks + tb; }
    else if (dir == 2) { i = b.is + ta; j = face; k = b.ks + tb; }
    else { i = b.is + ta; k = face; j = b.js + tb; }
    lF[((long)(dir - 1) * NV + var) * b.nc + (long)k * b.s3 + (long)j * b.s2 + i] += w;
}

// Transpose of the periodic ghost fill: each interior zone collects the
// adjoints held by its periodic images (up to 3 per axis offset, corners
// included); ghosts are cleared by adj_clear_ghosts afterwards.
// The interior zones that have periodic images: those within ng of a periodic
// face, enumerated by the outermost axis at which they are (k, then j, then i).
// Valid when every periodic axis holds at least 2 ng zones (else the two edges
// overlap and adj_fold_kernel walks the whole interior).
struct FoldShell {
    int p0, p1, p2;             // periodic axes with images
    long n1, n2, n3;            // zones in regions 1 (k edge), 2 (j edge), 3 (i edge)
};
inline FoldShell fold_shell(const BlockDims& b) {
    FoldShell f;
    f.p0 = b.periodic[0];
    f.p1 = b.ndim > 1 && b.periodic[1];
    f.p2 = b.ndim > 2 && b.periodic[2];
    const long nx1 = b.nx1, nx2 = b.ndim > 1 ? b.nx2 : 1, nx3 = b.ndim > 2 ? b.nx3 : 1;
    const long mk = nx3 - (f.p2 ? 2 * b.ng : 0), mj = nx2 - (f.p1 ? 2 * b.ng : 0);
    f.n1 = f.p2 ? 2L * b.ng * nx2 * nx1 : 0;
    f.n2 = f.p1 ? mk * 2L * b.ng * nx1 : 0;
    f.n3 = f.p0 ? mk * mj * 2L * b.ng : 0;
    return f;
}
inline bool fold_shell_ok(const BlockDims& b) {
    const FoldShell f = fold_shell(b);
    return (!f.p0 || b.nx1 >= 2 * b.ng) && (!f.p1 || b.nx2 >= 2 * b.ng) && (!f.p2 || b.nx3 >= 2 * b.ng);
}
__device__ __forceinline__ void fold_shell_zone(const BlockDims& b, const FoldShell& f, long t, int& k, int& j, int& i) {
    const int nx1 = b.nx1, nx2 = b.ndim > 1 ? b.nx2 : 1;
    auto edge = [&](int e, int s0, int e0) { return e < b.ng ? s0 + e : e0 - (e - b.ng); };
    const int k0 = b.ks + (f.p2 ? b.ng : 0), j0 = b.js + (f.p1 ? b.ng : 0);
    const int mj = nx2 - (f.p1 ? 2 * b.ng : 0);
    if (t < f.n1) {
        i = b.is + (int)(t % nx1); t /= nx1;
        j = b.js + (int)(t % nx2); t /= nx2;
        k = edge((int)t, b.ks, b.ke);
        return;
    }
    t -= f.n1;
    if (t < f.n2) {
        i = b.is + (int)(t % nx1); t /= nx1;
        j = edge((int)(t % (2 * b.ng)), b.js, b.je); t /= 2 * b.ng;
        k = k0 + (int)t;
        return;
    }
    t -= f.n2;
    i = edge((int)(t % (2 * b.ng)), b.is, b.ie); t /= 2 * b.ng;
    j = j0 + (int)(t % mj); t /= mj;
    k = k0 + (int)t;
}

__global__ void __launch_bounds__(ABS) adj_fold_kernel(BlockDims b, double* __restrict__ A, int nvar,
                                                       const int* __restrict__ gate, FoldShell f, int shell) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = shell ? f.n1 + f.n2 + f.n3 : (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n || (gate && *gate == 0)) return;
    int k, j, i;
    if (shell) fold_shell_zone(b, f, t, k, j, i);
    else interior_zone(b, t, k, j, i);
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

// Count of zones with a non-zero inversion flag, or floor flag where fflag is
// given (the adjoint covers converged recoveries, and floors only where the
// stage's U* is kept).
__global__ void __launch_bounds__(ABS) adj_count_flags_kernel(BlockDims b, const int* __restrict__ pflag,
                                                              const int* __restrict__ fflag, int* __restrict__ err) {
    // fflag null: floored zones are supported and only failed inversions count
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    int bad = 0;
    if (t < n) {
        int k, j, i;
        interior_zone(b, t, k, j, i);
        const long x = (long)k * b.s3 + (long)j * b.s2 + i;
        bad = (pflag[x] != 0) || (fflag && fflag[x] != 0);
    }
    if (bad) atomicAdd(&err[0], 1);
}

// lP += (dS/dP)^T lS at each interior zone (geometric source; lS = lam dudt).
__global__ void __launch_bounds__(ABS)
adj_geo_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ P, const double* __restrict__ ld,
               double* __restrict__ lP) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n) return;
    int k, j, i;
    interior_zone(b, t, k, j, i);
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    double Pz[NV], out[NV] = {0., 0., 0., 0., 0., 0., 0., 0.}, lS[4];
#pragma unroll
    for (int v = 0; v < NV; ++v) Pz[v] = P[v * b.nc + x];
#pragma unroll
    for (int v = 0; v < 4; ++v) lS[v] = ld[(UU + v) * b.nc + x];
    const long q = (long)j * gd.gn1 + i;
    CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, q);
    geo_source_vjp(g, gd, q, Pz, c.gam, lS, out);
#pragma unroll
    for (int v = 0; v < NV; ++v) lP[v * b.nc + x] += out[v];
}

// lPin = (dfloors/dP at Pin)^T lPout at each interior zone (B passes through).
template <int GEO, int FRAME>
__global__ void __launch_bounds__(ABS)
adj_floors_kernel(BlockDims b, StepConsts c, GeomDev gd, const double* __restrict__ Pin,
                  const double* __restrict__ lPout, double* __restrict__ lPin) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n) return;
    int k, j, i;
    interior_zone(b, t, k, j, i);
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    double Pz[NV], lo[NV], li[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { Pz[v] = Pin[v * b.nc + x]; lo[v] = lPout[v * b.nc + x]; }
    if constexpr (GEO) {
        const long q = (long)j * gd.gn1 + i;
        CachedGeom<double> g(gd.g[GLOC_CENTER], gd.np, q);
        floors_vjp<FRAME>(g, Pz, c.gam, c.floors, gd.flr[q], gd.flr[gd.np + q], lo, li);
    } else {
        MinkowskiGeom<double> g;
        floors_vjp<FRAME>(g, Pz, c.gam, c.floors, c.floors.rho_min_const, c.floors.u_min_const, lo, li);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) lPin[v * b.nc + x] = li[v];
}

// lP <- (dP_fixed/dP)^T lP for the fluid primitives of the fixup alone (task-level
// check of adj_fixup_kernel): failed zones lose their fluid adjoint to lPfx's
// neighbours, the others add lPfx.
__global__ void __launch_bounds__(ABS)
adj_fixup_apply_kernel(BlockDims b, const int* __restrict__ pf, const double* __restrict__ lPfx, double* __restrict__ lP) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
    if (t >= n) return;
    int k, j, i;
    interior_zone(b, t, k, j, i);
    const long x = (long)k * b.s3 + (long)j * b.s2 + i;
    const bool failed = pf[x] > 0;
#pragma unroll
    for (int v = 0; v < 5; ++v) lP[v * b.nc + x] = failed ? 0. : lP[v * b.nc + x] + lPfx[v * b.nc + x];
}

long ninterior_l(const BlockDims& b) {
    return (long)b.nx1 * (b.ndim > 1 ? b.nx2 : 1) * (b.ndim > 2 ? b.nx3 : 1);
}

template <int GEO>
hipError_t launch_adj_flux_t(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P,
                             const double* lF, double* lamP, hipStream_t s) {
    const int jlo = (b.ndim > 1) ? b.js - 1 : b.js, jhi = (b.ndim > 1) ? b.je + 1 : b.je;
    const int klo = (b.ndim > 2) ? b.ks - 1 : b.ks, khi = (b.ndim > 2) ? b.ke + 1 : b.ke;
    const int nj = jhi - jlo + 1, nk = khi - klo + 1;
    for (int dir = 1; dir <= b.ndim; ++dir) {
        const double* lFd = lF + (long)(dir - 1) * NV * b.nc;
        if (dir == 1) {
            const long slots = (long)(b.nx1 + 10) * nj * nk;
            hipLaunchKernelGGL((adj_flux_x1_fused_kernel<GEO>), dim3((unsigned)((slots + 249) / 250)), dim3(256), 0,
                               s, b, c, gd, P, lFd, lamP, jlo, nj, klo, nk);
        } else {
            const int no = (dir == 2) ? nk : nj;
            const int o0 = (dir == 2) ? klo : jlo;
            const int ncol = (b.nx1 + 2) * no;
            constexpr int TX = adj_xt_tx<GEO>(), TY = adj_xt_ty<GEO>();
            const int nout = ((dir == 2) ? b.nx2 : b.nx3) + 6;
            const dim3 grid((unsigned)((ncol + TX - 1) / TX), (unsigned)((nout + TY - 7) / (TY - 6)));
            if (dir == 2)
                hipLaunchKernelGGL((adj_flux_xt_fused_kernel<2, GEO>), grid, dim3(TX, TY), 0, s, b, c, gd, P, lFd, lamP, o0, no);
            else
                hipLaunchKernelGGL((adj_flux_xt_fused_kernel<3, GEO>), grid, dim3(TX, TY), 0, s, b, c, gd, P, lFd, lamP, o0, no);
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
                           const double* lF, double* lamP, hipStream_t s) {
    if (b.geo) return launch_adj_flux_t<1>(b, c, gd, P, lF, lamP, s);
    return launch_adj_flux_t<0>(b, c, gd, P, lF, lamP, s);
}

hipError_t launch_adj_divct(const BlockDims& b, const double* ld, double* lF, double* lemf, hipStream_t s,
                            FluxSeed seed) {
    const dim3 g = agrid(b.nc);
    if (b.ndim == 3) {
        hipLaunchKernelGGL(adj_divct_kernel<3>, g, dim3(ABS), 0, s, b, ld, lF, lemf, seed);
        hipLaunchKernelGGL(adj_emf_kernel<3>, g, dim3(ABS), 0, s, b, lemf, lF);
    } else if (b.ndim == 2) {
        hipLaunchKernelGGL(adj_divct_kernel<2>, g, dim3(ABS), 0, s, b, ld, lF, lemf, seed);
        hipLaunchKernelGGL(adj_emf_kernel<2>, g, dim3(ABS), 0, s, b, lemf, lF);
    } else {
        hipLaunchKernelGGL(adj_divct_kernel<1>, g, dim3(ABS), 0, s, b, ld, lF, lemf, seed);
    }
    return hipGetLastError();
}

hipError_t launch_adj_stage(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* Pout,
                            const double* lPout, const double* lUout, double g0, double g1, double beta_dt, bool same,
                            bool accumulate, double* lUsub, double* lUfull, double* ld, int* err, const double* Ustar,
                            const double* Psub, const int* ffs, double* lPgeo, const int* pfs, const int* gate,
                            const double* lPfx, double* Pinv, double* lPinv, double* lPtot, hipStream_t s) {
    const dim3 g = agrid(ninterior_l(b));
    const int frame = ffs ? c.floors.frame : -1;
    if (frame == 1 || (ffs && (!Ustar || !Pinv || !lPinv || !lPtot))) return hipErrorInvalidValue;
    const dim3 g8 = agrid(8 * ninterior_l(b));
#define KA_PREP(GEO, FR)                                                                                     \
    hipLaunchKernelGGL((adj_floors_prep_kernel<GEO>), g, dim3(ABS), 0, s, b, c, gd, Pout, lPout, lUout,       \
                       Ustar, Psub, ffs, pfs, gate, lPfx, Pinv, lPtot, err);                                 \
    hipLaunchKernelGGL((adj_floors_cols_kernel<GEO, FR>), g8, dim3(ABS), 0, s, b, c, gd, ffs, pfs, Pinv,     \
                       lPtot, lPinv)
#define KA_STAGE(GEO, FR)                                                                                      \
    hipLaunchKernelGGL((adj_stage_kernel<GEO, FR>), g, dim3(ABS), 0, s, b, c, gd, Pout, lPout, lUout, g0, g1,   \
                       beta_dt, (int)same, (int)accumulate, lUsub, lUfull, ld, err, Pinv, lPinv, Psub, ffs, lPgeo, \
                       pfs, gate, lPfx)
    if (b.geo) {
        if (frame == 0) { KA_PREP(1, 0); KA_STAGE(1, 0); }
        else if (frame == 2) { KA_PREP(1, 2); KA_STAGE(1, 2); }
        else KA_STAGE(1, -1);
    } else {
        if (frame == 0) { KA_PREP(0, 0); KA_STAGE(0, 0); }
        else if (frame == 2) { KA_PREP(0, 2); KA_STAGE(0, 2); }
        else KA_STAGE(0, -1);
    }
#undef KA_STAGE
#undef KA_PREP
    return hipGetLastError();
}

hipError_t launch_adj_fixup(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* Pout,
                            const double* lPout, const double* lUout, const int* pfs, const int* gate, double* lPfx,
                            hipStream_t s) {
    if (b.geo) hipLaunchKernelGGL(adj_fixup_kernel<1>, agrid(b.nc), dim3(ABS), 0, s, b, c, gd, Pout, lPout, lUout, pfs, gate, lPfx);
    else hipLaunchKernelGGL(adj_fixup_kernel<0>, agrid(b.nc), dim3(ABS), 0, s, b, c, gd, Pout, lPout, lUout, pfs, gate, lPfx);
    return hipGetLastError();
}

hipError_t launch_adj_fixup_apply(const BlockDims& b, const int* pfs, const double* lPfx, double* lP, hipStream_t s) {
    hipLaunchKernelGGL(adj_fixup_apply_kernel, agrid(ninterior_l(b)), dim3(ABS), 0, s, b, pfs, lPfx, lP);
    return hipGetLastError();
}

hipError_t launch_adj_bounds(const BlockDims& b, const GeomDev& gd, const double* P, double* A, double* C,
                             hipStream_t s) {
    if (b.periodic[0] && (b.ndim < 2 || b.periodic[1])) return hipSuccess;
    if (b.ng > 4) return hipErrorInvalidValue;
    const long ng = ghost_slab1(b) + ghost_slab2(b);
    if (b.geo) hipLaunchKernelGGL(adj_bounds_map_kernel<1>, agrid(ng), dim3(ABS), 0, s, b, gd, P, A, C);
    else hipLaunchKernelGGL(adj_bounds_map_kernel<0>, agrid(ng), dim3(ABS), 0, s, b, gd, P, A, C);
    const long n = bounds_band1(b) + bounds_band2(b);
    hipLaunchKernelGGL(adj_bounds_kernel, agrid(n), dim3(ABS), 0, s, b, C, A);
    return hipGetLastError();
}

long x1_flux_rows(const BlockDims& b) {
    if (b.periodic[0]) return 0;
    return 2L * ((b.ndim > 1) ? b.nx2 + 2 : 1) * ((b.ndim > 2) ? b.nx3 + 2 : 1);
}

hipError_t launch_save_x1_flux(const BlockDims& b, const double* F, double* x1f, hipStream_t s) {
    const long n = x1_flux_rows(b);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(save_x1_flux_kernel, agrid(n), dim3(ABS), 0, s, b, F, x1f);
    return hipGetLastError();
}

hipError_t launch_adj_fix_flux(const BlockDims& b, double* lF, const double* x1f, hipStream_t s) {
    const long npolar = (b.ndim > 1 && !b.periodic[1]) ? 2L * (b.nx1 + 2) * ((b.ndim > 2) ? b.nx3 + 2 : 1) : 0;
    const long n = npolar + x1_flux_rows(b);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(adj_fix_flux_kernel, agrid(n), dim3(ABS), 0, s, b, lF, x1f);
    return hipGetLastError();
}

hipError_t launch_adj_seed(const BlockDims& b, int dir, int var, int face, double w, double* lF, hipStream_t s) {
    const int na = (dir == 1) ? (b.ndim > 1 ? b.nx2 : 1) : b.nx1;
    const int nb = (dir == 3) ? (b.ndim > 1 ? b.nx2 : 1) : (b.ndim > 2 ? b.nx3 : 1);
    hipLaunchKernelGGL(adj_seed_kernel, agrid((long)na * nb), dim3(ABS), 0, s, b, dir, var, face, w, lF);
    return hipGetLastError();
}

hipError_t launch_adj_ptou(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P,
                           const double* lU, double* lP, hipStream_t s) {
    const dim3 g = agrid(ninterior_l(b));
    if (b.geo) hipLaunchKernelGGL(adj_ptou_kernel<1>, g, dim3(ABS), 0, s, b, c, gd, P, lU, lP);
    else hipLaunchKernelGGL(adj_ptou_kernel<0>, g, dim3(ABS), 0, s, b, c, gd, P, lU, lP);
    return hipGetLastError();
}

// Fold over the periodic shell (or the whole interior when an axis is too short).
static void fold_launch(const BlockDims& b, double* A, int nvar, const int* gate, hipStream_t s) {
    const FoldShell f = fold_shell(b);
    const bool shell = fold_shell_ok(b);
    const long n = shell ? f.n1 + f.n2 + f.n3 : ninterior_l(b);
    if (n > 0) hipLaunchKernelGGL(adj_fold_kernel, agrid(n), dim3(ABS), 0, s, b, A, nvar, gate, f, (int)shell);
}

hipError_t launch_adj_fold(const BlockDims& b, double* A, int nvar, hipStream_t s) {
    const bool periodic = b.periodic[0] || (b.ndim > 1 && b.periodic[1]) || (b.ndim > 2 && b.periodic[2]);
    if (periodic) fold_launch(b, A, nvar, nullptr, s);
    hipLaunchKernelGGL(adj_clear_ghosts_kernel, agrid(b.nc), dim3(ABS), 0, s, b, A, nvar);
    return hipGetLastError();
}

hipError_t launch_adj_fold_interior(const BlockDims& b, double* A, int nvar, const int* gate, hipStream_t s) {
    // periodic axes with images (collapsed axes have none)
    const bool periodic = b.periodic[0] || (b.ndim > 1 && b.periodic[1]) || (b.ndim > 2 && b.periodic[2]);
    if (!periodic) return hipSuccess;
    fold_launch(b, A, nvar, gate, s);
    return hipGetLastError();
}

hipError_t launch_adj_geo(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P, const double* ld,
                          double* lP, hipStream_t s) {
    if (!b.geo) return hipSuccess;
    hipLaunchKernelGGL(adj_geo_kernel, agrid(ninterior_l(b)), dim3(ABS), 0, s, b, c, gd, P, ld, lP);
    return hipGetLastError();
}

hipError_t launch_adj_floors(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* Pin,
                             const double* lPout, double* lPin, hipStream_t s) {
    const dim3 g = agrid(ninterior_l(b));
    const int fr = c.floors.frame;
    if (fr == 1) return hipErrorInvalidValue;
    if (b.geo) {
        if (fr == 0) hipLaunchKernelGGL((adj_floors_kernel<1, 0>), g, dim3(ABS), 0, s, b, c, gd, Pin, lPout, lPin);
        else hipLaunchKernelGGL((adj_floors_kernel<1, 2>), g, dim3(ABS), 0, s, b, c, gd, Pin, lPout, lPin);
    } else {
        if (fr == 0) hipLaunchKernelGGL((adj_floors_kernel<0, 0>), g, dim3(ABS), 0, s, b, c, gd, Pin, lPout, lPin);
        else hipLaunchKernelGGL((adj_floors_kernel<0, 2>), g, dim3(ABS), 0, s, b, c, gd, Pin, lPout, lPin);
    }
    return hipGetLastError();
}

hipError_t launch_adj_count_flags(const BlockDims& b, const int* pflag, const int* fflag, int* err, hipStream_t s) {
    hipLaunchKernelGGL(adj_count_flags_kernel, agrid(ninterior_l(b)), dim3(ABS), 0, s, b, pflag, fflag, err);
    return hipGetLastError();
}

}  // namespace kh
