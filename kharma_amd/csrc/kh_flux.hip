// kh_flux.hip -- fused face-flux kernel: reconstruction + reconstruction floors
// + prim_to_flux + signal speeds + LLF/HLLE in one pass per direction.
//
// Replaces the four kernels of Flux::GetFlux<Recon,dir> (flux/get_flux.hpp:60-373:
// calc_flux_recon, calc_flux_left, calc_flux_right, flux_llf), which stage
// Pl/Pr/Ul/Ur/Fl/Fr through HBM scratch.  Here each work-item owns one face:
// it reads the 6-point stencil of the 8 primitives along `dir` (coalesced
// across the wavefront along x1), keeps both face states in registers and
// writes only the 8 fluxes and the two signal speeds.
#include "kh_internal.hpp"

namespace kh {

template <int DIR, int RECON, typename R>
__global__ void __launch_bounds__(256)
flux_kernel(BlockDims b, StepConsts c, const double* __restrict__ P, R* __restrict__ F,
            R* __restrict__ cmax_out, R* __restrict__ cmin_out,
            int i0, int ni, int j0, int nj, int k0, int nk) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned nij = (unsigned)ni * (unsigned)nj;
    if (t >= nij * (unsigned)nk) return;
    const unsigned kk = t / nij;
    const unsigned rem = t - kk * nij;
    const unsigned jj = rem / (unsigned)ni;
    const int i = i0 + (int)(rem - jj * (unsigned)ni);
    const int j = j0 + (int)jj;
    const int k = k0 + (int)kk;
    const long ci = (long)k * b.s3 + (long)j * b.s2 + i;
    const long st = (DIR == 1) ? 1 : ((DIR == 2) ? b.s2 : b.s3);

    const R gam = R(c.gam);
    R Pl[NV], Pr[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double* q = P + v * b.nc + ci;
        R x[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) x[o] = R(q[(o - 3) * st]);
        face_states<RECON, R>(x, Pl[v], Pr[v]);
    }
    // get_flux.hpp:170-182 (Cartesian geometric floors on both face states)
    if (c.recon_floors) {
        apply_geo_floors<R>(Pl, R(c.floors.rho_min_const), R(c.floors.u_min_const));
        apply_geo_floors<R>(Pr, R(c.floors.rho_min_const), R(c.floors.u_min_const));
    }

    MinkowskiGeom<R> g;
    R Ul[NV], Fl[NV], Ur[NV], Fr[NV];
    R cmaxL, cminL, cmaxR, cminR;
    {
        FourVec<R> D;
        calc_4vecs<R>(g, Pl, D);
        prim_to_u_and_flux<R, MinkowskiGeom<R>, DIR>(g, Pl, D, gam, Ul, Fl);
        vchar<R, MinkowskiGeom<R>, DIR>(g, Pl, D, gam, cmaxL, cminL);
    }
    {
        FourVec<R> D;
        calc_4vecs<R>(g, Pr, D);
        prim_to_u_and_flux<R, MinkowskiGeom<R>, DIR>(g, Pr, D, gam, Ur, Fr);
        vchar<R, MinkowskiGeom<R>, DIR>(g, Pr, D, gam, cmaxR, cminR);
    }
    // get_flux.hpp:275-276 and 331-332: cmin is stored as a positive speed
    const R cmx = smax(smax(R(0), cmaxL), cmaxR);
    const R cmn = -smin(smin(R(0), cminL), cminR);
    cmax_out[ci] = cmx;
    cmin_out[ci] = cmn;
    if (c.use_hlle) {
        const R inv = R(1) / (cmx + cmn);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            F[v * b.nc + ci] = (cmx * Fl[v] + cmn * Fr[v] - cmx * cmn * (Ur[v] - Ul[v])) * inv;
    } else {
        const R ctop = smax(cmx, cmn);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            F[v * b.nc + ci] = R(0.5) * (Fl[v] + Fr[v] - ctop * (Ur[v] - Ul[v]));
    }
}

template <int DIR, typename R>
static hipError_t launch_dir_prec(const BlockDims& b, const StepConsts& c, int recon, const double* P, R* F,
                                  R* cmax, R* cmin, hipStream_t s) {
    // Face range consumed by flux-CT and the divergence (see header).
    int i0 = b.is, i1 = b.ie, j0 = b.js, j1 = b.je, k0 = b.ks, k1 = b.ke;
    if (DIR == 1) { i1 += 1; } else { i0 -= 1; i1 += 1; }
    if (b.ndim > 1) { if (DIR == 2) { j1 += 1; } else { j0 -= 1; j1 += 1; } }
    if (b.ndim > 2) { if (DIR == 3) { k1 += 1; } else { k0 -= 1; k1 += 1; } }
    const int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
    const long n = (long)ni * nj * nk;
    const int bs = 256;
    const dim3 grid((unsigned)((n + bs - 1) / bs));
    switch (recon) {
    case RECON_DONOR:
        hipLaunchKernelGGL((flux_kernel<DIR, RECON_DONOR, R>), grid, dim3(bs), 0, s, b, c, P, F, cmax, cmin, i0, ni, j0, nj, k0, nk);
        break;
    case RECON_LINEAR_MC:
        hipLaunchKernelGGL((flux_kernel<DIR, RECON_LINEAR_MC, R>), grid, dim3(bs), 0, s, b, c, P, F, cmax, cmin, i0, ni, j0, nj, k0, nk);
        break;
    case RECON_PPM:
        hipLaunchKernelGGL((flux_kernel<DIR, RECON_PPM, R>), grid, dim3(bs), 0, s, b, c, P, F, cmax, cmin, i0, ni, j0, nj, k0, nk);
        break;
    default:
        hipLaunchKernelGGL((flux_kernel<DIR, RECON_WENO5, R>), grid, dim3(bs), 0, s, b, c, P, F, cmax, cmin, i0, ni, j0, nj, k0, nk);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_flux(const BlockDims& b, const StepConsts& c, int dir, int recon, int prec,
                       const double* P, void* F, void* cmax, void* cmin, hipStream_t s) {
    if (dir == 2 && b.ndim < 2) return hipSuccess;
    if (dir == 3 && b.ndim < 3) return hipSuccess;
    if (prec == PREC_SPLIT) {
        float *f = (float*)F, *mx = (float*)cmax, *mn = (float*)cmin;
        if (dir == 1) return launch_dir_prec<1, float>(b, c, recon, P, f, mx, mn, s);
        if (dir == 2) return launch_dir_prec<2, float>(b, c, recon, P, f, mx, mn, s);
        return launch_dir_prec<3, float>(b, c, recon, P, f, mx, mn, s);
    }
    double *f = (double*)F, *mx = (double*)cmax, *mn = (double*)cmin;
    if (dir == 1) return launch_dir_prec<1, double>(b, c, recon, P, f, mx, mn, s);
    if (dir == 2) return launch_dir_prec<2, double>(b, c, recon, P, f, mx, mn, s);
    return launch_dir_prec<3, double>(b, c, recon, P, f, mx, mn, s);
}

}  // namespace kh
