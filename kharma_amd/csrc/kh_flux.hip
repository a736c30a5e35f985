// kh_flux.hip -- fused face-flux kernels: reconstruction + reconstruction floors
// + prim_to_flux + signal speeds + LLF/HLLE in one pass per direction.
//
// Replaces the four kernels of Flux::GetFlux<Recon,dir> (flux/get_flux.hpp:60-373:
// calc_flux_recon, calc_flux_left, calc_flux_right, flux_llf), which stage
// Pl/Pr/Ul/Ur/Fl/Fr through HBM scratch.  Here each work-item reconstructs ONE
// cell (both of its face values, sharing the smoothness indicators), publishes
// the right-face value through LDS, and then computes the flux through its
// cell's left face from its own left-face value and its neighbour's right-face
// value.  Only the 8 fluxes and 2 signal speeds per face reach HBM.
//
// X1: 256-thread blocks walk the cells of consecutive x1 rows (flattened,
// consecutive blocks overlap by one cell); loads and stores are contiguous.
// X2/X3: 64 x TY blocks; x is a flattened transverse index with x1 fastest (so
// a wavefront reads 64 consecutive values), y walks TY cells along the face
// normal, consecutive tiles overlapping by one cell.
#include "kh_internal.hpp"

namespace kh {

// The FP32 kernels, which are VALU-bound, address their arrays through 32-bit
// offsets (at32: 5-6 % fewer VALU instructions, 6-7 % faster X2/X3); the FP64
// kernels keep 64-bit indices, which measured faster for them (X1 8 %).
template <typename R> struct A32Of { static constexpr bool value = sizeof(R) == 4; };
template <bool A32, class T> __device__ __forceinline__ T& at_idx(T* p, long e) {
    if constexpr (A32) return at32(p, (unsigned)e);
    else return p[e];
}

template <int RECON, typename R>
__device__ __forceinline__ void cell_states(R x1, R x2, R x3, R x4, R x5, R& lout, R& rout) {
    if constexpr (RECON == RECON_WENO5) {
        weno5_cell(x1, x2, x3, x4, x5, lout, rout);
    } else if constexpr (RECON == RECON_LINEAR_MC) {
        const R dq = mc_lim(x3 - x2, x4 - x3) * (x4 - x3);
        rout = x3 + R(0.5) * dq;
        lout = x3 - R(0.5) * dq;
    } else if constexpr (RECON == RECON_PPM) {
        ppm(x1, x2, x3, x4, x5, lout, rout);
    } else {
        lout = x3;
        rout = x3;
    }
}

// Flux through one face from its two states (get_flux.hpp:237-367).
// SPEEDS: SPEEDS_BOTH writes cmax and cmin (Flux.cmax/cmin), SPEEDS_CTOP writes
// max(cmax, cmin) (all the timestep uses) into cmax_out, SPEEDS_NONE writes none.
template <int DIR, int SPEEDS, typename R, class G>
__device__ __forceinline__ void face_flux_g(const StepConsts& c, const BlockDims& b, const G& g, R rflr, R uflr,
                                            R Pl[NV], R Pr[NV], long ci,
                                            const Vars<R>& F, R* __restrict__ cmax_out, R* __restrict__ cmin_out) {
    constexpr bool A32 = A32Of<R>::value;
    const R gam = R(c.gam);
    // get_flux.hpp:170-182: geometric floors of the face's cell centre on both states
    if (c.recon_floors) {
        apply_geo_floors<R>(Pl, rflr, uflr);
        apply_geo_floors<R>(Pr, rflr, uflr);
    }
    R Ul[NV], Fl[NV], Ur[NV], Fr[NV];
    R cmaxL, cminL, cmaxR, cminR;
    {
        FourVec<R> D;
        calc_4vecs<R>(g, Pl, D);
        prim_to_u_and_flux<R, G, DIR>(g, Pl, D, gam, Ul, Fl);
        vchar<R, G, DIR>(g, Pl, D, gam, cmaxL, cminL);
    }
    {
        FourVec<R> D;
        calc_4vecs<R>(g, Pr, D);
        prim_to_u_and_flux<R, G, DIR>(g, Pr, D, gam, Ur, Fr);
        vchar<R, G, DIR>(g, Pr, D, gam, cmaxR, cminR);
    }
    // get_flux.hpp:275-276 and 331-332: cmin is stored as a positive speed
    const R cmx = smax(smax(R(0), cmaxL), cmaxR);
    const R cmn = -smin(smin(R(0), cminL), cminR);
    if constexpr (SPEEDS == SPEEDS_BOTH) {
        at_idx<A32>(cmax_out, ci) = cmx;
        at_idx<A32>(cmin_out, ci) = cmn;
    } else if constexpr (SPEEDS == SPEEDS_CTOP) {
        at_idx<A32>(cmax_out, ci) = smax(cmx, cmn);
    }
    if (c.use_hlle) {
        const R inv = rcp(cmx + cmn);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            at_idx<A32>(F[v], ci) = (cmx * Fl[v] + cmn * Fr[v] - cmx * cmn * (Ur[v] - Ul[v])) * inv;
    } else {
        const R ctop = smax(cmx, cmn);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            at_idx<A32>(F[v], ci) = R(0.5) * (Fl[v] + Fr[v] - ctop * (Ur[v] - Ul[v]));
    }
}

template <typename R> __device__ __forceinline__ const R* geo_table(const GeomDev& gd, int loc);
template <> __device__ __forceinline__ const double* geo_table<double>(const GeomDev& gd, int loc) { return gd.g[loc]; }
template <> __device__ __forceinline__ const float* geo_table<float>(const GeomDev& gd, int loc) { return gd.g32[loc]; }
template <typename R> __device__ __forceinline__ R geo_floor(const GeomDev& gd, int which, long q);
template <> __device__ __forceinline__ double geo_floor<double>(const GeomDev& gd, int which, long q) { return gd.flr[which * gd.np + q]; }
template <> __device__ __forceinline__ float geo_floor<float>(const GeomDev& gd, int which, long q) { return gd.flr32[which * gd.np + q]; }

// Face at entire-block indices (j, i): flat metric (GEO == 0) or the cached
// metric of the face locus (X1 face, X2 face, or the centre for X3 faces of the
// axisymmetric metric) with the floors of cell (j, i).
template <int DIR, int SPEEDS, typename R, int GEO>
__device__ __forceinline__ void face_flux(const StepConsts& c, const BlockDims& b, const GeomDev& gd, int j, int i,
                                          R Pl[NV], R Pr[NV], long ci,
                                          const Vars<R>& F, R* __restrict__ cmax_out, R* __restrict__ cmin_out) {
    if constexpr (GEO) {
        const long q = (long)j * gd.gn1 + i;
        const int loc = (DIR == 1) ? GLOC_FACE1 : (DIR == 2) ? GLOC_FACE2 : GLOC_CENTER;
        CachedGeom<R, A32Of<R>::value> g(geo_table<R>(gd, loc), gd.np, q);
        face_flux_g<DIR, SPEEDS, R>(c, b, g, geo_floor<R>(gd, 0, q), geo_floor<R>(gd, 1, q), Pl, Pr, ci, F, cmax_out, cmin_out);
    } else {
        MinkowskiGeom<R> g;
        face_flux_g<DIR, SPEEDS, R>(c, b, g, R(c.floors.rho_min_const), R(c.floors.u_min_const), Pl, Pr, ci, F, cmax_out, cmin_out);
    }
}

// X1 faces i in [is, ie+1] of rows j in [j0, j0+nj), k in [k0, k0+nk).
// Cells i in [is-1, ie+1] (L per row); block b handles flattened cells
// [255 b, 255 b + 256) and emits the faces of all but its first cell.
// Waves per SIMD the FP32 flux kernels are built for (the second launch bound):
// 6 for the X2/X3 kernels (68 VGPRs instead of 91: 11-14 % faster on the
// 192x128x128 torus) and for the X1 kernels (67-68 instead of 90-115).  The
// FP64 kernels keep their 140-142 VGPRs: a 4-wave bound spills 28-36 bytes per
// lane and is slower.
template <typename R, int SPEEDS, bool PK> struct X1Occ { static constexpr int w = sizeof(R) == 4 ? 6 : 1; };
template <typename R> struct XtOcc { static constexpr int w = sizeof(R) == 4 ? 6 : 1; };
template <int RECON, int SPEEDS, typename R, typename PT, int GEO, bool PK>
__global__ void __launch_bounds__(256, (X1Occ<R, SPEEDS, PK>::w))
flux_x1_kernel(BlockDims b, StepConsts c, GeomDev gd, Vars<const PT> P, Vars<R> F,
               R* __restrict__ cmax_out, R* __restrict__ cmin_out, int j0, int nj, int k0, int nk,
               const BlockEntry* __restrict__ pk) {
    __shared__ R sr[NV][256];
    if constexpr (PK) {
        const long blk = pack_select(pk, b, gd);
        P = pack_at(P, blk * NV * b.nc);
        F = pack_at(F, blk * 3 * NV * b.nc);
        cmax_out = pack_at(cmax_out, blk * 3 * b.nc);
        cmin_out = pack_at(cmin_out, blk * 3 * b.nc);
    }
    const int L = b.nx1 + 2;
    const int tid = threadIdx.x;
    // cell counts stay below 2^31 (the C ABI bounds a block's size), so the
    // flattened index is divided in 32 bits
    const unsigned g = blockIdx.x * 255u + (unsigned)tid;
    const bool valid = g < (unsigned)L * (unsigned)nj * (unsigned)nk;
    int ii = 0, jf = 0;
    long ci = 0;
    R lo[NV];
    if (valid) {
        const int row = (int)(g / (unsigned)L);
        ii = (int)(g - (unsigned)row * (unsigned)L);
        const int jj = (int)((unsigned)row % (unsigned)nj), kk = (int)((unsigned)row / (unsigned)nj);
        jf = j0 + jj;
        ci = (long)(k0 + kk) * b.s3 + (long)(j0 + jj) * b.s2 + (b.is - 1 + ii);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R ro;
            if constexpr (A32Of<R>::value) {
                const PT* __restrict__ q = P[v];
                const unsigned cu = (unsigned)ci;
                cell_states<RECON, R>(R(at32(q, cu - 2)), R(at32(q, cu - 1)), R(at32(q, cu)), R(at32(q, cu + 1)),
                                      R(at32(q, cu + 2)), lo[v], ro);
            } else {
                const PT* __restrict__ q = P[v] + ci;
                cell_states<RECON, R>(R(q[-2]), R(q[-1]), R(q[0]), R(q[1]), R(q[2]), lo[v], ro);
            }
            sr[v][tid] = ro;
        }
    }
    __syncthreads();
    if (!valid || tid == 0 || ii == 0) return;
    R Pl[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Pl[v] = sr[v][tid - 1];
    face_flux<1, SPEEDS, R, GEO>(c, b, gd, jf, b.is - 1 + ii, Pl, lo, ci, F, cmax_out, cmin_out);
}

// X2 / X3 faces.  Transverse columns: flattened (outer, i) with i in [is-1, ie+1]
// fastest; outer = k (X2) or j (X3) from o0.  Along the normal: faces s in
// [s0, s1]; tile y covers cells s0-1+(TY-1)y .. s0-2+TY+(TY-1)y.  TY matches the
// block to the waves a CU holds at the variant's registers: the curved-space
// FP64 variant (~142 VGPRs, 3 waves per SIMD) runs 64 x 12 blocks, one per CU at
// 12 waves (5 % faster than 64 x 8, whose one block per CU holds 8); the others
// (FP32 ~92, flat FP64 ~104 VGPRs) run 64 x 8, two blocks per CU (64 x 10 and
// 64 x 16 FP32 blocks measured 24 % and 2 % slower).
template <typename R, int GEO> struct XtTile { static constexpr int TY = (sizeof(R) == 8 && GEO) ? 12 : 8; };
template <int DIR, int RECON, int SPEEDS, typename R, typename PT, int GEO, bool PK>
__global__ void __launch_bounds__((64 * XtTile<R, GEO>::TY), XtOcc<R>::w)
flux_xt_kernel(BlockDims b, StepConsts c, GeomDev gd, Vars<const PT> P, Vars<R> F,
               R* __restrict__ cmax_out, R* __restrict__ cmin_out, int o0, int no, int s0, int s1,
               const BlockEntry* __restrict__ pk) {
    constexpr int TY = XtTile<R, GEO>::TY;
    __shared__ R sr[NV][TY][64];
    if constexpr (PK) {
        const long blk = pack_select(pk, b, gd);
        P = pack_at(P, blk * NV * b.nc);
        F = pack_at(F, blk * 3 * NV * b.nc);
        cmax_out = pack_at(cmax_out, blk * 3 * b.nc);
        cmin_out = pack_at(cmin_out, blk * 3 * b.nc);
    }
    const int ni = b.nx1 + 2;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int col = blockIdx.x * 64 + tx;
    const int s = s0 - 1 + (int)blockIdx.y * (TY - 1) + ty;
    const bool valid = col < ni * no && s <= s1;
    const long st = (DIR == 2) ? b.s2 : b.s3;
    long ci = 0;
    int jf = 0, if_ = 0;
    R lo[NV];
    if (valid) {
        const int outer = o0 + col / ni;
        const int i = b.is - 1 + col % ni;
        const int j = (DIR == 2) ? s : outer;
        const int k = (DIR == 2) ? outer : s;
        jf = j; if_ = i;
        ci = (long)k * b.s3 + (long)j * b.s2 + i;
        constexpr bool A = A32Of<R>::value;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R ro;
            if constexpr (A) {
                const PT* __restrict__ q = P[v];
                cell_states<RECON, R>(R(at32(q, unsigned(ci - 2 * st))), R(at32(q, unsigned(ci - st))),
                                      R(at32(q, unsigned(ci))), R(at32(q, unsigned(ci + st))),
                                      R(at32(q, unsigned(ci + 2 * st))), lo[v], ro);
            } else {
                const PT* __restrict__ q = P[v] + ci;
                cell_states<RECON, R>(R(q[-2 * st]), R(q[-st]), R(q[0]), R(q[st]), R(q[2 * st]), lo[v], ro);
            }
            sr[v][ty][tx] = ro;
        }
    }
    __syncthreads();
    if (!valid || ty == 0) return;
    R Pl[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Pl[v] = sr[v][ty - 1][tx];
    face_flux<DIR, SPEEDS, R, GEO>(c, b, gd, jf, if_, Pl, lo, ci, F, cmax_out, cmin_out);
}

template <int RECON, int SPEEDS, typename R, typename PT, int GEO, bool PK>
static hipError_t launch_recon(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir, const Vars<const PT>& P, const Vars<R>& F,
                               R* cmax, R* cmin, hipStream_t s, const Pack* pk) {
    const BlockEntry* pd = pack_dev(pk);
    // Face ranges consumed by flux-CT and the divergence: the face-normal index
    // covers [start, end+1]; transverse indices extend one zone each side.
    const int jlo = (b.ndim > 1) ? b.js - 1 : b.js, jhi = (b.ndim > 1) ? b.je + 1 : b.je;
    const int klo = (b.ndim > 2) ? b.ks - 1 : b.ks, khi = (b.ndim > 2) ? b.ke + 1 : b.ke;
    if (dir == 1) {
        const int nj = jhi - jlo + 1, nk = khi - klo + 1;
        const long cells = (long)(b.nx1 + 2) * nj * nk;
        const dim3 grid = pack_grid(dim3((unsigned)((cells - 1 + 254) / 255)), pk);
        hipLaunchKernelGGL((flux_x1_kernel<RECON, SPEEDS, R, PT, GEO, PK>), grid, dim3(256), 0, s, b, c, gd, P, F, cmax, cmin, jlo, nj, klo, nk, pd);
    } else if (dir == 2) {
        const int no = khi - klo + 1;
        const int ncol = (b.nx1 + 2) * no;
        const int nfaces = b.nx2 + 1;
        constexpr int TY = XtTile<R, GEO>::TY;
        const dim3 grid = pack_grid(dim3((unsigned)((ncol + 63) / 64), (unsigned)((nfaces + TY - 2) / (TY - 1))), pk);
        hipLaunchKernelGGL((flux_xt_kernel<2, RECON, SPEEDS, R, PT, GEO, PK>), grid, dim3(64, TY), 0, s, b, c, gd, P, F, cmax, cmin, klo, no, b.js, b.je + 1, pd);
    } else {
        const int no = jhi - jlo + 1;
        const int ncol = (b.nx1 + 2) * no;
        const int nfaces = b.nx3 + 1;
        constexpr int TY = XtTile<R, GEO>::TY;
        const dim3 grid = pack_grid(dim3((unsigned)((ncol + 63) / 64), (unsigned)((nfaces + TY - 2) / (TY - 1))), pk);
        hipLaunchKernelGGL((flux_xt_kernel<3, RECON, SPEEDS, R, PT, GEO, PK>), grid, dim3(64, TY), 0, s, b, c, gd, P, F, cmax, cmin, jlo, no, b.ks, b.ke + 1, pd);
    }
    return hipGetLastError();
}

template <int SPEEDS, typename R, typename PT, int GEO, bool PK>
static hipError_t launch_speeds(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir, int recon, const Vars<const PT>& P, const Vars<R>& F,
                                R* cmax, R* cmin, hipStream_t s, const Pack* pk) {
    switch (recon) {
    case RECON_DONOR: return launch_recon<RECON_DONOR, SPEEDS, R, PT, GEO, PK>(b, c, gd, dir, P, F, cmax, cmin, s, pk);
    case RECON_LINEAR_MC: return launch_recon<RECON_LINEAR_MC, SPEEDS, R, PT, GEO, PK>(b, c, gd, dir, P, F, cmax, cmin, s, pk);
    case RECON_PPM: return launch_recon<RECON_PPM, SPEEDS, R, PT, GEO, PK>(b, c, gd, dir, P, F, cmax, cmin, s, pk);
    default: return launch_recon<RECON_WENO5, SPEEDS, R, PT, GEO, PK>(b, c, gd, dir, P, F, cmax, cmin, s, pk);
    }
}

template <typename R, typename PT, int GEO, bool PK>
static hipError_t launch_geo(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir, int recon, int speeds,
                             const Vars<const PT>& P, const Vars<R>& F, R* cmax, R* cmin, hipStream_t s, const Pack* pk) {
    if (speeds == SPEEDS_CTOP) return launch_speeds<SPEEDS_CTOP, R, PT, GEO, PK>(b, c, gd, dir, recon, P, F, cmax, cmin, s, pk);
    if (speeds == SPEEDS_NONE) return launch_speeds<SPEEDS_NONE, R, PT, GEO, PK>(b, c, gd, dir, recon, P, F, cmax, cmin, s, pk);
    if constexpr (!PK) return launch_speeds<SPEEDS_BOTH, R, PT, GEO, PK>(b, c, gd, dir, recon, P, F, cmax, cmin, s, pk);
    return hipErrorInvalidValue;   // a pack launch is a stage's: no cmin array
}

template <typename R, typename PT, bool PK = false>
static hipError_t launch_prec(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir, int recon, int speeds,
                              const Vars<const PT>& P, const Vars<R>& F, R* cmax, R* cmin, hipStream_t s,
                              const Pack* pk = nullptr) {
    if (b.geo) return launch_geo<R, PT, 1, PK>(b, c, gd, dir, recon, speeds, P, F, cmax, cmin, s, pk);
    return launch_geo<R, PT, 0, PK>(b, c, gd, dir, recon, speeds, P, F, cmax, cmin, s, pk);
}

hipError_t launch_flux(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir, int recon, int prec,
                       const Vars<const double>& P, const Vars<void>& F, void* cmax, void* cmin, hipStream_t s) {
    if (dir == 2 && b.ndim < 2) return hipSuccess;
    if (dir == 3 && b.ndim < 3) return hipSuccess;
    if (prec == PREC_SPLIT)
        return launch_prec<float, double>(b, c, gd, dir, recon, SPEEDS_BOTH, P, vars_cast<float>(F), (float*)cmax,
                                          (float*)cmin, s);
    return launch_prec<double, double>(b, c, gd, dir, recon, SPEEDS_BOTH, P, vars_cast<double>(F), (double*)cmax,
                                       (double*)cmin, s);
}

hipError_t launch_flux_stage(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir, int recon, int prec,
                             int speeds, const double* P, const float* P32, void* F, void* ctop, hipStream_t s,
                             const Pack* pk) {
    if (dir == 2 && b.ndim < 2) return hipSuccess;
    if (dir == 3 && b.ndim < 3) return hipSuccess;
    if (pk) {
        if (prec == PREC_SPLIT)
            return launch_prec<float, float, true>(b, c, gd, dir, recon, speeds, vars_of(P32, b.nc),
                                                   vars_of((float*)F, b.nc), (float*)ctop, nullptr, s, pk);
        return launch_prec<double, double, true>(b, c, gd, dir, recon, speeds, vars_of(P, b.nc),
                                                 vars_of((double*)F, b.nc), (double*)ctop, nullptr, s, pk);
    }
    if (prec == PREC_SPLIT)
        return launch_prec<float, float>(b, c, gd, dir, recon, speeds, vars_of(P32, b.nc), vars_of((float*)F, b.nc),
                                         (float*)ctop, nullptr, s);
    return launch_prec<double, double>(b, c, gd, dir, recon, speeds, vars_of(P, b.nc), vars_of((double*)F, b.nc),
                                       (double*)ctop, nullptr, s);
}

// ---------------------------------------------- first-order flux correction
// Flux::FOFC (flux/fofc.cpp:113-226) with its defaults (LLF, speeds from both
// states, cell-centred B): each face of the consumed ranges next to a marked
// zone gets the LLF flux of the two zones' own primitives, and its speeds.
// `mark` holds the interior marks; a ghost zone reads the mark of its periodic
// image or, across a physical face, of the nearest interior zone -- what the
// outflow and reflecting fills of the synced flag give for the one ghost layer
// a consumed face touches.
__device__ __forceinline__ int fofc_source(int v, int s, int e, int n, int periodic) {
    if (v >= s && v <= e) return v;
    if (periodic) return v < s ? v + n : v - n;
    return v < s ? s : e;
}
template <int DIR, int SPEEDS, int GEO>
__global__ void __launch_bounds__(256)
fofc_flux_kernel(BlockDims b, StepConsts c, GeomDev gd, Vars<const double> P, const int* __restrict__ mark,
                 Vars<double> F, double* __restrict__ cmax_out, double* __restrict__ cmin_out, bool ghost_marks) {
    const int m1 = (DIR == 1) ? b.nx1 + 1 : b.nx1 + 2;
    const int m2 = (b.ndim < 2) ? 1 : (DIR == 2) ? b.nx2 + 1 : b.nx2 + 2;
    const int m3 = (b.ndim < 3) ? 1 : (DIR == 3) ? b.nx3 + 1 : b.nx3 + 2;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;   // face counts stay below 2^31
    if (t >= (unsigned)m1 * (unsigned)m2 * (unsigned)m3) return;
    const int i = ((DIR == 1) ? b.is : b.is - 1) + (int)(t % (unsigned)m1);
    const unsigned r = t / (unsigned)m1;
    const int j = ((b.ndim < 2 || DIR == 2) ? b.js : b.js - 1) + (int)(r % (unsigned)m2);
    const int k = ((b.ndim < 3 || DIR == 3) ? b.ks : b.ks - 1) + (int)(r / (unsigned)m2);
    // a ghost of a shared face keeps its own (synced) mark when ghost_marks is set
    auto src = [&](int v, int s, int e, int n, int d) {
        if (ghost_marks && ((v < s && b.img[d][0]) || (v > e && b.img[d][1]))) return v;
        return fofc_source(v, s, e, n, b.periodic[d]);
    };
    auto marked = [&](int kk, int jj, int ii) {
        ii = src(ii, b.is, b.ie, b.nx1, 0);
        if (b.ndim > 1) jj = src(jj, b.js, b.je, b.nx2, 1);
        if (b.ndim > 2) kk = src(kk, b.ks, b.ke, b.nx3, 2);
        return mark[(long)kk * b.s3 + (long)jj * b.s2 + ii] != 0;
    };
    if (!marked(k, j, i) && !marked(DIR == 3 ? k - 1 : k, DIR == 2 ? j - 1 : j, DIR == 1 ? i - 1 : i)) return;
    const long ci = (long)k * b.s3 + (long)j * b.s2 + i;
    const long st = (DIR == 1) ? 1 : (DIR == 2) ? b.s2 : b.s3;
    double Pl[NV], Pr[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        Pl[v] = P[v][ci - st];
        Pr[v] = P[v][ci];
    }
    face_flux<DIR, SPEEDS, double, GEO>(c, b, gd, j, i, Pl, Pr, ci, F, cmax_out, cmin_out);
}

template <int DIR, int SPEEDS>
static hipError_t launch_fofc_dir(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& P,
                                  const int* mark, const Vars<double>& F, double* cmax, double* cmin, hipStream_t s,
                                  bool gm) {
    const long m1 = (DIR == 1) ? b.nx1 + 1 : b.nx1 + 2;
    const long m2 = (b.ndim < 2) ? 1 : (DIR == 2) ? b.nx2 + 1 : b.nx2 + 2;
    const long m3 = (b.ndim < 3) ? 1 : (DIR == 3) ? b.nx3 + 1 : b.nx3 + 2;
    const dim3 grid((unsigned)((m1 * m2 * m3 + 255) / 256));
    if (b.geo) fofc_flux_kernel<DIR, SPEEDS, 1><<<grid, 256, 0, s>>>(b, c, gd, P, mark, F, cmax, cmin, gm);
    else fofc_flux_kernel<DIR, SPEEDS, 0><<<grid, 256, 0, s>>>(b, c, gd, P, mark, F, cmax, cmin, gm);
    return hipGetLastError();
}

template <int SPEEDS>
static hipError_t launch_fofc_speeds(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir,
                                     const Vars<const double>& P, const int* mark, const Vars<double>& F, double* cmax,
                                     double* cmin, hipStream_t s, bool gm) {
    if (dir == 1) return launch_fofc_dir<1, SPEEDS>(b, c, gd, P, mark, F, cmax, cmin, s, gm);
    if (dir == 2) return launch_fofc_dir<2, SPEEDS>(b, c, gd, P, mark, F, cmax, cmin, s, gm);
    return launch_fofc_dir<3, SPEEDS>(b, c, gd, P, mark, F, cmax, cmin, s, gm);
}

hipError_t launch_fofc_flux(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int speeds,
                            const Vars<const double>& P, const int* mark, const FluxVars<double>& F, double* cmax,
                            double* cmin, hipStream_t s, bool ghost_marks) {
    StepConsts cl = c;   // first-order states: no reconstruction floors; LLF whatever flux/type says
    cl.recon_floors = 0;
    cl.use_hlle = 0;
    for (int dir = 1; dir <= b.ndim; ++dir) {
        double* cx = cmax ? cmax + (long)(dir - 1) * b.nc : nullptr;
        double* cn = cmin ? cmin + (long)(dir - 1) * b.nc : nullptr;
        hipError_t e;
        if (speeds == SPEEDS_CTOP) e = launch_fofc_speeds<SPEEDS_CTOP>(b, cl, gd, dir, P, mark, F[dir - 1], cx, cn, s, ghost_marks);
        else if (speeds == SPEEDS_BOTH) e = launch_fofc_speeds<SPEEDS_BOTH>(b, cl, gd, dir, P, mark, F[dir - 1], cx, cn, s, ghost_marks);
        else e = launch_fofc_speeds<SPEEDS_NONE>(b, cl, gd, dir, P, mark, F[dir - 1], cx, cn, s, ghost_marks);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace kh
