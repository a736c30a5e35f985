// kh_internal.hpp -- shared host/device declarations between the kernel
// translation units and the C-ABI layer.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "kharma_amd.h"
#include "kh_device.hpp"

namespace kh {

// Index space of one meshblock.  Fields are SoA [var][k][j][i] over the entire
// block (interior + ng ghosts per side in each active dimension), i fastest:
// the layout of KHARMA's VariablePack(v, k, j, i).
struct BlockDims {
    int n1, n2, n3;        // entire sizes
    int nx1, nx2, nx3;     // interior sizes
    int is, ie, js, je, ks, ke;
    int ng, ndim;
    int periodic[3];
    long s2, s3;           // strides of j and k
    long nc;               // stride between variables
    double dx1, dx2, dx3;
    int geo;               // 0: flat Cartesian metric; 1: metric read from the geometry cache
    int bc[4];             // boundary types of the inner/outer x1 and inner/outer x2 faces (KH_BC_*; 0 if periodic)
    int check_inflow[2];   // inner/outer x1
    int zero_polar_flux, fix_polar_flux, fix_corner_inner;
    // Physical boundary fill of driver "simple" (primitives are the synced
    // fields): each filled zone's U.B is recomputed from its P.B = U.B / sqrt(-g).
    int prim_fill;
    // Faces of the block (axis d, side 0 low / 1 high).  img: the face's ghosts
    // are images of interior zones -- of this block (a periodic axis of a
    // one-block mesh) or of the neighbouring block of a mesh of blocks -- rather
    // than a physical boundary.  nbd: index offset of the neighbouring block
    // across the face in the mesh's block order (0 for the block itself).  A
    // block's arrays sit at a fixed stride per block, so an image zone lies at
    // the same offset as in a one-block mesh plus nbd times that stride.
    int img[3][2];
    int nbd[3][2];
    // Interior offset of this block in the mesh (zones per axis); coordinates are
    // x = xmin(mesh) + (i - is + goff + 1/2) dx with the mesh's dx.
    int goff[3];
    // Face-centred constrained transport (B_CT): face and edge arrays are
    // [3][n3+1][n2+1][n1+1], index (k fs3 + j fs2 + i) for the lower face / edge of
    // cell (k, j, i), component stride nf.
    int face_ct, ct_scheme;
    long fs2, fs3, nf;
};
// The block has a physical face (outflow / reflecting ghosts to fill).
__host__ __device__ inline bool has_physical_face(const BlockDims& b) {
    return !b.img[0][0] || !b.img[0][1] || (b.ndim > 1 && (!b.img[1][0] || !b.img[1][1]));
}

// Scalars that change every step live on the device so a run of steps needs no
// host round trip (dt is produced and consumed on the GPU).
struct DeviceState {
    double dt, dt_last, time;
    double last_divb;
    double last_ndt;
    int fail_count[2];                 // failed inversions per stage of the current step
    int ncycle;
    int total_fails;
    int divb_over_cycle;               // first cycle whose divB exceeded kill_on_divb_over (0: none)
    double divb_over_value;            // its divB
};

// Run-wide constants of the step.
struct StepConsts {
    double gam, cfl, dt_min, max_dt_increase, tlim;
    double gam0[2], gam1[2], beta[2];   // integrator weights per stage
    int recon_floors;
    int iter_max;
    double err_tol;
    int fix_average, fix_atmo;
    int use_hlle;
    FloorParams floors;
    int kill_divb;                      // b_field/kill_on_large_divb
    double divb_over;                   // b_field/kill_on_divb_over
    int inverter;                       // inverter/type: 0 onedw, 1 kastaun
};

enum Precision { PREC_FP64 = 0, PREC_SPLIT = 1 };

// Per-variable base pointers of one 8-variable cell container, in the order
// rho, u, u1, u2, u3, B1, B2, B3; each points at a [k][j][i] array over the
// entire block.  KHARMA allocates its variable groups as separate fields, so
// the task kernels address every variable through its own pointer; a
// contiguous [8][nc] array is the case v[n] = base + n nc.
template <typename T> struct Vars {
    T* v[NV];
    __host__ __device__ __forceinline__ T* operator[](int n) const { return v[n]; }
};
template <typename T> __host__ __device__ inline Vars<T> vars_of(T* base, long nc) {
    Vars<T> r;
    for (int n = 0; n < NV; ++n) r.v[n] = base ? base + (long)n * nc : nullptr;
    return r;
}
template <typename T, typename S> inline Vars<T> vars_cast(const Vars<S>& s) {
    Vars<T> r;
    for (int n = 0; n < NV; ++n) r.v[n] = (T*)s.v[n];
    return r;
}
// The flux containers of the three directions (index dir - 1).
template <typename T> struct FluxVars {
    Vars<T> d[3];
    __host__ __device__ __forceinline__ const Vars<T>& operator[](int n) const { return d[n]; }
};
template <typename T> __host__ __device__ inline FluxVars<T> flux_vars_of(T* base, long nc) {
    FluxVars<T> r;
    for (int n = 0; n < 3; ++n) r.d[n] = vars_of(base ? base + (long)n * NV * nc : nullptr, nc);
    return r;
}
template <typename T, typename S> inline FluxVars<T> flux_vars_cast(const FluxVars<S>& s) {
    FluxVars<T> r;
    for (int n = 0; n < 3; ++n) r.d[n] = vars_cast<T>(s.d[n]);
    return r;
}
// The flux containers of a contiguous [3][8][nc] array whose element type is
// the precision's (float for PREC_SPLIT, double otherwise).
inline FluxVars<void> flux_vars_prec(void* F, int prec, long nc) {
    if (prec == PREC_SPLIT) return flux_vars_cast<void>(flux_vars_of((float*)F, nc));
    return flux_vars_cast<void>(flux_vars_of((double*)F, nc));
}

// ---- geometry cache (kh_geom.cpp) ----
// Loci of the cache; the X3 face of an axisymmetric metric equals the centre.
enum { GLOC_CENTER = 0, GLOC_FACE1 = 1, GLOC_FACE2 = 2, GLOC_CORNER = 3, GLOC_N = 4 };
// Components per point: gcov (10, upper triangle row-major), gcon (10), gdet, lapse.
enum { GEO_GCOV = 0, GEO_GCON = 10, GEO_GDET = 20, GEO_ALPHA = 21, GEO_NCOMP = 22 };

struct GeomHost {
    int gn1 = 0, gn2 = 0;   // (n1 + 1) x (n2 + 1) points
    long np = 0;
    std::vector<double> g[GLOC_N];   // [GEO_NCOMP][gn2][gn1] per locus
    std::vector<double> conn;        // [64][gn2][gn1]: gdet * Gamma^mu_{nu lam} at centres, index mu*16+nu*4+lam
    std::vector<double> flr;         // [2][gn2][gn1]: geometric rho and u floors at centres
    std::vector<double> r;           // [gn2][gn1]: radius of the cell centre
};
void build_geometry(const kh_params& p, const BlockDims& b, GeomHost& G);
// Parameter checks and index ranges shared by every entry point (kh_capi.cpp).
int check_params(const kh_params* p);
void make_dims(const kh_params* p, BlockDims& b);
int api_fail(int code, const std::string& msg);
void locus_coords(const kh_params& p, const BlockDims& b, int loc, int j, int i, double X[4]);
double embed_r(const kh_params& p, const double X[4]);
void embed_coords(const kh_params& p, const double X[4], double Xe[4]);
void native_dxdX(const kh_params& p, const double X[4], double M[4][4]);
void native_gcov(const kh_params& p, const double X[4], double g[4][4]);
double invert_metric(const double g[4][4], double gi[4][4]);

// Device copy of the cache.  The flux kernels of the split form read the FP32 copies.
struct GeomDev {
    const double* g[3] = {nullptr, nullptr, nullptr};   // centre, face1, face2
    const float* g32[3] = {nullptr, nullptr, nullptr};
    const double* conn = nullptr;
    unsigned long long conn_mask = 0;   // bit n: component n of conn is non-zero somewhere
    const double* flr = nullptr;
    const float* flr32 = nullptr;
    int gn1 = 0;
    long np = 0;
};
enum SpeedsOut { SPEEDS_NONE = 0, SPEEDS_BOTH = 1, SPEEDS_CTOP = 2 };

// ---- launches over all blocks of a mesh ----
// KHARMA launches its kernels over a MeshBlockPack, every block of the mesh in
// one launch.  Here a launch given a Pack runs gridDim.z = nblk copies of the
// one-block grid: copy z reads block z's dims and geometry from the device
// table and addresses block z's part of every container, which the mesh keeps
// at a fixed stride per block ([block][...]: NV nc for a state, 3 NV nc for the
// fluxes, 3 nc for the signal speeds, nc for a flag, one DeviceState, one
// partial per work-group).  The launch arguments are block 0's.
struct BlockEntry {
    BlockDims dims;
    GeomDev geo;
};
struct Pack {
    const BlockEntry* dev = nullptr;   // nblk entries in device memory
    int nblk = 1;
    long fix_flux_n = 0;               // threads of the largest block's physical-face flux fix
    long bounds_n = 0;                 // and of its physical boundary fill
};
// This work-group's block of a pack launch (pk null: the one-block launch).
__device__ __forceinline__ long pack_select(const BlockEntry* pk, BlockDims& b, GeomDev& gd) {
    if (!pk) return 0;
    const int blk = blockIdx.z;
    b = pk[blk].dims;
    gd = pk[blk].geo;
    return blk;
}
__device__ __forceinline__ long pack_select(const BlockEntry* pk, BlockDims& b) {
    if (!pk) return 0;
    const int blk = blockIdx.z;
    b = pk[blk].dims;
    return blk;
}
template <typename T> __host__ __device__ __forceinline__ T* pack_at(T* p, long off) { return p ? p + off : p; }
template <typename T> __host__ __device__ __forceinline__ Vars<T> pack_at(const Vars<T>& V, long off) {
    Vars<T> r;
    for (int n = 0; n < NV; ++n) r.v[n] = pack_at(V.v[n], off);
    return r;
}
template <typename T> __host__ __device__ __forceinline__ FluxVars<T> pack_at(const FluxVars<T>& F, long off) {
    FluxVars<T> r;
    for (int n = 0; n < 3; ++n) r.d[n] = pack_at(F.d[n], off);
    return r;
}
inline dim3 pack_grid(dim3 g, const Pack* pk) { return dim3(g.x, g.y, pk ? (unsigned)pk->nblk : 1u); }
// The thread counts of the physical-face kernels: the largest over the blocks.
void pack_set_sizes(Pack& pk, const BlockDims* dims, int nblk);
inline const BlockEntry* pack_dev(const Pack* pk) { return pk ? pk->dev : nullptr; }

// ---- launchers (kh_flux.hip) ----
// Face fluxes through the faces normal to `dir` that the step consumes:
// X1 faces i in [is, ie+1], j in [js-1, je+1], k in [ks-1, ke+1] (collapsed in
// inactive dimensions), cyclically for X2/X3.  F/cmax/cmin point at this
// direction's arrays ([8][nc] and [nc]); element type float when prec == PREC_SPLIT.
hipError_t launch_flux(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir, int recon, int prec,
                       const Vars<const double>& P, const Vars<void>& F, void* cmax, void* cmin, hipStream_t s);
// Stage-driver variant: speeds = SPEEDS_NONE or SPEEDS_CTOP (ctop = max(cmax, cmin)
// into ctop[dir-1]); the split form reads the FP32 shadow P32 of the primitives.
hipError_t launch_flux_stage(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int dir, int recon, int prec,
                             int speeds, const double* P, const float* P32, void* F, void* ctop, hipStream_t s,
                             const Pack* pk = nullptr);

// ---- launchers (kh_step.hip) ----
hipError_t launch_emf(const BlockDims& b, int prec, const FluxVars<void>& F, void* emf, hipStream_t s);
hipError_t launch_flux_ct_rewrite(const BlockDims& b, int prec, const FluxVars<void>& F, const void* emf, hipStream_t s);
hipError_t launch_flux_divergence(const BlockDims& b, int prec, const FluxVars<void>& F, const Vars<double>& dUdt,
                                  hipStream_t s);
hipError_t launch_geo_source(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& P,
                             const Vars<double>& dUdt, hipStream_t s);
// Fused stage update (flux-CT EMFs, divergence, RK update, U->P, floors, P->U,
// ghost images).  ctop: [3][nc] max signal speeds, read when do_dt (the last
// stage), which then writes one CFL minimum per block into dt_part.
// First-order flux correction (kh_step.hip / kh_flux.hip).  mark[nc]: interior
// marks; fixed[np] (or null): zones marked whatever the trial state.  images:
// each mark is also stored at its images (the ghosts of shared faces, in the
// neighbouring blocks' mark arrays at a stride of nc): the synced flag of a mesh.
hipError_t launch_fofc_mark(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, const double* F,
                            const double* Usub, const double* Ufull, const double* Psub, const int* fixed,
                            const DeviceState* ds, int* mark, hipStream_t s, bool images = false);
// speeds: SPEEDS_NONE, SPEEDS_CTOP (max(cmax, cmin) into cmax[dir-1]) or SPEEDS_BOTH
// (cmax/cmin [3][nc]).  ghost_marks: the ghosts of shared faces hold their
// marks (launch_fofc_mark with images); otherwise they are read at their
// periodic image.
hipError_t launch_fofc_flux(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int speeds,
                            const Vars<const double>& P, const int* mark, const FluxVars<double>& F, double* cmax,
                            double* cmin, hipStream_t s, bool ghost_marks = false);
// Floors::DetermineGRMHDFloors over the entire block (fflag |= floor bits).
hipError_t launch_determine_floors(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& P,
                                   int* fflag, hipStream_t s);
// Flux::MarkFOFC over the entire block: mark = fflag != 0 or fixed[j][i] (fixed may be null).
hipError_t launch_mark_fofc(const BlockDims& b, const GeomDev& gd, const int* fflag, const int* fixed, int* mark,
                            hipStream_t s);
hipError_t launch_flags_to_real(const int* f, double* r, long n, hipStream_t s);
hipError_t launch_flags_from_real(const double* r, int* f, long n, hipStream_t s);
hipError_t launch_update(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int prec, int stage, bool do_dt,
                         const void* F, const void* ctop, const double* Usub, const double* Ufull, const double* Psub,
                         double* Uout, double* Pout, float* Pout32, int* pflag, int* fflag,
                         int* fail_list, DeviceState* ds, double* dt_part, hipStream_t s, const Pack* pk = nullptr);
// The two halves of launch_update on the curved-space FP64 path: parts bit 1
// runs the combine (U* = g0 U_sub + g1 U_full + beta dt (div F + S) into Uout),
// bit 2 the recovery (U* -> P, floors, P -> U).  With Ustar non-null the
// recovery also stores the interior zones' U* there as it reads them (the
// adjoint's recompute uses it).
hipError_t launch_update_parts(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, int parts,
                               const double* F, const double* Usub, const double* Ufull, const double* Psub, double* Uout,
                               double* Pout, int* pflag, int* fflag, int* fail_list, DeviceState* ds, hipStream_t s,
                               double* Ustar = nullptr);
// launch_update on the FP64 path with the stage's U* also stored to Ustar
// (curved space: the recovery stores the interior zones' U* as it reads them;
// flat space has no separate U* and runs the fused update).
hipError_t launch_update_recorded(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, bool do_dt,
                                  const void* F, const void* ctop, const double* Usub, const double* Ufull,
                                  const double* Psub, double* Uout, double* Pout, int* pflag, int* fflag,
                                  int* fail_list, DeviceState* ds, double* dt_part, double* Ustar, hipStream_t s);
// acc += w dt sum_faces F_dir[var] dA (one stage of a face-flux objective).
hipError_t launch_face_flux_objective(const BlockDims& b, int prec, const void* F, int dir, int var, int face, double w,
                                      const DeviceState* ds, double* acc, hipStream_t s);
// ---- face-centred constrained transport (kh_ct.hip) ----
// Cell-centre EMFs [3][nc] over the entire block from the primitives (Minkowski).
hipError_t launch_ct_cemf(const BlockDims& b, const Vars<const double>& P, double* cemf, hipStream_t s);
// Edge EMFs [3][nf] from the face fluxes (element type by precision) and cemf
// (unused by bs99), scheme b.ct_scheme.
hipError_t launch_ct_emf(const BlockDims& b, int prec, const FluxVars<const void>& F, const double* cemf, double* emf,
                         hipStream_t s);
// Faces from the edge circulation.  ds == nullptr: fBout = dB_Uf/dt on the
// interior face ranges; otherwise fBout = 1 (g0 fBsub + g1 fBfull) + beta dt dB/dt
// with dt from ds (same: the stage reads fBfull alone).
hipError_t launch_ct_face_update(const BlockDims& b, const double* emf, const double* fBsub, const double* fBfull,
                                 double g0, double g1, bool same, const DeviceState* ds, double beta, double* fBout,
                                 hipStream_t s);
hipError_t launch_ct_weighted_sum(const BlockDims& b, const double* x, const double* y, double w1, double w2, double* z,
                                  hipStream_t s);
hipError_t launch_ct_sync(const BlockDims& b, double* fB, hipStream_t s);
hipError_t launch_ct_utop(const BlockDims& b, const double* fB, const Vars<double>& U, const Vars<double>& P,
                          hipStream_t s);
// divb_blocks(b) partial maxima of the face divergence of the interior cells.
hipError_t launch_ct_max_divb(const BlockDims& b, const double* fB, double* part, hipStream_t s);
// The flat stage update with the cell B taken from the stage's new faces fB.
hipError_t launch_update_faces(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int prec, int stage,
                               bool do_dt, const void* F, const void* ctop, const double* Usub, const double* Ufull,
                               const double* Psub, double* Uout, double* Pout, float* Pout32, int* pflag, int* fflag,
                               int* fail_list, DeviceState* ds, double* dt_part, const double* fB, hipStream_t s);
long update_blocks(const BlockDims& b);
long divb_blocks(const BlockDims& b);
hipError_t launch_fixup(const BlockDims& b, const StepConsts& c, const GeomDev& gd, int stage, double* U, double* P,
                        float* P32, const int* pflag, const int* fail_list, DeviceState* ds, hipStream_t s,
                        const Pack* pk = nullptr);
// Flux fixes at physical boundaries, on the Riemann fluxes before the update:
// B_FluxCT::ZeroBoundaryFlux on polar faces (the update kernel's flux-CT then
// sees mirrored B2 fluxes), the no-inflow mass-flux clamp on X1 faces and the
// zeroed polar fluxes of KBoundaries::FixFlux.
hipError_t launch_fix_flux(const BlockDims& b, int prec, const FluxVars<void>& F, hipStream_t s);
hipError_t launch_fix_flux(const BlockDims& b, int prec, void* F, hipStream_t s, const Pack* pk = nullptr);
// Physical boundary ghost zones of U, P (and the FP32 shadow if non-null) from
// the final interior state: KBoundaries::ApplyBoundary for each x1/x2 face.
hipError_t launch_bounds(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<double>& U,
                         const Vars<double>& P, float* P32, hipStream_t s);
hipError_t launch_bounds(const BlockDims& b, const StepConsts& c, const GeomDev& gd, double* U, double* P, float* P32,
                         hipStream_t s, const Pack* pk = nullptr);
// End-of-step reductions and clock; ds: nds block states (a mesh), the first
// carrying the clock and totals.
hipError_t launch_finalize(const StepConsts& c, DeviceState* ds, const double* dt_part, int n_dt,
                           const double* divb_part, int n_divb, double* dt_hist, int hist_cap, hipStream_t s,
                           int nds = 1);
hipError_t launch_max_divb(const BlockDims& b, const Vars<const double>& U, double* part, hipStream_t s);
hipError_t launch_max_divb(const BlockDims& b, const double* U, double* part, hipStream_t s, const Pack* pk = nullptr);
hipError_t launch_reduce(const double* part, int n, int is_max, double* out, hipStream_t s);
hipError_t launch_convert_f32(const double* A, float* B, long n, hipStream_t s);
hipError_t launch_reset_state(DeviceState* ds, double dt, hipStream_t s);
hipError_t launch_set_clock(DeviceState* ds, double time, double dt, int ncycle, hipStream_t s);
// Task-level kernels over the entire block (the reference's IndexDomain::entire).
// Uout may alias Usub or Ufull (an in-place update).
hipError_t launch_weighted_sum(const BlockDims& b, const Vars<const double>& Usub, const Vars<const double>& Ufull,
                               const Vars<const double>& dUdt, double gam0, double gam1, double beta_dt,
                               const Vars<double>& Uout, hipStream_t s);
hipError_t launch_utop(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& U,
                       const Vars<double>& P, int* pflag, int* fflag, hipStream_t s);
hipError_t launch_floors(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<double>& U,
                         const Vars<double>& P, int* pflag, int* fflag, hipStream_t s);
hipError_t launch_fixup_entire(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<double>& U,
                               const Vars<double>& P, const int* pflag, hipStream_t s);
hipError_t launch_ptou(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const Vars<const double>& P,
                       const Vars<double>& U, hipStream_t s);
hipError_t launch_ptou(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P, double* U,
                       hipStream_t s);
// Periodic ghost fill: nvar <= 8 variables through their pointers, or any
// number of variables of a contiguous [nvar][nc] array.
hipError_t launch_sync(const BlockDims& b, const Vars<double>& A, int nvar, hipStream_t s);
hipError_t launch_sync(const BlockDims& b, double* A, int nvar, hipStream_t s);
hipError_t launch_sync_int(const BlockDims& b, int* A, hipStream_t s);

// ---- mesh of several blocks (kh_mesh.hip) ----
// Extents of a uniform mesh of nb[0] x nb[1] x nb[2] equal blocks: N interior
// zones per axis (nx = N / nb per block), Ne = N + 2 ng entire (1 for inactive
// axes); n, nc: one block's entire extent and zone count; ncm: the mesh's.
struct MeshDims {
    int N[3], Ne[3], nb[3], nx[3], n[3];
    int ng, ndim, nblk;
    int periodic[3];
    long nc, ncm;
};
// Mesh layout [8][Ne3][Ne2][Ne1] -> every block's [8][nc] (entire extents,
// periodic ghosts from their images), and back (interior from the owning
// block, physical ghosts from the edge block's own ghosts).
hipError_t launch_mesh_scatter(const MeshDims& m, const double* Pm, double* Pb, hipStream_t s);
hipError_t launch_mesh_gather(const MeshDims& m, const double* Pb, double* Pm, hipStream_t s);

// ---- adjoint (kh_adjoint.hip) ----
hipError_t launch_adj_fill(double* a, long n, double v, hipStream_t s);
// lamP += (d F / d P)^T lF over all directions.
hipError_t launch_adj_flux(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P,
                           const double* lF, double* lamP, hipStream_t s);
// A face-flux objective's seed: lF_dir[var] += w on the faces of index `face`
// along dir over the interior transverse range (dir 0: none).
struct FluxSeed { int dir = 0, var = 0, face = 0; double w = 0.; };
// lF ([3][8][nc]) = (d dudt / d F)^T ld (+ seed); lemf is [3][nc] scratch.
hipError_t launch_adj_divct(const BlockDims& b, const double* ld, double* lF, double* lemf, hipStream_t s,
                            FluxSeed seed = FluxSeed());
// accumulate = false overwrites lUsub / lUfull on the interior instead of adding to them.
// ffs (per-zone floor flags of the stage, or null: no zone floored) with Ustar
// (the stage's U before recovery), Psub (its input primitives, the recovery's
// guess) and scratch Pinv, lPinv, lPtot ([8][nc]); lPgeo (or null) receives the
// geometric source's transpose; pfs (inversion flags, or null), gate (device
// count of the stage's failed inversions) and lPfx (launch_adj_fixup's output).
hipError_t launch_adj_stage(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* Pout,
                            const double* lPout, const double* lUout, double g0, double g1, double beta_dt, bool same,
                            bool accumulate, double* lUsub, double* lUfull, double* ld, int* err, const double* Ustar,
                            const double* Psub, const int* ffs, double* lPgeo, const int* pfs, const int* gate,
                            const double* lPfx, double* Pinv, double* lPinv, double* lPtot, hipStream_t s);
// Transpose of the inversion fixup (fixup_list_kernel): lPfx[8][nc] (every zone;
// fold periodic images afterwards) = the fluid adjoints that zones failing to
// invert (pfs > 0) pass to the neighbours they were averaged from.  The stage
// transpose then adds lPfx (pfs, lPfx non-null) and gives failed zones only
// their P.B = U*.B / gdet dependence.
hipError_t launch_adj_fixup(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* Pout,
                            const double* lPout, const double* lUout, const int* pfs, const int* gate, double* lPfx,
                            hipStream_t s);
// Periodic images of A folded onto the interior (no ghost clearing), skipped
// when *gate == 0 (gate may be null).
hipError_t launch_adj_fold_interior(const BlockDims& b, double* A, int nvar, const int* gate, hipStream_t s);
// lP (interior fluid part) <- failed zones 0, others + lPfx.
hipError_t launch_adj_fixup_apply(const BlockDims& b, const int* pfs, const double* lPfx, double* lP, hipStream_t s);
// Transpose of the physical-boundary ghost fill of P (bounds_kernel), gathered
// into the zones it reads; ghosts are left for launch_adj_fold to clear.  C is
// [8][nc] scratch (its ghost entries are written and read).
hipError_t launch_adj_bounds(const BlockDims& b, const GeomDev& gd, const double* P, double* A, double* C,
                             hipStream_t s);
// X1 rows of fix_flux: x1_flux_rows(b) saved mass fluxes per stage.
long x1_flux_rows(const BlockDims& b);
hipError_t launch_save_x1_flux(const BlockDims& b, const double* F, double* x1f, hipStream_t s);
hipError_t launch_adj_fix_flux(const BlockDims& b, double* lF, const double* x1f, hipStream_t s);
hipError_t launch_adj_seed(const BlockDims& b, int dir, int var, int face, double w, double* lF, hipStream_t s);
hipError_t launch_adj_ptou(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P,
                           const double* lU, double* lP, hipStream_t s);
hipError_t launch_adj_fold(const BlockDims& b, double* A, int nvar, hipStream_t s);
// lP += (dS/dP)^T ld for the geometric source (curved geometry; interior).
hipError_t launch_adj_geo(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* P, const double* ld,
                          double* lP, hipStream_t s);
// lPin = (dfloors/dP at Pin)^T lPout (interior; frames 0 and 2).
hipError_t launch_adj_floors(const BlockDims& b, const StepConsts& c, const GeomDev& gd, const double* Pin,
                             const double* lPout, double* lPin, hipStream_t s);
// err[0] += zones with pflag != 0, or (fflag non-null) fflag != 0.
hipError_t launch_adj_count_flags(const BlockDims& b, const int* pflag, const int* fflag, int* err, hipStream_t s);
hipError_t launch_estimate_dt(const BlockDims& b, const StepConsts& c, int prec, const void* cmax, const void* cmin,
                              double* part, hipStream_t s);
// kh_estimate_timestep_noalloc: the estimate with its partials in a static device
// array of KH_NOALLOC_PARTIALS doubles (one per work-group of BS interior
// zones) and the clipped timestep written to the device double ndt; no
// allocation, no synchronisation.  hipErrorInvalidValue when the block needs
// more partials than the array holds.
constexpr long KH_NOALLOC_PARTIALS = 1L << 16;
constexpr int KH_MAX_DEVICES = 64;
hipError_t launch_estimate_dt_noalloc(const BlockDims& b, const StepConsts& c, int prec, const void* cmax,
                                      const void* cmin, double dt_last, double* ndt, hipStream_t s);

}  // namespace kh
