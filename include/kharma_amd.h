/*
 * kharma_amd.h -- C ABI of the MI355X (gfx950) GRMHD step.
 *
 * The entry points replace, one for one, the task functions KHARMA's driver
 * schedules for one explicit GRMHD stage (driver/kharma_step.cpp:77-319).  Each
 * declaration cites the reference interface it stands in for.  All data
 * pointers are NON-OWNING device pointers laid out like a Parthenon
 * VariablePack of one meshblock:
 *
 *   cell field A[var][k][j][i] over the entire block (interior + ng ghosts on each
 *   side of every active dimension), i fastest; var stride = n1*n2*n3 ("nc").
 *   Primitive / conserved order: rho, u, u1, u2, u3, B1, B2, B3.
 *   Fluxes F[dir-1][var][k][j][i]; face index i is the LEFT face of cell i.
 *   Signal speeds cmax/cmin[dir-1][k][j][i] (cmin stored positive).
 *   Flags pflag/fflag: int32 [k][j][i] (the reference stores them as Real fields).
 *
 * With precision KH_PREC_SPLIT the flux-like arrays (F, cmax, cmin, emf) hold
 * float instead of double; U and P are always double.
 *
 * Every function returns 0 (KH_OK) or a nonzero status; no C++ exception
 * crosses this boundary.  kh_last_error() describes the last failure of the
 * calling thread.  Work is enqueued on the block's stream and is asynchronous
 * unless a function says it synchronises.
 */
#ifndef KHARMA_AMD_H
#define KHARMA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KH_OK 0
#define KH_ERR_INVALID 1
#define KH_ERR_HIP 2
#define KH_ERR_UNSUPPORTED 3
#define KH_ERR_DIVB 4                  /* corner divB above b_field/kill_on_divb_over (kh_driver_check) */

#define KH_PREC_FP64 0
#define KH_PREC_SPLIT 1

#define KH_RECON_DONOR_CELL_C 0
#define KH_RECON_LINEAR_MC 1
#define KH_RECON_WENO5 2
#define KH_RECON_PPM 3

#define KH_COORDS_CARTESIAN_MINKOWSKI 0
#define KH_COORDS_KS 1                 /* spherical Kerr-Schild base */

#define KH_TRANSFORM_NULL 0
#define KH_TRANSFORM_MKS 1             /* "modified" */
#define KH_TRANSFORM_FMKS 2            /* "funky" */

#define KH_BC_OUTFLOW 1
#define KH_BC_REFLECT 2

/* driver/type (driver/kharma_driver.cpp:60-70): the task ordering of a stage.
 * "kharma" is MakeDefaultTaskCollection (driver/kharma_step.cpp:77-319), "simple"
 * MakeSimpleTaskCollection (driver/simple_step.cpp:39-168).  On one block both
 * orderings invert, floor and fix the same zones from the same values; they
 * differ only in the physical boundary fill, which "simple" makes from the
 * synced primitives (see kh_apply_bounds). */
#define KH_DRIVER_KHARMA 0
#define KH_DRIVER_SIMPLE 1

/* inverter/type (inverter/inverter.cpp:53-70): Inverter::u_to_p<onedw>
 * (inverter/onedw.hpp:89-233) or u_to_p<kastaun> (inverter/kastaun.hpp:83-400),
 * the reference's default, whose own defaults are err_tol 1e-12, iter_max 25. */
#define KH_INVERTER_ONEDW 0
#define KH_INVERTER_KASTAUN 1
#define KH_BSOLVER_FLUX_CT 0
#define KH_BSOLVER_FACE_CT 1
/* b_field/ct_scheme (b_ct.cpp:66-68): bs99 (flux average), gs05_0, gs05_c
 * (upwinded cell-centre EMF correction; sg07 is the same scheme) */
#define KH_CT_BS99 0
#define KH_CT_GS05_0 1
#define KH_CT_GS05_C 2
#define KH_CT_SG07 3

/* Run parameters of one meshblock: the .par options this path reads
 * (GRMHD/*, flux/*, floors/*, inverter/*, parthenon/time/*, parthenon/mesh/*). */
typedef struct kh_params {
    int nx1, nx2, nx3, ng;          /* interior zones (nx3 == 1 for 2D, nx2 == nx3 == 1 for 1D); ghost depth */
    double x1min, x1max, x2min, x2max, x3min, x3max;
    int periodic1, periodic2, periodic3;
    double gam;                     /* GRMHD/gamma */
    double cfl, dt_min, dt_start, max_dt_increase, tlim;
    int recon;                      /* KH_RECON_* (flux/reconstruction) */
    int use_hlle;                   /* flux/type == hlle */
    int floors_on;                  /* Floors package loaded */
    int recon_floors;               /* flux/reconstruction_floors */
    int floor_frame;                /* 0 fluid, 1 normal, 2 drift (floors/frame) */
    double rho_min_geom, u_min_geom, rho_min_const, u_min_const, r_char;
    double bsq_over_rho_max, bsq_over_u_max, u_over_rho_max, ktot_max, gamma_max;
    int use_r_char, temp_adjust_u;
    int iter_max;                   /* inverter/iter_max (onedw) */
    double err_tol;                 /* inverter/err_tol */
    int fix_average, fix_atmo;      /* inverter/fix_average_neighbors, fix_atmosphere */
    int integrator;                 /* 0 rk2, 1 vl2 (parthenon/time/integrator) */
    int coords;                     /* KH_COORDS_* (coordinates/base) */
    int precision;                  /* KH_PREC_FP64 or KH_PREC_SPLIT */
    /* Kerr-Schild geometry (coordinates/a, transform, hslope, mks_smooth, poly_xt, poly_alpha,
     * fmks_zero_point); x1 is log(r) and x2 the stretched polar angle for MKS/FMKS. */
    double a;
    int transform;                  /* KH_TRANSFORM_* */
    double hslope, mks_smooth, poly_xt, poly_alpha, fmks_zero_point;
    /* Physical boundaries of non-periodic x1/x2 faces (boundaries/inner_x1 ...): KH_BC_* */
    int bc_inner_x1, bc_outer_x1, bc_inner_x2, bc_outer_x2;
    int check_inflow_inner_x1, check_inflow_outer_x1;  /* boundaries/check_inflow_* */
    int zero_polar_flux;            /* boundaries/zero_polar_flux */
    int fix_polar_flux;             /* b_field/fix_polar_flux */
    int fix_corner_inner;           /* boundaries/fix_corner_inner */
    int driver_type;                /* KH_DRIVER_* (driver/type) */
    int kill_on_large_divb;         /* b_field/kill_on_large_divb (the reference's default: on) */
    double kill_on_divb_over;       /* b_field/kill_on_divb_over (default 1e-3) */
    int inverter;                   /* KH_INVERTER_* (inverter/type) */
    /* First-order flux correction (fofc/on, polar_cells, eh_buffer; flux/flux.cpp:175-215)
     * with the floors' own prescription, LLF speeds from both states and no source term
     * in the trial update.  FP64 form and the default driver only. */
    int fofc;
    int fofc_polar_cells;
    double fofc_eh_buffer;
    /* b_field/solver: KH_BSOLVER_FLUX_CT (B_FluxCT, cell-centred B) or
     * KH_BSOLVER_FACE_CT (B_CT, face-centred B, b_ct/b_ct.cpp), and
     * b_field/ct_scheme (KH_CT_*) for the latter. */
    int b_solver;
    int ct_scheme;
} kh_params;

typedef struct kh_block kh_block;
typedef struct kh_driver kh_driver;

const char* kh_last_error(void);
int kh_version(void);

/* Geometry and index ranges of one meshblock (replaces GRCoordinates'
 * construction, coordinates/gr_coordinates.cpp:64-77, for the block).
 * Allocates only a small reduction scratch area. */
int kh_block_create(const kh_params* p, kh_block** out);
void kh_block_destroy(kh_block* b);
/* hipStream_t, or NULL for the default stream (one stream per call sequence) */
int kh_block_set_stream(kh_block* b, void* hip_stream);
/* entire sizes and the variable stride nc */
int kh_block_shape(const kh_block* b, int* n1, int* n2, int* n3, long* nc);

/* Flux::GetFlux<Recon,dir> (flux/get_flux.hpp:60-373): reconstruct P to the faces
 * normal to dir (1..3), apply reconstruction floors, compute the face fluxes
 * (LLF or HLLE) and the signal speeds.  F_dir / cmax_dir / cmin_dir point at this
 * direction's [var][nc] and [nc] arrays.  Faces written: those the step consumes,
 * i.e. X1: i in [is, ie+1], j in [js-1, je+1], k in [ks-1, ke+1] (cyclic for X2/X3). */
int kh_get_flux(kh_block* b, int dir, const double* P, void* F_dir, void* cmax_dir, void* cmin_dir);

/* Packages::FixFlux (kharma_package.cpp:43-56): B_FluxCT::FixFlux
 * (b_flux_ct/b_flux_ct.cpp:264-370: polar ZeroBoundaryFlux, then FluxCT) and
 * KBoundaries::FixFlux (boundaries/boundaries.cpp:611-690: no mass inflow through
 * x1 faces, zero flux through polar faces).  Corner EMFs into emf[3][nc] and
 * the rewritten fluxes in F[3][8][nc]. */
int kh_flux_ct(kh_block* b, void* F, void* emf);

/* KHARMADriver::FluxDivergence (driver/kharma_driver.hpp:208-231): dUdt[8][nc] on the interior. */
int kh_flux_divergence(kh_block* b, const void* F, double* dUdt);

/* Flux::AddGeoSource (flux/flux.cpp:395-444): dUdt += sqrt(-g) T^mu_nu Gamma terms
 * on the interior, from the primitives P[8][nc] (no-op in Cartesian Minkowski). */
int kh_add_geo_source(kh_block* b, const double* P, double* dUdt);

/* Update::WeightedSumData pair of KHARMADriver::AddStateUpdate
 * (driver/kharma_driver.cpp:339-387), entire block:
 * Uout = gam0*Usub + gam1*Ufull; Uout = Uout + beta_dt*dUdt.
 * Uout may be Usub or Ufull (the reference updates in place, :366-370). */
int kh_weighted_sum(kh_block* b, const double* Usub, const double* Ufull, const double* dUdt,
                    double gam0, double gam1, double beta_dt, double* Uout);

/* Inverter::BlockUtoP (inverter/inverter.cpp:131-200) with the solver of
 * kh_params.inverter (inverter/onedw.hpp:89-233 or inverter/kastaun.hpp:83-400)
 * over the physical range (interior, plus the ghosts of periodic dimensions),
 * which is what BlockUtoP sweeps whatever domain it is given
 * (inverter.cpp:162), then B_FluxCT::BlockUtoP (b_flux_ct.cpp:195-217) over
 * the entire block.  P holds the guess on entry. */
int kh_utop(kh_block* b, const double* U, double* P, int* pflag, int* fflag);

/* Floors::ApplyGRMHDFloors (floors/floors.cpp:233-250, floors_impl.hpp:45-128), entire block. */
int kh_apply_floors(kh_block* b, double* U, double* P, int* pflag, int* fflag);

/* Inverter::FixUtoP (inverter/fixup.cpp:47-156), entire block. */
int kh_fix_utop(kh_block* b, double* U, double* P, const int* pflag);

/* Flux::MeshPtoU (flux/flux.cpp:276-322), entire block. */
int kh_ptou(kh_block* b, const double* P, double* U);

/* Physical boundaries of a block with outflow / reflecting x1 or x2 faces: the
 * ghost zones of those faces in U and P[8][nc] from the interior, as
 * KBoundaries::ApplyBoundary leaves them (boundaries/boundaries.cpp:369-589):
 * outflow or reflecting fill, check_inflow on x1 faces, B = U.B / sqrt(-g) and
 * the fluid conserved variables from P.  Faces in the order inner/outer x1,
 * inner/outer x2, then inner x1 again when fix_corner_inner.  The result is the
 * state after the stage's repeated boundary passes (kharma_step.cpp:240-314),
 * whose fluid conserved variables use the updated ghost B.  With driver_type
 * KH_DRIVER_SIMPLE the fill is that of the synced primitives
 * (boundaries.cpp:561-571): all 8 conserved variables of a filled zone come
 * from its primitives, so U.B = (U.B_src / sqrt(-g)) sqrt(-g) there. */
int kh_apply_bounds(kh_block* b, double* U, double* P);

/* Boundary exchange of nvar cell variables of a single periodic block with itself
 * (the parthenon exchange behind KHARMADriver::AddBoundarySync, kharma_driver.cpp:138-183). */
int kh_sync_bounds(kh_block* b, double* A, int nvar);

/* GRMHD::EstimateTimestep in-loop branch (grmhd/grmhd.cpp:310-339): CFL minimum over
 * the interior from cmax/cmin[3][nc], clipped to [dt_min, max_dt_increase*dt_last].
 * Synchronises the stream. */
int kh_estimate_timestep(kh_block* b, const void* cmax, const void* cmin, double dt_last, double* ndt);

/* B_FluxCT::MaxDivB (b_flux_ct.cpp:566-607).  Synchronises the stream. */
int kh_max_divb(kh_block* b, const double* U, double* out);


/* kh_estimate_timestep for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The same estimate and clip, with the
 * per-work-group partials in a static device array of 65536 doubles (512 KiB
 * per GPU, held until the process ends) instead of the block's buffer: the
 * block's interior may hold at most 65536 * 256 zones (KH_ERR_INVALID beyond).
 * The clipped timestep is written to the device double *ndt_device; the call
 * neither allocates nor synchronises.  Calls that overlap on one GPU share the
 * array and race. */
int kh_estimate_timestep_noalloc(kh_block* b, const void* cmax, const void* cmin, double dt_last, double* ndt_device);

/* ---- Pack view: the variables where KHARMA keeps them ---------------------
 * KHARMA holds each variable group as its own Parthenon field: prims.rho,
 * prims.u, prims.uvec (3 components), prims.B (3) and the cons.* equivalents
 * (grmhd/grmhd.cpp:170-185, b_flux_ct/b_flux_ct.cpp:116-118), each with its
 * own flux arrays, and reaches them through the VarMap indices of a pack
 * (types.hpp:126-218, grmhd/pack.hpp:51-89).  A kh_vars names the 8 variables
 * of one container by base pointer, in the order rho, u, u1, u2, u3, B1, B2,
 * B3; each points at a [k][j][i] array over the entire block.  Nothing
 * relates the 8 pointers to each other.  The *_v entry points below read and
 * write the fields in place; the contiguous entry points above are the case
 * var[n] = base + n * nc and run the same kernels.  Flux containers are one
 * kh_vars per direction (F[dir-1]), float in the split form. */
typedef struct kh_vars { void* var[8]; } kh_vars;

/* A pack's VarMap (types.hpp:126-218): the pack index of rho, u, the first
 * velocity and the first field component (U2 = U1 + 1, ..., B3 = B1 + 2). */
typedef struct kh_varmap { int RHO, UU, U1, B1; } kh_varmap;

/* kh_vars of a pack given as its per-variable base pointers (pack_var[v] is the
 * device array of pack variable v, v < nvar) and its VarMap; KH_ERR_INVALID
 * when an index is absent (-1) or out of range.  Host only. */
int kh_vars_from_pack(void* const* pack_var, int nvar, const kh_varmap* m, kh_vars* out);

int kh_get_flux_v(kh_block* b, int dir, const kh_vars* P, const kh_vars* F_dir, void* cmax_dir, void* cmin_dir);
int kh_flux_ct_v(kh_block* b, const kh_vars F[3], void* emf);
int kh_flux_divergence_v(kh_block* b, const kh_vars F[3], const kh_vars* dUdt);
int kh_add_geo_source_v(kh_block* b, const kh_vars* P, const kh_vars* dUdt);
/* Uout may be Usub or Ufull (AddStateUpdate updates in place, kharma_driver.cpp:366-370). */
int kh_weighted_sum_v(kh_block* b, const kh_vars* Usub, const kh_vars* Ufull, const kh_vars* dUdt, double gam0,
                      double gam1, double beta_dt, const kh_vars* Uout);
int kh_utop_v(kh_block* b, const kh_vars* U, const kh_vars* P, int* pflag, int* fflag);
int kh_apply_floors_v(kh_block* b, const kh_vars* U, const kh_vars* P, int* pflag, int* fflag);
int kh_fix_utop_v(kh_block* b, const kh_vars* U, const kh_vars* P, const int* pflag);
int kh_ptou_v(kh_block* b, const kh_vars* P, const kh_vars* U);
int kh_apply_bounds_v(kh_block* b, const kh_vars* U, const kh_vars* P);
/* the first nvar (1..8) variables of A */
int kh_sync_bounds_v(kh_block* b, const kh_vars* A, int nvar);
int kh_max_divb_v(kh_block* b, const kh_vars* U, double* out);

/* pflag / fflag between this library's int32 arrays and the reference's
 * Real-valued fields (inverter.cpp:104-108, floors.cpp:111-116), n entries;
 * Real -> int truncates as the reference's static_cast<int> reads do. */
int kh_flags_to_real(kh_block* b, const int* flag, double* real_flag, long n);
int kh_flags_from_real(kh_block* b, const double* real_flag, int* flag, long n);

/* ---- First-order flux correction at the task level ------------------------
 * The tasks of KHARMADriver::AddFOFC (driver/kharma_driver.cpp:292-337) after
 * the trial update (kh_flux_divergence_v of the raw fluxes, kh_weighted_sum_v
 * into the guess, kh_utop_v of the guess from the stage's primitives).  Need
 * kh_params.fofc; FP64 form. */
/* Floors::DetermineGRMHDFloors (floors/floors.cpp:197-231), entire block:
 * fflag |= the floor bits of P's zones, P unchanged.  No-op without floors. */
int kh_determine_floors_v(kh_block* b, const kh_vars* P, int* fflag);
/* Flux::MarkFOFC (flux/fofc.cpp:45-111), entire block: fofcflag = 1 where
 * fflag != 0, within fofc_eh_buffer of the horizon, or in the first
 * fofc_polar_cells rows at physical X2 faces (spherical coordinates); else 0. */
int kh_mark_fofc(kh_block* b, const int* fflag, int* fofcflag);
/* Flux::FOFC (flux/fofc.cpp:113-226) with fofc/use_glf off and cell-centred B:
 * every face the step consumes (as kh_get_flux) next to a marked zone gets the
 * LLF flux of the two zones' own primitives P, and those states' signal speeds
 * in cmax/cmin ([3][nc], cmin stored positive).  A ghost zone's mark is read
 * from its periodic image or, across a physical face, from the nearest interior
 * zone (the synced fofcflag of the reference). */
int kh_fofc_v(kh_block* b, const kh_vars* P, const int* fofcflag, const kh_vars F[3], double* cmax, double* cmin);

/* ---- Face-centred constrained transport (the B_CT package, b_ct/b_ct.cpp) ----
 * With kh_params.b_solver = KH_BSOLVER_FACE_CT the magnetic field is
 * cons.fB = sqrt(-g) B^d on cell faces: face and edge arrays are
 * [3][n3+1][n2+1][n1+1] doubles (components F1/F2/F3 or E1/E2/E3; element
 * (k, j, i) is the lower face / edge of cell (k, j, i)), nf = (n1+1)(n2+1)(n3+1)
 * per component.  Built for periodic Cartesian Minkowski blocks in 2-D and 3-D
 * (the reference's boundary EMF treatments, b_ct_boundaries.cpp, rest on
 * Parthenon's face-field fills), the default driver, both precisions. */
int kh_block_face_count(const kh_block* b, long* nf);
/* B_CT::CalculateEMF (b_ct.cpp:263-416): edge EMFs emf[3][nf] over i in [is, ie+1],
 * j in [js, je+1], k in [ks, ke+1] from the B fluxes F[3] (flux element type by
 * precision) and, for gs05_0 / gs05_c / sg07, the cell-centre EMF of P. */
int kh_ct_calculate_emf_v(kh_block* b, const kh_vars* P, const kh_vars F[3], double* emf);
int kh_ct_calculate_emf(kh_block* b, const double* P, const void* F, double* emf);
/* B_CT::AddSource (b_ct.cpp:418-470): dfB[3][nf] = dB_Uf/dt on the interior faces
 * (F_d's own index to e+1; F3 one layer in 2-D) from the edge circulation. */
int kh_ct_add_source(kh_block* b, const double* emf, double* dfB);
/* KHARMADriver::WeightedSumDataFace (driver/kharma_driver.hpp:167-190): z = w1 x + w2 y
 * over [3][nf]; z may alias x or y. */
int kh_ct_weighted_sum_face(kh_block* b, const double* x, const double* y, double w1, double w2, double* z);
/* The periodic face-field exchange of the block with itself (the fB part of the
 * boundary sync, kharma_step.cpp:129-138): ghost faces from their images. */
int kh_ct_sync_faces(kh_block* b, double* fB);
/* B_CT::BlockUtoP (b_ct.cpp:173-205) over the entire block: P.B from the face
 * averages, U.B = sqrt(-g) P.B (ghost faces must be current). */
int kh_ct_utop_v(kh_block* b, const double* fB, const kh_vars* U, const kh_vars* P);
int kh_ct_utop(kh_block* b, const double* fB, double* U, double* P);
/* B_CT::MaxDivB (b_ct.cpp:472-497): max |face divergence| over the interior.
 * Synchronises the stream. */
int kh_ct_max_divb(kh_block* b, const double* fB, double* out);

/* ---- Problem set-up and geometry (host only; no device work) -------------- */

/* Fishbone-Moncrief torus with a SANE seed field (torus/rin, rmax, kappa;
 * perturbation/u_jitter, rng_seed; b_field/beta_min, min_A). */
typedef struct kh_torus_params {
    double rin, rmax, kappa;
    double u_jitter;                /* u *= 1 + U(-u_jitter/2, u_jitter/2) */
    int rng_seed;                   /* std::mt19937 seed (31337 + block gid in the reference) */
    double beta_min;                /* max(P_gas) / max(P_B) after normalisation */
    double min_A;
} kh_torus_params;

/* Initial primitives of the torus problem into the host array P[8][nc]:
 * InitializeFMTorus (prob/fm_torus.cpp:41-197) with its initial floors, PerturbU
 * (prob/utils/perturbation.hpp:49-92, serial path), SeedBField<sane> with the
 * corner-averaged curl and NormalizeBField (prob/seed_B.cpp:81-489).  Interior
 * zones only; kh_driver_init_prims fills the ghosts. */
int kh_torus_prims(const kh_params* p, const kh_torus_params* t, double* P);

/* One component of the geometry cache of the block `p` describes
 * (GRCoordinates cache, coordinates/gr_coordinates.cpp:96-205), evaluated on the
 * host into out[(n2+1)*(n1+1)] (j-major).  loc: 0 centre, 1 X1 face, 2 X2 face,
 * 3 corner.  comp: 0..9 gcov (00 01 02 03 11 12 13 22 23 33), 10..19 gcon, 20 gdet,
 * 21 lapse, 22..85 gdet*Gamma^mu_{nu lam} at centres (index 22 + 16 mu + 4 nu + lam),
 * 86/87 geometric rho/u floors at centres, 88 radius of the centre. */
#define KH_GEOM_CONN 22
#define KH_GEOM_FLOOR_RHO 86
#define KH_GEOM_FLOOR_U 87
#define KH_GEOM_RADIUS 88
int kh_geometry_host(const kh_params* p, int loc, int comp, double* out);

/* ---- Stage driver --------------------------------------------------------
 * KHARMADriver::MakeDefaultTaskCollection (driver/kharma_step.cpp:77-319) for a
 * single periodic meshblock, with Parthenon's RK2/VL2 stage containers and
 * KHARMA's dt control (kharma_driver.cpp:439-460).  The driver owns its
 * containers (U, P for the base and stage-1 states, fluxes, flags).  dt is
 * computed and consumed on the device; kh_driver_step does not synchronise. */
int kh_driver_create(const kh_params* p, kh_driver** out);
void kh_driver_destroy(kh_driver* d);
int kh_driver_set_stream(kh_driver* d, void* hip_stream);
/* Initial primitives P[8][nc] (device pointer): sync, P->U over the entire
 * block, dt = dt_start (KHARMA::PostInitialize, prob/post_initialize.cpp:55-193). */
int kh_driver_init_prims(kh_driver* d, const double* P);
/* The face_ct driver's initial state: primitives P[8][nc] (B ignored) and face
 * fields fB[3][nf] as seeded (prob/seed_B.cpp:137-191, 336-353); face sync,
 * B_CT::BlockUtoP, then what kh_driver_init_prims does.  kh_driver_get_faces
 * copies the base container's face fields out (ghost faces synced).  The step
 * runs B_CT::CalculateEMF and AddSource in place of the flux-CT rewrite
 * (kharma_step.cpp:175-202) and the per-step divB is B_CT::MaxDivB. */
int kh_driver_init_faces(kh_driver* d, const double* P, const double* fB);
int kh_driver_get_faces(kh_driver* d, double* fB);
int kh_driver_step(kh_driver* d, int nsteps);
/* Set the clock as a restart does (Parthenon SimTime time / dt / ncycle restored,
 * KHARMA::ResetGlobals, kharma.cpp:107-119): the next step takes dt. */
int kh_driver_set_clock(kh_driver* d, double time, double dt, int ncycle);
/* Copy the base-container state out (device pointers; either may be NULL). */
int kh_driver_get_state(kh_driver* d, double* P, double* U);
/* Restart dump and reload: the fields KHARMA marks Metadata::Restart -- the
 * primitive and conserved variables (driver/kharma_driver.cpp:114-118,
 * grmhd/grmhd.cpp:155-156) -- over the entire block, and the clock (time, dt,
 * ncycle; KHARMA::ResetGlobals, kharma.cpp:107-119).  Parthenon writes these
 * to HDF5 (.rhdf); this is a flat binary file: a header (magic "KHAMDRST",
 * version 1, entire extents, ncycle, time, dt, the run's kh_params) followed
 * by P and U as [8][n3][n2][n1] doubles.  Loading checks the mesh, geometry and
 * gamma against the driver's, restores U and P exactly (not U from P) and the
 * clock, so the run continues bit for bit.  Both synchronise. */
int kh_driver_save_restart(kh_driver* d, const char* path);
int kh_driver_load_restart(kh_driver* d, const char* path);
/* Inversion / floor flags of the last stage, int32 [nc] device arrays (either may be NULL):
 * the reference's pflag / fflag fields (inverter/inverter.cpp:104-108, floors/floors.cpp:111-116). */
int kh_driver_get_flags(kh_driver* d, int* pflag, int* fflag);
/* PrintGlobalMaxDivB's stop (b_flux_ct/b_flux_ct.cpp:620-643): with
 * kill_on_large_divb, each step's maximum corner divB is compared on the device
 * with kill_on_divb_over and the first step above it is latched, so steps stay
 * asynchronous.  kh_driver_check synchronises and returns KH_ERR_DIVB (with the
 * cycle and value in kh_last_error) once a step has exceeded it, KH_OK
 * otherwise.  kh_driver_init_prims clears the latch. */
int kh_driver_check(kh_driver* d);
/* Synchronises; any output may be NULL. */
int kh_driver_get_scalars(kh_driver* d, double* time, double* dt, double* last_divb, int* ncycle, int* total_fails);
/* dt taken by each of the first n steps since init (host array). Synchronises. */
int kh_driver_get_dt_history(kh_driver* d, double* out, int n);
int kh_driver_block(kh_driver* d, kh_block** out);
/* Device pointers to the driver's base-container arrays (U, P: [8][nc] doubles). */
int kh_driver_arrays(kh_driver* d, double** U, double** P);
/* Per-kernel timing with HIP events recorded on the driver's stream around every
 * launch of kh_driver_step while enabled (no counterpart in the reference; it
 * plays the role of the Kokkos-tools kernel timer hooked in run.sh:11-91).
 * Kernel kinds: 0..2 face fluxes X1..X3, 3 physical-boundary flux fixes and ghost fills, 4 stage-1 update,
 * 5 stage-2 update, 6 inversion fix-up, 7 divB, 8 timestep finalize. */
#define KH_NKERNEL_KINDS 9
int kh_driver_set_kernel_timing(kh_driver* d, int enable);
/* Face-flux objective accumulated by kh_driver_step (no counterpart in the
 * reference; the differentiation form's time-integrated flux, cf. the horizon
 * fluxes of grmhd.cpp:221-233): J += weight * sum_stages w_s dt sum F_dir[var] dA
 * over the faces of index `face` (entire-block index along dir) and the interior
 * transverse range, from the fluxes after the physical-boundary fixes; w_s are
 * the RK weights of the stages' fluxes in the step (1/2, 1/2 for rk2; 0, 1 for
 * vl2), dA the product of the two transverse zone widths.  var in 0..4;
 * dir = 0 switches it off. */
int kh_driver_set_face_flux_objective(kh_driver* d, int dir, int var, int face, double weight);
/* Step records for the differentiation form (no counterpart in the reference,
 * whose tasks keep no per-stage history).  kh_driver_stage_record_bytes gives
 * the size of one step's record.  With a record set (a device buffer of that
 * size, 256-byte aligned), every following kh_driver_step step of the FP64
 * driver writes into it what kh_adjoint_step_recorded needs in place of a
 * recompute: stage 1's output primitives, each stage's U* (curved space),
 * inversion and floor flags, clamped X1 mass fluxes and failed-inversion
 * count.  Each step overwrites the record; NULL switches recording off. */
int kh_driver_stage_record_bytes(kh_driver* d, size_t* bytes);
int kh_driver_set_stage_record(kh_driver* d, void* rec);
/* The accumulated objective (0 if never set); reset != 0 zeroes it.  Synchronises. */
int kh_driver_objective(kh_driver* d, double* value, int reset);
/* Sum of launch durations (ms) and launch counts per kind since the last call;
 * synchronises and resets the accumulators. */
int kh_driver_kernel_times(kh_driver* d, double* total_ms, int* launches, int nkinds);

/* ---- Several meshblocks on one GPU ---------------------------------------
 * A uniform mesh of nb1 x nb2 x nb3 equal meshblocks (parthenon/mesh/nx* split
 * into parthenon/meshblock/nx*), advanced by MakeDefaultTaskCollection
 * (driver/kharma_step.cpp:77-319) with the ghost zones of shared faces kept
 * current as the boundary exchange does (KHARMADriver::AddBoundarySync,
 * kharma_driver.cpp:138-183) and the timestep and divB reduced over all blocks
 * (kharma_driver.cpp:439-460, b_flux_ct.cpp:566-643).  p describes the whole
 * mesh (nx1..3 the mesh's zones, the mesh's bounds and boundaries); each block
 * holds nx/nb zones per axis, at least 2 ng along a split axis.  Per face a
 * block has either the mesh's physical boundary or its neighbour's images, the
 * neighbour of the last block of a periodic axis being the first.  The blocks'
 * states sit in one allocation per container, [block][8][nc], block index b1
 * fastest; each block's update stores the zones it shares into its
 * neighbours' ghosts, so no exchange pass runs.  The mesh state reads and
 * writes as one block of the whole mesh ([8][N3+2ng][N2+2ng][N1+2ng]).  The
 * same arithmetic per zone as the one-block driver: a mesh of several blocks
 * gives the one-block result bit for bit.  With fofc on, each block's marks
 * are stored with their images in its neighbours' ghosts (the flag's boundary
 * sync, kharma_driver.cpp:328-331) before any block replaces its fluxes, and
 * the polar rows are marked on the blocks of the physical X2 faces only. */
typedef struct kh_mesh kh_mesh;
int kh_mesh_create(const kh_params* p, int nb1, int nb2, int nb3, kh_mesh** out);
void kh_mesh_destroy(kh_mesh* m);
int kh_mesh_set_stream(kh_mesh* m, void* hip_stream);
/* block count and one block's entire sizes and zone count */
int kh_mesh_shape(const kh_mesh* m, int* nblocks, int* n1, int* n2, int* n3, long* nc);
/* Initial primitives of the whole mesh (device, mesh layout; the interior is
 * read, with the ghosts of physical faces refilled): P->U, boundaries, dt_start. */
int kh_mesh_init_prims(kh_mesh* m, const double* P);
int kh_mesh_step(kh_mesh* m, int nsteps);
/* Replay the mesh's step as a HIP graph (no counterpart in the reference, whose
 * task lists are rebuilt each cycle): with enable != 0 the next kh_mesh_step
 * captures one step (its kernel launches over all blocks) on a private stream
 * and every step after that launches the captured graph on the mesh's stream,
 * which removes the per-launch host cost of a mesh of small blocks.  The graph
 * holds the mesh's own buffers, so it stays valid for the mesh's life; enable
 * = 0 drops it.  Results are those of the eager step, bit for bit. */
int kh_mesh_set_graph(kh_mesh* m, int enable);
/* How the mesh step launches its kernels.  enable != 0 (the default): each
 * kernel once over all blocks, as KHARMA launches over a MeshBlockPack
 * (driver/kharma_step.cpp:77-319 with a pack of every block; the pack loops of
 * kharma_driver.cpp:339-387 and flux/get_flux.hpp:60-373 run one kernel over
 * (b, k, j, i)); enable = 0: once per block.  Results are the same bit for bit;
 * a change drops a captured graph. */
int kh_mesh_set_packed(kh_mesh* m, int enable);
int kh_mesh_set_clock(kh_mesh* m, double time, double dt, int ncycle);
/* The base-container state in the mesh layout (device; either may be NULL):
 * interior zones from their blocks, periodic ghosts as images, the ghosts of
 * physical faces from the edge blocks. */
int kh_mesh_get_state(kh_mesh* m, double* P, double* U);
int kh_mesh_get_scalars(kh_mesh* m, double* time, double* dt, double* last_divb, int* ncycle, int* total_fails);
int kh_mesh_get_dt_history(kh_mesh* m, double* out, int n);
int kh_mesh_check(kh_mesh* m);
/* Restart dump and reload of the mesh in the mesh layout: the same file as a
 * one-block run of the mesh (kh_driver_save_restart), readable by either. */
int kh_mesh_save_restart(kh_mesh* m, const char* path);
int kh_mesh_load_restart(kh_mesh* m, const char* path);
/* Block b's base-container arrays ([8][nc] each) and its context, for the task
 * entry points on one block of the mesh. */
int kh_mesh_block_arrays(kh_mesh* m, int b, double** U, double** P);
int kh_mesh_block(kh_mesh* m, int b, kh_block** out);

/* ------------------------------------------------------------------------
 * Differentiation form: the reverse-mode adjoint of one driver step.  No
 * counterpart in the reference (KHARMA has no adjoint); it transposes the
 * forward step of kh_driver_step stage by stage: WENO5 + LLF face fluxes
 * (flux/get_flux.hpp:60-373), flux-CT and divergence (b_flux_ct.cpp:307-370,
 * kharma_driver.hpp:208-231), the RK update (kharma_driver.cpp:339-387) and the
 * onedw recovery (inverter/onedw.hpp:89-233) by the implicit-function theorem
 * (lam_U = (dU/dP)^-T lam_P at the converged primitives), the geometric source
 * (flux/flux.cpp:395-444), the floors (floors/floors_impl.hpp:45-128; fluid and
 * drift frames), the physical-boundary flux fixes (b_flux_ct.cpp:372-420,
 * boundaries.cpp:611-690) and ghost fills with the inflow check
 * (boundaries.cpp:369-589, boundaries.hpp:101-136).  The timestep is a fixed
 * input of each step.  Zones that fail to invert are covered through the
 * transpose of their fixup (inverter/fixup.cpp:47-156).  Supported: FP64 form,
 * WENO5, LLF; Cartesian Minkowski (floored zones unsupported there) or
 * Kerr-Schild metric; any boundaries of the forward step.  A step with a zone
 * outside this returns KH_ERR_UNSUPPORTED.
 * ---------------------------------------------------------------------- */
typedef struct kh_adjoint kh_adjoint;
int kh_adjoint_create(const kh_params* p, kh_adjoint** out);
void kh_adjoint_destroy(kh_adjoint* a);
int kh_adjoint_set_stream(kh_adjoint* a, void* hip_stream);
/* lam_P <- (dP_{n+1}/dP_n)^T lam_P for the step P_n -> P_{n+1} taken with timestep dt.
 * P: the step's starting primitives ([8][nc] doubles, device; the interior is
 * used, ghosts are refilled).  P_next: the step's result P_{n+1} (the next
 * checkpoint; interior used), or NULL to recompute it.  lam_P ([8][nc],
 * device): in, dL/dP_{n+1} on the interior (ghost entries ignored); out,
 * dL/dP_n on the interior, zero ghosts.  Stage 1 is recomputed from P
 * (per-step checkpointing).  P_next is used on periodic Minkowski blocks only,
 * where the caller vouches that no zone of the step's second stage was floored
 * or failed to invert (the driver's flags after the step); otherwise stage 2 is
 * recomputed too.  With a face-flux objective set, lam_P also receives
 * dJ_n/dP_n for the step's share J_n of the objective.  Synchronises. */
int kh_adjoint_step(kh_adjoint* a, const double* P, const double* P_next, double dt, double* lam_P);
/* kh_adjoint_step from the record of the step from P to P_next
 * (kh_driver_set_stage_record, a driver with the same parameters) instead of
 * recomputing its stages.  P and P_next are read as the driver holds its
 * states (kh_driver_get_state): ghost zones filled, periodic images included.
 * Same coverage and errors as kh_adjoint_step.  Synchronises. */
int kh_adjoint_step_recorded(kh_adjoint* a, const double* P, const void* rec, const double* P_next, double dt,
                             double* lam_P);
/* Deferred checks (on != 0): the following adjoint steps neither synchronise nor
 * report zones they do not cover; the counts accumulate on the device until
 * kh_adjoint_check, which synchronises, returns the status a step would have
 * returned for them, and clears them.  kh_adjoint_last_times is not updated
 * while deferred.  Switching either way clears the counts. */
int kh_adjoint_set_deferred_checks(kh_adjoint* a, int on);
int kh_adjoint_check(kh_adjoint* a);
/* The objective of kh_driver_set_face_flux_objective, same arguments: its
 * derivative is added by every following kh_adjoint_step (dir = 0: none). */
int kh_adjoint_set_face_flux_objective(kh_adjoint* a, int dir, int var, int face, double weight);
/* Task-level transposes (device arrays over the entire block, FP64):
 *   kh_adjoint_flux:      lam_P  = sum_d (dF_d/dP)^T lam_F_d   (kh_get_flux, all dirs; lam_F [3][8][nc])
 *   kh_adjoint_flux_div:  lam_F  = (d dUdt/dF)^T lam_dUdt      (kh_flux_ct + kh_flux_divergence)
 *   kh_adjoint_utop:      lam_U  = (dU/dP)^-T lam_P            (kh_utop at converged P, interior)
 *   kh_adjoint_ptou:      lam_P += (dU/dP)^T lam_U             (kh_ptou, interior)
 *   kh_adjoint_bounds:    lam_P on ghost zones moved onto the zones the ghost fill reads
 *                         (physical faces as kh_apply_bounds fills P, with P.B from U.B =
 *                         gdet P.B, then periodic images); ghosts zeroed, in place
 *   kh_adjoint_fix_flux:  lam_F = (dF_fixed/dF)^T lam_F, in place, for the physical-boundary
 *                         flux fixes kh_flux_ct applies before its CT; F_fixed: the fixed
 *                         fluxes [3][8][nc] (their X1 mass fluxes decide the inflow clamp)
 *   kh_adjoint_geo_source: lam_P += (dS/dP)^T lam_dUdt          (kh_add_geo_source, interior)
 *   kh_adjoint_floors:    lam_P_in = (dP_out/dP_in)^T lam_P_out (kh_apply_floors' floors and
 *                         ceilings of each interior zone at P_in, fluid or drift frame)
 *   kh_adjoint_fixup:     lam_P <- (dP_fixed/dP)^T lam_P on the fluid primitives, in place, for
 *                         kh_fix_utop's replacement of the zones with pflag > 0 (int32 [nc])
 *                         by the weighted average of their neighbours; P_fixed: its output
 * Asynchronous on the adjoint's stream; outputs are overwritten except where noted. */
int kh_adjoint_flux(kh_adjoint* a, const double* P, const double* lam_F, double* lam_P);
int kh_adjoint_flux_div(kh_adjoint* a, const double* lam_dUdt, double* lam_F);
int kh_adjoint_utop(kh_adjoint* a, const double* P, const double* lam_P, double* lam_U);
int kh_adjoint_ptou(kh_adjoint* a, const double* P, const double* lam_U, double* lam_P);
int kh_adjoint_bounds(kh_adjoint* a, const double* P, double* lam_P);
int kh_adjoint_fix_flux(kh_adjoint* a, const double* F_fixed, double* lam_F);
int kh_adjoint_geo_source(kh_adjoint* a, const double* P, const double* lam_dUdt, double* lam_P);
int kh_adjoint_floors(kh_adjoint* a, const double* P_in, const double* lam_P_out, double* lam_P_in);
int kh_adjoint_fixup(kh_adjoint* a, const double* P_fixed, const int* pflag, double* lam_P);
/* Milliseconds of the last kh_adjoint_step spent recomputing the forward step
 * and in the adjoint kernels (HIP events on the adjoint's stream). */
int kh_adjoint_last_times(kh_adjoint* a, double* forward_ms, double* adjoint_ms);

#ifdef __cplusplus
}
#endif

#endif /* KHARMA_AMD_H */
