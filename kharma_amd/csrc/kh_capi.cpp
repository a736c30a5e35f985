// kh_capi.cpp -- the C ABI (include/kharma_amd.h) and the C++ stage driver.
//
// Host code only: it validates arguments, fills the index/constant structs,
// and enqueues the gfx950 kernels of kh_flux.hip / kh_step.hip on one stream.
// No exception crosses the ABI; every HIP status is checked.
#include "kharma_amd.h"

#include <algorithm>
#include <cfloat>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kh_host.hpp"

using namespace kh;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

int kh::api_fail(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define KH_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return fail(KH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace



namespace {

bool bc_ok(int t) { return t == KH_BC_OUTFLOW || t == KH_BC_REFLECT; }

}  // namespace

int kh::check_params(const kh_params* p) {
    if (!p) return fail(KH_ERR_INVALID, "null params");
    if (p->nx1 < 1 || p->nx2 < 1 || p->nx3 < 1) return fail(KH_ERR_INVALID, "zone counts must be >= 1");
    if (p->ng < 3) return fail(KH_ERR_INVALID, "need at least 3 ghost zones for a 5-point reconstruction");
    if (p->nx3 > 1 && p->nx2 == 1) return fail(KH_ERR_INVALID, "3D blocks need nx2 > 1");
    if (p->coords != KH_COORDS_CARTESIAN_MINKOWSKI && p->coords != KH_COORDS_KS)
        return fail(KH_ERR_UNSUPPORTED, "coordinates must be cartesian_minkowski or spherical Kerr-Schild");
    if (p->coords == KH_COORDS_KS) {
        if (p->transform < KH_TRANSFORM_NULL || p->transform > KH_TRANSFORM_FMKS) return fail(KH_ERR_INVALID, "unknown coordinate transform");
        if (!(p->a >= 0. && p->a < 1.)) return fail(KH_ERR_INVALID, "black hole spin must be in [0, 1)");
    }
    if (p->nx3 > 1 && !p->periodic3) return fail(KH_ERR_UNSUPPORTED, "x3 boundaries must be periodic");
    if (!p->periodic1 && (!bc_ok(p->bc_inner_x1) || !bc_ok(p->bc_outer_x1)))
        return fail(KH_ERR_INVALID, "non-periodic x1 needs outflow or reflecting faces");
    if (p->nx2 > 1 && !p->periodic2 && (!bc_ok(p->bc_inner_x2) || !bc_ok(p->bc_outer_x2)))
        return fail(KH_ERR_INVALID, "non-periodic x2 needs outflow or reflecting faces");
    if (p->nx2 == 1 && !p->periodic2) return fail(KH_ERR_UNSUPPORTED, "a 1D block is periodic in x2");
    auto short_axis = [&](int nx, int periodic) { return nx > 1 && nx < (periodic ? 2 * p->ng : p->ng); };
    if (short_axis(p->nx1, p->periodic1) || short_axis(p->nx2, p->periodic2) || short_axis(p->nx3, p->periodic3))
        return fail(KH_ERR_INVALID, "each axis needs nx >= 2 ng zones (periodic) or nx >= ng (physical faces)");
    if (p->nx1 == 1) return fail(KH_ERR_INVALID, "nx1 must exceed 1");
    {
        // the kernels address one variable of a block with 32-bit byte offsets
        auto ext = [&](int nx) { return nx > 1 ? (double)nx + 2. * p->ng + 1. : 2.; };
        if (ext(p->nx1) * ext(p->nx2) * ext(p->nx3) * 8. >= 4294967296.)
            return fail(KH_ERR_UNSUPPORTED, "a block's variable must stay below 4 GiB (split the mesh into blocks)");
    }
    if (p->recon < 0 || p->recon > 3) return fail(KH_ERR_INVALID, "unknown reconstruction");
    if (p->precision != KH_PREC_FP64 && p->precision != KH_PREC_SPLIT) return fail(KH_ERR_INVALID, "unknown precision");
    if (p->integrator != 0 && p->integrator != 1) return fail(KH_ERR_INVALID, "unknown integrator");
    if (p->driver_type != KH_DRIVER_KHARMA && p->driver_type != KH_DRIVER_SIMPLE)
        return fail(KH_ERR_INVALID, "unknown driver type");
    if (p->inverter != KH_INVERTER_ONEDW && p->inverter != KH_INVERTER_KASTAUN)
        return fail(KH_ERR_INVALID, "unknown inverter type");
    if (p->fofc) {
        if (p->precision != KH_PREC_FP64) return fail(KH_ERR_UNSUPPORTED, "fofc runs in the FP64 form only");
        if (p->driver_type != KH_DRIVER_KHARMA) return fail(KH_ERR_UNSUPPORTED, "fofc is a step of the default (kharma) driver");
        if (p->fofc_polar_cells < 0 || !(p->fofc_eh_buffer >= 0.)) return fail(KH_ERR_INVALID, "fofc polar_cells and eh_buffer must be >= 0");
    }
    if (p->b_solver != KH_BSOLVER_FLUX_CT && p->b_solver != KH_BSOLVER_FACE_CT)
        return fail(KH_ERR_INVALID, "unknown b_field solver");
    if (p->b_solver == KH_BSOLVER_FACE_CT) {
        if (p->ct_scheme < KH_CT_BS99 || p->ct_scheme > KH_CT_SG07) return fail(KH_ERR_INVALID, "unknown ct_scheme");
        // The EMF treatments B_CT applies at physical boundaries (b_ct_boundaries.cpp)
        // rest on Parthenon's face-field boundary fills, which are not part of this path.
        if (p->coords != KH_COORDS_CARTESIAN_MINKOWSKI || !p->periodic1 || (p->nx2 > 1 && !p->periodic2))
            return fail(KH_ERR_UNSUPPORTED, "face_ct is built for periodic Cartesian Minkowski blocks");
        if (p->nx2 == 1) return fail(KH_ERR_UNSUPPORTED, "face_ct needs a 2-D or 3-D block");
        if (p->driver_type != KH_DRIVER_KHARMA)
            return fail(KH_ERR_UNSUPPORTED, "face_ct runs with the default (kharma) driver (kharma_driver.hpp:94)");
        if (p->fofc) return fail(KH_ERR_UNSUPPORTED, "fofc with face_ct is not built");
    }
    if (p->kill_on_large_divb && !(p->kill_on_divb_over > 0.))
        return fail(KH_ERR_INVALID, "kill_on_divb_over must be positive when kill_on_large_divb is set");
    if (p->floor_frame < 0 || p->floor_frame > 2) return fail(KH_ERR_INVALID, "unknown floor frame");
    if (p->iter_max < 0) return fail(KH_ERR_INVALID, "iter_max < 0");
    const double n = (double)(p->nx1 + 2 * p->ng) * (p->nx2 + 2 * p->ng) * (p->nx3 + 2 * p->ng);
    if (n * 8 > 2.0e9) return fail(KH_ERR_INVALID, "block too large for 32-bit zone indexing");
    return KH_OK;
}

void kh::make_dims(const kh_params* p, BlockDims& b) {
    b.ndim = (p->nx3 > 1) ? 3 : ((p->nx2 > 1) ? 2 : 1);
    b.ng = p->ng;
    b.nx1 = p->nx1; b.nx2 = p->nx2; b.nx3 = p->nx3;
    b.n1 = p->nx1 + 2 * p->ng;
    b.n2 = (b.ndim > 1) ? p->nx2 + 2 * p->ng : 1;
    b.n3 = (b.ndim > 2) ? p->nx3 + 2 * p->ng : 1;
    b.is = p->ng; b.ie = p->ng + p->nx1 - 1;
    b.js = (b.ndim > 1) ? p->ng : 0; b.je = (b.ndim > 1) ? p->ng + p->nx2 - 1 : 0;
    b.ks = (b.ndim > 2) ? p->ng : 0; b.ke = (b.ndim > 2) ? p->ng + p->nx3 - 1 : 0;
    b.periodic[0] = p->periodic1; b.periodic[1] = (b.ndim > 1) ? p->periodic2 : 1; b.periodic[2] = 1;
    b.s2 = b.n1;
    b.s3 = (long)b.n1 * b.n2;
    b.nc = (long)b.n1 * b.n2 * b.n3;
    b.dx1 = (p->x1max - p->x1min) / p->nx1;
    b.dx2 = (p->x2max - p->x2min) / p->nx2;
    b.dx3 = (p->x3max - p->x3min) / p->nx3;
    b.geo = (p->coords == KH_COORDS_KS);
    b.bc[0] = b.periodic[0] ? 0 : p->bc_inner_x1;
    b.bc[1] = b.periodic[0] ? 0 : p->bc_outer_x1;
    b.bc[2] = b.periodic[1] ? 0 : p->bc_inner_x2;
    b.bc[3] = b.periodic[1] ? 0 : p->bc_outer_x2;
    b.check_inflow[0] = !b.periodic[0] && p->check_inflow_inner_x1;
    b.check_inflow[1] = !b.periodic[0] && p->check_inflow_outer_x1;
    b.zero_polar_flux = !b.periodic[1] && p->zero_polar_flux;
    b.fix_polar_flux = !b.periodic[1] && p->fix_polar_flux;
    b.fix_corner_inner = !b.periodic[0] && !b.periodic[1] && p->fix_corner_inner;
    b.prim_fill = p->driver_type == KH_DRIVER_SIMPLE;
    for (int d = 0; d < 3; ++d) {
        b.img[d][0] = b.img[d][1] = b.periodic[d];
        b.nbd[d][0] = b.nbd[d][1] = 0;
        b.goff[d] = 0;
    }
    b.face_ct = p->b_solver == KH_BSOLVER_FACE_CT;
    b.ct_scheme = p->ct_scheme;
    b.fs2 = b.n1 + 1;
    b.fs3 = (long)(b.n1 + 1) * (b.n2 + 1);
    b.nf = b.fs3 * (b.n3 + 1);
}

namespace {

int validate(const kh_params* p) { return check_params(p); }
void fill_dims(const kh_params* p, BlockDims& b) { make_dims(p, b); }

void fill_consts(const kh_params* p, StepConsts& c) {
    c.gam = p->gam; c.cfl = p->cfl; c.dt_min = p->dt_min; c.max_dt_increase = p->max_dt_increase; c.tlim = p->tlim;
    if (p->integrator == 1) {  // vl2
        c.gam0[0] = 0.; c.gam1[0] = 1.; c.beta[0] = 0.5;
        c.gam0[1] = 0.; c.gam1[1] = 1.; c.beta[1] = 1.0;
    } else {                   // rk2
        c.gam0[0] = 0.; c.gam1[0] = 1.; c.beta[0] = 1.0;
        c.gam0[1] = 0.5; c.gam1[1] = 0.5; c.beta[1] = 0.5;
    }
    c.recon_floors = p->recon_floors;
    c.iter_max = p->iter_max;
    c.err_tol = p->err_tol;
    c.fix_average = p->fix_average;
    c.fix_atmo = p->fix_atmo;
    c.use_hlle = p->use_hlle;
    c.kill_divb = p->kill_on_large_divb;
    c.inverter = p->inverter;
    c.divb_over = p->kill_on_divb_over;
    FloorParams& f = c.floors;
    f.rho_min_geom = p->rho_min_geom; f.u_min_geom = p->u_min_geom;
    f.rho_min_const = p->rho_min_const; f.u_min_const = p->u_min_const; f.r_char = p->r_char;
    f.bsq_over_rho_max = p->bsq_over_rho_max; f.bsq_over_u_max = p->bsq_over_u_max;
    f.u_over_rho_max = p->u_over_rho_max; f.ktot_max = p->ktot_max; f.gamma_max = p->gamma_max;
    f.use_r_char = p->use_r_char; f.temp_adjust_u = p->temp_adjust_u; f.frame = p->floor_frame;
    f.on = p->floors_on;
}


// Build the block's metric cache on the host and copy it to the device: FP64
// tables for the centre and the X1/X2 faces (the flux kernels of the split form
// read FP32 copies), sqrt(-g) Gamma at centres with a mask of its non-zero
// components, and the geometric floors.
hipError_t upload_geometry(kh_block* b) {
    GeomHost G;
    build_geometry(b->p, b->dims, G);
    GeomDev& d = b->geo;
    d.gn1 = G.gn1;
    d.np = G.np;
    auto up = [&](const void* src, size_t bytes, void** dst) -> hipError_t {
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) return e;
        b->geo_bufs.push_back(*dst);
        return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    const size_t ntab = (size_t)GEO_NCOMP * G.np;
    for (int loc = 0; loc < 3; ++loc) {
        void* q = nullptr;
        hipError_t e = up(G.g[loc].data(), sizeof(double) * ntab, &q);
        if (e != hipSuccess) return e;
        d.g[loc] = (const double*)q;
        std::vector<float> f(G.g[loc].begin(), G.g[loc].end());
        e = up(f.data(), sizeof(float) * ntab, &q);
        if (e != hipSuccess) return e;
        d.g32[loc] = (const float*)q;
    }
    unsigned long long mask = 0;
    for (int n = 0; n < 64; ++n)
        for (long q = 0; q < G.np; ++q)
            if (G.conn[(size_t)n * G.np + q] != 0.) { mask |= 1ull << n; break; }
    d.conn_mask = mask;
    void* q = nullptr;
    hipError_t e = up(G.conn.data(), sizeof(double) * G.conn.size(), &q);
    if (e != hipSuccess) return e;
    d.conn = (const double*)q;
    e = up(G.flr.data(), sizeof(double) * G.flr.size(), &q);
    if (e != hipSuccess) return e;
    d.flr = (const double*)q;
    std::vector<float> f32(G.flr.begin(), G.flr.end());
    e = up(f32.data(), sizeof(float) * f32.size(), &q);
    if (e != hipSuccess) return e;
    d.flr32 = (const float*)q;
    return hipSuccess;
}

// MarkFOFC's zones marked whatever the trial state (flux/fofc.cpp:61-107):
// cell centres within eh_buffer of the horizon, and the first polar_cells rows
// at physical X2 faces, in spherical coordinates.  Laid out as the geometry
// tables, [gn2][gn1] with gn1 = n1 + 1; absent in Cartesian coordinates.
hipError_t upload_fofc_fixed(kh_block* blk) {
    const kh_params& pp = blk->p;
    if (!pp.fofc || pp.coords != KH_COORDS_KS) return hipSuccess;
    const BlockDims& bd = blk->dims;
    const long np = blk->geo.np, gn1 = blk->geo.gn1;
    std::vector<int> fixed(np, 0);
    const double r_eh = 1. + std::sqrt(1. - pp.a * pp.a);
    for (int j = 0; j < bd.n2; ++j)
        for (int i = 0; i < bd.n1; ++i) {
            double X[4];
            locus_coords(pp, bd, GLOC_CENTER, j, i, X);
            bool m = embed_r(pp, X) < r_eh + pp.fofc_eh_buffer;
            // the polar rows of a block's physical X2 faces only (fofc.cpp:82-85)
            const int pc = pp.fofc_polar_cells;
            if (bd.ndim > 1 && pc > 0 &&
                ((!bd.img[1][0] && j >= bd.js && j < bd.js + pc) || (!bd.img[1][1] && j <= bd.je && j > bd.je - pc)))
                m = true;
            fixed[(size_t)j * gn1 + i] = m;
        }
    hipError_t e = hipMalloc(&blk->fofc_fixed, sizeof(int) * np);
    if (e != hipSuccess) return e;
    return hipMemcpy(blk->fofc_fixed, fixed.data(), sizeof(int) * np, hipMemcpyHostToDevice);
}

// kh_vars -> device pointer tables; every one of the 8 pointers must be set.
bool vars_ok(const kh_vars* v) {
    if (!v) return false;
    for (int n = 0; n < NV; ++n)
        if (!v->var[n]) return false;
    return true;
}
bool flux_vars_ok(const kh_block* b, const kh_vars* F) {
    if (!F) return false;
    for (int d = 0; d < b->dims.ndim; ++d)
        if (!vars_ok(&F[d])) return false;
    return true;
}
template <typename T> Vars<T> vars_in(const kh_vars* v) {
    Vars<T> r;
    for (int n = 0; n < NV; ++n) r.v[n] = (T*)v->var[n];
    return r;
}
// Directions beyond the block's dimensionality are never read or written.
FluxVars<void> flux_vars_in(const kh_block* b, const kh_vars* F) {
    FluxVars<void> r;
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < NV; ++n) r.d[d].v[n] = (d < b->dims.ndim) ? F[d].var[n] : nullptr;
    return r;
}
kh_vars contiguous_vars(const void* base, long nc, size_t elem) {
    kh_vars r;
    for (int n = 0; n < NV; ++n) r.var[n] = base ? (void*)((const char*)base + (size_t)n * nc * elem) : nullptr;
    return r;
}
void contiguous_fluxes(const void* base, long nc, size_t elem, kh_vars F[3]) {
    for (int d = 0; d < 3; ++d) F[d] = contiguous_vars((const char*)base + (size_t)d * NV * nc * elem, nc, elem);
}

}  // namespace

extern "C" {

const char* kh_last_error(void) { return g_err.c_str(); }
int kh_version(void) { return 1; }

int kh_block_create(const kh_params* p, kh_block** out) {
    if (!out) return fail(KH_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int st = validate(p);
    if (st != KH_OK) return st;
    kh_block* b = new kh_block();
    b->p = *p;
    fill_dims(p, b->dims);
    fill_consts(p, b->consts);
    b->npartials = std::max(update_blocks(b->dims), divb_blocks(b->dims)) + 1;
    hipError_t e = hipMalloc(&b->scratch, sizeof(DeviceState));
    if (e == hipSuccess) e = hipMalloc(&b->partials, sizeof(double) * b->npartials);
    if (e == hipSuccess && b->dims.geo) e = upload_geometry(b);
    if (e == hipSuccess) e = upload_fofc_fixed(b);
    if (e == hipSuccess && b->dims.face_ct) e = hipMalloc(&b->ct_cemf, sizeof(double) * 3 * b->dims.nc);
    if (e != hipSuccess) {
        kh_block_destroy(b);
        return fail(KH_ERR_HIP, std::string("device allocation / copy: ") + hipGetErrorString(e));
    }
    *out = b;
    return KH_OK;
}

}  // extern "C"

// One block of a uniform mesh of nb[0] x nb[1] x nb[2] blocks (block index
// bidx): the block's own extents, the mesh's zone widths and coordinates
// (offset by its place in the mesh, so its geometry is the one-block mesh's at
// the same zones), and per face either the mesh's physical boundary or images
// of the neighbouring block (periodic axes wrap around the mesh).
int kh::block_create_in_mesh(const kh_params* mesh, const int nb[3], const int bidx[3], kh_block** out) {
    if (!out) return fail(KH_ERR_INVALID, "null output pointer");
    *out = nullptr;
    kh_params pb = *mesh;
    pb.nx1 = mesh->nx1 / nb[0];
    pb.nx2 = mesh->nx2 / nb[1];
    pb.nx3 = mesh->nx3 / nb[2];
    int st = validate(&pb);
    if (st != KH_OK) return st;
    kh_block* b = new kh_block();
    b->p = pb;
    BlockDims& bd = b->dims;
    fill_dims(&pb, bd);
    bd.dx1 = (mesh->x1max - mesh->x1min) / mesh->nx1;
    bd.dx2 = (mesh->x2max - mesh->x2min) / mesh->nx2;
    bd.dx3 = (mesh->x3max - mesh->x3min) / mesh->nx3;
    const int stride[3] = {1, nb[0], nb[0] * nb[1]};
    const int nx[3] = {bd.nx1, bd.nx2, bd.nx3};
    const int per[3] = {mesh->periodic1, bd.ndim > 1 ? mesh->periodic2 : 1, 1};
    for (int d = 0; d < 3; ++d) {
        bd.goff[d] = (d < bd.ndim) ? bidx[d] * nx[d] : 0;
        if (d >= bd.ndim || nb[d] == 1) continue;   // as a one-block mesh along this axis
        const bool first = bidx[d] == 0, last = bidx[d] == nb[d] - 1;
        bd.img[d][0] = !first || per[d];
        bd.nbd[d][0] = first ? (nb[d] - 1) * stride[d] : -stride[d];
        bd.img[d][1] = !last || per[d];
        bd.nbd[d][1] = last ? -(nb[d] - 1) * stride[d] : stride[d];
        bd.periodic[d] = bd.img[d][0] && bd.img[d][1];
    }
    bd.bc[0] = bd.img[0][0] ? 0 : mesh->bc_inner_x1;
    bd.bc[1] = bd.img[0][1] ? 0 : mesh->bc_outer_x1;
    bd.bc[2] = bd.img[1][0] ? 0 : mesh->bc_inner_x2;
    bd.bc[3] = bd.img[1][1] ? 0 : mesh->bc_outer_x2;
    bd.check_inflow[0] = !bd.img[0][0] && mesh->check_inflow_inner_x1;
    bd.check_inflow[1] = !bd.img[0][1] && mesh->check_inflow_outer_x1;
    const bool polar = bd.ndim > 1 && (!bd.img[1][0] || !bd.img[1][1]);
    bd.zero_polar_flux = polar && mesh->zero_polar_flux;
    bd.fix_polar_flux = polar && mesh->fix_polar_flux;
    bd.fix_corner_inner = !bd.img[0][0] && polar && mesh->fix_corner_inner;
    fill_consts(&pb, b->consts);
    b->npartials = std::max(update_blocks(bd), divb_blocks(bd)) + 1;
    hipError_t e = hipMalloc(&b->scratch, sizeof(DeviceState));
    if (e == hipSuccess) e = hipMalloc(&b->partials, sizeof(double) * b->npartials);
    if (e == hipSuccess && bd.geo) e = upload_geometry(b);
    if (e == hipSuccess) e = upload_fofc_fixed(b);
    if (e != hipSuccess) {
        kh_block_destroy(b);
        return fail(KH_ERR_HIP, std::string("device allocation / copy: ") + hipGetErrorString(e));
    }
    *out = b;
    return KH_OK;
}

extern "C" {

void kh_block_destroy(kh_block* b) {
    if (!b) return;
    if (b->scratch) (void)hipFree(b->scratch);
    if (b->partials) (void)hipFree(b->partials);
    for (void* q : b->geo_bufs) (void)hipFree(q);
    if (b->fofc_fixed) (void)hipFree(b->fofc_fixed);
    if (b->ct_cemf) (void)hipFree(b->ct_cemf);
    delete b;
}

int kh_block_set_stream(kh_block* b, void* s) {
    if (!b) return fail(KH_ERR_INVALID, "null block");
    b->stream = (hipStream_t)s;
    return KH_OK;
}

int kh_block_shape(const kh_block* b, int* n1, int* n2, int* n3, long* nc) {
    if (!b) return fail(KH_ERR_INVALID, "null block");
    if (n1) *n1 = b->dims.n1;
    if (n2) *n2 = b->dims.n2;
    if (n3) *n3 = b->dims.n3;
    if (nc) *nc = b->dims.nc;
    return KH_OK;
}

int kh_vars_from_pack(void* const* pack_var, int nvar, const kh_varmap* m, kh_vars* out) {
    if (!pack_var || !m || !out) return fail(KH_ERR_INVALID, "null argument");
    const int first[NV] = {m->RHO, m->UU, m->U1, m->U1, m->U1, m->B1, m->B1, m->B1};
    const int off[NV] = {0, 0, 0, 1, 2, 0, 1, 2};
    for (int n = 0; n < NV; ++n) {
        if (first[n] < 0) return fail(KH_ERR_INVALID, "VarMap index absent (-1)");
        const int v = first[n] + off[n];
        if (v >= nvar) return fail(KH_ERR_INVALID, "VarMap index beyond the pack");
        if (!pack_var[v]) return fail(KH_ERR_INVALID, "null pack variable");
        out->var[n] = pack_var[v];
    }
    return KH_OK;
}

int kh_get_flux_v(kh_block* b, int dir, const kh_vars* P, const kh_vars* F_dir, void* cmax_dir, void* cmin_dir) {
    if (!b || !vars_ok(P) || !vars_ok(F_dir) || !cmax_dir || !cmin_dir) return fail(KH_ERR_INVALID, "null argument");
    if (dir < 1 || dir > 3) return fail(KH_ERR_INVALID, "dir must be 1, 2 or 3");
    KH_HIP(launch_flux(b->dims, b->consts, b->geo, dir, b->p.recon, b->p.precision, vars_in<const double>(P),
                       vars_in<void>(F_dir), cmax_dir, cmin_dir, b->stream));
    return KH_OK;
}

int kh_get_flux(kh_block* b, int dir, const double* P, void* F_dir, void* cmax_dir, void* cmin_dir) {
    if (!b || !P || !F_dir) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Pv = contiguous_vars(P, b->dims.nc, sizeof(double));
    const kh_vars Fv = contiguous_vars(F_dir, b->dims.nc, flux_elem(b));
    return kh_get_flux_v(b, dir, &Pv, &Fv, cmax_dir, cmin_dir);
}

int kh_flux_ct_v(kh_block* b, const kh_vars F[3], void* emf) {
    if (!b || !flux_vars_ok(b, F) || !emf) return fail(KH_ERR_INVALID, "null argument");
    const FluxVars<void> Fv = flux_vars_in(b, F);
    KH_HIP(launch_fix_flux(b->dims, b->p.precision, Fv, b->stream));
    KH_HIP(launch_emf(b->dims, b->p.precision, Fv, emf, b->stream));
    KH_HIP(launch_flux_ct_rewrite(b->dims, b->p.precision, Fv, emf, b->stream));
    return KH_OK;
}

int kh_flux_ct(kh_block* b, void* F, void* emf) {
    if (!b || !F) return fail(KH_ERR_INVALID, "null argument");
    kh_vars Fv[3];
    contiguous_fluxes(F, b->dims.nc, flux_elem(b), Fv);
    return kh_flux_ct_v(b, Fv, emf);
}

int kh_flux_divergence_v(kh_block* b, const kh_vars F[3], const kh_vars* dUdt) {
    if (!b || !flux_vars_ok(b, F) || !vars_ok(dUdt)) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_flux_divergence(b->dims, b->p.precision, flux_vars_in(b, F), vars_in<double>(dUdt), b->stream));
    return KH_OK;
}

int kh_flux_divergence(kh_block* b, const void* F, double* dUdt) {
    if (!b || !F || !dUdt) return fail(KH_ERR_INVALID, "null argument");
    kh_vars Fv[3];
    contiguous_fluxes(F, b->dims.nc, flux_elem(b), Fv);
    const kh_vars Dv = contiguous_vars(dUdt, b->dims.nc, sizeof(double));
    return kh_flux_divergence_v(b, Fv, &Dv);
}

int kh_add_geo_source_v(kh_block* b, const kh_vars* P, const kh_vars* dUdt) {
    if (!b || !vars_ok(P) || !vars_ok(dUdt)) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_geo_source(b->dims, b->consts, b->geo, vars_in<const double>(P), vars_in<double>(dUdt), b->stream));
    return KH_OK;
}

int kh_add_geo_source(kh_block* b, const double* P, double* dUdt) {
    if (!b || !P || !dUdt) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Pv = contiguous_vars(P, b->dims.nc, sizeof(double)), Dv = contiguous_vars(dUdt, b->dims.nc, sizeof(double));
    return kh_add_geo_source_v(b, &Pv, &Dv);
}

int kh_weighted_sum_v(kh_block* b, const kh_vars* Usub, const kh_vars* Ufull, const kh_vars* dUdt, double gam0,
                      double gam1, double beta_dt, const kh_vars* Uout) {
    if (!b || !vars_ok(Usub) || !vars_ok(Ufull) || !vars_ok(dUdt) || !vars_ok(Uout))
        return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_weighted_sum(b->dims, vars_in<const double>(Usub), vars_in<const double>(Ufull),
                               vars_in<const double>(dUdt), gam0, gam1, beta_dt, vars_in<double>(Uout), b->stream));
    return KH_OK;
}

int kh_weighted_sum(kh_block* b, const double* Usub, const double* Ufull, const double* dUdt,
                    double gam0, double gam1, double beta_dt, double* Uout) {
    if (!b || !Usub || !Ufull || !dUdt || !Uout) return fail(KH_ERR_INVALID, "null argument");
    const long nc = b->dims.nc;
    const kh_vars a = contiguous_vars(Usub, nc, 8), f = contiguous_vars(Ufull, nc, 8), d = contiguous_vars(dUdt, nc, 8),
                  o = contiguous_vars(Uout, nc, 8);
    return kh_weighted_sum_v(b, &a, &f, &d, gam0, gam1, beta_dt, &o);
}

int kh_utop_v(kh_block* b, const kh_vars* U, const kh_vars* P, int* pflag, int* fflag) {
    if (!b || !vars_ok(U) || !vars_ok(P) || !pflag || !fflag) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_utop(b->dims, b->consts, b->geo, vars_in<const double>(U), vars_in<double>(P), pflag, fflag, b->stream));
    return KH_OK;
}

int kh_utop(kh_block* b, const double* U, double* P, int* pflag, int* fflag) {
    if (!b || !U || !P) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Uv = contiguous_vars(U, b->dims.nc, 8), Pv = contiguous_vars(P, b->dims.nc, 8);
    return kh_utop_v(b, &Uv, &Pv, pflag, fflag);
}

int kh_apply_floors_v(kh_block* b, const kh_vars* U, const kh_vars* P, int* pflag, int* fflag) {
    if (!b || !vars_ok(U) || !vars_ok(P) || !pflag || !fflag) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_floors(b->dims, b->consts, b->geo, vars_in<double>(U), vars_in<double>(P), pflag, fflag, b->stream));
    return KH_OK;
}

int kh_apply_floors(kh_block* b, double* U, double* P, int* pflag, int* fflag) {
    if (!b || !U || !P) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Uv = contiguous_vars(U, b->dims.nc, 8), Pv = contiguous_vars(P, b->dims.nc, 8);
    return kh_apply_floors_v(b, &Uv, &Pv, pflag, fflag);
}

int kh_fix_utop_v(kh_block* b, const kh_vars* U, const kh_vars* P, const int* pflag) {
    if (!b || !vars_ok(U) || !vars_ok(P) || !pflag) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_fixup_entire(b->dims, b->consts, b->geo, vars_in<double>(U), vars_in<double>(P), pflag, b->stream));
    return KH_OK;
}

int kh_fix_utop(kh_block* b, double* U, double* P, const int* pflag) {
    if (!b || !U || !P) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Uv = contiguous_vars(U, b->dims.nc, 8), Pv = contiguous_vars(P, b->dims.nc, 8);
    return kh_fix_utop_v(b, &Uv, &Pv, pflag);
}

int kh_ptou_v(kh_block* b, const kh_vars* P, const kh_vars* U) {
    if (!b || !vars_ok(P) || !vars_ok(U)) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_ptou(b->dims, b->consts, b->geo, vars_in<const double>(P), vars_in<double>(U), b->stream));
    return KH_OK;
}

int kh_ptou(kh_block* b, const double* P, double* U) {
    if (!b || !P || !U) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Pv = contiguous_vars(P, b->dims.nc, 8), Uv = contiguous_vars(U, b->dims.nc, 8);
    return kh_ptou_v(b, &Pv, &Uv);
}

int kh_apply_bounds_v(kh_block* b, const kh_vars* U, const kh_vars* P) {
    if (!b || !vars_ok(U) || !vars_ok(P)) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_bounds(b->dims, b->consts, b->geo, vars_in<double>(U), vars_in<double>(P), nullptr, b->stream));
    return KH_OK;
}

int kh_apply_bounds(kh_block* b, double* U, double* P) {
    if (!b || !U || !P) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Uv = contiguous_vars(U, b->dims.nc, 8), Pv = contiguous_vars(P, b->dims.nc, 8);
    return kh_apply_bounds_v(b, &Uv, &Pv);
}

int kh_sync_bounds_v(kh_block* b, const kh_vars* A, int nvar) {
    if (!b || !A || nvar < 1 || nvar > NV) return fail(KH_ERR_INVALID, "bad argument");
    for (int n = 0; n < nvar; ++n)
        if (!A->var[n]) return fail(KH_ERR_INVALID, "null variable");
    kh_vars a = *A;
    for (int n = nvar; n < NV; ++n) a.var[n] = nullptr;
    KH_HIP(launch_sync(b->dims, vars_in<double>(&a), nvar, b->stream));
    return KH_OK;
}

int kh_sync_bounds(kh_block* b, double* A, int nvar) {
    if (!b || !A || nvar < 1) return fail(KH_ERR_INVALID, "bad argument");
    KH_HIP(launch_sync(b->dims, A, nvar, b->stream));
    return KH_OK;
}

int kh_estimate_timestep(kh_block* b, const void* cmax, const void* cmin, double dt_last, double* ndt) {
    if (!b || !cmax || !cmin || !ndt) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_estimate_dt(b->dims, b->consts, b->p.precision, cmax, cmin, b->partials, b->stream));
    KH_HIP(launch_reduce(b->partials, (int)update_blocks(b->dims), 0, &b->scratch->last_ndt, b->stream));
    double m = 0.;
    KH_HIP(hipMemcpyAsync(&m, &b->scratch->last_ndt, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    KH_HIP(hipStreamSynchronize(b->stream));
    const double hi = b->p.max_dt_increase * dt_last;
    double v = m * b->p.cfl;
    v = (v < b->p.dt_min) ? b->p.dt_min : v;   // min(max(lower, n), upper)
    v = (hi < v) ? hi : v;
    *ndt = v;
    return KH_OK;
}

int kh_estimate_timestep_noalloc(kh_block* b, const void* cmax, const void* cmin, double dt_last, double* ndt) {
    if (!b || !cmax || !cmin || !ndt) return fail(KH_ERR_INVALID, "null argument");
    if (update_blocks(b->dims) > KH_NOALLOC_PARTIALS)
        return fail(KH_ERR_INVALID, "block interior exceeds the static scratch of kh_estimate_timestep_noalloc (" +
                                        std::to_string(KH_NOALLOC_PARTIALS * 256) + " zones)");
    KH_HIP(launch_estimate_dt_noalloc(b->dims, b->consts, b->p.precision, cmax, cmin, dt_last, ndt, b->stream));
    return KH_OK;
}

int kh_max_divb_v(kh_block* b, const kh_vars* U, double* out) {
    if (!b || !vars_ok(U) || !out) return fail(KH_ERR_INVALID, "null argument");
    if (b->dims.ndim < 2) { *out = 0.; return KH_OK; }
    KH_HIP(launch_max_divb(b->dims, vars_in<const double>(U), b->partials, b->stream));
    KH_HIP(launch_reduce(b->partials, (int)divb_blocks(b->dims), 1, &b->scratch->last_divb, b->stream));
    KH_HIP(hipMemcpyAsync(out, &b->scratch->last_divb, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    KH_HIP(hipStreamSynchronize(b->stream));
    return KH_OK;
}

int kh_max_divb(kh_block* b, const double* U, double* out) {
    if (!b || !U) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Uv = contiguous_vars(U, b->dims.nc, 8);
    return kh_max_divb_v(b, &Uv, out);
}

// ---------------------------------------------------------------- B_CT tasks
namespace {
int need_face_ct(const kh_block* b) {
    if (!b) return fail(KH_ERR_INVALID, "null block");
    if (!b->dims.face_ct) return fail(KH_ERR_INVALID, "the block's b_solver is not face_ct");
    return KH_OK;
}
}  // namespace

int kh_block_face_count(const kh_block* b, long* nf) {
    if (!b || !nf) return fail(KH_ERR_INVALID, "null argument");
    *nf = b->dims.nf;
    return KH_OK;
}

int kh_ct_calculate_emf_v(kh_block* b, const kh_vars* P, const kh_vars F[3], double* emf) {
    int st = need_face_ct(b);
    if (st != KH_OK) return st;
    if (!vars_ok(P) || !flux_vars_ok(b, F) || !emf) return fail(KH_ERR_INVALID, "null argument");
    if (b->dims.ct_scheme != KH_CT_BS99)
        KH_HIP(launch_ct_cemf(b->dims, vars_in<const double>(P), b->ct_cemf, b->stream));
    KH_HIP(launch_ct_emf(b->dims, b->p.precision, flux_vars_cast<const void>(flux_vars_in(b, F)), b->ct_cemf, emf,
                         b->stream));
    return KH_OK;
}

int kh_ct_calculate_emf(kh_block* b, const double* P, const void* F, double* emf) {
    if (!b || !P || !F) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Pv = contiguous_vars(P, b->dims.nc, 8);
    kh_vars Fv[3];
    contiguous_fluxes(F, b->dims.nc, flux_elem(b), Fv);
    return kh_ct_calculate_emf_v(b, &Pv, Fv, emf);
}

int kh_ct_add_source(kh_block* b, const double* emf, double* dfB) {
    int st = need_face_ct(b);
    if (st != KH_OK) return st;
    if (!emf || !dfB) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_ct_face_update(b->dims, emf, nullptr, nullptr, 0., 0., false, nullptr, 0., dfB, b->stream));
    return KH_OK;
}

int kh_ct_weighted_sum_face(kh_block* b, const double* x, const double* y, double w1, double w2, double* z) {
    int st = need_face_ct(b);
    if (st != KH_OK) return st;
    if (!x || !y || !z) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_ct_weighted_sum(b->dims, x, y, w1, w2, z, b->stream));
    return KH_OK;
}

int kh_ct_sync_faces(kh_block* b, double* fB) {
    int st = need_face_ct(b);
    if (st != KH_OK) return st;
    if (!fB) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_ct_sync(b->dims, fB, b->stream));
    return KH_OK;
}

int kh_ct_utop_v(kh_block* b, const double* fB, const kh_vars* U, const kh_vars* P) {
    int st = need_face_ct(b);
    if (st != KH_OK) return st;
    if (!fB || !vars_ok(U) || !vars_ok(P)) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_ct_utop(b->dims, fB, vars_in<double>(U), vars_in<double>(P), b->stream));
    return KH_OK;
}

int kh_ct_utop(kh_block* b, const double* fB, double* U, double* P) {
    if (!b || !U || !P) return fail(KH_ERR_INVALID, "null argument");
    const kh_vars Uv = contiguous_vars(U, b->dims.nc, 8), Pv = contiguous_vars(P, b->dims.nc, 8);
    return kh_ct_utop_v(b, fB, &Uv, &Pv);
}

int kh_ct_max_divb(kh_block* b, const double* fB, double* out) {
    int st = need_face_ct(b);
    if (st != KH_OK) return st;
    if (!fB || !out) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_ct_max_divb(b->dims, fB, b->partials, b->stream));
    KH_HIP(launch_reduce(b->partials, (int)divb_blocks(b->dims), 1, &b->scratch->last_divb, b->stream));
    KH_HIP(hipMemcpyAsync(out, &b->scratch->last_divb, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    KH_HIP(hipStreamSynchronize(b->stream));
    return KH_OK;
}

int kh_flags_to_real(kh_block* b, const int* flag, double* real_flag, long n) {
    if (!b || !flag || !real_flag || n < 0) return fail(KH_ERR_INVALID, "bad argument");
    if (n) KH_HIP(launch_flags_to_real(flag, real_flag, n, b->stream));
    return KH_OK;
}

int kh_flags_from_real(kh_block* b, const double* real_flag, int* flag, long n) {
    if (!b || !flag || !real_flag || n < 0) return fail(KH_ERR_INVALID, "bad argument");
    if (n) KH_HIP(launch_flags_from_real(real_flag, flag, n, b->stream));
    return KH_OK;
}

int kh_determine_floors_v(kh_block* b, const kh_vars* P, int* fflag) {
    if (!b || !vars_ok(P) || !fflag) return fail(KH_ERR_INVALID, "null argument");
    KH_HIP(launch_determine_floors(b->dims, b->consts, b->geo, vars_in<const double>(P), fflag, b->stream));
    return KH_OK;
}

int kh_mark_fofc(kh_block* b, const int* fflag, int* fofcflag) {
    if (!b || !fflag || !fofcflag) return fail(KH_ERR_INVALID, "null argument");
    if (!b->p.fofc) return fail(KH_ERR_INVALID, "the block was created without fofc");
    KH_HIP(launch_mark_fofc(b->dims, b->geo, fflag, b->fofc_fixed, fofcflag, b->stream));
    return KH_OK;
}

int kh_fofc_v(kh_block* b, const kh_vars* P, const int* fofcflag, const kh_vars F[3], double* cmax, double* cmin) {
    if (!b || !vars_ok(P) || !fofcflag || !flux_vars_ok(b, F) || !cmax || !cmin)
        return fail(KH_ERR_INVALID, "null argument");
    if (!b->p.fofc) return fail(KH_ERR_INVALID, "the block was created without fofc");
    KH_HIP(launch_fofc_flux(b->dims, b->consts, b->geo, SPEEDS_BOTH, vars_in<const double>(P), fofcflag,
                            flux_vars_cast<double>(flux_vars_in(b, F)), cmax, cmin, b->stream));
    return KH_OK;
}

}  // extern "C"

// ----------------------------------------------------------------- restart
// The reference's restart dump holds the fields marked Metadata::Restart --
// the conserved and primitive variables (driver/kharma_driver.cpp:114-118,
// grmhd/grmhd.cpp:155-156) -- and Parthenon's SimTime.  Parthenon writes them to
// HDF5 (.rhdf); this library writes the same content to a flat binary file.
int kh::write_restart(const char* path, const kh_params& p, const int n[3], double time, double dt, int ncycle,
                      const double* P, const double* U, hipStream_t s, const double* fB) {
    if (!path) return fail(KH_ERR_INVALID, "null path");
    const size_t cells = (size_t)NV * n[0] * n[1] * n[2];
    const size_t faces = fB ? (size_t)3 * (n[0] + 1) * (n[1] + 1) * (n[2] + 1) : 0;
    std::vector<double> host(2 * cells + faces);
    KH_HIP(hipMemcpyAsync(host.data(), P, sizeof(double) * cells, hipMemcpyDeviceToHost, s));
    KH_HIP(hipMemcpyAsync(host.data() + cells, U, sizeof(double) * cells, hipMemcpyDeviceToHost, s));
    if (fB) KH_HIP(hipMemcpyAsync(host.data() + 2 * cells, fB, sizeof(double) * faces, hipMemcpyDeviceToHost, s));
    KH_HIP(hipStreamSynchronize(s));
    RestartHeader hd{};
    std::memcpy(hd.magic, "KHAMDRST", 8);
    hd.version = 1;
    hd.header_bytes = sizeof(RestartHeader);
    for (int d = 0; d < 3; ++d) hd.n[d] = n[d];
    hd.ncycle = ncycle;
    hd.time = time;
    hd.dt = dt;
    hd.params = p;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(KH_ERR_INVALID, std::string("cannot open ") + path + " for writing");
    const bool ok = std::fwrite(&hd, sizeof(hd), 1, f) == 1 && std::fwrite(host.data(), sizeof(double), host.size(), f) == host.size();
    if (std::fclose(f) != 0 || !ok) return fail(KH_ERR_INVALID, std::string("short write to ") + path);
    return KH_OK;
}

int kh::read_restart(const char* path, const kh_params& p, const int n[3], double* time, double* dt, int* ncycle,
                     double* P, double* U, hipStream_t s, double* fB) {
    if (!path) return fail(KH_ERR_INVALID, "null path");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(KH_ERR_INVALID, std::string("cannot open ") + path);
    RestartHeader hd{};
    const bool hok = std::fread(&hd, sizeof(hd), 1, f) == 1;
    auto bad = [&](const std::string& why) { std::fclose(f); return fail(KH_ERR_INVALID, std::string(path) + ": " + why); };
    if (!hok || std::memcmp(hd.magic, "KHAMDRST", 8) != 0 || hd.version != 1 || hd.header_bytes != sizeof(RestartHeader))
        return bad("not a restart file of this library");
    const kh_params& q = hd.params;
    if (hd.n[0] != n[0] || hd.n[1] != n[1] || hd.n[2] != n[2] || q.nx1 != p.nx1 || q.nx2 != p.nx2 || q.nx3 != p.nx3 ||
        q.ng != p.ng)
        return bad("the mesh's zones differ from the run's");
    if (q.coords != p.coords || q.transform != p.transform || q.a != p.a || q.gam != p.gam || q.x1min != p.x1min ||
        q.x1max != p.x1max || q.x2min != p.x2min || q.x2max != p.x2max || q.x3min != p.x3min || q.x3max != p.x3max)
        return bad("the geometry or gamma differ from the run's");
    if (q.b_solver != p.b_solver || (fB == nullptr) != (q.b_solver != KH_BSOLVER_FACE_CT))
        return bad("the B field solver (flux_ct / face_ct) differs from the run's");
    const size_t cells = (size_t)NV * n[0] * n[1] * n[2];
    const size_t faces = fB ? (size_t)3 * (n[0] + 1) * (n[1] + 1) * (n[2] + 1) : 0;
    std::vector<double> host(2 * cells + faces);
    if (std::fread(host.data(), sizeof(double), host.size(), f) != host.size()) return bad("truncated");
    std::fclose(f);
    KH_HIP(hipMemcpyAsync(P, host.data(), sizeof(double) * cells, hipMemcpyHostToDevice, s));
    KH_HIP(hipMemcpyAsync(U, host.data() + cells, sizeof(double) * cells, hipMemcpyHostToDevice, s));
    if (fB) KH_HIP(hipMemcpyAsync(fB, host.data() + 2 * cells, sizeof(double) * faces, hipMemcpyHostToDevice, s));
    KH_HIP(hipStreamSynchronize(s));
    *time = hd.time;
    *dt = hd.dt;
    *ncycle = hd.ncycle;
    return KH_OK;
}

extern "C" {

// ------------------------------------------------------------------ driver
int kh_driver_create(const kh_params* p, kh_driver** out) {
    if (!out) return fail(KH_ERR_INVALID, "null output pointer");
    *out = nullptr;
    kh_block* blk = nullptr;
    int st = kh_block_create(p, &blk);
    if (st != KH_OK) return st;
    kh_driver* d = new kh_driver();
    d->blk = blk;
    const long nc = blk->dims.nc;
    const size_t st_bytes = sizeof(double) * NV * nc;
    const size_t fe = flux_elem(blk);
    d->hist_cap = 1 << 16;
    auto cleanup = [&](hipError_t e, const char* what) {
        kh_driver_destroy(d);
        return fail(KH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc(&d->U0, st_bytes)) != hipSuccess) return cleanup(e, "hipMalloc U0");
    if ((e = hipMalloc(&d->P0, st_bytes)) != hipSuccess) return cleanup(e, "hipMalloc P0");
    if ((e = hipMalloc(&d->U1, st_bytes)) != hipSuccess) return cleanup(e, "hipMalloc U1");
    if ((e = hipMalloc(&d->P1, st_bytes)) != hipSuccess) return cleanup(e, "hipMalloc P1");
    if ((e = hipMalloc(&d->F, fe * 3 * NV * nc)) != hipSuccess) return cleanup(e, "hipMalloc F");
    if ((e = hipMalloc(&d->ctop, fe * 3 * nc)) != hipSuccess) return cleanup(e, "hipMalloc ctop");
    if (blk->p.precision == KH_PREC_SPLIT) {
        if ((e = hipMalloc(&d->P0f, sizeof(float) * NV * nc)) != hipSuccess) return cleanup(e, "hipMalloc P0f");
        if ((e = hipMalloc(&d->P1f, sizeof(float) * NV * nc)) != hipSuccess) return cleanup(e, "hipMalloc P1f");
    }
    if ((e = hipMalloc(&d->dt_part, sizeof(double) * (update_blocks(blk->dims) + 1))) != hipSuccess) return cleanup(e, "hipMalloc dt_part");
    if ((e = hipMalloc(&d->divb_part, sizeof(double) * (divb_blocks(blk->dims) + 1))) != hipSuccess) return cleanup(e, "hipMalloc divb_part");
    if ((e = hipMalloc(&d->pflag, sizeof(int) * nc)) != hipSuccess) return cleanup(e, "hipMalloc pflag");
    if ((e = hipMalloc(&d->fflag, sizeof(int) * nc)) != hipSuccess) return cleanup(e, "hipMalloc fflag");
    if ((e = hipMalloc(&d->fail_list, sizeof(int) * nc)) != hipSuccess) return cleanup(e, "hipMalloc fail_list");
    if ((e = hipMalloc(&d->ds, sizeof(DeviceState))) != hipSuccess) return cleanup(e, "hipMalloc state");
    if ((e = hipMalloc(&d->dt_hist, sizeof(double) * d->hist_cap)) != hipSuccess) return cleanup(e, "hipMalloc dt_hist");
    if (blk->p.fofc) {
        if ((e = hipMalloc(&d->fofc_mark, sizeof(int) * nc)) != hipSuccess) return cleanup(e, "hipMalloc fofc_mark");
        if ((e = hipMemset(d->fofc_mark, 0, sizeof(int) * nc)) != hipSuccess) return cleanup(e, "hipMemset");
    }
    if (blk->dims.face_ct) {
        const size_t fb = sizeof(double) * 3 * blk->dims.nf;
        for (double** q : {&d->fB0, &d->fB1, &d->emf}) {
            if ((e = hipMalloc(q, fb)) != hipSuccess) return cleanup(e, "hipMalloc face fields");
            if ((e = hipMemset(*q, 0, fb)) != hipSuccess) return cleanup(e, "hipMemset");
        }
        if ((e = hipMemset(blk->ct_cemf, 0, sizeof(double) * 3 * nc)) != hipSuccess) return cleanup(e, "hipMemset");
    }
    // Zero everything once at creation (outside any captured work): the signal
    // speeds of an inactive direction must read 0, as the reference's zero-initialised
    // Flux.cmax/cmin do for X3 in 2D.
    if ((e = hipMemset(d->U0, 0, st_bytes)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->P0, 0, st_bytes)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->U1, 0, st_bytes)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->P1, 0, st_bytes)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->F, 0, fe * 3 * NV * nc)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->ctop, 0, fe * 3 * nc)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->divb_part, 0, sizeof(double) * (divb_blocks(blk->dims) + 1))) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->pflag, 0, sizeof(int) * nc)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->fflag, 0, sizeof(int) * nc)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipMemset(d->dt_hist, 0, sizeof(double) * d->hist_cap)) != hipSuccess) return cleanup(e, "hipMemset");
    if ((e = hipDeviceSynchronize()) != hipSuccess) return cleanup(e, "hipDeviceSynchronize");
    *out = d;
    return KH_OK;
}

void kh_driver_destroy(kh_driver* d) {
    if (!d) return;
    (void)hipDeviceSynchronize();
    for (hipEvent_t e : d->ev_pool) (void)hipEventDestroy(e);
    for (void* ptr : {(void*)d->U0, (void*)d->P0, (void*)d->U1, (void*)d->P1, (void*)d->P0f, (void*)d->P1f, d->F, d->ctop,
                      (void*)d->pflag, (void*)d->fflag, (void*)d->fail_list, (void*)d->ds, (void*)d->dt_hist,
                      (void*)d->dt_part, (void*)d->divb_part, (void*)d->obj, (void*)d->fofc_mark,
                      (void*)d->fB0, (void*)d->fB1, (void*)d->emf})
        if (ptr) (void)hipFree(ptr);
    kh_block_destroy(d->blk);
    delete d;
}

int kh_driver_set_stream(kh_driver* d, void* s) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    return kh_block_set_stream(d->blk, s);
}

int kh_driver_block(kh_driver* d, kh_block** out) {
    if (!d || !out) return fail(KH_ERR_INVALID, "null argument");
    *out = d->blk;
    return KH_OK;
}

int kh_driver_arrays(kh_driver* d, double** U, double** P) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    if (U) *U = d->U0;
    if (P) *P = d->P0;
    return KH_OK;
}

int kh_driver_init_prims(kh_driver* d, const double* P) {
    if (!d || !P) return fail(KH_ERR_INVALID, "null argument");
    if (d->blk->dims.face_ct) return fail(KH_ERR_INVALID, "a face_ct driver starts from kh_driver_init_faces");
    kh_block* b = d->blk;
    const size_t st_bytes = sizeof(double) * NV * b->dims.nc;
    KH_HIP(hipMemcpyAsync(d->P0, P, st_bytes, hipMemcpyDeviceToDevice, b->stream));
    KH_HIP(launch_sync(b->dims, d->P0, NV, b->stream));
    KH_HIP(launch_ptou(b->dims, b->consts, b->geo, d->P0, d->U0, b->stream));
    KH_HIP(launch_sync(b->dims, d->U0, NV, b->stream));
    KH_HIP(launch_bounds(b->dims, b->consts, b->geo, d->U0, d->P0, nullptr, b->stream));
    KH_HIP(hipMemcpyAsync(d->P1, d->P0, st_bytes, hipMemcpyDeviceToDevice, b->stream));
    KH_HIP(hipMemcpyAsync(d->U1, d->U0, st_bytes, hipMemcpyDeviceToDevice, b->stream));
    if (d->P0f) {
        KH_HIP(launch_convert_f32(d->P0, d->P0f, NV * b->dims.nc, b->stream));
        KH_HIP(launch_convert_f32(d->P0, d->P1f, NV * b->dims.nc, b->stream));
    }
    KH_HIP(launch_reset_state(d->ds, b->p.dt_start, b->stream));
    return KH_OK;
}

int kh_driver_init_faces(kh_driver* d, const double* P, const double* fB) {
    if (!d || !P || !fB) return fail(KH_ERR_INVALID, "null argument");
    kh_block* b = d->blk;
    const BlockDims& bd = b->dims;
    if (!bd.face_ct) return fail(KH_ERR_INVALID, "kh_driver_init_faces needs b_solver = KH_BSOLVER_FACE_CT");
    const size_t st_bytes = sizeof(double) * NV * bd.nc, fb = sizeof(double) * 3 * bd.nf;
    hipStream_t s = b->stream;
    KH_HIP(hipMemcpyAsync(d->P0, P, st_bytes, hipMemcpyDeviceToDevice, s));
    KH_HIP(hipMemcpyAsync(d->fB0, fB, fb, hipMemcpyDeviceToDevice, s));
    KH_HIP(launch_ct_sync(bd, d->fB0, s));
    KH_HIP(launch_sync(bd, d->P0, NV, s));
    // B_CT::BlockUtoP over the entire block (prob/seed_B.cpp:191), then PostInitialize's P->U
    KH_HIP(launch_ct_utop(bd, d->fB0, vars_of(d->U0, bd.nc), vars_of(d->P0, bd.nc), s));
    KH_HIP(launch_ptou(bd, b->consts, b->geo, d->P0, d->U0, s));
    KH_HIP(launch_sync(bd, d->U0, NV, s));
    KH_HIP(hipMemcpyAsync(d->P1, d->P0, st_bytes, hipMemcpyDeviceToDevice, s));
    KH_HIP(hipMemcpyAsync(d->U1, d->U0, st_bytes, hipMemcpyDeviceToDevice, s));
    KH_HIP(hipMemcpyAsync(d->fB1, d->fB0, fb, hipMemcpyDeviceToDevice, s));
    if (d->P0f) {
        KH_HIP(launch_convert_f32(d->P0, d->P0f, NV * bd.nc, s));
        KH_HIP(launch_convert_f32(d->P0, d->P1f, NV * bd.nc, s));
    }
    KH_HIP(launch_reset_state(d->ds, b->p.dt_start, s));
    return KH_OK;
}

int kh_driver_get_faces(kh_driver* d, double* fB) {
    if (!d || !fB) return fail(KH_ERR_INVALID, "null argument");
    const BlockDims& bd = d->blk->dims;
    if (!bd.face_ct) return fail(KH_ERR_INVALID, "the driver has no face fields (b_solver is flux_ct)");
    KH_HIP(launch_ct_sync(bd, d->fB0, d->blk->stream));
    KH_HIP(hipMemcpyAsync(fB, d->fB0, sizeof(double) * 3 * bd.nf, hipMemcpyDeviceToDevice, d->blk->stream));
    return KH_OK;
}

// One full step: two stages of the default task collection, then the divB
// diagnostic and the device-side timestep update.
// Launch `expr` with optional start/stop events for kernel kind `kind`.
#define KH_TIMED(kind, expr)                                                   \
    do {                                                                        \
        size_t ia_ = 0;                                                         \
        hipEvent_t ev_;                                                         \
        if (d->timing) { ia_ = d->ev_used; KH_HIP(d->mark(&ev_)); }            \
        KH_HIP(expr);                                                           \
        if (d->timing) { KH_HIP(d->mark(&ev_)); d->recs.push_back({(kind), ia_, d->ev_used - 1}); } \
    } while (0)

}  // extern "C"

int kh::check_face_objective(const BlockDims& b, int dir, int var, int face) {
    if (dir < 1 || dir > b.ndim) return fail(KH_ERR_INVALID, "objective direction outside the block's dimensions");
    if (var < 0 || var >= B1) return fail(KH_ERR_INVALID, "objective variable must be a fluid variable (0..4)");
    const int lo = (dir == 1) ? b.is : (dir == 2) ? b.js : b.ks;
    const int hi = (dir == 1) ? b.ie + 1 : (dir == 2) ? b.je + 1 : b.ke + 1;
    if (face < lo || face > hi) return fail(KH_ERR_INVALID, "objective face outside the block's faces");
    return KH_OK;
}

// One stage of the default task collection on the driver's buffers: fluxes,
// flux fixes, the fused update (+ inversion, floors, P->U), fixups, boundaries.
int kh::driver_stage(kh_driver* d, int stage) {
    kh_block* b = d->blk;
    const BlockDims& bd = b->dims;
    const StepConsts& c = b->consts;
    const int prec = b->p.precision;
    const size_t fe = flux_elem(b);
    hipStream_t s = b->stream;
    char* F = (char*)d->F;
    char* ct = (char*)d->ctop;
    const bool physical = has_physical_face(bd);
    const double* Psub = (stage == 1) ? d->P0 : d->P1;
    const float* Psubf = (stage == 1) ? d->P0f : d->P1f;
    const double* Usub = (stage == 1) ? d->U0 : d->U1;
    double* Pout = (stage == 1) ? d->P1 : d->P0;
    float* Poutf = (stage == 1) ? d->P1f : d->P0f;
    double* Uout = (stage == 1) ? d->U1 : d->U0;
    // signal speeds only feed the timestep estimate of the last stage
    const int speeds = (stage == 2) ? SPEEDS_CTOP : SPEEDS_NONE;
    for (int dir = 1; dir <= bd.ndim; ++dir)
        KH_TIMED(dir - 1, launch_flux_stage(bd, c, b->geo, dir, b->p.recon, prec, speeds, Psub, Psubf,
                                            F + fe * (dir - 1) * NV * bd.nc, ct + fe * (dir - 1) * bd.nc, s));
    if (b->p.fofc) {
        // KHARMADriver::AddFOFC (kharma_step.cpp:168-173): on the raw fluxes, before the flux fixes
        KH_TIMED(3, launch_fofc_mark(bd, c, b->geo, stage, (const double*)d->F, Usub, d->U0, Psub, b->fofc_fixed,
                                     d->ds, d->fofc_mark, s));
        KH_TIMED(3, launch_fofc_flux(bd, c, b->geo, speeds, vars_of(Psub, bd.nc), d->fofc_mark,
                                     flux_vars_of((double*)d->F, bd.nc), (double*)d->ctop, nullptr, s));
    }
    if (physical) KH_TIMED(3, launch_fix_flux(bd, prec, d->F, s));
    if (d->obj_dir)
        KH_HIP(launch_face_flux_objective(bd, prec, d->F, d->obj_dir, d->obj_var, d->obj_face,
                                          d->obj_weight * stage_flux_weight(c, stage), d->ds, d->obj, s));
    const RecordLayout L = record_layout(bd);
    const int q = stage - 1;
    if (d->rec && physical)
        KH_HIP(launch_save_x1_flux(bd, (const double*)d->F, (double*)(d->rec + L.x1f[q]), s));
    if (bd.face_ct) {
        // B_CT::CalculateEMF (kharma_step.cpp:183-190) on the stage's fluxes and
        // primitives, B_CT::AddSource with the face half of AddStateUpdate, then
        // the cell update with the cell B from the new faces
        double* fBsub = (stage == 1) ? d->fB0 : d->fB1;
        double* fBout = (stage == 1) ? d->fB1 : d->fB0;
        const bool same = (c.gam0[stage - 1] == 0.) && (fBsub == d->fB0);
        if (bd.ct_scheme != KH_CT_BS99) KH_TIMED(3, launch_ct_cemf(bd, vars_of(Psub, bd.nc), b->ct_cemf, s));
        KH_TIMED(3, launch_ct_emf(bd, prec, flux_vars_cast<const void>(flux_vars_prec(d->F, prec, bd.nc)), b->ct_cemf,
                                  d->emf, s));
        KH_TIMED(3, launch_ct_face_update(bd, d->emf, fBsub, d->fB0, c.gam0[stage - 1], c.gam1[stage - 1], same, d->ds,
                                          c.beta[stage - 1], fBout, s));
        KH_TIMED(3 + stage, launch_update_faces(bd, c, b->geo, prec, stage, stage == 2, d->F, d->ctop, Usub, d->U0, Psub,
                                                Uout, Pout, Poutf, d->pflag, d->fflag, d->fail_list, d->ds, d->dt_part,
                                                fBout, s));
    } else if (d->rec) {
        KH_TIMED(3 + stage, launch_update_recorded(bd, c, b->geo, stage, stage == 2, d->F, d->ctop, Usub, d->U0, Psub,
                                                   Uout, Pout, d->pflag, d->fflag, d->fail_list, d->ds, d->dt_part,
                                                   (double*)(d->rec + L.Ustar[q]), s));
    } else {
        KH_TIMED(3 + stage, launch_update(bd, c, b->geo, prec, stage, stage == 2, d->F, d->ctop, Usub, d->U0, Psub,
                                          Uout, Pout, Poutf, d->pflag, d->fflag, d->fail_list, d->ds, d->dt_part, s));
    }
    if (c.fix_average || c.fix_atmo)
        KH_TIMED(6, launch_fixup(bd, c, b->geo, stage, Uout, Pout, Poutf, d->pflag, d->fail_list, d->ds, s));
    if (physical) KH_TIMED(3, launch_bounds(bd, c, b->geo, Uout, Pout, Poutf, s));
    if (d->rec) {
        // the stage's flags (after the fix-up), failed-inversion count and, for stage 1, output
        const size_t fb = sizeof(int) * bd.nc;
        KH_HIP(hipMemcpyAsync(d->rec + L.pflag[q], d->pflag, fb, hipMemcpyDeviceToDevice, s));
        KH_HIP(hipMemcpyAsync(d->rec + L.fflag[q], d->fflag, fb, hipMemcpyDeviceToDevice, s));
        KH_HIP(hipMemcpyAsync(d->rec + L.fails + sizeof(int) * q,
                              reinterpret_cast<const char*>(d->ds) + offsetof(DeviceState, fail_count) + sizeof(int) * q,
                              sizeof(int), hipMemcpyDeviceToDevice, s));
        if (stage == 1)
            KH_HIP(hipMemcpyAsync(d->rec + L.P1, Pout, sizeof(double) * NV * bd.nc, hipMemcpyDeviceToDevice, s));
    }
    return KH_OK;
}

extern "C" {

int kh_driver_step(kh_driver* d, int nsteps) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    kh_block* b = d->blk;
    const BlockDims& bd = b->dims;
    const StepConsts& c = b->consts;
    hipStream_t s = b->stream;
    const int n_dt = (int)update_blocks(bd), n_divb = (int)divb_blocks(bd);
    for (int n = 0; n < nsteps; ++n) {
        for (int stage = 1; stage <= 2; ++stage) {
            const int st = driver_stage(d, stage);
            if (st != KH_OK) return st;
        }
        if (bd.face_ct) KH_TIMED(7, launch_ct_max_divb(bd, d->fB0, d->divb_part, s));
        else KH_TIMED(7, launch_max_divb(bd, d->U0, d->divb_part, s));
        KH_TIMED(8, launch_finalize(c, d->ds, d->dt_part, n_dt, d->divb_part, bd.ndim > 1 ? n_divb : 0,
                                    d->dt_hist, d->hist_cap, s));
    }
    return KH_OK;
}

int kh_driver_get_state(kh_driver* d, double* P, double* U) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    kh_block* b = d->blk;
    const size_t st_bytes = sizeof(double) * NV * b->dims.nc;
    if (P) KH_HIP(hipMemcpyAsync(P, d->P0, st_bytes, hipMemcpyDeviceToDevice, b->stream));
    if (U) KH_HIP(hipMemcpyAsync(U, d->U0, st_bytes, hipMemcpyDeviceToDevice, b->stream));
    return KH_OK;
}

int kh_driver_save_restart(kh_driver* d, const char* path) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    kh_block* b = d->blk;
    DeviceState h;
    KH_HIP(hipMemcpyAsync(&h, d->ds, sizeof(h), hipMemcpyDeviceToHost, b->stream));
    KH_HIP(hipStreamSynchronize(b->stream));
    const int n[3] = {b->dims.n1, b->dims.n2, b->dims.n3};
    if (b->dims.face_ct) KH_HIP(launch_ct_sync(b->dims, d->fB0, b->stream));
    return write_restart(path, b->p, n, h.time, h.dt, h.ncycle, d->P0, d->U0, b->stream, d->fB0);
}

int kh_driver_load_restart(kh_driver* d, const char* path) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    kh_block* b = d->blk;
    const int n[3] = {b->dims.n1, b->dims.n2, b->dims.n3};
    double time = 0., dt = 0.;
    int ncycle = 0;
    const int st = read_restart(path, b->p, n, &time, &dt, &ncycle, d->P0, d->U0, b->stream, d->fB0);
    if (st != KH_OK) return st;
    const size_t st_bytes = sizeof(double) * NV * b->dims.nc;
    if (d->fB0)
        KH_HIP(hipMemcpyAsync(d->fB1, d->fB0, sizeof(double) * 3 * b->dims.nf, hipMemcpyDeviceToDevice, b->stream));
    KH_HIP(hipMemcpyAsync(d->P1, d->P0, st_bytes, hipMemcpyDeviceToDevice, b->stream));
    KH_HIP(hipMemcpyAsync(d->U1, d->U0, st_bytes, hipMemcpyDeviceToDevice, b->stream));
    if (d->P0f) {
        KH_HIP(launch_convert_f32(d->P0, d->P0f, NV * b->dims.nc, b->stream));
        KH_HIP(launch_convert_f32(d->P0, d->P1f, NV * b->dims.nc, b->stream));
    }
    KH_HIP(launch_reset_state(d->ds, dt, b->stream));
    KH_HIP(launch_set_clock(d->ds, time, dt, ncycle, b->stream));
    return KH_OK;
}

int kh_driver_get_flags(kh_driver* d, int* pflag, int* fflag) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    const size_t bytes = sizeof(int) * d->blk->dims.nc;
    if (pflag) KH_HIP(hipMemcpyAsync(pflag, d->pflag, bytes, hipMemcpyDeviceToDevice, d->blk->stream));
    if (fflag) KH_HIP(hipMemcpyAsync(fflag, d->fflag, bytes, hipMemcpyDeviceToDevice, d->blk->stream));
    return KH_OK;
}

int kh_driver_check(kh_driver* d) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    DeviceState h;
    KH_HIP(hipMemcpyAsync(&h, d->ds, sizeof(h), hipMemcpyDeviceToHost, d->blk->stream));
    KH_HIP(hipStreamSynchronize(d->blk->stream));
    if (h.divb_over_cycle > 0) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "DivB exceeds maximum! max corner divB %.6g > %.6g after cycle %d",
                      h.divb_over_value, d->blk->consts.divb_over, h.divb_over_cycle);
        return fail(KH_ERR_DIVB, msg);
    }
    return KH_OK;
}

int kh_driver_get_scalars(kh_driver* d, double* time, double* dt, double* last_divb, int* ncycle, int* total_fails) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    DeviceState h;
    KH_HIP(hipMemcpyAsync(&h, d->ds, sizeof(h), hipMemcpyDeviceToHost, d->blk->stream));
    KH_HIP(hipStreamSynchronize(d->blk->stream));
    if (time) *time = h.time;
    if (dt) *dt = h.dt;
    if (last_divb) *last_divb = h.last_divb;
    if (ncycle) *ncycle = h.ncycle;
    if (total_fails) *total_fails = h.total_fails;
    return KH_OK;
}

int kh_driver_set_clock(kh_driver* d, double time, double dt, int ncycle) {
    if (!d || !(dt > 0.) || ncycle < 0) return fail(KH_ERR_INVALID, "bad argument");
    KH_HIP(launch_set_clock(d->ds, time, dt, ncycle, d->blk->stream));
    return KH_OK;
}

int kh_driver_stage_record_bytes(kh_driver* d, size_t* bytes) {
    if (!d || !bytes) return fail(KH_ERR_INVALID, "null argument");
    *bytes = record_layout(d->blk->dims).total;
    return KH_OK;
}

int kh_driver_set_stage_record(kh_driver* d, void* rec) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    if (rec && d->blk->p.precision != KH_PREC_FP64)
        return fail(KH_ERR_UNSUPPORTED, "step records are kept by the FP64 driver (the adjoint's form)");
    if (rec && d->blk->dims.face_ct) return fail(KH_ERR_UNSUPPORTED, "the adjoint does not cover face_ct");
    if (rec && (reinterpret_cast<uintptr_t>(rec) & 255))
        return fail(KH_ERR_INVALID, "a step record must be 256-byte aligned");
    d->rec = static_cast<char*>(rec);
    return KH_OK;
}

int kh_driver_set_face_flux_objective(kh_driver* d, int dir, int var, int face, double weight) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    if (dir == 0) { d->obj_dir = 0; return KH_OK; }
    const int st = check_face_objective(d->blk->dims, dir, var, face);
    if (st != KH_OK) return st;
    if (!d->obj) {
        KH_HIP(hipMalloc(&d->obj, sizeof(double)));
        KH_HIP(hipMemsetAsync(d->obj, 0, sizeof(double), d->blk->stream));
    }
    d->obj_dir = dir; d->obj_var = var; d->obj_face = face; d->obj_weight = weight;
    return KH_OK;
}

int kh_driver_objective(kh_driver* d, double* value, int reset) {
    if (!d || !value) return fail(KH_ERR_INVALID, "null argument");
    *value = 0.;
    if (!d->obj) return KH_OK;
    KH_HIP(hipMemcpyAsync(value, d->obj, sizeof(double), hipMemcpyDeviceToHost, d->blk->stream));
    if (reset) KH_HIP(hipMemsetAsync(d->obj, 0, sizeof(double), d->blk->stream));
    KH_HIP(hipStreamSynchronize(d->blk->stream));
    return KH_OK;
}

int kh_driver_set_kernel_timing(kh_driver* d, int enable) {
    if (!d) return fail(KH_ERR_INVALID, "null driver");
    d->timing = enable != 0;
    return KH_OK;
}

int kh_driver_kernel_times(kh_driver* d, double* total_ms, int* launches, int nkinds) {
    if (!d || !total_ms || !launches || nkinds < KH_NKERNEL_KINDS) return fail(KH_ERR_INVALID, "bad argument");
    KH_HIP(hipStreamSynchronize(d->blk->stream));
    for (int k = 0; k < nkinds; ++k) { total_ms[k] = 0.; launches[k] = 0; }
    for (const auto& r : d->recs) {
        float ms = 0.f;
        KH_HIP(hipEventElapsedTime(&ms, d->ev_pool[r.a], d->ev_pool[r.b]));
        total_ms[r.kind] += ms;
        launches[r.kind] += 1;
    }
    d->recs.clear();
    d->ev_used = 0;
    return KH_OK;
}

int kh_driver_get_dt_history(kh_driver* d, double* out, int n) {
    if (!d || !out || n < 0) return fail(KH_ERR_INVALID, "bad argument");
    if (n > d->hist_cap) return fail(KH_ERR_INVALID, "history request exceeds capacity");
    KH_HIP(hipMemcpyAsync(out, d->dt_hist, sizeof(double) * n, hipMemcpyDeviceToHost, d->blk->stream));
    KH_HIP(hipStreamSynchronize(d->blk->stream));
    return KH_OK;
}

}  // extern "C"
