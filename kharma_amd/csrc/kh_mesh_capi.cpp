// kh_mesh_capi.cpp -- the stage driver over a uniform mesh of several meshblocks on
// one GPU (include/kharma_amd.h, kh_mesh_*).
//
// KHARMA runs a mesh of equal meshblocks through MakeDefaultTaskCollection
// (driver/kharma_step.cpp:77-319), exchanging the ghost zones of shared faces
// after each stage (KHARMADriver::AddBoundarySync, kharma_driver.cpp:138-183)
// and reducing the timestep and divB over all blocks (kharma_driver.cpp:439-460,
// b_flux_ct.cpp:566-643).  Here all blocks of the mesh live in one allocation
// per container, [block][8][nc], and there is no exchange pass: each block's
// update (and inversion fix-up) stores every interior zone within ng of a
// face shared with another block into that block's ghosts as well (the image
// stores of kh_step.hip).  Per stage: the face fluxes of every block, then the
// physical-face flux fixes, the updates, the fix-ups and the physical boundary
// fills, each over every block, so that no block reads a ghost zone before the
// block owning it has written it; per step one divB and one finalize over the
// partials of all blocks.  By default every one of these kernels is one launch
// over all blocks (the blocks' dims and geometry in a device table, kh::Pack),
// as the reference launches over a MeshBlockPack; kh_mesh_set_packed(m, 0)
// launches once per block instead, with the same results bit for bit.
#include "kharma_amd.h"

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "kh_host.hpp"

using namespace kh;

struct kh_mesh {
    kh_params p;
    MeshDims md{};
    std::vector<kh_block*> blk;
    hipStream_t stream = nullptr;
    double *U0 = nullptr, *P0 = nullptr, *U1 = nullptr, *P1 = nullptr;
    float *P0f = nullptr, *P1f = nullptr;
    void *F = nullptr, *ctop = nullptr;
    int *pflag = nullptr, *fflag = nullptr, *fail_list = nullptr;
    int* fofc_mark = nullptr;               // fofc/on: marks of all blocks, ghosts of shared faces synced
    DeviceState* ds = nullptr;              // one per block; the first carries the clock
    double *dt_part = nullptr, *divb_part = nullptr, *dt_hist = nullptr;
    long n_dt = 0, n_divb = 0;              // partials per block
    int hist_cap = 1 << 16;
    // kh_mesh_set_graph: one step captured once (on a private stream) as a HIP
    // graph and replayed on the mesh's stream
    bool use_graph = false;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // one launch per kernel over all blocks (kh_mesh_set_packed)
    bool packed = true;
    Pack pack;
    BlockEntry* pack_dev = nullptr;
};

namespace {

#define KM_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return api_fail(KH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int mesh_stage(kh_mesh* m, int stage, hipStream_t s) {
    const MeshDims& md = m->md;
    const long nc = md.nc;
    const int prec = m->p.precision;
    const size_t fe = (prec == KH_PREC_SPLIT) ? sizeof(float) : sizeof(double);
    double* Psub = (stage == 1) ? m->P0 : m->P1;
    float* Psubf = (stage == 1) ? m->P0f : m->P1f;
    double* Usub = (stage == 1) ? m->U0 : m->U1;
    double* Pout = (stage == 1) ? m->P1 : m->P0;
    float* Poutf = (stage == 1) ? m->P1f : m->P0f;
    double* Uout = (stage == 1) ? m->U1 : m->U0;
    const int speeds = (stage == 2) ? SPEEDS_CTOP : SPEEDS_NONE;
    const long st = NV * nc;   // one block's state
    auto F = [&](int b) { return (char*)m->F + fe * 3 * st * b; };
    auto ct = [&](int b) { return (char*)m->ctop + fe * 3 * nc * b; };
    auto f32 = [&](float* a, int b) { return a ? a + st * b : nullptr; };
    const StepConsts& c = m->blk[0]->consts;
    // KHARMADriver::AddFOFC (kharma_driver.cpp:291-337) after every block's raw
    // fluxes: the marks of all blocks, stored with their images in the
    // neighbours' ghosts (the flag's boundary sync), then the replaced fluxes.
    auto fofc = [&]() -> int {
        if (!m->p.fofc) return KH_OK;
        for (int b = 0; b < md.nblk; ++b) {
            const kh_block* k = m->blk[b];
            KM_HIP(launch_fofc_mark(k->dims, k->consts, k->geo, stage, (const double*)F(b), Usub + st * b,
                                    m->U0 + st * b, Psub + st * b, k->fofc_fixed, m->ds + b, m->fofc_mark + nc * b,
                                    s, true));
        }
        for (int b = 0; b < md.nblk; ++b) {
            const kh_block* k = m->blk[b];
            KM_HIP(launch_fofc_flux(k->dims, k->consts, k->geo, speeds, vars_of((const double*)(Psub + st * b), nc),
                                    m->fofc_mark + nc * b, flux_vars_of((double*)F(b), nc), (double*)ct(b), nullptr,
                                    s, true));
        }
        return KH_OK;
    };
    if (m->packed) {
        // every kernel once over all blocks, with block 0's arguments
        const kh_block* k = m->blk[0];
        const Pack* pk = &m->pack;
        for (int dir = 1; dir <= k->dims.ndim; ++dir)
            KM_HIP(launch_flux_stage(k->dims, k->consts, k->geo, dir, k->p.recon, prec, speeds, Psub, Psubf,
                                     F(0) + fe * (dir - 1) * st, ct(0) + fe * (dir - 1) * nc, s, pk));
        if (const int r = fofc(); r != KH_OK) return r;
        if (pk->fix_flux_n) KM_HIP(launch_fix_flux(k->dims, prec, F(0), s, pk));
        KM_HIP(launch_update(k->dims, k->consts, k->geo, prec, stage, stage == 2, F(0), ct(0), Usub, m->U0, Psub, Uout,
                             Pout, Poutf, m->pflag, m->fflag, m->fail_list, m->ds, m->dt_part, s, pk));
        if (c.fix_average || c.fix_atmo)
            KM_HIP(launch_fixup(k->dims, k->consts, k->geo, stage, Uout, Pout, Poutf, m->pflag, m->fail_list, m->ds, s,
                                pk));
        if (pk->bounds_n) KM_HIP(launch_bounds(k->dims, k->consts, k->geo, Uout, Pout, Poutf, s, pk));
        return KH_OK;
    }
    for (int b = 0; b < md.nblk; ++b) {
        const kh_block* k = m->blk[b];
        for (int dir = 1; dir <= k->dims.ndim; ++dir)
            KM_HIP(launch_flux_stage(k->dims, k->consts, k->geo, dir, k->p.recon, prec, speeds, Psub + st * b,
                                     f32(Psubf, b), F(b) + fe * (dir - 1) * st, ct(b) + fe * (dir - 1) * nc, s));
    }
    if (const int r = fofc(); r != KH_OK) return r;
    for (int b = 0; b < md.nblk; ++b) {
        const kh_block* k = m->blk[b];
        if (has_physical_face(k->dims)) KM_HIP(launch_fix_flux(k->dims, prec, F(b), s));
    }
    for (int b = 0; b < md.nblk; ++b) {
        const kh_block* k = m->blk[b];
        KM_HIP(launch_update(k->dims, k->consts, k->geo, prec, stage, stage == 2, F(b), ct(b), Usub + st * b,
                             m->U0 + st * b, Psub + st * b, Uout + st * b, Pout + st * b, f32(Poutf, b),
                             m->pflag + nc * b, m->fflag + nc * b, m->fail_list + nc * b, m->ds + b,
                             m->dt_part + m->n_dt * b, s));
    }
    if (c.fix_average || c.fix_atmo)
        for (int b = 0; b < md.nblk; ++b) {
            const kh_block* k = m->blk[b];
            KM_HIP(launch_fixup(k->dims, k->consts, k->geo, stage, Uout + st * b, Pout + st * b, f32(Poutf, b),
                                m->pflag + nc * b, m->fail_list + nc * b, m->ds + b, s));
        }
    for (int b = 0; b < md.nblk; ++b) {
        const kh_block* k = m->blk[b];
        if (has_physical_face(k->dims))
            KM_HIP(launch_bounds(k->dims, k->consts, k->geo, Uout + st * b, Pout + st * b, f32(Poutf, b), s));
    }
    return KH_OK;
}

// One full step on stream s: both stages over all blocks, then divB and the
// finalize over the partials of all blocks.
int mesh_step_body(kh_mesh* m, hipStream_t s) {
    const MeshDims& md = m->md;
    const long st = NV * md.nc;
    const StepConsts& c = m->blk[0]->consts;
    for (int stage = 1; stage <= 2; ++stage) {
        const int r = mesh_stage(m, stage, s);
        if (r != KH_OK) return r;
    }
    if (md.ndim > 1 && m->packed)
        KM_HIP(launch_max_divb(m->blk[0]->dims, m->U0, m->divb_part, s, &m->pack));
    else if (md.ndim > 1)
        for (int b = 0; b < md.nblk; ++b)
            KM_HIP(launch_max_divb(m->blk[b]->dims, m->U0 + st * b, m->divb_part + m->n_divb * b, s));
    KM_HIP(launch_finalize(c, m->ds, m->dt_part, (int)(m->n_dt * md.nblk), m->divb_part,
                           md.ndim > 1 ? (int)(m->n_divb * md.nblk) : 0, m->dt_hist, m->hist_cap, s, md.nblk));
    return KH_OK;
}

void drop_graph(kh_mesh* m) {
    if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->gexec = nullptr;
    m->graph = nullptr;
}

}  // namespace

extern "C" {

int kh_mesh_create(const kh_params* p, int nb1, int nb2, int nb3, kh_mesh** out) {
    if (!out) return api_fail(KH_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int st = check_params(p);
    if (st != KH_OK) return st;
    const int ndim = (p->nx3 > 1) ? 3 : ((p->nx2 > 1) ? 2 : 1);
    const int nb[3] = {nb1, nb2, nb3};
    const int N[3] = {p->nx1, p->nx2, p->nx3};
    for (int d = 0; d < 3; ++d) {
        if (nb[d] < 1) return api_fail(KH_ERR_INVALID, "block counts must be >= 1");
        if (d >= ndim && nb[d] != 1) return api_fail(KH_ERR_INVALID, "an inactive axis holds one block");
        if (N[d] % nb[d]) return api_fail(KH_ERR_INVALID, "the mesh's zones must divide evenly into blocks");
        if (nb[d] > 1 && N[d] / nb[d] < 2 * p->ng)
            return api_fail(KH_ERR_INVALID, "a block shared along an axis needs nx >= 2 ng zones along it");
    }
    if (p->b_solver == KH_BSOLVER_FACE_CT)
        return api_fail(KH_ERR_UNSUPPORTED, "face_ct runs on one-block meshes (kh_driver_*) only");
    kh_mesh* m = new kh_mesh();
    m->p = *p;
    MeshDims& md = m->md;
    md.ng = p->ng;
    md.ndim = ndim;
    md.nblk = nb1 * nb2 * nb3;
    md.periodic[0] = p->periodic1;
    md.periodic[1] = ndim > 1 ? p->periodic2 : 1;
    md.periodic[2] = 1;
    md.ncm = 1;
    for (int d = 0; d < 3; ++d) {
        md.nb[d] = nb[d];
        md.N[d] = N[d];
        md.nx[d] = N[d] / nb[d];
        md.Ne[d] = (d < ndim) ? N[d] + 2 * p->ng : 1;
        md.n[d] = (d < ndim) ? md.nx[d] + 2 * p->ng : 1;
        md.ncm *= md.Ne[d];
    }
    md.nc = (long)md.n[0] * md.n[1] * md.n[2];
    auto cleanup = [&](int code, const std::string& what) {
        kh_mesh_destroy(m);
        return api_fail(code, what);
    };
    for (int b3 = 0; b3 < nb3; ++b3)
        for (int b2 = 0; b2 < nb2; ++b2)
            for (int b1 = 0; b1 < nb1; ++b1) {
                const int bidx[3] = {b1, b2, b3};
                kh_block* k = nullptr;
                st = block_create_in_mesh(p, nb, bidx, &k);
                if (st != KH_OK) {
                    const std::string msg = kh_last_error();
                    return cleanup(st, msg);
                }
                m->blk.push_back(k);
            }
    // The curved-space tables depend on (x1, x2) only, so every column of blocks
    // along x3 reads the tables of its first block (one copy per column in the
    // caches instead of one per block).
    if (m->blk[0]->dims.geo)
        for (int b = nb1 * nb2; b < md.nblk; ++b) {
            kh_block* k = m->blk[b];
            for (void* q : k->geo_bufs) (void)hipFree(q);
            k->geo_bufs.clear();
            k->geo = m->blk[b % (nb1 * nb2)]->geo;
        }
    const BlockDims& bd = m->blk[0]->dims;
    m->n_dt = update_blocks(bd);
    m->n_divb = divb_blocks(bd);
    const long nc = md.nc, nblk = md.nblk;
    const size_t sb = sizeof(double) * NV * nc * nblk;
    const size_t fe = (p->precision == KH_PREC_SPLIT) ? sizeof(float) : sizeof(double);
    hipError_t e = hipSuccess;
    auto alloc = [&](void** q, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(q, bytes);
        if (e == hipSuccess) e = hipMemset(*q, 0, bytes);
    };
    alloc((void**)&m->U0, sb);
    alloc((void**)&m->P0, sb);
    alloc((void**)&m->U1, sb);
    alloc((void**)&m->P1, sb);
    if (p->precision == KH_PREC_SPLIT) {
        alloc((void**)&m->P0f, sizeof(float) * NV * nc * nblk);
        alloc((void**)&m->P1f, sizeof(float) * NV * nc * nblk);
    }
    alloc(&m->F, fe * 3 * NV * nc * nblk);
    alloc(&m->ctop, fe * 3 * nc * nblk);
    alloc((void**)&m->pflag, sizeof(int) * nc * nblk);
    alloc((void**)&m->fflag, sizeof(int) * nc * nblk);
    alloc((void**)&m->fail_list, sizeof(int) * nc * nblk);
    if (p->fofc) alloc((void**)&m->fofc_mark, sizeof(int) * nc * nblk);
    alloc((void**)&m->ds, sizeof(DeviceState) * nblk);
    alloc((void**)&m->dt_part, sizeof(double) * (m->n_dt * nblk + 1));
    alloc((void**)&m->divb_part, sizeof(double) * (m->n_divb * nblk + 1));
    alloc((void**)&m->dt_hist, sizeof(double) * m->hist_cap);
    alloc((void**)&m->pack_dev, sizeof(BlockEntry) * nblk);
    if (e == hipSuccess) {
        std::vector<BlockEntry> tab(nblk);
        std::vector<BlockDims> dims(nblk);
        for (long b = 0; b < nblk; ++b) {
            tab[b].dims = dims[b] = m->blk[b]->dims;
            tab[b].geo = m->blk[b]->geo;
        }
        e = hipMemcpy(m->pack_dev, tab.data(), sizeof(BlockEntry) * nblk, hipMemcpyHostToDevice);
        m->pack.dev = m->pack_dev;
        m->pack.nblk = (int)nblk;
        pack_set_sizes(m->pack, dims.data(), (int)nblk);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return cleanup(KH_ERR_HIP, std::string("device allocation: ") + hipGetErrorString(e));
    *out = m;
    return KH_OK;
}

void kh_mesh_destroy(kh_mesh* m) {
    if (!m) return;
    (void)hipDeviceSynchronize();
    for (void* q : {(void*)m->U0, (void*)m->P0, (void*)m->U1, (void*)m->P1, (void*)m->P0f, (void*)m->P1f, m->F,
                    m->ctop, (void*)m->pflag, (void*)m->fflag, (void*)m->fail_list, (void*)m->fofc_mark, (void*)m->ds,
                    (void*)m->dt_part, (void*)m->divb_part, (void*)m->dt_hist, (void*)m->pack_dev})
        if (q) (void)hipFree(q);
    for (kh_block* b : m->blk) kh_block_destroy(b);
    drop_graph(m);
    if (m->cap) (void)hipStreamDestroy(m->cap);
    delete m;
}

int kh_mesh_set_stream(kh_mesh* m, void* s) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    m->stream = (hipStream_t)s;
    for (kh_block* b : m->blk) b->stream = (hipStream_t)s;
    return KH_OK;
}

int kh_mesh_shape(const kh_mesh* m, int* nblocks, int* n1, int* n2, int* n3, long* nc) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    if (nblocks) *nblocks = m->md.nblk;
    if (n1) *n1 = m->md.n[0];
    if (n2) *n2 = m->md.n[1];
    if (n3) *n3 = m->md.n[2];
    if (nc) *nc = m->md.nc;
    return KH_OK;
}

int kh_mesh_init_prims(kh_mesh* m, const double* P) {
    if (!m || !P) return api_fail(KH_ERR_INVALID, "null argument");
    const MeshDims& md = m->md;
    const long st = NV * md.nc;
    hipStream_t s = m->stream;
    KM_HIP(launch_mesh_scatter(md, P, m->P0, s));
    for (int b = 0; b < md.nblk; ++b) {
        const kh_block* k = m->blk[b];
        KM_HIP(launch_ptou(k->dims, k->consts, k->geo, m->P0 + st * b, m->U0 + st * b, s));
    }
    for (int b = 0; b < md.nblk; ++b) {
        const kh_block* k = m->blk[b];
        KM_HIP(launch_bounds(k->dims, k->consts, k->geo, m->U0 + st * b, m->P0 + st * b, nullptr, s));
    }
    const size_t sb = sizeof(double) * st * md.nblk;
    KM_HIP(hipMemcpyAsync(m->P1, m->P0, sb, hipMemcpyDeviceToDevice, s));
    KM_HIP(hipMemcpyAsync(m->U1, m->U0, sb, hipMemcpyDeviceToDevice, s));
    if (m->P0f) {
        KM_HIP(launch_convert_f32(m->P0, m->P0f, st * md.nblk, s));
        KM_HIP(launch_convert_f32(m->P0, m->P1f, st * md.nblk, s));
    }
    for (int b = 0; b < md.nblk; ++b) KM_HIP(launch_reset_state(m->ds + b, m->p.dt_start, s));
    return KH_OK;
}

int kh_mesh_step(kh_mesh* m, int nsteps) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    if (!m->use_graph) {
        for (int n = 0; n < nsteps; ++n) {
            const int r = mesh_step_body(m, m->stream);
            if (r != KH_OK) return r;
        }
        return KH_OK;
    }
    if (!m->gexec) {
        if (!m->cap) KM_HIP(hipStreamCreate(&m->cap));
        KM_HIP(hipStreamBeginCapture(m->cap, hipStreamCaptureModeThreadLocal));
        const int r = mesh_step_body(m, m->cap);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(m->cap, &g);
        if (r != KH_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return r != KH_OK ? r : api_fail(KH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        }
        m->graph = g;
        KM_HIP(hipGraphInstantiate(&m->gexec, g, nullptr, nullptr, 0));
    }
    for (int n = 0; n < nsteps; ++n) KM_HIP(hipGraphLaunch(m->gexec, m->stream));
    return KH_OK;
}

int kh_mesh_set_packed(kh_mesh* m, int enable) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    if (m->packed != (enable != 0)) drop_graph(m);
    m->packed = enable != 0;
    return KH_OK;
}

int kh_mesh_set_graph(kh_mesh* m, int enable) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    m->use_graph = enable != 0;
    if (!m->use_graph) drop_graph(m);
    return KH_OK;
}

int kh_mesh_set_clock(kh_mesh* m, double time, double dt, int ncycle) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    for (int b = 0; b < m->md.nblk; ++b) KM_HIP(launch_set_clock(m->ds + b, time, dt, ncycle, m->stream));
    return KH_OK;
}

int kh_mesh_get_state(kh_mesh* m, double* P, double* U) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    if (P) KM_HIP(launch_mesh_gather(m->md, m->P0, P, m->stream));
    if (U) KM_HIP(launch_mesh_gather(m->md, m->U0, U, m->stream));
    return KH_OK;
}

// The restart file of the mesh: the state in the mesh layout (the file of a
// one-block run of the same mesh, which either driver reads).
int kh_mesh_save_restart(kh_mesh* m, const char* path) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    const MeshDims& md = m->md;
    double* buf = nullptr;
    KM_HIP(hipMalloc(&buf, sizeof(double) * 2 * NV * md.ncm));
    hipError_t e = launch_mesh_gather(md, m->P0, buf, m->stream);
    if (e == hipSuccess) e = launch_mesh_gather(md, m->U0, buf + NV * md.ncm, m->stream);
    DeviceState h{};
    if (e == hipSuccess) e = hipMemcpyAsync(&h, m->ds, sizeof(h), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    int st = (e == hipSuccess) ? KH_OK : api_fail(KH_ERR_HIP, std::string("restart gather: ") + hipGetErrorString(e));
    if (st == KH_OK) st = write_restart(path, m->p, md.Ne, h.time, h.dt, h.ncycle, buf, buf + NV * md.ncm, m->stream);
    (void)hipFree(buf);
    return st;
}

int kh_mesh_load_restart(kh_mesh* m, const char* path) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    const MeshDims& md = m->md;
    double* buf = nullptr;
    KM_HIP(hipMalloc(&buf, sizeof(double) * 2 * NV * md.ncm));
    double time = 0., dt = 0.;
    int ncycle = 0;
    int st = read_restart(path, m->p, md.Ne, &time, &dt, &ncycle, buf, buf + NV * md.ncm, m->stream);
    hipError_t e = hipSuccess;
    if (st == KH_OK) e = launch_mesh_scatter(md, buf, m->P0, m->stream);
    if (st == KH_OK && e == hipSuccess) e = launch_mesh_scatter(md, buf + NV * md.ncm, m->U0, m->stream);
    if (st == KH_OK && e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(buf);
    if (st != KH_OK) return st;
    if (e != hipSuccess) return api_fail(KH_ERR_HIP, std::string("restart scatter: ") + hipGetErrorString(e));
    const size_t sb = sizeof(double) * NV * md.nc * md.nblk;
    KM_HIP(hipMemcpyAsync(m->P1, m->P0, sb, hipMemcpyDeviceToDevice, m->stream));
    KM_HIP(hipMemcpyAsync(m->U1, m->U0, sb, hipMemcpyDeviceToDevice, m->stream));
    if (m->P0f) {
        KM_HIP(launch_convert_f32(m->P0, m->P0f, NV * md.nc * md.nblk, m->stream));
        KM_HIP(launch_convert_f32(m->P0, m->P1f, NV * md.nc * md.nblk, m->stream));
    }
    for (int b = 0; b < md.nblk; ++b) {
        KM_HIP(launch_reset_state(m->ds + b, dt, m->stream));
        KM_HIP(launch_set_clock(m->ds + b, time, dt, ncycle, m->stream));
    }
    return KH_OK;
}

int kh_mesh_block_arrays(kh_mesh* m, int b, double** U, double** P) {
    if (!m || b < 0 || b >= m->md.nblk) return api_fail(KH_ERR_INVALID, "bad block index");
    const long st = NV * m->md.nc;
    if (U) *U = m->U0 + st * b;
    if (P) *P = m->P0 + st * b;
    return KH_OK;
}

int kh_mesh_block(kh_mesh* m, int b, kh_block** out) {
    if (!m || !out || b < 0 || b >= m->md.nblk) return api_fail(KH_ERR_INVALID, "bad block index");
    *out = m->blk[b];
    return KH_OK;
}

int kh_mesh_get_scalars(kh_mesh* m, double* time, double* dt, double* last_divb, int* ncycle, int* total_fails) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    DeviceState h;
    KM_HIP(hipMemcpyAsync(&h, m->ds, sizeof(h), hipMemcpyDeviceToHost, m->stream));
    KM_HIP(hipStreamSynchronize(m->stream));
    if (time) *time = h.time;
    if (dt) *dt = h.dt;
    if (last_divb) *last_divb = h.last_divb;
    if (ncycle) *ncycle = h.ncycle;
    if (total_fails) *total_fails = h.total_fails;
    return KH_OK;
}

int kh_mesh_get_dt_history(kh_mesh* m, double* out, int n) {
    if (!m || !out || n < 0) return api_fail(KH_ERR_INVALID, "bad argument");
    if (n > m->hist_cap) return api_fail(KH_ERR_INVALID, "history request exceeds capacity");
    KM_HIP(hipMemcpyAsync(out, m->dt_hist, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
    KM_HIP(hipStreamSynchronize(m->stream));
    return KH_OK;
}

int kh_mesh_check(kh_mesh* m) {
    if (!m) return api_fail(KH_ERR_INVALID, "null mesh");
    DeviceState h;
    KM_HIP(hipMemcpyAsync(&h, m->ds, sizeof(h), hipMemcpyDeviceToHost, m->stream));
    KM_HIP(hipStreamSynchronize(m->stream));
    if (h.divb_over_cycle > 0) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "DivB exceeds maximum! max corner divB %.6g > %.6g after cycle %d",
                      h.divb_over_value, m->blk[0]->consts.divb_over, h.divb_over_cycle);
        return api_fail(KH_ERR_DIVB, msg);
    }
    return KH_OK;
}

}  // extern "C"
