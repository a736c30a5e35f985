// kh_host.hpp -- host-side objects shared by the C ABI translation units:
// the block context and the stage driver (include/kharma_amd.h).
#pragma once
#include <cstdint>
#include <vector>

#include "kh_internal.hpp"

struct kh_block {
    kh_params p;
    kh::BlockDims dims;
    kh::StepConsts consts;
    kh::GeomDev geo;                     // device geometry cache (curvilinear metrics only)
    std::vector<void*> geo_bufs;     // owned device allocations behind geo
    hipStream_t stream = nullptr;
    kh::DeviceState* scratch = nullptr;  // reductions of the task-level entry points
    double* partials = nullptr;      // one value per work-group of a reduction
    long npartials = 0;
    // First-order flux correction (kh_params.fofc): zones MarkFOFC marks whatever
    // the trial state, [gn2][gn1] like the geometry tables (spherical coordinates only)
    int* fofc_fixed = nullptr;
    // Face-centred CT (b_solver == KH_BSOLVER_FACE_CT): cell-centre EMFs [3][nc]
    double* ct_cemf = nullptr;
};

struct kh_driver {
    kh_block* blk = nullptr;
    double *U0 = nullptr, *P0 = nullptr, *U1 = nullptr, *P1 = nullptr;
    float *P0f = nullptr, *P1f = nullptr;   // FP32 shadows of P (split precision only)
    void *F = nullptr, *ctop = nullptr;
    int *pflag = nullptr, *fflag = nullptr, *fail_list = nullptr;
    kh::DeviceState* ds = nullptr;
    double *dt_part = nullptr, *divb_part = nullptr;
    double* dt_hist = nullptr;
    int hist_cap = 0;
    // face-flux objective (kh_driver_set_face_flux_objective): dir 0 = off
    int obj_dir = 0, obj_var = 0, obj_face = 0;
    double obj_weight = 0.;
    double* obj = nullptr;   // device accumulator
    // step record (kh_driver_set_stage_record): caller's device buffer, or null
    char* rec = nullptr;
    // first-order flux correction (kh_params.fofc): marks [nc]
    int* fofc_mark = nullptr;
    // face-centred CT: face fields of the base and stage-1 containers and the
    // edge EMFs, [3][nf] each
    double *fB0 = nullptr, *fB1 = nullptr, *emf = nullptr;
    // kernel timing (kh_driver_set_kernel_timing)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct Rec { int kind; size_t a, b; };
    std::vector<Rec> recs;

    hipError_t mark(hipEvent_t* out) {
        if (ev_used == ev_pool.size()) {
            hipEvent_t e;
            hipError_t st = hipEventCreate(&e);
            if (st != hipSuccess) return st;
            ev_pool.push_back(e);
        }
        *out = ev_pool[ev_used++];
        return hipEventRecord(*out, blk->stream);
    }
};

namespace kh {
// One block of a uniform mesh of nb[0] x nb[1] x nb[2] blocks (kh_capi.cpp).
int block_create_in_mesh(const kh_params* mesh, const int nb[3], const int bidx[3], kh_block** out);

// Restart file (kh_capi.cpp): a fixed header, then P and U as [8][n3][n2][n1]
// doubles over the entire extent of a one-block mesh.
struct RestartHeader {
    char magic[8];          // "KHAMDRST"
    uint32_t version;       // 1
    uint32_t header_bytes;  // sizeof(RestartHeader)
    int32_t n[3];           // entire extents n1, n2, n3 of the arrays
    int32_t ncycle;
    double time, dt;
    kh_params params;       // the run's parameters (the reader checks the mesh against them)
};
// Write / read the clock and the state; the device arrays hold 8 n1 n2 n3 doubles.
// With face-centred B the face fields [3][n3+1][n2+1][n1+1] follow U (fB non-null
// exactly when the run's b_solver is face_ct).
int write_restart(const char* path, const kh_params& p, const int n[3], double time, double dt, int ncycle,
                  const double* P, const double* U, hipStream_t s, const double* fB = nullptr);
int read_restart(const char* path, const kh_params& p, const int n[3], double* time, double* dt, int* ncycle,
                 double* P, double* U, hipStream_t s, double* fB = nullptr);
// Bytes per element of the face-flux arrays (FP32 in the split form).
inline size_t flux_elem(const kh_block* b) { return b->p.precision == KH_PREC_SPLIT ? sizeof(float) : sizeof(double); }
// One RK stage (1 or 2) of the driver's step on its current buffers.
int driver_stage(kh_driver* d, int stage);
// Weight of stage `stage`'s fluxes in the step's update U_n -> U_{n+1}
// (the RK weights: beta_1 gam0_2 for stage 1, beta_2 for stage 2).
inline double stage_flux_weight(const StepConsts& c, int stage) {
    return stage == 1 ? c.beta[0] * c.gam0[1] : c.beta[1];
}
// Checks a face-flux objective's (dir, var, face) against the block.
int check_face_objective(const BlockDims& b, int dir, int var, int face);

// Byte offsets of one step's record (kh_driver_set_stage_record) in its buffer:
// stage 1's output primitives, each stage's U* (curved-space blocks), inversion
// and floor flags, clamped X1 mass fluxes (physical X1 faces) and failed-
// inversion count.  Sections start on 256-byte boundaries.
struct RecordLayout {
    size_t P1, Ustar[2], pflag[2], fflag[2], x1f[2], fails, total;
};
inline RecordLayout record_layout(const BlockDims& b) {
    RecordLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) { const size_t at = o; o += (bytes + 255) / 256 * 256; return at; };
    const size_t sb = sizeof(double) * NV * b.nc;
    L.P1 = take(sb);
    for (int q = 0; q < 2; ++q) L.Ustar[q] = take(b.geo ? sb : 0);
    for (int q = 0; q < 2; ++q) L.pflag[q] = take(sizeof(int) * b.nc);
    for (int q = 0; q < 2; ++q) L.fflag[q] = take(sizeof(int) * b.nc);
    for (int q = 0; q < 2; ++q) L.x1f[q] = take(sizeof(double) * x1_flux_rows(b));
    L.fails = take(2 * sizeof(int));
    L.total = o;
    return L;
}
}  // namespace kh
