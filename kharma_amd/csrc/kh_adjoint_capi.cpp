// kh_adjoint_capi.cpp -- C ABI of the adjoint step (include/kharma_amd.h).
//
// kh_adjoint_step recomputes the forward step from its checkpointed starting
// primitives on an internal driver (stage 1 -> P1, U1; stage 2 -> P2, U2 of
// the adjoint's own buffers),
// keeping what the transposes need: each stage's U* (the update before
// recovery) and floor flags on the curved-space path, and the clamped X1 mass
// fluxes where X1 is physical.  Then it runs the stage transposes of
// kh_adjoint.hip in reverse order:
//   stage 2: lU*_2 = (dU/dP)^-T (dfloors/dP)^T (...) lP_2;  lU_1 += g0 lU*_2, lU_0 += g1 lU*_2;
//            lP_1 = (dF/dP at P_1)^T (dfix/dF)^T (ddudt/dF)^T (beta dt lU*_2) + (dS/dP_1)^T beta dt lU*_2
//   stage 1: likewise from lP_1 and lU_1 into lP_0, lU_0
//   then     lP_0 += (dU/dP at P_0)^T lU_0        (U_0 = p_to_u(P_0))
// with the ghost zones' adjoints mapped back through the physical-boundary
// fill and the periodic images after each flux transpose, and a face-flux
// objective's seed added to the flux adjoints of each stage.
#include "kharma_amd.h"

#include <cstddef>
#include <cstdint>
#include <string>

#include "kh_host.hpp"

using namespace kh;

struct kh_adjoint {
    kh_driver* drv = nullptr;
    double* P2 = nullptr;      // stage 2's recomputed primitives and conserved state (the driver's
    double* U2 = nullptr;      // P0, U0 keep the step's starting state)
    double* lPin = nullptr;    // adjoint of the stage-2 input primitives (P1)
    double* lU0 = nullptr;     // adjoint of U_0 (the step's starting conserved state)
    double* lU1 = nullptr;     // adjoint of U_1
    double* ld = nullptr;      // adjoint of dU/dt
    double* lF = nullptr;      // [3][8][nc] adjoint of the face fluxes
    double* lemf = nullptr;    // [3][nc] adjoint of the corner EMFs
    int* err = nullptr;        // [0] flagged zones, [1] singular Jacobians (device)
    double* Ustar[2] = {nullptr, nullptr};   // U* of each stage (curved-space path)
    int* ffs[2] = {nullptr, nullptr};        // floor flags of each stage (curved-space path with floors)
    double* x1f[2] = {nullptr, nullptr};     // clamped X1 mass fluxes of each stage (physical X1)
    int* pfs[2] = {nullptr, nullptr};        // inversion flags of each stage (with the fixup)
    double* lPfx = nullptr;                  // [8][nc] fixup transpose
    double* Pinv = nullptr;                  // [8][nc] recovered primitives of floored zones
    double* lPinv = nullptr;                 // [8][nc] their adjoints through the floors
    double* lPtot = nullptr;                 // [8][nc] their output adjoints, P->U included
    double* gC = nullptr;                    // [8][nc] ghost-fill transpose scratch (physical boundaries)
    bool geo = false, physical = false, floors = false, fixup = false;
    int obj_dir = 0, obj_var = 0, obj_face = 0;   // face-flux objective (kh_adjoint_set_face_flux_objective)
    double obj_weight = 0.;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float fwd_ms = 0.f, adj_ms = 0.f;
    bool deferred = false;   // kh_adjoint_set_deferred_checks: counts accumulate, no per-step sync
};

namespace {

#define KA_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return api_fail(KH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// One forward stage on the driver's buffers, as kh::driver_stage runs it,
// keeping U* (between the update's combine and recovery halves), the floor
// flags and the clamped X1 mass fluxes of the stage.
int recompute_stage(kh_adjoint* a, int stage) {
    kh_driver* d = a->drv;
    kh_block* b = d->blk;
    const BlockDims& bd = b->dims;
    const StepConsts& c = b->consts;
    hipStream_t s = b->stream;
    const double* Psub = (stage == 1) ? d->P0 : d->P1;
    const double* Usub = (stage == 1) ? d->U0 : d->U1;
    double* Pout = (stage == 1) ? d->P1 : a->P2;
    double* Uout = (stage == 1) ? d->U1 : a->U2;
    // the stage's flags go straight to the buffers the transpose reads
    int* pflag = a->fixup ? a->pfs[stage - 1] : d->pflag;
    int* fflag = a->floors ? a->ffs[stage - 1] : d->fflag;
    double* F = (double*)d->F;
    double* ct = (double*)d->ctop;
    for (int dir = 1; dir <= bd.ndim; ++dir)
        KA_HIP(launch_flux_stage(bd, c, b->geo, dir, b->p.recon, b->p.precision, SPEEDS_NONE, Psub, nullptr,
                                 F + (long)(dir - 1) * NV * bd.nc, ct + (long)(dir - 1) * bd.nc, s));
    if (a->physical) {
        KA_HIP(launch_fix_flux(bd, b->p.precision, F, s));
        KA_HIP(launch_save_x1_flux(bd, F, a->x1f[stage - 1], s));
    }
    if (a->geo) {
        KA_HIP(launch_update_parts(bd, c, b->geo, stage, 3, F, Usub, d->U0, Psub, Uout, Pout, pflag, fflag,
                                   d->fail_list, d->ds, s, a->Ustar[stage - 1]));
    } else {
        KA_HIP(launch_update(bd, c, b->geo, b->p.precision, stage, false, F, ct, Usub, d->U0, Psub, Uout, Pout,
                             nullptr, pflag, fflag, d->fail_list, d->ds, d->dt_part, s));
    }
    if (a->fixup) {
        KA_HIP(launch_fixup(bd, c, b->geo, stage, Uout, Pout, nullptr, pflag, d->fail_list, d->ds, s));
        if (!a->floors) KA_HIP(launch_adj_count_flags(bd, fflag, nullptr, a->err, s));
    } else {
        KA_HIP(launch_adj_count_flags(bd, pflag, a->floors ? nullptr : fflag, a->err, s));
    }
    if (a->physical) KA_HIP(launch_bounds(bd, c, b->geo, Uout, Pout, nullptr, s));
    return KH_OK;
}

// Transpose of one stage.  Pin, Pout: the stage's input and output primitives.
// lPout holds lam(P_out) (interior); lPin receives lam(P_in) (interior, ghosts
// zero).  lUout may be null (no adjoint on the stage's output conserved
// state).  accumulate = false: lUsub / lUfull are overwritten, not added to.
// clear_ghosts = false leaves ghost entries of lPin stale (interior consumers).
int stage_transpose(kh_adjoint* a, int stage, const double* Pin, const double* Pout, const double* lPout,
                    const double* lUout, double* lUsub, double* lUfull, bool same, bool accumulate, double dt,
                    double* lPin, bool clear_ghosts, bool fixup) {
    kh_block* b = a->drv->blk;
    const BlockDims& bd = b->dims;
    const StepConsts& c = b->consts;
    hipStream_t s = b->stream;
    const double g0 = c.gam0[stage - 1], g1 = c.gam1[stage - 1], beta_dt = c.beta[stage - 1] * dt;
    // the stage's count of failed inversions (device), which gates the fixup transpose
    const int* gate = reinterpret_cast<const int*>(reinterpret_cast<const char*>(a->drv->ds) +
                                                   offsetof(DeviceState, fail_count)) + (stage - 1);
    if (fixup) {
        KA_HIP(launch_adj_fixup(bd, c, b->geo, Pout, lPout, lUout, a->pfs[stage - 1], gate, a->lPfx, s));
        KA_HIP(launch_adj_fold_interior(bd, a->lPfx, NV, gate, s));
    }
    KA_HIP(launch_adj_fill(lPin, (long)NV * bd.nc, 0., s));
    KA_HIP(launch_adj_stage(bd, c, b->geo, Pout, lPout, lUout, g0, g1, beta_dt, same, accumulate, lUsub, lUfull,
                            a->ld, a->err, a->geo ? a->Ustar[stage - 1] : nullptr, Pin,
                            a->floors ? a->ffs[stage - 1] : nullptr, a->geo ? lPin : nullptr,
                            fixup ? a->pfs[stage - 1] : nullptr, fixup ? gate : nullptr,
                            fixup ? a->lPfx : nullptr, a->Pinv, a->lPinv, a->lPtot, s));
    FluxSeed seed;
    if (a->obj_dir) {
        const int dir = a->obj_dir;
        const double dA = (dir == 1) ? bd.dx2 * bd.dx3 : (dir == 2) ? bd.dx1 * bd.dx3 : bd.dx1 * bd.dx2;
        seed.dir = dir; seed.var = a->obj_var; seed.face = a->obj_face;
        seed.w = a->obj_weight * stage_flux_weight(c, stage) * dt * dA;
    }
    KA_HIP(launch_adj_divct(bd, a->ld, a->lF, a->lemf, s, seed));
    if (a->physical) KA_HIP(launch_adj_fix_flux(bd, a->lF, a->x1f[stage - 1], s));
    KA_HIP(launch_adj_flux(bd, c, b->geo, Pin, a->lF, lPin, s));
    if (a->physical) KA_HIP(launch_adj_bounds(bd, b->geo, Pin, lPin, a->gC, s));
    if (clear_ghosts) {
        KA_HIP(launch_adj_fold(bd, lPin, NV, s));
    } else {
        // the next consumer reads the interior only
        KA_HIP(launch_adj_fold_interior(bd, lPin, NV, nullptr, s));
    }
    return KH_OK;
}

}  // namespace

extern "C" {

int kh_adjoint_create(const kh_params* p, kh_adjoint** out) {
    if (!out) return api_fail(KH_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int st = check_params(p);
    if (st != KH_OK) return st;
    if (p->precision != KH_PREC_FP64) return api_fail(KH_ERR_UNSUPPORTED, "the adjoint runs in the FP64 form");
    if (p->recon != KH_RECON_WENO5) return api_fail(KH_ERR_UNSUPPORTED, "the adjoint supports WENO5 reconstruction");
    if (p->use_hlle) return api_fail(KH_ERR_UNSUPPORTED, "the adjoint supports the LLF flux");
    if (p->fofc) return api_fail(KH_ERR_UNSUPPORTED, "the adjoint does not differentiate the first-order flux correction");
    if (p->b_solver == KH_BSOLVER_FACE_CT) return api_fail(KH_ERR_UNSUPPORTED, "the adjoint differentiates flux-CT only");
    if (p->inverter != KH_INVERTER_ONEDW)
        return api_fail(KH_ERR_UNSUPPORTED, "the adjoint supports the onedw inverter (Kastaun's inverter floors "
                                            "act inside the recovery, outside the implicit-function transpose)");
    if (p->floors_on && p->floor_frame == 1)
        return api_fail(KH_ERR_UNSUPPORTED, "the adjoint supports the fluid and drift floor frames");
    kh_adjoint* a = new kh_adjoint();
    st = kh_driver_create(p, &a->drv);
    if (st != KH_OK) { delete a; return st; }
    const BlockDims& bd = a->drv->blk->dims;
    const long nc = bd.nc;
    a->geo = bd.geo != 0;
    a->physical = !bd.periodic[0] || (bd.ndim > 1 && !bd.periodic[1]);
    a->floors = a->geo && a->drv->blk->consts.floors.on;
    a->fixup = a->drv->blk->consts.fix_average || a->drv->blk->consts.fix_atmo;
    const size_t sb = sizeof(double) * NV * nc;
    hipError_t e = hipSuccess;
    auto alloc = [&](double** q, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(q, bytes);
        if (e == hipSuccess) e = hipMemset(*q, 0, bytes);
    };
    alloc(&a->P2, sb);
    alloc(&a->U2, sb);
    alloc(&a->lPin, sb);
    alloc(&a->lU0, sb);
    alloc(&a->lU1, sb);
    alloc(&a->ld, sb);
    alloc(&a->lF, 3 * sb);
    alloc(&a->lemf, sizeof(double) * 3 * nc);
    for (int q = 0; q < 2; ++q) {
        if (a->geo) alloc(&a->Ustar[q], sb);
        if (a->floors && e == hipSuccess) e = hipMalloc(&a->ffs[q], sizeof(int) * nc);
        if (a->physical && x1_flux_rows(bd) > 0) alloc(&a->x1f[q], sizeof(double) * x1_flux_rows(bd));
        if (a->fixup && e == hipSuccess) e = hipMalloc(&a->pfs[q], sizeof(int) * nc);
        if (a->fixup && e == hipSuccess) e = hipMemset(a->pfs[q], 0, sizeof(int) * nc);
    }
    if (a->fixup) alloc(&a->lPfx, sb);
    if (a->physical) alloc(&a->gC, sb);
    if (a->floors) {
        alloc(&a->Pinv, sb);
        alloc(&a->lPinv, sb);
        alloc(&a->lPtot, sb);
    }
    if (e == hipSuccess) e = hipMalloc(&a->err, 2 * sizeof(int));
    for (int q = 0; q < 3 && e == hipSuccess; ++q) e = hipEventCreate(&a->ev[q]);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        kh_adjoint_destroy(a);
        return api_fail(KH_ERR_HIP, std::string("adjoint allocation: ") + hipGetErrorString(e));
    }
    *out = a;
    return KH_OK;
}

void kh_adjoint_destroy(kh_adjoint* a) {
    if (!a) return;
    if (a->drv) (void)hipStreamSynchronize(a->drv->blk->stream);
    for (double* q : {a->P2, a->U2, a->lPin, a->lU0, a->lU1, a->ld, a->lF, a->lemf, a->Ustar[0], a->Ustar[1],
                      a->x1f[0], a->x1f[1], a->lPfx, a->Pinv, a->lPinv, a->lPtot, a->gC})
        if (q) (void)hipFree(q);
    for (int* q : {a->ffs[0], a->ffs[1], a->pfs[0], a->pfs[1]})
        if (q) (void)hipFree(q);
    if (a->err) (void)hipFree(a->err);
    for (hipEvent_t ev : a->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (a->drv) kh_driver_destroy(a->drv);
    delete a;
}

int kh_adjoint_set_stream(kh_adjoint* a, void* stream) {
    if (!a) return api_fail(KH_ERR_INVALID, "null adjoint");
    return kh_driver_set_stream(a->drv, stream);
}

// The device error counts (after a synchronisation) as a status.
static int report_errors(kh_adjoint* a, const int err[2]) {
    if (err[0] > 0)
        return api_fail(KH_ERR_UNSUPPORTED,
                        std::to_string(err[0]) + " zone-stages " +
                            (a->floors ? "failed to invert with the fixup off" : a->fixup ? "were floored" :
                             "were floored or failed to invert with the fixup off") +
                            "; the adjoint does not cover them on this block");
    if (err[1] > 0)
        return api_fail(KH_ERR_UNSUPPORTED, std::to_string(err[1]) + " zones have a singular dU/dP");
    return KH_OK;
}

// Start of an adjoint step: clear the error counts unless they accumulate.
static int begin_step(kh_adjoint* a) {
    if (!a->deferred) KA_HIP(hipMemsetAsync(a->err, 0, 2 * sizeof(int), a->drv->blk->stream));
    return KH_OK;
}

// End of an adjoint step: timing marks, then the device error counts (deferred
// mode: left on the device for kh_adjoint_check).
static int finish_step(kh_adjoint* a) {
    hipStream_t s = a->drv->blk->stream;
    KA_HIP(hipEventRecord(a->ev[2], s));
    if (a->deferred) return KH_OK;
    int err[2] = {0, 0};
    KA_HIP(hipMemcpyAsync(err, a->err, sizeof(err), hipMemcpyDeviceToHost, s));
    KA_HIP(hipStreamSynchronize(s));
    KA_HIP(hipEventElapsedTime(&a->fwd_ms, a->ev[0], a->ev[1]));
    KA_HIP(hipEventElapsedTime(&a->adj_ms, a->ev[1], a->ev[2]));
    return report_errors(a, err);
}

int kh_adjoint_set_deferred_checks(kh_adjoint* a, int on) {
    if (!a) return api_fail(KH_ERR_INVALID, "null adjoint");
    a->deferred = on != 0;
    KA_HIP(hipMemsetAsync(a->err, 0, 2 * sizeof(int), a->drv->blk->stream));
    return KH_OK;
}

int kh_adjoint_check(kh_adjoint* a) {
    if (!a) return api_fail(KH_ERR_INVALID, "null adjoint");
    hipStream_t s = a->drv->blk->stream;
    int err[2] = {0, 0};
    KA_HIP(hipMemcpyAsync(err, a->err, sizeof(err), hipMemcpyDeviceToHost, s));
    KA_HIP(hipStreamSynchronize(s));
    KA_HIP(hipMemsetAsync(a->err, 0, 2 * sizeof(int), s));
    return report_errors(a, err);
}

int kh_adjoint_step(kh_adjoint* a, const double* P, const double* P_next, double dt, double* lam_P) {
    if (!a || !P || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    if (!(dt > 0.)) return api_fail(KH_ERR_INVALID, "dt must be positive");
    kh_driver* d = a->drv;
    kh_block* b = d->blk;
    const BlockDims& bd = b->dims;
    const StepConsts& c = b->consts;
    hipStream_t s = b->stream;
    const size_t sb = sizeof(double) * NV * bd.nc;
    if (const int st0 = begin_step(a)) return st0;
    KA_HIP(hipEventRecord(a->ev[0], s));
    // ---- forward recompute from the checkpoint: P_0 (ghosts as checkpointed; periodic
    // images refilled), U_0 = p_to_u(P_0), stage 1
    KA_HIP(hipMemcpyAsync(d->P0, P, sb, hipMemcpyDeviceToDevice, s));
    if (bd.periodic[0] || (bd.ndim > 1 && bd.periodic[1]) || (bd.ndim > 2 && bd.periodic[2]))
        KA_HIP(launch_sync(bd, d->P0, NV, s));
    KA_HIP(launch_ptou(bd, c, b->geo, d->P0, d->U0, s));
    KA_HIP(launch_reset_state(d->ds, dt, s));
    int st = recompute_stage(a, 1);                          // -> P1, U1
    if (st != KH_OK) return st;
    const double* P0 = d->P0;
    const double* P2 = P_next;
    // the curved-space and physical-boundary transposes need stage 2's U*, flags
    // and clamped fluxes: recompute it unless a plain periodic step can use P_next
    // stage 2 taken from P_next: the caller vouches that no zone of it failed, so
    // its transpose has no fixup to undo
    const bool recompute2 = !P2 || a->geo || a->physical;
    if (recompute2) {
        if ((st = recompute_stage(a, 2)) != KH_OK) return st;   // -> a->P2, a->U2
        P2 = a->P2;
    }
    KA_HIP(hipEventRecord(a->ev[1], s));
    // ---- transpose.  stage 2: P1 -> P2 with U_sub = U1, U_full = U0 (writes lU1, lU0)
    if ((st = stage_transpose(a, 2, d->P1, P2, lam_P, nullptr, a->lU1, a->lU0, false, false, dt, a->lPin, false,
                              a->fixup && recompute2)) != KH_OK)
        return st;
    // stage 1: P0 -> P1 with U_sub = U_full = U0 (adds to lU0); lam(P_0) lands in lam_P
    if ((st = stage_transpose(a, 1, P0, d->P1, a->lPin, a->lU1, a->lU0, a->lU0, true, true, dt, lam_P, true,
                              a->fixup)) != KH_OK)
        return st;
    // U_0 = p_to_u(P_0)
    KA_HIP(launch_adj_ptou(bd, c, b->geo, P0, a->lU0, lam_P, s));
    return finish_step(a);
}

int kh_adjoint_step_recorded(kh_adjoint* a, const double* P, const void* rec, const double* P_next, double dt,
                             double* lam_P) {
    if (!a || !P || !rec || !P_next || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    if (!(dt > 0.)) return api_fail(KH_ERR_INVALID, "dt must be positive");
    if (reinterpret_cast<uintptr_t>(rec) & 255) return api_fail(KH_ERR_INVALID, "a step record must be 256-byte aligned");
    kh_driver* d = a->drv;
    kh_block* b = d->blk;
    const BlockDims& bd = b->dims;
    const StepConsts& c = b->consts;
    hipStream_t s = b->stream;
    const RecordLayout L = record_layout(bd);
    const char* r = static_cast<const char*>(rec);
    auto at_d = [&](size_t off) { return reinterpret_cast<double*>(const_cast<char*>(r + off)); };
    auto at_i = [&](size_t off) { return reinterpret_cast<int*>(const_cast<char*>(r + off)); };
    if (const int st0 = begin_step(a)) return st0;
    KA_HIP(hipEventRecord(a->ev[0], s));
    // the stages' failed-inversion counts (the fix-up transposes' gates) and
    // the count of zones the adjoint does not cover, by the rule of the
    // recompute; P is read as the driver holds its state, ghosts filled
    KA_HIP(hipMemcpyAsync(reinterpret_cast<char*>(d->ds) + offsetof(DeviceState, fail_count), r + L.fails,
                          2 * sizeof(int), hipMemcpyDeviceToDevice, s));
    for (int q = 0; q < 2; ++q) {
        if (a->fixup) {
            if (!a->floors) KA_HIP(launch_adj_count_flags(bd, at_i(L.fflag[q]), nullptr, a->err, s));
        } else {
            KA_HIP(launch_adj_count_flags(bd, at_i(L.pflag[q]), a->floors ? nullptr : at_i(L.fflag[q]), a->err, s));
        }
    }
    KA_HIP(hipEventRecord(a->ev[1], s));
    // the transposes read the recorded intermediates in place of the recompute's
    struct Swap {
        kh_adjoint* a;
        double* Ustar[2]; int* pfs[2]; int* ffs[2]; double* x1f[2];
        ~Swap() {
            for (int q = 0; q < 2; ++q) { a->Ustar[q] = Ustar[q]; a->pfs[q] = pfs[q]; a->ffs[q] = ffs[q]; a->x1f[q] = x1f[q]; }
        }
    } keep{a, {a->Ustar[0], a->Ustar[1]}, {a->pfs[0], a->pfs[1]}, {a->ffs[0], a->ffs[1]}, {a->x1f[0], a->x1f[1]}};
    for (int q = 0; q < 2; ++q) {
        if (a->geo) a->Ustar[q] = at_d(L.Ustar[q]);
        if (a->fixup) a->pfs[q] = at_i(L.pflag[q]);
        if (a->floors) a->ffs[q] = at_i(L.fflag[q]);
        if (a->physical) a->x1f[q] = at_d(L.x1f[q]);
    }
    const double* P1 = at_d(L.P1);
    int st;
    if ((st = stage_transpose(a, 2, P1, P_next, lam_P, nullptr, a->lU1, a->lU0, false, false, dt, a->lPin, false,
                              a->fixup)) != KH_OK)
        return st;
    if ((st = stage_transpose(a, 1, P, P1, a->lPin, a->lU1, a->lU0, a->lU0, true, true, dt, lam_P, true,
                              a->fixup)) != KH_OK)
        return st;
    KA_HIP(launch_adj_ptou(bd, c, b->geo, P, a->lU0, lam_P, s));
    return finish_step(a);
}

int kh_adjoint_flux(kh_adjoint* a, const double* P, const double* lam_F, double* lam_P) {
    if (!a || !P || !lam_F || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    KA_HIP(launch_adj_fill(lam_P, (long)NV * b->dims.nc, 0., b->stream));
    KA_HIP(launch_adj_flux(b->dims, b->consts, b->geo, P, lam_F, lam_P, b->stream));
    return KH_OK;
}

int kh_adjoint_flux_div(kh_adjoint* a, const double* lam_dUdt, double* lam_F) {
    if (!a || !lam_dUdt || !lam_F) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    KA_HIP(launch_adj_divct(b->dims, lam_dUdt, lam_F, a->lemf, b->stream));
    return KH_OK;
}

int kh_adjoint_utop(kh_adjoint* a, const double* P, const double* lam_P, double* lam_U) {
    if (!a || !P || !lam_P || !lam_U) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    // lam_U* with no base-state or dU/dt consumers: the stage transpose with
    // g0 = 0, g1 = 1, same = true accumulates lam_U* into lam_U (cleared first).
    KA_HIP(launch_adj_fill(lam_U, (long)NV * b->dims.nc, 0., b->stream));
    if (const int st0 = begin_step(a)) return st0;
    KA_HIP(launch_adj_stage(b->dims, b->consts, b->geo, P, lam_P, nullptr, 0., 1., 0., true, true, lam_U, lam_U,
                            a->ld, a->err, nullptr, P, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, b->stream));
    return KH_OK;
}

int kh_adjoint_ptou(kh_adjoint* a, const double* P, const double* lam_U, double* lam_P) {
    if (!a || !P || !lam_U || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    KA_HIP(launch_adj_ptou(b->dims, b->consts, b->geo, P, lam_U, lam_P, b->stream));
    return KH_OK;
}

int kh_adjoint_set_face_flux_objective(kh_adjoint* a, int dir, int var, int face, double weight) {
    if (!a) return api_fail(KH_ERR_INVALID, "null adjoint");
    if (dir == 0) { a->obj_dir = 0; return KH_OK; }
    const int st = check_face_objective(a->drv->blk->dims, dir, var, face);
    if (st != KH_OK) return st;
    a->obj_dir = dir; a->obj_var = var; a->obj_face = face; a->obj_weight = weight;
    return KH_OK;
}

int kh_adjoint_bounds(kh_adjoint* a, const double* P, double* lam_P) {
    if (!a || !P || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    if (a->gC) KA_HIP(launch_adj_bounds(b->dims, b->geo, P, lam_P, a->gC, b->stream));
    KA_HIP(launch_adj_fold(b->dims, lam_P, NV, b->stream));
    return KH_OK;
}

int kh_adjoint_fix_flux(kh_adjoint* a, const double* F_fixed, double* lam_F) {
    if (!a || !F_fixed || !lam_F) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    double* x1f = a->x1f[0];
    if (x1_flux_rows(b->dims) > 0) {
        if (!x1f) return api_fail(KH_ERR_INVALID, "no X1 flux rows on this block");
        KA_HIP(launch_save_x1_flux(b->dims, F_fixed, x1f, b->stream));
    }
    KA_HIP(launch_adj_fix_flux(b->dims, lam_F, x1f, b->stream));
    return KH_OK;
}

int kh_adjoint_geo_source(kh_adjoint* a, const double* P, const double* lam_dUdt, double* lam_P) {
    if (!a || !P || !lam_dUdt || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    KA_HIP(launch_adj_geo(b->dims, b->consts, b->geo, P, lam_dUdt, lam_P, b->stream));
    return KH_OK;
}

int kh_adjoint_floors(kh_adjoint* a, const double* P_in, const double* lam_P_out, double* lam_P_in) {
    if (!a || !P_in || !lam_P_out || !lam_P_in) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    if (!b->consts.floors.on) return api_fail(KH_ERR_INVALID, "floors are off on this block");
    KA_HIP(launch_adj_floors(b->dims, b->consts, b->geo, P_in, lam_P_out, lam_P_in, b->stream));
    return KH_OK;
}

int kh_adjoint_fixup(kh_adjoint* a, const double* P_fixed, const int* pflag, double* lam_P) {
    if (!a || !P_fixed || !pflag || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    if (!a->fixup) return api_fail(KH_ERR_INVALID, "the fixup is off on this block");
    kh_block* b = a->drv->blk;
    KA_HIP(launch_adj_fixup(b->dims, b->consts, b->geo, P_fixed, lam_P, nullptr, pflag, nullptr, a->lPfx, b->stream));
    KA_HIP(launch_adj_fold(b->dims, a->lPfx, NV, b->stream));
    KA_HIP(launch_adj_fixup_apply(b->dims, pflag, a->lPfx, lam_P, b->stream));
    return KH_OK;
}

int kh_adjoint_last_times(kh_adjoint* a, double* forward_ms, double* adjoint_ms) {
    if (!a) return api_fail(KH_ERR_INVALID, "null adjoint");
    if (forward_ms) *forward_ms = a->fwd_ms;
    if (adjoint_ms) *adjoint_ms = a->adj_ms;
    return KH_OK;
}

}  // extern "C"
