// kh_adjoint_capi.cpp -- C ABI of the adjoint step (include/kharma_amd.h).
//
// kh_adjoint_step recomputes the forward step from its checkpointed starting
// primitives on an internal driver (stage 1 -> P1, U1; stage 2 -> P2, U2), then
// runs the stage transposes of kh_adjoint.hip in reverse order:
//   stage 2: lU*_2 = (dU/dP)^-T lP_2;  lU_1 += g0 lU*_2, lU_0 += g1 lU*_2;
//            lP_1 = (dF/dP at P_1)^T (ddudt/dF)^T (beta dt lU*_2)
//   stage 1: lU*_1 = (dU/dP)^-T lP_1 + lU_1;  lU_0 += lU*_1;
//            lP_0 = (dF/dP at P_0)^T (ddudt/dF)^T (beta dt lU*_1)
//   then     lP_0 += (dU/dP at P_0)^T lU_0        (U_0 = p_to_u(P_0))
// with the periodic ghost images folded back onto their interior zones after
// each flux transpose.
#include "kharma_amd.h"

#include <string>

#include "kh_host.hpp"

using namespace kh;

struct kh_adjoint {
    kh_driver* drv = nullptr;
    double* P0 = nullptr;      // copy of the step's starting primitives when stage 2 is recomputed
    double* lPin = nullptr;    // adjoint of the stage-2 input primitives (P1)
    double* lU0 = nullptr;     // adjoint of U_0 (the step's starting conserved state)
    double* lU1 = nullptr;     // adjoint of U_1
    double* ld = nullptr;      // adjoint of dU/dt
    double* lF = nullptr;      // [3][8][nc] adjoint of the face fluxes
    double* lemf = nullptr;    // [3][nc] adjoint of the corner EMFs
    double* lLR = nullptr;     // [2][8][nc] adjoints of the reconstructed face values
    int* err = nullptr;        // [0] flagged zones, [1] singular Jacobians (device)
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float fwd_ms = 0.f, adj_ms = 0.f;
};

namespace {

#define KA_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return api_fail(KH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Transpose of one stage.  Pin, Pout: the stage's input and output primitives.
// lPout holds lam(P_out) (interior); lPin receives lam(P_in) (interior, ghosts
// zero).  lUout may be null (no adjoint on the stage's output conserved
// state).  accumulate = false: lUsub / lUfull are overwritten, not added to.
int stage_transpose(kh_adjoint* a, int stage, const double* Pin, const double* Pout, const double* lPout,
                    const double* lUout, double* lUsub, double* lUfull, bool same, bool accumulate, double dt,
                    double* lPin) {
    kh_block* b = a->drv->blk;
    const BlockDims& bd = b->dims;
    const StepConsts& c = b->consts;
    hipStream_t s = b->stream;
    const double g0 = c.gam0[stage - 1], g1 = c.gam1[stage - 1], beta_dt = c.beta[stage - 1] * dt;
    KA_HIP(launch_adj_stage(bd, c, b->geo, Pout, lPout, lUout, g0, g1, beta_dt, same, accumulate, lUsub, lUfull,
                            a->ld, a->err, s));
    KA_HIP(launch_adj_divct(bd, a->ld, a->lF, a->lemf, s));
    KA_HIP(launch_adj_fill(lPin, (long)NV * bd.nc, 0., s));
    KA_HIP(launch_adj_flux(bd, c, b->geo, Pin, a->lF, a->lLR, lPin, s));
    KA_HIP(launch_adj_fold(bd, lPin, NV, s));
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
    if (!p->periodic1 || (p->nx2 > 1 && !p->periodic2))
        return api_fail(KH_ERR_UNSUPPORTED, "the adjoint needs every axis periodic");
    kh_adjoint* a = new kh_adjoint();
    st = kh_driver_create(p, &a->drv);
    if (st != KH_OK) { delete a; return st; }
    const long nc = a->drv->blk->dims.nc;
    const size_t sb = sizeof(double) * NV * nc;
    hipError_t e = hipSuccess;
    auto alloc = [&](double** q, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(q, bytes);
        if (e == hipSuccess) e = hipMemset(*q, 0, bytes);
    };
    alloc(&a->P0, sb);
    alloc(&a->lPin, sb);
    alloc(&a->lU0, sb);
    alloc(&a->lU1, sb);
    alloc(&a->ld, sb);
    alloc(&a->lF, 3 * sb);
    alloc(&a->lemf, sizeof(double) * 3 * nc);
    alloc(&a->lLR, 2 * sb);
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
    for (double* q : {a->P0, a->lPin, a->lU0, a->lU1, a->ld, a->lF, a->lemf, a->lLR})
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

int kh_adjoint_step(kh_adjoint* a, const double* P, const double* P_next, double dt, double* lam_P) {
    if (!a || !P || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    if (!(dt > 0.)) return api_fail(KH_ERR_INVALID, "dt must be positive");
    kh_driver* d = a->drv;
    kh_block* b = d->blk;
    const BlockDims& bd = b->dims;
    const StepConsts& c = b->consts;
    hipStream_t s = b->stream;
    const size_t sb = sizeof(double) * NV * bd.nc;
    KA_HIP(hipMemsetAsync(a->err, 0, 2 * sizeof(int), s));
    KA_HIP(hipEventRecord(a->ev[0], s));
    // ---- forward recompute from the checkpoint: P_0 (ghosts refilled), U_0 = p_to_u(P_0), stage 1
    KA_HIP(hipMemcpyAsync(d->P0, P, sb, hipMemcpyDeviceToDevice, s));
    KA_HIP(launch_sync(bd, d->P0, NV, s));
    KA_HIP(launch_ptou(bd, c, b->geo, d->P0, d->U0, s));
    KA_HIP(launch_reset_state(d->ds, dt, s));
    int st = driver_stage(d, 1);                             // -> P1, U1
    if (st != KH_OK) return st;
    KA_HIP(launch_adj_count_flags(bd, d->pflag, d->fflag, a->err, s));
    const double* P0 = d->P0;
    const double* P2 = P_next;
    if (!P2) {
        // stage 2 writes P2, U2 over the driver's P0, U0: keep P_0
        KA_HIP(hipMemcpyAsync(a->P0, d->P0, sb, hipMemcpyDeviceToDevice, s));
        P0 = a->P0;
        if ((st = driver_stage(d, 2)) != KH_OK) return st;
        KA_HIP(launch_adj_count_flags(bd, d->pflag, d->fflag, a->err, s));
        P2 = d->P0;
    }
    KA_HIP(hipEventRecord(a->ev[1], s));
    // ---- transpose.  stage 2: P1 -> P2 with U_sub = U1, U_full = U0 (writes lU1, lU0)
    if ((st = stage_transpose(a, 2, d->P1, P2, lam_P, nullptr, a->lU1, a->lU0, false, false, dt, a->lPin)) != KH_OK)
        return st;
    // stage 1: P0 -> P1 with U_sub = U_full = U0 (adds to lU0); lam(P_0) lands in lam_P
    if ((st = stage_transpose(a, 1, P0, d->P1, a->lPin, a->lU1, a->lU0, a->lU0, true, true, dt, lam_P)) != KH_OK)
        return st;
    // U_0 = p_to_u(P_0)
    KA_HIP(launch_adj_ptou(bd, c, b->geo, P0, a->lU0, lam_P, s));
    KA_HIP(hipEventRecord(a->ev[2], s));
    int err[2] = {0, 0};
    KA_HIP(hipMemcpyAsync(err, a->err, sizeof(err), hipMemcpyDeviceToHost, s));
    KA_HIP(hipStreamSynchronize(s));
    KA_HIP(hipEventElapsedTime(&a->fwd_ms, a->ev[0], a->ev[1]));
    KA_HIP(hipEventElapsedTime(&a->adj_ms, a->ev[1], a->ev[2]));
    if (err[0] > 0)
        return api_fail(KH_ERR_UNSUPPORTED, std::to_string(err[0]) + " zone-stages were floored or failed to invert; "
                                            "the adjoint covers unfloored, converged steps only");
    if (err[1] > 0)
        return api_fail(KH_ERR_UNSUPPORTED, std::to_string(err[1]) + " zones have a singular dU/dP");
    return KH_OK;
}

int kh_adjoint_flux(kh_adjoint* a, const double* P, const double* lam_F, double* lam_P) {
    if (!a || !P || !lam_F || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    KA_HIP(launch_adj_fill(lam_P, (long)NV * b->dims.nc, 0., b->stream));
    KA_HIP(launch_adj_flux(b->dims, b->consts, b->geo, P, lam_F, a->lLR, lam_P, b->stream));
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
    KA_HIP(hipMemsetAsync(a->err, 0, 2 * sizeof(int), b->stream));
    KA_HIP(launch_adj_stage(b->dims, b->consts, b->geo, P, lam_P, nullptr, 0., 1., 0., true, true, lam_U, lam_U,
                            a->ld, a->err, b->stream));
    return KH_OK;
}

int kh_adjoint_ptou(kh_adjoint* a, const double* P, const double* lam_U, double* lam_P) {
    if (!a || !P || !lam_U || !lam_P) return api_fail(KH_ERR_INVALID, "null argument");
    kh_block* b = a->drv->blk;
    KA_HIP(launch_adj_ptou(b->dims, b->consts, b->geo, P, lam_U, lam_P, b->stream));
    return KH_OK;
}

int kh_adjoint_last_times(kh_adjoint* a, double* forward_ms, double* adjoint_ms) {
    if (!a) return api_fail(KH_ERR_INVALID, "null adjoint");
    if (forward_ms) *forward_ms = a->fwd_ms;
    if (adjoint_ms) *adjoint_ms = a->adj_ms;
    return KH_OK;
}

}  // extern "C"
