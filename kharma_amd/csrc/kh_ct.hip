// kh_ct.hip -- face-centred constrained transport (KHARMA's B_CT package,
// b_ct/b_ct.cpp): the magnetic field lives on cell faces as sqrt(-g) B^d
// (cons.fB); each stage its edge EMFs are formed from the Riemann B fluxes
// (BS99 average, with the GS05 cell-centre corrections), every face is updated
// by the circulation of its edges, and the cell-centred B the fluid sees is the
// average of a cell's two faces.
//
// Face and edge arrays are [3][n3+1][n2+1][n1+1] (F1/F2/F3, E1/E2/E3) with
// index f = k fs3 + j fs2 + i for the lower face / edge of cell (k, j, i):
// F1 at x1 = face i, E3 along x3 at the (i, j) corner, and so on.  The
// element sizes are Parthenon's UniformCartesian ones: an edge along x_d is dx_d
// long and a face normal to x_d has area dx_e dx_f, wherever it lies.
#include "kh_internal.hpp"

namespace kh {

static constexpr int CBS = 256;

__device__ __forceinline__ double ct_block_max(double v) {
    __shared__ double red[CBS / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int q = 1; q < CBS / 64; ++q) m = fmax(m, red[q]);
    return m;
}

// Cell-centre EMF -(v x B) sqrt(-g) over the entire block, b_ct.cpp:318-335.
__global__ void __launch_bounds__(CBS) ct_cemf_kernel(BlockDims b, Vars<const double> P, double* __restrict__ cemf) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc) return;
    double Pc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Pc[v] = P[v][c];
    MinkowskiGeom<double> g;
    FourVec<double> D;
    calc_4vecs<double>(g, Pc, D);
    const double gdet = g.gdet();
    cemf[c] = (D.bcon[2] * D.ucon[3] - D.bcon[3] * D.ucon[2]) * gdet;
    cemf[b.nc + c] = (D.bcon[3] * D.ucon[1] - D.bcon[1] * D.ucon[3]) * gdet;
    cemf[2 * b.nc + c] = (D.bcon[1] * D.ucon[2] - D.bcon[2] * D.ucon[1]) * gdet;
}

// Edge EMFs, B_CT::CalculateEMF (b_ct.cpp:263-416), one thread per edge
// position (k, j, i) of GetRange(interior, 0, 1): i in [is, ie+1], j in
// [js, je+1], k in [ks, ke+1] (k = ks in 2-D), all three components.
// SCHEME: 0 bs99, 1 gs05_0, 2 gs05_c / sg07.  Fluxes are read in their own
// precision and the EMFs formed in FP64.
template <typename R, int SCHEME, int ND>
__global__ void __launch_bounds__(CBS) ct_emf_kernel(BlockDims b, FluxVars<const R> F, const double* __restrict__ cemf,
                                                      double* __restrict__ emf) {
    const int ni = b.nx1 + 1, nj = b.nx2 + 1;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;   // edge / face counts stay below 2^31
    const unsigned ne = (unsigned)ni * (unsigned)nj * (unsigned)(ND > 2 ? b.nx3 + 1 : 1);
    if (t >= ne) return;
    const int i = b.is + (int)(t % (unsigned)ni);
    const int j = b.js + (int)((t / (unsigned)ni) % (unsigned)nj);
    const int k = b.ks + (int)(t / ((unsigned)ni * (unsigned)nj));
    const long s2 = b.s2, s3 = b.s3, nc = b.nc;
    const long c = (long)k * s3 + (long)j * s2 + i;
    auto f = [&](int dir, int v, long cc) -> double { return (double)F[dir - 1][v][cc]; };
    auto C = [&](int v, long cc) -> double { return cemf[(long)v * nc + cc]; };
    double e1, e2, e3;
    if (ND > 2) {
        e1 = 0.25 * (-f(2, B3, c - s3) - f(2, B3, c) + f(3, B2, c - s2) + f(3, B2, c));
        e2 = 0.25 * (-f(3, B1, c - 1) - f(3, B1, c) + f(1, B3, c - s3) + f(1, B3, c));
    } else {
        e1 = -f(2, B3, c);
        e2 = f(1, B3, c);
    }
    e3 = 0.25 * (-f(1, B2, c - s2) - f(1, B2, c) + f(2, B1, c - 1) + f(2, B1, c));
    if constexpr (SCHEME == 1) {
        // gs05_0, b_ct.cpp:337-356 (its E1 average repeats (k, j-1) and omits (k-1, j))
        if (ND > 2) {
            e1 = 2 * e1 - 0.25 * (C(0, c) + C(0, c - s2) + C(0, c - s2) + C(0, c - s3 - s2));
            e2 = 2 * e2 - 0.25 * (C(1, c) + C(1, c - 1) + C(1, c - s3) + C(1, c - s3 - 1));
        }
        e3 = 2 * e3 - 0.25 * (C(2, c) + C(2, c - 1) + C(2, c - s2) + C(2, c - s2 - 1));
    } else if constexpr (SCHEME == 2) {
        // gs05_c / sg07, b_ct.cpp:357-412: upwinded by the sign of the mass flux
        if (ND > 2) {
            const double l3 = (f(2, RHO, c - s3) >= 0.0) ? f(3, B2, c - s2) - C(0, c - s3 - s2) : f(3, B2, c) - C(0, c - s3);
            const double r3 = (f(2, RHO, c) >= 0.0) ? f(3, B2, c - s2) - C(0, c - s2) : f(3, B2, c) - C(0, c);
            const double l2 = (f(3, RHO, c - s2) >= 0.0) ? -f(2, B3, c - s3) - C(0, c - s3 - s2) : -f(2, B3, c) - C(0, c - s2);
            const double r2 = (f(3, RHO, c) >= 0.0) ? -f(2, B3, c - s3) - C(0, c - s3) : -f(2, B3, c) - C(0, c);
            e1 += 0.25 * (l3 + r3 + l2 + r2);
            const double m3 = (f(1, RHO, c - s3) >= 0.0) ? -f(3, B1, c - 1) - C(1, c - s3 - 1) : -f(3, B1, c) - C(1, c - s3);
            const double q3 = (f(1, RHO, c) >= 0.0) ? -f(3, B1, c - 1) - C(1, c - 1) : -f(3, B1, c) - C(1, c);
            const double m1 = (f(3, RHO, c - 1) >= 0.0) ? f(1, B3, c - s3) - C(1, c - s3 - 1) : f(1, B3, c) - C(1, c - 1);
            const double q1 = (f(3, RHO, c) >= 0.0) ? f(1, B3, c - s3) - C(1, c - s3) : f(1, B3, c) - C(1, c);
            e2 += 0.25 * (m3 + q3 + m1 + q1);
        }
        const double l2 = (f(1, RHO, c - s2) >= 0.0) ? f(2, B1, c - 1) - C(2, c - s2 - 1) : f(2, B1, c) - C(2, c - s2);
        const double r2 = (f(1, RHO, c) >= 0.0) ? f(2, B1, c - 1) - C(2, c - 1) : f(2, B1, c) - C(2, c);
        const double l1 = (f(2, RHO, c - 1) >= 0.0) ? -f(1, B2, c - s2) - C(2, c - s2 - 1) : -f(1, B2, c) - C(2, c - 1);
        const double r1 = (f(2, RHO, c) >= 0.0) ? -f(1, B2, c - s2) - C(2, c - s2) : -f(1, B2, c) - C(2, c);
        e3 += 0.25 * (l2 + r2 + l1 + r1);
    }
    const long fi = (long)k * b.fs3 + (long)j * b.fs2 + i;
    emf[fi] = e1;
    emf[b.nf + fi] = e2;
    emf[2 * b.nf + fi] = e3;
}

// Face update from the edge circulation, B_CT::AddSource (b_ct.cpp:418-470),
// one thread per face position of the union of the three interior face ranges:
// F1 i in [is, ie+1], F2 j in [js, je+1], F3 k in [ks, ke+1] (k = ks in 2-D),
// the other two indices over the interior.  OUT_RATE writes dB_Uf/dt; otherwise
// the face half of AddStateUpdate (kharma_driver.cpp:353-373) follows in place:
// out = 1 (g0 sub + g1 full) + beta dt dB/dt (sub read once when SAME: 0 x + 1 y == y).
template <int ND, bool OUT_RATE, bool SAME>
__global__ void __launch_bounds__(CBS)
ct_face_kernel(BlockDims b, const double* __restrict__ emf, const double* fBsub, const double* fBfull, double g0,
               double g1, const DeviceState* __restrict__ ds, double beta, double* fBout) {
    const int ni = b.nx1 + 1, nj = b.nx2 + 1;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;   // edge / face counts stay below 2^31
    const unsigned nface = (unsigned)ni * (unsigned)nj * (unsigned)(ND > 2 ? b.nx3 + 1 : 1);
    if (t >= nface) return;
    const int i = b.is + (int)(t % (unsigned)ni);
    const int j = b.js + (int)((t / (unsigned)ni) % (unsigned)nj);
    const int k = b.ks + (int)(t / ((unsigned)ni * (unsigned)nj));
    const long f = (long)k * b.fs3 + (long)j * b.fs2 + i;
    const long fs2 = b.fs2, fs3 = b.fs3, nf = b.nf;
    const double dx1 = b.dx1, dx2 = b.dx2, dx3 = b.dx3;
    const double* E1 = emf;
    const double* E2 = emf + nf;
    const double* E3 = emf + 2 * nf;
    const bool iin = i <= b.ie, jin = j <= b.je, kin = k <= b.ke;
    double d[3];
    bool on[3];
    on[0] = jin && kin;
    on[1] = iin && kin;
    on[2] = iin && jin;
    if (on[0]) {
        d[0] = (-dx3 * E3[f + fs2] + dx3 * E3[f]);
        if (ND > 2) d[0] += (dx2 * E2[f + fs3] - dx2 * E2[f]);
        d[0] /= (dx2 * dx3);
    }
    if (on[1]) {
        d[1] = (dx3 * E3[f + 1] - dx3 * E3[f]);
        if (ND > 2) d[1] += (-dx1 * E1[f + fs3] + dx1 * E1[f]);
        d[1] /= (dx1 * dx3);
    }
    if (on[2]) d[2] = (-dx2 * E2[f + 1] + dx2 * E2[f] + dx1 * E1[f + fs2] - dx1 * E1[f]) / (dx1 * dx2);
    const double bdt = OUT_RATE ? 0. : beta * ds->dt;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (!on[q]) continue;
        const long fq = (long)q * nf + f;
        if (OUT_RATE) {
            fBout[fq] = d[q];
        } else {
            const double base = SAME ? fBfull[fq] : g0 * fBsub[fq] + g1 * fBfull[fq];
            fBout[fq] = 1.0 * base + bdt * d[q];
        }
    }
}

// KHARMADriver::WeightedSumDataFace (driver/kharma_driver.hpp:167-190): z = w1 x + w2 y.
__global__ void __launch_bounds__(CBS) ct_weighted_sum_kernel(long n, const double* x, const double* y, double w1,
                                                               double w2, double* z) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) z[q] = w1 * x[q] + w2 * y[q];
}

// Periodic exchange of the face fields of one block with itself: along a face's
// own axis its faces s..e+1 are the block's and the others take the value N
// faces in; across, the cells wrap as a cell field's do.  Every slot reads an
// interior position, so the kernel writes nothing it reads.
__device__ __forceinline__ int ct_wrap(int x, int lo, int hi, int N) {
    return (x < lo) ? x + N : (x > hi) ? x - N : x;
}
__global__ void __launch_bounds__(CBS) ct_sync_kernel(BlockDims b, double* fB) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;   // face positions stay below 2^31
    const int f = blockIdx.y;
    if (t >= (unsigned)b.nf) return;
    const int i = (int)(t % (unsigned)b.fs2), j = (int)((t / (unsigned)b.fs2) % (unsigned)(b.n2 + 1)),
              k = (int)(t / (unsigned)b.fs3);
    const int x[3] = {i, j, k};
    const int n[3] = {b.n1, b.n2, b.n3};
    const int lo[3] = {b.is, b.js, b.ks}, hi[3] = {b.ie, b.je, b.ke};
    const int N[3] = {b.nx1, b.nx2, b.nx3};
    int y[3];
    bool ghost = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const bool along = (d == f) && d < b.ndim;
        const int top = along ? n[d] : n[d] - 1;
        if (x[d] > top) return;   // a slot that holds no face
        y[d] = x[d];
        if (d < b.ndim && b.periodic[d]) {
            y[d] = ct_wrap(x[d], lo[d], along ? hi[d] + 1 : hi[d], N[d]);
            ghost |= y[d] != x[d];
        }
    }
    if (!ghost) return;
    fB[(long)f * b.nf + t] = fB[(long)f * b.nf + (long)y[2] * b.fs3 + (long)y[1] * b.fs2 + y[0]];
}

// B_CT::BlockUtoP (b_ct.cpp:173-205) over the entire block: primitive B of each
// cell from its two faces (each over its sqrt(-g) = 1), conserved B = B sqrt(-g).
__global__ void __launch_bounds__(CBS) ct_utop_kernel(BlockDims b, const double* __restrict__ fB, Vars<double> U,
                                                       Vars<double> P) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.nc) return;
    const int i = (int)(c % b.n1), j = (int)((c / b.n1) % b.n2), k = (int)(c / b.s3);
    const long f = (long)k * b.fs3 + (long)j * b.fs2 + i;
    const double* F1 = fB;
    const double* F2 = fB + b.nf;
    const double* F3 = fB + 2 * b.nf;
    const double p1 = (F1[f] / 1. + F1[f + 1] / 1.) / 2;
    const double p2 = (b.ndim > 1) ? (F2[f] / 1. + F2[f + b.fs2] / 1.) / 2 : F2[f] / 1.;
    const double p3 = (b.ndim > 2) ? (F3[f] / 1. + F3[f + b.fs3] / 1.) / 2 : F3[f] / 1.;
    P[B1][c] = p1;
    P[B2][c] = p2;
    P[B3][c] = p3;
    U[B1][c] = p1 * 1.;
    U[B2][c] = p2 * 1.;
    U[B3][c] = p3 * 1.;
}

// B_CT::MaxDivB (b_ct.cpp:472-497): |face_div| (b_ct_functions.hpp:47-56) of the
// interior cells, one partial maximum per work-group over the corner-sized grid
// of divb_blocks (threads past the interior contribute 0).
__global__ void __launch_bounds__(CBS) ct_max_divb_kernel(BlockDims b, const double* __restrict__ fB,
                                                           double* __restrict__ part) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nin = (long)b.nx1 * b.nx2 * (b.ndim > 2 ? b.nx3 : 1);
    double m = 0.;
    if (t < nin) {
        const int i = b.is + (int)(t % b.nx1);
        const int j = b.js + (int)((t / b.nx1) % b.nx2);
        const int k = b.ks + (int)(t / ((long)b.nx1 * b.nx2));
        const long f = (long)k * b.fs3 + (long)j * b.fs2 + i;
        double du = (fB[f + 1] - fB[f]) / b.dx1;
        du += (fB[b.nf + f + b.fs2] - fB[b.nf + f]) / b.dx2;
        if (b.ndim > 2) du += (fB[2 * b.nf + f + b.fs3] - fB[2 * b.nf + f]) / b.dx3;
        m = fabs(du);
    }
    m = ct_block_max(m);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// ------------------------------------------------------------------ launchers
static inline dim3 ct_grid(long n) { return dim3((unsigned)((n + CBS - 1) / CBS)); }

hipError_t launch_ct_cemf(const BlockDims& b, const Vars<const double>& P, double* cemf, hipStream_t s) {
    hipLaunchKernelGGL(ct_cemf_kernel, ct_grid(b.nc), dim3(CBS), 0, s, b, P, cemf);
    return hipGetLastError();
}

template <typename R>
static void launch_ct_emf_t(const BlockDims& b, const FluxVars<const R>& F, const double* cemf, double* emf,
                            hipStream_t s) {
    const long ne = (long)(b.nx1 + 1) * (b.nx2 + 1) * (b.ndim > 2 ? b.nx3 + 1 : 1);
    const int sch = b.ct_scheme == 3 ? 2 : b.ct_scheme;
#define KH_CTE(SC, ND) hipLaunchKernelGGL((ct_emf_kernel<R, SC, ND>), ct_grid(ne), dim3(CBS), 0, s, b, F, cemf, emf)
    if (b.ndim > 2) {
        if (sch == 0) KH_CTE(0, 3); else if (sch == 1) KH_CTE(1, 3); else KH_CTE(2, 3);
    } else {
        if (sch == 0) KH_CTE(0, 2); else if (sch == 1) KH_CTE(1, 2); else KH_CTE(2, 2);
    }
#undef KH_CTE
}

hipError_t launch_ct_emf(const BlockDims& b, int prec, const FluxVars<const void>& F, const double* cemf, double* emf,
                         hipStream_t s) {
    if (prec == PREC_SPLIT) launch_ct_emf_t<float>(b, flux_vars_cast<const float>(F), cemf, emf, s);
    else launch_ct_emf_t<double>(b, flux_vars_cast<const double>(F), cemf, emf, s);
    return hipGetLastError();
}

hipError_t launch_ct_face_update(const BlockDims& b, const double* emf, const double* fBsub, const double* fBfull,
                                 double g0, double g1, bool same, const DeviceState* ds, double beta, double* fBout,
                                 hipStream_t s) {
    const long n = (long)(b.nx1 + 1) * (b.nx2 + 1) * (b.ndim > 2 ? b.nx3 + 1 : 1);
#define KH_CTF(ND, RATE, SAME) \
    hipLaunchKernelGGL((ct_face_kernel<ND, RATE, SAME>), ct_grid(n), dim3(CBS), 0, s, b, emf, fBsub, fBfull, g0, g1, ds, beta, fBout)
    if (!ds) {
        if (b.ndim > 2) KH_CTF(3, true, false); else KH_CTF(2, true, false);
    } else if (same) {
        if (b.ndim > 2) KH_CTF(3, false, true); else KH_CTF(2, false, true);
    } else {
        if (b.ndim > 2) KH_CTF(3, false, false); else KH_CTF(2, false, false);
    }
#undef KH_CTF
    return hipGetLastError();
}

hipError_t launch_ct_weighted_sum(const BlockDims& b, const double* x, const double* y, double w1, double w2, double* z,
                                  hipStream_t s) {
    const long n = 3 * b.nf;
    hipLaunchKernelGGL(ct_weighted_sum_kernel, ct_grid(n), dim3(CBS), 0, s, n, x, y, w1, w2, z);
    return hipGetLastError();
}

hipError_t launch_ct_sync(const BlockDims& b, double* fB, hipStream_t s) {
    hipLaunchKernelGGL(ct_sync_kernel, dim3((unsigned)((b.nf + CBS - 1) / CBS), 3), dim3(CBS), 0, s, b, fB);
    return hipGetLastError();
}

hipError_t launch_ct_utop(const BlockDims& b, const double* fB, const Vars<double>& U, const Vars<double>& P,
                          hipStream_t s) {
    hipLaunchKernelGGL(ct_utop_kernel, ct_grid(b.nc), dim3(CBS), 0, s, b, fB, U, P);
    return hipGetLastError();
}

hipError_t launch_ct_max_divb(const BlockDims& b, const double* fB, double* part, hipStream_t s) {
    hipLaunchKernelGGL(ct_max_divb_kernel, dim3((unsigned)divb_blocks(b)), dim3(CBS), 0, s, b, fB, part);
    return hipGetLastError();
}

}  // namespace kh
