// kh_mesh.hip -- moving state between the layout of a whole mesh and the
// layout of its meshblocks.
//
// A mesh of nb1 x nb2 x nb3 equal meshblocks keeps every container as
// [block][8][nc] (block index b1 fastest), each block's [8][nc] laid out as a
// one-block mesh's (entire extent, ng ghosts per active side).  The mesh
// layout is the one-block layout of the whole mesh, [8][N3e][N2e][N1e].  The
// update kernels keep the blocks' ghosts current themselves (every interior
// zone within ng of a face shared with another block is also stored into that
// block's ghosts, kh_step.hip's image stores); these kernels only load a
// state into the blocks and read one back.
#include "kh_internal.hpp"

namespace kh {

__device__ __forceinline__ int mesh_wrap(int x, int ng, int n) {
    int r = (x - ng) % n;
    if (r < 0) r += n;
    return ng + r;
}

// Every zone of every block's entire extent from the mesh array: the mesh zone
// at the same place, a periodic axis's ghosts taken from their images (what the
// periodic exchange of a one-block mesh gives), a physical face's ghosts as the
// mesh array holds them (the boundary fill overwrites them).
__global__ void __launch_bounds__(256) mesh_scatter_kernel(MeshDims m, const double* __restrict__ Pm,
                                                           double* __restrict__ Pb) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int blk = blockIdx.y;
    if (t >= m.nc) return;
    const int b1 = blk % m.nb[0], b2 = (blk / m.nb[0]) % m.nb[1], b3 = blk / (m.nb[0] * m.nb[1]);
    const int k = (int)(t / ((long)m.n[0] * m.n[1]));
    const int j = (int)((t / m.n[0]) % m.n[1]);
    const int i = (int)(t % m.n[0]);
    int I = i + b1 * m.nx[0], J = j + b2 * m.nx[1], K = k + b3 * m.nx[2];
    if (m.periodic[0] && (I < m.ng || I >= m.ng + m.N[0])) I = mesh_wrap(I, m.ng, m.N[0]);
    if (m.ndim > 1 && m.periodic[1] && (J < m.ng || J >= m.ng + m.N[1])) J = mesh_wrap(J, m.ng, m.N[1]);
    if (m.ndim > 2 && m.periodic[2] && (K < m.ng || K >= m.ng + m.N[2])) K = mesh_wrap(K, m.ng, m.N[2]);
    const long src = ((long)K * m.Ne[1] + J) * m.Ne[0] + I;
    double* dst = Pb + (long)blk * NV * m.nc + t;
#pragma unroll
    for (int v = 0; v < NV; ++v) dst[(long)v * m.nc] = Pm[(long)v * m.ncm + src];
}

// Block and local index along one axis of mesh-entire index X: interior zones
// from the block that holds them, a periodic axis's ghosts from their images,
// a physical face's ghosts from the edge block's own ghosts.
__device__ __forceinline__ void mesh_owner(int X, int ng, int N, int nx, int nb, int periodic, int& blk, int& x) {
    if ((X < ng || X >= ng + N) && periodic) X = mesh_wrap(X, ng, N);
    if (X < ng) blk = 0;
    else if (X >= ng + N) blk = nb - 1;
    else blk = (X - ng) / nx;
    x = X - blk * nx;
}

__global__ void __launch_bounds__(256) mesh_gather_kernel(MeshDims m, const double* __restrict__ Pb,
                                                          double* __restrict__ Pm) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.ncm) return;
    const int K = (int)(t / ((long)m.Ne[0] * m.Ne[1]));
    const int J = (int)((t / m.Ne[0]) % m.Ne[1]);
    const int I = (int)(t % m.Ne[0]);
    int b1, b2 = 0, b3 = 0, i, j = J, k = K;
    mesh_owner(I, m.ng, m.N[0], m.nx[0], m.nb[0], m.periodic[0], b1, i);
    if (m.ndim > 1) mesh_owner(J, m.ng, m.N[1], m.nx[1], m.nb[1], m.periodic[1], b2, j);
    if (m.ndim > 2) mesh_owner(K, m.ng, m.N[2], m.nx[2], m.nb[2], m.periodic[2], b3, k);
    const int blk = (b3 * m.nb[1] + b2) * m.nb[0] + b1;
    const long src = (long)blk * NV * m.nc + ((long)k * m.n[1] + j) * m.n[0] + i;
#pragma unroll
    for (int v = 0; v < NV; ++v) Pm[(long)v * m.ncm + t] = Pb[src + (long)v * m.nc];
}

hipError_t launch_mesh_scatter(const MeshDims& m, const double* Pm, double* Pb, hipStream_t s) {
    const dim3 grid((unsigned)((m.nc + 255) / 256), (unsigned)m.nblk);
    hipLaunchKernelGGL(mesh_scatter_kernel, grid, dim3(256), 0, s, m, Pm, Pb);
    return hipGetLastError();
}

hipError_t launch_mesh_gather(const MeshDims& m, const double* Pb, double* Pm, hipStream_t s) {
    hipLaunchKernelGGL(mesh_gather_kernel, dim3((unsigned)((m.ncm + 255) / 256)), dim3(256), 0, s, m, Pb, Pm);
    return hipGetLastError();
}

}  // namespace kh
