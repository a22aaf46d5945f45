/*
 * Adjoint of the cartesian 2-D time loop (gfx950), for the gradient of a
 * trace misfit with respect to squared slowness.
 *
 * With c_n the field at time n after its first sponge (the stored current
 * field), G the sponge (diagonal), M the interior mask, L the 2*NOP-order
 * Laplacian and V = (v dt)^2, the forward step of the reference loop
 * (MAWEFDcart2Dcpu.c:297-383, ts_kernel.c:49-161) is
 *
 *   c_{n+1} = G [ M (2 c_n - p_n + V.L c_n) + (I - M) p_n + f_n ],
 *   p_{n+1} = G c_n,   trace row n = R G c_n,   f_n = w_n W V (sources).
 *
 * Its adjoint, backwards in n with t = G alpha_{n+1}:
 *
 *   alpha_n = G R^T r_n + 2 M t + L (V M t) + G beta_{n+1}
 *   beta_n  = (I - 2M) t
 *   dJ/dV  += M (L c_n) t  (+ w_n W t at the source nodes)
 *
 * k_adj2d_step evaluates alpha_n, beta_n and the M (L c_n) t term for every
 * node; k_adj2d_rec adds G R^T r_n from a per-node list of receiver weights
 * (each node's contributions summed by one thread, in receiver order);
 * k_adj2d_src adds the injection term; k_adj2d_chain maps dJ/dV to dJ/ds on
 * the model grid (dV/ds = -V v^2, summed over the padded copies of a node).
 */
#include <hip/hip_runtime.h>
#include "awefd_hip.h"
#include "awefd_internal.h"

namespace {

using namespace awefd;

/* V_j M_j G_j alpha_{n+1,j}: zero off the interior */
template <int NOP>
__device__ __forceinline__ float vmt(const awefd_adj2d_args &a, int jz, int jx)
{
    if (jz < NOP || jz >= a.nz - NOP || jx < NOP || jx >= a.nx - NOP) return 0.f;
    const long j = (long)jx * a.ld + jz;
    const float gj = a.dabc ? a.sp1[jz] * a.sp2[jx] : 1.f;
    return a.V[j] * (gj * a.A1[j]);
}

template <int NOP>
__global__ __launch_bounds__(256) void k_adj2d_step(const awefd_adj2d_args a)
{
    const int iz = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ix = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int nz = a.nz, nx = a.nx, ld = a.ld;
    if (iz >= nz || ix >= nx) return;
    const long i = (long)ix * ld + iz;
    const bool zin = iz >= NOP && iz < nz - NOP, xin = ix >= NOP && ix < nx - NOP;
    const bool inter = zin && xin;
    const float gi = a.dabc ? a.sp1[iz] * a.sp2[ix] : 1.f;
    const float t = gi * a.A1[i];
    /* L (V M t) at node i: neighbours j contribute V_j G_j A1_j when interior */
    float ly = a.c0 * vmt<NOP>(a, iz, ix);
#pragma unroll
    for (int k = 1; k <= NOP; k++)
        ly += a.cz[k - 1] * (vmt<NOP>(a, iz + k, ix) + vmt<NOP>(a, iz - k, ix))
              + a.cx[k - 1] * (vmt<NOP>(a, iz, ix + k) + vmt<NOP>(a, iz, ix - k));
    a.A0[i] = (inter ? 2.f * t : 0.f) + ly + gi * a.B1[i];
    a.B0[i] = inter ? -t : t;
    if (inter) {
        const float *c = a.Cn;
        float lc = a.c0 * c[i];
#pragma unroll
        for (int k = 1; k <= NOP; k++)
            lc += a.cz[k - 1] * (c[i + k] + c[i - k]) + a.cx[k - 1] * (c[i + (long)k * ld] + c[i - (long)k * ld]);
        a.gV[i] += lc * t;
    }
}

/* alpha_n += G R^T r_n: node k of the receiver list gathers its weights */
__global__ __launch_bounds__(256) void k_adj2d_rec(const awefd_adj2d_args a)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= a.nrnode) return;
    const long i = a.rnode[k];
    float s = 0.f;
    for (int e = a.rbeg[k]; e < a.rbeg[k + 1]; e++) s += a.rw[e] * a.res[a.rid[e]];
    const int iz = (int)(i % a.ld), ix = (int)(i / a.ld);
    const float gi = a.dabc ? a.sp1[iz] * a.sp2[ix] : 1.f;
    a.A0[i] += gi * s;
}

/* dJ/dV += w_n W t at the injection nodes (t = G alpha_{n+1}) */
__global__ __launch_bounds__(256) void k_adj2d_src(const awefd_adj2d_args a)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= a.nsnode) return;
    const long i = a.snode[k];
    const int iz = (int)(i % a.ld), ix = (int)(i / a.ld);
    const float gi = a.dabc ? a.sp1[iz] * a.sp2[ix] : 1.f;
    float s = 0.f;
    for (int e = a.sbeg[k]; e < a.sbeg[k + 1]; e++) s += a.wav[(long)a.it * a.wstride + a.sid[e]] * a.sw[e];
    a.gV[i] += s * (gi * a.A1[i]);
}

/* dJ/ds on the model grid: sum over the padded nodes copied from model node
 * (kz, kx) of dJ/dV * dV/ds, dV/ds = -V v^2 = -V (V / dt^2) (0 where a free
 * surface fixed V = 0) */
__global__ __launch_bounds__(256) void k_adj2d_chain(const awefd_adj2d_args a, float *gs)
{
    const int kz = blockIdx.x * 64 + (threadIdx.x & 63);
    const int kx = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (kz >= a.mz || kx >= a.mx) return;
    const int nb = a.nb;
    const int z0 = kz == 0 ? 0 : kz + nb, z1 = kz == a.mz - 1 ? a.nz : kz + nb + 1;
    const int x0 = kx == 0 ? 0 : kx + nb, x1 = kx == a.mx - 1 ? a.nx : kx + nb + 1;
    float acc = 0.f;
    for (int ix = x0; ix < x1; ix++)
        for (int iz = z0; iz < z1; iz++) {
            const long i = (long)ix * a.ld + iz;
            const float v = a.V[i];
            acc += a.gV[i] * (-v * (v * a.idt2));
        }
    gs[(long)kx * a.mz + kz] = acc;
}

template <int NOP>
int launch_step(const awefd_adj2d_args *a, hipStream_t s)
{
    const dim3 g((a->nz + 63) / 64, (a->nx + 3) / 4);
    hipLaunchKernelGGL(k_adj2d_step<NOP>, g, dim3(256), 0, s, *a);
    return (int)hipGetLastError();
}

} // namespace

extern "C" int awefd_launch_adj2d_step(const awefd_adj2d_args *a, int nop, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    int e;
    switch (nop) {
    case 1: e = launch_step<1>(a, s); break;
    case 2: e = launch_step<2>(a, s); break;
    case 3: e = launch_step<3>(a, s); break;
    case 4: e = launch_step<4>(a, s); break;
    case 5: e = launch_step<5>(a, s); break;
    case 6: e = launch_step<6>(a, s); break;
    default: return (int)hipErrorInvalidValue;
    }
    if (e) return e;
    if (a->nrnode > 0) hipLaunchKernelGGL(k_adj2d_rec, dim3((a->nrnode + 255) / 256), dim3(256), 0, s, *a);
    if (a->nsnode > 0) hipLaunchKernelGGL(k_adj2d_src, dim3((a->nsnode + 255) / 256), dim3(256), 0, s, *a);
    return (int)hipGetLastError();
}

extern "C" int awefd_launch_adj2d_chain(const awefd_adj2d_args *a, float *gs, void *stream)
{
    const dim3 g((a->mz + 63) / 64, (a->mx + 3) / 4);
    hipLaunchKernelGGL(k_adj2d_chain, g, dim3(256), 0, (hipStream_t)stream, *a, gs);
    return (int)hipGetLastError();
}
