/*
 * Cartesian acoustic FD step kernels for gfx950 (MI355X).
 *
 * One launch advances the whole padded grid one time step and, in the same
 * launch, extracts the receiver traces of the step (extra blocks).  It fuses
 * what the reference CPU loop does in separate passes (MAWEFDcart2Dcpu.c:303-378,
 * MAWEFDcart3Dcpu.c:302-360):
 *   step_forward_2d/3d   (ts_kernel.c:49-314)   2N-order Laplacian + leapfrog
 *   lint*_inject*_with_vv (fdutil.c:1580-1740)  injection scaled by (v dt)^2
 *   abcone3d_apply, z faces (fdutil.c:2085-2120) in-register recurrence
 *   sponge*_apply        (fdutil.c:2206-2310)   strips, z then x then y factors
 *   lint*_extract        (fdutil.c:1596-1670)
 *
 * "Lazy" second sponge: the reference multiplies BOTH fields by the sponge
 * every step.  The field that becomes "previous" is only read once more, at
 * its own node, so its second multiplication is applied when it is read
 * (here, and by the extraction, which reads it as the reference's u0).  The
 * stored "previous" field therefore always lacks exactly one sponge factor.
 * Every multiplication happens in the reference order, so the result is bit
 * for bit the reference's.
 *
 * Layout: z fastest (RSF axis 1), then x, then y; row length n1.
 * Blocks: 256 threads = 4 waves; a wave covers up to 64 z nodes of 4 x rows;
 * a block covers a (tz <= 64) x 16 tile and streams it along y (2.5-D
 * blocking) with a per-thread register queue of 2*NOP+1 planes.  The current
 * plane of the tile plus NOP-wide halos is staged in LDS (double buffered,
 * one barrier per plane).  Each field value is read from HBM once per step.
 *
 * Compiled with -ffp-contract=off: no FMA contraction, so every add and
 * multiply rounds as in the reference's SSE (or scalar) build.
 */
#include <hip/hip_runtime.h>
#include "awefd_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int WAVE = 64;
constexpr int NWAVE = 4;
constexpr int RX = 4;
constexpr int TX = NWAVE * RX;
static_assert(TX == AWEFD_CART_TX, "tile width");

/* Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
 * Give each XCD a contiguous range of tiles so that tiles sharing halos
 * share an L2. */
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    const int per = n >> 3, rem = n & 7, xcd = b & 7, slot = b >> 3;
    return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

template <int NDIM>
__device__ __forceinline__ float sponge(float v, int iz, int ix, int iy, const awefd_cart_args &a)
{
    v = v * a.sp1[iz];
    v = v * a.sp2[ix];
    if (NDIM == 3) v = v * a.sp3[iy];
    return v;
}

template <int NDIM>
__device__ __forceinline__ bool in_sponge(int iz, int ix, int iy, const awefd_cart_args &a)
{
    bool s = iz < a.nb || iz >= a.n1 - a.nb || ix < a.nb || ix >= a.n2 - a.nb;
    if (NDIM == 3) s = s || iy < a.nb || iy >= a.n3 - a.nb;
    return s;
}

/* lint2d_extract / lint3d_extract of the reference's u0, i.e. the current
 * field with its pending sponge factor. */
template <int NDIM>
__device__ void cart_extract(const awefd_cart_args &a, int r)
{
    if (r >= a.nr) return;
    const long sx = a.n1, sy = (long)a.n1 * a.n2;
    const int jz = a.rj[3 * r], jx = a.rj[3 * r + 1], jy = a.rj[3 * r + 2];
    constexpr int NC = NDIM == 3 ? 8 : 4;
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < NC; q++) {
        const int z = jz + (q & 1), x = jx + ((q >> 1) & 1), y = NDIM == 3 ? jy + ((q >> 2) & 1) : 0;
        float u = a.C[y * sy + x * sx + z];
        if (a.dabc) u = sponge<NDIM>(u, z, x, y, a);
        const float t = u * a.rw[8 * r + q];
        s = q == 0 ? t : s + t;
    }
    a.trow[r] = s;
}

template <int NOP, int NDIM, bool SSE>
__global__ __launch_bounds__(256) void k_cart_step(const awefd_cart_args a)
{
    const int tid = threadIdx.x;
    if ((int)blockIdx.x >= a.nblk) {
        cart_extract<NDIM>(a, ((int)blockIdx.x - a.nblk) * 256 + tid);
        return;
    }
    const int lb = xcd_remap(blockIdx.x, a.nblk);
    const int tzi = lb % a.ntz;
    const int rr = lb / a.ntz;
    const int txi = rr % a.ntx;
    const int tyi = rr / a.ntx;
    const int lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int z0 = tzi * a.tz;
    const int tzw = min(a.tz, a.n1 - z0);
    const int x0 = txi * TX;
    const int y0 = NDIM == 3 ? tyi * a.ychunk : 0;
    const int y1 = NDIM == 3 ? min(y0 + a.ychunk, a.n3) : 1;
    const int iz = z0 + lane;
    const bool zin = lane < tzw;
    const long sx = a.n1, sy = (long)a.n1 * a.n2;

    constexpr int LW = WAVE + 2 * NOP;
    constexpr int LR = TX + 2 * NOP;
    __shared__ float lds[2][LR * LW];

    int ixr[RX];
    bool xin[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        ixr[j] = x0 + wv * RX + j;
        xin[j] = ixr[j] < a.n2;
    }

    float q[RX][2 * NOP + 1];
    if (NDIM == 3) {
#pragma unroll
        for (int k = 0; k < 2 * NOP; k++) {
            const int yy = y0 - NOP + k;
#pragma unroll
            for (int j = 0; j < RX; j++)
                q[j][k + 1] = (zin && xin[j] && yy >= 0 && yy < a.n3)
                                  ? a.C[yy * sy + ixr[j] * sx + iz] : 0.f;
        }
    }

    /* injection entries of this tile column, restricted to planes >= y0 */
    const int col = tzi + a.ntz * txi;
    int ib = a.ibeg[col];
    const int ie = a.ibeg[col + 1];
    while (ib < ie && a.inj[ib].y < y0) ib++;

    const bool zlow = NDIM == 3 && a.abc && tzi == 0 && !a.fsrf;
    const bool zhigh = NDIM == 3 && a.abc && tzi == a.ntz - 1;
    int buf = 0;
    for (int y = y0; y < y1; y++, buf ^= 1) {
        float *L = lds[buf];
        float cen[RX];
        if (NDIM == 3) {
            const int yn = y + NOP;
#pragma unroll
            for (int j = 0; j < RX; j++) {
#pragma unroll
                for (int k = 0; k < 2 * NOP; k++) q[j][k] = q[j][k + 1];
                q[j][2 * NOP] = (zin && xin[j] && yn < a.n3) ? a.C[yn * sy + ixr[j] * sx + iz] : 0.f;
                cen[j] = q[j][NOP];
            }
        } else {
#pragma unroll
            for (int j = 0; j < RX; j++) cen[j] = (zin && xin[j]) ? a.C[ixr[j] * sx + iz] : 0.f;
        }
        const long yb = (long)y * sy;
        float pv[RX], vv[RX];
#pragma unroll
        for (int j = 0; j < RX; j++) {
            pv[j] = 0.f;
            vv[j] = 0.f;
            if (zin && xin[j]) {
                pv[j] = a.P[yb + ixr[j] * sx + iz];
                vv[j] = a.vel[yb + ixr[j] * sx + iz];
            }
        }
        /* stage the current plane (tile + NOP halos) in LDS */
        if (zin) {
#pragma unroll
            for (int j = 0; j < RX; j++) L[(wv * RX + j + NOP) * LW + lane + NOP] = cen[j];
        }
        for (int h = wv; h < 2 * NOP; h += NWAVE) {
            const int xr = h < NOP ? x0 - NOP + h : x0 + TX + (h - NOP);
            const int lrow = h < NOP ? h : TX + NOP + (h - NOP);
            if (zin) {
                float v = 0.f;
                if (xr >= 0 && xr < a.n2) v = a.C[yb + xr * sx + iz];
                L[lrow * LW + lane + NOP] = v;
            }
        }
        for (int e = tid; e < 2 * NOP * TX; e += 256) {
            const int row = e / (2 * NOP), s = e % (2 * NOP);
            const int xr = x0 + row;
            const int zz = s < NOP ? z0 - NOP + s : z0 + tzw + (s - NOP);
            const int lcol = s < NOP ? s : NOP + tzw + (s - NOP);
            float v = 0.f;
            if (xr < a.n2 && zz >= 0 && zz < a.n1) v = a.C[yb + xr * sx + zz];
            L[(row + NOP) * LW + lcol] = v;
        }
        __syncthreads();

        const bool yint = NDIM == 2 || (y >= NOP && y < a.n3 - NOP);
        const bool zint = iz >= NOP && iz < a.n1 - NOP;
        float nv[RX];
#pragma unroll
        for (int j = 0; j < RX; j++) {
            const int ix = ixr[j];
            float p = pv[j];
            if (a.dabc && zin && xin[j] && in_sponge<NDIM>(iz, ix, y, a)) p = sponge<NDIM>(p, iz, ix, y, a);
            const bool interior = zint && yint && ix >= NOP && ix < a.n2 - NOP;
            const float *Lr = L + (wv * RX + j + NOP) * LW + lane + NOP;
            float v = p;
            if (zin && xin[j] && interior) {
                const float c = cen[j];
                float lap = c * a.c0;
                if (SSE) {
#pragma unroll
                    for (int k = 1; k <= NOP; k++) {
                        lap = lap + a.cz[k - 1] * (Lr[k] + Lr[-k]);
                        lap = lap + a.cx[k - 1] * (Lr[k * LW] + Lr[-k * LW]);
                        if (NDIM == 3) lap = lap + a.cy[k - 1] * (q[j][NOP + k] + q[j][NOP - k]);
                    }
                    const float t = 2.f * c - p;
                    v = t + vv[j] * lap;
                } else {
#pragma unroll
                    for (int k = 1; k <= NOP; k++) {
                        float t = (Lr[-k] + Lr[k]) * a.cz[k - 1] + (Lr[-k * LW] + Lr[k * LW]) * a.cx[k - 1];
                        if (NDIM == 3) t = t + (q[j][NOP - k] + q[j][NOP + k]) * a.cy[k - 1];
                        lap = lap + t;
                    }
                    const float vl = vv[j] * lap;
                    v = (float)(2.0 * (double)c - (double)p + (double)vl);
                }
            }
            nv[j] = v;
        }

        /* source injection into the new field, in source order */
        int e = ib;
        for (; e < ie && a.inj[e].y == y; e++) {
            const awefd_inj en = a.inj[e];
            const float amp = a.wav[(long)a.it * a.wstride + (a.wexpl ? 0 : en.src)];
#pragma unroll
            for (int j = 0; j < RX; j++)
                if (zin && en.z == iz && en.x == ixr[j]) nv[j] = nv[j] + amp * en.w * vv[j];
        }
        ib = e;

        /* one-way ABC on the z faces: serial outward recurrence along z,
         * carried across lanes (abcone3d_apply, fdutil.c:2097-2120) */
        if (zlow || zhigh) {
#pragma unroll
            for (int j = 0; j < RX; j++) {
                if (!xin[j]) continue;
                const int ix = ixr[j];
                const float *Lc = L + (wv * RX + j + NOP) * LW + NOP;
                if (zlow) {
                    const float b = a.bzl[y * a.n2 + ix];
                    for (int k = 0; k < NOP; k++) {
                        const int zt = NOP - k;
                        const float up = __shfl_down(nv[j], 1);
                        if (lane == zt) nv[j] = Lc[zt + 1] + (Lc[zt] - up) * b;
                    }
                }
                if (zhigh) {
                    const float b = a.bzh[y * a.n2 + ix];
                    for (int k = 0; k < NOP; k++) {
                        const int zt = a.n1 - NOP + k - 1 - z0;
                        const float dn = __shfl_up(nv[j], 1);
                        if (lane == zt) nv[j] = Lc[zt - 1] + (Lc[zt] - dn) * b;
                    }
                }
            }
        }

        /* sponge of the new field and store; ABC x/y shell nodes are left
         * unsponged for the face kernels */
#pragma unroll
        for (int j = 0; j < RX; j++) {
            if (!(zin && xin[j])) continue;
            const int ix = ixr[j];
            float v = nv[j];
            if (a.dabc && in_sponge<NDIM>(iz, ix, y, a)) {
                bool defer = false;
                if (NDIM == 3 && a.abc) {
                    const bool xsh = ix <= NOP + 1 || ix >= a.n2 - NOP - 2;
                    const bool ysh = y <= NOP + 1 || y >= a.n3 - NOP - 2;
                    defer = xsh || ysh;
                }
                if (!defer) v = sponge<NDIM>(v, iz, ix, y, a);
            }
            a.P[yb + ix * sx + iz] = v;
        }
    }
}

/* One-way ABC on the x faces, then the y faces (abcone3d_apply,
 * fdutil.c:2122-2176), each followed by the sponge of the shell nodes it
 * finalises.  One thread per line; lines along z are coalesced. */
template <int NOP>
__global__ __launch_bounds__(256) void k_cart_abc_x(const awefd_cart_args a)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)a.n1 * a.n3) return;
    const int iz = (int)(t % a.n1), iy = (int)(t / a.n1);
    const long sx = a.n1, sy = (long)a.n1 * a.n2;
    const long b = iy * sy + iz;
    const float bl = a.bxl[(long)iy * a.n1 + iz], bh = a.bxh[(long)iy * a.n1 + iz];
    float *P = a.P;
    const float *C = a.C;
    for (int k = 0; k < NOP; k++) {
        long ix = NOP - k;
        P[b + ix * sx] = C[b + (ix + 1) * sx] + (C[b + ix * sx] - P[b + (ix + 1) * sx]) * bl;
        ix = a.n2 - NOP + k - 1;
        P[b + ix * sx] = C[b + (ix - 1) * sx] + (C[b + ix * sx] - P[b + (ix - 1) * sx]) * bh;
    }
    if (!a.dabc) return;
    const bool ysh = iy <= NOP + 1 || iy >= a.n3 - NOP - 2;
    if (ysh) return;
    /* x shell: [0, NOP+2) and [n2-NOP-2, n2), each node once */
    const int hi = max(NOP + 2, a.n2 - NOP - 2);
    for (int ix = 0; ix < a.n2; ix = ix + 1 == NOP + 2 ? hi : ix + 1)
        if (in_sponge<3>(iz, ix, iy, a)) P[b + ix * sx] = sponge<3>(P[b + ix * sx], iz, ix, iy, a);
}

template <int NOP>
__global__ __launch_bounds__(256) void k_cart_abc_y(const awefd_cart_args a)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)a.n1 * a.n2) return;
    const int iz = (int)(t % a.n1), ix = (int)(t / a.n1);
    const long sx = a.n1, sy = (long)a.n1 * a.n2;
    const long b = ix * sx + iz;
    const float bl = a.byl[(long)ix * a.n1 + iz], bh = a.byh[(long)ix * a.n1 + iz];
    float *P = a.P;
    const float *C = a.C;
    for (int k = 0; k < NOP; k++) {
        long iy = NOP - k;
        P[b + iy * sy] = C[b + (iy + 1) * sy] + (C[b + iy * sy] - P[b + (iy + 1) * sy]) * bl;
        iy = a.n3 - NOP + k - 1;
        P[b + iy * sy] = C[b + (iy - 1) * sy] + (C[b + iy * sy] - P[b + (iy - 1) * sy]) * bh;
    }
    if (!a.dabc) return;
    /* y shell: [0, NOP+2) and [n3-NOP-2, n3), each node once */
    const int hi = max(NOP + 2, a.n3 - NOP - 2);
    for (int iy = 0; iy < a.n3; iy = iy + 1 == NOP + 2 ? hi : iy + 1)
        if (in_sponge<3>(iz, ix, iy, a)) P[b + iy * sy] = sponge<3>(P[b + iy * sy], iz, ix, iy, a);
}

__global__ void k_fill(float *p, long n, float v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = v;
}

template <int NOP, int NDIM, bool SSE>
int launch_step(const awefd_cart_args *a, hipStream_t s)
{
    const int nextract = (a->nr + 255) / 256;
    hipLaunchKernelGGL((k_cart_step<NOP, NDIM, SSE>), dim3(a->nblk + nextract), dim3(256), 0, s, *a);
    return (int)hipGetLastError();
}

template <int NOP>
int launch_step_nop(const awefd_cart_args *a, int ndim, int sse, hipStream_t s)
{
    if (ndim == 2) return sse ? launch_step<NOP, 2, true>(a, s) : launch_step<NOP, 2, false>(a, s);
    return sse ? launch_step<NOP, 3, true>(a, s) : launch_step<NOP, 3, false>(a, s);
}

template <int NOP>
int launch_abc(const awefd_cart_args *a, hipStream_t s)
{
    const long nxl = (long)a->n1 * a->n3, nyl = (long)a->n1 * a->n2;
    hipLaunchKernelGGL((k_cart_abc_x<NOP>), dim3((unsigned)((nxl + 255) / 256)), dim3(256), 0, s, *a);
    hipLaunchKernelGGL((k_cart_abc_y<NOP>), dim3((unsigned)((nyl + 255) / 256)), dim3(256), 0, s, *a);
    return (int)hipGetLastError();
}

} // namespace

extern "C" int awefd_launch_cart_step(const awefd_cart_args *a, int nop, int ndim, int sse,
                                      int varden, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (varden) return (int)hipErrorInvalidValue;
    switch (nop) {
    case 1: return launch_step_nop<1>(a, ndim, sse, s);
    case 2: return launch_step_nop<2>(a, ndim, sse, s);
    case 3: return launch_step_nop<3>(a, ndim, sse, s);
    case 4: return launch_step_nop<4>(a, ndim, sse, s);
    case 5: return launch_step_nop<5>(a, ndim, sse, s);
    case 6: return launch_step_nop<6>(a, ndim, sse, s);
    default: return (int)hipErrorInvalidValue;
    }
}

extern "C" int awefd_launch_cart_abc_xy(const awefd_cart_args *a, int nop, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    switch (nop) {
    case 1: return launch_abc<1>(a, s);
    case 2: return launch_abc<2>(a, s);
    case 3: return launch_abc<3>(a, s);
    case 4: return launch_abc<4>(a, s);
    case 5: return launch_abc<5>(a, s);
    case 6: return launch_abc<6>(a, s);
    default: return (int)hipErrorInvalidValue;
    }
}

extern "C" int awefd_launch_fill(float *p, long n, float v, void *stream)
{
    if (n <= 0) return 0;
    long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, p, n, v);
    return (int)hipGetLastError();
}
