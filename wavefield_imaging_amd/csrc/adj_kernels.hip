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
 * Its adjoint, backwards in n, with t_n = G alpha_{n+1} and
 * beta_{n+1} = (I - 2M) t_{n+1} = (I - 2M) G alpha_{n+2}:
 *
 *   alpha_n = G R^T r_n + 2 M t_n + L (V M t_n) + G (I - 2M) G alpha_{n+2}
 *   dJ/dV  += M (L c_n) t_n  (+ w_n W t_n at the source nodes)
 *
 * so three alpha fields carry the whole adjoint state; four rotate, so that
 * a pass of two steps (alpha_n and alpha_{n-1} from alpha_{n+1} and
 * alpha_{n+2}) never writes a field another of its tiles still reads.
 *
 * k_adj2d_step does one backward step for every node in one launch.  Its
 * layout follows the forward step kernel: a block is NW waves, lane l of a
 * wave holds z = z0 + l of RX x rows, z neighbours move across lanes by DPP
 * shifts with the 2*NOP values beyond the tile edge fetched by the edge
 * lanes, and x neighbours come from the wave's own rows or NOP halo rows on
 * each side (served by L2).  Receiver residuals and the injection term are
 * applied in the same pass from per-tile buckets.  Per node and step it
 * moves alpha_{n+1}, V, c_n, alpha_{n+2} and dJ/dV in, alpha_n and dJ/dV out.
 *
 * k_adj2d_chain maps dJ/dV to dJ/ds on the model grid (dV/ds = -V v^2,
 * summed over the padded copies of a node).
 */
#include <hip/hip_runtime.h>
#include "awefd_hip.h"
#include "awefd_internal.h"

namespace {

using namespace awefd;
constexpr int NWA = AWEFD_ADJ_NWAVE;
constexpr int RXA = AWEFD_ADJ_RX;

/* z neighbours at distance 1..NOP of a value held one node per lane; the
 * edge lanes hold the values beyond the tile in `halo` (lane k < NOP:
 * z0-1-k, lane 63-k: z0+64+k).  Every lane of the wave must be active. */
template <int NOP>
__device__ __forceinline__ void znb(float c, float halo, float (&zm)[NOP], float (&zp)[NOP])
{
    float hlo = halo, hhi = halo;
    zm[0] = shr1(hlo, c);
    zp[0] = shl1(hhi, c);
#pragma unroll
    for (int k = 1; k < NOP; k++) {
        hlo = shl1(hlo, hlo);
        hhi = shr1(hhi, hhi);
        zm[k] = shr1(hlo, zm[k - 1]);
        zp[k] = shl1(hhi, zp[k - 1]);
    }
}

/* Entry `lane` of a bucket slice (node and amplitude; lanes past the end
 * get a node no lane matches): receivers carry r_n(receiver) * weight,
 * sources w_n(column) * weight. */
__device__ __forceinline__ void rec_entry(const awefd_adj2d_args &a, const float *res, int b0, int cnt, int lane,
                                          int &z, int &x, float &amp)
{
    z = x = -1;
    amp = 0.f;
    if (lane < cnt) {
        const awefd_inj en = a.rinj[b0 + lane];
        z = en.z;
        x = en.x;
        amp = res[en.src] * en.w;
    }
}

__device__ __forceinline__ void src_entry(const awefd_adj2d_args &a, int n, int b0, int cnt, int lane, int &z, int &x,
                                          float &amp)
{
    z = x = -1;
    amp = 0.f;
    if (lane < cnt) {
        const awefd_inj en = a.sinj[b0 + lane];
        z = en.z;
        x = en.x;
        amp = a.wav[(long)n * a.wstride + en.src] * en.w;
    }
}

/* One backward step n of one grid: alpha_n into A[n % 4], dJ/dV
 * accumulated. */
template <int NOP, int RX>
__device__ __forceinline__ void adj_body(const awefd_adj2d_args &a, int bid, int n)
{
    float *A0 = a.A[n & 3];
    const float *A1 = a.A[(n + 1) & 3], *A2 = a.A[(n + 2) & 3];
    const int k = n / a.K, n0 = k * a.K;
    const float *Cn = n == n0 ? a.ck + 2L * k * a.nal : a.seg + (long)(n - n0 - 1) * a.nal;
    const float *res = a.res + (long)n * a.nr;
    constexpr int H = RX + 2 * NOP;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / WAVE);
    const int lb = xcd_remap(bid, a.nblk);
    const int tzi = lb % a.ntz, txi = lb / a.ntz;
    const int nz = a.nz, nx = a.nx;
    const int z0 = tzi * WAVE, iz = z0 + lane;
    const int x0 = txi * (NWA * RX) + wv * RX;          /* wave-uniform */
    int zz = -1;
    if (lane < NOP) zz = z0 - 1 - lane;
    else if (lane >= WAVE - NOP) zz = z0 + WAVE + (WAVE - 1 - lane);
    const bool zv = iz < nz, zhv = zz >= 0 && zz < nz;
    const unsigned vz = zv ? (unsigned)iz * 4u : OOB;
    const unsigned vzh = zhv ? (unsigned)zz * 4u : OOB;
    const bool zi = iz >= NOP && iz < nz - NOP;
    const bool zhi = zz >= NOP && zz < nz - NOP;
    /* sponge factors (tables of ones without a sponge), read at clamped
     * indices so that no load waits on a branch; lanes and rows outside the
     * grid never use theirs */
    const float gz = a.sp1[min(iz, nz - 1)];
    const float gzh = a.sp1[min(max(zz, 0), nz - 1)];
    float gxv[H];
#pragma unroll
    for (int r = 0; r < H; r++) gxv[r] = a.sp2[min(max(x0 - NOP + r, 0), nx - 1)];
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned nbytes = (unsigned)a.ld * (unsigned)nx * 4u;
    const rsrc_t rA1 = mk_rsrc(A1, nbytes), rA2 = mk_rsrc(A2, nbytes), rV = mk_rsrc(a.V, nbytes);
    const rsrc_t rC = mk_rsrc(Cn, nbytes), rA0 = mk_rsrc(A0, nbytes), rG = mk_rsrc(a.gV, nbytes);


    /* the own rows' alpha_{n+2} and dJ/dV first: every load is in flight
     * before the first store (buffer accesses keep program order) */
    float a2v[RX], g0[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const unsigned so = (unsigned)(x0 + j) * rowb;   /* rows >= nx: beyond the range */
        a2v[j] = bld(rA2, vz, so);
        g0[j] = bld(rG, vz, so);
    }

    /* y = V M G alpha_{n+1} and c_n on the RX rows and NOP halo rows each
     * side; rows outside the grid load zeros (offsets beyond the range) */
    float y[H], c[H], yh[RX], chl[RX], t[RX], g[RX];
#pragma unroll
    for (int r = 0; r < H; r++) {
        const int xr = x0 - NOP + r;
        const unsigned so = (unsigned)xr * rowb;
        const bool xi = xr >= NOP && xr < nx - NOP;
        const float gx = gxv[r];
        const float a1 = bld(rA1, vz, so), v = bld(rV, vz, so);
        const float gg = gz * gx;
        const float tt = gg * a1;
        y[r] = (zi && xi) ? v * tt : 0.f;
        c[r] = bld(rC, vz, so);
        if (r >= NOP && r < NOP + RX) {
            const int j = r - NOP;
            t[j] = tt;
            g[j] = gg;
            const float a1h = bld(rA1, vzh, so), vh = bld(rV, vzh, so);
            yh[j] = (zhi && xi) ? vh * ((gzh * gx) * a1h) : 0.f;
            chl[j] = bld(rC, vzh, so);
        }
    }

    /* receiver and injection entries of this tile column, one per lane
     * (64 at a time): the first 64 requested behind the field loads, so
     * that their latency hides under the stencil work */
    const int col = tzi + a.ntz * txi;
    const int rb0 = a.rbeg[col], rn = a.rbeg[col + 1] - rb0;
    const int sb0 = a.sbeg[col], sn = a.sbeg[col + 1] - sb0;
    int rz, rx, sz, sx;
    float ramp, samp;
    rec_entry(a, res, rb0, rn, lane, rz, rx, ramp);
    src_entry(a, n, sb0, sn, lane, sz, sx, samp);

    float al[RX], gv[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int r = NOP + j, ix = x0 + j;
        const bool inter = zi && ix >= NOP && ix < nx - NOP;
        float zm[NOP], zp[NOP];
        znb<NOP>(y[r], yh[j], zm, zp);
        float ly = a.c0 * y[r];
#pragma unroll
        for (int k = 1; k <= NOP; k++)
            ly = ly + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]) + a.cx[k - 1] * (y[r + k] + y[r - k]);
        znb<NOP>(c[r], chl[j], zm, zp);
        float lc = a.c0 * c[r];
#pragma unroll
        for (int k = 1; k <= NOP; k++)
            lc = lc + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]) + a.cx[k - 1] * (c[r + k] + c[r - k]);
        const float tb = g[j] * a2v[j];                 /* t_{n+1} = G alpha_{n+2} */
        al[j] = (inter ? 2.f * t[j] : 0.f) + ly + g[j] * (inter ? -tb : tb);
        gv[j] = inter ? lc * t[j] : 0.f;
    }

    /* receivers (G R^T r_n) and the injection term, entry after entry in
     * bucket order, broadcast from the lane that loaded it */
    for (int q = 0; q < rn; q++) {
        if (q && (q & (WAVE - 1)) == 0) rec_entry(a, res, rb0 + q, rn - q, lane, rz, rx, ramp);
        const int l = q & (WAVE - 1);
        const int qz = __builtin_amdgcn_readlane(rz, l), qx = __builtin_amdgcn_readlane(rx, l);
        const float qa = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ramp), l));
#pragma unroll
        for (int j = 0; j < RX; j++) al[j] = (qz == iz && qx == x0 + j) ? al[j] + g[j] * qa : al[j];
    }
    for (int q = 0; q < sn; q++) {
        if (q && (q & (WAVE - 1)) == 0) src_entry(a, n, sb0 + q, sn - q, lane, sz, sx, samp);
        const int l = q & (WAVE - 1);
        const int qz = __builtin_amdgcn_readlane(sz, l), qx = __builtin_amdgcn_readlane(sx, l);
        const float qa = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, samp), l));
#pragma unroll
        for (int j = 0; j < RX; j++) gv[j] = (qz == iz && qx == x0 + j) ? gv[j] + qa * t[j] : gv[j];
    }

#pragma unroll
    for (int j = 0; j < RX; j++) {
        const unsigned so = (unsigned)(x0 + j) * rowb;   /* rows >= nx: beyond the range */
        bst(rA0, vz, so, al[j]);
        bst(rG, vz, so, g0[j] + gv[j]);
    }
}

template <int NOP>
__global__ __launch_bounds__(NWA * WAVE) void k_adj2d_step(const awefd_adj2d_args a, int n)
{
    adj_body<NOP, RXA>(a, blockIdx.x, n);
}

/* Every shot of a batch in one launch: block b belongs to shot bshot[b] and
 * is that shot's block b - blk0.  Tiles of AWEFD_ADJ_RXB rows per wave (4:
 * a batched gradient 4.5 % faster than with the single-grid kernel's 2; the
 * single-grid loop is faster with 2). */
template <int NOP>
__global__ __launch_bounds__(NWA * WAVE) void k_adj2d_batch(const awefd_adj2d_args *__restrict__ shots,
                                                            const int *__restrict__ bshot, int n)
{
    const int s = __builtin_amdgcn_readfirstlane(bshot[blockIdx.x]);
    const awefd_adj2d_args &a = shots[s];
    adj_body<NOP, AWEFD_ADJ_RXB>(a, (int)blockIdx.x - a.blk0, n);
}

/* Two backward steps n and n-1 of one grid in one pass (k_adj2d_batch2).
 *
 * A block of NW waves covers 64 z (one per lane) x ROWS = NW*RX rows, offset
 * so that neighbouring tiles overlap by 2 NOP in z and x; its outputs are
 * the inner (64 - 2 NOP) x (ROWS - 2 NOP) nodes.  Step n is the single step
 * on every lane and row of the block (each wave's NOP halo rows and the
 * edge lanes' z halo from L2, as adj_body reads them), with the receivers
 * and sources of the block's whole extent; step n-1 needs alpha_n only
 * through t_{n-1} = G alpha_n and V M t_{n-1}, whose x neighbours the waves
 * exchange through LDS and whose z neighbours are lane shifts, so it is
 * exact on the inner nodes.  alpha_{n+1} is read once for both steps (the
 * G alpha_{n+1} of step n-1 is step n's t), and dJ/dV is read and written
 * once: (g + dJ_n) + dJ_{n-1} in the order two single steps add them.  Every
 * product, sum and selection is adj_body's, in its order, so the outputs
 * are those of two single steps bit for bit.  Per node it moves alpha_{n+1},
 * alpha_{n+2}, V, c_n, c_{n-1} and dJ/dV in, alpha_n, alpha_{n-1} and dJ/dV
 * out: 36 bytes for two steps instead of 56. */
template <int NOP, int RX, int NW>
__device__ __forceinline__ void adj2_body(const awefd_adj2d_args &a, int bid, int n)
{
    constexpr int ROWS = NW * RX;
    constexpr int SZ = WAVE - 2 * NOP, SX = ROWS - 2 * NOP;
    constexpr int H = RX + 2 * NOP;
    static_assert(SX > 0, "two-step adjoint tile too narrow for the stencil");
    float *An = a.A[n & 3], *Am = a.A[(n - 1) & 3];
    const float *A1 = a.A[(n + 1) & 3], *A2 = a.A[(n + 2) & 3];
    const int k = n / a.K, n0 = k * a.K;            /* n - 1 >= n0 */
    const float *ckc = a.ck + 2L * k * a.nal;
    const float *Cn = n == n0 ? ckc : a.seg + (long)(n - n0 - 1) * a.nal;
    const float *Cm = n - 1 == n0 ? ckc : a.seg + (long)(n - n0 - 2) * a.nal;
    const float *res = a.res + (long)n * a.nr;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / WAVE);
    const int lb = xcd_remap(bid, a.nblk2);
    const int tzi = lb % a.ntz2, txi = lb / a.ntz2;
    const int nz = a.nz, nx = a.nx;
    const int z0 = tzi * SZ - NOP, iz = z0 + lane;
    const int bx0 = txi * SX - NOP;                 /* the block's first row */
    const int x0 = bx0 + wv * RX;                   /* wave-uniform          */
    int zz = -1;
    if (lane < NOP) zz = z0 - 1 - lane;
    else if (lane >= WAVE - NOP) zz = z0 + WAVE + (WAVE - 1 - lane);
    const bool zv = iz >= 0 && iz < nz, zhv = zz >= 0 && zz < nz;
    const unsigned vz = zv ? (unsigned)iz * 4u : OOB;
    const unsigned vzh = zhv ? (unsigned)zz * 4u : OOB;
    /* the lanes whose nodes are outputs of the tile */
    const unsigned vzo = (lane >= NOP && lane < WAVE - NOP) ? vz : OOB;
    const bool zi = iz >= NOP && iz < nz - NOP;
    const bool zhi = zz >= NOP && zz < nz - NOP;
    const float gz = a.sp1[min(max(iz, 0), nz - 1)];
    const float gzh = a.sp1[min(max(zz, 0), nz - 1)];
    float gxv[H];
#pragma unroll
    for (int r = 0; r < H; r++) gxv[r] = a.sp2[min(max(x0 - NOP + r, 0), nx - 1)];
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned nbytes = (unsigned)a.ld * (unsigned)nx * 4u;
    const rsrc_t rA1 = mk_rsrc(A1, nbytes), rA2 = mk_rsrc(A2, nbytes), rV = mk_rsrc(a.V, nbytes);
    const rsrc_t rC = mk_rsrc(Cn, nbytes), rCm = mk_rsrc(Cm, nbytes), rG = mk_rsrc(a.gV, nbytes);
    const rsrc_t rAn = mk_rsrc(An, nbytes), rAm = mk_rsrc(Am, nbytes);

    /* alpha_{n+2} on every node of the wave's rows; dJ/dV on the outputs
     * only (rows outside the grid: offsets beyond the range) */
    float a2v[RX], g0[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int r = wv * RX + j;
        const unsigned so = (unsigned)(x0 + j) * rowb;
        a2v[j] = bld(rA2, vz, so);
        g0[j] = bld(rG, (r >= NOP && r < ROWS - NOP) ? vzo : OOB, so);
    }

    /* step n: y = V M G alpha_{n+1} and c_n on the RX rows and NOP halo rows
     * each side, as adj_body; c_{n-1} on the same rows for step n-1 */
    float y[H], c[H], cm[H], yh[RX], chl[RX], t[RX], g[RX], vo[RX];
#pragma unroll
    for (int r = 0; r < H; r++) {
        const int xr = x0 - NOP + r;
        const unsigned so = (unsigned)xr * rowb;
        const bool xi = xr >= NOP && xr < nx - NOP;
        const float gx = gxv[r];
        const float a1 = bld(rA1, vz, so), v = bld(rV, vz, so);
        const float gg = gz * gx;
        const float tt = gg * a1;
        y[r] = (zi && xi) ? v * tt : 0.f;
        c[r] = bld(rC, vz, so);
        cm[r] = bld(rCm, vz, so);
        if (r >= NOP && r < NOP + RX) {
            const int j = r - NOP;
            t[j] = tt;
            g[j] = gg;
            vo[j] = v;
            const float a1h = bld(rA1, vzh, so), vh = bld(rV, vzh, so);
            yh[j] = (zhi && xi) ? vh * ((gzh * gx) * a1h) : 0.f;
            chl[j] = bld(rC, vzh, so);
        }
    }

    const int col = tzi + a.ntz2 * txi;
    const int rb0 = a.rbeg2[col], rn = a.rbeg2[col + 1] - rb0;
    const int sb0 = a.sbeg2[col], sn = a.sbeg2[col + 1] - sb0;
    /* the first 64 receiver and injection entries of the tile, one per
     * lane, requested behind the field loads */
    int rz = -1, rx = -1, sz = -1, sx = -1;
    float ramp = 0.f, samp = 0.f;
    if (lane < rn) {
        const awefd_inj en = a.rinj2[rb0 + lane];
        rz = en.z;
        rx = en.x;
        ramp = res[en.src] * en.w;
    }
    if (lane < sn) {
        const awefd_inj en = a.sinj2[sb0 + lane];
        sz = en.z;
        sx = en.x;
        samp = a.wav[(long)n * a.wstride + en.src] * en.w;
    }

    float al[RX], gv[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int r = NOP + j, ix = x0 + j;
        const bool inter = zi && ix >= NOP && ix < nx - NOP;
        float zm[NOP], zp[NOP];
        znb<NOP>(y[r], yh[j], zm, zp);
        float ly = a.c0 * y[r];
#pragma unroll
        for (int k = 1; k <= NOP; k++)
            ly = ly + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]) + a.cx[k - 1] * (y[r + k] + y[r - k]);
        znb<NOP>(c[r], chl[j], zm, zp);
        float lc = a.c0 * c[r];
#pragma unroll
        for (int k = 1; k <= NOP; k++)
            lc = lc + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]) + a.cx[k - 1] * (c[r + k] + c[r - k]);
        const float tb = g[j] * a2v[j];                 /* t_{n+1} = G alpha_{n+2} */
        al[j] = (inter ? 2.f * t[j] : 0.f) + ly + g[j] * (inter ? -tb : tb);
        gv[j] = inter ? lc * t[j] : 0.f;
    }
    for (int q = 0; q < rn; q++) {
        if (q && (q & (WAVE - 1)) == 0) {
            rz = rx = -1;
            ramp = 0.f;
            if (lane < rn - q) {
                const awefd_inj en = a.rinj2[rb0 + q + lane];
                rz = en.z;
                rx = en.x;
                ramp = res[en.src] * en.w;
            }
        }
        const int l = q & (WAVE - 1);
        const int qz = __builtin_amdgcn_readlane(rz, l), qx = __builtin_amdgcn_readlane(rx, l);
        const float qa = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ramp), l));
#pragma unroll
        for (int j = 0; j < RX; j++) al[j] = (qz == iz && qx == x0 + j) ? al[j] + g[j] * qa : al[j];
    }
    for (int q = 0; q < sn; q++) {
        if (q && (q & (WAVE - 1)) == 0) {
            sz = sx = -1;
            samp = 0.f;
            if (lane < sn - q) {
                const awefd_inj en = a.sinj2[sb0 + q + lane];
                sz = en.z;
                sx = en.x;
                samp = a.wav[(long)n * a.wstride + en.src] * en.w;
            }
        }
        const int l = q & (WAVE - 1);
        const int qz = __builtin_amdgcn_readlane(sz, l), qx = __builtin_amdgcn_readlane(sx, l);
        const float qa = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, samp), l));
#pragma unroll
        for (int j = 0; j < RX; j++) gv[j] = (qz == iz && qx == x0 + j) ? gv[j] + qa * t[j] : gv[j];
    }

    /* step n-1: y' = V M G alpha_n on the wave's rows, its x neighbours in
     * the other waves' rows through LDS */
    __shared__ float ys[ROWS][WAVE];
    float t2[RX], y2[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int ix = x0 + j;
        const bool xi = ix >= NOP && ix < nx - NOP;
        t2[j] = g[j] * al[j];
        y2[j] = (zi && xi) ? vo[j] * t2[j] : 0.f;
        ys[wv * RX + j][lane] = y2[j];
    }
    __syncthreads();
    const float *resm = res - a.nr;
    float al2[RX], gv2[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int r = NOP + j, rw = wv * RX + j, ix = x0 + j;
        const bool inter = zi && ix >= NOP && ix < nx - NOP;
        float zm[NOP], zp[NOP];
        znb<NOP>(y2[j], 0.f, zm, zp);                   /* edge lanes are not outputs */
        float ly = a.c0 * y2[j];
#pragma unroll
        for (int k = 1; k <= NOP; k++) {
            const float yp = j + k < RX ? y2[j + k] : ys[min(rw + k, ROWS - 1)][lane];
            const float ym = j - k >= 0 ? y2[j - k] : ys[max(rw - k, 0)][lane];
            ly = ly + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]) + a.cx[k - 1] * (yp + ym);
        }
        znb<NOP>(cm[r], 0.f, zm, zp);
        float lc = a.c0 * cm[r];
#pragma unroll
        for (int k = 1; k <= NOP; k++)
            lc = lc + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]) + a.cx[k - 1] * (cm[r + k] + cm[r - k]);
        const float tb = t[j];                          /* G alpha_{n+1}, as step n computed it */
        al2[j] = (inter ? 2.f * t2[j] : 0.f) + ly + g[j] * (inter ? -tb : tb);
        gv2[j] = inter ? lc * t2[j] : 0.f;
    }
    for (int q0 = 0; q0 < rn; q0 += WAVE) {
        rz = rx = -1;
        ramp = 0.f;
        if (lane < rn - q0) {
            const awefd_inj en = a.rinj2[rb0 + q0 + lane];
            rz = en.z;
            rx = en.x;
            ramp = resm[en.src] * en.w;
        }
        const int qn = min(rn - q0, WAVE);
        for (int l = 0; l < qn; l++) {
            const int qz = __builtin_amdgcn_readlane(rz, l), qx = __builtin_amdgcn_readlane(rx, l);
            const float qa = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ramp), l));
#pragma unroll
            for (int j = 0; j < RX; j++) al2[j] = (qz == iz && qx == x0 + j) ? al2[j] + g[j] * qa : al2[j];
        }
    }
    for (int q0 = 0; q0 < sn; q0 += WAVE) {
        sz = sx = -1;
        samp = 0.f;
        if (lane < sn - q0) {
            const awefd_inj en = a.sinj2[sb0 + q0 + lane];
            sz = en.z;
            sx = en.x;
            samp = a.wav[(long)(n - 1) * a.wstride + en.src] * en.w;
        }
        const int qn = min(sn - q0, WAVE);
        for (int l = 0; l < qn; l++) {
            const int qz = __builtin_amdgcn_readlane(sz, l), qx = __builtin_amdgcn_readlane(sx, l);
            const float qa = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, samp), l));
#pragma unroll
            for (int j = 0; j < RX; j++) gv2[j] = (qz == iz && qx == x0 + j) ? gv2[j] + qa * t2[j] : gv2[j];
        }
    }

    /* the tile's inner nodes: alpha_n, alpha_{n-1}, dJ/dV */
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int r = wv * RX + j;
        if (r < NOP || r >= ROWS - NOP) continue;       /* wave-uniform */
        const unsigned so = (unsigned)(x0 + j) * rowb;  /* rows >= nx: beyond the range */
        bst(rAn, vzo, so, al[j]);
        bst(rAm, vzo, so, al2[j]);
        bst(rG, vzo, so, (g0[j] + gv[j]) + gv2[j]);
    }
}

template <int NOP>
__global__ __launch_bounds__(AWEFD_ADJ2_NWAVE * WAVE) void k_adj2d_batch2(const awefd_adj2d_args *__restrict__ shots,
                                                                          const int *__restrict__ bshot, int n)
{
    const int s = __builtin_amdgcn_readfirstlane(bshot[blockIdx.x]);
    const awefd_adj2d_args &a = shots[s];
    adj2_body<NOP, AWEFD_ADJ2_RX, AWEFD_ADJ2_NWAVE>(a, (int)blockIdx.x - a.blk02, n);
}

/* dJ/ds on the model grid, summed over the shots in shot order: for each
 * shot, the sum over the padded nodes copied from model node (kz, kx) of
 * dJ/dV * dV/ds, dV/ds = -V v^2 = -V (V / dt^2) (0 where a free surface
 * fixed V = 0).  All shots share the grid.  An edge node has nb + 1 copies
 * per side (a corner (nb + 1)^2), added x outer, z inner.  A thread takes SG
 * shots at a time, one accumulator each, and issues the loads of 8 copies of
 * all SG shots before their adds, which keep each shot's order; the shots'
 * sums are then added in shot order. */
template <int SG>
__global__ __launch_bounds__(256) void k_adj2d_chain(const awefd_adj2d_args *__restrict__ shots, int nshot,
                                                     float *gs, int mz, int mx, int nz, int nx, int ld, int nb)
{
    constexpr int B = 8;
    const int kz = blockIdx.x * 64 + (threadIdx.x & 63);
    const int kx = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (kz >= mz || kx >= mx) return;
    const int z0 = kz == 0 ? 0 : kz + nb, z1 = kz == mz - 1 ? nz : kz + nb + 1;
    const int x0 = kx == 0 ? 0 : kx + nb, x1 = kx == mx - 1 ? nx : kx + nb + 1;
    const int ncopy = (z1 - z0) * (x1 - x0);
    float tot = 0.f;
    for (int s0 = 0; s0 < nshot; s0 += SG) {
        const int ns = min(SG, nshot - s0);
        const float *V[SG], *gV[SG];
        float idt2[SG], acc[SG];
#pragma unroll
        for (int q = 0; q < SG; q++) {
            const awefd_adj2d_args &a = shots[s0 + (q < ns ? q : 0)];
            V[q] = a.V;
            gV[q] = a.gV;
            idt2[q] = a.idt2;
            acc[q] = 0.f;
        }
        int ix = x0, iz = z0, j = 0;
        for (; j + B <= ncopy; j += B) {
            long idx[B];
#pragma unroll
            for (int k = 0; k < B; k++) {
                idx[k] = (long)ix * ld + iz;
                if (++iz == z1) {
                    iz = z0;
                    ix++;
                }
            }
            float v[SG][B], g[SG][B];
#pragma unroll
            for (int q = 0; q < SG; q++)
                if (q < ns) {
#pragma unroll
                    for (int k = 0; k < B; k++) {
                        v[q][k] = V[q][idx[k]];
                        g[q][k] = gV[q][idx[k]];
                    }
                }
#pragma unroll
            for (int q = 0; q < SG; q++)
                if (q < ns) {
#pragma unroll
                    for (int k = 0; k < B; k++) acc[q] += g[q][k] * (-v[q][k] * (v[q][k] * idt2[q]));
                }
        }
        for (; j < ncopy; j++) {
            const long i = (long)ix * ld + iz;
            float v[SG], g[SG];
#pragma unroll
            for (int q = 0; q < SG; q++)
                if (q < ns) {
                    v[q] = V[q][i];
                    g[q] = gV[q][i];
                }
#pragma unroll
            for (int q = 0; q < SG; q++)
                if (q < ns) acc[q] += g[q] * (-v[q] * (v[q] * idt2[q]));
            if (++iz == z1) {
                iz = z0;
                ix++;
            }
        }
#pragma unroll
        for (int q = 0; q < SG; q++)
            if (q < ns) tot = s0 + q == 0 ? acc[q] : tot + acc[q];
    }
    gs[(long)kx * mz + kz] = tot;
}

/* Table of copies (or zero fills): block b copies chunk b - cbeg[j] of job j,
 * 1024 floats, 4 per thread. */
__global__ __launch_bounds__(256) void k_copy_jobs(const awefd_copy_job *__restrict__ jobs,
                                                   const long *__restrict__ cbeg, int njob)
{
    const long b = blockIdx.x;
    int lo = 0, hi = njob - 1;
    while (lo < hi) {                    /* last job with cbeg[j] <= b */
        const int mid = (lo + hi + 1) >> 1;
        if (cbeg[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const awefd_copy_job jb = jobs[lo];
    const long base = (b - cbeg[lo]) * 1024;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const long i = base + q * 256 + threadIdx.x;
        if (i < jb.n) jb.dst[i] = jb.src ? jb.src[i] : 0.f;
    }
}

template <int NOP>
int launch_step(const awefd_adj2d_args *a, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj2d_step<NOP>, dim3(a->nblk), dim3(NWA * WAVE), 0, s, *a, n);
    return (int)hipGetLastError();
}

template <int NOP>
int launch_batch(const awefd_adj2d_args *shots, const int *bshot, int nblk, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj2d_batch<NOP>, dim3(nblk), dim3(NWA * WAVE), 0, s, shots, bshot, n);
    return (int)hipGetLastError();
}

template <int NOP>
int launch_batch2(const awefd_adj2d_args *shots, const int *bshot, int nblk, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj2d_batch2<NOP>, dim3(nblk), dim3(AWEFD_ADJ2_NWAVE * WAVE), 0, s, shots, bshot, n);
    return (int)hipGetLastError();
}

} // namespace

extern "C" int awefd_launch_adj2d_batch2(const awefd_adj2d_args *shots, const int *bshot, int nblk, int n, int nop,
                                         void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (n < 1) return (int)hipErrorInvalidValue;
    switch (nop) {
    case 1: return launch_batch2<1>(shots, bshot, nblk, n, s);
    case 2: return launch_batch2<2>(shots, bshot, nblk, n, s);
    case 3: return launch_batch2<3>(shots, bshot, nblk, n, s);
    case 4: return launch_batch2<4>(shots, bshot, nblk, n, s);
    case 5: return launch_batch2<5>(shots, bshot, nblk, n, s);
    case 6: return launch_batch2<6>(shots, bshot, nblk, n, s);
    default: return (int)hipErrorInvalidValue;
    }
}

extern "C" int awefd_launch_adj2d_step(const awefd_adj2d_args *a, int n, int nop, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (a->nblk != a->ntz * a->ntx || a->ntz * 64 < a->nz || a->ntx * AWEFD_ADJ_TX < a->nx || a->K < 1)
        return (int)hipErrorInvalidValue;
    switch (nop) {
    case 1: return launch_step<1>(a, n, s);
    case 2: return launch_step<2>(a, n, s);
    case 3: return launch_step<3>(a, n, s);
    case 4: return launch_step<4>(a, n, s);
    case 5: return launch_step<5>(a, n, s);
    case 6: return launch_step<6>(a, n, s);
    default: return (int)hipErrorInvalidValue;
    }
}

extern "C" int awefd_launch_adj2d_batch(const awefd_adj2d_args *shots, const int *bshot, int nblk, int n, int nop,
                                        void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    switch (nop) {
    case 1: return launch_batch<1>(shots, bshot, nblk, n, s);
    case 2: return launch_batch<2>(shots, bshot, nblk, n, s);
    case 3: return launch_batch<3>(shots, bshot, nblk, n, s);
    case 4: return launch_batch<4>(shots, bshot, nblk, n, s);
    case 5: return launch_batch<5>(shots, bshot, nblk, n, s);
    case 6: return launch_batch<6>(shots, bshot, nblk, n, s);
    default: return (int)hipErrorInvalidValue;
    }
}

extern "C" int awefd_launch_adj2d_chain(const awefd_adj2d_args *h, const awefd_adj2d_args *shots, int nshot,
                                        float *gout, void *stream)
{
    const dim3 g((h->mz + 63) / 64, (h->mx + 3) / 4);
    if (nshot >= 4)
        hipLaunchKernelGGL(k_adj2d_chain<4>, g, dim3(256), 0, (hipStream_t)stream, shots, nshot, gout, h->mz, h->mx,
                           h->nz, h->nx, h->ld, h->nb);
    else
        hipLaunchKernelGGL(k_adj2d_chain<1>, g, dim3(256), 0, (hipStream_t)stream, shots, nshot, gout, h->mz, h->mx,
                           h->nz, h->nx, h->ld, h->nb);
    return (int)hipGetLastError();
}

extern "C" int awefd_launch_copy_jobs(const awefd_copy_job *jobs, const long *cbeg, int njob, long nchunk,
                                      void *stream)
{
    if (njob < 1 || nchunk < 1 || nchunk > 0x7fffffffL) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy_jobs, dim3((unsigned)nchunk), dim3(256), 0, (hipStream_t)stream, jobs, cbeg, njob);
    return (int)hipGetLastError();
}
