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
 * Layout: z fastest (RSF axis 1), then x, then y.  Rows are padded to ld
 * nodes, a multiple of 32 (128 bytes), so that a z tile of 64 nodes is two
 * whole cache lines: no line is shared between tiles.
 *
 * Step kernel: NWAVE waves per block.  Lane l of a wave holds z = z0 + l of
 * RX x rows; a block covers 64 z x NWAVE*RX x nodes and streams them along y
 * (2.5-D blocking).  z neighbours are moved across lanes by DPP wave shifts,
 * the 2*NOP values beyond the tile edges arriving with one masked load per
 * row; x neighbours come from the wave's own rows or from loads of the
 * adjacent rows that hit L1/L2; y neighbours sit in a register ring.  Each
 * field value is read from HBM once per step.
 *
 * Addressing: AMD buffer instructions.  The per-lane part of an offset (z)
 * is loop-invariant; the row/plane part is a scalar offset.  Lanes or rows
 * that must not touch memory get offsets beyond the descriptor's range,
 * which the hardware turns into zero loads and dropped stores.
 *
 * Compiled with -ffp-contract=off: no FMA contraction, so every add and
 * multiply rounds as in the reference's SSE (or scalar) build.
 */
#include <hip/hip_runtime.h>
#include "awefd_hip.h"
#include "awefd_internal.h"

#pragma clang fp contract(off)

namespace {

using namespace awefd;
constexpr int NWAVE0 = AWEFD_NWAVE;   /* waves per block of the single-grid kernel */
constexpr int NTHR0 = NWAVE0 * WAVE;
constexpr int RX0 = AWEFD_RX;         /* x rows per wave of the single-grid kernel */
/* cache policy of the streamed arrays (previous field, (v dt)^2, stores):
 * non-temporal, so that L2 keeps the current field's rows for their reuse
 * as halos */
#ifndef AWEFD_NT
#define AWEFD_NT 2
#endif
constexpr int NT = AWEFD_NT;
static_assert(NWAVE0 * RX0 == AWEFD_CART_TX, "tile width");
/* tile of the single-grid step kernel: the narrow 2-D one, or 4 waves */
template <bool NARROW>
constexpr int step_nw() { return NARROW ? AWEFD_NWAVE2 : NWAVE0; }
template <bool NARROW>
constexpr int step_rx() { return NARROW ? AWEFD_RX2 : RX0; }
static_assert(AWEFD_NWAVE2 * AWEFD_RX2 == AWEFD_CART_TX2, "2-D tile width");

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
 * field with its pending sponge factor; with sinc taps sinc2d_extract /
 * sinc3d_extract (fdutil.c:1015-1051,1301-1332): taps y, x, z in order,
 * g += ((u wy) wx) wz (2-D (u wx) wz) from g = 0. */
template <int NDIM>
__device__ void cart_extract(const awefd_cart_args &a, const float *C, float *trow, int r)
{
    if (r >= a.nr) return;
    const long sx = a.ld, sy = (long)a.ld * a.n2;
    if (a.rsinc) {
        const awefd_sinc_pt &p = a.rsinc[r];
        float g = 0.f;
        for (int ty = p.f[2]; ty < p.f[2] + p.n[2]; ty++) {
            const int y = p.c[2] - 4 + ty;
            for (int tx = p.f[1]; tx < p.f[1] + p.n[1]; tx++) {
                const int x = p.c[1] - 4 + tx;
                for (int tz = p.f[0]; tz < p.f[0] + p.n[0]; tz++) {
                    const int z = p.c[0] - 4 + tz;
                    float u = C[y * sy + x * sx + z];
                    if (a.dabc) u = sponge<NDIM>(u, z, x, y, a);
                    g = NDIM == 3 ? g + u * p.w[2][ty] * p.w[1][tx] * p.w[0][tz] : g + u * p.w[1][tx] * p.w[0][tz];
                }
            }
        }
        trow[r] = g;
        return;
    }
    const int jz = a.rj[3 * r], jx = a.rj[3 * r + 1], jy = a.rj[3 * r + 2];
    constexpr int NC = NDIM == 3 ? 8 : 4;
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < NC; q++) {
        const int z = jz + (q & 1), x = jx + ((q >> 1) & 1), y = NDIM == 3 ? jy + ((q >> 2) & 1) : 0;
        float u = C[y * sy + x * sx + z];
        if (a.dabc) u = sponge<NDIM>(u, z, x, y, a);
        const float t = u * a.rw[8 * r + q];
        s = q == 0 ? t : s + t;
    }
    trow[r] = s;
}

/* x neighbour at offset `off` of row j of a wave: its own rows, else the
 * halo rows (h < NOP: x0-NOP+h, h >= NOP: x0+RX+h-NOP).  j and off are
 * compile-time constants after unrolling. */
template <int NOP, int RX>
__device__ __forceinline__ float xnb(const float (&cen)[RX], const float (&hal)[2 * NOP], int j, int off)
{
    const int r = j + off;
    if (r >= 0 && r < RX) return cen[r];
    return r < 0 ? hal[NOP + r] : hal[NOP + r - RX];
}

/* byte offset of row x at plane y relative to the block window (wave-uniform;
 * rows/planes below the window wrap to offsets beyond its range) */
__device__ __forceinline__ unsigned soff(int x, int y, int ybase, unsigned rowb, unsigned psz)
{
    return (unsigned)x * rowb + (unsigned)(y - ybase) * psz;
}

/* x halo row h of a wave (h < NOP: x0-NOP+h, h >= NOP: x0+RX+h-NOP) lies
 * in the block's own rows [bx0, bx0 + TX): then it comes from LDS */
template <int NOP, int RX>
__device__ __forceinline__ bool halo_in_block(int h, int x0, int bx0, int tx)
{
    const int hx = h < NOP ? x0 - NOP + h : x0 + RX + (h - NOP);
    return hx >= bx0 && hx < bx0 + tx;
}

/* Loads of one plane: previous field, (v dt)^2 and the z-halo value of the
 * wave's rows, and the x-halo rows outside the block (the others are
 * exchanged through LDS; their loads get an out-of-range offset). */
template <int NOP, int RX>
__device__ __forceinline__ void load_pvh(float (&p)[RX], float (&v)[RX], float (&zh)[RX],
                                         float (&hal)[2 * NOP], rsrc_t rP, rsrc_t rV, rsrc_t rC,
                                         unsigned vz, unsigned vzh, int x0, int bx0, int tx, int y, int ybase,
                                         unsigned rowb, unsigned psz)
{
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const unsigned s = soff(x0 + j, y, ybase, rowb, psz);
        p[j] = bld<NT>(rP, vz, s);
        v[j] = bld<NT>(rV, vz, s);
        zh[j] = bld(rC, vzh, s);
    }
#pragma unroll
    for (int h = 0; h < 2 * NOP; h++) {
        const int hx = h < NOP ? x0 - NOP + h : x0 + RX + (h - NOP);
        hal[h] = bld(rC, halo_in_block<NOP, RX>(h, x0, bx0, tx) ? OOB : vz, soff(hx, y, ybase, rowb, psz));
    }
}

/* Variable-density term of ts_kernel.c:101-135 (2-D) and 228-280 (3-D):
 * lap - (du . drho) / rho, du from the stencil's own neighbour values, drho
 * precomputed per node (g = {rho, drho_z, drho_x, drho_y}). */
template <int NOP, int NDIM, bool SSE, int QN, int RX>
__device__ __forceinline__ float den_term(float lap, const float (&zp)[NOP], const float (&zm)[NOP],
                                          const float (&cen)[RX], const float (&hal)[2 * NOP],
                                          const float (&qj)[QN], int S0, int j, float4 g,
                                          const awefd_cart_args &a)
{
    float duz = 0.f, dux = 0.f, duy = 0.f;
#pragma unroll
    for (int k = 1; k <= NOP; k++) {
        duz = duz + (zp[k - 1] - zm[k - 1]) * a.bz[k - 1];
        dux = dux + (xnb<NOP, RX>(cen, hal, j, k) - xnb<NOP, RX>(cen, hal, j, -k)) * a.bx[k - 1];
        if (NDIM == 3) duy = duy + (qj[(S0 + NOP + k) % QN] - qj[(S0 + NOP - k) % QN]) * a.by[k - 1];
    }
    float dd;
    if (NDIM == 2) dd = duz * g.y + dux * g.z;
    else if (SSE) dd = duz * g.y + (dux * g.z + duy * g.w);
    else dd = duz * g.y + dux * g.z + duy * g.w;
    return lap - dd / g.x;
}

/* Density record of a node (variable-density runs): rho and its three
 * time-invariant first differences, four adjacent dwords (the b128 buffer
 * builtin of this compiler loads only the first dword, so the record is read
 * as four 32-bit loads, which the compiler merges). */
__device__ __forceinline__ float4 bld4(rsrc_t r, unsigned voff, unsigned soff)
{
    return make_float4(bld<NT>(r, voff, soff), bld<NT>(r, voff + 4u, soff), bld<NT>(r, voff + 8u, soff),
                       bld<NT>(r, voff + 12u, soff));
}

template <bool VD, int RX>
__device__ __forceinline__ void load_den(float4 (&g)[RX], rsrc_t rR, unsigned vz4, int x0, int y, int ybase,
                                         unsigned rowb, unsigned psz)
{
    if (!VD) return;
#pragma unroll
    for (int j = 0; j < RX; j++) g[j] = bld4(rR, vz4, 4u * soff(x0 + j, y, ybase, rowb, psz));
}

/* The step kernel.  Planes of the current field live in a ring of
 * QN = 2*NOP+2 register slots; the previous field, (v dt)^2 and the halo
 * values of the next plane in the second of two register buffers; the y loop
 * is unrolled QN times so that every slot index is a compile-time constant.
 * The loads of plane y+1 are in flight while plane y is computed.  No LDS and
 * no barriers: waves advance independently. */
template <int NOP, int NDIM, bool SSE, bool VD, int RX, int NW>
__device__ __forceinline__ void cart_step_body(const awefd_cart_args &a, float *Pout, const float *Cin, int it,
                                               float *trow, int bid, float *Pdst)
{
    constexpr int QN = 2 * NOP + 2;
    constexpr int TX = NW * RX;
    const int tid = threadIdx.x;
    if (bid >= a.nblk) {
        cart_extract<NDIM>(a, Cin, trow, (bid - a.nblk) * (NW * WAVE) + tid);
        return;
    }
    const int lb = xcd_remap(bid, a.nblk);
    const int tzi = lb % a.ntz;
    const int rr = lb / a.ntz;
    const int txi = rr % a.ntx;
    const int tyi = rr / a.ntx;
    const int lane = tid & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / WAVE);
    const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
    const int z0 = tzi * WAVE;
    const int iz = z0 + lane;
    const bool zin = iz < n1;
    const int x0 = txi * TX + wv * RX;                   /* wave-uniform */
    /* y chunks of balanced length (nty chunks over n3 planes) */
    const int y0 = NDIM == 3 ? (int)((long)tyi * n3 / a.nty) : 0;
    const int y1 = NDIM == 3 ? (int)((long)(tyi + 1) * n3 / a.nty) : 1;
    const long sy = (long)a.ld * n2;
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned psz = (unsigned)(sy * 4);

    /* this block's window of planes [ybase, ytop), < 1 GiB (host-checked) */
    const int ybase = NDIM == 3 ? max(y0 - NOP, 0) : 0;
    const int ytop = NDIM == 3 ? min(y1 + NOP, n3) : 1;
    const unsigned nbytes = (unsigned)(ytop - ybase) * psz;
    const rsrc_t rC = mk_rsrc(Cin + ybase * sy, nbytes);
    const rsrc_t rP = mk_rsrc(Pout + ybase * sy, nbytes);
    const rsrc_t rN = mk_rsrc(Pdst + ybase * sy, nbytes);
    const rsrc_t rV = mk_rsrc(a.vel + ybase * sy, nbytes);
    /* density records: 4 floats per node, same layout (window < 4 GiB) */
    const rsrc_t rR = mk_rsrc(VD ? (const void *)(a.rho + 4 * ybase * sy) : (const void *)a.vel,
                              VD ? 4u * nbytes : 0u);

    /* per-lane offsets: the node itself, and the z value beyond the tile edge
     * this lane fetches (lane k < NOP: z0-1-k; lane 63-k: z0+64+k) */
    const unsigned vz = zin ? (unsigned)iz * 4u : OOB;
    const unsigned vz4 = zin ? (unsigned)iz * 16u : OOB;   /* density records */
    int zz = -1;
    if (lane < NOP) zz = z0 - 1 - lane;
    else if (lane >= WAVE - NOP) zz = z0 + WAVE + (WAVE - 1 - lane);
    const unsigned vzh = (zz >= 0 && zz < n1) ? (unsigned)zz * 4u : OOB;

    /* per-lane constants */
    const bool zint = iz >= NOP && iz < n1 - NOP;
    const bool zsp = a.dabc && zin && (iz < a.nb || iz >= n1 - a.nb);
    const float spz = zsp ? a.sp1[iz] : 1.f;
    const bool zsp_w = a.dabc && (z0 < a.nb || z0 + WAVE > n1 - a.nb);   /* wave-uniform */
    const bool zdef = a.zsep && iz >= n1 - NOP - 2;                      /* z-high shell */
    /* per-row (wave-uniform) constants, as bit masks over the RX rows */
    float spx[RX];
    unsigned xint_m = 0, xdef_m = 0, xin_m = 0;
    bool xsp_w = false;
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int ix = x0 + j;
        const bool xsp = a.dabc && ix < n2 && (ix < a.nb || ix >= n2 - a.nb);
        xsp_w = xsp_w || xsp;
        spx[j] = xsp ? a.sp2[ix] : 1.f;
        xin_m |= ix < n2 ? 1u << j : 0u;
        xint_m |= (ix >= NOP && ix < n2 - NOP) ? 1u << j : 0u;
        xdef_m |= (NDIM == 3 && a.abc && (ix <= NOP + 1 || ix >= n2 - NOP - 2)) ? 1u << j : 0u;
    }
    const bool sp_xz = zsp_w || xsp_w;

    /* ring: current-field plane p in slot (p - y0 + NOP) % QN */
    float q[RX][QN];
    float bp_[2][RX], bv_[2][RX], bz_[2][RX], bh_[2][2 * NOP];
    float4 bg_[2][RX];
    if (NDIM == 3) {
#pragma unroll
        for (int k = 0; k <= 2 * NOP; k++)
#pragma unroll
            for (int j = 0; j < RX; j++) q[j][k] = bld(rC, vz, soff(x0 + j, y0 - NOP + k, ybase, rowb, psz));
    } else {
#pragma unroll
        for (int j = 0; j < RX; j++) q[j][0] = bld(rC, vz, soff(x0 + j, 0, 0, rowb, psz));
    }
    const int bx0 = txi * TX;                            /* the block's first row */
    load_pvh<NOP, RX>(bp_[0], bv_[0], bz_[0], bh_[0], rP, rV, rC, vz, vzh, x0, bx0, TX, y0, ybase, rowb, psz);
    /* the centre plane's rows of every wave of the block, for the x
     * neighbours inside the block (two planes, one barrier per plane) */
    __shared__ float xrow[2][TX][WAVE];
    load_den<VD, RX>(bg_[0], rR, vz4, x0, y0, ybase, rowb, psz);

    /* injection entries of this tile column, restricted to planes >= y0 */
    const int col = tzi + a.ntz * txi;
    int ib = a.ibeg[col];
    const int ie = a.ibeg[col + 1];
    while (ib < ie && a.inj[ib].y < y0) ib++;

    const bool zlow = NDIM == 3 && a.abc && tzi == 0 && !a.fsrf;
    const bool zhigh = NDIM == 3 && a.abc && !a.zsep && tzi == a.ntz - 1;
    /* the z-face one-way coefficients of the block's rows and planes, staged
     * in LDS when the chunk has at most ZB_PL planes: read from memory in the
     * plane loop, each was a load the recurrence waited on (6 % of a cart3d
     * step).  The loop's first barrier orders these writes before the reads. */
    constexpr int ZB_PL = 32;
    __shared__ float zbt[NDIM == 3 ? 2 : 1][NDIM == 3 ? ZB_PL : 1][TX];
    const bool zbs = NDIM == 3 && (zlow || zhigh) && y1 - y0 <= ZB_PL;   /* block-uniform */
    if (NDIM == 3 && zbs) {
        for (int i = tid; i < (y1 - y0) * TX; i += NW * WAVE) {
            const int yy = i / TX, xx = bx0 + i % TX;
            const long o = (long)(y0 + yy) * n2 + min(xx, n2 - 1);
            zbt[0][yy][i % TX] = zlow ? a.bzl[o] : 0.f;
            zbt[1][yy][i % TX] = zhigh ? a.bzh[o] : 0.f;
        }
    }

    for (int yg = y0; yg < y1; yg += QN) {
#pragma unroll
        for (int R = 0; R < QN; R++) {
            const int y = yg + R;
            if (y >= y1) break;
            const int S0 = NDIM == 3 ? R : 0;     /* slot of plane y-NOP */
            const int B = R & 1;                  /* buffer of plane y   */
            /* loads of plane y+1: head y+1+NOP into the slot of plane
             * y-NOP-1, previous field / vel / halos into the other buffer */
            if (y + 1 < y1) {
                if (NDIM == 3) {
#pragma unroll
                    for (int j = 0; j < RX; j++)
                        q[j][(S0 + 2 * NOP + 1) % QN] = bld(rC, vz, soff(x0 + j, y + 1 + NOP, ybase, rowb, psz));
                }
                load_pvh<NOP, RX>(bp_[B ^ 1], bv_[B ^ 1], bz_[B ^ 1], bh_[B ^ 1], rP, rV, rC, vz, vzh, x0, bx0, TX,
                              y + 1, ybase, rowb, psz);
                load_den<VD, RX>(bg_[B ^ 1], rR, vz4, x0, y + 1, ybase, rowb, psz);
            }
            float cen[RX];
#pragma unroll
            for (int j = 0; j < RX; j++) cen[j] = NDIM == 3 ? q[j][(S0 + NOP) % QN] : q[j][0];
            /* x halo rows inside the block from the other waves' centre rows */
#pragma unroll
            for (int j = 0; j < RX; j++) xrow[B][wv * RX + j][lane] = cen[j];
            __syncthreads();
            float hal[2 * NOP];
#pragma unroll
            for (int h = 0; h < 2 * NOP; h++) {
                const int hr = h < NOP ? wv * RX - NOP + h : wv * RX + RX + (h - NOP);
                const float l = xrow[B][min(max(hr, 0), TX - 1)][lane];
                hal[h] = halo_in_block<NOP, RX>(h, x0, bx0, TX) ? l : bh_[B][h];
            }

            const bool yint = NDIM == 2 || (y >= NOP && y < n3 - NOP);
            const bool ysp = NDIM == 3 && a.dabc && (y < a.nb || y >= n3 - a.nb);
            const float spy = ysp ? a.sp3[y] : 1.f;
            const bool ysh = NDIM == 3 && a.abc && (y <= NOP + 1 || y >= n3 - NOP - 2);

            float nv[RX], zp1[RX], zm1[RX];
#pragma unroll
            for (int j = 0; j < RX; j++) {
                /* z neighbours: adjacent lanes, the edge lanes taking the
                 * fetched values beyond the tile (moved to lane 0 / 63) */
                float zm[NOP], zp[NOP];
                float hlo = bz_[B][j], hhi = bz_[B][j];
                zm[0] = shr1(hlo, cen[j]);
                zp[0] = shl1(hhi, cen[j]);
#pragma unroll
                for (int k = 1; k < NOP; k++) {
                    hlo = shl1(hlo, hlo);
                    hhi = shr1(hhi, hhi);
                    zm[k] = shr1(hlo, zm[k - 1]);
                    zp[k] = shl1(hhi, zp[k - 1]);
                }
                zp1[j] = zp[0];
                zm1[j] = zm[0];
                /* the pending sponge factor of the previous field */
                float p = bp_[B][j];
                if (sp_xz || ysp) {   /* wave-uniform: no divergence */
                    p = p * spz;
                    p = p * spx[j];
                    if (NDIM == 3) p = p * spy;
                }
                /* every lane evaluates the stencil, then selects */
                const float c = cen[j];
                float lap = c * a.c0, v;
                if (SSE) {
#pragma unroll
                    for (int k = 1; k <= NOP; k++) {
                        const float xpk = xnb<NOP, RX>(cen, hal, j, k), xmk = xnb<NOP, RX>(cen, hal, j, -k);
                        lap = lap + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]);
                        lap = lap + a.cx[k - 1] * (xpk + xmk);
                        if (NDIM == 3)
                            lap = lap + a.cy[k - 1] * (q[j][(S0 + NOP + k) % QN] + q[j][(S0 + NOP - k) % QN]);
                    }
                    if (VD) lap = den_term<NOP, NDIM, SSE>(lap, zp, zm, cen, hal, q[j], S0, j, bg_[B][j], a);
                    const float t = 2.f * c - p;
                    v = t + bv_[B][j] * lap;
                } else {
#pragma unroll
                    for (int k = 1; k <= NOP; k++) {
                        const float xpk = xnb<NOP, RX>(cen, hal, j, k), xmk = xnb<NOP, RX>(cen, hal, j, -k);
                        float t = (zm[k - 1] + zp[k - 1]) * a.cz[k - 1] + (xmk + xpk) * a.cx[k - 1];
                        if (NDIM == 3)
                            t = t + (q[j][(S0 + NOP - k) % QN] + q[j][(S0 + NOP + k) % QN]) * a.cy[k - 1];
                        lap = lap + t;
                    }
                    if (VD) lap = den_term<NOP, NDIM, SSE>(lap, zp, zm, cen, hal, q[j], S0, j, bg_[B][j], a);
                    const float vl = bv_[B][j] * lap;
                    v = (float)(2.0 * (double)c - (double)p + (double)vl);
                }
                nv[j] = (zint && ((xint_m >> j) & 1u) && yint) ? v : p;
            }

            /* source injection into the new field, in source order */
            int e = ib;
            for (; e < ie && a.inj[e].y == y; e++) {
                const awefd_inj en = a.inj[e];
                const float amp = a.wav[(long)it * a.wstride + (a.wexpl ? 0 : en.src)];
#pragma unroll
                for (int j = 0; j < RX; j++) {
                    const float inj = nv[j] + amp * en.w * en.w2 * en.w3 * bv_[B][j];
                    nv[j] = (zin && en.z == iz && en.x == x0 + j) ? inj : nv[j];
                }
            }
            ib = e;

            /* one-way ABC on the z faces: serial outward recurrence along z,
             * carried across lanes (abcone3d_apply, fdutil.c:2097-2120); um
             * is the current field (this lane's centre and z+-1 neighbours) */
            if (zlow || zhigh) {
#pragma unroll
                for (int j = 0; j < RX; j++) {
                    if (!((xin_m >> j) & 1u)) continue;
                    if (zlow) {
                        float b;
                        if (zbs) b = zbt[0][y - y0][x0 - bx0 + j];
                        else b = a.bzl[y * n2 + x0 + j];
#pragma unroll
                        for (int k = 0; k < NOP; k++) {
                            const int zt = NOP - k;
                            const float up = shl1(nv[j], nv[j]);
                            const float w = zp1[j] + (cen[j] - up) * b;
                            nv[j] = iz == zt ? w : nv[j];
                        }
                    }
                    if (zhigh) {
                        float b;
                        if (zbs) b = zbt[1][y - y0][x0 - bx0 + j];
                        else b = a.bzh[y * n2 + x0 + j];
#pragma unroll
                        for (int k = 0; k < NOP; k++) {
                            const int zt = n1 - NOP + k - 1;
                            const float dn = shr1(nv[j], nv[j]);
                            const float w = zm1[j] + (cen[j] - dn) * b;
                            nv[j] = iz == zt ? w : nv[j];
                        }
                    }
                }
            }

            /* sponge of the new field and store; nodes another face pass
             * still has to update are stored unsponged */
#pragma unroll
            for (int j = 0; j < RX; j++) {
                if (!((xin_m >> j) & 1u)) continue;
                float v = nv[j];
                if ((sp_xz || ysp) && !(((xdef_m >> j) & 1u) || ysh)) {   /* wave-uniform */
                    float w = v * spz;
                    w = w * spx[j];
                    if (NDIM == 3) w = w * spy;
                    v = zdef ? v : w;
                }
                bst<NT>(rN, vz, soff(x0 + j, y, ybase, rowb, psz), v);
            }
        }
    }
}

/* ---- The 3-D step in columns ----
 *
 * A block is one wave: 64 z nodes (one per lane) of RX x rows (or, for the
 * last z tile when it holds at most 16 nodes, 16 z nodes of 4 x tiles of RX
 * rows: four 16-lane segments, z neighbours by row shifts), streaming a
 * y range of the grid (the whole axis unless the grid is too small to fill
 * the chip with columns).  Every per-plane input (previous field, (v dt)^2,
 * the current field's head plane, its z- and x-halo values, density
 * records) is loaded PD planes ahead into rotating register slots, so a
 * wave keeps PD planes of loads in flight and no computation waits on a load
 * issued in its own iteration.  The x-halo rows come from L2: the
 * neighbouring columns stream the same planes at about the same time.  The
 * y sponge factors and the z-face one-way coefficients of the block's planes
 * are staged in LDS once.  No barrier in the plane loop; with whole-axis
 * columns each current-field plane is read from HBM once (no chunk
 * prologues).  The arithmetic, operation order and selections are the
 * tiled body's (cart_step_body), so the fields are the same bit for bit. */

/* ring slots of the current field: a multiple of PD + 1 holding planes
 * y - NOP .. y + NOP + PD */
template <int NOP, int PD>
constexpr int col_ring()
{
    return ((2 * NOP + 1 + PD + PD) / (PD + 1)) * (PD + 1);
}

/* loads of plane y into stage slot S: previous field, (v dt)^2, the z-halo
 * value of each row and the 2 NOP x-halo rows of the current field */
template <int NOP, int RX, bool VD, int SEG, int NH>
__device__ __forceinline__ void col_stage(float (&p)[RX], float (&v)[RX], float (&zh)[RX], float (&hal)[2 * NOP],
                                          float4 (&g)[RX], rsrc_t rP, rsrc_t rV, rsrc_t rC, rsrc_t rR, unsigned vz,
                                          unsigned vzh, unsigned vz4, const unsigned (&vhx)[NH], int x0, int y,
                                          int ybase, unsigned rowb, unsigned psz)
{
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const unsigned s = soff(x0 + j, y, ybase, rowb, psz);
        p[j] = bld<NT>(rP, vz, s);
        v[j] = bld<NT>(rV, vz, s);
        zh[j] = bld(rC, vzh, s);
        if (VD) g[j] = bld4(rR, vz4, 4u * s);
    }
#pragma unroll
    for (int h = 0; h < 2 * NOP; h++) {
        if (SEG == WAVE) {
            const int hx = h < NOP ? x0 - NOP + h : x0 + RX + (h - NOP);
            hal[h] = bld(rC, vz, soff(hx, y, ybase, rowb, psz));
        } else {
            hal[h] = bld(rC, vhx[h], (unsigned)(y - ybase) * psz);
        }
    }
}

/* z shifts by one node within the wave's z segments: the whole wave
 * (SEG 64) or each 16-lane row (SEG 16); the lane without a source keeps
 * `old` */
template <int SEG>
__device__ __forceinline__ float zshr(float old, float v)
{
    if (SEG == WAVE) return shr1(old, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, false));
}
template <int SEG>
__device__ __forceinline__ float zshl(float old, float v)
{
    if (SEG == WAVE) return shl1(old, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), 0x101, 0xf, 0xf, false));
}

/* One column chunk: z tile tzi, x tile txi (RX rows), y chunk tyi.  SEG 64:
 * one z segment of 64 lanes.  SEG 16: the last z tile of a grid with
 * n1 % 64 <= 16, four 16-lane z segments, segment s taking x tile txi + s
 * (s < nsub); every address then carries the segment's row offset. */
template <int NOP, bool SSE, bool VD, int RX, int PD, int SEG>
__device__ __forceinline__ void col_tile(const awefd_cart_args &a, float *Pout, const float *Cin, int it, int tzi,
                                         int txi, int tyi, int nsub, float *lds)
{
    constexpr int QR = col_ring<NOP, PD>();
    constexpr int NS = PD + 1;
    constexpr int NSEG = WAVE / SEG;
    constexpr int RXT = RX * NSEG;                     /* x rows of the wave */
    const int lane = threadIdx.x;
    const int zoff = lane % SEG, sub = lane / SEG;
    const bool subok = sub < nsub;
    const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
    const int z0 = tzi * WAVE;
    const int iz = z0 + zoff;
    const bool zin = iz < n1 && subok;
    const int x0 = txi * RX;                           /* first row of the wave */
    const int xl = x0 + sub * RX;                      /* first row of this lane's segment */
    const int y0 = (int)((long)tyi * n3 / a.nty);
    const int y1 = (int)((long)(tyi + 1) * n3 / a.nty);
    const long sy = (long)a.ld * n2;
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned psz = (unsigned)(sy * 4);
    const int ybase = max(y0 - NOP, 0);
    const int ytop = min(y1 + NOP, n3);
    const unsigned nbytes = (unsigned)(ytop - ybase) * psz;
    const rsrc_t rC = mk_rsrc(Cin + ybase * sy, nbytes);
    const rsrc_t rP = mk_rsrc(Pout + ybase * sy, nbytes);
    const rsrc_t rV = mk_rsrc(a.vel + ybase * sy, nbytes);
    const rsrc_t rR = mk_rsrc(VD ? (const void *)(a.rho + 4 * ybase * sy) : (const void *)a.vel,
                              VD ? 4u * nbytes : 0u);
    const unsigned xo = (unsigned)(sub * RX) * rowb;  /* the segment's row offset */
    const unsigned vz = zin ? (unsigned)iz * 4u + xo : OOB;
    const unsigned vz4 = zin ? (unsigned)iz * 16u + 4u * xo : OOB;
    int zz = -1;
    if (zoff < NOP) zz = z0 - 1 - zoff;
    else if (zoff >= SEG - NOP) zz = z0 + SEG + (SEG - 1 - zoff);
    const unsigned vzh = (subok && zz >= 0 && zz < n1) ? (unsigned)zz * 4u + xo : OOB;
    /* SEG 16: the x-halo rows by lane (a row below the grid is dropped) */
    unsigned vhx[SEG == WAVE ? 1 : 2 * NOP] = {OOB};
    if (SEG != WAVE) {
#pragma unroll
        for (int h = 0; h < 2 * NOP; h++) {
            const int hx = h < NOP ? xl - NOP + h : xl + RX + (h - NOP);
            vhx[h] = (zin && hx >= 0) ? (unsigned)hx * rowb + (unsigned)iz * 4u : OOB;
        }
    }

    const bool zint = iz >= NOP && iz < n1 - NOP;
    const bool zsp = a.dabc && zin && (iz < a.nb || iz >= n1 - a.nb);
    const float spz = zsp ? a.sp1[iz] : 1.f;
    const bool zsp_w = a.dabc && (z0 < a.nb || z0 + SEG > n1 - a.nb);
    const bool zdef = a.zsep && iz >= n1 - NOP - 2;
    float spx[RX];
    unsigned xint_m = 0, xdef_m = 0, xin_m = 0;
    bool xsp_w = false;
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int ix = xl + j;
        const bool xsp = a.dabc && subok && ix < n2 && (ix < a.nb || ix >= n2 - a.nb);
        spx[j] = xsp ? a.sp2[ix] : 1.f;
        xin_m |= (subok && ix < n2) ? 1u << j : 0u;
        xint_m |= (ix >= NOP && ix < n2 - NOP) ? 1u << j : 0u;
        xdef_m |= (a.abc && (ix <= NOP + 1 || ix >= n2 - NOP - 2)) ? 1u << j : 0u;
    }
    /* wave-uniform: any row of any segment in an x sponge strip */
#pragma unroll
    for (int k = 0; k < RXT; k++) {
        const int ix = x0 + k;
        xsp_w = xsp_w || (a.dabc && k < nsub * RX && ix < n2 && (ix < a.nb || ix >= n2 - a.nb));
    }
    const bool sp_xz = zsp_w || xsp_w;
    const bool zlow = a.abc && z0 == 0 && !a.fsrf;
    const bool zhigh = a.abc && !a.zsep && z0 + SEG >= n1;

    /* LDS: the y sponge factor of each plane of the range, then (z-edge
     * columns) the z-low and z-high one-way coefficients of each row */
    const int ny = y1 - y0;
    float *ysp_l = lds;
    float *zb_l = lds + ny;
    for (int i = lane; i < ny; i += WAVE) ysp_l[i] = a.dabc ? a.sp3[y0 + i] : 1.f;
    if (zlow || zhigh)
        for (int i = lane; i < ny * RXT; i += WAVE) {
            const long o = (long)(y0 + i / RXT) * n2 + min(x0 + i % RXT, n2 - 1);
            zb_l[2 * i] = zlow ? a.bzl[o] : 0.f;
            zb_l[2 * i + 1] = zhigh ? a.bzh[o] : 0.f;
        }
    __syncthreads();

    /* injection entries of each segment's column, from plane y0 on */
    int ib[NSEG], ie[NSEG];
#pragma unroll
    for (int g = 0; g < NSEG; g++) {
        const int col = tzi + a.ntz * (txi + g);
        ib[g] = g < nsub ? a.ibeg[col] : 0;
        ie[g] = g < nsub ? a.ibeg[col + 1] : 0;
        while (ib[g] < ie[g] && a.inj[ib[g]].y < y0) ib[g]++;
    }

    /* current field ring (plane p in slot (p - y0 + NOP) % QR) and stage
     * slots (plane y in slot (y - y0) % NS) */
    float q[RX][QR];
    float sp_[NS][RX], sv_[NS][RX], sz_[NS][RX], sh_[NS][2 * NOP];
    float4 sg_[NS][RX];
#pragma unroll
    for (int k = 0; k < 2 * NOP + PD; k++)
#pragma unroll
        for (int j = 0; j < RX; j++) q[j][k] = bld(rC, vz, soff(x0 + j, y0 - NOP + k, ybase, rowb, psz));
#pragma unroll
    for (int k = 0; k < PD; k++)
        col_stage<NOP, RX, VD, SEG, (SEG == WAVE ? 1 : 2 * NOP)>(sp_[k], sv_[k], sz_[k], sh_[k], sg_[k], rP, rV, rC, rR, vz, vzh, vz4, vhx,
                                    x0, y0 + k, ybase, rowb, psz);

    for (int yg = y0; yg < y1; yg += QR) {
#pragma unroll
        for (int R = 0; R < QR; R++) {
            const int y = yg + R;
            if (y >= y1) break;
            const int B = R % NS;                         /* stage slot of plane y */
            const int C0 = (R + NOP) % QR;                /* ring slot of plane y  */
            if (y + PD < y1) {
                const int Bn = (R + PD) % NS;
#pragma unroll
                for (int j = 0; j < RX; j++)
                    q[j][(R + 2 * NOP + PD) % QR] = bld(rC, vz, soff(x0 + j, y + NOP + PD, ybase, rowb, psz));
                col_stage<NOP, RX, VD, SEG, (SEG == WAVE ? 1 : 2 * NOP)>(sp_[Bn], sv_[Bn], sz_[Bn], sh_[Bn], sg_[Bn], rP, rV, rC, rR, vz, vzh,
                                            vz4, vhx, x0, y + PD, ybase, rowb, psz);
            }
            float cen[RX];
#pragma unroll
            for (int j = 0; j < RX; j++) cen[j] = q[j][C0];
            const bool yint = y >= NOP && y < n3 - NOP;
            const bool ysp = a.dabc && (y < a.nb || y >= n3 - a.nb);
            const float spy = ysp ? ysp_l[y - y0] : 1.f;
            const bool ysh = a.abc && (y <= NOP + 1 || y >= n3 - NOP - 2);

            float nv[RX], zp1[RX], zm1[RX];
#pragma unroll
            for (int j = 0; j < RX; j++) {
                float zm[NOP], zp[NOP];
                float hlo = sz_[B][j], hhi = sz_[B][j];
                zm[0] = zshr<SEG>(hlo, cen[j]);
                zp[0] = zshl<SEG>(hhi, cen[j]);
#pragma unroll
                for (int k = 1; k < NOP; k++) {
                    hlo = zshl<SEG>(hlo, hlo);
                    hhi = zshr<SEG>(hhi, hhi);
                    zm[k] = zshr<SEG>(hlo, zm[k - 1]);
                    zp[k] = zshl<SEG>(hhi, zp[k - 1]);
                }
                zp1[j] = zp[0];
                zm1[j] = zm[0];
                float p = sp_[B][j];
                if (sp_xz || ysp) {
                    p = p * spz;
                    p = p * spx[j];
                    p = p * spy;
                }
                const float c = cen[j];
                float lap = c * a.c0, v;
                if (SSE) {
#pragma unroll
                    for (int k = 1; k <= NOP; k++) {
                        const float xpk = xnb<NOP, RX>(cen, sh_[B], j, k), xmk = xnb<NOP, RX>(cen, sh_[B], j, -k);
                        lap = lap + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]);
                        lap = lap + a.cx[k - 1] * (xpk + xmk);
                        lap = lap + a.cy[k - 1] * (q[j][(C0 + k) % QR] + q[j][(C0 + QR - k) % QR]);
                    }
                    if (VD) lap = den_term<NOP, 3, SSE>(lap, zp, zm, cen, sh_[B], q[j], (C0 + QR - NOP) % QR, j,
                                                       sg_[B][j], a);
                    const float t = 2.f * c - p;
                    v = t + sv_[B][j] * lap;
                } else {
#pragma unroll
                    for (int k = 1; k <= NOP; k++) {
                        const float xpk = xnb<NOP, RX>(cen, sh_[B], j, k), xmk = xnb<NOP, RX>(cen, sh_[B], j, -k);
                        float t = (zm[k - 1] + zp[k - 1]) * a.cz[k - 1] + (xmk + xpk) * a.cx[k - 1];
                        t = t + (q[j][(C0 + QR - k) % QR] + q[j][(C0 + k) % QR]) * a.cy[k - 1];
                        lap = lap + t;
                    }
                    if (VD) lap = den_term<NOP, 3, SSE>(lap, zp, zm, cen, sh_[B], q[j], (C0 + QR - NOP) % QR, j,
                                                       sg_[B][j], a);
                    const float vl = sv_[B][j] * lap;
                    v = (float)(2.0 * (double)c - (double)p + (double)vl);
                }
                nv[j] = (zint && ((xint_m >> j) & 1u) && yint) ? v : p;
            }

#pragma unroll
            for (int g = 0; g < NSEG; g++) {
                int e = ib[g];
                for (; e < ie[g] && a.inj[e].y == y; e++) {
                    const awefd_inj en = a.inj[e];
                    const float amp = a.wav[(long)it * a.wstride + (a.wexpl ? 0 : en.src)];
#pragma unroll
                    for (int j = 0; j < RX; j++) {
                        const float inj = nv[j] + amp * en.w * en.w2 * en.w3 * sv_[B][j];
                        nv[j] = (zin && en.z == iz && en.x == xl + j) ? inj : nv[j];
                    }
                }
                ib[g] = e;
            }

            if (zlow || zhigh) {
#pragma unroll
                for (int j = 0; j < RX; j++) {
                    if (!((xin_m >> j) & 1u)) continue;
                    if (zlow) {
                        const float b = zb_l[2 * ((y - y0) * RXT + sub * RX + j)];
#pragma unroll
                        for (int k = 0; k < NOP; k++) {
                            const int zt = NOP - k;
                            const float up = zshl<SEG>(nv[j], nv[j]);
                            const float w = zp1[j] + (cen[j] - up) * b;
                            nv[j] = iz == zt ? w : nv[j];
                        }
                    }
                    if (zhigh) {
                        const float b = zb_l[2 * ((y - y0) * RXT + sub * RX + j) + 1];
#pragma unroll
                        for (int k = 0; k < NOP; k++) {
                            const int zt = n1 - NOP + k - 1;
                            const float dn = zshr<SEG>(nv[j], nv[j]);
                            const float w = zm1[j] + (cen[j] - dn) * b;
                            nv[j] = iz == zt ? w : nv[j];
                        }
                    }
                }
            }

#pragma unroll
            for (int j = 0; j < RX; j++) {
                if (!((xin_m >> j) & 1u)) continue;
                float v = nv[j];
                if ((sp_xz || ysp) && !(((xdef_m >> j) & 1u) || ysh)) {
                    float w = v * spz;
                    w = w * spx[j];
                    w = w * spy;
                    v = zdef ? v : w;
                }
                bst<NT>(rP, vz, soff(x0 + j, y, ybase, rowb, psz), v);
            }
        }
    }
}

template <int NOP, bool SSE, bool VD, int RX, int PD>
__device__ __forceinline__ void cart_col_body(const awefd_cart_args &a, float *Pout, const float *Cin, int it,
                                              float *trow, int bid, float *lds)
{
    const int lane = threadIdx.x;
    if (bid >= a.nblk) {
        cart_extract<3>(a, Cin, trow, (bid - a.nblk) * WAVE + lane);
        return;
    }
    /* Block b runs on XCD b % 8.  XCD k owns the columns of x tiles
     * [ntx k / 8, ntx (k + 1) / 8) and takes them z tile fastest, then x,
     * then y chunk: the columns that share x-halo rows, and the chunks that
     * share prologue planes, run on one XCD close together in time, so those
     * re-reads hit its L2, and the eight XCDs stream the same planes at a
     * time.  Blocks past an XCD's share (when 8 does not divide ntx) exit. */
    const int xcd = bid & 7, slot = bid >> 3;
    const int xb = (int)((long)a.ntx * xcd / 8), nxk = (int)((long)a.ntx * (xcd + 1) / 8) - xb;
    if (nxk == 0) return;
    if (!a.ztail) {
        const int tzi = slot % a.ntz;
        const int txi = xb + (slot / a.ntz) % nxk;
        const int tyi = slot / (a.ntz * nxk);
        if (tyi >= a.nty) return;
        col_tile<NOP, SSE, VD, RX, PD, WAVE>(a, Pout, Cin, it, tzi, txi, tyi, 1, lds);
        return;
    }
    /* groups of 4 x tiles: their 4 (ntz - 1) full z tiles, then one wave for
     * the last z tile of all 4 (in the same launch: as a kernel of its own,
     * launched after the full tiles, it cost more than it saved) */
    const int ntzm = a.ntz - 1, gsz = 4 * ntzm + 1, per = ((nxk + 3) / 4) * gsz;
    const int tyi = slot / per;
    if (tyi >= a.nty) return;
    const int i = slot % per, g = i / gsz, r = i % gsz;
    if (r < 4 * ntzm) {
        const int txi = xb + 4 * g + r / ntzm;
        if (txi >= xb + nxk) return;
        col_tile<NOP, SSE, VD, RX, PD, WAVE>(a, Pout, Cin, it, r % ntzm, txi, tyi, 1, lds);
    } else {
        const int txi = xb + 4 * g;
        col_tile<NOP, SSE, VD, RX, PD, 16>(a, Pout, Cin, it, ntzm, txi, tyi, min(4, xb + nxk - txi), lds);
    }
}

template <int NOP, bool SSE, bool VD>
__global__ __launch_bounds__(WAVE) void k_cart_col(const awefd_cart_args a)
{
    extern __shared__ float col_lds[];
    cart_col_body<NOP, SSE, VD, AWEFD_COL_RX, AWEFD_COL_PD>(a, a.P, a.C, a.it, a.trow, blockIdx.x, col_lds);
}

/* The single-grid step kernel: one launch per time step of one model. */
template <int NOP, int NDIM, bool SSE, bool VD, bool NARROW>
__global__ __launch_bounds__(step_nw<NARROW>() * WAVE) void k_cart_step(const awefd_cart_args a)
{
    static_assert(NDIM == 2 || !NARROW, "narrow tiles are 2-D");
    float *dst = NDIM == 2 && a.Pn ? a.Pn : a.P;
    cart_step_body<NOP, NDIM, SSE, VD, step_rx<NARROW>(), step_nw<NARROW>()>(a, a.P, a.C, a.it, a.trow, blockIdx.x,
                                                                            dst);
}

/* The batched 2-D step kernel: one launch advances every shot of a batch one
 * time step.  Each shot's arguments live in device memory; block b belongs
 * to shot bshot[b] and is that shot's block b - blk0 (stencil tiles, then
 * its extraction blocks).  The two field buffers of a shot alternate roles
 * with the parity of the step, and its trace row is traces + it*nr.  Tile
 * shape (RX rows per wave, NW waves per block) is a template parameter, chosen
 * per batch by the host autotuner. */
template <int NOP, bool SSE, int RX, int NW, bool RC>
__global__ __launch_bounds__(NW * WAVE) void k_cart_batch2(const awefd_cart_args *__restrict__ shots,
                                                           const int *__restrict__ bshot, int it)
{
    const int s = __builtin_amdgcn_readfirstlane(bshot[blockIdx.x]);
    const awefd_cart_args &a = shots[s];
    const int bid = (int)blockIdx.x - a.blk0;
    if (!RC) {
        const bool odd = it & 1;
        float *P = odd ? (float *)a.C : a.P;
        const float *C = odd ? a.P : a.C;
        cart_step_body<NOP, 2, SSE, false, RX, NW>(a, P, C, it, a.trow + (long)it * a.nr, bid, P);
        return;
    }
    /* recompute of step it inside its checkpoint segment, out of place;
     * the fields are those of the forward run, so no traces are written */
    if (bid >= a.nblk) return;
    const int k = it / a.K, n0 = k * a.K;
    const float *ckc = a.ck + 2L * k * a.nal;
    float *cur = it == n0 ? (float *)ckc : a.seg + (long)(it - n0 - 1) * a.nal;
    float *prev = it == n0 ? (float *)ckc + a.nal : (it - 1 == n0 ? (float *)ckc : a.seg + (long)(it - n0 - 2) * a.nal);
    float *next = a.seg + (long)(it - n0) * a.nal;
    cart_step_body<NOP, 2, SSE, false, RX, NW>(a, prev, cur, it, a.trow, bid, next);
}

/* ---- Two time steps per pass (batched 2-D shots) ----
 *
 * A block of NW waves covers 64 z (one per lane) x ROWS = NW*RX rows of the
 * first step; its outputs are the inner (64 - 2 NOP) x (ROWS - 2 NOP) nodes,
 * so neighbouring tiles overlap by 2 NOP in z and x.  The first step is the
 * single-step body's arithmetic on every lane and row (the z values beyond
 * the lanes come from the edge lanes' halo loads, the x rows beyond the
 * block from global loads, the block's own rows through LDS); its result
 * goes to LDS and is the current field of the second step, whose previous
 * field is the first step's current field with its pending sponge factor.
 * Both results are stored for the tile's inner nodes: the pair is the next
 * pass's input.  Every product, sum and selection is the single-step one, in
 * the same order, so the fields are those of two single steps bit for bit
 * (a factor of exactly 1 stands for "no sponge here"). */

/* the 2*NOP z neighbours of v (one per lane), halo values in the edge lanes */
template <int NOP>
__device__ __forceinline__ void zneigh(float v, float halo, float (&zm)[NOP], float (&zp)[NOP])
{
    float hlo = halo, hhi = halo;
    zm[0] = shr1(hlo, v);
    zp[0] = shl1(hhi, v);
#pragma unroll
    for (int k = 1; k < NOP; k++) {
        hlo = shl1(hlo, hlo);
        hhi = shr1(hhi, hhi);
        zm[k] = shr1(hlo, zm[k - 1]);
        zp[k] = shl1(hhi, zp[k - 1]);
    }
}

/* One leapfrog update of a node from its current value c, z neighbours,
 * x neighbours (xm[k-1], xp[k-1] at distance k) and sponged previous value
 * p, as the single-step body evaluates it (SSE or scalar order). */
template <int NOP, bool SSE>
__device__ __forceinline__ float leap2d(const awefd_cart_args &a, float c, const float (&zm)[NOP],
                                        const float (&zp)[NOP], const float (&xm)[NOP], const float (&xp)[NOP],
                                        float p, float vv)
{
    float lap = c * a.c0;
    if (SSE) {
#pragma unroll
        for (int k = 1; k <= NOP; k++) {
            lap = lap + a.cz[k - 1] * (zp[k - 1] + zm[k - 1]);
            lap = lap + a.cx[k - 1] * (xp[k - 1] + xm[k - 1]);
        }
        const float t = 2.f * c - p;
        return t + vv * lap;
    }
#pragma unroll
    for (int k = 1; k <= NOP; k++) {
        const float t = (zm[k - 1] + zp[k - 1]) * a.cz[k - 1] + (xm[k - 1] + xp[k - 1]) * a.cx[k - 1];
        lap = lap + t;
    }
    const float vl = vv * lap;
    return (float)(2.0 * (double)c - (double)p + (double)vl);
}

template <int NOP, bool SSE, int RX, int NW>
__device__ __forceinline__ void cart_step2_body(const awefd_cart_args &a, int bid, int it, int roles, int two)
{
    constexpr int ROWS = NW * RX;
    constexpr int SZ = WAVE - 2 * NOP;            /* z outputs per tile */
    constexpr int SX = ROWS - 2 * NOP;            /* x outputs per tile */
    static_assert(SX > 0, "two-step tile too narrow for the stencil");
    const float *Pin = a.F[roles & 3], *Cin = a.F[(roles >> 2) & 3];
    float *D1 = a.F[(roles >> 4) & 3], *D2 = a.F[(roles >> 6) & 3];
    const int tid = threadIdx.x;
    if (bid >= a.nblk2) {
        /* trace rows it (current field) and it-1 (previous field) */
        const int r = (bid - a.nblk2) * (NW * WAVE) + tid;
        cart_extract<2>(a, Cin, a.trow + (long)it * a.nr, r);
        if (it > 0) cart_extract<2>(a, Pin, a.trow + (long)(it - 1) * a.nr, r);
        return;
    }
    const int lane = tid & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / WAVE);
    const int lb = xcd_remap(bid, a.nblk2);
    const int tzi = lb % a.ntz2, txi = lb / a.ntz2;
    const int n1 = a.n1, n2 = a.n2;
    const int z0 = tzi * SZ - NOP;                /* z of lane 0           */
    const int iz = z0 + lane;
    const bool zin = iz >= 0 && iz < n1;
    const int bx0 = txi * SX - NOP;               /* the block's first row */
    const int x0 = bx0 + wv * RX;                 /* wave-uniform          */
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned nbytes = (unsigned)a.ld * (unsigned)n2 * 4u;
    const rsrc_t rC = mk_rsrc(Cin, nbytes), rP = mk_rsrc(Pin, nbytes), rV = mk_rsrc(a.vel, nbytes);
    const rsrc_t rD1 = mk_rsrc(D1, nbytes), rD2 = mk_rsrc(D2, nbytes);
    const unsigned vz = zin ? (unsigned)iz * 4u : OOB;
    int zz = -1;
    if (lane < NOP) zz = z0 - 1 - lane;
    else if (lane >= WAVE - NOP) zz = z0 + WAVE + (WAVE - 1 - lane);
    const unsigned vzh = (zz >= 0 && zz < n1) ? (unsigned)zz * 4u : OOB;
    const bool zint = iz >= NOP && iz < n1 - NOP;
    const bool zsp = a.dabc && zin && (iz < a.nb || iz >= n1 - a.nb);
    const float spz = zsp ? a.sp1[iz] : 1.f;
    /* this lane's node is an output of the tile */
    const unsigned vzo = (lane >= NOP && lane < WAVE - NOP) ? vz : OOB;

    float c[RX], p[RX], vv[RX], zh[RX], spx[RX];
    unsigned xint_m = 0;
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int ix = x0 + j;
        const unsigned so = (unsigned)ix * rowb;  /* rows outside the grid: beyond the range */
        c[j] = bld(rC, vz, so);
        p[j] = bld<NT>(rP, vz, so);
        vv[j] = bld<NT>(rV, vz, so);
        zh[j] = bld(rC, vzh, so);
        const bool xsp = a.dabc && ix >= 0 && ix < n2 && (ix < a.nb || ix >= n2 - a.nb);
        spx[j] = xsp ? a.sp2[ix] : 1.f;
        xint_m |= (ix >= NOP && ix < n2 - NOP) ? 1u << j : 0u;
    }
    /* x rows beyond the block, for the edge waves (the others get an
     * out-of-range offset) */
    float hg[2 * NOP];
#pragma unroll
    for (int h = 0; h < 2 * NOP; h++) {
        const int hx = h < NOP ? x0 - NOP + h : x0 + RX + (h - NOP);
        const bool outside = hx < bx0 || hx >= bx0 + ROWS;
        hg[h] = bld(rC, outside ? vz : OOB, (unsigned)hx * rowb);
    }
    /* injection entries whose node lies in this tile's first-step region */
    const int col = tzi + a.ntz2 * txi;
    const int ib = a.ibeg2[col], ie = a.ibeg2[col + 1];

    __shared__ float xr[2][ROWS][WAVE];
#pragma unroll
    for (int j = 0; j < RX; j++) xr[0][wv * RX + j][lane] = c[j];
    __syncthreads();

    /* first step (it) on every lane and row */
    float s1[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        float zm[NOP], zp[NOP], xm[NOP], xp[NOP];
        zneigh<NOP>(c[j], zh[j], zm, zp);
#pragma unroll
        for (int k = 1; k <= NOP; k++) {
            const int rm = wv * RX + j - k, rp = wv * RX + j + k;   /* block rows */
            const int hm = j - k < 0 ? NOP + (j - k) : -1;          /* halo slot  */
            const int hp = j + k >= RX ? NOP + (j + k - RX) : -1;
            float vm, vp;
            if (j - k >= 0) vm = c[j - k];
            else vm = rm >= 0 ? xr[0][rm][lane] : hg[hm];
            if (j + k < RX) vp = c[j + k];
            else vp = rp < ROWS ? xr[0][rp][lane] : hg[hp];
            xm[k - 1] = vm;
            xp[k - 1] = vp;
        }
        float pp = p[j] * spz;
        pp = pp * spx[j];
        const float v = leap2d<NOP, SSE>(a, c[j], zm, zp, xm, xp, pp, vv[j]);
        float nv = (zint && ((xint_m >> j) & 1u)) ? v : pp;
        for (int e = ib; e < ie; e++) {
            const awefd_inj en = a.inj2[e];
            const float amp = a.wav[(long)it * a.wstride + (a.wexpl ? 0 : en.src)];
            const float inj = nv + amp * en.w * en.w2 * en.w3 * vv[j];
            nv = (zin && en.z == iz && en.x == x0 + j) ? inj : nv;
        }
        float w = nv * spz;
        w = w * spx[j];
        s1[j] = w;
        const int r = wv * RX + j;
        if (r >= NOP && r < ROWS - NOP) bst<NT>(rD1, vzo, (unsigned)(x0 + j) * rowb, w);
    }
    if (!two) return;                             /* block-uniform */

#pragma unroll
    for (int j = 0; j < RX; j++) xr[1][wv * RX + j][lane] = s1[j];
    __syncthreads();

    /* second step (it + 1): current field s1, previous field c */
#pragma unroll
    for (int j = 0; j < RX; j++) {
        float zm[NOP], zp[NOP], xm[NOP], xp[NOP];
        zneigh<NOP>(s1[j], 0.f, zm, zp);          /* edge lanes are not outputs */
#pragma unroll
        for (int k = 1; k <= NOP; k++) {
            const int rm = wv * RX + j - k, rp = wv * RX + j + k;
            xm[k - 1] = j - k >= 0 ? s1[j - k] : xr[1][max(rm, 0)][lane];
            xp[k - 1] = j + k < RX ? s1[j + k] : xr[1][min(rp, ROWS - 1)][lane];
        }
        float pp = c[j] * spz;
        pp = pp * spx[j];
        const float v = leap2d<NOP, SSE>(a, s1[j], zm, zp, xm, xp, pp, vv[j]);
        float nv = (zint && ((xint_m >> j) & 1u)) ? v : pp;
        for (int e = ib; e < ie; e++) {
            const awefd_inj en = a.inj2[e];
            const float amp = a.wav[(long)(it + 1) * a.wstride + (a.wexpl ? 0 : en.src)];
            const float inj = nv + amp * en.w * en.w2 * en.w3 * vv[j];
            nv = (zin && en.z == iz && en.x == x0 + j) ? inj : nv;
        }
        float w = nv * spz;
        w = w * spx[j];
        const int r = wv * RX + j;
        if (r >= NOP && r < ROWS - NOP) bst<NT>(rD2, vzo, (unsigned)(x0 + j) * rowb, w);
    }
}

template <int NOP, bool SSE, int RX, int NW>
__global__ __launch_bounds__(NW * WAVE) void k_cart_batch_t2(const awefd_cart_args *__restrict__ shots,
                                                             const int *__restrict__ bshot, int it, int roles,
                                                             int two)
{
    const int s = __builtin_amdgcn_readfirstlane(bshot[blockIdx.x]);
    const awefd_cart_args &a = shots[s];
    cart_step2_body<NOP, SSE, RX, NW>(a, (int)blockIdx.x - a.blk02, it, roles, two);
}

/* Three steps per pass (batched 2-D).  A block covers 64 z x NW*RX x rows
 * and computes the first step on all of them, the second on the nodes at
 * least nop inside, the third on those at least 2 nop inside: outputs are
 * (64 - 4 nop) x (NW*RX - 4 nop) nodes.  Each step's field passes to the
 * next through LDS.  The pass reads the previous and current field and V
 * once and writes the second and third steps' fields (D1, D2); the first
 * step's trace row, it + 1, is extracted in the block from LDS by the tile
 * that owns the receiver's first corner (rbeg3/rx3), the others by the
 * following passes' extraction blocks.  nst = 1 takes one step (writing D1)
 * with the same tiles, for the end of the loop. */
template <int NOP, bool SSE, int RX, int NW, bool RC>
__device__ __forceinline__ void cart_step3_body(const awefd_cart_args &a, int bid, int it, const float *Pin,
                                                const float *Cin, float *D0, float *D1, float *D2, int nst)
{
    constexpr int ROWS = NW * RX;
    constexpr int H = 2 * NOP;                    /* tile halo for three steps */
    constexpr int SZ = WAVE - 2 * H;
    constexpr int SX = ROWS - 2 * H;
    static_assert(SX > 0 && SZ > 0, "three-step tile too small for the stencil");
    const int tid = threadIdx.x;
    if (bid >= a.nblk2) {
        if (RC) return;                           /* a recompute writes no traces */
        /* trace rows it (current field) and it-1 (previous field) */
        const int r = (bid - a.nblk2) * (NW * WAVE) + tid;
        cart_extract<2>(a, Cin, a.trow + (long)it * a.nr, r);
        if (it > 0) cart_extract<2>(a, Pin, a.trow + (long)(it - 1) * a.nr, r);
        return;
    }
    const int lane = tid & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / WAVE);
    const int lb = xcd_remap(bid, a.nblk2);
    const int tzi = lb % a.ntz2, txi = lb / a.ntz2;
    const int n1 = a.n1, n2 = a.n2;
    const int z0 = tzi * SZ - H;                  /* z of lane 0           */
    const int iz = z0 + lane;
    const bool zin = iz >= 0 && iz < n1;
    const int bx0 = txi * SX - H;                 /* the block's first row */
    const int x0 = bx0 + wv * RX;                 /* wave-uniform          */
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned nbytes = (unsigned)a.ld * (unsigned)n2 * 4u;
    const rsrc_t rC = mk_rsrc(Cin, nbytes), rP = mk_rsrc(Pin, nbytes), rV = mk_rsrc(a.vel, nbytes);
    const rsrc_t rD1 = mk_rsrc(D1, nbytes), rD2 = mk_rsrc(D2, nbytes);
    const unsigned vz = zin ? (unsigned)iz * 4u : OOB;
    int zz = -1;
    if (lane < NOP) zz = z0 - 1 - lane;
    else if (lane >= WAVE - NOP) zz = z0 + WAVE + (WAVE - 1 - lane);
    const unsigned vzh = (zz >= 0 && zz < n1) ? (unsigned)zz * 4u : OOB;
    const bool zint = iz >= NOP && iz < n1 - NOP;
    const bool zsp = a.dabc && zin && (iz < a.nb || iz >= n1 - a.nb);
    const float spz = zsp ? a.sp1[iz] : 1.f;
    /* this lane's node is an output of the tile */
    const unsigned vzo = (lane >= H && lane < WAVE - H) ? vz : OOB;

    float c[RX], p[RX], vv[RX], zh[RX], spx[RX];
    unsigned xint_m = 0;
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int ix = x0 + j;
        const unsigned so = (unsigned)ix * rowb;  /* rows outside the grid: beyond the range */
        c[j] = bld(rC, vz, so);
        p[j] = bld<NT>(rP, vz, so);
        vv[j] = bld<NT>(rV, vz, so);
        zh[j] = bld(rC, vzh, so);
        const bool xsp = a.dabc && ix >= 0 && ix < n2 && (ix < a.nb || ix >= n2 - a.nb);
        spx[j] = xsp ? a.sp2[ix] : 1.f;
        xint_m |= (ix >= NOP && ix < n2 - NOP) ? 1u << j : 0u;
    }
    float hg[2 * NOP];
#pragma unroll
    for (int h = 0; h < 2 * NOP; h++) {
        const int hx = h < NOP ? x0 - NOP + h : x0 + RX + (h - NOP);
        const bool outside = hx < bx0 || hx >= bx0 + ROWS;
        hg[h] = bld(rC, outside ? vz : OOB, (unsigned)hx * rowb);
    }
    const int col = tzi + a.ntz2 * txi;
    const int ib = a.ibeg2[col], ie = a.ibeg2[col + 1];
    /* this tile's receivers for the in-block trace row, read early (none in
     * a recompute) */
    const int kb = RC ? 0 : a.rbeg3[col], ke = RC ? 0 : a.rbeg3[col + 1];

    /* xr[0]: the current field, then (once every wave is past the first
     * step) the second step's field; xr[1]: the first step's field */
    __shared__ float xr[2][ROWS][WAVE];
#pragma unroll
    for (int j = 0; j < RX; j++) xr[0][wv * RX + j][lane] = c[j];
    __syncthreads();

    /* first step (it) on every lane and row */
    float s1[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        float zm[NOP], zp[NOP], xm[NOP], xp[NOP];
        zneigh<NOP>(c[j], zh[j], zm, zp);
#pragma unroll
        for (int k = 1; k <= NOP; k++) {
            const int rm = wv * RX + j - k, rp = wv * RX + j + k;
            const int hm = j - k < 0 ? NOP + (j - k) : -1;
            const int hp = j + k >= RX ? NOP + (j + k - RX) : -1;
            float vm, vp;
            if (j - k >= 0) vm = c[j - k];
            else vm = rm >= 0 ? xr[0][rm][lane] : hg[hm];
            if (j + k < RX) vp = c[j + k];
            else vp = rp < ROWS ? xr[0][rp][lane] : hg[hp];
            xm[k - 1] = vm;
            xp[k - 1] = vp;
        }
        float pp = p[j] * spz;
        pp = pp * spx[j];
        const float v = leap2d<NOP, SSE>(a, c[j], zm, zp, xm, xp, pp, vv[j]);
        float nv = (zint && ((xint_m >> j) & 1u)) ? v : pp;
        for (int e = ib; e < ie; e++) {
            const awefd_inj en = a.inj2[e];
            const float amp = a.wav[(long)it * a.wstride + (a.wexpl ? 0 : en.src)];
            const float inj = nv + amp * en.w * en.w2 * en.w3 * vv[j];
            nv = (zin && en.z == iz && en.x == x0 + j) ? inj : nv;
        }
        float w = nv * spz;
        w = w * spx[j];
        s1[j] = w;
        /* a recompute keeps every step's field: the first one too */
        const int r = wv * RX + j;
        if (RC && r >= H && r < ROWS - H) bst<NT>(mk_rsrc(D0, nbytes), vzo, (unsigned)(x0 + j) * rowb, w);
    }
    if (!RC && nst == 1) {                        /* block-uniform */
#pragma unroll
        for (int j = 0; j < RX; j++) {
            const int r = wv * RX + j;
            if (r >= H && r < ROWS - H) bst<NT>(rD1, vzo, (unsigned)(x0 + j) * rowb, s1[j]);
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < RX; j++) xr[1][wv * RX + j][lane] = s1[j];
    __syncthreads();

    /* second step (it + 1): current field s1, previous field c */
    float s2[RX];
#pragma unroll
    for (int j = 0; j < RX; j++) {
        float zm[NOP], zp[NOP], xm[NOP], xp[NOP];
        zneigh<NOP>(s1[j], 0.f, zm, zp);          /* edge lanes are not used */
#pragma unroll
        for (int k = 1; k <= NOP; k++) {
            const int rm = wv * RX + j - k, rp = wv * RX + j + k;
            xm[k - 1] = j - k >= 0 ? s1[j - k] : xr[1][max(rm, 0)][lane];
            xp[k - 1] = j + k < RX ? s1[j + k] : xr[1][min(rp, ROWS - 1)][lane];
        }
        float pp = c[j] * spz;
        pp = pp * spx[j];
        const float v = leap2d<NOP, SSE>(a, s1[j], zm, zp, xm, xp, pp, vv[j]);
        float nv = (zint && ((xint_m >> j) & 1u)) ? v : pp;
        for (int e = ib; e < ie; e++) {
            const awefd_inj en = a.inj2[e];
            const float amp = a.wav[(long)(it + 1) * a.wstride + (a.wexpl ? 0 : en.src)];
            const float inj = nv + amp * en.w * en.w2 * en.w3 * vv[j];
            nv = (zin && en.z == iz && en.x == x0 + j) ? inj : nv;
        }
        float w = nv * spz;
        w = w * spx[j];
        s2[j] = w;
        const int r = wv * RX + j;
        if (r >= H && r < ROWS - H) bst<NT>(rD1, vzo, (unsigned)(x0 + j) * rowb, w);
    }
#pragma unroll
    for (int j = 0; j < RX; j++) xr[0][wv * RX + j][lane] = s2[j];
    __syncthreads();

    /* trace row it + 1 (the first step's field, with its pending sponge
     * factor) for the receivers whose first corner is an output of this tile */
    for (int k = kb + tid; k < ke; k += NW * WAVE) {
        const awefd_rx3 rc = a.rx3[k];
        float g = 0.f;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int z = rc.jz + (q & 1), x = rc.jx + ((q >> 1) & 1);
            float u = xr[1][x - bx0][z - z0];
            if (a.dabc) {            /* sponge<2>: the z factor, then the x factor */
                u = u * rc.sz[q & 1];
                u = u * rc.sx[(q >> 1) & 1];
            }
            const float t = u * rc.w[q];
            g = q == 0 ? t : g + t;
        }
        a.trow[(long)(it + 1) * a.nr + rc.r] = g;
    }

    /* third step (it + 2): current field s2, previous field s1 */
#pragma unroll
    for (int j = 0; j < RX; j++) {
        float zm[NOP], zp[NOP], xm[NOP], xp[NOP];
        zneigh<NOP>(s2[j], 0.f, zm, zp);
#pragma unroll
        for (int k = 1; k <= NOP; k++) {
            const int rm = wv * RX + j - k, rp = wv * RX + j + k;
            xm[k - 1] = j - k >= 0 ? s2[j - k] : xr[0][max(rm, 0)][lane];
            xp[k - 1] = j + k < RX ? s2[j + k] : xr[0][min(rp, ROWS - 1)][lane];
        }
        float pp = s1[j] * spz;
        pp = pp * spx[j];
        const float v = leap2d<NOP, SSE>(a, s2[j], zm, zp, xm, xp, pp, vv[j]);
        float nv = (zint && ((xint_m >> j) & 1u)) ? v : pp;
        for (int e = ib; e < ie; e++) {
            const awefd_inj en = a.inj2[e];
            const float amp = a.wav[(long)(it + 2) * a.wstride + (a.wexpl ? 0 : en.src)];
            const float inj = nv + amp * en.w * en.w2 * en.w3 * vv[j];
            nv = (zin && en.z == iz && en.x == x0 + j) ? inj : nv;
        }
        float w = nv * spz;
        w = w * spx[j];
        const int r = wv * RX + j;
        if (r >= H && r < ROWS - H) bst<NT>(rD2, vzo, (unsigned)(x0 + j) * rowb, w);
    }
}

/* RC: three steps it, it+1, it+2 of a checkpoint segment's recompute, out of
 * place: from the segment's fields of steps it-1 and it (the checkpoint pair
 * at the segment's start) into its buffers of steps it+1, it+2, it+3, which
 * the host keeps inside the segment; the fields are those of the forward
 * run, so no traces are written. */
template <int NOP, bool SSE, int RX, int NW, bool RC>
__global__ __launch_bounds__(NW * WAVE) void k_cart_batch_t3(const awefd_cart_args *__restrict__ shots,
                                                             const int *__restrict__ bshot, int it, int roles,
                                                             int nst)
{
    const int s = __builtin_amdgcn_readfirstlane(bshot[blockIdx.x]);
    const awefd_cart_args &a = shots[s];
    const int bid = (int)blockIdx.x - a.blk02;
    if (!RC) {
        cart_step3_body<NOP, SSE, RX, NW, false>(a, bid, it, a.F[roles & 3], a.F[(roles >> 2) & 3], nullptr,
                                                 a.F[(roles >> 4) & 3], a.F[(roles >> 6) & 3], nst);
        return;
    }
    if (bid >= a.nblk2) return;
    const int k = it / a.K, n0 = k * a.K;
    const float *ckc = a.ck + 2L * k * a.nal;
    const float *cur = it == n0 ? ckc : a.seg + (long)(it - n0 - 1) * a.nal;
    const float *prev = it == n0 ? ckc + a.nal : (it - 1 == n0 ? ckc : a.seg + (long)(it - n0 - 2) * a.nal);
    float *d = a.seg + (long)(it - n0) * a.nal;
    cart_step3_body<NOP, SSE, RX, NW, true>(a, bid, it, prev, cur, d, d + a.nal, d + 2 * a.nal, 3);
}

/* z-high one-way ABC as a separate pass, for grids whose last z tile is too
 * narrow to hold the recurrence (abcone3d_apply, fdutil.c:2110-2118).  One
 * thread per (x, y) column. */
template <int NOP>
__global__ __launch_bounds__(256) void k_cart_abc_zhigh(const awefd_cart_args a)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)a.n2 * a.n3) return;
    const int ix = (int)(t % a.n2), iy = (int)(t / a.n2);
    const long b = (long)iy * a.ld * a.n2 + (long)ix * a.ld;
    const float bh = a.bzh[(long)iy * a.n2 + ix];
    float *P = a.P;
    const float *C = a.C;
    for (int k = 0; k < NOP; k++) {
        const int iz = a.n1 - NOP + k - 1;
        P[b + iz] = C[b + iz - 1] + (C[b + iz] - P[b + iz - 1]) * bh;
    }
    if (!a.dabc) return;
    const bool xsh = ix <= NOP + 1 || ix >= a.n2 - NOP - 2;
    const bool ysh = iy <= NOP + 1 || iy >= a.n3 - NOP - 2;
    if (xsh || ysh) return;
    for (int iz = a.n1 - NOP - 2; iz < a.n1; iz++)
        if (in_sponge<3>(iz, ix, iy, a)) P[b + iz] = sponge<3>(P[b + iz], iz, ix, iy, a);
}

/* One-way ABC on the x faces, then the y faces (abcone3d_apply,
 * fdutil.c:2122-2176), each followed by the sponge of the shell nodes it
 * finalises.  One thread per line; lines along z are coalesced.  A line's
 * nop + 2 nodes at each face (the step kernel stored them unsponged) are
 * loaded once, the serial recurrence uom[i] = um[i+-1] + (um[i] - uo[i+-1]) b
 * and the sponge run in registers, and they are stored once.  The 3-D
 * set-up guarantees n >= 2 nop + 4, so the two faces' node sets are
 * disjoint. */
template <int NOP>
__device__ __forceinline__ void abc_line(float *__restrict__ P, const float *__restrict__ C, long b, long st, int n,
                                         float bl, float bh, bool sp, int iz, int ix, int iy, int ax,
                                         const awefd_cart_args &a)
{
    constexpr int S = NOP + 2;
    float ul[S], uh[S], cl[S], ch[S];
#pragma unroll
    for (int i = 0; i < S; i++) {
        ul[i] = P[b + i * st];
        cl[i] = C[b + i * st];
        uh[i] = P[b + (long)(n - S + i) * st];
        ch[i] = C[b + (long)(n - S + i) * st];
    }
#pragma unroll
    for (int k = 0; k < NOP; k++) {
        const int l = NOP - k;                  /* low face: node NOP - k */
        ul[l] = cl[l + 1] + (cl[l] - ul[l + 1]) * bl;
        const int h = k + 1;                    /* high face: node n - NOP + k - 1 */
        uh[h] = ch[h - 1] + (ch[h] - uh[h - 1]) * bh;
    }
    if (sp) {
#pragma unroll
        for (int i = 0; i < S; i++) {
            const int jl = i, jh = n - S + i;
            const int zl = iz, xl = ax == 1 ? jl : ix, yl = ax == 2 ? jl : iy;
            const int xh = ax == 1 ? jh : ix, yh = ax == 2 ? jh : iy;
            if (in_sponge<3>(zl, xl, yl, a)) ul[i] = sponge<3>(ul[i], zl, xl, yl, a);
            if (in_sponge<3>(zl, xh, yh, a)) uh[i] = sponge<3>(uh[i], zl, xh, yh, a);
        }
    }
#pragma unroll
    for (int i = 0; i < S; i++) {
        P[b + i * st] = ul[i];
        P[b + (long)(n - S + i) * st] = uh[i];
    }
}

template <int NOP>
__global__ __launch_bounds__(256) void k_cart_abc_x(const awefd_cart_args a)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)a.n1 * a.n3) return;
    const int iz = (int)(t % a.n1), iy = (int)(t / a.n1);
    const long sx = a.ld, sy = (long)a.ld * a.n2;
    const float bl = a.bxl[(long)iy * a.n1 + iz], bh = a.bxh[(long)iy * a.n1 + iz];
    /* lines in the y shell are sponged by the y pass */
    const bool ysh = iy <= NOP + 1 || iy >= a.n3 - NOP - 2;
    abc_line<NOP>(a.P, a.C, iy * sy + iz, sx, a.n2, bl, bh, a.dabc && !ysh, iz, 0, iy, 1, a);
}

template <int NOP>
__global__ __launch_bounds__(256) void k_cart_abc_y(const awefd_cart_args a)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)a.n1 * a.n2) return;
    const int iz = (int)(t % a.n1), ix = (int)(t / a.n1);
    const long sx = a.ld, sy = (long)a.ld * a.n2;
    const float bl = a.byl[(long)ix * a.n1 + iz], bh = a.byh[(long)ix * a.n1 + iz];
    abc_line<NOP>(a.P, a.C, ix * sx + iz, sy, a.n3, bl, bh, a.dabc, iz, ix, 0, 2, a);
}

/* One face of one line of the hybrid boundary (apply_abc,
 * MAWEFDcart2Dcpu.c:553-614, MAWEFDcart3Dcpu.c:534-630).  Node q of the face
 * (q = 0 on the face, q+1 its inner neighbour) is at a0 + q st.  The nop
 * outermost nodes take the one-way value, from q = nop-1 outward, starting
 * from the stencil value at q = nop; then the nodes q = nb-1 .. nop blend the
 * one-way value with their stencil value by damp[nb-1-q], each using its
 * blended inner neighbour.  The inner neighbour's new value stays in a
 * register; the stencil values are read before they are overwritten. */
__device__ __forceinline__ void hyb_face(float *__restrict__ P, const float *__restrict__ C, long a0, long st,
                                         int nop, int nb, const float *__restrict__ damp, float k)
{
    float in = P[a0 + nop * st];
    for (int q = nop - 1; q >= 0; q--) {
        const long i = a0 + q * st;
        const float v = C[i + st] + (C[i] - in) * k;
        P[i] = v;
        in = v;
    }
    in = P[a0 + (long)nb * st];
    for (int q = nb - 1; q >= nop; q--) {
        const long i = a0 + q * st;
        const float w = damp[nb - 1 - q];
        const float bc = C[i + st] + (C[i] - in) * k;
        const float v = bc * (1.f - w) + P[i] * w;
        P[i] = v;
        in = v;
    }
}

/* Hybrid boundary along axis AX (0 z, 1 x, 2 y) of the new field P, C the
 * current one.  Thread t < nline sweeps the low face of line t, the next
 * nline threads the high faces, so neighbouring threads take neighbouring
 * lines (consecutive z for the x and y lines).  The one-way tables are
 * indexed by line: bz*[iy n2 + ix], bx*[iy n1 + iz], by*[ix n1 + iz].  A free
 * surface zeroes z in [0, nb] of each z line after its low face. */
template <int AX>
__global__ __launch_bounds__(256) void k_cart_hybrid(const awefd_cart_args a, int nop)
{
    const long sx = a.ld, sy = (long)a.ld * a.n2;
    const long nline = AX == 0 ? (long)a.n2 * a.n3 : AX == 1 ? (long)a.n1 * a.n3 : (long)a.n1 * a.n2;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= 2 * nline) return;
    const bool hi = t >= nline;
    const long line = hi ? t - nline : t;
    long b, st;
    int n;
    const float *tab;
    if (AX == 0) {
        b = (line / a.n2) * sy + (line % a.n2) * sx;
        st = 1;
        n = a.n1;
        tab = hi ? a.bzh : a.bzl;
    } else if (AX == 1) {
        b = (line / a.n1) * sy + line % a.n1;
        st = sx;
        n = a.n2;
        tab = hi ? a.bxh : a.bxl;
    } else {
        b = (line / a.n1) * sx + line % a.n1;
        st = sy;
        n = a.n3;
        tab = hi ? a.byh : a.byl;
    }
    const float k = tab[line];
    if (hi) hyb_face(a.P, a.C, b + (long)(n - 1) * st, -st, nop, a.nb, a.damp, k);
    else hyb_face(a.P, a.C, b, st, nop, a.nb, a.damp, k);
    if (AX == 0 && !hi && a.fsrf)
        for (int iz = 0; iz <= a.nb; iz++) a.P[b + iz] = 0.f;
}

/* the padded (v dt)^2 of awefd_set_model_dev: node (i1, i2, i3) takes model
 * node (clamp(i1 - nb1), clamp(i2 - nb2), clamp(i3 - nb3)) (expand2d/3d,
 * fdutil.c:435-571), then v (v dt2) as MAWEFDcart2Dcpu.c:289 forms it and
 * the free surface's zeros (MAWEFDcart2Dcpu.c:291-295) */
__global__ void k_set_vel(const float *vin, float *vout, int n1, int n2, int n3, int nb1, int nb2, int nb3, int ld,
                          float dt2, int fsnb)
{
    const int p1 = n1 + 2 * nb1, p2 = n2 + 2 * nb2, p3 = n3 + 2 * nb3;
    const long np = (long)p1 * p2 * p3;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (long)gridDim.x * blockDim.x) {
        const int i1 = (int)(i % p1);
        const long q = i / p1;
        const int i2 = (int)(q % p2), i3 = (int)(q / p2);
        const int k1 = min(max(i1 - nb1, 0), n1 - 1), k2 = min(max(i2 - nb2, 0), n2 - 1);
        const int k3 = min(max(i3 - nb3, 0), n3 - 1);
        const float v = vin[((long)k3 * n2 + k2) * n1 + k1];
        float w = v * (v * dt2);
        if (i1 <= fsnb) w = 0.f;
        vout[((long)i3 * p2 + i2) * ld + i1] = w;
    }
}

__global__ void k_fill(float *p, long n, float v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = v;
}

template <int NOP, int NDIM, bool SSE>
int launch_step(const awefd_cart_args *a, bool vd, hipStream_t s)
{
    if (NDIM == 3 && a->col) {
        /* LDS: the y sponge factor and two z-face coefficients per row of
         * each plane of the longest chunk (4 x tiles of rows for the waves
         * that take the last z tile of 4 x tiles) */
        const size_t lds = sizeof(float) * (size_t)a->ychunk * (1 + 2 * AWEFD_COL_RX * (a->ztail ? 4 : 1));
        const int nx = (a->nr + WAVE - 1) / WAVE;
        if (vd)
            hipLaunchKernelGGL((k_cart_col<NOP, SSE, true>), dim3(a->nblk + nx), dim3(WAVE), lds, s, *a);
        else
            hipLaunchKernelGGL((k_cart_col<NOP, SSE, false>), dim3(a->nblk + nx), dim3(WAVE), lds, s, *a);
        return (int)hipGetLastError();
    }
    if (NDIM == 2 && a->narrow) {
        constexpr int nthr = step_nw<true>() * WAVE;
        const int nextract = (a->nr + nthr - 1) / nthr;
        if (vd)
            hipLaunchKernelGGL((k_cart_step<NOP, 2, SSE, true, true>), dim3(a->nblk + nextract), dim3(nthr), 0, s, *a);
        else
            hipLaunchKernelGGL((k_cart_step<NOP, 2, SSE, false, true>), dim3(a->nblk + nextract), dim3(nthr), 0, s, *a);
        return (int)hipGetLastError();
    }
    constexpr int nthr = step_nw<false>() * WAVE;
    const int nextract = (a->nr + nthr - 1) / nthr;
    if (vd)
        hipLaunchKernelGGL((k_cart_step<NOP, NDIM, SSE, true, false>), dim3(a->nblk + nextract), dim3(nthr), 0, s, *a);
    else
        hipLaunchKernelGGL((k_cart_step<NOP, NDIM, SSE, false, false>), dim3(a->nblk + nextract), dim3(nthr), 0, s, *a);
    return (int)hipGetLastError();
}

template <int NOP, int NDIM, bool SSE>
int occupancy_step()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cart_step<NOP, NDIM, SSE, false, false>, NTHR0, 0) !=
        hipSuccess)
        return 0;
    return nb;
}

template <int NOP>
int occupancy_nop(int ndim, int sse)
{
    if (ndim == 2) return sse ? occupancy_step<NOP, 2, true>() : occupancy_step<NOP, 2, false>();
    return sse ? occupancy_step<NOP, 3, true>() : occupancy_step<NOP, 3, false>();
}

template <int NOP>
int launch_step_nop(const awefd_cart_args *a, int ndim, int sse, bool vd, hipStream_t s)
{
    if (ndim == 2) return sse ? launch_step<NOP, 2, true>(a, vd, s) : launch_step<NOP, 2, false>(a, vd, s);
    return sse ? launch_step<NOP, 3, true>(a, vd, s) : launch_step<NOP, 3, false>(a, vd, s);
}

template <int NOP>
int launch_abc(const awefd_cart_args *a, hipStream_t s)
{
    if (a->zsep) {
        const long nzl = (long)a->n2 * a->n3;
        hipLaunchKernelGGL((k_cart_abc_zhigh<NOP>), dim3((unsigned)((nzl + 255) / 256)), dim3(256), 0, s, *a);
    }
    const long nxl = (long)a->n1 * a->n3, nyl = (long)a->n1 * a->n2;
    hipLaunchKernelGGL((k_cart_abc_x<NOP>), dim3((unsigned)((nxl + 255) / 256)), dim3(256), 0, s, *a);
    hipLaunchKernelGGL((k_cart_abc_y<NOP>), dim3((unsigned)((nyl + 255) / 256)), dim3(256), 0, s, *a);
    return (int)hipGetLastError();
}

template <int NOP, bool SSE, bool RC>
int launch_batch_sse(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int rx, hipStream_t s)
{
    constexpr int NW = AWEFD_BATCH_NWAVE;
    switch (rx) {
    case 2: hipLaunchKernelGGL((k_cart_batch2<NOP, SSE, 2, NW, RC>), dim3(nblk), dim3(NW * WAVE), 0, s, shots, bshot, it); break;
    case 4: hipLaunchKernelGGL((k_cart_batch2<NOP, SSE, 4, NW, RC>), dim3(nblk), dim3(NW * WAVE), 0, s, shots, bshot, it); break;
    case 8: hipLaunchKernelGGL((k_cart_batch2<NOP, SSE, 8, NW, RC>), dim3(nblk), dim3(NW * WAVE), 0, s, shots, bshot, it); break;
    case 16: hipLaunchKernelGGL((k_cart_batch2<NOP, SSE, 16, NW, RC>), dim3(nblk), dim3(NW * WAVE), 0, s, shots, bshot, it); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

template <int NOP, bool SSE>
int launch_t2_sse(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int roles, int two, int rx,
                  hipStream_t s)
{
    constexpr int NW = AWEFD_BATCH_NWAVE;
    switch (rx) {
    case 4: hipLaunchKernelGGL((k_cart_batch_t2<NOP, SSE, 4, NW>), dim3(nblk), dim3(NW * WAVE), 0, s, shots, bshot, it, roles, two); break;
    case 8: hipLaunchKernelGGL((k_cart_batch_t2<NOP, SSE, 8, NW>), dim3(nblk), dim3(NW * WAVE), 0, s, shots, bshot, it, roles, two); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

template <int NOP, bool SSE, bool RC>
int launch_t3_sse(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int roles, int nst, int rx,
                  hipStream_t s)
{
    constexpr int N4 = AWEFD_T3_NWAVE(4), N8 = AWEFD_T3_NWAVE(8);
    switch (rx) {
    case 4: hipLaunchKernelGGL((k_cart_batch_t3<NOP, SSE, 4, N4, RC>), dim3(nblk), dim3(N4 * WAVE), 0, s, shots, bshot, it, roles, nst); break;
    case 8: hipLaunchKernelGGL((k_cart_batch_t3<NOP, SSE, 8, N8, RC>), dim3(nblk), dim3(N8 * WAVE), 0, s, shots, bshot, it, roles, nst); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

template <bool RC>
int launch_t3(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int roles, int nst, int nop, int sse,
              int rx, hipStream_t s)
{
    switch (nop) {
    case 1: return sse ? launch_t3_sse<1, true, RC>(shots, bshot, nblk, it, roles, nst, rx, s) : launch_t3_sse<1, false, RC>(shots, bshot, nblk, it, roles, nst, rx, s);
    case 2: return sse ? launch_t3_sse<2, true, RC>(shots, bshot, nblk, it, roles, nst, rx, s) : launch_t3_sse<2, false, RC>(shots, bshot, nblk, it, roles, nst, rx, s);
    case 3: return sse ? launch_t3_sse<3, true, RC>(shots, bshot, nblk, it, roles, nst, rx, s) : launch_t3_sse<3, false, RC>(shots, bshot, nblk, it, roles, nst, rx, s);
    case 4: return sse ? launch_t3_sse<4, true, RC>(shots, bshot, nblk, it, roles, nst, rx, s) : launch_t3_sse<4, false, RC>(shots, bshot, nblk, it, roles, nst, rx, s);
    case 5: return sse ? launch_t3_sse<5, true, RC>(shots, bshot, nblk, it, roles, nst, rx, s) : launch_t3_sse<5, false, RC>(shots, bshot, nblk, it, roles, nst, rx, s);
    case 6: return sse ? launch_t3_sse<6, true, RC>(shots, bshot, nblk, it, roles, nst, rx, s) : launch_t3_sse<6, false, RC>(shots, bshot, nblk, it, roles, nst, rx, s);
    default: return (int)hipErrorInvalidValue;
    }
}

template <int NOP>
int launch_t2_nop(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int roles, int two, int sse,
                  int rx, hipStream_t s)
{
    return sse ? launch_t2_sse<NOP, true>(shots, bshot, nblk, it, roles, two, rx, s)
               : launch_t2_sse<NOP, false>(shots, bshot, nblk, it, roles, two, rx, s);
}

template <int NOP, bool RC>
int launch_batch_nop(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int sse, int rx, hipStream_t s)
{
    return sse ? launch_batch_sse<NOP, true, RC>(shots, bshot, nblk, it, rx, s)
               : launch_batch_sse<NOP, false, RC>(shots, bshot, nblk, it, rx, s);
}

template <bool RC>
int launch_batch(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int nop, int sse, int rx,
                 hipStream_t s)
{
    switch (nop) {
    case 1: return launch_batch_nop<1, RC>(shots, bshot, nblk, it, sse, rx, s);
    case 2: return launch_batch_nop<2, RC>(shots, bshot, nblk, it, sse, rx, s);
    case 3: return launch_batch_nop<3, RC>(shots, bshot, nblk, it, sse, rx, s);
    case 4: return launch_batch_nop<4, RC>(shots, bshot, nblk, it, sse, rx, s);
    case 5: return launch_batch_nop<5, RC>(shots, bshot, nblk, it, sse, rx, s);
    case 6: return launch_batch_nop<6, RC>(shots, bshot, nblk, it, sse, rx, s);
    default: return (int)hipErrorInvalidValue;
    }
}

} // namespace

extern "C" int awefd_launch_cart_batch2(const awefd_cart_args *shots, const int *bshot, int nblk, int it,
                                        int nop, int sse, int rx, void *stream)
{
    return launch_batch<false>(shots, bshot, nblk, it, nop, sse, rx, (hipStream_t)stream);
}

extern "C" int awefd_launch_cart_batch_t2(const awefd_cart_args *shots, const int *bshot, int nblk, int it,
                                          int roles, int two, int nop, int sse, int rx, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    switch (nop) {
    case 1: return launch_t2_nop<1>(shots, bshot, nblk, it, roles, two, sse, rx, s);
    case 2: return launch_t2_nop<2>(shots, bshot, nblk, it, roles, two, sse, rx, s);
    case 3: return launch_t2_nop<3>(shots, bshot, nblk, it, roles, two, sse, rx, s);
    case 4: return launch_t2_nop<4>(shots, bshot, nblk, it, roles, two, sse, rx, s);
    case 5: return launch_t2_nop<5>(shots, bshot, nblk, it, roles, two, sse, rx, s);
    case 6: return launch_t2_nop<6>(shots, bshot, nblk, it, roles, two, sse, rx, s);
    default: return (int)hipErrorInvalidValue;
    }
}

extern "C" int awefd_launch_cart_batch_t3(const awefd_cart_args *shots, const int *bshot, int nblk, int it,
                                          int roles, int nst, int nop, int sse, int rx, void *stream)
{
    return launch_t3<false>(shots, bshot, nblk, it, roles, nst, nop, sse, rx, (hipStream_t)stream);
}

extern "C" int awefd_launch_cart_batch_t3rc(const awefd_cart_args *shots, const int *bshot, int nblk, int n,
                                            int nop, int sse, int rx, void *stream)
{
    return launch_t3<true>(shots, bshot, nblk, n, 0, 3, nop, sse, rx, (hipStream_t)stream);
}

extern "C" int awefd_launch_set_vel(const float *d_vel, float *dvel, const int n[3], const int nbs[3], int ld,
                                    float dt2, int fsnb, void *stream)
{
    const long np = (long)(n[0] + 2 * nbs[0]) * (n[1] + 2 * nbs[1]) * (n[2] + 2 * nbs[2]);
    const unsigned g = (unsigned)min((np + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_set_vel, dim3(g), dim3(256), 0, (hipStream_t)stream, d_vel, dvel, n[0], n[1], n[2], nbs[0],
                       nbs[1], nbs[2], ld, dt2, fsnb);
    return (int)hipGetLastError();
}

extern "C" int awefd_launch_cart_batch_rc(const awefd_cart_args *shots, const int *bshot, int nblk, int n,
                                          int nop, int sse, int rx, void *stream)
{
    return launch_batch<true>(shots, bshot, nblk, n, nop, sse, rx, (hipStream_t)stream);
}

extern "C" int awefd_launch_cart_step(const awefd_cart_args *a, int nop, int ndim, int sse,
                                      int varden, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const bool vd = varden != 0;
    if (vd && !a->rho) return (int)hipErrorInvalidValue;
    if (a->Pn && ndim != 2) return (int)hipErrorInvalidValue;
    switch (nop) {
    case 1: return launch_step_nop<1>(a, ndim, sse, vd, s);
    case 2: return launch_step_nop<2>(a, ndim, sse, vd, s);
    case 3: return launch_step_nop<3>(a, ndim, sse, vd, s);
    case 4: return launch_step_nop<4>(a, ndim, sse, vd, s);
    case 5: return launch_step_nop<5>(a, ndim, sse, vd, s);
    case 6: return launch_step_nop<6>(a, ndim, sse, vd, s);
    default: return (int)hipErrorInvalidValue;
    }
}

extern "C" int awefd_cart_step_occupancy(int nop, int ndim, int sse)
{
    switch (nop) {
    case 1: return occupancy_nop<1>(ndim, sse);
    case 2: return occupancy_nop<2>(ndim, sse);
    case 3: return occupancy_nop<3>(ndim, sse);
    case 4: return occupancy_nop<4>(ndim, sse);
    case 5: return occupancy_nop<5>(ndim, sse);
    case 6: return occupancy_nop<6>(ndim, sse);
    default: return 0;
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

extern "C" int awefd_launch_cart_hybrid(const awefd_cart_args *a, int nop, int ndim, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const long nl[3] = {(long)a->n2 * a->n3, (long)a->n1 * a->n3, (long)a->n1 * a->n2};
    hipLaunchKernelGGL(k_cart_hybrid<0>, dim3((unsigned)((2 * nl[0] + 255) / 256)), dim3(256), 0, s, *a, nop);
    hipLaunchKernelGGL(k_cart_hybrid<1>, dim3((unsigned)((2 * nl[1] + 255) / 256)), dim3(256), 0, s, *a, nop);
    if (ndim == 3)
        hipLaunchKernelGGL(k_cart_hybrid<2>, dim3((unsigned)((2 * nl[2] + 255) / 256)), dim3(256), 0, s, *a, nop);
    return (int)hipGetLastError();
}

extern "C" int awefd_launch_fill(float *p, long n, float v, void *stream)
{
    if (n <= 0) return 0;
    long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, p, n, v);
    return (int)hipGetLastError();
}
