/*
 * Polar (2-D) and spherical (3-D) acoustic FD kernels for gfx950 (MI355X).
 *
 * One time step of the reference CPU loop (MAWEFDspher2Dcpu.c:264-317,
 * MAWEFDspher3Dcpu.c:302-359) is, in order: inject into the current field,
 * solve for the new field, shift the three fields, one-way boundary on the new
 * field, sponge on the new and on the (now) previous field, free surface,
 * extract.  Here it is four launches:
 *
 *   k_spher_inject   inject_sources_{2D,3D}[_const] (spher_utils.c:106-322):
 *                    one thread per injected node, contributions in source order
 *   k_spher_step3 /  solve_3D / solve_2D (spher_utils.c:329-460), fused with the
 *   k_spher_step2    sponge of the new field and the pending sponge of the
 *                    previous one; the new field overwrites the previous one in
 *                    place (it is read at its own node only), so the shift is a
 *                    pointer swap
 *   k_spher_shell    onewayBC_{2D,3D} (spher_utils.c:518-637), then the sponge
 *                    and the free surface, on the nodes near the faces
 *   k_spher_extract  extract_{2D,3D} (spher_utils.c:789-856)
 *
 * Sponge of the previous field: spongeBC scales both fields every step
 * (MAWEFDspher3Dcpu.c:333-334).  The previous field is read once more, at its
 * own node, by the next solve, so its factor is applied there ("pending").
 *
 * One-way boundary: the reference sweeps the whole grid serially (r outer,
 * then theta, then phi), and each rule reads the new field at the node's
 * inward neighbour.  Rules towards a low face read a neighbour the sweep has
 * not reached yet (its value before the boundary); rules towards a high face
 * read one the sweep has finished (its final value).  The last rule that
 * applies at a node decides its value.  The step kernel therefore stores the
 * pre-boundary values of every node within AWEFD_SPHER_SHELL layers of a face
 * in a separate array S, and the shell kernel evaluates each node's final
 * value by following its chain of high-face reads inward from S: no node's
 * result depends on the order in which the GPU visits nodes, and the values
 * are those of the serial sweep.
 *
 * Arithmetic: the reference's float expressions in its order (every divide an
 * IEEE divide, -ffp-contract=off), the double metric terms of solve_3D from
 * per-(theta, r) double tables computed on the host with the reference's
 * expressions, the sponge as (float)((double)p * damp).  The results are bit
 * for bit the reference's.
 */
#include <hip/hip_runtime.h>
#include "awefd_hip.h"
#include "awefd_internal.h"

#pragma clang fp contract(off)

namespace {

using namespace awefd;
constexpr int NWAVE = AWEFD_SPHER_NWAVE;
constexpr int NTHR = NWAVE * WAVE;
constexpr int RX = AWEFD_SPHER_RX;          /* theta rows per wave */
constexpr int TX = NWAVE * RX;
constexpr int W = AWEFD_SPHER_SHELL;
constexpr int NOP = AWEFD_SPHER_NOP;
constexpr int NT = 2;                       /* non-temporal policy */
constexpr int DTAB_LDS = 512;               /* sponge table entries staged in LDS */
static_assert(TX == AWEFD_SPHER_TX, "tile width");

/* Sponge depth of a node (spongeBC_2D/3D, spher_utils.c:641-740), -1 outside
 * the strips.  Low faces first; within a side the first of r, phi, theta. */
template <int NDIM>
__device__ __forceinline__ int damp_index(int ir, int ith, int iph, int n1, int n2, int n3, int nb)
{
    const bool plo = NDIM == 3 && iph < nb, phi = NDIM == 3 && iph > n3 - nb;
    if (ir < nb || plo || ith < nb) return ir < nb ? nb - ir : (plo ? nb - iph : nb - ith);
    if (ir > n1 - nb || phi || ith > n2 - nb)
        return ir > n1 - nb ? ir - (n1 - nb) : (phi ? iph - (n3 - nb) : ith - (n2 - nb));
    return -1;
}

__device__ __forceinline__ float damp(float v, int i, const double *tab)
{
    return i < 0 ? v : (float)((double)v * tab[i]);
}

template <int NDIM>
__device__ __forceinline__ bool in_shell(int ir, int ith, int iph, int n1, int n2, int n3)
{
    bool s = ir < W || ir >= n1 - W || ith < W || ith >= n2 - W;
    if (NDIM == 3) s = s || iph < W || iph >= n3 - W;
    return s;
}

/* Last one-way rule at a node: 0 none, 1 theta-low, 2 theta-high, 3 r-low,
 * 4 r-high, 5 phi-low, 6 phi-high.  iop is the depth into the first face (r,
 * theta, phi) the node lies within NOP of; the rules then test the layer
 * (exactly in 2-D, spher_utils.c:538-563; up to the depth in 3-D,
 * spher_utils.c:593-631). */
template <int NDIM>
__device__ __forceinline__ int oneway_rule(int ir, int ith, int iph, int n1, int n2, int n3)
{
    int iop;
    if (ir < NOP) iop = ir;
    else if (ir > n1 - NOP) iop = n1 - ir;
    else if (ith < NOP) iop = ith;
    else if (ith > n2 - NOP) iop = n2 - ith;
    else if (NDIM == 3 && iph < NOP) iop = iph;
    else if (NDIM == 3 && iph > n3 - NOP) iop = n3 - iph;
    else iop = 0;
    int r = 0;
    if (NDIM == 2) {
        if (ith == NOP - iop) r = 1;
        if (ith == n2 - NOP + iop - 1) r = 2;
        if (ir == NOP - iop) r = 3;
        if (ir == n1 - NOP + iop - 1) r = 4;
    } else {
        if (ith <= NOP - iop) r = 1;
        if (ith >= n2 - NOP + iop - 1) r = 2;
        if (ir <= NOP - iop) r = 3;
        if (ir >= n1 - NOP + iop - 1) r = 4;
        if (iph <= NOP - iop) r = 5;
        if (iph >= n3 - NOP + iop - 1) r = 6;
    }
    return r;
}

struct node3 {
    int r, t, p;
};

__device__ __forceinline__ node3 inward(node3 c, int rule)
{
    switch (rule) {
    case 1: c.t++; break;
    case 2: c.t--; break;
    case 3: c.r++; break;
    case 4: c.r--; break;
    case 5: c.p++; break;
    default: c.p--; break;
    }
    return c;
}

__device__ __forceinline__ float oneway_coef(const awefd_spher_args &a, node3 c, int rule)
{
    switch (rule) {
    case 1: return a.bthl[(long)c.p * a.n1 + c.r];
    case 2: return a.bthh[(long)c.p * a.n1 + c.r];
    case 3: return a.brl[(long)c.p * a.n2 + c.t];
    case 4: return a.brh[(long)c.p * a.n2 + c.t];
    case 5: return a.bphl[(long)c.t * a.n1 + c.r];
    default: return a.bphh[(long)c.t * a.n1 + c.r];
    }
}

__device__ __forceinline__ long offs(const awefd_spher_args &a, node3 c)
{
    return ((long)c.p * a.n2 + c.t) * a.ld + c.r;
}

/* Final (pre-sponge) value of a shell node after the serial one-way sweep.
 * uo(a) = um(n) + (um(a) - uo(n)) * b with n the inward neighbour of the
 * node's last rule; uo(n) is n's pre-boundary value S(n) for a low-face rule
 * and n's final value for a high-face rule.  The chain of high-face rules is
 * walked inward to its end, then evaluated outward in the reference's order.
 * This form walks the chain again for every level; shell_final below keeps
 * the levels in registers and falls back to it for chains deeper than that. */
template <int NDIM>
__device__ __forceinline__ float shell_final_walk(const awefd_spher_args &a, node3 c0)
{
    const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
    int rule = oneway_rule<NDIM>(c0.r, c0.t, c0.p, n1, n2, n3);
    node3 c = c0;
    int depth = 0;
    while ((rule == 2 || rule == 4 || rule == 6) && depth < AWEFD_SPHER_MAXCHAIN) {
        c = inward(c, rule);
        rule = oneway_rule<NDIM>(c.r, c.t, c.p, n1, n2, n3);
        depth++;
    }
    float x;
    if (rule == 0) {
        x = a.S[offs(a, c)];
    } else {
        const node3 n = inward(c, rule);
        x = a.P[offs(a, n)] + (a.P[offs(a, c)] - a.S[offs(a, n)]) * oneway_coef(a, c, rule);
    }
    for (int l = depth - 1; l >= 0; l--) {
        node3 cl = c0;
        int rl = oneway_rule<NDIM>(cl.r, cl.t, cl.p, n1, n2, n3);
        for (int m = 0; m < l; m++) {
            cl = inward(cl, rl);
            rl = oneway_rule<NDIM>(cl.r, cl.t, cl.p, n1, n2, n3);
        }
        const node3 n = inward(cl, rl);
        x = a.P[offs(a, n)] + (a.P[offs(a, cl)] - x) * oneway_coef(a, cl, rl);
    }
    return x;
}

/* The same value, the chain's levels kept in registers: level l holds the
 * current field at chain node c_l, at its inward neighbour c_(l+1) and the
 * one-way coefficient of c_l's rule, all loaded on the way in (no load waits
 * for another), then x = P(c_(l+1)) + (P(c_l) - x) b_l outward from the
 * chain's end, the same operations in the same order.  Chains are at most 8
 * deep, and on the 460 x 572 x 572 shell 99.8 % of the nodes have at most
 * 2 levels (6.66 M none, 1.67 M one, 1.67 M two, 20 K deeper); a deeper
 * chain takes the walk above. */
constexpr int CHAIN_REG = 2;

__device__ __forceinline__ bool high_rule(int r) { return r == 2 || r == 4 || r == 6; }

template <int NDIM>
__device__ __forceinline__ float shell_final(const awefd_spher_args &a, node3 c0)
{
    const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
    int rule = oneway_rule<NDIM>(c0.r, c0.t, c0.p, n1, n2, n3);
    node3 c = c0;
    float pc[CHAIN_REG], pn[CHAIN_REG], bc[CHAIN_REG];
    float pcur = rule != 0 ? a.P[offs(a, c0)] : 0.f;
    int depth = 0;
#pragma unroll
    for (int l = 0; l < CHAIN_REG; l++) {
        if (!high_rule(rule)) break;
        const node3 n = inward(c, rule);
        pc[l] = pcur;
        bc[l] = oneway_coef(a, c, rule);
        pcur = a.P[offs(a, n)];
        pn[l] = pcur;
        c = n;
        rule = oneway_rule<NDIM>(c.r, c.t, c.p, n1, n2, n3);
        depth = l + 1;
    }
    if (high_rule(rule)) return shell_final_walk<NDIM>(a, c0);
    float x;
    if (rule == 0) {
        x = a.S[offs(a, c)];
    } else {
        const node3 n = inward(c, rule);
        x = a.P[offs(a, n)] + (pcur - a.S[offs(a, n)]) * oneway_coef(a, c, rule);
    }
#pragma unroll
    for (int l = CHAIN_REG - 1; l >= 0; l--)
        if (l < depth) x = pn[l] + (pc[l] - x) * bc[l];
    return x;
}

/* The shell: every node within W layers of a face, enumerated as the r-face
 * slabs, then the theta-face slabs without r-face nodes, then the phi-face
 * slabs without either; r is the fastest index in each group. */
struct span {
    int lo, hs, n;   /* [0, lo) and [hs, n) form the set */
    __device__ int count() const { return lo + (n - hs); }
    __device__ int at(int k) const { return k < lo ? k : hs + (k - lo); }
};

__device__ __forceinline__ span face_span(int n)
{
    span s;
    s.lo = min(W, n);
    s.hs = max(W, n - W);
    s.n = n;
    return s;
}

/* n / d for n < 2^31 by the multiplier the host prepared (awefd_magic_div) */
__device__ __forceinline__ unsigned mdiv(unsigned n, awefd_udiv d)
{
    return (__umulhi(n, d.m) + n) >> d.l;
}

/* Shell node number t (r-face group, then theta-face, then phi-face). */
template <int NDIM>
__device__ __forceinline__ void shell_node(const awefd_spher_args &a, unsigned t)
{
    const span R = face_span(a.n1), T = face_span(a.n2), F = face_span(a.n3);
    node3 c;
    if (t < a.nsr) {
        const unsigned q = mdiv(t, a.dv_rc);            /* / R.count() */
        const unsigned p = mdiv(q, a.dv_n2);
        c.r = R.at((int)(t - q * (unsigned)R.count()));
        c.t = (int)(q - p * (unsigned)a.n2);
        c.p = (int)p;
    } else if ((t -= a.nsr) < a.nst) {
        const unsigned q = mdiv(t, a.dv_rin);           /* / (n1 - R.count()) */
        const unsigned p = mdiv(q, a.dv_tc);            /* / T.count() */
        c.r = R.lo + (int)(t - q * (unsigned)(a.n1 - R.count()));
        c.t = T.at((int)(q - p * (unsigned)T.count()));
        c.p = (int)p;
    } else if (NDIM == 3 && (t -= a.nst) < a.nsp) {
        const unsigned q = mdiv(t, a.dv_rin);
        const unsigned p = mdiv(q, a.dv_tin);           /* / (n2 - T.count()) */
        c.r = R.lo + (int)(t - q * (unsigned)(a.n1 - R.count()));
        c.t = T.lo + (int)(q - p * (unsigned)(a.n2 - T.count()));
        c.p = F.at((int)p);
    } else {
        return;
    }
    float v;
    if (NDIM == 3 && a.fsrf) {
        v = 0.f;  /* freeBC_3D zeroes the whole field (spher_utils.c:765-785) */
    } else {
        v = shell_final<NDIM>(a, c);
        v = damp(v, damp_index<NDIM>(c.r, c.t, c.p, a.n1, a.n2, a.n3, a.nb), a.damp);
        /* freeSurf_2D (spher_utils.c:745-763): r > n1 - nb */
        if (NDIM == 2 && a.fsrf && c.r > a.n1 - a.nb) v = 0.f;
    }
    a.Q[offs(a, c)] = v;
}

template <int NDIM>
__global__ __launch_bounds__(256) void k_spher_shell(const awefd_spher_args a)
{
    shell_node<NDIM>(a, blockIdx.x * 256u + threadIdx.x);
}

/* Injection of step it into field F at injection node i; a node's
 * contributions are added in the reference's source order. */
__device__ __forceinline__ void inject_node(const awefd_spher_args &a, float *F, int it, int i)
{
    const long o = a.inode[i];
    const float wa = a.wav[it];
    float v = F[o];
    for (int e = a.ibeg[i]; e < a.ibeg[i + 1]; e++) {
        float inc = wa * a.iw[2 * e];
        inc = inc * a.iw[2 * e + 1];   /* reflectivity (1 for constant injection) */
        v = v + inc;
    }
    F[o] = v;
}

/* Injection into the current field before the solve, one thread per node. */
__global__ __launch_bounds__(256) void k_spher_inject(const awefd_spher_args a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.ninode) return;
    inject_node(a, a.P, a.it, i);
}

/* Receiver gather from the finished field of the step. */
template <int NDIM>
__global__ __launch_bounds__(256) void k_spher_extract(const awefd_spher_args a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= a.nr) return;
    const long sy = (long)a.ld * a.n2;
    const long b = (long)a.rj[3 * r + 2] * sy + (long)a.rj[3 * r + 1] * a.ld + a.rj[3 * r];
    constexpr int NC = NDIM == 3 ? 8 : 4;
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < NC; q++) {
        const long o = b + (q & 1) * (long)a.ld + ((q >> 1) & 1) + ((q >> 2) & 1) * sy;
        const float term = a.Q[o] * a.rw[8 * r + q];
        s = q == 0 ? term : s + term;
    }
    a.trow[r] = s;
}

/* x / c for a launch divisor c.  FD: (float)((double)x * rc) with rc = 1/c in
 * double, which the host has shown to give the IEEE quotient for every float
 * x (awefd__exact_recip): one double multiplication and two conversions
 * instead of the division's scale, reciprocal, four fused multiply-adds and
 * fix-up. */
template <bool FD>
__device__ __forceinline__ float cdiv(float x, float c, double rc)
{
    if (FD) return (float)((double)x * rc);
    return x / c;
}

/* 2-D polar step: one thread per node (polar grids are small). */
__global__ __launch_bounds__(256) void k_spher_step2(const awefd_spher_args a)
{
    const int n1 = a.n1, n2 = a.n2;
    const int ir = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ith = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ir >= n1 || ith >= n2) return;
    const long o = (long)ith * a.ld + ir;
    const float c = a.P[o];
    float lap = 0.f;
    if (ir >= NOP && ir < n1 - NOP && ith >= NOP && ith < n2 - NOP) {
        const float *pc = a.pc + 5 * ir;
        const float cra = (pc[0] * a.P[o - 1] + pc[1] * c) + pc[2] * a.P[o + 1];
        const float cth = (pc[3] * a.P[o - a.ld] + pc[4] * c) + pc[3] * a.P[o + a.ld];
        lap = cra + cth;
    }
    const int di = damp_index<2>(ir, ith, 0, n1, n2, 1, a.nb);
    const float pp = damp(a.Q[o], di, a.damp);
    float f = (a.vv[o] * lap + 2.f * c) - pp;
    if (in_shell<2>(ir, ith, 0, n1, n2, 1)) {
        a.S[o] = f;
        return;
    }
    f = damp(f, di, a.damp);
    if (a.fsrf && ir > n1 - a.nb) f = 0.f;
    a.Q[o] = f;
}

/* 3-D spherical step.  A block covers 64 r nodes (one per lane) x TX theta
 * rows (RX per wave) and streams along phi, holding planes phi-1, phi, phi+1
 * of the current field in a register ring of 4 slots (the 4th receives the
 * next plane's loads while this one is computed).  r neighbours move across
 * lanes by DPP; lanes 0 and 63 fetch the nodes beyond the tile.  theta
 * neighbours are the wave's own rows or the two rows beyond them. */
__device__ __forceinline__ unsigned soff(int t, int p, int pbase, unsigned rowb, unsigned psz)
{
    return (unsigned)t * rowb + (unsigned)(p - pbase) * psz;
}

__device__ __forceinline__ void load_plane(float (&q)[RX], float (&v)[RX], float (&zh)[RX], float (&hal)[2],
                                           rsrc_t rQ, rsrc_t rV, rsrc_t rP, unsigned vz, unsigned vzh,
                                           int x0, int p, int pbase, unsigned rowb, unsigned psz)
{
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const unsigned s = soff(x0 + j, p, pbase, rowb, psz);
        q[j] = bld<NT>(rQ, vz, s);
        v[j] = bld<NT>(rV, vz, s);
        zh[j] = bld(rP, vzh, s);
    }
    hal[0] = bld(rP, vz, soff(x0 - 1, p, pbase, rowb, psz));
    hal[1] = bld(rP, vz, soff(x0 + RX, p, pbase, rowb, psz));
}

/* Six waves per SIMD: 75 VGPRs and no spill (the metric factors in LDS);
 * seven (72 VGPRs) or eight (64) spill and run 17 % and 50 % slower on the
 * 460 x 572 x 572 shell. */
template <bool FD>
__global__ __launch_bounds__(NTHR, 6) void k_spher_step3(const awefd_spher_args a)
{
    constexpr int QN = 4;
    const int tid = threadIdx.x;
    const int lb = xcd_remap(blockIdx.x, a.nblk);
    const int tzi = lb % a.ntz;
    const int rr = lb / a.ntz;
    const int txi = rr % a.ntx;
    const int tyi = rr / a.ntx;
    const int lane = tid & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / WAVE);
    const int n1 = a.n1, n2 = a.n2, n3 = a.n3, nb = a.nb;
    const int z0 = tzi * WAVE;
    const int ir = z0 + lane;
    const bool zin = ir < n1;
    const int x0 = txi * TX + wv * RX;
    const int y0 = (int)((long)tyi * n3 / a.nty);
    const int y1 = (int)((long)(tyi + 1) * n3 / a.nty);
    const long sy = (long)a.ld * n2;
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned psz = (unsigned)(sy * 4);
    const int pbase = max(y0 - 1, 0);
    const int ptop = min(y1 + 1, n3);
    const unsigned nbytes = (unsigned)(ptop - pbase) * psz;
    const rsrc_t rP = mk_rsrc(a.P + pbase * sy, nbytes);
    const rsrc_t rQ = mk_rsrc(a.Q + pbase * sy, nbytes);
    const rsrc_t rS = mk_rsrc(a.S + pbase * sy, nbytes);
    const rsrc_t rV = mk_rsrc(a.vv + pbase * sy, nbytes);

    const unsigned vz = zin ? (unsigned)ir * 4u : OOB;
    int zz = -1;
    if (lane == 0) zz = z0 - 1;
    else if (lane == WAVE - 1) zz = z0 + WAVE;
    const unsigned vzh = (zz >= 0 && zz < n1) ? (unsigned)zz * 4u : OOB;

    /* per-lane and per-row constants */
    const bool rint = ir >= NOP && ir < n1 - NOP;
    const bool rsh = ir < W || ir >= n1 - W;
    const float m2 = zin ? a.mr2[ir] : 0.f;
    const float m1 = zin ? a.mr1[ir] : 0.f;
    /* the double metric factors of this lane's nodes, read once per block
     * into LDS slots only this thread uses: 16 VGPRs fewer than holding
     * them in registers, for six waves per SIMD without spills */
    __shared__ double mts[2][TX][WAVE];
    /* the sponge table (nb + 1 doubles) in LDS when it fits: a node's factor
     * is a load whose index depends on the node, and from L2 its latency
     * cost a quarter of the step */
    __shared__ double dtab[DTAB_LDS];
    const bool dlds = a.nb + 1 <= DTAB_LDS;                 /* block-uniform */
    if (dlds)
        for (int i = tid; i <= a.nb; i += NTHR) dtab[i] = a.damp[i];
    __syncthreads();
    unsigned tint_m = 0, tsh_m = 0, tin_m = 0;
#pragma unroll
    for (int j = 0; j < RX; j++) {
        const int ith = x0 + j;
        const bool ok = zin && ith < n2;
        mts[0][wv * RX + j][lane] = ok ? a.mt1[(long)ith * n1 + ir] : 0.0;
        mts[1][wv * RX + j][lane] = ok ? a.mt2[(long)ith * n1 + ir] : 0.0;
        tin_m |= ith < n2 ? 1u << j : 0u;
        tint_m |= (ith >= NOP && ith < n2 - NOP) ? 1u << j : 0u;
        tsh_m |= (ith < W || ith >= n2 - W) ? 1u << j : 0u;
    }
    /* wave-uniform: does the tile touch a sponge strip in r or theta? */
    const bool sp_rt = z0 < nb || z0 + WAVE - 1 > n1 - nb || x0 < nb || x0 + RX - 1 > n2 - nb;
    const float tdr = a.tdr, dr2 = a.dr2, tdt = a.tdt, dt2 = a.dt2, dp2 = a.dp2;

    float q[RX][QN];
    float bq_[2][RX], bv_[2][RX], bz_[2][RX], bh_[2][2];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < RX; j++) q[j][k] = bld(rP, vz, soff(x0 + j, y0 - 1 + k, pbase, rowb, psz));
    load_plane(bq_[0], bv_[0], bz_[0], bh_[0], rQ, rV, rP, vz, vzh, x0, y0, pbase, rowb, psz);

    for (int yg = y0; yg < y1; yg += QN) {
#pragma unroll
        for (int R = 0; R < QN; R++) {
            const int y = yg + R;
            if (y >= y1) break;
            const int B = R & 1;
            if (y + 1 < y1) {
#pragma unroll
                for (int j = 0; j < RX; j++)
                    q[j][(R + 3) % QN] = bld(rP, vz, soff(x0 + j, y + 2, pbase, rowb, psz));
                load_plane(bq_[B ^ 1], bv_[B ^ 1], bz_[B ^ 1], bh_[B ^ 1], rQ, rV, rP, vz, vzh, x0, y + 1,
                           pbase, rowb, psz);
            }
            const bool pint = y >= NOP && y < n3 - NOP;
            const bool psh = y < W || y >= n3 - W;
            const bool spw = sp_rt || y < nb || y > n3 - nb;   /* wave-uniform */
#pragma unroll
            for (int j = 0; j < RX; j++) {
                if (!((tin_m >> j) & 1u)) continue;
                const float c = q[j][(R + 1) % QN];
                const float rm = shr1(bz_[B][j], c);
                const float rp = shl1(bz_[B][j], c);
                const float tm = j > 0 ? q[j - 1][(R + 1) % QN] : bh_[B][0];
                const float tp = j < RX - 1 ? q[j + 1][(R + 1) % QN] : bh_[B][1];
                const float fm = q[j][R % QN];
                const float fp = q[j][(R + 2) % QN];
                /* solve_3D (spher_utils.c:412-448), every lane, then select */
                const float pra = cdiv<FD>(rp - rm, tdr, a.rtdr);
                const float ppra = cdiv<FD>((rp - 2.f * c) + rm, dr2, a.rdr2);
                const float pth = cdiv<FD>(tp - tm, tdt, a.rtdt);
                const float ppth = cdiv<FD>((tp - 2.f * c) + tm, dt2, a.rdt2);
                const float ppph = cdiv<FD>((fp - 2.f * c) + fm, dp2, a.rdp2);
                const float s1 = m2 * pra + ppra;
                double acc = (double)s1 + mts[0][wv * RX + j][lane] * (double)pth;
                acc = acc + (double)(m1 * ppth);
                acc = acc + mts[1][wv * RX + j][lane] * (double)ppph;
                const bool inter = rint && ((tint_m >> j) & 1u) && pint;
                const float lap = inter ? (float)acc : 0.f;
                float pp = bq_[B][j];
                int di = -1;
                if (spw) {
                    di = damp_index<3>(ir, x0 + j, y, n1, n2, n3, nb);
                    pp = dlds ? damp(pp, di, dtab) : damp(pp, di, a.damp);
                }
                const float f0 = (bv_[B][j] * lap + 2.f * c) - pp;
                float f = spw ? (dlds ? damp(f0, di, dtab) : damp(f0, di, a.damp)) : f0;
                if (a.fsrf) f = 0.f;   /* freeBC_3D zeroes the whole field */
                const bool sh = rsh || ((tsh_m >> j) & 1u) || psh;
                const unsigned s = soff(x0 + j, y, pbase, rowb, psz);
                /* shell nodes keep their pre-boundary value in S */
                bst<NT>(rQ, sh ? OOB : vz, s, f);
                bst<NT>(rS, sh ? vz : OOB, s, f0);
            }
        }
    }
}

/* 3-D spherical step in columns: one wave per block owns 64 r nodes (one per
 * lane) of CRX theta rows and streams a chunk of phi planes, with every
 * per-plane input (previous field, (dt v)^2, the r-halo value of each row,
 * the two theta-halo rows) loaded CPD planes ahead into rotating register
 * slots.  Block b runs on XCD b % 8; XCD k owns the theta tiles
 * [ntx k / 8, ntx (k + 1) / 8), taken r tile fastest, then theta, then phi
 * chunk, so the theta-halo rows and the chunk prologues of neighbouring
 * columns are re-read from that XCD's L2.  The arithmetic, its order and the
 * selections are k_spher_step3's, node for node. */
#ifndef AWEFD_SPHER_CPD
#define AWEFD_SPHER_CPD 3
#endif
constexpr int CRX = AWEFD_SPHER_CRX;
constexpr int CPD = AWEFD_SPHER_CPD;
constexpr int CNS = CPD + 1;                               /* stage slots */
constexpr int CQR = ((3 + CPD + CPD) / CNS) * CNS;         /* ring slots: phi-1 .. phi+1+CPD */

__device__ __forceinline__ void col_plane(float (&q)[CRX], float (&v)[CRX], float (&zh)[CRX], float (&hal)[2],
                                          rsrc_t rQ, rsrc_t rV, rsrc_t rP, unsigned vz, unsigned vzh, int x0, int p,
                                          int pbase, unsigned rowb, unsigned psz)
{
#pragma unroll
    for (int j = 0; j < CRX; j++) {
        const unsigned s = soff(x0 + j, p, pbase, rowb, psz);
        q[j] = bld<NT>(rQ, vz, s);
        v[j] = bld<NT>(rV, vz, s);
        zh[j] = bld(rP, vzh, s);
    }
    hal[0] = bld(rP, vz, soff(x0 - 1, p, pbase, rowb, psz));
    hal[1] = bld(rP, vz, soff(x0 + CRX, p, pbase, rowb, psz));
}

template <bool FD>
__global__ __launch_bounds__(WAVE) void k_spher_col3(const awefd_spher_args a)
{
    extern __shared__ double col_dtab[];
    const int lane = threadIdx.x;
    const int bid = blockIdx.x;
    const int xcd = bid & 7, slot = bid >> 3;
    const int xb = (int)((long)a.ntx * xcd / 8), nxk = (int)((long)a.ntx * (xcd + 1) / 8) - xb;
    if (nxk == 0) return;
    const int tzi = slot % a.ntz;
    const int txi = xb + (slot / a.ntz) % nxk;
    const int tyi = slot / (a.ntz * nxk);
    if (tyi >= a.nty) return;
    const int n1 = a.n1, n2 = a.n2, n3 = a.n3, nb = a.nb;
    const int z0 = tzi * WAVE;
    const int ir = z0 + lane;
    const bool zin = ir < n1;
    const int x0 = txi * CRX;
    const int y0 = (int)((long)tyi * n3 / a.nty);
    const int y1 = (int)((long)(tyi + 1) * n3 / a.nty);
    const long sy = (long)a.ld * n2;
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned psz = (unsigned)(sy * 4);
    const int pbase = max(y0 - 1, 0);
    const int ptop = min(y1 + 1, n3);
    const unsigned nbytes = (unsigned)(ptop - pbase) * psz;
    const rsrc_t rP = mk_rsrc(a.P + pbase * sy, nbytes);
    const rsrc_t rQ = mk_rsrc(a.Q + pbase * sy, nbytes);
    const rsrc_t rS = mk_rsrc(a.S + pbase * sy, nbytes);
    const rsrc_t rV = mk_rsrc(a.vv + pbase * sy, nbytes);

    const unsigned vz = zin ? (unsigned)ir * 4u : OOB;
    int zz = -1;
    if (lane == 0) zz = z0 - 1;
    else if (lane == WAVE - 1) zz = z0 + WAVE;
    const unsigned vzh = (zz >= 0 && zz < n1) ? (unsigned)zz * 4u : OOB;

    const bool rint = ir >= NOP && ir < n1 - NOP;
    const bool rsh = ir < W || ir >= n1 - W;
    const float m2 = zin ? a.mr2[ir] : 0.f;
    const float m1 = zin ? a.mr1[ir] : 0.f;
    double mt1[CRX], mt2[CRX];
    unsigned tint_m = 0, tsh_m = 0, tin_m = 0;
#pragma unroll
    for (int j = 0; j < CRX; j++) {
        const int ith = x0 + j;
        const bool ok = zin && ith < n2;
        mt1[j] = ok ? a.mt1[(long)ith * n1 + ir] : 0.0;
        mt2[j] = ok ? a.mt2[(long)ith * n1 + ir] : 0.0;
        tin_m |= ith < n2 ? 1u << j : 0u;
        tint_m |= (ith >= NOP && ith < n2 - NOP) ? 1u << j : 0u;
        tsh_m |= (ith < W || ith >= n2 - W) ? 1u << j : 0u;
    }
    /* the sponge table (nb + 1 doubles) in LDS */
    for (int i = lane; i <= nb; i += WAVE) col_dtab[i] = a.damp[i];
    __syncthreads();
    const bool sp_rt = z0 < nb || z0 + WAVE - 1 > n1 - nb || x0 < nb || x0 + CRX - 1 > n2 - nb;
    const float tdr = a.tdr, dr2 = a.dr2, tdt = a.tdt, dt2 = a.dt2, dp2 = a.dp2;

    /* current field: plane p in ring slot (p - y0 + 1) % CQR; stage slot of
     * plane p: (p - y0) % CNS */
    float q[CRX][CQR];
    float sq_[CNS][CRX], sv_[CNS][CRX], sz_[CNS][CRX], sh_[CNS][2];
#pragma unroll
    for (int k = 0; k < 2 + CPD; k++)
#pragma unroll
        for (int j = 0; j < CRX; j++) q[j][k] = bld(rP, vz, soff(x0 + j, y0 - 1 + k, pbase, rowb, psz));
#pragma unroll
    for (int k = 0; k < CPD; k++)
        col_plane(sq_[k], sv_[k], sz_[k], sh_[k], rQ, rV, rP, vz, vzh, x0, y0 + k, pbase, rowb, psz);

    for (int yg = y0; yg < y1; yg += CQR) {
#pragma unroll
        for (int R = 0; R < CQR; R++) {
            const int y = yg + R;
            if (y >= y1) break;
            const int B = R % CNS;
            const int C0 = (R + 1) % CQR;
            if (y + CPD < y1) {
                const int Bn = (R + CPD) % CNS;
#pragma unroll
                for (int j = 0; j < CRX; j++)
                    q[j][(R + 2 + CPD) % CQR] = bld(rP, vz, soff(x0 + j, y + 1 + CPD, pbase, rowb, psz));
                col_plane(sq_[Bn], sv_[Bn], sz_[Bn], sh_[Bn], rQ, rV, rP, vz, vzh, x0, y + CPD, pbase, rowb, psz);
            }
            const bool pint = y >= NOP && y < n3 - NOP;
            const bool psh = y < W || y >= n3 - W;
            const bool spw = sp_rt || y < nb || y > n3 - nb;   /* wave-uniform */
#pragma unroll
            for (int j = 0; j < CRX; j++) {
                if (!((tin_m >> j) & 1u)) continue;
                const float c = q[j][C0];
                const float rm = shr1(sz_[B][j], c);
                const float rp = shl1(sz_[B][j], c);
                const float tm = j > 0 ? q[j - 1][C0] : sh_[B][0];
                const float tp = j < CRX - 1 ? q[j + 1][C0] : sh_[B][1];
                const float fm = q[j][(C0 + CQR - 1) % CQR];
                const float fp = q[j][(C0 + 1) % CQR];
                const float pra = cdiv<FD>(rp - rm, tdr, a.rtdr);
                const float ppra = cdiv<FD>((rp - 2.f * c) + rm, dr2, a.rdr2);
                const float pth = cdiv<FD>(tp - tm, tdt, a.rtdt);
                const float ppth = cdiv<FD>((tp - 2.f * c) + tm, dt2, a.rdt2);
                const float ppph = cdiv<FD>((fp - 2.f * c) + fm, dp2, a.rdp2);
                const float s1 = m2 * pra + ppra;
                double acc = (double)s1 + mt1[j] * (double)pth;
                acc = acc + (double)(m1 * ppth);
                acc = acc + mt2[j] * (double)ppph;
                const bool inter = rint && ((tint_m >> j) & 1u) && pint;
                const float lap = inter ? (float)acc : 0.f;
                float pp = sq_[B][j];
                int di = -1;
                if (spw) {
                    di = damp_index<3>(ir, x0 + j, y, n1, n2, n3, nb);
                    pp = damp(pp, di, col_dtab);
                }
                const float f0 = (sv_[B][j] * lap + 2.f * c) - pp;
                float f = spw ? damp(f0, di, col_dtab) : f0;
                if (a.fsrf) f = 0.f;   /* freeBC_3D zeroes the whole field */
                const bool sh = rsh || ((tsh_m >> j) & 1u) || psh;
                const unsigned s = soff(x0 + j, y, pbase, rowb, psz);
                bst<NT>(rQ, sh ? OOB : vz, s, f);
                bst<NT>(rS, sh ? vz : OOB, s, f0);
            }
        }
    }
}

} // namespace

extern "C" int awefd_launch_spher_inject(const awefd_spher_args *a, void *stream)
{
    if (a->ninode <= 0) return 0;
    hipLaunchKernelGGL(k_spher_inject, dim3((a->ninode + 255) / 256), dim3(256), 0, (hipStream_t)stream, *a);
    return (int)hipGetLastError();
}

extern "C" int awefd_launch_spher_step(const awefd_spher_args *a, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (a->ndim == 3 && a->col) {
        if (a->fastdiv)
            hipLaunchKernelGGL(k_spher_col3<true>, dim3(a->nblk), dim3(WAVE), sizeof(double) * (size_t)(a->nb + 1), s, *a);
        else
            hipLaunchKernelGGL(k_spher_col3<false>, dim3(a->nblk), dim3(WAVE), sizeof(double) * (size_t)(a->nb + 1), s, *a);
    } else if (a->ndim == 3) {
        if (a->fastdiv) hipLaunchKernelGGL(k_spher_step3<true>, dim3(a->nblk), dim3(NTHR), 0, s, *a);
        else hipLaunchKernelGGL(k_spher_step3<false>, dim3(a->nblk), dim3(NTHR), 0, s, *a);
    } else {
        dim3 g((a->n1 + 63) / 64, (a->n2 + 3) / 4);
        hipLaunchKernelGGL(k_spher_step2, g, dim3(256), 0, s, *a);
    }
    return (int)hipGetLastError();
}

extern "C" int awefd_launch_spher_shell(const awefd_spher_args *a, void *stream)
{
    const long n = a->nsr + a->nst + a->nsp;
    if (n <= 0) return 0;
    const dim3 g((unsigned)((n + 255) / 256));
    if (a->ndim == 3) hipLaunchKernelGGL(k_spher_shell<3>, g, dim3(256), 0, (hipStream_t)stream, *a);
    else hipLaunchKernelGGL(k_spher_shell<2>, g, dim3(256), 0, (hipStream_t)stream, *a);
    return (int)hipGetLastError();
}

extern "C" int awefd_launch_spher_extract(const awefd_spher_args *a, void *stream)
{
    if (a->nr <= 0) return 0;
    const dim3 g((a->nr + 255) / 256);
    if (a->ndim == 3) hipLaunchKernelGGL(k_spher_extract<3>, g, dim3(256), 0, (hipStream_t)stream, *a);
    else hipLaunchKernelGGL(k_spher_extract<2>, g, dim3(256), 0, (hipStream_t)stream, *a);
    return (int)hipGetLastError();
}

extern "C" int awefd_spher_step_occupancy(int ndim)
{
    int nb = 0;
    if (ndim != 3) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spher_step3<true>, NTHR, 0) != hipSuccess) return 0;
    return nb;
}
