/*
 * Polar (2-D) and spherical (3-D) back end of libawefd (host C).
 *
 * Set-up mirrors MAWEFDspher2Dcpu.c / MAWEFDspher3Dcpu.c: edge-replicating
 * pad of the velocity, one-way coefficients from the raw padded velocity, the
 * reflectivity of each source node (expl=y), the metric terms of solve_3D
 * evaluated once per (theta, r) with the reference's expressions, and the
 * sponge factors.  The time loop launches the kernels of spher_kernels.hip.
 *
 * Axes: 1 = r (fastest), 2 = theta, 3 = phi.  The index map puts sources and
 * receivers on (z = theta, x = r, y = phi), as pt2d/pt3d do in the reference
 * (MAWEFDspher3Dcpu.c:108-116).
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "awefd_model.h"
#include "awefd_setup.h"

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return awefd__fail(AWEFD_EHIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                               __FILE__, __LINE__);                                         \
    } while (0)

/* spherical state hung off awefd_model.spher */
typedef struct {
    float *dS;                       /* shell values before the one-way boundary */
    float *d_mr2, *d_mr1;            /* 2/r, 1/(r*r) per r                        */
    double *d_mt1, *d_mt2;           /* metric terms per (theta, r)               */
    float *d_pc;                     /* polar coefficients, 5 per r               */
    float *d_ow[6];                  /* one-way tables bthl bthh brl brh bphl bphh */
    long *d_inode;
    int *d_ibeg;
    float *d_iw;
    int ninode;
    int inj_dirty;
    int *d_rrtp;                     /* receiver map in (r, theta, phi) order     */
    float *h_velp;                   /* padded raw velocity                       */
    unsigned nsr, nst, nsp;
    float tdr, dr2, tdt, dt2, dp2;
    double rtdr, rdr2, rtdt, rdt2, rdp2;
    int fastdiv;
} spher_state;

static spher_state *st(awefd_model *m) { return (spher_state *)m->spher; }
static int build_injection(awefd_model *m);

static void dfree(void *pp)
{
    void **p = (void **)pp;
    if (*p) hipFree(*p);
    *p = NULL;
}

static int up(void *dst, const void *src, size_t bytes)
{
    void **d = (void **)dst;
    dfree(d);
    if (!bytes) return AWEFD_OK;
    HIPCHK(hipMalloc(d, bytes));
    HIPCHK(hipMemcpy(*d, src, bytes, hipMemcpyHostToDevice));
    return AWEFD_OK;
}

void awefd__spher_free(awefd_model *m)
{
    spher_state *s = st(m);
    if (!s) return;
    dfree(&s->dS);
    dfree(&s->d_mr2);
    dfree(&s->d_mr1);
    dfree(&s->d_mt1);
    dfree(&s->d_mt2);
    dfree(&s->d_pc);
    for (int i = 0; i < 6; i++) dfree(&s->d_ow[i]);
    dfree(&s->d_inode);
    dfree(&s->d_ibeg);
    dfree(&s->d_iw);
    dfree(&s->d_rrtp);
    free(s->h_velp);
    free(s);
    m->spher = NULL;
}

/* nodes in [0, min(W, n)) and [max(W, n - W), n) */
static int face_count(int n)
{
    const int W = AWEFD_SPHER_SHELL;
    const int lo = n < W ? n : W;
    const int hs = n - W > W ? n - W : W;
    return lo + (n - hs);
}

/* Step tiling: 64 r x TX theta columns, phi chunks sized for ~target blocks
 * (about 20 planes for the 460 x 572 x 572 shell at 8K blocks).  Shorter
 * chunks keep more blocks resident but re-read more halo planes; measured
 * step times on that shell: 4K blocks 0.609 ms, 6K-12K 0.595, 16K 0.608,
 * 32K 0.637. */
static void spher_tiling(awefd_model *m)
{
    const int n1 = m->np[0], n2 = m->np[1], n3 = m->np[2];
    m->ntz = (n1 + 63) / 64;
    const char *ce = getenv("AWEFD_COL");
    m->col = m->ndim == 3 && !(ce && atoi(ce) == 0);
    if (m->col) {
        /* columns of 64 r x AWEFD_SPHER_CRX theta, phi chunks of about
         * AWEFD_SPHER_CHUNK planes (default 12: 0.563 ms per step on the
         * 460 x 572 x 572 shell; 6-8 planes 0.577, 16 0.567); blocks: 8 XCD
         * shares of ceil(ntx / 8) theta tiles */
        const char *ch = getenv("AWEFD_SPHER_CHUNK");
        const long chunk = ch && atol(ch) > 0 ? atol(ch) : 12;
        m->ntx = (n2 + AWEFD_SPHER_CRX - 1) / AWEFD_SPHER_CRX;
        long nty = (n3 + chunk - 1) / chunk;
        m->nty = (int)(nty < 1 ? 1 : nty > n3 ? n3 : nty);
        m->nblk = 8 * m->ntz * ((m->ntx + 7) / 8) * m->nty;
        return;
    }
    m->ntx = (n2 + AWEFD_SPHER_TX - 1) / AWEFD_SPHER_TX;
    m->nty = 1;
    if (m->ndim == 3) {
        int target = m->target_blocks > 0 ? m->target_blocks : 8192;
        const char *e = getenv("AWEFD_SPHER_BLOCKS");
        if (e && atoi(e) > 0) target = atoi(e);
        const long cols = (long)m->ntz * m->ntx;
        long nty = (target + cols - 1) / cols;
        if (nty > n3) nty = n3;
        if (nty < 1) nty = 1;
        m->nty = (int)nty;
    }
    m->nblk = m->ntz * m->ntx * m->nty;
}

int awefd__spher_create(awefd_model *m)
{
    spher_state *s = calloc(1, sizeof *s);
    if (!s) return awefd__fail(AWEFD_ENOMEM, "spherical state");
    m->spher = s;
    const int n1 = m->np[0], n2 = m->np[1], n3 = m->np[2];
    const int W = AWEFD_SPHER_SHELL;
    if (n1 < 2 * W || n2 < 2 * W || (m->ndim == 3 && n3 < 2 * W))
        return awefd__fail(AWEFD_EINVAL, "polar/spherical grids need >= %d padded nodes per axis", 2 * W);
    if (m->ndim == 3 && n2 - AWEFD_SPHER_NOP - 1 >= n3)
        /* MAWEFDspher3Dcpu.c:269 builds the phi-high one-way table at phi layer
         * n_theta_pad - NOP - 1, which must exist */
        return awefd__fail(AWEFD_EINVAL, "padded n_theta=%d needs padded n_phi > %d (one-way phi table)",
                           n2, n2 - AWEFD_SPHER_NOP - 1);
    spher_tiling(m);
    const long plane = (long)m->ld * n2 * 4;
    if (m->ndim == 3 && (n3 / m->nty + 3) * plane >= (1L << 30))
        return awefd__fail(AWEFD_EINVAL, "a %dx%d plane is too large for the step kernel's window", n1, n2);
    HIPCHK(hipMalloc((void **)&s->dS, sizeof(float) * (size_t)m->nalloc));

    /* shell groups, as k_spher_shell enumerates them (32-bit node numbers) */
    const int cr = face_count(n1), ct = face_count(n2);
    const long nsr = (long)cr * n2 * n3, nst = (long)(n1 - cr) * ct * n3;
    const long nsp = m->ndim == 3 ? (long)(n1 - cr) * (n2 - ct) * face_count(n3) : 0;
    if (nsr + nst + nsp >= (1L << 31) - 256)
        return awefd__fail(AWEFD_EINVAL, "the %dx%dx%d grid has too many shell nodes", n1, n2, n3);
    s->nsr = (unsigned)nsr;
    s->nst = (unsigned)nst;
    s->nsp = (unsigned)nsp;

    /* derivative divisors as the reference forms them (spher_utils.c:413-437) */
    const float dra = m->p.d[0], dth = m->p.d[1], dph = m->p.d[2];
    s->tdr = 2 * dra;
    s->dr2 = dra * dra;
    s->tdt = 2 * dth;
    s->dt2 = dth * dth;
    s->dp2 = dph * dph;
    s->rtdr = awefd__exact_recip(s->tdr);
    s->rdr2 = awefd__exact_recip(s->dr2);
    s->rtdt = awefd__exact_recip(s->tdt);
    s->rdt2 = awefd__exact_recip(s->dt2);
    s->rdp2 = awefd__exact_recip(s->dp2);
    const char *ev = getenv("AWEFD_SPHER_IEEEDIV");
    s->fastdiv = s->rtdr != 0.0 && s->rdr2 != 0.0 && s->rtdt != 0.0 && s->rdt2 != 0.0 && s->rdp2 != 0.0 &&
                 !(ev && atoi(ev) == 1);
    if (getenv("AWEFD_VERBOSE"))
        fprintf(stderr, "awefd: spherical derivatives %s\n",
                s->fastdiv ? "as double reciprocal products" : "by IEEE float division");

    const float ora = m->opad[0], oth = m->opad[1];
    int rc = AWEFD_OK;
    if (m->ndim == 3) {
        /* solve_3D (spher_utils.c:400-445): ra = dra*ira + ora, th = dth*ith + oth
         * in float; (2/ra) and (1/(ra*ra)) in float; cos(th)/(ra*ra*sin(th)) and
         * 1/(ra*ra*sin(th)*sin(th)) in double */
        float *mr2 = malloc(sizeof(float) * n1), *mr1 = malloc(sizeof(float) * n1);
        double *mt1 = malloc(sizeof(double) * (size_t)n1 * n2);
        double *mt2 = malloc(sizeof(double) * (size_t)n1 * n2);
        if (!mr2 || !mr1 || !mt1 || !mt2) rc = awefd__fail(AWEFD_ENOMEM, "metric tables");
        for (int ir = 0; !rc && ir < n1; ir++) {
            const float ra = dra * ir + ora;
            mr2[ir] = 2 / ra;
            mr1[ir] = 1 / (ra * ra);
        }
        for (int ith = 0; !rc && ith < n2; ith++) {
            const float th = dth * ith + oth;
            const double c = cos(th), sn = sin(th);
            for (int ir = 0; ir < n1; ir++) {
                const float ra = dra * ir + ora;
                const float r2 = ra * ra;
                mt1[(long)ith * n1 + ir] = c / (r2 * sn);
                mt2[(long)ith * n1 + ir] = 1 / (r2 * sn * sn);
            }
        }
        if (!rc) rc = up(&s->d_mr2, mr2, sizeof(float) * n1);
        if (!rc) rc = up(&s->d_mr1, mr1, sizeof(float) * n1);
        if (!rc) rc = up(&s->d_mt1, mt1, sizeof(double) * (size_t)n1 * n2);
        if (!rc) rc = up(&s->d_mt2, mt2, sizeof(double) * (size_t)n1 * n2);
        free(mr2); free(mr1); free(mt1); free(mt2);
    } else {
        /* solve_2D (spher_utils.c:355-364): per-r coefficients in float */
        float *pc = malloc(sizeof(float) * 5 * n1);
        if (!pc) return awefd__fail(AWEFD_ENOMEM, "polar coefficients");
        for (int ir = 0; ir < n1; ir++) {
            const float ra = dra * ir + ora;
            pc[5 * ir + 0] = (1 / (dra * dra)) + (1 / (2 * ra * dra));
            pc[5 * ir + 1] = (-2 / (dra * dra));
            pc[5 * ir + 2] = (1 / (dra * dra)) - (1 / (2 * ra * dra));
            pc[5 * ir + 3] = (1 / (ra * ra * dth * dth));
            pc[5 * ir + 4] = (-2 / (ra * ra * dth * dth));
        }
        rc = up(&s->d_pc, pc, sizeof(float) * 5 * n1);
        free(pc);
    }
    if (rc) return rc;

    /* sponge factors for depths 0..nb (spher_utils.c:665-667) */
    const int nb = m->nb;
    double *damp = malloc(sizeof(double) * (nb + 1));
    if (!damp) return awefd__fail(AWEFD_ENOMEM, "sponge table");
    for (int i = 0; i <= nb; i++) damp[i] = awefd_spher_damp(i, nb);
    rc = up(&m->d_damp, damp, sizeof(double) * (nb + 1));
    free(damp);
    return rc;
}

/* Model: pad, (dt*dt)*(v*v), one-way tables (MAWEFDspher2Dcpu.c:190-220,
 * MAWEFDspher3Dcpu.c:196-272). */
int awefd__spher_set_model(awefd_model *m, const float *vel)
{
    spher_state *s = st(m);
    const int n1 = m->np[0], n2 = m->np[1], n3 = m->np[2];
    const int NOP = AWEFD_SPHER_NOP;
    const long nc = m->ncell;
    float *vp = malloc(sizeof(float) * nc), *vv = malloc(sizeof(float) * nc);
    if (!vp || !vv) {
        free(vp); free(vv);
        return awefd__fail(AWEFD_ENOMEM, "padded model");
    }
    const int nbs[3] = {m->nb, m->nb, m->ndim == 3 ? m->nb : 0};
    awefd_pad(vel, vp, m->p.n, nbs);
    const float dt = m->p.dt;
    for (long i = 0; i < nc; i++) vv[i] = (dt * dt) * (vp[i] * vp[i]);
    hipError_t e = hipMemcpy2D(m->dvel, sizeof(float) * m->ld, vv, sizeof(float) * n1, sizeof(float) * n1,
                               (size_t)n2 * n3, hipMemcpyHostToDevice);
    free(vv);
    if (e != hipSuccess) {
        free(vp);
        return awefd__fail(AWEFD_EHIP, "velocity upload: %s", hipGetErrorString(e));
    }
    const float dra = m->p.d[0], dth = m->p.d[1], dph = m->p.d[2];
    const long Z = n1, Y = (long)n1 * n2;
    float *t[6];
    size_t sz[6];
    if (m->ndim == 2) {
        sz[0] = sz[1] = n1;           /* theta faces, per r */
        sz[2] = sz[3] = n2;           /* r faces, per theta */
        sz[4] = sz[5] = 1;
    } else {
        sz[0] = sz[1] = (size_t)n1 * n3;
        sz[2] = sz[3] = (size_t)n2 * n3;
        sz[4] = sz[5] = (size_t)n1 * n2;
    }
    int rc = AWEFD_OK;
    for (int i = 0; i < 6; i++)
        if (!(t[i] = malloc(sizeof(float) * sz[i]))) rc = awefd__fail(AWEFD_ENOMEM, "one-way tables");
    if (!rc && m->ndim == 2) {
        for (int ir = 0; ir < n1; ir++) {
            t[0][ir] = awefd_oneway_spher(vp[NOP * Z + ir], dt, dth);
            t[1][ir] = awefd_oneway_spher(vp[(n2 - NOP - 1) * Z + ir], dt, dth);
        }
        /* onewayBC_2D receives (bral, brah) for its (r-low, r-high) rules */
        for (int ith = 0; ith < n2; ith++) {
            t[2][ith] = awefd_oneway_spher(vp[ith * Z + NOP], dt, dra);
            t[3][ith] = awefd_oneway_spher(vp[ith * Z + n1 - NOP - 1], dt, dra);
        }
        t[4][0] = t[5][0] = 0.f;
    } else if (!rc) {
        for (int ir = 0; ir < n1; ir++)
            for (int iph = 0; iph < n3; iph++) {
                t[0][(long)iph * n1 + ir] = awefd_oneway_spher(vp[iph * Y + NOP * Z + ir], dt, dth);
                t[1][(long)iph * n1 + ir] = awefd_oneway_spher(vp[iph * Y + (n2 - NOP - 1) * Z + ir], dt, dth);
            }
        /* onewayBC_3D receives (one_brah, one_bral) for its (r-low, r-high)
         * rules (MAWEFDspher3Dcpu.c:328-330); the r-high rule reads the table it
         * is given (the reference indexes the theta-high table there, past its
         * end when n_theta_pad > n_r_pad) */
        for (int ith = 0; ith < n2; ith++)
            for (int iph = 0; iph < n3; iph++) {
                t[2][(long)iph * n2 + ith] = awefd_oneway_spher(vp[iph * Y + ith * Z + n1 - NOP - 1], dt, dra);
                t[3][(long)iph * n2 + ith] = awefd_oneway_spher(vp[iph * Y + ith * Z + NOP], dt, dra);
            }
        /* phi-high table at phi layer n_theta_pad - NOP - 1 (MAWEFDspher3Dcpu.c:269) */
        for (int ith = 0; ith < n2; ith++)
            for (int ir = 0; ir < n1; ir++) {
                t[4][(long)ith * n1 + ir] = awefd_oneway_spher(vp[NOP * Y + ith * Z + ir], dt, dph);
                t[5][(long)ith * n1 + ir] = awefd_oneway_spher(vp[(long)(n2 - NOP - 1) * Y + ith * Z + ir], dt, dph);
            }
    }
    for (int i = 0; i < 6; i++) {
        if (!rc) rc = up(&s->d_ow[i], t[i], sizeof(float) * sz[i]);
        free(t[i]);
    }
    free(s->h_velp);
    s->h_velp = vp;
    s->inj_dirty = 1;
    if (!rc && m->src_set) rc = build_injection(m);
    return rc;
}

/* The injection list needs the sources and (expl=y) the padded velocity:
 * built by whichever of the two setters comes last, so a run allocates
 * nothing. */
int awefd__spher_set_sources(awefd_model *m)
{
    spher_state *s = st(m);
    s->inj_dirty = 1;
    return s->h_velp ? build_injection(m) : AWEFD_OK;
}

/* Receiver map in (r, theta, phi) order for the extraction kernel. */
int awefd__spher_set_receivers(awefd_model *m)
{
    spher_state *s = st(m);
    int *h = malloc(sizeof(int) * 3 * (m->nr > 0 ? m->nr : 1));
    if (!h) return awefd__fail(AWEFD_ENOMEM, "receiver map");
    for (int r = 0; r < m->nr; r++) {
        h[3 * r] = m->h_rj[3 * r + 1];
        h[3 * r + 1] = m->h_rj[3 * r];
        h[3 * r + 2] = m->ndim == 3 ? m->h_rj[3 * r + 2] : 0;
    }
    int rc = up(&s->d_rrtp, h, sizeof(int) * 3 * (m->nr > 0 ? m->nr : 1));
    free(h);
    return rc;
}

/* Reflectivity |v0 - vk| / (v0 + vk) of the first neighbour k whose velocity
 * differs from the source node's, in the reference's search order
 * (spher_utils.c:155-170 in 2-D, 250-304 in 3-D); 0 if none differs. */
static float reflectivity(const awefd_model *m, const float *vp, int jr, int jt, int jp)
{
    static const int o2[7][2] = {{0, 1}, {1, 0}, {1, 1}, {-1, 0}, {-1, -1}, {-1, 1}, {1, -1}};
    /* (dr, dtheta, dphi) */
    static const int o3[27][3] = {
        {0, 1, 0},  {0, 0, 1},  {1, 0, 0},   {0, -1, 0}, {0, 0, -1}, {-1, 0, 0},
        {1, 1, 0},  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0},
        {0, 1, 1},  {0, -1, -1}, {0, -1, 1}, {0, 1, -1},
        {1, 0, 1},  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1},
        {-1, 0, 1},
        {1, 1, 1},  {-1, 1, 1},  {1, -1, 1}, {-1, -1, 1},
        {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1},
    };
    const long Z = m->np[0], Y = (long)m->np[0] * m->np[1];
    const long a0 = jp * Y + jt * Z + jr;
    const float v0 = vp[a0];
    const int n = m->ndim == 3 ? 27 : 7;
    for (int k = 0; k < n; k++) {
        const int dr = m->ndim == 3 ? o3[k][0] : o2[k][0];
        const int dt = m->ndim == 3 ? o3[k][1] : o2[k][1];
        const int dp = m->ndim == 3 ? o3[k][2] : 0;
        const int r = jr + dr, t = jt + dt, p = jp + dp;
        if (r < 0 || r >= m->np[0] || t < 0 || t >= m->np[1] || p < 0 || p >= m->np[2]) continue;
        const float vk = vp[p * Y + t * Z + r];
        if (v0 != vk) return (float)fabs((v0 - vk) / (v0 + vk));
    }
    return 0.f;
}

float awefd__spher_src_factor(const awefd_model *m, int src)
{
    const spher_state *s = (const spher_state *)m->spher;
    if (!m->p.expl) return 1.f;
    if (!s || !s->h_velp) return -1.f;
    const int jt = m->h_sj[3 * src], jr = m->h_sj[3 * src + 1], jp = m->ndim == 3 ? m->h_sj[3 * src + 2] : 0;
    return reflectivity(m, s->h_velp, jr, jt, jp);
}

typedef struct {
    long node;
    int order;
    float w, r;
} inj_entry;

static int inj_cmp(const void *pa, const void *pb)
{
    const inj_entry *a = pa, *b = pb;
    if (a->node != b->node) return a->node < b->node ? -1 : 1;
    return a->order < b->order ? -1 : (a->order > b->order);
}

/* Injection list: one group per node, contributions in (source, corner)
 * order, as the reference's serial loop adds them. */
static int build_injection(awefd_model *m)
{
    spher_state *s = st(m);
    const int nc = m->ndim == 3 ? 8 : 4;
    const long n = (long)m->ns * nc;
    inj_entry *e = malloc(sizeof(*e) * (n > 0 ? n : 1));
    if (!e) return awefd__fail(AWEFD_ENOMEM, "injection list");
    const long sy = (long)m->ld * m->np[1];
    for (int src = 0; src < m->ns; src++) {
        /* set_points stores (z = theta, x = r, y = phi) */
        const int jt = m->h_sj[3 * src], jr = m->h_sj[3 * src + 1], jp = m->ndim == 3 ? m->h_sj[3 * src + 2] : 0;
        const float r = m->p.expl ? reflectivity(m, s->h_velp, jr, jt, jp) : 1.f;
        for (int q = 0; q < nc; q++) {
            inj_entry *x = &e[(long)src * nc + q];
            x->node = (long)(jp + ((q >> 2) & 1)) * sy + (long)(jt + (q & 1)) * m->ld + jr + ((q >> 1) & 1);
            x->order = src * nc + q;
            x->w = m->h_sw[8 * src + q];
            x->r = r;
        }
    }
    qsort(e, (size_t)n, sizeof *e, inj_cmp);
    long *node = malloc(sizeof(long) * (n > 0 ? n : 1));
    int *ibeg = malloc(sizeof(int) * (n + 1));
    float *iw = malloc(sizeof(float) * 2 * (n > 0 ? n : 1));
    int rc = (!node || !ibeg || !iw) ? awefd__fail(AWEFD_ENOMEM, "injection list") : AWEFD_OK;
    int nn = 0;
    for (long i = 0; !rc && i < n; i++) {
        if (i == 0 || e[i].node != e[i - 1].node) {
            node[nn] = e[i].node;
            ibeg[nn] = (int)i;
            nn++;
        }
        iw[2 * i] = e[i].w;
        iw[2 * i + 1] = e[i].r;
    }
    if (!rc) {
        ibeg[nn] = (int)n;
        s->ninode = nn;
        rc = up(&s->d_inode, node, sizeof(long) * (nn > 0 ? nn : 1));
        if (!rc) rc = up(&s->d_ibeg, ibeg, sizeof(int) * (nn + 1));
        if (!rc) rc = up(&s->d_iw, iw, sizeof(float) * 2 * (n > 0 ? n : 1));
    }
    free(e); free(node); free(ibeg); free(iw);
    if (!rc) s->inj_dirty = 0;
    return rc;
}

static void spher_args(const awefd_model *m, awefd_spher_args *a)
{
    const spher_state *s = (const spher_state *)m->spher;
    memset(a, 0, sizeof *a);
    a->S = s->dS;
    a->vv = m->dvel;
    a->n1 = m->np[0];
    a->n2 = m->np[1];
    a->n3 = m->np[2];
    a->ld = m->ld;
    a->ndim = m->ndim;
    a->tdr = s->tdr;
    a->dr2 = s->dr2;
    a->tdt = s->tdt;
    a->dt2 = s->dt2;
    a->dp2 = s->dp2;
    a->rtdr = s->rtdr;
    a->rdr2 = s->rdr2;
    a->rtdt = s->rtdt;
    a->rdt2 = s->rdt2;
    a->rdp2 = s->rdp2;
    a->fastdiv = s->fastdiv;
    a->mr2 = s->d_mr2;
    a->mr1 = s->d_mr1;
    a->mt1 = s->d_mt1;
    a->mt2 = s->d_mt2;
    a->pc = s->d_pc;
    a->nb = m->nb;
    a->fsrf = m->p.fsrf;
    a->damp = m->d_damp;
    a->bthl = s->d_ow[0];
    a->bthh = s->d_ow[1];
    a->brl = s->d_ow[2];
    a->brh = s->d_ow[3];
    a->bphl = s->d_ow[4];
    a->bphh = s->d_ow[5];
    a->ntz = m->ntz;
    a->ntx = m->ntx;
    a->nty = m->nty;
    a->nblk = m->nblk;
    a->col = m->col;
    a->nsr = s->nsr;
    a->nst = s->nst;
    a->nsp = s->nsp;
    {
        const int cr = face_count(a->n1), ct = face_count(a->n2);
        a->dv_rc = awefd_magic_div((unsigned)cr);
        a->dv_n2 = awefd_magic_div((unsigned)a->n2);
        a->dv_rin = awefd_magic_div((unsigned)(a->n1 - cr > 0 ? a->n1 - cr : 1));
        a->dv_tc = awefd_magic_div((unsigned)ct);
        a->dv_tin = awefd_magic_div((unsigned)(a->n2 - ct > 0 ? a->n2 - ct : 1));
    }
    a->ninode = s->ninode;
    a->inode = s->d_inode;
    a->ibeg = s->d_ibeg;
    a->iw = s->d_iw;
    a->wav = m->d_wav;
    a->nr = m->nr;
    a->rj = m->d_rj;
    a->rw = m->d_rw;
}

/* The time loop (MAWEFDspher2Dcpu.c:264-317, MAWEFDspher3Dcpu.c:302-359).
 * m->dC holds the current field, m->dP the previous one. */
int awefd__spher_loop(awefd_model *m, float *dtr, int jsnap, awefd_snap_fn cb, void *user)
{
    spher_state *s = st(m);
    int rc;
    if (s->inj_dirty && (rc = build_injection(m))) return rc;
    if (!s->d_rrtp) return awefd__fail(AWEFD_ESTATE, "receivers not set");
    awefd_spher_args a;
    spher_args(m, &a);
    a.rj = s->d_rrtp;
    float *host = NULL;
    int npairs = 0;
    const int nt = m->p.nt;
    if (m->prof && m->nev < 2 * nt) {
        hipEvent_t *ev = realloc(m->ev, sizeof(hipEvent_t) * 2 * nt);
        if (!ev) return awefd__fail(AWEFD_ENOMEM, "event array");
        m->ev = ev;
        for (int i = m->nev; i < 2 * nt; i++) hipEventCreate(&m->ev[i]);
        m->nev = 2 * nt;
    }
    for (int it = 0; it < nt && !rc; it++) {
        a.P = m->dC;
        a.Q = m->dP;
        a.it = it;
        a.trow = dtr ? dtr + (long)it * m->nr : m->d_traces + (long)it * m->nr;
        int e = awefd_launch_spher_inject(&a, m->stream);
        if (m->prof) hipEventRecord(m->ev[2 * npairs], m->stream);
        if (!e) e = awefd_launch_spher_step(&a, m->stream);
        if (m->prof) hipEventRecord(m->ev[2 * npairs + 1], m->stream);
        if (m->prof) npairs++;
        if (!e) e = awefd_launch_spher_shell(&a, m->stream);
        if (!e) e = awefd_launch_spher_extract(&a, m->stream);
        if (e) {
            rc = awefd__fail(AWEFD_EHIP, "step %d launch: %s", it, hipGetErrorString((hipError_t)e));
            break;
        }
        /* shift: the new field becomes current, the current one previous */
        float *t = m->dC;
        m->dC = m->dP;
        m->dP = t;
        m->steps_done = it + 1;
        if (cb && jsnap > 0 && it % jsnap == 0) {
            if (!host && !(host = malloc(sizeof(float) * m->ncell))) {
                rc = awefd__fail(AWEFD_ENOMEM, "snapshot buffer");
                break;
            }
            hipError_t he = hipStreamSynchronize(m->stream);
            if (he == hipSuccess)
                he = hipMemcpy2D(host, sizeof(float) * m->np[0], m->dC, sizeof(float) * m->ld,
                                 sizeof(float) * m->np[0], (size_t)m->np[1] * m->np[2], hipMemcpyDeviceToHost);
            if (he != hipSuccess) {
                rc = awefd__fail(AWEFD_EHIP, "snapshot copy: %s", hipGetErrorString(he));
                break;
            }
            cb(user, it, host, m->np);
        }
    }
    free(host);
    if (!rc && m->prof && npairs) {
        hipError_t he = hipEventSynchronize(m->ev[2 * npairs - 1]);
        for (int i = 0; he == hipSuccess && i < npairs; i++) {
            float ms = 0.f;
            he = hipEventElapsedTime(&ms, m->ev[2 * i], m->ev[2 * i + 1]);
            m->prof_ms += ms;
        }
        m->prof_launches += npairs;
        if (he != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "profiling events: %s", hipGetErrorString(he));
    }
    hipStreamSynchronize(m->stream);
    return rc;
}

int awefd__spher_get_field(awefd_model *m, float *host)
{
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy2D(host, sizeof(float) * m->np[0], m->dC, sizeof(float) * m->ld, sizeof(float) * m->np[0],
                       (size_t)m->np[1] * m->np[2], hipMemcpyDeviceToHost));
    return AWEFD_OK;
}
