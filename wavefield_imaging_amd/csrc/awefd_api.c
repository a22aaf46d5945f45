/*
 * libawefd C ABI (include/awefd.h): lifecycle, device memory, the time loop.
 *
 * Host C over the HIP runtime API; the kernels live in cart_kernels.hip and
 * spher_kernels.hip and are reached through the launchers of
 * awefd_internal.h.  Nothing here does field arithmetic on the CPU except
 * the set-up tables (awefd_setup.c) and the pending sponge factor applied to
 * snapshot copies.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "awefd.h"
#include "awefd_internal.h"
#include "awefd_model.h"
#include "awefd_setup.h"

static __thread char g_err[512];

int awefd__fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *awefd_last_error(void) { return g_err; }

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return awefd__fail(AWEFD_EHIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                               __FILE__, __LINE__);                                         \
    } while (0)

int awefd_params_default(int kind, awefd_params *p)
{
    if (!p || kind < AWEFD_CART2D || kind > AWEFD_SPHER3D)
        return awefd__fail(AWEFD_EINVAL, "awefd_params_default: bad kind %d", kind);
    memset(p, 0, sizeof *p);
    p->kind = kind;
    p->n[0] = p->n[1] = p->n[2] = 1;
    p->d[0] = p->d[1] = p->d[2] = 1.f;
    p->nb = -1;
    p->fdorder = 4;
    p->sse = 1;
    p->jdata = 1;
    if (kind == AWEFD_CART2D || kind == AWEFD_CART3D) {
        /* MAWEFDcart2Dcpu.c:116-128: every flag defaults to n */
        p->expl = 0;
        p->cden = 0;
    } else {
        /* MAWEFDspher2Dcpu.c:56-60, MAWEFDspher3Dcpu.c:57-61: expl defaults to y */
        p->expl = 1;
    }
    return AWEFD_OK;
}

static int is_cart(const awefd_model *m) { return m->p.kind <= AWEFD_CART3D; }

static void free_dev(void **p)
{
    if (*p) hipFree(*p);
    *p = NULL;
}

void awefd_destroy(awefd_model *m)
{
    if (!m) return;
    hipSetDevice(m->dev);
    void **ptrs[] = {(void **)&m->dP,     (void **)&m->dC,      (void **)&m->dF,
                     (void **)&m->dvel,   (void **)&m->drho,    (void **)&m->dsp[0],
                     (void **)&m->dsp[1], (void **)&m->dsp[2],  (void **)&m->dtab[0],
                     (void **)&m->dtab[1], (void **)&m->dtab[2], (void **)&m->dtab[3],
                     (void **)&m->dtab[4], (void **)&m->dtab[5], (void **)&m->d_rj,
                     (void **)&m->d_rw,   (void **)&m->d_inj,   (void **)&m->d_ibeg,
                     (void **)&m->d_wav,  (void **)&m->d_traces, (void **)&m->d_sj,
                     (void **)&m->d_sw,   (void **)&m->d_srefl, (void **)&m->d_damp,
                     (void **)&m->d_metric, (void **)&m->d_abcrules, (void **)&m->d_damp_h,
                     (void **)&m->d_wavr, (void **)&m->d_rsinc};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free_dev(ptrs[i]);
    if (m->ev) {
        for (int i = 0; i < m->nev; i++) hipEventDestroy(m->ev[i]);
        free(m->ev);
    }
    if (m->stream) hipStreamDestroy(m->stream);
    free(m->h_sj); free(m->h_sw); free(m->h_rj); free(m->h_rw);
    free(m->h_sp[0]); free(m->h_sp[1]); free(m->h_sp[2]);
    free(m->h_srefl);
    free(m->h_wav);
    free(m->h_ssinc);
    free(m->h_rsinc);
    awefd__spher_free(m);
    awefd__adjoint_free(m);
    if (m->cgexec) hipGraphExecDestroy(m->cgexec);
    if (m->cgraph) hipGraphDestroy(m->cgraph);
    free(m);
}

/* Tiling of the cartesian step kernel: z tiles of balanced width
 * <= 64 - 2*nop (a wave also holds nop halo lanes on each side), x tiles of
 * 16 rows (2-D grids too small for 3 such blocks per CU: 3 rows), y chunks
 * sized so that the launch has >= ~2048 blocks (8 per CU on 256 CUs). */
static void cart_tiling(awefd_model *m)
{
    const int n1 = m->np[0], n2 = m->np[1], n3 = m->np[2];
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m->dev);
    if (m->target_blocks <= 0) {
        /* about 16K blocks (12-13 rounds of resident blocks on 256 CUs):
         * short y chunks keep neighbouring tiles in step, so halo rows are
         * still in L2 when they are re-read, and the last round is small.
         * AWEFD_ROUNDS=r asks for r whole rounds instead. */
        const char *e = getenv("AWEFD_TARGET_BLOCKS");
        const char *r = getenv("AWEFD_ROUNDS");
        int per = awefd_cart_step_occupancy(m->nop, m->ndim, m->p.sse);
        if (per <= 0) per = 4;
        m->target_blocks = e ? atoi(e) : r ? atoi(r) * per * cus : 16384;
    }
    /* z tiles of 64 nodes on 256-byte boundaries */
    m->ntz = (n1 + 63) / 64;
    m->tz = 64;
    /* the z-high ABC recurrence needs its nop+2 nodes in the last tile */
    m->zsep = m->sponge && m->ndim == 3 && m->ntz > 1 && n1 - 64 * (m->ntz - 1) < m->nop + 2;
    const char *ce = getenv("AWEFD_COL");
    m->col = m->ndim == 3 && !(ce && atoi(ce) == 0);
    /* AWEFD_NARROW=0/1 forces the 2-D tile */
    const char *ne = getenv("AWEFD_NARROW");
    const long wide = (long)m->ntz * ((n2 + AWEFD_CART_TX - 1) / AWEFD_CART_TX);
    m->narrow = m->ndim == 2 && (ne ? atoi(ne) != 0 : wide < 3L * cus);
    m->tx = m->col ? AWEFD_COL_RX : m->narrow ? AWEFD_CART_TX2 : AWEFD_CART_TX;
    m->ntx = (n2 + m->tx - 1) / m->tx;
    if (m->col) {
        /* one wave per (z tile, RX rows) column chunk of about
         * AWEFD_COL_CHUNK planes (default 8) */
        const char *ch = getenv("AWEFD_COL_CHUNK");
        const long chunk = ch && atol(ch) > 0 ? atol(ch) : 8;
        long nty = (n3 + chunk - 1) / chunk;
        if (nty > n3) nty = n3;
        if (nty < 1) nty = 1;
        const long plane = (long)m->ld * n2 * 4;
        while (nty < n3 && ((n3 + nty - 1) / nty + 2L * m->nop) * plane >= (1L << 30)) nty++;
        /* the chunk's LDS tables (5 floats per plane) stay within 20 KB */
        while (nty < n3 && (n3 + nty - 1) / nty > 1024) nty++;
        m->nty = (int)nty;
        m->ychunk = (int)((n3 + nty - 1) / nty);
    } else if (m->ndim == 3) {
        int cols = m->ntz * m->ntx;
        int nty = (m->target_blocks + cols / 2) / cols;
        if (nty > n3) nty = n3;
        if (nty < 1) nty = 1;
        /* a block addresses (ychunk + 2 nop) planes through 32-bit buffer
         * offsets: keep that window below 1 GiB */
        const long plane = (long)m->ld * n2 * 4;
        while (nty < n3 && ((n3 + nty - 1) / nty + 2L * m->nop) * plane >= (1L << 30)) nty++;
        m->nty = nty;
        m->ychunk = (n3 + nty - 1) / nty;
    } else {
        m->ychunk = 1;
        m->nty = 1;
    }
    m->nblk = m->ntz * m->ntx * m->nty;
    /* the column kernel's blocks: 8 XCD shares of ceil(ntx / 8) x tiles.
     * When the last z tile holds at most 16 nodes (n1 % 64 <= 16: 336 =
     * 5 x 64 + 16), a wave would run 48 or more idle lanes there; instead
     * one wave takes that tile for 4 x tiles at once, 16 lanes per x tile
     * (AWEFD_COL_TAIL=0: one x tile per wave), and an XCD's share is laid out
     * in groups of 4 x tiles: their 4 (ntz - 1) full tiles, then the tail. */
    const char *te = getenv("AWEFD_COL_TAIL");
    m->ztail = m->col && n1 % 64 != 0 && n1 % 64 <= 16 && !(te && atoi(te) == 0);
    if (m->col && !m->ztail) m->nblk = 8 * m->ntz * ((m->ntx + 7) / 8) * m->nty;
    if (m->ztail) m->nblk = 8 * (4 * (m->ntz - 1) + 1) * (((m->ntx + 7) / 8 + 3) / 4) * m->nty;
    if (getenv("AWEFD_VERBOSE"))
        fprintf(stderr, "awefd: step kernel%s tiles %d(z) x %d(x) x %d(y, <= %d planes) = %d blocks, "
                "target %d\n",
                m->col ? (m->ztail ? " (columns, last z tile in 16-lane rows)" : " (columns)")
                       : m->narrow ? " (one-wave tiles)" : "",
                m->ntz, m->ntx, m->nty, m->ychunk, m->nblk, m->target_blocks);
}

int awefd_create(const awefd_params *p, awefd_model **out)
{
    if (!p || !out) return awefd__fail(AWEFD_EINVAL, "awefd_create: NULL argument");
    *out = NULL;
    if (p->kind < AWEFD_CART2D || p->kind > AWEFD_SPHER3D)
        return awefd__fail(AWEFD_EINVAL, "awefd_create: bad kind %d", p->kind);
    awefd_model *m = calloc(1, sizeof *m);
    if (!m) return awefd__fail(AWEFD_ENOMEM, "awefd_create: out of host memory");
    m->p = *p;
    m->ndim = (p->kind == AWEFD_CART3D || p->kind == AWEFD_SPHER3D) ? 3 : 2;
    if (m->ndim == 2) {
        m->p.n[2] = 1;
        m->p.d[2] = 1.f;
        m->p.o[2] = 0.f;
    }
    for (int a = 0; a < m->ndim; a++)
        if (p->n[a] < 1 || !(p->d[a] > 0.f)) {
            free(m);
            return awefd__fail(AWEFD_EINVAL, "awefd_create: axis %d has n=%d d=%g", a + 1,
                               p->n[a], (double)p->d[a]);
        }
    if (p->nt < 1) {
        free(m);
        return awefd__fail(AWEFD_EINVAL, "awefd_create: nt=%d", p->nt);
    }
    int rc;
    if (is_cart(m)) {
        const int fdorder = p->fdorder > 0 ? p->fdorder : 4;
        m->nop = fdorder / 2;
        if (fdorder % 2 || m->nop < 1 || m->nop > AWEFD_SETUP_MAX_NOP) {
            free(m);
            return awefd__fail(AWEFD_EINVAL, "fdorder=%d unsupported (even, 2..12)", fdorder);
        }
        /* MAWEFDcart2Dcpu.c:211-212: nb < nop or absent -> nop; the hybrid
         * boundary widens nb <= nop to 2 nop */
        m->nb = p->nb < m->nop ? m->nop : p->nb;
        m->hyb = p->dabc && p->hybrid;
        m->sponge = p->dabc && !p->hybrid;
        if (m->hyb && m->nb <= m->nop) m->nb = 2 * m->nop;
        m->sinc = p->sinc != 0;
        m->adj2 = p->adj && m->ndim == 2;
        m->inj_rec = m->adj2 && !p->expl;
        m->nsteps = p->nt;
        if (m->adj2) {
            /* the loop images at it = nt-kt and stops (MAWEFDcart2Dcpu.c:353-370);
             * outside 1..nt it never does and writes nothing */
            if (p->kt < 1 || p->kt > p->nt) {
                free(m);
                return awefd__fail(AWEFD_EINVAL, "adj=y needs 1 <= kt <= nt (kt=%d, nt=%d): the image is "
                                   "the field at step nt-kt", p->kt, p->nt);
            }
            m->nsteps = p->nt - p->kt + 1;
        }
    } else {
        m->nop = 4; /* NOP of spher_utils.c:326 */
        m->nb = p->nb < 4 ? 4 : p->nb;
        m->nsteps = p->nt;
    }
    m->ncell = 1;
    for (int a = 0; a < 3; a++) {
        const int nba = a < m->ndim ? m->nb : 0;
        m->np[a] = m->p.n[a] + 2 * nba;
        /* fdutil.c:293-297: padded origin o - nb*d in float */
        m->opad[a] = m->p.o[a] - (float)nba * m->p.d[a];
        m->ncell *= m->np[a];
    }
    /* device rows padded to 128 bytes: a 64-node tile is two whole cache lines */
    m->ld = (m->np[0] + 31) / 32 * 32;
    m->nalloc = (long)m->ld * m->np[1] * m->np[2];
    if (m->ncell > (1L << 33)) {
        free(m);
        return awefd__fail(AWEFD_EINVAL, "grid of %ld nodes too large", m->ncell);
    }
    m->dev = p->device;
    hipError_t e = hipSetDevice(m->dev);
    if (e != hipSuccess) {
        free(m);
        return awefd__fail(AWEFD_EHIP, "hipSetDevice(%d): %s", p->device, hipGetErrorString(e));
    }
    e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free(m);
        return awefd__fail(AWEFD_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    if (is_cart(m)) {
        const float d[3] = {m->p.d[0], m->p.d[1], m->p.d[2]};
        awefd_fdcoef(m->nop, m->ndim, d, p->optfd, 2, m->c2);
        awefd_fdcoef(m->nop, m->ndim, d, p->optfd, 1, m->c1);
        cart_tiling(m);
        {
            const long plane = (long)m->ld * m->np[1] * 4;
            const long win = m->ndim == 3 ? (m->ychunk + 2L * m->nop) * plane : plane;
            if (win >= (1L << 30)) {
                awefd_destroy(m);
                return awefd__fail(AWEFD_EINVAL, "a %dx%d plane is too large for the 1 GiB "
                                   "buffer window of the step kernel", m->np[0], m->np[1]);
            }
        }
        if (m->ndim == 3 && m->sponge && (m->np[0] < 2 * m->nop + 4 || m->np[1] < 2 * m->nop + 4
                                        || m->np[2] < 2 * m->nop + 4)) {
            awefd_destroy(m);
            return awefd__fail(AWEFD_EINVAL, "3-D one-way ABC needs >= %d padded nodes per axis",
                               2 * m->nop + 4);
        }
    }
    const size_t bytes = sizeof(float) * (size_t)m->nalloc;
    if ((e = hipMalloc((void **)&m->dP, bytes)) != hipSuccess
        || (e = hipMalloc((void **)&m->dC, bytes)) != hipSuccess
        || (e = hipMalloc((void **)&m->dvel, bytes)) != hipSuccess) {
        awefd_destroy(m);
        return awefd__fail(AWEFD_ENOMEM, "hipMalloc of 3 x %zu bytes: %s", bytes,
                           hipGetErrorString(e));
    }
    if (!is_cart(m) && (rc = awefd__spher_create(m)) != 0) {
        awefd_destroy(m);
        return rc;
    }
    *out = m;
    return AWEFD_OK;
}

int awefd_padded_dims(const awefd_model *m, int dims[3], int *nb)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "NULL model");
    if (dims)
        for (int a = 0; a < 3; a++) dims[a] = m->np[a];
    if (nb) *nb = m->nb;
    return AWEFD_OK;
}

int awefd_device(const awefd_model *m) { return m ? m->dev : -1; }

static int upload(awefd_model *m, void **dst, const void *src, size_t bytes)
{
    free_dev(dst);
    if (bytes == 0) return AWEFD_OK;
    HIPCHK(hipMalloc(dst, bytes));
    HIPCHK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    (void)m;
    return AWEFD_OK;
}

/* Variable density (cden=n with den=, MAWEFDcart2Dcpu.c:137-145,264-268): the
 * padded density and its first differences along z, x, y with the 1st-order
 * coefficients, as step_forward_2d/3d form them (ts_kernel.c:101-135,228-280).
 * They do not change with time, so they are computed once here and stored as
 * one 16-byte record per node {rho, drho_z, drho_x, drho_y}. */
static int cart_set_density(awefd_model *m, const float *den)
{
    const int nbs[3] = {m->nb, m->nb, m->ndim == 3 ? m->nb : 0};
    const int n1 = m->np[0], n2 = m->np[1], n3 = m->np[2], nop = m->nop;
    float *r = malloc(sizeof(float) * m->ncell);
    float *g = calloc((size_t)m->ncell * 4, sizeof(float));
    if (!r || !g) {
        free(r); free(g);
        return awefd__fail(AWEFD_ENOMEM, "density buffers");
    }
    awefd_pad(den, r, m->p.n, nbs);
    const long sx = n1, sy = (long)n1 * n2;
    const float *bz = m->c1, *bx = m->c1 + nop, *by = m->c1 + 2 * nop;
    for (long i = 0; i < m->ncell; i++) g[4 * i] = r[i];
    const int y0 = m->ndim == 3 ? nop : 0, y1 = m->ndim == 3 ? n3 - nop : 1;
    for (int iy = y0; iy < y1; iy++)
        for (int ix = nop; ix < n2 - nop; ix++)
            for (int iz = nop; iz < n1 - nop; iz++) {
                const long i = iy * sy + ix * sx + iz;
                float dz = 0.f, dx = 0.f, dy = 0.f;
                for (int k = 1; k <= nop; k++) {
                    dz = dz + (r[i + k] - r[i - k]) * bz[k];
                    dx = dx + (r[i + k * sx] - r[i - k * sx]) * bx[k];
                    if (m->ndim == 3) dy = dy + (r[i + k * sy] - r[i - k * sy]) * by[k];
                }
                g[4 * i + 1] = dz;
                g[4 * i + 2] = dx;
                g[4 * i + 3] = dy;
            }
    free(r);
    free_dev((void **)&m->drho);
    hipError_t e = hipMalloc((void **)&m->drho, sizeof(float) * 4 * (size_t)m->nalloc);
    if (e == hipSuccess)
        e = hipMemcpy2D(m->drho, sizeof(float) * 4 * m->ld, g, sizeof(float) * 4 * n1, sizeof(float) * 4 * n1,
                        (size_t)n2 * n3, hipMemcpyHostToDevice);
    free(g);
    if (e != hipSuccess) return awefd__fail(AWEFD_EHIP, "density upload: %s", hipGetErrorString(e));
    return AWEFD_OK;
}

/* Cartesian model set-up: MAWEFDcart2Dcpu.c:258-295, MAWEFDcart3Dcpu.c:254-293. */
/* sponge_make (fdutil.c) per axis: the factor of every padded node along
 * each axis, 1 outside the strips */
static int cart_sponge_tables(awefd_model *m)
{
    if (m->nb > 4096) return awefd__fail(AWEFD_EINVAL, "nb=%d too large", m->nb);
    float w[4096];
    awefd_sponge_weights(m->nb, w);
    int rc = AWEFD_OK;
    for (int a = 0; a < 3 && rc == 0; a++) {
        free(m->h_sp[a]);
        m->h_sp[a] = malloc(sizeof(float) * m->np[a]);
        if (!m->h_sp[a]) return awefd__fail(AWEFD_ENOMEM, "sponge factors");
        if (a < m->ndim) awefd_sponge_axis(w, m->nb, m->np[a], m->h_sp[a]);
        else m->h_sp[a][0] = 1.f;
        rc = upload(m, (void **)&m->dsp[a], m->h_sp[a], sizeof(float) * m->np[a]);
    }
    return rc;
}

static int cart_set_model(awefd_model *m, const float *vel, const float *den)
{
    const int nbs[3] = {m->nb, m->nb, m->ndim == 3 ? m->nb : 0};
    float *v = malloc(sizeof(float) * m->ncell);
    if (!v) return awefd__fail(AWEFD_ENOMEM, "host pad buffer");
    awefd_pad(vel, v, m->p.n, nbs);
    const int n1 = m->np[0], n2 = m->np[1], n3 = m->np[2];
    int rc = AWEFD_OK;
    if ((m->sponge && m->ndim == 3) || m->hyb) {
        /* abcone2d_make/abcone3d_make(nb, dt, vel, ...) on the raw velocity,
         * MAWEFDcart2Dcpu.c:276, MAWEFDcart3Dcpu.c:270; the 2-D tables are
         * bz*[n2] and bx*[n1] (the 3-D layout with n3 = 1) */
        const long sx = n1, sy = (long)n1 * n2;
        const int nb = m->nb;
        float *t[6];
        size_t sz[6] = {(size_t)n2 * n3, (size_t)n2 * n3, (size_t)n1 * n3,
                        (size_t)n1 * n3, (size_t)n1 * n2, (size_t)n1 * n2};
        for (int i = 0; i < 6; i++) t[i] = malloc(sizeof(float) * sz[i]);
        if (!t[0] || !t[1] || !t[2] || !t[3] || !t[4] || !t[5]) {
            for (int i = 0; i < 6; i++) free(t[i]);
            free(v);
            return awefd__fail(AWEFD_ENOMEM, "one-way tables");
        }
        for (int iy = 0; iy < n3; iy++)
            for (int ix = 0; ix < n2; ix++) {
                t[0][iy * n2 + ix] = awefd_oneway_cart(v[iy * sy + ix * sx + nb], m->p.dt, m->p.d[0]);
                t[1][iy * n2 + ix] = awefd_oneway_cart(v[iy * sy + ix * sx + n1 - nb - 1], m->p.dt, m->p.d[0]);
            }
        for (int iy = 0; iy < n3; iy++)
            for (int iz = 0; iz < n1; iz++) {
                t[2][iy * n1 + iz] = awefd_oneway_cart(v[iy * sy + nb * sx + iz], m->p.dt, m->p.d[1]);
                t[3][iy * n1 + iz] = awefd_oneway_cart(v[iy * sy + (n2 - nb - 1) * sx + iz], m->p.dt, m->p.d[1]);
            }
        for (int ix = 0; ix < n2 && m->ndim == 3; ix++)
            for (int iz = 0; iz < n1; iz++) {
                t[4][ix * n1 + iz] = awefd_oneway_cart(v[(long)nb * sy + ix * sx + iz], m->p.dt, m->p.d[2]);
                t[5][ix * n1 + iz] = awefd_oneway_cart(v[(long)(n3 - nb - 1) * sy + ix * sx + iz], m->p.dt, m->p.d[2]);
            }
        for (int i = 0; i < (m->ndim == 3 ? 6 : 4) && rc == 0; i++)
            rc = upload(m, (void **)&m->dtab[i], t[i], sizeof(float) * sz[i]);
        for (int i = 0; i < 6; i++) free(t[i]);
    }
    if (m->p.dabc && m->nb > 4096) {
        free(v);
        return awefd__fail(AWEFD_EINVAL, "nb=%d too large", m->nb);
    }
    if (m->hyb && rc == 0) {
        /* damp_make(nb - nop), MAWEFDcart2Dcpu.c:407-421: the sponge_make
         * table over nb - nop nodes */
        float w[4096];
        awefd_sponge_weights(m->nb - m->nop, w);
        rc = upload(m, (void **)&m->d_damp_h, w, sizeof(float) * (m->nb - m->nop));
    }
    if (m->sponge && rc == 0) rc = cart_sponge_tables(m);
    /* vel <- vel * (vel * dt^2) ; free surface zeroes z in [0, nb], except
     * under hybridbc (MAWEFDcart2Dcpu.c:289-295) */
    const float dt2 = m->p.dt * m->p.dt;
    for (long i = 0; i < m->ncell; i++) v[i] *= v[i] * dt2;
    if (m->p.fsrf && !m->p.hybrid)
        for (long c = 0; c < (long)n2 * n3; c++)
            for (int iz = 0; iz <= m->nb && iz < n1; iz++) v[c * n1 + iz] = 0.f;
    if (rc == 0) {
        hipError_t e = hipMemcpy2D(m->dvel, sizeof(float) * m->ld, v, sizeof(float) * n1,
                                   sizeof(float) * n1, (size_t)n2 * n3, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "vel upload: %s", hipGetErrorString(e));
    }
    free(v);
    if (rc == 0 && den && !m->p.cden) rc = cart_set_density(m, den);
    else free_dev((void **)&m->drho);     /* constant density */
    return rc;
}

static int setter_done(awefd_model *m);

int awefd_set_model(awefd_model *m, const float *vel, const float *den)
{
    if (!m || !vel) return awefd__fail(AWEFD_EINVAL, "awefd_set_model: NULL argument");
    HIPCHK(hipSetDevice(m->dev));
    int rc = is_cart(m) ? cart_set_model(m, vel, den) : awefd__spher_set_model(m, vel);
    if (rc == 0) m->model_set = 1;
    const int rd = setter_done(m);
    return rc ? rc : rd;
}

/* The velocity from device memory: the same padded (v dt)^2 and free-surface
 * zeros as awefd_set_model(m, vel, NULL), formed by a kernel, so a model
 * update in an inversion loop does not pass through the host.  Only for the
 * set-ups whose other tables do not depend on the velocity: cartesian, not
 * hybridbc=y, and in 3-D not dabc=y (both build one-way boundary tables from
 * the velocity on the host). */
int awefd_set_model_dev(awefd_model *m, const float *d_vel)
{
    if (!m || !d_vel) return awefd__fail(AWEFD_EINVAL, "awefd_set_model_dev: NULL argument");
    if (!is_cart(m))
        return awefd__fail(AWEFD_EINVAL, "awefd_set_model_dev: cartesian models only (the spherical set-up "
                                         "derives its one-way tables and reflectivities on the host)");
    if ((m->sponge && m->ndim == 3) || m->hyb)
        return awefd__fail(AWEFD_EINVAL, "awefd_set_model_dev: %s builds one-way boundary tables from the "
                                         "velocity on the host; use awefd_set_model",
                           m->hyb ? "hybridbc=y" : "3-D dabc=y");
    HIPCHK(hipSetDevice(m->dev));
    int rc = m->sponge ? cart_sponge_tables(m) : AWEFD_OK;
    if (!rc) {
        const float dt = m->p.dt;
        const int nbs[3] = {m->nb, m->nb, m->ndim == 3 ? m->nb : 0};
        const int e = awefd_launch_set_vel(d_vel, m->dvel, m->p.n, nbs, m->ld, dt * dt,
                                           m->p.fsrf && !m->p.hybrid ? m->nb : -1, m->stream);
        hipError_t he = e ? (hipError_t)e : hipStreamSynchronize(m->stream);
        if (he != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "awefd_set_model_dev: %s", hipGetErrorString(he));
    }
    if (!rc) {
        free_dev((void **)&m->drho);      /* constant density */
        m->model_set = 1;
    }
    const int rd = setter_done(m);
    return rc ? rc : rd;
}

static int set_points(awefd_model *m, int which, int n, const float *coords)
{
    if (!m || n < 0 || (n > 0 && !coords)) return awefd__fail(AWEFD_EINVAL, "set_points: bad argument");
    int *j = malloc(sizeof(int) * 3 * (n > 0 ? n : 1));
    float *w = malloc(sizeof(float) * 8 * (n > 0 ? n : 1));
    if (!j || !w) {
        free(j);
        free(w);
        return awefd__fail(AWEFD_ENOMEM, "set_points: %d points", n);
    }
    /* grid in (z, x, y) order; spherical: z = theta, x = r, y = phi */
    int npad[3];
    float opad[3], d[3];
    if (is_cart(m)) {
        for (int a = 0; a < 3; a++) {
            npad[a] = m->np[a];
            opad[a] = m->opad[a];
            d[a] = m->p.d[a];
        }
    } else {
        const int perm[3] = {1, 0, 2};
        for (int a = 0; a < 3; a++) {
            npad[a] = m->np[perm[a]];
            opad[a] = m->opad[perm[a]];
            d[a] = m->p.d[perm[a]];
        }
    }
    awefd_lint(m->ndim, n, coords, npad, opad, d, j, w);
    awefd_sinc_pt *sp = NULL;
    if (m->sinc) {
        /* sinc=y: the 9-node taps replace the bilinear corners */
        sp = malloc(sizeof(*sp) * (n > 0 ? n : 1));
        const int oom = sp == NULL;
        if (oom || awefd_sinc_points(m->ndim, n, coords, npad, opad, d, sp)) {
            free(j);
            free(w);
            free(sp);
            return oom ? awefd__fail(AWEFD_ENOMEM, "sinc records")
                       : awefd__fail(AWEFD_EINVAL, "a %s's sinc taps (4 nodes each way) leave the padded "
                                     "grid", which == 0 ? "source" : "receiver");
        }
    }
    /* every interpolation corner must be a padded node: the kernels index
     * with these values unchecked */
    for (int a = 0; a < n && !m->sinc; a++)
        for (int ax = 0; ax < m->ndim; ax++)
            if (j[3 * a + ax] < 0 || j[3 * a + ax] + 1 >= npad[ax]) {
                free(j);
                free(w);
                return awefd__fail(AWEFD_EINVAL, "%s %d maps outside the padded grid",
                                   which == 0 ? "source" : "receiver", a);
            }
    if (which == 0) {
        free(m->h_ssinc);
        m->h_ssinc = sp;
    } else {
        free(m->h_rsinc);
        m->h_rsinc = sp;
    }
    if (which == 0) {
        free(m->h_sj); free(m->h_sw);
        m->h_sj = j; m->h_sw = w; m->ns = n;
    } else {
        free(m->h_rj); free(m->h_rw);
        m->h_rj = j; m->h_rw = w; m->nr = n;
    }
    return AWEFD_OK;
}

/* Injection buckets of the cartesian step kernel (one per z-tile/x-tile
 * column of a tile tz x tx), holding every (node, point, weight)
 * contribution of the injected points (the sources, or the receivers in a
 * 2-D adj=y loop without expl); host arrays, freed by the caller. */
int awefd__cart_buckets(const awefd_model *m, int tz, int ntz, int tx, awefd_inj **inj, int **ibeg_out,
                        int *nbucket_out, long *ninj)
{
    const int nc = m->ndim == 3 ? 8 : 4;
    const int npt = m->inj_rec ? m->nr : m->ns;
    const int *pj = m->inj_rec ? m->h_rj : m->h_sj;
    const float *pw = m->inj_rec ? m->h_rw : m->h_sw;
    const awefd_sinc_pt *psp = m->inj_rec ? m->h_rsinc : m->h_ssinc;
    long n = (long)npt * nc;
    if (m->sinc) {
        n = 0;
        for (int s = 0; s < npt; s++) n += (long)psp[s].n[0] * psp[s].n[1] * psp[s].n[2];
    }
    const int ntx = (m->np[1] + tx - 1) / tx;
    const int nbucket = ntz * ntx;
    awefd_inj_key *keys = malloc(sizeof(*keys) * (n > 0 ? n : 1));
    awefd_inj *e = malloc(sizeof(*e) * (n > 0 ? n : 1));
    int *ibeg = malloc(sizeof(int) * (nbucket + 1));
    if (!keys || !e || !ibeg) {
        free(keys); free(e); free(ibeg);
        return awefd__fail(AWEFD_ENOMEM, "injection buckets");
    }
    for (int s = 0; s < npt && !m->sinc; s++)
        for (int q = 0; q < nc; q++) {
            awefd_inj_key *k = &keys[(long)s * nc + q];
            k->e.z = pj[3 * s] + (q & 1);
            k->e.x = pj[3 * s + 1] + ((q >> 1) & 1);
            k->e.y = m->ndim == 3 ? pj[3 * s + 2] + ((q >> 2) & 1) : 0;
            k->e.src = s;
            k->e.w = pw[8 * s + q];
            k->e.w2 = 1.f;
            k->e.w3 = 1.f;
            k->bucket = (k->e.z / tz) + ntz * (k->e.x / tx);
        }
    /* sinc taps: node (c - 4 + t) per axis receives ((amp wy) wx) wz v in
     * 3-D, (amp wx) wz v in 2-D (sinc*d_inject*_with_vv, fdutil.c:888-1012,
     * 1184-1298) */
    long q = 0;
    for (int s = 0; s < npt && m->sinc; s++) {
        const awefd_sinc_pt *p = &psp[s];
        for (int ty = p->f[2]; ty < p->f[2] + p->n[2]; ty++)
            for (int tx_ = p->f[1]; tx_ < p->f[1] + p->n[1]; tx_++)
                for (int tz_ = p->f[0]; tz_ < p->f[0] + p->n[0]; tz_++) {
                    awefd_inj_key *k = &keys[q++];
                    k->e.z = p->c[0] - 4 + tz_;
                    k->e.x = p->c[1] - 4 + tx_;
                    k->e.y = m->ndim == 3 ? p->c[2] - 4 + ty : 0;
                    k->e.src = s;
                    k->e.w = m->ndim == 3 ? p->w[2][ty] : p->w[1][tx_];
                    k->e.w2 = m->ndim == 3 ? p->w[1][tx_] : p->w[0][tz_];
                    k->e.w3 = m->ndim == 3 ? p->w[0][tz_] : 1.f;
                    k->bucket = (k->e.z / tz) + ntz * (k->e.x / tx);
                }
    }
    awefd_sort_inj(keys, n, nbucket, e, ibeg);
    free(keys);
    *inj = e;
    *ibeg_out = ibeg;
    *nbucket_out = nbucket;
    *ninj = n;
    return AWEFD_OK;
}

static int cart_build_inj(awefd_model *m)
{
    awefd_inj *e;
    int *ibeg, nbucket;
    long n;
    int rc = awefd__cart_buckets(m, m->tz, m->ntz, m->tx, &e, &ibeg, &nbucket, &n);
    if (rc) return rc;
    rc = upload(m, (void **)&m->d_inj, e, sizeof(*e) * (n > 0 ? n : 1));
    if (!rc) rc = upload(m, (void **)&m->d_ibeg, ibeg, sizeof(int) * (nbucket + 1));
    m->ninj = n;
    free(e); free(ibeg);
    return rc;
}

int awefd_set_sources(awefd_model *m, int ns, const float *coords)
{
    int rc = set_points(m, 0, ns, coords);
    if (rc) return rc;
    HIPCHK(hipSetDevice(m->dev));
    if (!is_cart(m)) rc = awefd__spher_set_sources(m);
    else if (!m->inj_rec) rc = cart_build_inj(m);
    if (!rc) m->src_set = 1;
    const int rd = setter_done(m);
    return rc ? rc : rd;
}

int awefd_set_receivers(awefd_model *m, int nr, const float *coords)
{
    int rc = set_points(m, 1, nr, coords);
    if (rc) return rc;
    HIPCHK(hipSetDevice(m->dev));
    rc = upload(m, (void **)&m->d_rj, m->h_rj, sizeof(int) * 3 * (nr > 0 ? nr : 1));
    if (!rc) rc = upload(m, (void **)&m->d_rw, m->h_rw, sizeof(float) * 8 * (nr > 0 ? nr : 1));
    if (!rc && m->sinc)
        rc = upload(m, (void **)&m->d_rsinc, m->h_rsinc, sizeof(awefd_sinc_pt) * (nr > 0 ? nr : 1));
    if (!rc && m->inj_rec) rc = cart_build_inj(m);
    if (!rc && !is_cart(m)) rc = awefd__spher_set_receivers(m);
    if (!rc) m->rec_set = 1;
    const int rd = setter_done(m);
    return rc ? rc : rd;
}

int awefd_set_wavelet(awefd_model *m, long nsamples, const float *wav)
{
    if (!m || nsamples < 1 || !wav) return awefd__fail(AWEFD_EINVAL, "awefd_set_wavelet: bad argument");
    HIPCHK(hipSetDevice(m->dev));
    int rc = upload(m, (void **)&m->d_wav, wav, sizeof(float) * nsamples);
    if (!rc && is_cart(m) && m->p.adj) {
        /* adj=y reads the rows back to front; the row length depends on the
         * points, so the reversed copy is made once the set-up is complete */
        free(m->h_wav);
        m->h_wav = malloc(sizeof(float) * nsamples);
        if (!m->h_wav) rc = awefd__fail(AWEFD_ENOMEM, "wavelet copy");
        else memcpy(m->h_wav, wav, sizeof(float) * nsamples);
    }
    if (!rc) {
        m->nwav = nsamples;
        m->wav_set = 1;
    }
    const int rd = setter_done(m);
    return rc ? rc : rd;
}

int awefd_debug_points(const awefd_model *m, int which, int *idx, float *w)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "NULL model");
    if (which < 0 || which > 2) return awefd__fail(AWEFD_EINVAL, "awefd_debug_points: which=%d", which);
    const int n = which != 1 ? m->ns : m->nr;
    const int *j = which != 1 ? m->h_sj : m->h_rj;
    const float *ww = which != 1 ? m->h_sw : m->h_rw;
    if (!j) return awefd__fail(AWEFD_ESTATE, "points not set");
    if (which == 2) {
        /* per-source injection factor: spherical expl=y reflectivity, else 1 */
        for (int a = 0; a < n; a++) {
            float f = 1.f;
            if (!is_cart(m)) {
                f = awefd__spher_src_factor(m, a);
                if (f < 0.f) return awefd__fail(AWEFD_ESTATE, "awefd_debug_points: model not set");
            }
            for (int q = 0; q < 8; q++) w[8 * a + q] = q == 0 ? f : 0.f;
            idx[3 * a] = idx[3 * a + 1] = idx[3 * a + 2] = 0;
        }
        return AWEFD_OK;
    }
    /* report (fastest, middle, slowest) axis indices of the model */
    for (int a = 0; a < n; a++) {
        if (is_cart(m)) {
            idx[3 * a] = j[3 * a];
            idx[3 * a + 1] = j[3 * a + 1];
            idx[3 * a + 2] = j[3 * a + 2];
        } else { /* spherical: (r, theta, phi) from the (theta, r, phi) map */
            idx[3 * a] = j[3 * a + 1];
            idx[3 * a + 1] = j[3 * a];
            idx[3 * a + 2] = j[3 * a + 2];
        }
        memcpy(w + 8 * a, ww + 8 * a, 8 * sizeof(float));
    }
    return AWEFD_OK;
}

int awefd_reset(awefd_model *m)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "NULL model");
    HIPCHK(hipSetDevice(m->dev));
    int e = awefd_launch_fill(m->dP, m->nalloc, 0.f, m->stream);
    if (!e) e = awefd_launch_fill(m->dC, m->nalloc, 0.f, m->stream);
    if (!e && m->dF) e = awefd_launch_fill(m->dF, m->nalloc, 0.f, m->stream);
    if (e) return awefd__fail(AWEFD_EHIP, "reset: %s", hipGetErrorString((hipError_t)e));
    m->steps_done = 0;
    return AWEFD_OK;
}

int awefd_profile(awefd_model *m, int enable)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "NULL model");
    m->prof = enable;
    m->prof_launches = 0;
    m->prof_ms = 0.0;
    return AWEFD_OK;
}

int awefd_kernel_stats(awefd_model *m, long *launches, double *total_ms)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "NULL model");
    if (launches) *launches = m->prof_launches;
    if (total_ms) *total_ms = m->prof_ms;
    return AWEFD_OK;
}

/* Events around the dominant kernel: pairs (start, stop) per launch, read
 * back after the loop. */
static int prof_reserve(awefd_model *m, int n)
{
    if (m->nev >= n) return AWEFD_OK;
    hipEvent_t *ev = realloc(m->ev, sizeof(hipEvent_t) * n);
    if (!ev) return awefd__fail(AWEFD_ENOMEM, "event array");
    m->ev = ev;
    for (int i = m->nev; i < n; i++) HIPCHK(hipEventCreate(&m->ev[i]));
    m->nev = n;
    return AWEFD_OK;
}

static int prof_collect(awefd_model *m, int npairs)
{
    HIPCHK(hipEventSynchronize(m->ev[2 * npairs - 1]));
    for (int i = 0; i < npairs; i++) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, m->ev[2 * i], m->ev[2 * i + 1]));
        m->prof_ms += ms;
    }
    m->prof_launches += npairs;
    return AWEFD_OK;
}

/* Device field (row stride ld) to a dense host array (row stride np[0]). */
static hipError_t copy_field_out(const awefd_model *m, float *host, const float *dev)
{
    return hipMemcpy2D(host, sizeof(float) * m->np[0], dev, sizeof(float) * m->ld,
                       sizeof(float) * m->np[0], (size_t)m->np[1] * m->np[2], hipMemcpyDeviceToHost);
}

/* Apply the pending sponge factor to a host copy of the stored field (the
 * reference's u0 view), in the reference's z, x, y order. */
static void cart_apply_pending(const awefd_model *m, float *u)
{
    if (!m->sponge) return;
    const int n1 = m->np[0], n2 = m->np[1], n3 = m->np[2], nb = m->nb;
    for (int iy = 0; iy < n3; iy++)
        for (int ix = 0; ix < n2; ix++)
            for (int iz = 0; iz < n1; iz++) {
                const int s = iz < nb || iz >= n1 - nb || ix < nb || ix >= n2 - nb
                              || (m->ndim == 3 && (iy < nb || iy >= n3 - nb));
                if (!s) continue;
                float *p = &u[((long)iy * n2 + ix) * n1 + iz];
                float v = *p * m->h_sp[0][iz];
                v = v * m->h_sp[1][ix];
                if (m->ndim == 3) v = v * m->h_sp[2][iy];
                *p = v;
            }
}

int awefd_get_field(awefd_model *m, float *host)
{
    if (!m || !host) return awefd__fail(AWEFD_EINVAL, "awefd_get_field: NULL argument");
    HIPCHK(hipSetDevice(m->dev));
    if (!is_cart(m)) return awefd__spher_get_field(m, host);
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(copy_field_out(m, host, m->dP));
    cart_apply_pending(m, host);
    return AWEFD_OK;
}

void awefd__cart_args(const awefd_model *m, awefd_cart_args *a)
{
    memset(a, 0, sizeof *a);
    a->vel = m->dvel;
    a->rho = m->p.cden ? NULL : m->drho;
    a->n1 = m->np[0];
    a->n2 = m->np[1];
    a->n3 = m->np[2];
    a->ld = m->ld;
    a->zsep = m->zsep;
    const int nop = m->nop;
    a->c0 = m->c2[0];
    for (int k = 1; k <= nop; k++) {
        a->cz[k - 1] = m->c2[k];
        a->cx[k - 1] = m->c2[nop + k];
        a->cy[k - 1] = m->ndim == 3 ? m->c2[2 * nop + k] : 0.f;
        a->bz[k - 1] = m->c1[k];
        a->bx[k - 1] = m->c1[nop + k];
        a->by[k - 1] = m->ndim == 3 ? m->c1[2 * nop + k] : 0.f;
    }
    a->tz = m->tz;
    a->ntz = m->ntz;
    a->ntx = m->ntx;
    a->ychunk = m->ychunk;
    a->nty = m->nty;
    a->col = m->col;
    a->narrow = m->narrow;
    a->ztail = m->ztail;
    a->nblk = m->nblk;
    a->dabc = m->sponge;
    a->nb = m->nb;
    a->sp1 = m->dsp[0];
    a->sp2 = m->dsp[1];
    a->sp3 = m->dsp[2];
    a->abc = m->sponge && m->ndim == 3;
    a->fsrf = m->p.fsrf;
    a->bzl = m->dtab[0];
    a->bzh = m->dtab[1];
    a->bxl = m->dtab[2];
    a->bxh = m->dtab[3];
    a->byl = m->dtab[4];
    a->byh = m->dtab[5];
    a->damp = m->d_damp_h;
    a->ibeg = m->d_ibeg;
    a->inj = m->d_inj;
    a->wav = m->p.adj ? m->d_wavr : m->d_wav;
    a->wexpl = m->p.expl;
    a->wstride = m->p.expl ? 1 : m->inj_rec ? m->nr : m->ns;
    a->nr = m->adj2 ? 0 : m->nr;     /* the 2-D adj=y loop extracts no traces */
    a->rj = m->d_rj;
    a->rw = m->d_rw;
    a->rsinc = m->sinc ? m->d_rsinc : NULL;
}

/* samples of the wavelet the loop reads */
static long wavelet_need(const awefd_model *m, long *row_out)
{
    const long row = m->p.expl ? 1 : m->inj_rec ? m->nr : m->ns;
    if (row_out) *row_out = row;
    return is_cart(m) ? (long)m->p.nt * row : m->p.nt;
}

/* adj=y: step it reads row nt-1-it (MAWEFDcart2Dcpu.c:305-316,
 * MAWEFDcart3Dcpu.c:304-315); the kernels read row it of this copy.  The row
 * length depends on the points, so the copy is made by the last setter that
 * completes the set-up (and again by a run if it is stale). */
static int build_wrev(awefd_model *m)
{
    long row;
    const long need = wavelet_need(m, &row);
    if (m->nwav < need) return AWEFD_OK;   /* the run reports the short wavelet */
    float *r = malloc(sizeof(float) * (size_t)(need > 0 ? need : 1));
    if (!r) return awefd__fail(AWEFD_ENOMEM, "reversed wavelet");
    for (int it = 0; it < m->p.nt; it++)
        memcpy(r + (long)it * row, m->h_wav + (long)(m->p.nt - 1 - it) * row, sizeof(float) * row);
    free_dev((void **)&m->d_wavr);
    int rc = upload(m, (void **)&m->d_wavr, r, sizeof(float) * (size_t)(need > 0 ? need : 1));
    free(r);
    if (rc) return rc;
    m->wrev_gen = m->gen;
    m->wrev_ok = 1;
    return AWEFD_OK;
}

/* end of every setter: with the set-up complete, build what a run would
 * otherwise build, so that runs allocate nothing */
static int setter_done(awefd_model *m)
{
    m->gen++;
    if (is_cart(m) && m->p.adj && m->model_set && m->src_set && m->rec_set && m->wav_set) return build_wrev(m);
    return AWEFD_OK;
}

/* may_alloc = 0: report a stale derived table instead of rebuilding it */
static int check_ready(awefd_model *m, int may_alloc)
{
    if (!m->model_set) return awefd__fail(AWEFD_ESTATE, "velocity not set");
    if (!m->src_set) return awefd__fail(AWEFD_ESTATE, "sources not set");
    if (!m->rec_set) return awefd__fail(AWEFD_ESTATE, "receivers not set");
    if (!m->wav_set) return awefd__fail(AWEFD_ESTATE, "wavelet not set");
    const long need = wavelet_need(m, NULL);
    if (m->nwav < need)
        return awefd__fail(AWEFD_EINVAL, "wavelet has %ld samples, the loop reads %ld", m->nwav, need);
    if (is_cart(m) && m->p.adj && (!m->wrev_ok || m->wrev_gen != m->gen)) {
        if (!may_alloc) return awefd__fail(AWEFD_ESTATE, "adj=y: the reversed wavelet is not built");
        return build_wrev(m);
    }
    return AWEFD_OK;
}

int awefd__ensure_traces(awefd_model *m)
{
    const long n = (long)m->p.nt * m->nr;
    if (m->ntraces >= n && m->d_traces) return AWEFD_OK;
    free_dev((void **)&m->d_traces);
    HIPCHK(hipMalloc((void **)&m->d_traces, sizeof(float) * (n > 0 ? n : 1)));
    m->ntraces = n;
    return AWEFD_OK;
}

/* The cartesian time loop on the device. */
static int cart_loop(awefd_model *m, float *dtr, int jsnap, awefd_snap_fn cb, void *user)
{
    awefd_cart_args a;
    awefd__cart_args(m, &a);
    const int nt = m->nsteps;
    float *P = m->dP, *C = m->dC;
    float *host = NULL;
    int npairs = 0, rc = AWEFD_OK;
    if (m->prof && (rc = prof_reserve(m, 2 * nt))) return rc;
    for (int it = 0; it < nt; it++) {
        a.P = P;
        a.C = C;
        a.it = it;
        a.trow = dtr ? dtr + (long)it * m->nr : m->d_traces + (long)it * m->nr;
        if (m->prof) HIPCHK(hipEventRecord(m->ev[2 * npairs], m->stream));
        int e = awefd_launch_cart_step(&a, m->nop, m->ndim, m->p.sse, a.rho != NULL, m->stream);
        if (m->prof) HIPCHK(hipEventRecord(m->ev[2 * npairs + 1], m->stream));
        if (m->prof) npairs++;
        if (!e && a.abc) e = awefd_launch_cart_abc_xy(&a, m->nop, m->stream);
        if (!e && m->hyb) e = awefd_launch_cart_hybrid(&a, m->nop, m->ndim, m->stream);
        if (e) return awefd__fail(AWEFD_EHIP, "step %d launch: %s", it, hipGetErrorString((hipError_t)e));
        float *t = P;
        P = C;
        C = t;
        m->dP = P;
        m->dC = C;
        m->steps_done = it + 1;
        if (cb && jsnap > 0 && it % jsnap == 0) {
            if (!host && !(host = malloc(sizeof(float) * m->ncell))) return awefd__fail(AWEFD_ENOMEM, "snapshot buffer");
            HIPCHK(hipStreamSynchronize(m->stream));
            HIPCHK(copy_field_out(m, host, P));
            cart_apply_pending(m, host);
            cb(user, it, host, m->np);
        }
    }
    free(host);
    if (m->prof && npairs) rc = prof_collect(m, npairs);
    return rc;
}

/* The 2-D loop as one graph: a 2-D step is a few microseconds of work, so
 * per-launch overhead would dominate a stream of nt launches.  Captured when
 * a run repeats the previous one (same model generation and trace buffer;
 * a one-off run is cheaper launched directly), then launched per run; the
 * fields are zero at every run's start, so the captured buffer roles are
 * reused and the model's roles set to the captured final ones afterwards. */
static int cart_loop_graph(awefd_model *m, float *dtr)
{
    float *trow0 = dtr ? dtr : m->d_traces;
    if (!m->cvalid || m->cgen != m->gen || m->cdtr != trow0) {
        if (m->cgexec) hipGraphExecDestroy(m->cgexec);
        if (m->cgraph) hipGraphDestroy(m->cgraph);
        m->cgexec = NULL;
        m->cgraph = NULL;
        m->cvalid = 0;
        awefd_cart_args a;
        awefd__cart_args(m, &a);
        float *P = m->dP, *C = m->dC;
        HIPCHK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        int e = 0;
        for (int it = 0; it < m->nsteps && !e; it++) {
            a.P = P;
            a.C = C;
            a.it = it;
            a.trow = trow0 + (long)it * m->nr;
            e = awefd_launch_cart_step(&a, m->nop, m->ndim, m->p.sse, a.rho != NULL, m->stream);
            if (!e && m->hyb) e = awefd_launch_cart_hybrid(&a, m->nop, m->ndim, m->stream);
            float *t = P;
            P = C;
            C = t;
        }
        hipGraph_t g = NULL;
        hipError_t ce = hipStreamEndCapture(m->stream, &g);
        if (e || ce != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return awefd__fail(AWEFD_EHIP, "2-D loop capture: %s",
                               hipGetErrorString(e ? (hipError_t)e : ce));
        }
        m->cgraph = g;
        HIPCHK(hipGraphInstantiate(&m->cgexec, g, NULL, NULL, 0));
        m->cPf = P;
        m->cCf = C;
        m->cgen = m->gen;
        m->cdtr = trow0;
        m->cvalid = 1;
    }
    HIPCHK(hipGraphLaunch(m->cgexec, m->stream));
    m->dP = m->cPf;
    m->dC = m->cCf;
    m->steps_done = m->nsteps;
    return AWEFD_OK;
}

static int run_common(awefd_model *m, float *dtr, int jsnap, awefd_snap_fn cb, void *user)
{
    int rc = check_ready(m, 1);
    if (rc) return rc;
    HIPCHK(hipSetDevice(m->dev));
    if ((rc = awefd__ensure_traces(m))) return rc;
    if ((rc = awefd_reset(m))) return rc;
    const int plain2d = is_cart(m) && m->ndim == 2 && !m->prof && !(cb && jsnap > 0);
    float *trow0 = dtr ? dtr : m->d_traces;
    const int repeat = plain2d && ((m->cvalid && m->cgen == m->gen && m->cdtr == trow0) ||
                                   (m->eager_seen && m->eager_gen == m->gen && m->eager_dtr == trow0));
    if (repeat) {
        rc = cart_loop_graph(m, dtr);
    } else if (is_cart(m)) {
        rc = cart_loop(m, dtr, jsnap, cb, user);
        m->eager_seen = plain2d;
        m->eager_gen = m->gen;
        m->eager_dtr = trow0;
    } else {
        rc = awefd__spher_loop(m, dtr, jsnap, cb, user);
    }
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return AWEFD_OK;
}

int awefd_run(awefd_model *m, float *traces, int jsnap, awefd_snap_fn cb, void *user)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "NULL model");
    int rc = run_common(m, NULL, jsnap, cb, user);
    if (rc) return rc;
    if (traces && m->nr > 0 && !m->adj2)
        HIPCHK(hipMemcpy(traces, m->d_traces, sizeof(float) * (size_t)m->p.nt * m->nr,
                         hipMemcpyDeviceToHost));
    return AWEFD_OK;
}

int awefd_run_noalloc(awefd_model *m, float *traces)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "NULL model");
    if (m->prof) return awefd__fail(AWEFD_EINVAL, "awefd_run_noalloc: profiling is on (it allocates events)");
    int rc = check_ready(m, 0);
    if (rc) return rc;
    const size_t nrow = m->adj2 ? 0 : (size_t)m->nr;
    const size_t bytes = sizeof(float) * (size_t)m->p.nt * nrow;
    if (bytes > AWEFD_NOALLOC_TRACE_BYTES)
        return awefd__fail(AWEFD_EINVAL, "awefd_run_noalloc: %d steps x %zu receivers need %zu bytes of trace "
                           "staging; the scratch holds %u", m->p.nt, nrow, bytes, AWEFD_NOALLOC_TRACE_BYTES);
    HIPCHK(hipSetDevice(m->dev));
    float *scr = NULL;
    const int e = awefd__scratch(m->dev, &scr);
    if (e) return awefd__fail(AWEFD_EHIP, "device scratch: %s", hipGetErrorString((hipError_t)e));
    if ((rc = awefd_reset(m))) return rc;
    rc = is_cart(m) ? cart_loop(m, scr, 0, NULL, NULL) : awefd__spher_loop(m, scr, 0, NULL, NULL);
    if (rc) return rc;
    if (traces && bytes) HIPCHK(hipMemcpyAsync(traces, scr, bytes, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return AWEFD_OK;
}

int awefd_traces_dev(awefd_model *m, float **d_traces)
{
    if (!m || !d_traces) return awefd__fail(AWEFD_EINVAL, "awefd_traces_dev: NULL argument");
    HIPCHK(hipSetDevice(m->dev));
    int rc = awefd__ensure_traces(m);
    if (rc) return rc;
    *d_traces = m->d_traces;
    return AWEFD_OK;
}

int awefd_get_traces(awefd_model *m, float *host)
{
    if (!m || !host) return awefd__fail(AWEFD_EINVAL, "awefd_get_traces: NULL argument");
    if (!m->d_traces) return awefd__fail(AWEFD_ESTATE, "no traces yet");
    HIPCHK(hipSetDevice(m->dev));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (m->nr > 0)
        HIPCHK(hipMemcpy(host, m->d_traces, sizeof(float) * (size_t)m->p.nt * m->nr, hipMemcpyDeviceToHost));
    return AWEFD_OK;
}

int awefd_run_dev(awefd_model *m, float *d_traces)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "NULL model");
    return run_common(m, d_traces, 0, NULL, NULL);
}
