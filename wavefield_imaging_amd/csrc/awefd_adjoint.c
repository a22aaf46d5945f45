/*
 * Misfit gradient of the cartesian 2-D loop with respect to squared slowness
 * (host C).  No counterpart in the reference, whose adj=y mode only injects
 * data backwards in time (MAWEFDcart2Dcpu.c:305-316); this is the adjoint of
 * the discrete forward loop the library runs (adj_kernels.hip).
 *
 * awefd_run_ckpt runs the forward loop in place, as awefd_run_dev does, and
 * copies the (current, previous) pair of fields at steps 0, K, 2K, ... into
 * checkpoint memory.  awefd_adjoint_grad visits the segments last to first:
 * it recomputes the fields of a segment from its checkpoint into K-1 segment
 * buffers (out-of-place steps, so the checkpoint survives and no copies are
 * made) and steps the adjoint backwards through them (four adjoint fields
 * in rotation and the dJ/dV accumulator).  Fields held: 2 ceil(nt/K)
 * checkpoints + K-1 segment buffers + 5.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "awefd_model.h"
#include "awefd_setup.h"

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            rc = awefd__fail(AWEFD_EHIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                             __FILE__, __LINE__);                                           \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)

typedef struct {
    int K, nck, nt;
    unsigned gen;             /* model generation of the checkpoints      */
    int valid;
    float *ck;                /* nck pairs (c_{kK}, c_{kK-1})             */
    float *seg;               /* K-1 fields c_{n0+1} .. c_{n0+K-1}        */
    float *A[4], *gV;         /* alpha fields in rotation, dJ/dV (padded) */
    /* receiver and injection nodes bucketed per adjoint tile column: of the
     * single-grid kernel's tiles, (suffix b) of the batched kernel's and
     * (suffix 2) of the two-step pass's overlapping tiles */
    awefd_inj *rinj, *sinj, *rinjb, *sinjb, *rinj2, *sinj2;
    int *rbeg, *sbeg, *rbegb, *sbegb, *rbeg2, *sbeg2;
    unsigned lists_gen;
    int lists_ok;
    awefd_adj2d_args *d_args; /* device copy of the kernel arguments    */
    float *ones;              /* sponge table of ones (runs without dabc) */
} adj_state;

static void free_lists(adj_state *s)
{
    hipFree(s->rinj); hipFree(s->sinj); hipFree(s->rbeg); hipFree(s->sbeg);
    hipFree(s->rinjb); hipFree(s->sinjb); hipFree(s->rbegb); hipFree(s->sbegb);
    hipFree(s->rinj2); hipFree(s->sinj2); hipFree(s->rbeg2); hipFree(s->sbeg2);
    s->rinj = s->sinj = s->rinjb = s->sinjb = s->rinj2 = s->sinj2 = NULL;
    s->rbeg = s->sbeg = s->rbegb = s->sbegb = s->rbeg2 = s->sbeg2 = NULL;
    s->lists_ok = 0;
}

void awefd__adjoint_free(awefd_model *m)
{
    adj_state *s = m ? m->adj : NULL;
    if (!s) return;
    hipSetDevice(m->dev);
    hipFree(s->ck); hipFree(s->seg);
    hipFree(s->A[0]); hipFree(s->A[1]); hipFree(s->A[2]); hipFree(s->A[3]); hipFree(s->gV); hipFree(s->d_args);
    hipFree(s->ones);
    free_lists(s);
    free(s);
    m->adj = NULL;
}

void awefd_adjoint_release(awefd_model *m)
{
    awefd__adjoint_free(m);
}

static int min_i(int a, int b)
{
    return a < b ? a : b;
}

/* Interpolation points of a 2-D grid as (node, id, weight) entries bucketed
 * per tile, uploaded to the device.  id is the point index (receivers) or
 * the wavelet column (sources: 0 with expl).  ov = 0: tiles of 64 x tx
 * nodes, each node in one; ov = nop: the two-step pass's tiles, whose
 * outputs are (64 - 2 ov) x (tx - 2 ov) nodes and whose extent reaches ov
 * beyond them, each node in every tile whose extent holds it (at most two
 * per axis). */
static int point_buckets(const awefd_model *m, int n, const int *j, const float *w, int same_id, int ntz,
                         int ntx, int tx, int ov, awefd_inj **d_inj, int **d_beg)
{
    const long cap = (ov ? 16L : 4L) * n;
    const int nbucket = ntz * ntx, sz = 64 - 2 * ov, sx = tx - 2 * ov;
    awefd_inj_key *keys = malloc(sizeof(*keys) * (cap > 0 ? cap : 1));
    awefd_inj *e = malloc(sizeof(*e) * (cap > 0 ? cap : 1));
    int *beg = malloc(sizeof(int) * (nbucket + 1));
    int rc = (!keys || !e || !beg) ? awefd__fail(AWEFD_ENOMEM, "gradient point buckets") : AWEFD_OK;
    long ne = 0;
    if (!rc) {
        for (int a = 0; a < n; a++)
            for (int q = 0; q < 4; q++) {
                awefd_inj en;
                en.z = j[3 * a] + (q & 1);
                en.x = j[3 * a + 1] + ((q >> 1) & 1);
                en.y = 0;
                en.src = same_id ? 0 : a;
                en.w = w[8 * a + q];
                en.w2 = 1.f;
                en.w3 = 1.f;
                if (!ov) {
                    keys[ne].e = en;
                    keys[ne++].bucket = en.z / 64 + ntz * (en.x / tx);
                    continue;
                }
                /* tile t's extent: [t s - ov, t s - ov + width) */
                const int vz = en.z + ov - 63, vx = en.x + ov - (tx - 1);
                const int tz0 = vz <= 0 ? 0 : (vz + sz - 1) / sz, tz1 = min_i((en.z + ov) / sz, ntz - 1);
                const int tx0 = vx <= 0 ? 0 : (vx + sx - 1) / sx, tx1 = min_i((en.x + ov) / sx, ntx - 1);
                for (int bx = tx0; bx <= tx1; bx++)
                    for (int bz = tz0; bz <= tz1; bz++) {
                        keys[ne].e = en;
                        keys[ne++].bucket = bz + ntz * bx;
                    }
            }
        awefd_sort_inj(keys, ne, nbucket, e, beg);
        hipError_t er = hipMalloc((void **)d_inj, sizeof(*e) * (ne > 0 ? ne : 1));
        if (er == hipSuccess) er = hipMalloc((void **)d_beg, sizeof(int) * (nbucket + 1));
        if (er == hipSuccess && ne) er = hipMemcpy(*d_inj, e, sizeof(*e) * ne, hipMemcpyHostToDevice);
        if (er == hipSuccess) er = hipMemcpy(*d_beg, beg, sizeof(int) * (nbucket + 1), hipMemcpyHostToDevice);
        if (er != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "gradient point buckets: %s", hipGetErrorString(er));
    }
    (void)m;
    free(keys); free(e); free(beg);
    return rc;
}

static int check_model(awefd_model *m, const char *who)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "%s: NULL model", who);
    if (m->p.kind != AWEFD_CART2D) return awefd__fail(AWEFD_EINVAL, "%s: cartesian 2-D models only", who);
    if (!m->p.cden && m->drho)
        return awefd__fail(AWEFD_EINVAL, "%s: constant density only (cden=y or no density set)", who);
    if (m->p.adj || m->hyb || m->sinc)
        return awefd__fail(AWEFD_EINVAL, "%s: adj=y, hybridbc=y and sinc=y loops have no gradient", who);
    if (!m->model_set || !m->src_set || !m->rec_set || !m->wav_set)
        return awefd__fail(AWEFD_ESTATE, "%s: model not fully set up", who);
    const long need = (long)m->p.nt * (m->p.expl ? 1 : m->ns);
    if (m->nwav < need)
        return awefd__fail(AWEFD_EINVAL, "%s: wavelet has %ld samples, the loop reads %ld", who, m->nwav, need);
    return AWEFD_OK;
}

static int dev_alloc(float **p, size_t bytes)
{
    if (*p) return AWEFD_OK;
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = NULL;
        return awefd__fail(AWEFD_ENOMEM, "gradient buffers (%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    return AWEFD_OK;
}

static adj_state *state(const awefd_model *m)
{
    return (adj_state *)m->adj;
}

int awefd__adj_ckpt_setup(awefd_model *m, int ckpt)
{
    int rc = check_model(m, "checkpointed run");
    if (rc) return rc;
    const int nt = m->p.nt;
    const int K = ckpt > 0 ? (ckpt < nt ? ckpt : nt) : (int)ceil(sqrt((double)nt));
    const int nck = (nt + K - 1) / K;
    if ((size_t)m->nalloc * 4 >= 0x80000000u)
        return awefd__fail(AWEFD_EINVAL, "gradient: field of %ld nodes exceeds 2 GiB", m->nalloc);
    if (hipSetDevice(m->dev) != hipSuccess) return awefd__fail(AWEFD_EHIP, "hipSetDevice");
    adj_state *s = state(m);
    if (!s) {
        s = calloc(1, sizeof *s);
        if (!s) return awefd__fail(AWEFD_ENOMEM, "gradient state");
        m->adj = s;
    }
    s->valid = 0;
    if (s->K != K || s->nck != nck) {
        hipFree(s->ck);
        hipFree(s->seg);
        s->ck = s->seg = NULL;
        s->K = K;
        s->nck = nck;
    }
    return dev_alloc(&s->ck, sizeof(float) * (size_t)m->nalloc * 2 * nck);
}

float *awefd__adj_ckpt(const awefd_model *m, int k)
{
    return state(m)->ck + (size_t)2 * k * m->nalloc;
}

int awefd__adj_K(const awefd_model *m)
{
    return state(m) ? state(m)->K : 0;
}

void awefd__adj_mark(awefd_model *m)
{
    adj_state *s = state(m);
    s->nt = m->p.nt;
    s->gen = m->gen;
    s->valid = 1;
}

int awefd__adj_ready(const awefd_model *m, const char *who)
{
    int rc = check_model((awefd_model *)m, who);
    if (rc) return rc;
    const adj_state *s = state(m);
    if (!s || !s->valid || s->gen != m->gen || s->nt != m->p.nt)
        return awefd__fail(AWEFD_ESTATE, "%s: no checkpoints of the current model (run with checkpoints first)", who);
    return AWEFD_OK;
}

int awefd_run_ckpt(awefd_model *m, float *d_traces, int ckpt)
{
    int rc = awefd__adj_ckpt_setup(m, ckpt);
    if (rc) return rc;
    const adj_state *s = state(m);
    const int nt = m->p.nt, K = s->K;
    const size_t fb = sizeof(float) * (size_t)m->nalloc;
    awefd_cart_args a;
    if ((rc = awefd__ensure_traces(m))) goto done;
    if ((rc = awefd_reset(m))) goto done;
    awefd__cart_args(m, &a);
    for (int it = 0; it < nt; it++) {
        if (it % K == 0) {
            float *c = awefd__adj_ckpt(m, it / K);
            HIPCHK(hipMemcpyAsync(c, m->dC, fb, hipMemcpyDeviceToDevice, m->stream));
            HIPCHK(hipMemcpyAsync(c + m->nalloc, m->dP, fb, hipMemcpyDeviceToDevice, m->stream));
        }
        a.P = m->dP;
        a.C = m->dC;
        a.it = it;
        a.trow = (d_traces ? d_traces : m->d_traces) + (long)it * m->nr;
        int e = awefd_launch_cart_step(&a, m->nop, 2, m->p.sse, 0, m->stream);
        if (e) {
            rc = awefd__fail(AWEFD_EHIP, "step %d launch: %s", it, hipGetErrorString((hipError_t)e));
            goto done;
        }
        float *t = m->dP;
        m->dP = m->dC;
        m->dC = t;
        m->steps_done = it + 1;
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    awefd__adj_mark(m);
done:
    return rc;
}

/* Tiles of the two-step pass: (64 - 2 nop) x (AWEFD_ADJ2_ROWS - 2 nop)
 * outputs each. */
static void adj2_tiles(const awefd_model *m, int *ntz2, int *ntx2)
{
    const int sz = 64 - 2 * m->nop, sx = AWEFD_ADJ2_ROWS - 2 * m->nop;
    *ntz2 = (m->np[0] + sz - 1) / sz;
    *ntx2 = (m->np[1] + sx - 1) / sx;
}

/* Buffers and point buckets of the backward pass (kept across calls). */
int awefd__adj_prepare(awefd_model *m)
{
    adj_state *s = state(m);
    const size_t fb = sizeof(float) * (size_t)m->nalloc;
    int rc = AWEFD_OK;
    if (hipSetDevice(m->dev) != hipSuccess) return awefd__fail(AWEFD_EHIP, "hipSetDevice");
    if (s->K > 1) rc = dev_alloc(&s->seg, fb * (s->K - 1));
    for (int i = 0; i < 4 && !rc; i++) rc = dev_alloc(&s->A[i], fb);
    if (!rc) rc = dev_alloc(&s->gV, fb);
    if (!rc && !m->p.dabc && !s->ones) {
        const long n = m->np[0] > m->np[1] ? m->np[0] : m->np[1];
        rc = dev_alloc(&s->ones, sizeof(float) * n);
        if (!rc && awefd_launch_fill(s->ones, n, 1.f, m->stream))
            rc = awefd__fail(AWEFD_EHIP, "gradient sponge table");
    }
    if (!rc && !s->d_args) {
        hipError_t e = hipMalloc((void **)&s->d_args, sizeof(awefd_adj2d_args));
        if (e != hipSuccess) rc = awefd__fail(AWEFD_ENOMEM, "gradient arguments: %s", hipGetErrorString(e));
    }
    if (rc) return rc;
    if (!s->lists_ok || s->lists_gen != m->gen) {
        free_lists(s);
        const int ntz = (m->np[0] + 63) / 64, tx = AWEFD_ADJ_TX, txb = AWEFD_ADJ_TXB;
        const int ntx = (m->np[1] + tx - 1) / tx, ntxb = (m->np[1] + txb - 1) / txb;
        int ntz2, ntx2;
        adj2_tiles(m, &ntz2, &ntx2);
        const int ex = m->p.expl;
        if ((rc = point_buckets(m, m->nr, m->h_rj, m->h_rw, 0, ntz, ntx, tx, 0, &s->rinj, &s->rbeg))) return rc;
        if ((rc = point_buckets(m, m->ns, m->h_sj, m->h_sw, ex, ntz, ntx, tx, 0, &s->sinj, &s->sbeg))) return rc;
        if ((rc = point_buckets(m, m->nr, m->h_rj, m->h_rw, 0, ntz, ntxb, txb, 0, &s->rinjb, &s->rbegb))) return rc;
        if ((rc = point_buckets(m, m->ns, m->h_sj, m->h_sw, ex, ntz, ntxb, txb, 0, &s->sinjb, &s->sbegb)))
            return rc;
        if ((rc = point_buckets(m, m->nr, m->h_rj, m->h_rw, 0, ntz2, ntx2, AWEFD_ADJ2_ROWS, m->nop, &s->rinj2,
                                &s->rbeg2)))
            return rc;
        if ((rc = point_buckets(m, m->ns, m->h_sj, m->h_sw, ex, ntz2, ntx2, AWEFD_ADJ2_ROWS, m->nop, &s->sinj2,
                                &s->sbeg2)))
            return rc;
        s->lists_ok = 1;
        s->lists_gen = m->gen;
    }
    return AWEFD_OK;
}

/* Kernel arguments of the backward pass (after awefd__adj_prepare). */
void awefd__adj_args(const awefd_model *m, const float *d_res, awefd_adj2d_args *g)
{
    const adj_state *s = state(m);
    memset(g, 0, sizeof *g);
    for (int i = 0; i < 4; i++) g->A[i] = s->A[i];
    g->ck = s->ck;
    g->seg = s->seg;
    g->nal = m->nalloc;
    g->K = s->K;
    g->res = d_res;
    g->nr = m->nr;
    g->V = m->dvel;
    g->gV = s->gV;
    g->nz = m->np[0];
    g->nx = m->np[1];
    g->ld = m->ld;
    g->mz = m->p.n[0];
    g->mx = m->p.n[1];
    g->nb = m->nb;
    g->dabc = m->p.dabc;
    g->idt2 = (float)(1.0 / ((double)m->p.dt * m->p.dt));
    g->sp1 = m->p.dabc ? m->dsp[0] : s->ones;
    g->sp2 = m->p.dabc ? m->dsp[1] : s->ones;
    g->c0 = m->c2[0];
    for (int k = 1; k <= m->nop; k++) {
        g->cz[k - 1] = m->c2[k];
        g->cx[k - 1] = m->c2[m->nop + k];
    }
    g->ntz = (m->np[0] + 63) / 64;
    g->ntx = (m->np[1] + AWEFD_ADJ_TX - 1) / AWEFD_ADJ_TX;
    g->nblk = g->ntz * g->ntx;
    g->rbeg = s->rbeg;
    g->rinj = s->rinj;
    g->sbeg = s->sbeg;
    g->sinj = s->sinj;
    g->wav = m->d_wav;
    g->wstride = m->p.expl ? 1 : m->ns;
    adj2_tiles(m, &g->ntz2, &g->ntx2);
    g->nblk2 = g->ntz2 * g->ntx2;
    g->rbeg2 = s->rbeg2;
    g->rinj2 = s->rinj2;
    g->sbeg2 = s->sbeg2;
    g->sinj2 = s->sinj2;
}

/* The same for the batched kernel, whose tiles are 64 x AWEFD_ADJ_TXB. */
void awefd__adj_args_batch(const awefd_model *m, const float *d_res, awefd_adj2d_args *g)
{
    const adj_state *s = state(m);
    awefd__adj_args(m, d_res, g);
    g->ntx = (m->np[1] + AWEFD_ADJ_TXB - 1) / AWEFD_ADJ_TXB;
    g->nblk = g->ntz * g->ntx;
    g->rbeg = s->rbegb;
    g->rinj = s->rinjb;
    g->sbeg = s->sbegb;
    g->sinj = s->sinjb;
}

/* Checkpoint segment bases of the batched recompute. */
void awefd__adj_cart_ckpt(const awefd_model *m, awefd_cart_args *a)
{
    const adj_state *s = state(m);
    a->ck = s->ck;
    a->seg = s->seg;
    a->nal = m->nalloc;
    a->K = s->K;
}

/* The buffers a backward pass starts from zero: alpha_nt, alpha_{nt+1}
 * and dJ/dV (src NULL: fill). */
void awefd__adj_zero_jobs(const awefd_model *m, awefd_copy_job *jobs)
{
    const adj_state *s = state(m);
    const int nt = m->p.nt;
    float *z[3] = {s->A[nt & 3], s->A[(nt + 1) & 3], s->gV};
    for (int i = 0; i < 3; i++) {
        jobs[i].src = NULL;
        jobs[i].dst = z[i];
        jobs[i].n = m->nalloc;
    }
}

int awefd_adjoint_grad(awefd_model *m, const float *d_res, float *d_grad)
{
    int rc = awefd__adj_ready(m, "awefd_adjoint_grad");
    if (rc) return rc;
    if (!d_res || !d_grad) return awefd__fail(AWEFD_EINVAL, "awefd_adjoint_grad: NULL buffer");
    adj_state *s = state(m);
    const int nt = m->p.nt, K = s->K, nck = s->nck;
    awefd_cart_args a;
    awefd_adj2d_args g;
    if ((rc = awefd__adj_prepare(m))) goto done;
    awefd__adj_args(m, d_res, &g);
    HIPCHK(hipMemcpy(s->d_args, &g, sizeof g, hipMemcpyHostToDevice));
    {
        awefd_copy_job z[3];
        awefd__adj_zero_jobs(m, z);
        for (int i = 0; i < 3; i++) {
            int e = awefd_launch_fill(z[i].dst, z[i].n, 0.f, m->stream);
            if (e) {
                rc = awefd__fail(AWEFD_EHIP, "gradient reset: %s", hipGetErrorString((hipError_t)e));
                goto done;
            }
        }
    }
    /* recompute steps: out of place, no trace extraction */
    awefd__cart_args(m, &a);
    a.nr = 0;
    for (int k = nck - 1; k >= 0; k--) {
        const int n0 = k * K, n1 = n0 + K < nt ? n0 + K : nt;
        float *ckc = awefd__adj_ckpt(m, k), *ckp = ckc + m->nalloc;
#define FIELD(n) ((n) == n0 ? ckc : s->seg + (size_t)((n) - n0 - 1) * m->nalloc)
        for (int n = n0; n + 1 < n1; n++) {
            a.C = FIELD(n);
            a.P = n == n0 ? ckp : FIELD(n - 1);
            a.Pn = FIELD(n + 1);
            a.it = n;
            int e = awefd_launch_cart_step(&a, m->nop, 2, m->p.sse, 0, m->stream);
            if (e) {
                rc = awefd__fail(AWEFD_EHIP, "recompute step %d: %s", n, hipGetErrorString((hipError_t)e));
                goto done;
            }
        }
#undef FIELD
        for (int n = n1 - 1; n >= n0; n--) {
            int e = awefd_launch_adj2d_step(&g, n, m->nop, m->stream);
            if (e) {
                rc = awefd__fail(AWEFD_EHIP, "adjoint step %d: %s", n, hipGetErrorString((hipError_t)e));
                goto done;
            }
        }
    }
    {
        int e = awefd_launch_adj2d_chain(&g, s->d_args, 1, d_grad, m->stream);
        if (e) {
            rc = awefd__fail(AWEFD_EHIP, "gradient chain: %s", hipGetErrorString((hipError_t)e));
            goto done;
        }
    }
    HIPCHK(hipStreamSynchronize(m->stream));
done:
    if (rc) hipStreamSynchronize(m->stream);
    return rc;
}
