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
 * made) and steps the adjoint backwards through them.  Fields held:
 * 2 ceil(nt/K) checkpoints + K-1 segment buffers + 5 adjoint fields.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "awefd_model.h"

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
    float *A[2], *B[2], *gV;  /* adjoint fields and dJ/dV (padded)        */
    /* per-node receiver / source weight lists */
    long *rnode, *snode;
    int *rbeg, *rid, *sbeg, *sid;
    float *rw, *sw;
    int nrn, nsn;
    unsigned lists_gen;
    int lists_ok;
} adj_state;

static void free_lists(adj_state *s)
{
    hipFree(s->rnode); hipFree(s->rbeg); hipFree(s->rid); hipFree(s->rw);
    hipFree(s->snode); hipFree(s->sbeg); hipFree(s->sid); hipFree(s->sw);
    s->rnode = s->snode = NULL;
    s->rbeg = s->rid = s->sbeg = s->sid = NULL;
    s->rw = s->sw = NULL;
    s->lists_ok = 0;
}

void awefd__adjoint_free(awefd_model *m)
{
    adj_state *s = m ? m->adj : NULL;
    if (!s) return;
    hipSetDevice(m->dev);
    hipFree(s->ck); hipFree(s->seg);
    hipFree(s->A[0]); hipFree(s->A[1]); hipFree(s->B[0]); hipFree(s->B[1]); hipFree(s->gV);
    free_lists(s);
    free(s);
    m->adj = NULL;
}

void awefd_adjoint_release(awefd_model *m)
{
    awefd__adjoint_free(m);
}

typedef struct {
    long node;
    int order, id;
    float w;
} pt_entry;

static int pt_cmp(const void *pa, const void *pb)
{
    const pt_entry *a = pa, *b = pb;
    if (a->node != b->node) return a->node < b->node ? -1 : 1;
    return a->order < b->order ? -1 : (a->order > b->order);
}

/* Lists of (id, weight) per grid node of n interpolation points: nodes in
 * ascending order, each node's entries in point order. */
static int node_lists(int n, const int *j, const float *w, int ld, int same_id, long **d_node, int **d_beg,
                      int **d_id, float **d_w, int *nnode)
{
    const size_t ne = (size_t)4 * (n > 0 ? n : 1);
    pt_entry *e = malloc(sizeof(*e) * ne);
    long *node = malloc(sizeof(long) * ne);
    int *beg = malloc(sizeof(int) * (ne + 1)), *id = malloc(sizeof(int) * ne);
    float *wt = malloc(sizeof(float) * ne);
    int rc = (!e || !node || !beg || !id || !wt) ? awefd__fail(AWEFD_ENOMEM, "gradient point lists") : AWEFD_OK;
    int nn = 0;
    if (!rc) {
        for (int a = 0; a < n; a++)
            for (int q = 0; q < 4; q++) {
                pt_entry *x = &e[4 * a + q];
                x->node = (long)(j[3 * a + 1] + ((q >> 1) & 1)) * ld + j[3 * a] + (q & 1);
                x->order = 4 * a + q;
                x->id = same_id ? 0 : a;
                x->w = w[8 * a + q];
            }
        qsort(e, (size_t)4 * n, sizeof *e, pt_cmp);
        for (long i = 0; i < 4L * n; i++) {
            if (i == 0 || e[i].node != e[i - 1].node) {
                node[nn] = e[i].node;
                beg[nn] = (int)i;
                nn++;
            }
            id[i] = e[i].id;
            wt[i] = e[i].w;
        }
        beg[nn] = 4 * n;
        hipError_t er = hipMalloc((void **)d_node, sizeof(long) * (nn > 0 ? nn : 1));
        if (er == hipSuccess) er = hipMalloc((void **)d_beg, sizeof(int) * (nn + 1));
        if (er == hipSuccess) er = hipMalloc((void **)d_id, sizeof(int) * ne);
        if (er == hipSuccess) er = hipMalloc((void **)d_w, sizeof(float) * ne);
        if (er == hipSuccess && nn) er = hipMemcpy(*d_node, node, sizeof(long) * nn, hipMemcpyHostToDevice);
        if (er == hipSuccess) er = hipMemcpy(*d_beg, beg, sizeof(int) * (nn + 1), hipMemcpyHostToDevice);
        if (er == hipSuccess && n) er = hipMemcpy(*d_id, id, sizeof(int) * 4 * n, hipMemcpyHostToDevice);
        if (er == hipSuccess && n) er = hipMemcpy(*d_w, wt, sizeof(float) * 4 * n, hipMemcpyHostToDevice);
        if (er != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "gradient point lists: %s", hipGetErrorString(er));
    }
    *nnode = nn;
    free(e); free(node); free(beg); free(id); free(wt);
    return rc;
}

static int check_model(awefd_model *m, const char *who)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "%s: NULL model", who);
    if (m->p.kind != AWEFD_CART2D) return awefd__fail(AWEFD_EINVAL, "%s: cartesian 2-D models only", who);
    if (!m->p.cden && m->drho)
        return awefd__fail(AWEFD_EINVAL, "%s: constant density only (cden=y or no density set)", who);
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

int awefd_run_ckpt(awefd_model *m, float *d_traces, int ckpt)
{
    int rc = check_model(m, "awefd_run_ckpt");
    if (rc) return rc;
    const int nt = m->p.nt;
    const int K = ckpt > 0 ? (ckpt < nt ? ckpt : nt) : (int)ceil(sqrt((double)nt));
    const int nck = (nt + K - 1) / K;
    const size_t fb = sizeof(float) * (size_t)m->nalloc;
    awefd_cart_args a;
    HIPCHK(hipSetDevice(m->dev));
    adj_state *s = m->adj;
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
    if ((rc = dev_alloc(&s->ck, fb * 2 * nck))) goto done;
    if ((rc = awefd__ensure_traces(m))) goto done;
    if ((rc = awefd_reset(m))) goto done;
    awefd__cart_args(m, &a);
    for (int it = 0; it < nt; it++) {
        if (it % K == 0) {
            float *c = s->ck + (size_t)2 * (it / K) * m->nalloc;
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
    s->nt = nt;
    s->gen = m->gen;
    s->valid = 1;
done:
    return rc;
}

int awefd_adjoint_grad(awefd_model *m, const float *d_res, float *d_grad)
{
    int rc = check_model(m, "awefd_adjoint_grad");
    if (rc) return rc;
    if (!d_res || !d_grad) return awefd__fail(AWEFD_EINVAL, "awefd_adjoint_grad: NULL buffer");
    adj_state *s = m->adj;
    if (!s || !s->valid || s->gen != m->gen || s->nt != m->p.nt)
        return awefd__fail(AWEFD_ESTATE, "awefd_adjoint_grad: no checkpoints of the current model "
                                         "(call awefd_run_ckpt first)");
    const int nt = m->p.nt, K = s->K, nck = s->nck;
    const size_t fb = sizeof(float) * (size_t)m->nalloc;
    awefd_cart_args a;
    awefd_adj2d_args g;
    HIPCHK(hipSetDevice(m->dev));
    if (K > 1 && (rc = dev_alloc(&s->seg, fb * (K - 1)))) goto done;
    for (int i = 0; i < 2 && !rc; i++) {
        rc = dev_alloc(&s->A[i], fb);
        if (!rc) rc = dev_alloc(&s->B[i], fb);
    }
    if (!rc) rc = dev_alloc(&s->gV, fb);
    if (rc) goto done;
    if (!s->lists_ok || s->lists_gen != m->gen) {
        free_lists(s);
        if ((rc = node_lists(m->nr, m->h_rj, m->h_rw, m->ld, 0, &s->rnode, &s->rbeg, &s->rid, &s->rw, &s->nrn)))
            goto done;
        if ((rc = node_lists(m->ns, m->h_sj, m->h_sw, m->ld, m->p.expl, &s->snode, &s->sbeg, &s->sid, &s->sw,
                             &s->nsn)))
            goto done;
        s->lists_ok = 1;
        s->lists_gen = m->gen;
    }

    memset(&g, 0, sizeof g);
    g.V = m->dvel;
    g.nz = m->np[0];
    g.nx = m->np[1];
    g.ld = m->ld;
    g.mz = m->p.n[0];
    g.mx = m->p.n[1];
    g.nb = m->nb;
    g.dabc = m->p.dabc;
    g.idt2 = (float)(1.0 / ((double)m->p.dt * m->p.dt));
    g.sp1 = m->dsp[0];
    g.sp2 = m->dsp[1];
    g.c0 = m->c2[0];
    for (int k = 1; k <= m->nop; k++) {
        g.cz[k - 1] = m->c2[k];
        g.cx[k - 1] = m->c2[m->nop + k];
    }
    g.nrnode = s->nrn;
    g.rnode = s->rnode;
    g.rbeg = s->rbeg;
    g.rid = s->rid;
    g.rw = s->rw;
    g.nsnode = s->nsn;
    g.snode = s->snode;
    g.sbeg = s->sbeg;
    g.sid = s->sid;
    g.sw = s->sw;
    g.wav = m->d_wav;
    g.wstride = m->p.expl ? 1 : m->ns;
    g.gV = s->gV;
    {
        int e = awefd_launch_fill(s->A[1], m->nalloc, 0.f, m->stream);
        if (!e) e = awefd_launch_fill(s->B[1], m->nalloc, 0.f, m->stream);
        if (!e) e = awefd_launch_fill(s->gV, m->nalloc, 0.f, m->stream);
        if (e) {
            rc = awefd__fail(AWEFD_EHIP, "gradient reset: %s", hipGetErrorString((hipError_t)e));
            goto done;
        }
    }
    /* recompute steps: out of place, no trace extraction */
    awefd__cart_args(m, &a);
    a.nr = 0;
    int cur = 1;   /* A[cur], B[cur] hold alpha_{n+1}, beta_{n+1} */
    for (int k = nck - 1; k >= 0; k--) {
        const int n0 = k * K, n1 = n0 + K < nt ? n0 + K : nt;
        float *ckc = s->ck + (size_t)2 * k * m->nalloc, *ckp = ckc + m->nalloc;
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
        for (int n = n1 - 1; n >= n0; n--) {
            g.A1 = s->A[cur];
            g.B1 = s->B[cur];
            g.A0 = s->A[cur ^ 1];
            g.B0 = s->B[cur ^ 1];
            g.Cn = FIELD(n);
            g.res = d_res + (long)n * m->nr;
            g.it = n;
            int e = awefd_launch_adj2d_step(&g, m->nop, m->stream);
            if (e) {
                rc = awefd__fail(AWEFD_EHIP, "adjoint step %d: %s", n, hipGetErrorString((hipError_t)e));
                goto done;
            }
            cur ^= 1;
        }
#undef FIELD
    }
    {
        int e = awefd_launch_adj2d_chain(&g, d_grad, m->stream);
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
