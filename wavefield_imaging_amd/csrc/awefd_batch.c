/*
 * Batched shots (host C): many independent 2-D cartesian models advanced
 * together, one kernel launch per time step over every shot, the whole time
 * loop captured into one HIP graph and launched once.
 *
 * The reference runs one program per shot (byrne_awe.py loops over shots and
 * calls sfAWEFDcart2D for each, byrne_awe.py:163-250); a shot here is a
 * complete awefd_model with its own grid, model, sources, receivers and
 * wavelet, so every shot's results are those of its own run.
 *
 * Two tuning knobs, chosen by awefd_batch_tune on the batch itself:
 *   tile   x rows per wave of the batched kernel (compiled for 2, 4, 8, 16);
 *   group  shots are independent, so the loop may run a group of shots
 *          through all its time steps before the next group starts.  A group
 *          whose fields fit in the 256 MB Infinity Cache (MALL) re-reads them
 *          from the cache instead of HBM every step.  group = 0 runs every
 *          shot in every launch.
 * The injection buckets and block tables are rebuilt for each choice.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
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

static const int RX_CAND[4] = {2, 4, 8, 16};
#define NGC 5
static const int GROUP_CAND[NGC] = {0, 32, 64, 128, 192};   /* MiB of fields per group */

/* Copies that end a multi-step run: a field left in the batch's extra pair
 * goes back into the model's own buffer (the slot the other field does not
 * occupy); fin_p, fin_c: the slots (0: u0, 1: u1) of the previous and
 * current field afterwards. */
typedef struct {
    awefd_copy_job *d_jobs;
    long *d_cbeg, nch;
    int njob, fin_p, fin_c;
} final_copies;

struct awefd_batch {
    int nshot;
    awefd_model **m;
    float **u0, **u1;                /* each shot's two field buffers            */
    int nt, nop, sse, dev;
    hipStream_t stream;
    int rx;                          /* current tile: rows per wave              */
    int group_mb;                    /* current group budget (0: one group)      */
    int ngroup;
    long *goff;                      /* per group: first entry in d_bshot        */
    int *gnblk;                      /* per group: blocks of one launch          */
    awefd_cart_args *d_shots;
    int *d_bshot;
    awefd_inj **d_inj;               /* per shot, bucketed for the current tile  */
    int **d_ibeg;
    hipGraph_t graph;
    hipGraphExec_t gexec;
    int graph_ok;
    double tune_ms[4][NGC];
    int *grp;                        /* group of each shot                       */
    awefd_cart_args *h;              /* host copy of d_shots                     */
    /* gradient: checkpointed forward and backward pass (awefd_adjoint.c holds
     * each shot's buffers) */
    int K, nck;
    awefd_copy_job *d_jobs;          /* checkpoint copies per (group, k)         */
    long *d_cbeg;
    long *joff, *jnch;               /* per (group, k): first job, chunks        */
    awefd_adj2d_args *ha, *d_adj;    /* per-shot adjoint arguments               */
    int *d_abshot;                   /* adjoint launch: block -> shot            */
    long *agoff;
    int *agnblk;
    int *d_abshot2;                  /* two-step adjoint launch: block -> shot   */
    long *agoff2;
    int *agnblk2;
    long abshot2_cap;
    awefd_copy_job *d_zjobs;         /* zero fills starting the backward pass    */
    long *d_zcbeg, znch;
    long abshot_cap;
    float **fwd_ck;                  /* checkpoint bases the forward graph holds */
    /* two or three time steps per pass (depth 2, 3) */
    int depth;                       /* 1, 2 or 3                                */
    float **e0, **e1;                /* per shot: the extra field pair           */
    int *d_bshot2;                   /* two-step launch: block -> shot           */
    long *goff2;
    int *gnblk2;
    awefd_inj **d_inj2;              /* per shot: two-step injection buckets     */
    int **d_ibeg2;
    int **d_rbeg3;                   /* per shot: three-step receiver buckets    */
    awefd_rx3 **d_rx3;               /* and their records                        */
    int t2_ok;                       /* multi-step tables built for the tile     */
    int t2_depth;                    /* ... and for this depth                   */
    final_copies fin;                /* end of a multi-step run                  */
    final_copies kfin;               /* end of a three-step checkpointed run     */
    double tune2_ms[4][NGC], tune3_ms[4][NGC];
    hipGraph_t gfwd, gbwd;
    hipGraphExec_t xfwd, xbwd;
    int fwd_ok, bwd_ok;
    int fwd_depth, bwd_depth;        /* steps per pass the graphs were made for  */
    int bwd_adepth;                  /* ... and adjoint steps per pass           */
    float *bwd_gout;
    unsigned *mgen;                  /* each model's generation when the tables
                                        were built (setters bump it)           */
    int built;                       /* tables built for the current tile       */
};

static void dfree(void *pp)
{
    void **p = (void **)pp;
    if (*p) hipFree(*p);
    *p = NULL;
}

void awefd_batch_destroy(awefd_batch *b)
{
    if (!b) return;
    hipSetDevice(b->dev);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->xfwd) hipGraphExecDestroy(b->xfwd);
    if (b->gfwd) hipGraphDestroy(b->gfwd);
    if (b->xbwd) hipGraphExecDestroy(b->xbwd);
    if (b->gbwd) hipGraphDestroy(b->gbwd);
    dfree(&b->d_jobs);
    dfree(&b->d_cbeg);
    dfree(&b->d_adj);
    dfree(&b->d_abshot);
    dfree(&b->d_abshot2);
    dfree(&b->d_zjobs);
    dfree(&b->d_zcbeg);
    free(b->joff);
    free(b->jnch);
    free(b->ha);
    free(b->agoff);
    free(b->agnblk);
    free(b->agoff2);
    free(b->agnblk2);
    free(b->grp);
    free(b->h);
    free(b->fwd_ck);
    for (int s = 0; s < b->nshot; s++) {
        if (b->e0) dfree(&b->e0[s]);
        if (b->e1) dfree(&b->e1[s]);
        if (b->d_inj2) dfree(&b->d_inj2[s]);
        if (b->d_ibeg2) dfree(&b->d_ibeg2[s]);
        if (b->d_rbeg3) dfree(&b->d_rbeg3[s]);
        if (b->d_rx3) dfree(&b->d_rx3[s]);
    }
    free(b->e0);
    free(b->e1);
    free(b->d_inj2);
    free(b->d_ibeg2);
    free(b->d_rbeg3);
    free(b->d_rx3);
    dfree(&b->d_bshot2);
    dfree(&b->fin.d_jobs);
    dfree(&b->fin.d_cbeg);
    dfree(&b->kfin.d_jobs);
    dfree(&b->kfin.d_cbeg);
    free(b->goff2);
    free(b->gnblk2);
    for (int s = 0; s < b->nshot; s++) {
        if (b->d_inj) dfree(&b->d_inj[s]);
        if (b->d_ibeg) dfree(&b->d_ibeg[s]);
    }
    free(b->d_inj);
    free(b->d_ibeg);
    dfree(&b->d_shots);
    dfree(&b->d_bshot);
    if (b->stream) hipStreamDestroy(b->stream);
    free(b->m);
    free(b->mgen);
    free(b->u0);
    free(b->u1);
    free(b->goff);
    free(b->gnblk);
    free(b);
}

/* A shot the batched kernels can run: set up, plain 2-D, like shot 0. */
static int shot_ok(const awefd_model *m, const awefd_model *m0, int s)
{
    if (!m) return awefd__fail(AWEFD_EINVAL, "batch: shot %d is NULL", s);
    if (m->p.kind != AWEFD_CART2D)
        return awefd__fail(AWEFD_EINVAL, "batch: shot %d is not a cartesian 2-D model", s);
    if (m->p.nt != m0->p.nt || m->nop != m0->nop || m->p.sse != m0->p.sse || m->dev != m0->dev)
        return awefd__fail(AWEFD_EINVAL, "batch: shot %d differs from shot 0 in nt, fdorder, sse or device", s);
    if (m->drho && !m->p.cden)
        return awefd__fail(AWEFD_EINVAL, "batch: shot %d has variable density (not batched)", s);
    if (m->p.adj || m->hyb || m->sinc)
        return awefd__fail(AWEFD_EINVAL, "batch: shot %d uses adj=y, hybridbc=y or sinc=y (not batched)", s);
    if (!m->model_set || !m->src_set || !m->rec_set || !m->wav_set)
        return awefd__fail(AWEFD_ESTATE, "batch: shot %d is not fully set up", s);
    const long need = (long)m->p.nt * (m->p.expl ? 1 : m->ns);
    if (m->nwav < need) return awefd__fail(AWEFD_EINVAL, "batch: shot %d wavelet too short", s);
    return AWEFD_OK;
}

int awefd_batch_create(int nshot, awefd_model *const *models, awefd_batch **out)
{
    if (!out || nshot < 1 || !models) return awefd__fail(AWEFD_EINVAL, "awefd_batch_create: bad argument");
    *out = NULL;
    int rc0;
    for (int s = 0; s < nshot; s++)
        if ((rc0 = shot_ok(models[s], models[0], s))) return rc0;
    awefd_batch *b = calloc(1, sizeof *b);
    if (!b) return awefd__fail(AWEFD_ENOMEM, "batch");
    b->nshot = nshot;
    b->m = malloc(sizeof(*b->m) * nshot);
    b->mgen = calloc(nshot, sizeof(*b->mgen));
    b->d_inj = calloc(nshot, sizeof(*b->d_inj));
    b->d_ibeg = calloc(nshot, sizeof(*b->d_ibeg));
    b->u0 = malloc(sizeof(*b->u0) * nshot);
    b->u1 = malloc(sizeof(*b->u1) * nshot);
    if (!b->m || !b->mgen || !b->d_inj || !b->d_ibeg || !b->u0 || !b->u1) {
        awefd_batch_destroy(b);
        return awefd__fail(AWEFD_ENOMEM, "batch");
    }
    memcpy(b->m, models, sizeof(*b->m) * nshot);
    for (int s = 0; s < nshot; s++) {
        b->u0[s] = models[s]->dP;
        b->u1[s] = models[s]->dC;
    }
    const awefd_model *m0 = models[0];
    b->nt = m0->p.nt;
    b->nop = m0->nop;
    b->sse = m0->p.sse;
    b->dev = m0->dev;
    hipError_t e = hipSetDevice(b->dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        awefd_batch_destroy(b);
        return awefd__fail(AWEFD_EHIP, "batch stream: %s", hipGetErrorString(e));
    }
    /* every shot's traces live on the device */
    for (int s = 0; s < nshot; s++) {
        int rc = awefd__ensure_traces(models[s]);
        if (rc) {
            awefd_batch_destroy(b);
            return rc;
        }
    }
    *out = b;
    return AWEFD_OK;
}

/* Per-shot arguments, injection buckets, shot groups and the block -> shot
 * table for tile width TX = rx * AWEFD_BATCH_NWAVE and group budget group_mb. */
static int batch_build(awefd_batch *b, int rx, int group_mb)
{
    const int tx = rx * AWEFD_BATCH_NWAVE, nthr = AWEFD_BATCH_NWAVE * 64;
    awefd_cart_args *h = malloc(sizeof(*h) * b->nshot);
    int *grp = malloc(sizeof(int) * b->nshot), *nb = malloc(sizeof(int) * b->nshot);
    if (!h || !grp || !nb) {
        free(h); free(grp); free(nb);
        return awefd__fail(AWEFD_ENOMEM, "batch tables");
    }
    /* groups: consecutive shots while their fields fit the budget */
    const double budget = group_mb > 0 ? group_mb * 1048576.0 : 0.0;
    double acc = 0.0;
    int ng = 0;
    for (int s = 0; s < b->nshot; s++) {
        const double fp = 3.0 * 4.0 * (double)b->m[s]->nalloc;
        if (s > 0 && budget > 0 && acc + fp > budget) {
            ng++;
            acc = 0.0;
        }
        grp[s] = ng;
        acc += fp;
    }
    ng++;
    long *goff = calloc(ng + 1, sizeof(long));
    int *gnblk = calloc(ng, sizeof(int));
    int rc = (!goff || !gnblk) ? awefd__fail(AWEFD_ENOMEM, "batch groups") : AWEFD_OK;
    for (int s = 0; s < b->nshot && !rc; s++) {
        awefd_model *m = b->m[s];
        awefd__cart_args(m, &h[s]);
        h[s].ntx = (m->np[1] + tx - 1) / tx;
        h[s].nty = 1;
        h[s].nblk = h[s].ntz * h[s].ntx;
        h[s].P = b->u0[s];
        h[s].C = b->u1[s];
        h[s].trow = m->d_traces;
        nb[s] = h[s].nblk + (m->nr + nthr - 1) / nthr;
        h[s].blk0 = gnblk[grp[s]];            /* first block within its group's launch */
        if ((long)gnblk[grp[s]] + nb[s] > 0x7fffffffL) rc = awefd__fail(AWEFD_EINVAL, "batch: too many blocks");
        gnblk[grp[s]] += nb[s];
        /* injection buckets for this tile width */
        awefd_inj *inj = NULL;
        int *ibeg = NULL, nbucket = 0;
        long ninj = 0;
        if (!rc) rc = awefd__cart_buckets(m, m->tz, h[s].ntz, tx, &inj, &ibeg, &nbucket, &ninj);
        if (!rc) {
            dfree(&b->d_inj[s]);
            dfree(&b->d_ibeg[s]);
            hipError_t e = hipMalloc((void **)&b->d_inj[s], sizeof(*inj) * (ninj > 0 ? ninj : 1));
            if (e == hipSuccess) e = hipMalloc((void **)&b->d_ibeg[s], sizeof(int) * (nbucket + 1));
            if (e == hipSuccess && ninj > 0)
                e = hipMemcpy(b->d_inj[s], inj, sizeof(*inj) * ninj, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(b->d_ibeg[s], ibeg, sizeof(int) * (nbucket + 1), hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "batch injection upload: %s", hipGetErrorString(e));
            h[s].inj = b->d_inj[s];
            h[s].ibeg = b->d_ibeg[s];
        }
        free(inj);
        free(ibeg);
    }
    for (int g = 0; g < ng && !rc; g++) goff[g + 1] = goff[g] + gnblk[g];
    const long total = rc ? 0 : goff[ng];
    int *bshot = NULL;
    if (!rc && !(bshot = malloc(sizeof(int) * (total > 0 ? total : 1)))) rc = awefd__fail(AWEFD_ENOMEM, "batch block table");
    for (int s = 0; !rc && s < b->nshot; s++)
        for (int k = 0; k < nb[s]; k++) bshot[goff[grp[s]] + h[s].blk0 + k] = s;
    if (!rc) {
        dfree(&b->d_shots);
        dfree(&b->d_bshot);
        hipError_t e = hipMalloc((void **)&b->d_shots, sizeof(*h) * b->nshot);
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_bshot, sizeof(int) * total);
        if (e == hipSuccess) e = hipMemcpy(b->d_shots, h, sizeof(*h) * b->nshot, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_bshot, bshot, sizeof(int) * total, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "batch table upload: %s", hipGetErrorString(e));
    }
    free(bshot);
    free(nb);
    if (!rc) {
        free(b->goff);
        free(b->gnblk);
        free(b->h);
        free(b->grp);
        b->goff = goff;
        b->gnblk = gnblk;
        b->h = h;
        b->grp = grp;
        b->ngroup = ng;
        b->rx = rx;
        b->group_mb = group_mb;
        for (int s = 0; s < b->nshot; s++) b->mgen[s] = b->m[s]->gen;
        b->built = 1;
    } else {
        b->built = 0;
        free(goff);
        free(gnblk);
        free(h);
        free(grp);
    }
    /* captured graphs refer to the tables just replaced */
    b->graph_ok = b->fwd_ok = b->bwd_ok = 0;
    b->t2_ok = 0;
    if (!rc && b->depth >= 2 && rx != 4 && rx != 8) b->depth = 1;
    return rc;
}

/* The tables and the captured graphs hold each model's device pointers
 * (sponge tables, (v dt)^2, wavelet, points, trace buffer).  A setter that
 * ran on a model since they were built may have replaced those buffers (its
 * generation moved): check the shots again and rebuild everything for the
 * current tile before the next launch. */
static int shots_ready(awefd_batch *b)
{
    int rc;
    for (int s = 0; s < b->nshot; s++) {
        if ((rc = shot_ok(b->m[s], b->m[0], s))) return rc;
        if ((rc = awefd__ensure_traces(b->m[s]))) return rc;
    }
    return AWEFD_OK;
}

static int batch_sync(awefd_batch *b)
{
    int stale = !b->built, rc;
    if ((rc = shots_ready(b))) return rc;
    for (int s = 0; s < b->nshot; s++) stale |= b->mgen[s] != b->m[s]->gen;
    if (!stale) return AWEFD_OK;
    const int depth = b->depth;
    if ((rc = batch_build(b, b->rx ? b->rx : 8, b->group_mb))) return rc;
    if (depth >= 2 && (b->rx == 4 || b->rx == 8)) b->depth = depth;
    return AWEFD_OK;
}

static int batch_reset(awefd_batch *b)
{
    for (int s = 0; s < b->nshot; s++) {
        awefd_model *m = b->m[s];
        int e = awefd_launch_fill(b->u0[s], m->nalloc, 0.f, b->stream);
        if (!e) e = awefd_launch_fill(b->u1[s], m->nalloc, 0.f, b->stream);
        if (e) return awefd__fail(AWEFD_EHIP, "batch reset: %s", hipGetErrorString((hipError_t)e));
    }
    return AWEFD_OK;
}

/* nsteps steps of every shot, group after group */
static int batch_steps(awefd_batch *b, int nsteps)
{
    for (int g = 0; g < b->ngroup; g++)
        for (int it = 0; it < nsteps; it++) {
            int e = awefd_launch_cart_batch2(b->d_shots, b->d_bshot + b->goff[g], b->gnblk[g], it, b->nop, b->sse,
                                             b->rx, b->stream);
            if (e) return awefd__fail(AWEFD_EHIP, "batch step %d: %s", it, hipGetErrorString((hipError_t)e));
        }
    return AWEFD_OK;
}

/* ---- Two time steps per pass (depth 2) ----
 * Tiles of (64 - 2 nop) z x (4 rx - 2 nop) x outputs, each computing its
 * first step on 64 x 4 rx nodes; the shot's four field buffers rotate
 * (u0, u1 and the batch's e0, e1): a pass reads the previous and current
 * field and writes the next two into the other pair. */

static int t2_tile(const awefd_batch *b, int *sz, int *sx)
{
    *sz = 64 - 2 * b->nop;
    *sx = AWEFD_BATCH_NWAVE * b->rx - 2 * b->nop;
    return (b->rx == 4 || b->rx == 8) && *sx > 0;
}

/* Tile geometry of a multi-step pass: depth 2 as above; depth 3 blocks of
 * 64 z x AWEFD_T3_ROWS x rows (AWEFD_T3_NWAVE(rx) waves) with outputs 2 nop
 * inside on every side. */
static int tn_tile(const awefd_batch *b, int depth, int *sz, int *sx, int *rows, int *halo)
{
    const int nw = depth == 3 ? AWEFD_T3_NWAVE(b->rx) : AWEFD_BATCH_NWAVE;
    *halo = (depth - 1) * b->nop;
    *rows = nw * b->rx;
    *sz = 64 - 2 * *halo;
    *sx = *rows - 2 * *halo;
    return (depth == 2 || depth == 3) && (b->rx == 4 || b->rx == 8) && *sz > 0 && *sx > 0;
}

/* every (source node, tile) pair whose tile's first-step region holds the
 * node, bucketed per tile */
static int t2_buckets(const awefd_batch *b, const awefd_model *m, int depth, int ntz2, int ntx2, awefd_inj **inj,
                      int **ibeg, long *ninj)
{
    int sz, sx, rows, nop;
    tn_tile(b, depth, &sz, &sx, &rows, &nop);   /* nop: the tile halo */
    long cap = 16L * 4 * (m->ns > 0 ? m->ns : 1), n = 0;
    awefd_inj_key *keys = malloc(sizeof(*keys) * cap);
    int *beg = malloc(sizeof(int) * (ntz2 * ntx2 + 1));
    if (!keys || !beg) {
        free(keys); free(beg);
        return awefd__fail(AWEFD_ENOMEM, "two-step injection buckets");
    }
    for (int s = 0; s < m->ns; s++)
        for (int q = 0; q < 4; q++) {
            const int z = m->h_sj[3 * s] + (q & 1), x = m->h_sj[3 * s + 1] + ((q >> 1) & 1);
            for (int tz = 0; tz < ntz2; tz++) {
                if (z < tz * sz - nop || z >= tz * sz - nop + 64) continue;
                for (int tx = 0; tx < ntx2; tx++) {
                    if (x < tx * sx - nop || x >= tx * sx - nop + rows) continue;
                    if (n == cap) {
                        awefd_inj_key *k2 = realloc(keys, sizeof(*keys) * cap * 2);
                        if (!k2) {
                            free(keys); free(beg);
                            return awefd__fail(AWEFD_ENOMEM, "two-step injection buckets");
                        }
                        keys = k2;
                        cap *= 2;
                    }
                    awefd_inj_key *k = &keys[n++];
                    k->e.z = z;
                    k->e.x = x;
                    k->e.y = 0;
                    k->e.src = s;
                    k->e.w = m->h_sw[8 * s + q];
                    k->e.w2 = 1.f;
                    k->e.w3 = 1.f;
                    k->bucket = tz + ntz2 * tx;
                }
            }
        }
    awefd_inj *e = malloc(sizeof(*e) * (n > 0 ? n : 1));
    if (!e) {
        free(keys); free(beg);
        return awefd__fail(AWEFD_ENOMEM, "two-step injection buckets");
    }
    awefd_sort_inj(keys, n, ntz2 * ntx2, e, beg);
    free(keys);
    *inj = e;
    *ibeg = beg;
    *ninj = n;
    return AWEFD_OK;
}

/* Two-step tables for the current tile and groups. */
/* receivers bucketed by the tile whose outputs hold their first corner */
static int t3_receivers(const awefd_model *m, int sz, int sx, int ntz, int ntx, int **rbeg, awefd_rx3 **rx)
{
    const int nt = ntz * ntx, nr = m->nr;
    int *beg = calloc(nt + 1, sizeof(int)), *fill = malloc(sizeof(int) * (nt > 0 ? nt : 1));
    awefd_rx3 *list = calloc(nr > 0 ? nr : 1, sizeof(*list));
    if (!beg || !fill || !list) {
        free(beg); free(fill); free(list);
        return awefd__fail(AWEFD_ENOMEM, "three-step receiver buckets");
    }
    for (int r = 0; r < nr; r++) beg[m->h_rj[3 * r] / sz + ntz * (m->h_rj[3 * r + 1] / sx) + 1]++;
    for (int t = 0; t < nt; t++) beg[t + 1] += beg[t];
    memcpy(fill, beg, sizeof(int) * nt);
    for (int r = 0; r < nr; r++) {
        const int jz = m->h_rj[3 * r], jx = m->h_rj[3 * r + 1];
        awefd_rx3 *e = &list[fill[jz / sz + ntz * (jx / sx)]++];
        e->r = r;
        e->jz = jz;
        e->jx = jx;
        for (int q = 0; q < 4; q++) e->w[q] = m->h_rw[8 * r + q];
        /* the device sponge tables hold these same floats */
        for (int k = 0; k < 2; k++) {
            e->sz[k] = m->sponge ? m->h_sp[0][jz + k] : 1.f;
            e->sx[k] = m->sponge ? m->h_sp[1][jx + k] : 1.f;
        }
    }
    free(fill);
    *rbeg = beg;
    *rx = list;
    return AWEFD_OK;
}

static int t2_build(awefd_batch *b, int depth)
{
    int sz, sx, rows, halo;
    if (!tn_tile(b, depth, &sz, &sx, &rows, &halo))
        return awefd__fail(AWEFD_EINVAL, "%d steps per pass need tile rx 4 or 8 (and a tile wider than its halo)",
                           depth);
    /* every captured graph that launches the multi-step kernels holds the
     * block table about to be replaced */
    b->graph_ok = b->fwd_ok = b->bwd_ok = 0;
    b->t2_ok = 0;
    const int nthr = (depth == 3 ? AWEFD_T3_NWAVE(b->rx) : AWEFD_BATCH_NWAVE) * 64, ns = b->nshot, ng = b->ngroup;
    if (!b->e0) b->e0 = calloc(ns, sizeof(float *));
    if (!b->e1) b->e1 = calloc(ns, sizeof(float *));
    if (!b->d_inj2) b->d_inj2 = calloc(ns, sizeof(awefd_inj *));
    if (!b->d_ibeg2) b->d_ibeg2 = calloc(ns, sizeof(int *));
    if (!b->d_rbeg3) b->d_rbeg3 = calloc(ns, sizeof(int *));
    if (!b->d_rx3) b->d_rx3 = calloc(ns, sizeof(awefd_rx3 *));
    long *goff2 = calloc(ng + 1, sizeof(long));
    int *gnblk2 = calloc(ng, sizeof(int)), *nb = malloc(sizeof(int) * ns);
    int rc = (!b->e0 || !b->e1 || !b->d_inj2 || !b->d_ibeg2 || !b->d_rbeg3 || !b->d_rx3 || !goff2 || !gnblk2 || !nb)
                 ? awefd__fail(AWEFD_ENOMEM, "two-step tables") : AWEFD_OK;
    for (int s = 0; s < ns && !rc; s++) {
        awefd_model *m = b->m[s];
        awefd_cart_args *h = &b->h[s];
        hipError_t e = hipSuccess;
        if (!b->e0[s]) e = hipMalloc((void **)&b->e0[s], sizeof(float) * m->nalloc);
        if (e == hipSuccess && !b->e1[s]) e = hipMalloc((void **)&b->e1[s], sizeof(float) * m->nalloc);
        if (e != hipSuccess) {
            rc = awefd__fail(AWEFD_ENOMEM, "two-step fields: %s", hipGetErrorString(e));
            break;
        }
        h->F[0] = b->u0[s];
        h->F[1] = b->u1[s];
        h->F[2] = b->e0[s];
        h->F[3] = b->e1[s];
        h->ntz2 = (m->np[0] + sz - 1) / sz;
        h->ntx2 = (m->np[1] + sx - 1) / sx;
        h->nblk2 = h->ntz2 * h->ntx2;
        nb[s] = h->nblk2 + (m->nr + nthr - 1) / nthr;
        h->blk02 = gnblk2[b->grp[s]];
        if ((long)gnblk2[b->grp[s]] + nb[s] > 0x7fffffffL) rc = awefd__fail(AWEFD_EINVAL, "batch: too many blocks");
        gnblk2[b->grp[s]] += nb[s];
        awefd_inj *inj = NULL;
        int *ibeg = NULL;
        long ninj = 0;
        if (!rc) rc = t2_buckets(b, m, depth, h->ntz2, h->ntx2, &inj, &ibeg, &ninj);
        if (!rc) {
            dfree(&b->d_inj2[s]);
            dfree(&b->d_ibeg2[s]);
            e = hipMalloc((void **)&b->d_inj2[s], sizeof(*inj) * (ninj > 0 ? ninj : 1));
            if (e == hipSuccess) e = hipMalloc((void **)&b->d_ibeg2[s], sizeof(int) * (h->nblk2 + 1));
            if (e == hipSuccess && ninj > 0) e = hipMemcpy(b->d_inj2[s], inj, sizeof(*inj) * ninj, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(b->d_ibeg2[s], ibeg, sizeof(int) * (h->nblk2 + 1), hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "two-step injection upload: %s", hipGetErrorString(e));
            h->inj2 = b->d_inj2[s];
            h->ibeg2 = b->d_ibeg2[s];
        }
        free(inj);
        free(ibeg);
        h->rbeg3 = NULL;
        h->rx3 = NULL;
        if (!rc && depth == 3) {
            int *rbeg = NULL;
            awefd_rx3 *rx = NULL;
            rc = t3_receivers(m, sz, sx, h->ntz2, h->ntx2, &rbeg, &rx);
            if (!rc) {
                dfree(&b->d_rbeg3[s]);
                dfree(&b->d_rx3[s]);
                const size_t nbeg = (size_t)h->ntz2 * h->ntx2 + 1, nrl = m->nr > 0 ? m->nr : 1;
                e = hipMalloc((void **)&b->d_rbeg3[s], sizeof(int) * nbeg);
                if (e == hipSuccess) e = hipMalloc((void **)&b->d_rx3[s], sizeof(awefd_rx3) * nrl);
                if (e == hipSuccess) e = hipMemcpy(b->d_rbeg3[s], rbeg, sizeof(int) * nbeg, hipMemcpyHostToDevice);
                if (e == hipSuccess) e = hipMemcpy(b->d_rx3[s], rx, sizeof(awefd_rx3) * nrl, hipMemcpyHostToDevice);
                if (e != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "three-step receiver upload: %s", hipGetErrorString(e));
                h->rbeg3 = b->d_rbeg3[s];
                h->rx3 = b->d_rx3[s];
            }
            free(rbeg);
            free(rx);
        }
    }
    for (int g = 0; g < ng && !rc; g++) goff2[g + 1] = goff2[g] + gnblk2[g];
    int *bshot = rc ? NULL : malloc(sizeof(int) * (goff2[ng] > 0 ? goff2[ng] : 1));
    if (!rc && !bshot) rc = awefd__fail(AWEFD_ENOMEM, "two-step block table");
    for (int s = 0; s < ns && !rc; s++)
        for (int k = 0; k < nb[s]; k++) bshot[goff2[b->grp[s]] + b->h[s].blk02 + k] = s;
    if (!rc) {
        dfree(&b->d_bshot2);
        hipError_t e = hipMalloc((void **)&b->d_bshot2, sizeof(int) * goff2[ng]);
        if (e == hipSuccess) e = hipMemcpy(b->d_bshot2, bshot, sizeof(int) * goff2[ng], hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_shots, b->h, sizeof(*b->h) * ns, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "two-step table upload: %s", hipGetErrorString(e));
    }
    free(bshot);
    free(nb);
    if (!rc) {
        free(b->goff2);
        free(b->gnblk2);
        b->goff2 = goff2;
        b->gnblk2 = gnblk2;
        b->t2_ok = 1;
        b->t2_depth = depth;
    } else {
        free(goff2);
        free(gnblk2);
    }
    return rc;
}

/* nsteps steps of group g in passes of two steps, then one or two single
 * steps (each pass extracts the rows of its two input fields, so ending on
 * single steps extracts every row); final roles of the previous and current
 * field in *fp, *fc */
/* Passes of `depth` steps while at least one more step follows, then
 * single steps: a multi-step pass at `it` extracts rows it-1 and it (and,
 * with three steps, it+1 in-block), so every row is extracted.  With launch
 * set the passes of group g are issued; the final roles of the previous and
 * current field go to *fp, *fc. */
static int tn_schedule(awefd_batch *b, int depth, int g, int nsteps, int *fp, int *fc, int launch)
{
    int iP = 0, iC = 1, it = 0;
    const int nmulti = (nsteps - 1) / depth;
    for (int k = 0; k < nmulti; k++, it += depth) {
        int d[2], nd = 0;
        for (int i = 0; i < 4; i++)
            if (i != iP && i != iC) d[nd++] = i;
        const int roles = iP | iC << 2 | d[0] << 4 | d[1] << 6;
        if (launch) {
            int e = depth == 3 ? awefd_launch_cart_batch_t3(b->d_shots, b->d_bshot2 + b->goff2[g], b->gnblk2[g], it,
                                                            roles, 3, b->nop, b->sse, b->rx, b->stream)
                               : awefd_launch_cart_batch_t2(b->d_shots, b->d_bshot2 + b->goff2[g], b->gnblk2[g], it,
                                                            roles, 1, b->nop, b->sse, b->rx, b->stream);
            if (e) return awefd__fail(AWEFD_EHIP, "%d-step pass %d: %s", depth, it, hipGetErrorString((hipError_t)e));
        }
        iP = d[0];
        iC = d[1];
    }
    for (; it < nsteps; it++) {
        int d0 = 0;
        while (d0 == iP || d0 == iC) d0++;
        const int roles = iP | iC << 2 | d0 << 4 | d0 << 6;
        if (launch) {
            int e = depth == 3 ? awefd_launch_cart_batch_t3(b->d_shots, b->d_bshot2 + b->goff2[g], b->gnblk2[g], it,
                                                            roles, 1, b->nop, b->sse, b->rx, b->stream)
                               : awefd_launch_cart_batch_t2(b->d_shots, b->d_bshot2 + b->goff2[g], b->gnblk2[g], it,
                                                            roles, 0, b->nop, b->sse, b->rx, b->stream);
            if (e) return awefd__fail(AWEFD_EHIP, "one-step pass %d: %s", it, hipGetErrorString((hipError_t)e));
        }
        iP = iC;
        iC = d0;
    }
    *fp = iP;
    *fc = iC;
    return AWEFD_OK;
}

/* nsteps steps of group g in multi-step passes (the depth the tables were
 * built for); final roles of the previous and current field in *fp, *fc */
static int t2_steps(awefd_batch *b, int g, int nsteps, int *fp, int *fc)
{
    return tn_schedule(b, b->t2_depth, g, nsteps, fp, fc, 1);
}

/* After a multi-step run the last two fields (final roles fp, fc) may sit in
 * the batch's extra pair: the copies into the models' own buffers */
static int final_jobs(awefd_batch *b, int fp, int fc, final_copies *f)
{
    int tp = fp, tc = fc;
    if (tp >= 2) tp = tc == 0 ? 1 : 0;
    if (tc >= 2) tc = tp == 0 ? 1 : 0;
    f->fin_p = tp;
    f->fin_c = tc;
    const int per = (fp >= 2) + (fc >= 2);
    f->njob = per * b->nshot;
    dfree(&f->d_jobs);
    dfree(&f->d_cbeg);
    if (!per) return AWEFD_OK;
    awefd_copy_job *jobs = malloc(sizeof(*jobs) * f->njob);
    long *cbeg = malloc(sizeof(long) * f->njob);
    if (!jobs || !cbeg) {
        free(jobs); free(cbeg);
        return awefd__fail(AWEFD_ENOMEM, "final copies");
    }
    long c = 0;
    int j = 0;
    for (int s = 0; s < b->nshot; s++) {
        const awefd_cart_args *h = &b->h[s];
        const long n = b->m[s]->nalloc;
        if (fp >= 2) {
            jobs[j] = (awefd_copy_job){h->F[fp], h->F[tp], n};
            cbeg[j++] = c;
            c += (n + 1023) / 1024;
        }
        if (fc >= 2) {
            jobs[j] = (awefd_copy_job){h->F[fc], h->F[tc], n};
            cbeg[j++] = c;
            c += (n + 1023) / 1024;
        }
    }
    f->nch = c;
    hipError_t e = hipMalloc((void **)&f->d_jobs, sizeof(*jobs) * f->njob);
    if (e == hipSuccess) e = hipMalloc((void **)&f->d_cbeg, sizeof(long) * f->njob);
    if (e == hipSuccess) e = hipMemcpy(f->d_jobs, jobs, sizeof(*jobs) * f->njob, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(f->d_cbeg, cbeg, sizeof(long) * f->njob, hipMemcpyHostToDevice);
    free(jobs);
    free(cbeg);
    if (e != hipSuccess) return awefd__fail(AWEFD_EHIP, "final copies: %s", hipGetErrorString(e));
    return AWEFD_OK;
}

int awefd_batch_set_depth(awefd_batch *b, int depth)
{
    if (!b) return awefd__fail(AWEFD_EINVAL, "NULL batch");
    if (depth < 1 || depth > 3) return awefd__fail(AWEFD_EINVAL, "batch depth %d (1, 2 or 3)", depth);
    HIPCHK(hipSetDevice(b->dev));
    int rc;
    if ((rc = batch_sync(b))) return rc;
    int sz, sx, rows, halo;
    if (depth >= 2 && !tn_tile(b, depth, &sz, &sx, &rows, &halo))
        return awefd__fail(AWEFD_EINVAL, "%d steps per pass need tile rx 4 or 8 (tile is %d)", depth, b->rx);
    b->depth = depth;
    b->graph_ok = 0;
    return AWEFD_OK;
}

int awefd_batch_set_tile(awefd_batch *b, int rx, int group_mb)
{
    if (!b) return awefd__fail(AWEFD_EINVAL, "NULL batch");
    int ok = 0;
    for (int i = 0; i < 4; i++) ok |= RX_CAND[i] == rx;
    if (!ok) return awefd__fail(AWEFD_EINVAL, "batch tile: rx=%d (2, 4, 8 or 16)", rx);
    if (group_mb < 0) return awefd__fail(AWEFD_EINVAL, "batch group budget %d MiB", group_mb);
    HIPCHK(hipSetDevice(b->dev));
    int rc = shots_ready(b);
    return rc ? rc : batch_build(b, rx, group_mb);
}

int awefd_batch_tune(awefd_batch *b, int steps, int *rx_out, int *group_out, double *ms_out)
{
    if (!b) return awefd__fail(AWEFD_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->dev));
    if (steps < 1) steps = 1;
    if (steps > b->nt) steps = b->nt;
    int rc = shots_ready(b);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int bi = -1, bg = -1;
    double bestms = 0;
    for (int i = 0; i < 4 && !rc; i++)
        for (int g = 0; g < NGC && !rc; g++) {
            if ((rc = batch_build(b, RX_CAND[i], GROUP_CAND[g]))) break;
            if ((rc = batch_reset(b))) break;
            if ((rc = batch_steps(b, 1))) break;      /* warm the tables */
            if ((rc = (int)hipEventRecord(e0, b->stream))) break;
            if ((rc = batch_steps(b, steps))) break;
            if ((rc = (int)hipEventRecord(e1, b->stream))) break;
            if ((rc = (int)hipEventSynchronize(e1))) break;
            float ms = 0.f;
            if ((rc = (int)hipEventElapsedTime(&ms, e0, e1))) break;
            b->tune_ms[i][g] = ms / steps;
            if (bi < 0 || b->tune_ms[i][g] < bestms) {
                bi = i;
                bg = g;
                bestms = b->tune_ms[i][g];
            }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc > 0) rc = awefd__fail(AWEFD_EHIP, "batch tuning: %s", hipGetErrorString((hipError_t)rc));
    if (rc) return rc;
    if (rx_out) *rx_out = RX_CAND[bi];
    if (group_out) *group_out = GROUP_CAND[bg];
    if (ms_out)
        for (int i = 0; i < 4; i++)
            for (int g = 0; g < NGC; g++) ms_out[i * NGC + g] = b->tune_ms[i][g];
    return batch_build(b, RX_CAND[bi], GROUP_CAND[bg]);
}

/* ms per step of one (depth, tile, group) choice on the batch */
static int time_choice(awefd_batch *b, int depth, int rx, int group_mb, int steps, hipEvent_t e0, hipEvent_t e1,
                       double *ms_out)
{
    int rc = batch_build(b, rx, group_mb), fp, fc;
    b->depth = depth;
    /* whole passes plus the closing single step, as a run ends */
    if (depth >= 2) {
        const int k = (steps - 1) / depth > 0 ? (steps - 1) / depth : 1;
        steps = depth * k + 1 <= b->nt ? depth * k + 1 : b->nt;
    }
    if (!rc && depth >= 2) rc = t2_build(b, depth);
    if (!rc) rc = batch_reset(b);
    /* warm the tables */
    for (int g = 0; g < b->ngroup && !rc && depth >= 2; g++) rc = t2_steps(b, g, 1, &fp, &fc);
    if (!rc && depth == 1) rc = batch_steps(b, 1);
    if (!rc) rc = (int)hipEventRecord(e0, b->stream);
    for (int g = 0; g < b->ngroup && !rc && depth >= 2; g++) rc = t2_steps(b, g, steps, &fp, &fc);
    if (!rc && depth == 1) rc = batch_steps(b, steps);
    if (!rc) rc = (int)hipEventRecord(e1, b->stream);
    if (!rc) rc = (int)hipEventSynchronize(e1);
    float ms = 0.f;
    if (!rc) rc = (int)hipEventElapsedTime(&ms, e0, e1);
    if (rc > 0) rc = awefd__fail(AWEFD_EHIP, "batch tuning: %s", hipGetErrorString((hipError_t)rc));
    *ms_out = ms / steps;
    return rc;
}

int awefd_batch_tune_depth(awefd_batch *b, int steps, int *rx_out, int *group_out, int *depth_out, double *ms_out)
{
    if (!b) return awefd__fail(AWEFD_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->dev));
    if (steps < 2) steps = 2;
    if (steps > b->nt) steps = b->nt;
    int rc = shots_ready(b);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int bd = 0, bi = -1, bg = -1;
    double best = 0;
    memset(b->tune_ms, 0, sizeof b->tune_ms);
    memset(b->tune2_ms, 0, sizeof b->tune2_ms);
    memset(b->tune3_ms, 0, sizeof b->tune3_ms);
    for (int d = 1; d <= 3 && !rc; d++)
        for (int i = 0; i < 4 && !rc; i++) {
            int sz, sx, rows, halo;
            b->rx = RX_CAND[i];
            if (d >= 2 && !tn_tile(b, d, &sz, &sx, &rows, &halo)) continue;
            for (int g = 0; g < NGC && !rc; g++) {
                double ms;
                if ((rc = time_choice(b, d, RX_CAND[i], GROUP_CAND[g], steps, e0, e1, &ms))) break;
                (d == 1 ? b->tune_ms : d == 2 ? b->tune2_ms : b->tune3_ms)[i][g] = ms;
                if (bi < 0 || ms < best) {
                    bd = d;
                    bi = i;
                    bg = g;
                    best = ms;
                }
            }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    if (rx_out) *rx_out = RX_CAND[bi];
    if (group_out) *group_out = GROUP_CAND[bg];
    if (depth_out) *depth_out = bd;
    if (ms_out)
        for (int i = 0; i < 4; i++)
            for (int g = 0; g < NGC; g++) {
                ms_out[i * NGC + g] = b->tune_ms[i][g];
                ms_out[4 * NGC + i * NGC + g] = b->tune2_ms[i][g];
                ms_out[8 * NGC + i * NGC + g] = b->tune3_ms[i][g];
            }
    if ((rc = batch_build(b, RX_CAND[bi], GROUP_CAND[bg]))) return rc;
    b->depth = bd;
    return AWEFD_OK;
}

/* The time loop of every shot as one graph: field resets, then one batched
 * launch per step of each group. */
static int batch_capture(awefd_batch *b)
{
    if (b->gexec && b->graph_ok) return AWEFD_OK;
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = NULL;
    b->graph = NULL;
    int fp = 0, fc = 1, rc;
    if (b->depth >= 2) {
        if ((!b->t2_ok || b->t2_depth != b->depth) && (rc = t2_build(b, b->depth))) return rc;
        /* the final roles are those of any group's schedule */
        int i = 0, c = 1;
        tn_schedule(b, b->depth, 0, b->nt, &i, &c, 0);
        if ((rc = final_jobs(b, i, c, &b->fin))) return rc;
    }
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = batch_reset(b);
    if (b->depth >= 2) {
        for (int g = 0; g < b->ngroup && !rc; g++) rc = t2_steps(b, g, b->nt, &fp, &fc);
        if (!rc && b->fin.njob) {
            int e = awefd_launch_copy_jobs(b->fin.d_jobs, b->fin.d_cbeg, b->fin.njob, b->fin.nch, b->stream);
            if (e) rc = awefd__fail(AWEFD_EHIP, "final copies: %s", hipGetErrorString((hipError_t)e));
        }
    } else if (!rc) {
        rc = batch_steps(b, b->nt);
    }
    hipGraph_t g = NULL;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return awefd__fail(AWEFD_EHIP, "batch capture: %s", hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->gexec, g, NULL, NULL, 0));
    b->graph_ok = 1;
    return AWEFD_OK;
}

int awefd_batch_run(awefd_batch *b)
{
    if (!b) return awefd__fail(AWEFD_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->dev));
    int rc;
    if ((rc = batch_sync(b))) return rc;
    if ((rc = batch_capture(b))) return rc;
    HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    /* leave each model as a single run leaves it: step it writes buffer
     * (it even ? u0 : u1), so the reference's current field after the last
     * step (u0 after its swap) is buffer u1 for odd nt, u0 for even nt */
    for (int s = 0; s < b->nshot; s++) {
        awefd_model *m = b->m[s];
        if (b->depth >= 2) {
            m->dP = b->fin.fin_p ? b->u1[s] : b->u0[s];
            m->dC = b->fin.fin_c ? b->u1[s] : b->u0[s];
        } else {
            m->dP = b->nt % 2 ? b->u1[s] : b->u0[s];
            m->dC = b->nt % 2 ? b->u0[s] : b->u1[s];
        }
        m->steps_done = b->nt;
    }
    return AWEFD_OK;
}

/* ---- Gradient: checkpointed forward and backward pass of every shot ---- */

static long chunks(long n)
{
    return (n + 1023) / 1024;
}

/* Steps per pass of the checkpointed forward run and of the segment
 * recompute: three where the tile allows it (rx 4 or 8), else one.
 * AWEFD_CKPT_DEPTH=1 keeps one step per launch. */
static int ckpt_depth(const awefd_batch *b)
{
    int sz, sx, rows, halo;
    const char *e = getenv("AWEFD_CKPT_DEPTH");
    if (e && atoi(e) == 1) return 1;
    return tn_tile(b, 3, &sz, &sx, &rows, &halo) ? 3 : 1;
}

/* Adjoint steps per pass of the batched backward run: two (carrying dJ/dV
 * across both), or one with AWEFD_ADJ_DEPTH=1. */
static int adj_depth(void)
{
    const char *e = getenv("AWEFD_ADJ_DEPTH");
    return e && atoi(e) == 1 ? 1 : 2;
}

/* Launch the checkpoint copies of group g at checkpoint k */
static int ckpt_copies(awefd_batch *b, int g, int k)
{
    const long t = (long)g * b->nck + k;
    const int nj = (int)((t + 1 < (long)b->ngroup * b->nck ? b->joff[t + 1] : 2L * b->nshot * b->nck) - b->joff[t]);
    if (nj <= 0) return 0;
    return awefd_launch_copy_jobs(b->d_jobs + b->joff[t], b->d_cbeg + b->joff[t], nj, b->jnch[t], b->stream);
}

/* The checkpointed forward run of group g in passes of three steps: in each
 * segment [kK, min((k+1)K, nt)) the checkpoint copies of k, then three-step
 * passes while a pass ends inside the segment and at least one step of the
 * run follows it (the launch after a pass extracts its last field's trace
 * row), then single steps.  The fields rotate through the shot's four
 * buffers as in tn_schedule; the roles of the previous and current field at
 * each checkpoint go to ckp[k], ckc[k] (when not NULL), the final roles to
 * *fp, *fc.  With launch set the copies and passes are issued. */
static int ckpt_schedule3(awefd_batch *b, int g, int *ckp, int *ckc, int *fp, int *fc, int launch)
{
    int iP = 0, iC = 1;
    for (int k = 0; k < b->nck; k++) {
        const int n0 = k * b->K, n1 = n0 + b->K < b->nt ? n0 + b->K : b->nt;
        if (ckp) {
            ckp[k] = iP;
            ckc[k] = iC;
        }
        int e = launch ? ckpt_copies(b, g, k) : 0;
        if (e) return awefd__fail(AWEFD_EHIP, "checkpoint %d copies: %s", k, hipGetErrorString((hipError_t)e));
        int it = n0;
        for (; it + 3 <= n1 && it + 3 < b->nt; it += 3) {
            int d[2], nd = 0;
            for (int i = 0; i < 4; i++)
                if (i != iP && i != iC) d[nd++] = i;
            const int roles = iP | iC << 2 | d[0] << 4 | d[1] << 6;
            if (launch && (e = awefd_launch_cart_batch_t3(b->d_shots, b->d_bshot2 + b->goff2[g], b->gnblk2[g], it,
                                                          roles, 3, b->nop, b->sse, b->rx, b->stream)))
                return awefd__fail(AWEFD_EHIP, "checkpointed 3-step pass %d: %s", it, hipGetErrorString((hipError_t)e));
            iP = d[0];
            iC = d[1];
        }
        for (; it < n1; it++) {
            int d0 = 0;
            while (d0 == iP || d0 == iC) d0++;
            const int roles = iP | iC << 2 | d0 << 4 | d0 << 6;
            if (launch && (e = awefd_launch_cart_batch_t3(b->d_shots, b->d_bshot2 + b->goff2[g], b->gnblk2[g], it,
                                                          roles, 1, b->nop, b->sse, b->rx, b->stream)))
                return awefd__fail(AWEFD_EHIP, "checkpointed step %d: %s", it, hipGetErrorString((hipError_t)e));
            iP = iC;
            iC = d0;
        }
    }
    *fp = iP;
    *fc = iC;
    return AWEFD_OK;
}

/* Checkpoint copy tables: for group g and checkpoint k (step kK), every shot
 * of g copies its current and previous field into its checkpoint pair k.
 * One step per launch alternates the two buffers by step parity (as
 * k_cart_batch2 does); three steps per pass rotate four buffers, whose roles
 * at each checkpoint the schedule gives (the same for every group). */
static int ckpt_tables(awefd_batch *b, int depth)
{
    const int ng = b->ngroup, nck = b->nck;
    long njob = 2L * b->nshot * nck;
    awefd_copy_job *jobs = malloc(sizeof(*jobs) * njob);
    long *cbeg = malloc(sizeof(long) * njob);
    long *joff = malloc(sizeof(long) * ng * nck), *jnch = malloc(sizeof(long) * ng * nck);
    int *ckp = malloc(sizeof(int) * nck), *ckc = malloc(sizeof(int) * nck);
    if (!jobs || !cbeg || !joff || !jnch || !ckp || !ckc) {
        free(jobs); free(cbeg); free(joff); free(jnch); free(ckp); free(ckc);
        return awefd__fail(AWEFD_ENOMEM, "checkpoint tables");
    }
    int fp = 0, fc = 1;
    if (depth == 3) ckpt_schedule3(b, 0, ckp, ckc, &fp, &fc, 0);
    long j = 0;
    for (int g = 0; g < ng; g++)
        for (int k = 0; k < nck; k++) {
            const int it = k * b->K;
            joff[g * nck + k] = j;
            long c = 0;
            for (int s = 0; s < b->nshot; s++) {
                if (b->grp[s] != g) continue;
                awefd_model *m = b->m[s];
                float *ck = awefd__adj_ckpt(m, k);
                const float *cur = it & 1 ? b->u0[s] : b->u1[s];
                const float *prev = it & 1 ? b->u1[s] : b->u0[s];
                if (depth == 3) {
                    cur = b->h[s].F[ckc[k]];
                    prev = b->h[s].F[ckp[k]];
                }
                jobs[j] = (awefd_copy_job){cur, ck, m->nalloc};
                cbeg[j++] = c;
                c += chunks(m->nalloc);
                jobs[j] = (awefd_copy_job){prev, ck + m->nalloc, m->nalloc};
                cbeg[j++] = c;
                c += chunks(m->nalloc);
            }
            jnch[g * nck + k] = c;
        }
    free(ckp);
    free(ckc);
    dfree(&b->d_jobs);
    dfree(&b->d_cbeg);
    hipError_t e = hipMalloc((void **)&b->d_jobs, sizeof(*jobs) * njob);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_cbeg, sizeof(long) * njob);
    if (e == hipSuccess) e = hipMemcpy(b->d_jobs, jobs, sizeof(*jobs) * njob, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_cbeg, cbeg, sizeof(long) * njob, hipMemcpyHostToDevice);
    free(jobs);
    free(cbeg);
    free(b->joff);
    free(b->jnch);
    b->joff = joff;
    b->jnch = jnch;
    if (e != hipSuccess) return awefd__fail(AWEFD_EHIP, "checkpoint tables: %s", hipGetErrorString(e));
    if (depth == 3) return final_jobs(b, fp, fc, &b->kfin);
    return AWEFD_OK;
}

static int capture_fwd(awefd_batch *b, int depth)
{
    if (b->xfwd && b->fwd_ok && b->fwd_depth == depth) return AWEFD_OK;
    if (b->xfwd) hipGraphExecDestroy(b->xfwd);
    if (b->gfwd) hipGraphDestroy(b->gfwd);
    b->xfwd = NULL;
    b->gfwd = NULL;
    int rc = AWEFD_OK;
    if (depth == 3 && (!b->t2_ok || b->t2_depth != 3)) rc = t2_build(b, 3);
    if (!rc) rc = ckpt_tables(b, depth);
    if (rc) return rc;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = batch_reset(b);
    for (int g = 0; g < b->ngroup && !rc && depth == 3; g++) {
        int fp, fc;
        rc = ckpt_schedule3(b, g, NULL, NULL, &fp, &fc, 1);
    }
    if (!rc && depth == 3 && b->kfin.njob) {
        int e = awefd_launch_copy_jobs(b->kfin.d_jobs, b->kfin.d_cbeg, b->kfin.njob, b->kfin.nch, b->stream);
        if (e) rc = awefd__fail(AWEFD_EHIP, "checkpointed run, final copies: %s", hipGetErrorString((hipError_t)e));
    }
    for (int g = 0; g < b->ngroup && !rc && depth == 1; g++)
        for (int it = 0; it < b->nt && !rc; it++) {
            int e = it % b->K == 0 ? ckpt_copies(b, g, it / b->K) : 0;
            if (!e)
                e = awefd_launch_cart_batch2(b->d_shots, b->d_bshot + b->goff[g], b->gnblk[g], it, b->nop, b->sse,
                                             b->rx, b->stream);
            if (e) rc = awefd__fail(AWEFD_EHIP, "checkpointed batch step %d: %s", it, hipGetErrorString((hipError_t)e));
        }
    hipGraph_t gr = NULL;
    hipError_t e = hipStreamEndCapture(b->stream, &gr);
    if (rc) {
        if (gr) hipGraphDestroy(gr);
        return rc;
    }
    if (e != hipSuccess) return awefd__fail(AWEFD_EHIP, "checkpointed batch capture: %s", hipGetErrorString(e));
    b->gfwd = gr;
    HIPCHK(hipGraphInstantiate(&b->xfwd, gr, NULL, NULL, 0));
    b->fwd_ok = 1;
    b->fwd_depth = depth;
    return AWEFD_OK;
}

int awefd_batch_run_ckpt(awefd_batch *b, int ckpt)
{
    if (!b) return awefd__fail(AWEFD_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->dev));
    int rc;
    for (int s = 0; s < b->nshot; s++)
        if ((rc = awefd__adj_ckpt_setup(b->m[s], ckpt))) return rc;
    const int K = awefd__adj_K(b->m[0]);
    if ((rc = batch_sync(b))) return rc;
    if (K != b->K) {
        b->K = K;
        b->nck = (b->nt + K - 1) / K;
        b->fwd_ok = b->bwd_ok = 0;
    }
    /* the graph holds every shot's checkpoint base: recapture if one moved */
    if (!b->fwd_ck && !(b->fwd_ck = calloc(b->nshot, sizeof(float *))))
        return awefd__fail(AWEFD_ENOMEM, "batch checkpoint bases");
    for (int s = 0; s < b->nshot; s++)
        if (b->fwd_ck[s] != awefd__adj_ckpt(b->m[s], 0)) b->fwd_ok = 0;
    const int depth = ckpt_depth(b);
    if ((rc = capture_fwd(b, depth))) return rc;
    for (int s = 0; s < b->nshot; s++) b->fwd_ck[s] = awefd__adj_ckpt(b->m[s], 0);
    HIPCHK(hipGraphLaunch(b->xfwd, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int s = 0; s < b->nshot; s++) {
        awefd_model *m = b->m[s];
        if (depth == 3) {
            m->dP = b->kfin.fin_p ? b->u1[s] : b->u0[s];
            m->dC = b->kfin.fin_c ? b->u1[s] : b->u0[s];
        } else {
            m->dP = b->nt % 2 ? b->u1[s] : b->u0[s];
            m->dC = b->nt % 2 ? b->u0[s] : b->u1[s];
        }
        m->steps_done = b->nt;
        awefd__adj_mark(m);
    }
    return AWEFD_OK;
}

/* Adjoint tables: arguments per shot and the block -> shot map of each
 * group's launch. */
static int adj_tables(awefd_batch *b, const float *const *d_res)
{
    const int ns = b->nshot, ng = b->ngroup;
    awefd_adj2d_args *ha = malloc(sizeof(*ha) * ns);
    long *agoff = calloc(ng + 1, sizeof(long)), *agoff2 = calloc(ng + 1, sizeof(long));
    int *agnblk = calloc(ng, sizeof(int)), *agnblk2 = calloc(ng, sizeof(int));
    int rc = (!ha || !agoff || !agnblk || !agoff2 || !agnblk2) ? awefd__fail(AWEFD_ENOMEM, "adjoint tables")
                                                               : AWEFD_OK;
    for (int s = 0; s < ns && !rc; s++) {
        awefd__adj_args_batch(b->m[s], d_res[s], &ha[s]);
        const int g = b->grp[s];
        ha[s].blk0 = agnblk[g];
        ha[s].blk02 = agnblk2[g];
        if ((long)agnblk[g] + ha[s].nblk > 0x7fffffffL || (long)agnblk2[g] + ha[s].nblk2 > 0x7fffffffL)
            rc = awefd__fail(AWEFD_EINVAL, "adjoint: too many blocks");
        agnblk[g] += ha[s].nblk;
        agnblk2[g] += ha[s].nblk2;
    }
    for (int g = 0; g < ng && !rc; g++) {
        agoff[g + 1] = agoff[g] + agnblk[g];
        agoff2[g + 1] = agoff2[g] + agnblk2[g];
    }
    int *bshot = rc ? NULL : malloc(sizeof(int) * (agoff[ng] > 0 ? agoff[ng] : 1));
    int *bshot2 = rc ? NULL : malloc(sizeof(int) * (agoff2[ng] > 0 ? agoff2[ng] : 1));
    if (!rc && (!bshot || !bshot2)) rc = awefd__fail(AWEFD_ENOMEM, "adjoint block table");
    for (int s = 0; s < ns && !rc; s++) {
        for (int k = 0; k < ha[s].nblk; k++) bshot[agoff[b->grp[s]] + ha[s].blk0 + k] = s;
        for (int k = 0; k < ha[s].nblk2; k++) bshot2[agoff2[b->grp[s]] + ha[s].blk02 + k] = s;
    }
    /* zero fills of every shot */
    awefd_copy_job *z = rc ? NULL : malloc(sizeof(*z) * 3 * ns);
    long *zc = rc ? NULL : malloc(sizeof(long) * 3 * ns);
    if (!rc && (!z || !zc)) rc = awefd__fail(AWEFD_ENOMEM, "adjoint fills");
    long nch = 0;
    for (int s = 0; s < ns && !rc; s++) {
        awefd__adj_zero_jobs(b->m[s], z + 3 * s);
        for (int i = 0; i < 3; i++) {
            zc[3 * s + i] = nch;
            nch += chunks(z[3 * s + i].n);
        }
    }
    if (!rc) {
        hipError_t e = hipSuccess;
        /* allocated once per block-table size: a captured graph holds these
         * addresses */
        if (b->abshot_cap < agoff[ng]) {
            dfree(&b->d_abshot);
            b->abshot_cap = 0;
            b->bwd_ok = 0;
            e = hipMalloc((void **)&b->d_abshot, sizeof(int) * (agoff[ng] > 0 ? agoff[ng] : 1));
            if (e == hipSuccess) b->abshot_cap = agoff[ng];
        }
        if (e == hipSuccess && b->abshot2_cap < agoff2[ng]) {
            dfree(&b->d_abshot2);
            b->abshot2_cap = 0;
            b->bwd_ok = 0;
            e = hipMalloc((void **)&b->d_abshot2, sizeof(int) * (agoff2[ng] > 0 ? agoff2[ng] : 1));
            if (e == hipSuccess) b->abshot2_cap = agoff2[ng];
        }
        if (e == hipSuccess && !b->d_adj) {
            b->bwd_ok = 0;
            e = hipMalloc((void **)&b->d_adj, sizeof(*ha) * ns);
            if (e == hipSuccess) e = hipMalloc((void **)&b->d_zjobs, sizeof(*z) * 3 * ns);
            if (e == hipSuccess) e = hipMalloc((void **)&b->d_zcbeg, sizeof(long) * 3 * ns);
        }
        if (e == hipSuccess) e = hipMemcpy(b->d_adj, ha, sizeof(*ha) * ns, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_abshot, bshot, sizeof(int) * agoff[ng], hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_abshot2, bshot2, sizeof(int) * agoff2[ng], hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_zjobs, z, sizeof(*z) * 3 * ns, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_zcbeg, zc, sizeof(long) * 3 * ns, hipMemcpyHostToDevice);
        /* the recompute reads each shot's checkpoint and segment bases */
        for (int s = 0; s < ns; s++) awefd__adj_cart_ckpt(b->m[s], &b->h[s]);
        if (e == hipSuccess) e = hipMemcpy(b->d_shots, b->h, sizeof(*b->h) * ns, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = awefd__fail(AWEFD_EHIP, "adjoint tables: %s", hipGetErrorString(e));
    }
    free(bshot);
    free(bshot2);
    free(z);
    free(zc);
    if (!rc) {
        free(b->ha);
        free(b->agoff);
        free(b->agnblk);
        free(b->agoff2);
        free(b->agnblk2);
        b->ha = ha;
        b->agoff = agoff;
        b->agnblk = agnblk;
        b->agoff2 = agoff2;
        b->agnblk2 = agnblk2;
        b->znch = nch;
    } else {
        free(ha);
        free(agoff);
        free(agnblk);
        free(agoff2);
        free(agnblk2);
    }
    return rc;
}

/* The backward pass of every shot as one graph: zero fills, then per group
 * the segments last to first, each recomputed from its checkpoint (three
 * steps per pass at depth 3, while the pass's last field is one the segment
 * keeps, then single steps) and walked back by the adjoint steps (in pairs
 * n, n-1 inside the segment at adepth 2, a single step for an odd one
 * left); last the shots' dJ/dV summed in shot order into d_grad. */
static int capture_bwd(awefd_batch *b, float *d_grad, int depth, int adepth)
{
    if (b->xbwd && b->bwd_ok && b->bwd_gout == d_grad && b->bwd_depth == depth && b->bwd_adepth == adepth)
        return AWEFD_OK;
    if (b->xbwd) hipGraphExecDestroy(b->xbwd);
    if (b->gbwd) hipGraphDestroy(b->gbwd);
    b->xbwd = NULL;
    b->gbwd = NULL;
    int rc = AWEFD_OK;
    if (depth == 3 && (!b->t2_ok || b->t2_depth != 3) && (rc = t2_build(b, 3))) return rc;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int e = awefd_launch_copy_jobs(b->d_zjobs, b->d_zcbeg, 3 * b->nshot, b->znch, b->stream);
    for (int g = 0; g < b->ngroup && !e; g++)
        for (int k = b->nck - 1; k >= 0 && !e; k--) {
            const int n0 = k * b->K, n1 = n0 + b->K < b->nt ? n0 + b->K : b->nt;
            int n = n0;
            for (; depth == 3 && n + 3 <= n1 - 1 && !e; n += 3)
                e = awefd_launch_cart_batch_t3rc(b->d_shots, b->d_bshot2 + b->goff2[g], b->gnblk2[g], n, b->nop,
                                                 b->sse, b->rx, b->stream);
            for (; n + 1 < n1 && !e; n++)
                e = awefd_launch_cart_batch_rc(b->d_shots, b->d_bshot + b->goff[g], b->gnblk[g], n, b->nop, b->sse,
                                               b->rx, b->stream);
            int na = n1 - 1;
            for (; adepth == 2 && na - 1 >= n0 && !e; na -= 2)
                e = awefd_launch_adj2d_batch2(b->d_adj, b->d_abshot2 + b->agoff2[g], b->agnblk2[g], na, b->nop,
                                              b->stream);
            for (; na >= n0 && !e; na--)
                e = awefd_launch_adj2d_batch(b->d_adj, b->d_abshot + b->agoff[g], b->agnblk[g], na, b->nop,
                                             b->stream);
        }
    if (!e) e = awefd_launch_adj2d_chain(&b->ha[0], b->d_adj, b->nshot, d_grad, b->stream);
    if (e) rc = awefd__fail(AWEFD_EHIP, "batched backward pass: %s", hipGetErrorString((hipError_t)e));
    hipGraph_t gr = NULL;
    hipError_t he = hipStreamEndCapture(b->stream, &gr);
    if (rc) {
        if (gr) hipGraphDestroy(gr);
        return rc;
    }
    if (he != hipSuccess) return awefd__fail(AWEFD_EHIP, "batched backward capture: %s", hipGetErrorString(he));
    b->gbwd = gr;
    HIPCHK(hipGraphInstantiate(&b->xbwd, gr, NULL, NULL, 0));
    b->bwd_ok = 1;
    b->bwd_gout = d_grad;
    b->bwd_depth = depth;
    b->bwd_adepth = adepth;
    return AWEFD_OK;
}

int awefd_batch_adjoint_grad(awefd_batch *b, const float *const *d_res, float *d_grad)
{
    if (!b || !d_res || !d_grad) return awefd__fail(AWEFD_EINVAL, "awefd_batch_adjoint_grad: NULL argument");
    HIPCHK(hipSetDevice(b->dev));
    const awefd_model *m0 = b->m[0];
    int rc;
    for (int s = 0; s < b->nshot; s++) {
        awefd_model *m = b->m[s];
        if (!d_res[s]) return awefd__fail(AWEFD_EINVAL, "awefd_batch_adjoint_grad: residual %d is NULL", s);
        if ((rc = awefd__adj_ready(m, "awefd_batch_adjoint_grad"))) return rc;
        if (awefd__adj_K(m) != b->K)
            return awefd__fail(AWEFD_ESTATE, "awefd_batch_adjoint_grad: shot %d was not run by awefd_batch_run_ckpt", s);
        if (m->p.n[0] != m0->p.n[0] || m->p.n[1] != m0->p.n[1] || m->nb != m0->nb || m->ld != m0->ld)
            return awefd__fail(AWEFD_EINVAL, "awefd_batch_adjoint_grad: shot %d has another grid (the gradient "
                                             "sums over shots on one grid)", s);
        if ((rc = awefd__adj_prepare(m))) return rc;
    }
    /* the tables hold the residual pointers and the buffers just prepared;
     * the captured graph reads them when it runs */
    if ((rc = adj_tables(b, d_res))) return rc;
    if ((rc = capture_bwd(b, d_grad, ckpt_depth(b), adj_depth()))) return rc;
    HIPCHK(hipGraphLaunch(b->xbwd, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return AWEFD_OK;
}
