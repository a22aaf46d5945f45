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
    free(b->u0);
    free(b->u1);
    free(b->goff);
    free(b->gnblk);
    free(b);
}

int awefd_batch_create(int nshot, awefd_model *const *models, awefd_batch **out)
{
    if (!out || nshot < 1 || !models) return awefd__fail(AWEFD_EINVAL, "awefd_batch_create: bad argument");
    *out = NULL;
    const awefd_model *m0 = models[0];
    for (int s = 0; s < nshot; s++) {
        const awefd_model *m = models[s];
        if (!m) return awefd__fail(AWEFD_EINVAL, "batch: shot %d is NULL", s);
        if (m->p.kind != AWEFD_CART2D)
            return awefd__fail(AWEFD_EINVAL, "batch: shot %d is not a cartesian 2-D model", s);
        if (m->p.nt != m0->p.nt || m->nop != m0->nop || m->p.sse != m0->p.sse || m->dev != m0->dev)
            return awefd__fail(AWEFD_EINVAL, "batch: shot %d differs from shot 0 in nt, fdorder, sse or device", s);
        if (m->drho && !m->p.cden)
            return awefd__fail(AWEFD_EINVAL, "batch: shot %d has variable density (not batched)", s);
        if (!m->model_set || !m->src_set || !m->rec_set || !m->wav_set)
            return awefd__fail(AWEFD_ESTATE, "batch: shot %d is not fully set up", s);
        const long need = (long)m->p.nt * (m->p.expl ? 1 : m->ns);
        if (m->nwav < need) return awefd__fail(AWEFD_EINVAL, "batch: shot %d wavelet too short", s);
    }
    awefd_batch *b = calloc(1, sizeof *b);
    if (!b) return awefd__fail(AWEFD_ENOMEM, "batch");
    b->nshot = nshot;
    b->m = malloc(sizeof(*b->m) * nshot);
    b->d_inj = calloc(nshot, sizeof(*b->d_inj));
    b->d_ibeg = calloc(nshot, sizeof(*b->d_ibeg));
    b->u0 = malloc(sizeof(*b->u0) * nshot);
    b->u1 = malloc(sizeof(*b->u1) * nshot);
    if (!b->m || !b->d_inj || !b->d_ibeg || !b->u0 || !b->u1) {
        awefd_batch_destroy(b);
        return awefd__fail(AWEFD_ENOMEM, "batch");
    }
    memcpy(b->m, models, sizeof(*b->m) * nshot);
    for (int s = 0; s < nshot; s++) {
        b->u0[s] = models[s]->dP;
        b->u1[s] = models[s]->dC;
    }
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
    free(h);
    free(grp);
    free(nb);
    if (!rc) {
        free(b->goff);
        free(b->gnblk);
        b->goff = goff;
        b->gnblk = gnblk;
        b->ngroup = ng;
        b->rx = rx;
        b->group_mb = group_mb;
    } else {
        free(goff);
        free(gnblk);
    }
    b->graph_ok = 0;   /* a captured graph refers to the tables just replaced */
    return rc;
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

int awefd_batch_set_tile(awefd_batch *b, int rx, int group_mb)
{
    if (!b) return awefd__fail(AWEFD_EINVAL, "NULL batch");
    int ok = 0;
    for (int i = 0; i < 4; i++) ok |= RX_CAND[i] == rx;
    if (!ok) return awefd__fail(AWEFD_EINVAL, "batch tile: rx=%d (2, 4, 8 or 16)", rx);
    if (group_mb < 0) return awefd__fail(AWEFD_EINVAL, "batch group budget %d MiB", group_mb);
    HIPCHK(hipSetDevice(b->dev));
    return batch_build(b, rx, group_mb);
}

int awefd_batch_tune(awefd_batch *b, int steps, int *rx_out, int *group_out, double *ms_out)
{
    if (!b) return awefd__fail(AWEFD_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->dev));
    if (steps < 1) steps = 1;
    if (steps > b->nt) steps = b->nt;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int bi = -1, bg = -1, rc = AWEFD_OK;
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

/* The time loop of every shot as one graph: field resets, then one batched
 * launch per step of each group. */
static int batch_capture(awefd_batch *b)
{
    if (b->gexec && b->graph_ok) return AWEFD_OK;
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = NULL;
    b->graph = NULL;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_reset(b);
    if (!rc) rc = batch_steps(b, b->nt);
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
    if (!b->rx && (rc = batch_build(b, 8, 0))) return rc;
    if ((rc = batch_capture(b))) return rc;
    HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    /* leave each model as a single run leaves it: step it writes buffer
     * (it even ? u0 : u1), so the reference's current field after the last
     * step (u0 after its swap) is buffer u1 for odd nt, u0 for even nt */
    for (int s = 0; s < b->nshot; s++) {
        awefd_model *m = b->m[s];
        m->dP = b->nt % 2 ? b->u1[s] : b->u0[s];
        m->dC = b->nt % 2 ? b->u0[s] : b->u1[s];
        m->steps_done = b->nt;
    }
    return AWEFD_OK;
}
