/*
 * CPU oracle (test infrastructure only): FP64 tangent-linear model and
 * adjoint of the cartesian 2-D time loop, and the gradient of the L2 trace
 * misfit with respect to squared slowness.
 *
 * The forward scheme is the reference's (MAWEFDcart2Dcpu.c:297-383 with
 * ts_kernel.c:49-161, constant density, SSE operation order irrelevant in
 * FP64).  With c_n the field at time n after its first sponge and p_n the
 * previous field after its second (the reference's u1 and u0 at the start of
 * step n), G the sponge (diagonal, 1 without dabc), M the interior mask,
 * V = (v dt)^2 and f_n the injected term w_n W V:
 *
 *   c_{n+1} = G [ M (2 c_n - p_n + V.L c_n) + (I - M) p_n + f_n ]
 *   p_{n+1} = G c_n
 *   d_n     = R G c_n                       (trace row n)
 *
 * Adjoint, backwards from alpha_nt = beta_nt = 0, with t = G alpha_{n+1} and
 * residual r_n = d_n - d_obs_n:
 *
 *   alpha_n = G R^T r_n + 2 M t + L (V M t) + G beta_{n+1}
 *   beta_n  = (I - 2M) t
 *   dJ/dV  += M (L c_n) t + w_n W t          (injection nodes)
 *
 * and dJ/ds for squared slowness s = 1/v^2 on the model grid: dV/ds =
 * -V v^2, summed over the padded nodes each model node was copied to.
 * The tangent-linear model differentiates the forward scheme directly:
 *
 *   dc_{n+1} = G [ M (2 dc_n - dp_n + V.L dc_n + dV.L c_n) + (I-M) dp_n + w_n W dV ]
 *   dp_{n+1} = G dc_n,   dd_n = R G dc_n,   dV = -V v^2 ds (padded)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "awefd_oracle.h"

/* shared set-up (awefd_oracle.c) */
int orc_cart2d_setup_f64(const orc_par *p, const float *vel_in, const double *slow, const float *sou,
                         const float *rec, orc_cart2d_f64_t *g);
void orc_cart2d_free_f64(orc_cart2d_f64_t *g);

static void lap2d(const orc_cart2d_f64_t *g, const double *u, double *out)
{
    const int nz = g->nz, nx = g->nx, nop = g->nop;
    memset(out, 0, sizeof(double) * (size_t)nz * nx);
#pragma omp parallel for schedule(static)
    for (int ix = nop; ix < nx - nop; ix++)
        for (int iz = nop; iz < nz - nop; iz++) {
            const long i = (long)ix * nz + iz;
            double l = g->c2[0] * u[i];
            for (int k = 1; k <= nop; k++)
                l += g->c2[k] * (u[i + k] + u[i - k]) + g->c2[nop + k] * (u[i + (long)k * nz] + u[i - (long)k * nz]);
            out[i] = l;
        }
}

/* transpose of "keep interior rows of L": out = L^T y for y supported on the
 * interior.  L is symmetric node to node and y vanishes off the interior, so
 * out_i gathers c0 y_i and c_k times the y of each neighbour that exists in
 * the array (a z neighbour in the same column); each node is written once. */
static void lapT2d(const orc_cart2d_f64_t *g, const double *y, double *out)
{
    const int nz = g->nz, nx = g->nx, nop = g->nop;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            const long i = (long)ix * nz + iz;
            double o = g->c2[0] * y[i];
            for (int k = 1; k <= nop; k++) {
                if (iz - k >= 0) o += g->c2[k] * y[i - k];
                if (iz + k < nz) o += g->c2[k] * y[i + k];
                if (ix - k >= 0) o += g->c2[nop + k] * y[i - (long)k * nz];
                if (ix + k < nx) o += g->c2[nop + k] * y[i + (long)k * nz];
            }
            out[i] = o;
        }
}

static int interior(const orc_cart2d_f64_t *g, long i)
{
    const int iz = (int)(i % g->nz), ix = (int)(i / g->nz);
    return iz >= g->nop && iz < g->nz - g->nop && ix >= g->nop && ix < g->nx - g->nop;
}

/* forward: traces dat[nt*nr]; optionally every c_n (nt+1 fields) */
static void forward(const orc_cart2d_f64_t *g, const float *wav, double *dat, double *cs, const double *dV,
                    double *ddat)
{
    const long N = (long)g->nz * g->nx;
    double *c = calloc(N, sizeof(double)), *pp = calloc(N, sizeof(double)), *cn = malloc(sizeof(double) * N);
    double *lc = malloc(sizeof(double) * N), *dc = NULL, *dp = NULL, *dcn = NULL, *ldc = NULL;
    if (dV) {
        dc = calloc(N, sizeof(double));
        dp = calloc(N, sizeof(double));
        dcn = malloc(sizeof(double) * N);
        ldc = malloc(sizeof(double) * N);
    }
    for (int it = 0; it < g->nt; it++) {
        if (cs) memcpy(cs + (size_t)it * N, c, sizeof(double) * N);
        /* trace row it = R G c_it */
        for (int r = 0; r < g->nr; r++) {
            double s = 0, ds = 0;
            for (int q = 0; q < 4; q++) {
                const long node = g->rnode[4 * r + q];
                s += g->G[node] * c[node] * g->rw[4 * r + q];
                if (dV) ds += g->G[node] * dc[node] * g->rw[4 * r + q];
            }
            dat[(long)it * g->nr + r] = s;
            if (dV) ddat[(long)it * g->nr + r] = ds;
        }
        lap2d(g, c, lc);
        if (dV) lap2d(g, dc, ldc);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; i++) {
            double a = interior(g, i) ? 2 * c[i] - pp[i] + g->V[i] * lc[i] : pp[i];
            cn[i] = a;
            if (dV) dcn[i] = interior(g, i) ? 2 * dc[i] - dp[i] + g->V[i] * ldc[i] + dV[i] * lc[i] : dp[i];
        }
        for (int s = 0; s < g->ns; s++)
            for (int q = 0; q < 4; q++) {
                const long node = g->snode[4 * s + q];
                const double w = (double)wav[(long)it * g->wstride + (g->expl ? 0 : s)] * g->sw[4 * s + q];
                cn[node] += w * g->V[node];
                if (dV) dcn[node] += w * dV[node];
            }
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; i++) {
            pp[i] = g->G[i] * c[i];
            c[i] = g->G[i] * cn[i];
            if (dV) {
                dp[i] = g->G[i] * dc[i];
                dc[i] = g->G[i] * dcn[i];
            }
        }
    }
    free(c); free(pp); free(cn); free(lc); free(dc); free(dp); free(dcn); free(ldc);
}

/* dV/ds on the padded grid and the padding adjoint: s is on the model grid */
static void model_chain(const orc_cart2d_f64_t *g, const double *gV, double *gs)
{
    const int nz = g->nz, nx = g->nx, nb = g->nb, mz = g->mz, mx = g->mx;
    memset(gs, 0, sizeof(double) * (size_t)mz * mx);
    for (int ix = 0; ix < nx; ix++) {
        int kx = ix - nb; kx = kx < 0 ? 0 : (kx >= mx ? mx - 1 : kx);
        for (int iz = 0; iz < nz; iz++) {
            int kz = iz - nb; kz = kz < 0 ? 0 : (kz >= mz ? mz - 1 : kz);
            const long i = (long)ix * nz + iz;
            gs[(long)kx * mz + kz] += gV[i] * g->dVds[i];
        }
    }
}

int orc_cart2d_grad_f64(const orc_par *p, const float *vel, const double *slow, const float *sou,
                        const float *rec, const float *wav, const float *dobs, const double *res_in,
                        double *grad, double *misfit, double *dat_out)
{
    orc_cart2d_f64_t g;
    int rc = orc_cart2d_setup_f64(p, vel, slow, sou, rec, &g);
    if (rc) return rc;
    const long N = (long)g.nz * g.nx;
    double *cs = malloc(sizeof(double) * (size_t)N * g.nt);
    double *dat = malloc(sizeof(double) * (size_t)g.nt * g.nr);
    double *res = malloc(sizeof(double) * (size_t)g.nt * g.nr);
    double *al = calloc(N, sizeof(double)), *be = calloc(N, sizeof(double)), *t = malloc(sizeof(double) * N);
    double *y = malloc(sizeof(double) * N), *ly = malloc(sizeof(double) * N), *lc = malloc(sizeof(double) * N);
    double *gV = calloc(N, sizeof(double)), *an = malloc(sizeof(double) * N);
    if (!cs || !dat || !res || !al || !be || !t || !y || !ly || !lc || !gV || !an) {
        rc = -3;
        goto done;
    }
    forward(&g, wav, dat, cs, NULL, NULL);
    double J = 0;
    for (long k = 0; k < (long)g.nt * g.nr; k++) {
        res[k] = res_in ? res_in[k] : dat[k] - (dobs ? (double)dobs[k] : 0.0);
        J += 0.5 * res[k] * res[k];
    }
    if (misfit) *misfit = J;
    if (dat_out) memcpy(dat_out, dat, sizeof(double) * (size_t)g.nt * g.nr);
    for (int n = g.nt - 1; n >= 0; n--) {
        /* t = G alpha_{n+1} (alpha_nt = 0) */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; i++) {
            t[i] = g.G[i] * al[i];
            y[i] = interior(&g, i) ? g.V[i] * t[i] : 0.0;
        }
        lapT2d(&g, y, ly);
        lap2d(&g, cs + (size_t)n * N, lc);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; i++) {
            const int in = interior(&g, i);
            if (in) gV[i] += lc[i] * t[i];
            an[i] = (in ? 2 * t[i] : 0.0) + ly[i] + g.G[i] * be[i];
            be[i] = (in ? -1.0 : 1.0) * t[i];
        }
        /* injection at step n into c_{n+1} */
        for (int s = 0; s < g.ns; s++)
            for (int q = 0; q < 4; q++) {
                const long node = g.snode[4 * s + q];
                gV[node] += (double)wav[(long)n * g.wstride + (g.expl ? 0 : s)] * g.sw[4 * s + q] * t[node];
            }
        /* output d_n = R G c_n */
        for (int r = 0; r < g.nr; r++)
            for (int q = 0; q < 4; q++) {
                const long node = g.rnode[4 * r + q];
                an[node] += g.G[node] * g.rw[4 * r + q] * res[(long)n * g.nr + r];
            }
        memcpy(al, an, sizeof(double) * N);
    }
    model_chain(&g, gV, grad);
done:
    free(cs); free(dat); free(res); free(al); free(be); free(t); free(y); free(ly); free(lc); free(gV); free(an);
    orc_cart2d_free_f64(&g);
    return rc;
}

/* Tangent-linear model: traces dd of the perturbation ds (model grid) */
int orc_cart2d_tlm_f64(const orc_par *p, const float *vel, const double *slow, const float *sou,
                       const float *rec, const float *wav, const double *ds, double *dd)
{
    orc_cart2d_f64_t g;
    int rc = orc_cart2d_setup_f64(p, vel, slow, sou, rec, &g);
    if (rc) return rc;
    const long N = (long)g.nz * g.nx;
    double *dV = malloc(sizeof(double) * N), *dat = malloc(sizeof(double) * (size_t)g.nt * g.nr);
    if (!dV || !dat) {
        free(dV); free(dat);
        orc_cart2d_free_f64(&g);
        return -3;
    }
    for (int ix = 0; ix < g.nx; ix++) {
        int kx = ix - g.nb; kx = kx < 0 ? 0 : (kx >= g.mx ? g.mx - 1 : kx);
        for (int iz = 0; iz < g.nz; iz++) {
            int kz = iz - g.nb; kz = kz < 0 ? 0 : (kz >= g.mz ? g.mz - 1 : kz);
            const long i = (long)ix * g.nz + iz;
            dV[i] = g.dVds[i] * ds[(long)kx * g.mz + kz];
        }
    }
    forward(&g, wav, dat, NULL, dV, dd);
    free(dV); free(dat);
    orc_cart2d_free_f64(&g);
    return 0;
}

/* Plain FP64 forward of the same scheme (traces), for finite differences */
int orc_cart2d_fwd_f64(const orc_par *p, const float *vel, const double *slow, const float *sou,
                       const float *rec, const float *wav, double *dat)
{
    orc_cart2d_f64_t g;
    int rc = orc_cart2d_setup_f64(p, vel, slow, sou, rec, &g);
    if (rc) return rc;
    forward(&g, wav, dat, NULL, NULL, NULL);
    orc_cart2d_free_f64(&g);
    return 0;
}
