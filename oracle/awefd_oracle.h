/*
 * CPU oracle for the acoustic finite-difference time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is an independent C restatement of the
 * reference CPU programs sfAWEFDcart2Dcpu, sfAWEFDcart3Dcpu, sfAWEFDspher2Dcpu
 * and sfAWEFDspher3Dcpu (Nacnud04/wavefield_imaging, CODE/).  It is the checker
 * for the HIP library: only tests, the smoke check and the CPU-baseline leg of
 * the benchmark may load it.  The product never links, calls or falls back to it.
 *
 * Every routine cites the reference file:line it restates.  The FP32 entry
 * points (suffix _f32) follow the reference's C float/double evaluation order
 * exactly (built with -ffp-contract=off, no fast-math); the FP64 entry points
 * (suffix _f64) run the same algorithm in double and give the FP64 error.
 *
 * Pinning (tests/test_oracle.py; the reference ships no tests or golden data
 * for this path): FD coefficients against the closed form and the Taylor
 * moments (optimized tables by their wider accurate wavenumber band); sponge
 * tables against their closed forms; the index map and the spherical one-way
 * rule against hand-derived known answers (tests/golden/); every propagator,
 * FP64 and FP32, against a textbook float64 numpy leapfrog written from the
 * PDE (cartesian 2*nop-order Laplacian, spherical metric terms, the polar
 * d/dr sign), both away from and across the absorbing layers (sponge, one-way
 * faces, free surfaces); linearity, superposition, reciprocity, absorption
 * against a larger domain and the FP64 error.  Seeded mutations of the
 * stencil, metric, injection, swap, one-way, sponge, reflectivity, index map
 * and extraction code each fail at least one of these tests.
 */
#ifndef AWEFD_ORACLE_H
#define AWEFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameter block, mirroring the Madagascar parameters of the four programs.
 * Axis 1 is the fastest axis of the velocity file:
 *   cartesian: (z, x, y);  polar/spherical: (r, theta, phi). */
typedef struct {
    int   n[3];
    float o[3];
    float d[3];
    int   nt;
    float dt;
    int   nb;        /* "nb=" as given; < 0 means absent                       */
    int   fdorder;   /* cartesian only (default 4)                             */
    int   optfd;     /* cartesian: optimized FD coefficients                   */
    int   dabc;      /* cartesian: sponge (and 3-D one-way ABC)                */
    int   fsrf;      /* free surface                                           */
    int   expl;      /* cartesian: one wavelet for all sources;
                        spherical: reflectivity-scaled injection               */
    int   cden;      /* cartesian: constant density                            */
    int   sse;       /* cartesian: 1 = SSE operation order (ts_kernel.c:78-138),
                        0 = scalar order (ts_kernel.c:139-157)                 */
    int   ns, nr;
    int   snap, jsnap;
    int   adj;       /* cartesian adj=: wavelet read back to front; in 2-D the
                        loop stops at it = nt-kt (image) and, without expl,
                        injects nr samples per step at the receivers          */
    int   kt;        /* cartesian 2-D kt= (image step nt-kt)                   */
    int   hybrid;    /* cartesian hybridbc=: apply_abc instead of the sponge   */
    int   sinc;      /* cartesian sinc=: 8-point sinc injection and extraction */
} orc_par;

/* Snapshot callback: padded field (the reference's u0 / h_po after the step),
 * called when snap && it % jsnap == 0. */
typedef void (*orc_snap_f32)(void *ud, int it, const float *u);
typedef void (*orc_snap_f64)(void *ud, int it, const double *u);

/* Forward modelling.  Inputs are host arrays in RSF order:
 *   vel, den : unpadded model (den may be NULL)
 *   sou, rec : coordinate tuples as in the sou=/rec= files
 *              (cart2d (x,z); cart3d (x,y,z); spher2d (r,theta); spher3d (r,phi,theta))
 *   wav      : wavelet stream read sequentially by the time loop
 * Outputs:
 *   dat      : nt*nr traces (row it = receivers at step it)
 *   ufin     : padded final field (nullable)
 * Returns 0 on success, <0 on an unsupported option or bad argument. */
int orc_cart2d_f32(const orc_par *p, const float *vel, const float *den,
                   const float *sou, const float *rec, const float *wav,
                   float *dat, float *ufin, orc_snap_f32 cb, void *ud);
int orc_cart2d_f64(const orc_par *p, const float *vel, const float *den,
                   const float *sou, const float *rec, const float *wav,
                   double *dat, double *ufin, orc_snap_f64 cb, void *ud);
int orc_cart3d_f32(const orc_par *p, const float *vel, const float *den,
                   const float *sou, const float *rec, const float *wav,
                   float *dat, float *ufin, orc_snap_f32 cb, void *ud);
int orc_cart3d_f64(const orc_par *p, const float *vel, const float *den,
                   const float *sou, const float *rec, const float *wav,
                   double *dat, double *ufin, orc_snap_f64 cb, void *ud);
int orc_spher2d_f32(const orc_par *p, const float *vel,
                    const float *sou, const float *rec, const float *wav,
                    float *dat, float *ufin, orc_snap_f32 cb, void *ud);
int orc_spher2d_f64(const orc_par *p, const float *vel,
                    const float *sou, const float *rec, const float *wav,
                    double *dat, double *ufin, orc_snap_f64 cb, void *ud);
int orc_spher3d_f32(const orc_par *p, const float *vel,
                    const float *sou, const float *rec, const float *wav,
                    float *dat, float *ufin, orc_snap_f32 cb, void *ud);
int orc_spher3d_f64(const orc_par *p, const float *vel,
                    const float *sou, const float *rec, const float *wav,
                    double *dat, double *ufin, orc_snap_f64 cb, void *ud);

/* Wall time (s) of the last time loop, and the OpenMP thread count. */
double orc_last_loop_seconds(void);
int orc_last_imaged(void);

/* ints8r and the sinc point records of sinc2d_make/sinc3d_make (see
 * awefd_oracle.c) */
void orc_mksinc(float d, float *out);
void orc_ints8r(int nxin, float dxin, float fxin, const float *yin, float yinl, float yinr, int nxout,
                const float *xout, float *yout);
void orc_sinc_axis(float q, int zaxis, float *w, int *f, int *n);
int  orc_sinc_make(int ndim, int na, const float *c, const int np[3], const float opad[3], const float h[3],
                   int *rec, float *w);
int    orc_threads(void);

/* Building blocks exposed for the known-answer tests. */
int  orc_fdcoef(int nop, int ndim, const float *d, int optfd, int order,
                float *out /* ndim*nop+1 */);
void orc_sponge_table(int nb, float *w /* nb */);
void orc_lint2d(int na, const float *xz, int nzpad, int nxpad, float ozpad,
                float oxpad, float dz, float dx, int *jz, int *jx, float *w /* 4*na */);
void orc_lint3d(int na, const float *xyz, int nzpad, int nxpad, int nypad,
                float ozpad, float oxpad, float oypad, float dz, float dx, float dy,
                int *jz, int *jx, int *jy, float *w /* 8*na */);
double orc_spher_damp(int i, int nb);
/* ABC rule of the spherical one-way boundary for one cell (0 none, 1 th-low,
 * 2 th-high, 3 r-low, 4 r-high, 5 ph-low, 6 ph-high): the last rule that the
 * reference's serial sweep applies at that cell. */
int  orc_spher3d_rule(int ir, int ith, int iph, int nrpad, int nthpad, int nphpad);
/* Injection reflectivity of expl=y at padded node (jr, jth[, jph]) of the raw
 * padded velocity (r fastest, then theta, then phi): |v0-vk|/(v0+vk) of the
 * first differing neighbour in the reference's search order, 0 if none.
 * The node's whole neighbourhood must lie inside the array. */
float orc_refl_at(int ndim, const float *vel, int jr, int jth, int jph, int nrpad, int nthpad);

/* ---- FP64 tangent-linear model, adjoint and misfit gradient of the
 * cartesian 2-D loop (oracle_adj.c).  Constant density, no free surface.
 * Squared slowness s = 1/v^2 on the model grid (mz x mx, z fastest). ---- */
typedef struct {
    int nz, nx, mz, mx, nb, nop, nt, ns, nr, expl;
    long wstride;
    double c2[19];
    double *G, *V, *dVds;        /* padded: sponge product, (v dt)^2, dV/ds      */
    long *snode, *rnode;         /* 4 nodes per point                            */
    double *sw, *rw;
} orc_cart2d_f64_t;

/* The model is vel (float) or, when slow != NULL, the squared slowness slow
 * (double, model grid).  Gradient: J = 1/2 sum r^2 with r = res_in when given,
 * else d - dobs (dobs NULL: zero); dJ/ds into grad[mz*mx], J into *misfit,
 * forward traces into dat_out[nt*nr] (nullable). */
int orc_cart2d_grad_f64(const orc_par *p, const float *vel, const double *slow, const float *sou,
                        const float *rec, const float *wav, const float *dobs, const double *res_in,
                        double *grad, double *misfit, double *dat_out);
/* traces dd[nt*nr] of the model perturbation ds[mz*mx] (tangent-linear) */
int orc_cart2d_tlm_f64(const orc_par *p, const float *vel, const double *slow, const float *sou,
                       const float *rec, const float *wav, const double *ds, double *dd);
/* traces of the same FP64 scheme */
int orc_cart2d_fwd_f64(const orc_par *p, const float *vel, const double *slow, const float *sou,
                       const float *rec, const float *wav, double *dat);

#ifdef __cplusplus
}
#endif
#endif
