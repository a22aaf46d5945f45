/* Private state of an awefd_model handle (host C). */
#ifndef AWEFD_MODEL_H
#define AWEFD_MODEL_H

#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include "awefd.h"
#include "awefd_internal.h"

struct awefd_model {
    awefd_params p;
    int ndim;
    int nop;            /* stencil half-width (cartesian) / NOP=4 (spherical) */
    int nb;             /* effective padding                                 */
    int np[3];          /* padded dims, fastest first (model axis order)     */
    int ld;             /* device row stride in floats (>= np[0])            */
    long nalloc;        /* device floats per field: ld * np[1] * np[2]       */
    float opad[3];      /* padded origins                                    */
    long ncell;
    int model_set, src_set, rec_set, wav_set;
    float c2[3 * AWEFD_MAX_NOP + 1], c1[3 * AWEFD_MAX_NOP + 1];

    int dev;
    hipStream_t stream;
    float *dP, *dC, *dF;          /* fields (dF: third field, spherical)     */
    float *dvel, *drho;
    float *dsp[3];                /* cartesian sponge factors per axis       */
    float *h_sp[3];
    float *dtab[6];               /* one-way ABC tables                      */

    /* sources / receivers: index map (3 ints, grid z/x/y order) and weights */
    int ns, nr;
    int *h_sj, *h_rj;
    float *h_sw, *h_rw;
    int *d_sj, *d_rj;
    float *d_sw, *d_rw;
    float *h_srefl, *d_srefl;     /* spherical reflectivity scale per source */
    awefd_inj *d_inj;
    int *d_ibeg;
    long ninj;

    float *d_wav;
    long nwav;

    /* cartesian modes: sponge = dabc without hybridbc (the sponge, and in
     * 3-D the one-way faces); hyb = dabc with hybridbc (apply_abc); adj2 =
     * 2-D adj=y (image at nt-kt, no traces); inj_rec = the injected points
     * are the receivers (2-D adj=y without expl); nsteps = steps a run takes */
    int sponge, hyb, adj2, inj_rec, nsteps;
    int sinc;                     /* sinc=y: points on 8-point sinc taps     */
    awefd_sinc_pt *h_ssinc, *h_rsinc;   /* sinc records of sources/receivers */
    awefd_sinc_pt *d_rsinc;
    float *d_damp_h;              /* damp table, nb - nop entries (hyb)      */
    float *h_wav;                 /* adj=y: host copy of the wavelet, and    */
    float *d_wavr;                /* its rows back to front on the device,   */
    unsigned wrev_gen;            /* built for model generation wrev_gen     */
    int wrev_ok;
    float *d_traces;
    long ntraces;

    /* cartesian tiling */
    int tz, ntz, ntx, ychunk, nty, nblk;
    int tx;                       /* x rows per step-kernel tile (bucket width) */
    int col;                      /* 3-D step in columns (k_cart_col)        */
    int narrow;                   /* 2-D step in one-wave tiles              */
    int ztail;                    /* its last z tile in 16-lane rows         */
    int target_blocks;            /* stencil blocks the y chunking aims at   */
    int zsep;                     /* z-high one-way ABC as a separate pass   */

    /* spherical state (spher_host.c) */
    double *d_damp;
    void *d_metric;
    void *d_abcrules;
    void *spher;

    /* profiling */
    int prof;
    long prof_launches;
    double prof_ms;
    hipEvent_t *ev;
    int nev;

    int steps_done;

    /* 2-D time loop captured as one HIP graph (runs without snapshots or
     * profiling): valid for model generation cgen and trace buffer cdtr */
    hipGraph_t cgraph;
    hipGraphExec_t cgexec;
    unsigned cgen;
    float *cdtr;
    float *cPf, *cCf;             /* field roles after the captured loop      */
    int cvalid;
    unsigned eager_gen;           /* signature of the last eager 2-D run: the */
    float *eager_dtr;             /* loop is captured when a run repeats it   */
    int eager_seen;

    /* gradient state (awefd_adjoint.c): checkpoints of the last
     * awefd_run_ckpt, valid while gen is unchanged; setters bump gen */
    void *adj;
    unsigned gen;
};

int awefd__fail(int code, const char *fmt, ...);

/* cartesian helpers shared with the batched path (awefd_api.c) */
void awefd__cart_args(const awefd_model *m, awefd_cart_args *a);
int  awefd__ensure_traces(awefd_model *m);
int  awefd__cart_buckets(const awefd_model *m, int tz, int ntz, int tx, awefd_inj **inj, int **ibeg,
                         int *nbucket, long *ninj);

/* gradient internals shared with the batched path (awefd_adjoint.c) */
void   awefd__adjoint_free(awefd_model *m);
int    awefd__adj_ckpt_setup(awefd_model *m, int ckpt);
float *awefd__adj_ckpt(const awefd_model *m, int k);
int    awefd__adj_K(const awefd_model *m);
void   awefd__adj_mark(awefd_model *m);
int    awefd__adj_ready(const awefd_model *m, const char *who);
int    awefd__adj_prepare(awefd_model *m);
void   awefd__adj_args(const awefd_model *m, const float *d_res, awefd_adj2d_args *g);
void   awefd__adj_args_batch(const awefd_model *m, const float *d_res, awefd_adj2d_args *g);
void   awefd__adj_cart_ckpt(const awefd_model *m, awefd_cart_args *a);
void   awefd__adj_zero_jobs(const awefd_model *m, awefd_copy_job *jobs);

/* spherical back end (spher_host.c) */
int  awefd__spher_create(awefd_model *m);
int  awefd__spher_set_model(awefd_model *m, const float *vel);
int  awefd__spher_set_sources(awefd_model *m);
int  awefd__spher_set_receivers(awefd_model *m);
/* device address of the current device's fixed scratch (awefd_scratch.hip);
 * returns a hipError_t */
int  awefd__scratch(int dev, float **out);
/* injection factor of source src: the expl=y reflectivity, else 1 (< 0: no model) */
float awefd__spher_src_factor(const awefd_model *m, int src);
int  awefd__spher_loop(awefd_model *m, float *dtr, int jsnap, awefd_snap_fn cb, void *user);
int  awefd__spher_get_field(awefd_model *m, float *host);
void awefd__spher_free(awefd_model *m);

#endif
