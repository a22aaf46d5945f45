/*
 * awefd.h -- C ABI of libawefd, the MI355X (gfx950) acoustic finite-difference
 * time loop behind the Madagascar programs sfAWEFDcart2D, sfAWEFDcart3D,
 * sfAWEFDspher2D and sfAWEFDspher3D (Nacnud04/wavefield_imaging, CODE/).
 *
 * The reference has no library boundary: each program's main() reads RSF
 * files, sets up the grid and runs the time loop in-process.  libawefd is that
 * time loop behind plain C entry points (no torch or HIP types in the
 * signatures), so that the project's own C programs (the drop-in CLIs) and a
 * ctypes binding share it.  Each entry point names the part of the reference
 * main it replaces.
 *
 * Conventions
 *   - Return value 0 on success, a negative AWEFD_E* code on error; the message
 *     is available from awefd_last_error() (thread-local).  Nothing exits.
 *   - Host buffers belong to the caller.  Device buffers belong to the handle,
 *     except in the *_dev entry points, which take device pointers.
 *   - One handle = one grid on one GPU, one HIP stream.  A handle is not
 *     thread-safe; separate handles are independent.
 *   - Arrays are in RSF order, axis 1 fastest: cartesian (z, x, y),
 *     polar/spherical (r, theta, phi).
 */
#ifndef AWEFD_H
#define AWEFD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AWEFD_CART2D  0   /* sfAWEFDcart2D   MAWEFDcart2Dcpu.c  */
#define AWEFD_CART3D  1   /* sfAWEFDcart3D   MAWEFDcart3Dcpu.c  */
#define AWEFD_SPHER2D 2   /* sfAWEFDspher2D  MAWEFDspher2Dcpu.c */
#define AWEFD_SPHER3D 3   /* sfAWEFDspher3D  MAWEFDspher3Dcpu.c */

#define AWEFD_OK        0
#define AWEFD_EINVAL   -1   /* bad argument or unsupported option            */
#define AWEFD_EHIP     -2   /* HIP runtime error                             */
#define AWEFD_ENOMEM   -3   /* host or device allocation failed              */
#define AWEFD_ESTATE   -4   /* call out of order (e.g. run before model set) */

/* Program parameters (the getpar keys of the four mains, same meaning and
 * defaults; MAWEFDcart2Dcpu.c:116-128,167-211, MAWEFDcart3Dcpu.c:116-210,
 * MAWEFDspher2Dcpu.c:56-116, MAWEFDspher3Dcpu.c:57-133). */
typedef struct awefd_params {
    int   kind;      /* AWEFD_CART2D ... AWEFD_SPHER3D                        */
    int   n[3];      /* model axes n1,n2,n3 (unpadded); n3 = 1 in 2-D          */
    float o[3];      /* axis origins o1,o2,o3                                 */
    float d[3];      /* axis samplings d1,d2,d3                               */
    int   nt;        /* time steps (wavelet axis 2)                           */
    float dt;        /* time sampling                                         */
    int   nb;        /* nb=  (< 0: not given; clamped to the stencil half-width) */
    int   fdorder;   /* fdorder= (cartesian; 2..12, default 4)                */
    int   optfd;     /* optfd=  (cartesian)                                   */
    int   dabc;      /* dabc=   (cartesian sponge, plus the 3-D one-way ABC)   */
    int   fsrf;      /* free= / fsrf=                                         */
    int   expl;      /* expl=  cartesian: one wavelet sample for all sources;
                        spherical: reflectivity-scaled injection (default y)   */
    int   cden;      /* cden=  (cartesian constant density)                    */
    int   sse;       /* operation order of ts_kernel.c: 1 = 4-wide SSE build
                        (default), 0 = scalar build                           */
    int   jdata;     /* jdata= (receiver decimation; traces keep every step)   */
    int   device;    /* gpu=   HIP device ordinal                             */
    int   adj;       /* adj=   cartesian: the wavelet is read from its last row
                        back (MAWEFDcart2Dcpu.c:305-316, MAWEFDcart3Dcpu.c:304-315).
                        In 2-D the loop also stops after step nt-kt, whose field
                        is the image (awefd_get_field), extracts no traces, and
                        without expl injects nr samples per step at the receivers */
    int   kt;        /* kt=    cartesian 2-D adj: image step nt-kt, 1 <= kt <= nt */
    int   hybrid;    /* hybridbc= cartesian: with dabc, the hybrid one-way/damped
                        boundary (apply_abc) replaces the sponge (and the 3-D
                        one-way faces); nb <= nop widens to 2 nop; a free surface
                        no longer zeroes (v dt)^2 (MAWEFDcart2Dcpu.c:211-212,276-292) */
    int   sinc;      /* sinc=  cartesian: sources and receivers on 8-point sinc
                        taps, 9 nodes per axis (sinc2d_make/sinc3d_make,
                        fdutil.c:784-1360) instead of bilinear corners      */
} awefd_params;

typedef struct awefd_model awefd_model;

/* Fill *p with the reference defaults of program `kind` (MAWEFD*cpu.c getpar
 * calls).  n/o/d/nt/dt must still be set by the caller. */
int awefd_params_default(int kind, awefd_params *p);

/* Grid set-up: replaces fdutil_init/fdutil3d_init, compute_fdcoef and the
 * allocations of MAWEFDcart2Dcpu.c:209-287 (3-D :207-284, polar :114-168,
 * spherical :99-206).  Allocates the device wavefields on p->device. */
int awefd_create(const awefd_params *p, awefd_model **out);

/* Padded grid dimensions (fastest first) and the padding width nb. */
int awefd_padded_dims(const awefd_model *m, int dims[3], int *nb);

/* Velocity (and density, nullable) in RSF order, unpadded: replaces the
 * sf_floatread + expand/expand3d/expand_cpu_* and the (v dt)^2 scaling,
 * free-surface zeroing, sponge and one-way coefficient set-up
 * (MAWEFDcart2Dcpu.c:258-295, MAWEFDcart3Dcpu.c:254-293,
 * MAWEFDspher2Dcpu.c:153-220, MAWEFDspher3Dcpu.c:185-272). */
int awefd_set_model(awefd_model *m, const float *vel, const float *den);

/* The velocity from device memory (d_vel: RSF order, unpadded, on the
 * model's device, complete when called): the same padded (v dt)^2 and
 * free-surface zeros as awefd_set_model(m, vel, NULL) (MAWEFDcart2Dcpu.c:
 * 258-295 for a velocity that is already in HBM), so an inversion loop
 * updates the model without a host round trip.  Cartesian models whose
 * other tables do not depend on the velocity: not hybridbc=y, and in 3-D
 * not dabc=y (AWEFD_EINVAL otherwise: use awefd_set_model). */
int awefd_set_model_dev(awefd_model *m, const float *d_vel);

/* Source / receiver coordinates as stored in the sou= / rec= files:
 * cart2d (x,z), cart3d (x,y,z), spher2d (r,theta), spher3d (r,phi,theta).
 * Replaces pt2dread1/pt3dread1 + lint2d_make/lint3d_make
 * (MAWEFDcart2Dcpu.c:249-256, MAWEFDcart3Dcpu.c:245-252, fdutil.c:1363-1539). */
int awefd_set_sources(awefd_model *m, int ns, const float *coords);
int awefd_set_receivers(awefd_model *m, int nr, const float *coords);

/* Wavelet stream in the order the time loop reads it: cartesian reads
 * (expl ? 1 : ns) samples per step (MAWEFDcart2Dcpu.c:318-326); spherical
 * reads one trace of nt samples (MAWEFDspher3Dcpu.c:173-182).
 * nsamples = number of floats available. */
int awefd_set_wavelet(awefd_model *m, long nsamples, const float *wav);

/* Snapshot callback: the padded field the reference would write for step it
 * (u0 after the swap / h_po after the step), host memory, valid during the
 * call only. */
typedef void (*awefd_snap_fn)(void *user, int it, const float *field, const int dims[3]);

/* Run the whole time loop (MAWEFDcart2Dcpu.c:297-383, MAWEFDcart3Dcpu.c:295-366,
 * MAWEFDspher2Dcpu.c:264-317, MAWEFDspher3Dcpu.c:302-359).
 * traces: host nt*nr floats (row it = receivers at step it), may be NULL.
 * cb is called for it % jsnap == 0 when jsnap > 0 (cb may be NULL). */
int awefd_run(awefd_model *m, float *traces, int jsnap, awefd_snap_fn cb, void *user);

/* Traces of the last run (nt*nr floats, row it = receivers at step it) from
 * device memory; used after awefd_run_dev or a batched run.  Replaces the
 * per-step sf_floatwrite of the traces (MAWEFDcart2Dcpu.c:373-378). */
int awefd_get_traces(awefd_model *m, float *host);

/* Same loop with every buffer already resident in HBM: d_traces is a device
 * pointer to nt*nr floats (may be NULL); no host traffic, no snapshots. */
int awefd_run_dev(awefd_model *m, float *d_traces);

/* Device address of the handle's own trace buffer (nt*nr floats): the traces
 * of the last run that wrote there (awefd_run, awefd_run_dev without a
 * buffer, batched runs).  Valid until awefd_destroy. */
int awefd_traces_dev(awefd_model *m, float **d_traces);

/* Reset the wavefields to zero (the state before the first step). */
int awefd_reset(awefd_model *m);

/* Copy the current padded field, as the reference would write it after the
 * last completed step, into host memory (dims from awefd_padded_dims). */
int awefd_get_field(awefd_model *m, float *host);

/* Bit-exact view of the host index map (lint2d_make / lint3d_make,
 * fdutil.c:1363-1539): which = 0 sources, 1 receivers.  idx gets 3 ints per
 * point (j1, j2, j3 along the fastest..slowest axis), w gets 8 floats per
 * point in the reference corner order (4 used in 2-D).
 * which = 2: each source's injection factor in w[8*i] (the rest 0): the
 * spherical expl=y reflectivity of inject_sources_2D/3D
 * (spher_utils.c:153-173, 250-310), 1 otherwise; needs the model set. */
int awefd_debug_points(const awefd_model *m, int which, int *idx, float *w);

/* Per-kernel timing with HIP events on the handle's stream.  enable=1 starts
 * recording events around every launch of the dominant (stencil) kernel;
 * awefd_kernel_stats returns the launch count and total milliseconds. */
int awefd_profile(awefd_model *m, int enable);
int awefd_kernel_stats(awefd_model *m, long *launches, double *total_ms);

/* HIP device ordinal the handle runs on. */
int awefd_device(const awefd_model *m);

const char *awefd_last_error(void);
void awefd_destroy(awefd_model *m);

/* ---- Batched shots.  No counterpart in the reference, which runs one
 * program per shot (byrne_awe.py:163-250): for callers with many independent
 * 2-D cartesian shots of uneven size.  The shots' time loops run together,
 * one launch per step over every shot, captured into one HIP graph. ---- */
typedef struct awefd_batch awefd_batch;

/* Group fully set-up AWEFD_CART2D models (same nt, fdorder, sse and device;
 * constant density).  The models keep their buffers and must outlive the
 * batch. */
int awefd_batch_create(int nshot, awefd_model *const *models, awefd_batch **out);
/* Tile shape (x rows per wave: 2, 4, 8 or 16; default 8) and shot grouping:
 * group_mb > 0 runs consecutive shots whose fields total at most group_mb MiB
 * through all their steps before the next group (cache-resident groups);
 * 0 advances every shot in every launch. */
int awefd_batch_set_tile(awefd_batch *b, int rx, int group_mb);
/* Autotune both: time `steps` steps of every candidate (rx in 2, 4, 8, 16 x
 * group_mb in 0, 32, 64, 128, 192) on the batch and keep the fastest.
 * Outputs are nullable; ms_out[4*5] gets ms per step, row-major by rx. */
int awefd_batch_tune(awefd_batch *b, int steps, int *rx_out, int *group_mb_out, double *ms_out);
/* Run every shot's whole time loop (fields reset first).  Afterwards each
 * model holds what its own awefd_run_dev would leave: traces via
 * awefd_get_traces, the field via awefd_get_field. */
int awefd_batch_run(awefd_batch *b);
void awefd_batch_destroy(awefd_batch *b);
/* Steps per launch of awefd_batch_run: 1 (default), 2 or 3 (tile rx 4 or
 * 8).  With 2 or 3, each launch advances every shot that many steps on
 * overlapping tiles (each step's field kept in LDS for the next), reading
 * the fields once per launch; results are bit-identical. */
int awefd_batch_set_depth(awefd_batch *b, int depth);
/* awefd_batch_tune over the depth as well (depth 1: rx 2, 4, 8, 16; depths
 * 2 and 3: rx 4, 8) and keep the fastest; ms_out[3*4*5] holds ms per step,
 * depth-major then row-major by rx (0 for the choices not tried). */
int awefd_batch_tune_depth(awefd_batch *b, int steps, int *rx_out, int *group_mb_out, int *depth_out,
                           double *ms_out);
/* The batch's awefd_run_ckpt: every shot's forward loop with checkpoints
 * every ckpt steps (<= 0: ceil(sqrt(nt))), one graph.  Traces as after
 * awefd_batch_run. */
int awefd_batch_run_ckpt(awefd_batch *b, int ckpt);
/* Sum over the shots of awefd_adjoint_grad, every shot's backward pass in
 * one graph (one launch per step over a group of shots).  d_res: nshot
 * device pointers (float[nt][nr] of each shot); d_grad: device
 * float[n2][n1].  The shots must share the grid (n, nb) and come from the
 * last awefd_batch_run_ckpt; the sum runs in shot order. */
int awefd_batch_adjoint_grad(awefd_batch *b, const float *const *d_res, float *d_grad);

/* ---- Misfit gradient.  No counterpart in the reference, whose adj=y mode
 * (MAWEFDcart2Dcpu.c:305-316) back-propagates data but forms no gradient:
 * for inversion callers that need dJ/ds of J(s) = sum r * d(s) over the
 * traces d of a cartesian 2-D loop, s = 1/v^2 on the model grid. ---- */

/* The forward loop of awefd_run_dev (same traces, same final fields) that
 * also keeps the field pair of every ckpt-th step (<= 0: ceil(sqrt(nt))) in
 * device memory for awefd_adjoint_grad: 2 ceil(nt/ckpt) padded fields.
 * AWEFD_CART2D with constant density only. */
int awefd_run_ckpt(awefd_model *m, float *d_traces, int ckpt);
/* dJ/ds at the model of the last awefd_run_ckpt (AWEFD_ESTATE if a setter
 * ran since).  d_res: device float[nt][nr], the trace weights r = dJ/dd (for
 * an L2 misfit the residual d - d_obs).  d_grad: device float[n2][n1] (n1
 * fastest).  Recomputes each segment of fields from its checkpoint; holds
 * ckpt + 3 more padded fields (kept for later calls until
 * awefd_adjoint_release or awefd_destroy).  The model's fields keep the
 * forward run's final state. */
int awefd_adjoint_grad(awefd_model *m, const float *d_res, float *d_grad);
/* Free the checkpoint and adjoint buffers of a model. */
void awefd_adjoint_release(awefd_model *m);

/* ---- Allocation-free run.  No counterpart in the reference, whose
 * programs allocate per run: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). ---- */

/* Largest trace block (nt * nr * 4 bytes) awefd_run_noalloc can stage. */
#define AWEFD_NOALLOC_TRACE_BYTES (32u << 20)

/* awefd_run without snapshots (MAWEFDcart2Dcpu.c:297-383, MAWEFDcart3Dcpu.c:295-366,
 * MAWEFDspher2Dcpu.c:264-317, MAWEFDspher3Dcpu.c:302-359; same traces, same
 * final field), allocating nothing: the receiver traces are staged in the
 * library's fixed device scratch (one per GPU, AWEFD_NOALLOC_TRACE_BYTES,
 * held until the process ends) instead of the handle's trace buffer, the
 * steps are launched one by one (no graph), and every table the loop reads
 * was built by the setters.  traces: host nt*nr floats (may be NULL); they
 * are returned only here (awefd_get_traces does not see them).
 * AWEFD_EINVAL if nt*nr*4 exceeds the scratch or profiling is on.
 * Every call on one GPU uses that GPU's one scratch: calls that overlap on
 * one GPU (threads, separate handles) race and must be serialised by the
 * caller. */
int awefd_run_noalloc(awefd_model *m, float *traces);

#ifdef __cplusplus
}
#endif
#endif
