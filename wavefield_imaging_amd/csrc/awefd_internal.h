/*
 * Internal interface between the C host layer (awefd_api.c, awefd_setup.c)
 * and the HIP kernels (cart_kernels.hip, spher_kernels.hip).  Plain C structs
 * passed by value to the kernels.
 */
#ifndef AWEFD_INTERNAL_H
#define AWEFD_INTERNAL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AWEFD_MAX_NOP 6
#ifndef AWEFD_RX
#define AWEFD_RX 4                /* x rows per wave of the cartesian step kernel */
#endif
#ifndef AWEFD_NWAVE
#define AWEFD_NWAVE 4             /* waves per block of the cartesian step kernel */
#endif
#define AWEFD_CART_TX (AWEFD_NWAVE * AWEFD_RX) /* x rows per stencil block     */

/* One source-injection contribution: node (z, x, y) += amp(src) * w * vel. */
typedef struct {
    int   z, x, y;
    int   src;
    float w;
} awefd_inj;

/* Arguments of the cartesian step kernels. */
typedef struct {
    float       *P;          /* previous field, overwritten with the new one  */
    const float *C;          /* current field                                 */
    const float *vel;        /* (v dt)^2, padded                              */
    const float *rho;        /* density (variable-density runs) or NULL       */
    int n1, n2, n3;          /* padded z, x, y                                */
    int ld;                  /* row stride in floats (n1 rounded up to 32)    */
    int zsep;                /* z-high ABC done by a separate pass            */
    float c0;
    float cz[AWEFD_MAX_NOP], cx[AWEFD_MAX_NOP], cy[AWEFD_MAX_NOP];
    float bz[AWEFD_MAX_NOP], bx[AWEFD_MAX_NOP], by[AWEFD_MAX_NOP];
    /* tiling */
    int tz, ntz, ntx, ychunk, nty, nblk; /* ychunk: longest y chunk      */
    /* sponge: per-axis factors, 1 outside the strips */
    int dabc, nb;
    const float *sp1, *sp2, *sp3;
    /* 3-D one-way ABC */
    int abc, fsrf;
    const float *bzl, *bzh, *bxl, *bxh, *byl, *byh;
    /* injection, bucketed by (z tile, x tile) column, sorted by (y, z, x, src) */
    const int *ibeg;
    const awefd_inj *inj;
    const float *wav;
    long wstride;
    int wexpl;
    int it;
    /* extraction of the current field at this step */
    int nr;
    const int *rj;           /* 3 ints per receiver (jz, jx, jy)              */
    const float *rw;         /* 8 weights per receiver (4 used in 2-D)       */
    float *trow;             /* traces + it*nr                                */
} awefd_cart_args;

/* Launchers (cart_kernels.hip).  Return 0 or a HIP error code. */
int awefd_launch_cart_step(const awefd_cart_args *a, int nop, int ndim, int sse, int varden,
                           void *stream);
int awefd_launch_cart_abc_xy(const awefd_cart_args *a, int nop, void *stream);
int awefd_launch_fill(float *p, long n, float v, void *stream);
/* resident blocks per CU of the cartesian step kernel (0 if unknown) */
int awefd_cart_step_occupancy(int nop, int ndim, int sse);

#ifdef __cplusplus
}
#endif
#endif
