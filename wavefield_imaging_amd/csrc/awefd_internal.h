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
/* The narrow 2-D tile: one wave of 3 x rows per block, for grids whose
 * 4 x 4 tiling leaves fewer than 3 blocks per CU.  On the 500 x 900 grid
 * (456 blocks of 4 x 4) it runs 3.68 us per step against 4.34 us (1 wave x
 * 2 rows 3.86, x 4 4.10, x 5 3.98, x 6 4.21; 2 waves x 2 rows 4.24); from
 * 800 x 1440 (1170 blocks) on, 4 x 4 is 3-10 % faster. */
#ifndef AWEFD_RX2
#define AWEFD_RX2 3
#endif
#ifndef AWEFD_NWAVE2
#define AWEFD_NWAVE2 1
#endif
#define AWEFD_CART_TX2 (AWEFD_NWAVE2 * AWEFD_RX2)
#ifndef AWEFD_COL_RX
#define AWEFD_COL_RX 2            /* x rows per wave of the 3-D column step kernel */
#endif
#ifndef AWEFD_COL_PD
#define AWEFD_COL_PD 3            /* planes its loads run ahead                    */
#endif
#define AWEFD_BATCH_NWAVE 4      /* waves per block of the batched 2-D kernel     */
/* x rows per block of the three-step 2-D pass: 32, as 4 waves of 8 rows
 * (rx 8) or 8 waves of 4 rows (rx 4).  On 256 uneven shots, per step: rx 8 x
 * 4 waves 0.194-0.200 ms, x 6 waves 0.250, x 8 waves (64 rows) 0.203-0.212,
 * x 12 0.228, x 16 0.302; rx 4 x 8 waves 0.249-0.257, x 4 waves 0.286-0.291
 * (the steps' barriers wait on every wave of a block; smaller blocks read
 * more halo). */
#ifndef AWEFD_T3_ROWS
#define AWEFD_T3_ROWS 32
#endif
#define AWEFD_T3_NWAVE(rx) (AWEFD_T3_ROWS / (rx))

/* One source-injection contribution: node (z, x, y) += amp(src) * w * vel. */
typedef struct {
    int   z, x, y;
    int   src;
    float w, w2, w3;   /* the node receives ((amp w) w2) w3 v: bilinear
                          points w2 = w3 = 1; sinc points the per-axis tap
                          weights (2-D wx, wz, 1; 3-D wy, wx, wz) */
} awefd_inj;

/* 8-point sinc point (sinc2d_make / sinc3d_make, fdutil.c:784-842,1088-1145):
 * centre node, first tap (0..8) and tap count (9, or 1 on a node) per axis,
 * and the taps' weights; axis order z, x, y. */
typedef struct {
    int   c[3];
    int   f[3], n[3];
    float w[3][9];
} awefd_sinc_pt;

/* A receiver of the three-step pass's in-block extraction: its index, first
 * corner (z, x), its four bilinear weights and the sponge factors of the
 * corners' z and x (1 without the sponge), so that a block reads one record
 * instead of a chain of dependent loads. */
typedef struct {
    int   r, jz, jx, pad;
    float w[4];
    float sz[2], sx[2];
} awefd_rx3;

/* Arguments of the cartesian step kernels. */
typedef struct {
    float       *P;          /* previous field, overwritten with the new one  */
    const float *C;          /* current field                                 */
    float       *Pn;         /* 2-D only: new field out, P left as it was
                                (NULL: in place over P)                       */
    const float *vel;        /* (v dt)^2, padded                              */
    const float *rho;        /* variable density: {rho, drho_z, drho_x, drho_y}
                                per node (ld pitch), else NULL              */
    int n1, n2, n3;          /* padded z, x, y                                */
    int ld;                  /* row stride in floats (n1 rounded up to 32)    */
    int zsep;                /* z-high ABC done by a separate pass            */
    float c0;
    float cz[AWEFD_MAX_NOP], cx[AWEFD_MAX_NOP], cy[AWEFD_MAX_NOP];
    float bz[AWEFD_MAX_NOP], bx[AWEFD_MAX_NOP], by[AWEFD_MAX_NOP];
    /* tiling */
    int tz, ntz, ntx, ychunk, nty, nblk; /* ychunk: longest y chunk      */
    int col;                 /* 3-D: the column kernel (one wave per block)    */
    int narrow;              /* 2-D: the one-wave tile of AWEFD_RX2 x rows     */
    int ztail;               /* column kernel: the last z tile (n1 % 64 <= 16
                                nodes) as 16-lane rows, 4 x tiles a wave */
    /* sponge: per-axis factors, 1 outside the strips */
    int dabc, nb;
    const float *sp1, *sp2, *sp3;
    /* 3-D one-way ABC; the same tables (2-D: bz*, bx*) for the hybrid
     * boundary, whose transition strip blends by damp[nb - nop] */
    int abc, fsrf;
    const float *bzl, *bzh, *bxl, *bxh, *byl, *byh;
    const float *damp;
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
    const awefd_sinc_pt *rsinc;  /* sinc=y: the receivers' taps (else NULL)  */
    float *trow;             /* traces + it*nr (batched: the traces base)     */
    int blk0;                /* batched: first block of this shot             */
    /* batched recompute of a checkpoint segment: c_n of segment k is
     * ck + 2k nal at n = kK (c_{kK-1} at ck + (2k+1) nal), else
     * seg + (n - kK - 1) nal */
    const float *ck;
    float *seg;
    long nal;
    int K;
    /* batched two-step pass (2-D): the shot's four field buffers, its tiles
     * of (64 - 2 nop) z x (rows - 2 nop) x outputs and their injection
     * buckets (every source node inside a tile's first-step region) */
    float *F[4];
    int ntz2, ntx2, nblk2, blk02;
    const int *ibeg2;
    const awefd_inj *inj2;
    /* batched three-step pass: receivers per tile, by the tile owning their
     * first interpolation corner (trace row it + 1 is extracted in-block) */
    const int *rbeg3;
    const awefd_rx3 *rx3;
} awefd_cart_args;

/* Adjoint of the cartesian 2-D loop (adj_kernels.hip). */
#define AWEFD_ADJ_NWAVE 4        /* waves per block of the adjoint step kernel   */
#ifndef AWEFD_ADJ_RX
#define AWEFD_ADJ_RX 2           /* x rows per wave of the adjoint step kernel   */
#endif
#define AWEFD_ADJ_TX (AWEFD_ADJ_NWAVE * AWEFD_ADJ_RX)
#ifndef AWEFD_ADJ_RXB
#define AWEFD_ADJ_RXB 4          /* x rows per wave of the batched adjoint kernel */
#endif
#define AWEFD_ADJ_TXB (AWEFD_ADJ_NWAVE * AWEFD_ADJ_RXB)
/* The two-step adjoint pass: blocks of AWEFD_ADJ2_NWAVE waves x
 * AWEFD_ADJ2_RX rows over 64 z; a tile's outputs are its inner
 * (64 - 2 nop) x (rows - 2 nop) nodes.  4 x 8 (126 VGPRs, 4 waves per
 * SIMD): the backward pass of 16 shots of 1100 x 2100 nodes 12 % faster
 * than with one adjoint step per launch; 3 x 8 and 4 x 6 tie, 8 x 4 and
 * 8 x 8 gain nothing, 6-wave and 16-wave blocks lose 10-30 %. */
#ifndef AWEFD_ADJ2_NWAVE
#define AWEFD_ADJ2_NWAVE 4
#endif
#ifndef AWEFD_ADJ2_RX
#define AWEFD_ADJ2_RX 8
#endif
#define AWEFD_ADJ2_ROWS (AWEFD_ADJ2_NWAVE * AWEFD_ADJ2_RX)
typedef struct {
    /* per-step pointers follow from the step n: alpha_n lives in A[n % 4];
     * c_n is checkpoint k = n / K (ck + 2k nal) at n = kK, else segment
     * buffer seg + (n - kK - 1) nal; residual row n at res + n nr */
    float *A[4];
    const float *ck, *seg;
    long nal;                /* floats per field                              */
    int K;
    const float *res;        /* residual rows (nt x nr)                       */
    int nr;
    const float *V;          /* (v dt)^2                                      */
    float *gV;               /* dJ/dV accumulator (padded grid)               */
    float *gout;             /* dJ/ds output of the chain (model grid)        */
    int nz, nx, ld, mz, mx, nb, dabc;
    float idt2;              /* 1/dt^2                                        */
    const float *sp1, *sp2;  /* sponge factors per axis (ones without dabc)   */
    float c0, cz[AWEFD_MAX_NOP], cx[AWEFD_MAX_NOP];
    int ntz, ntx, nblk;      /* 64 x AWEFD_ADJ_TX tiles                       */
    /* receiver and injection nodes, bucketed per tile column like the
     * forward injection (src = receiver index, or wavelet column) */
    const int *rbeg, *sbeg;
    const awefd_inj *rinj, *sinj;
    const float *wav;
    long wstride;
    int blk0;                /* batched: first block of this shot             */
    /* the two-step pass: its tiles, the receiver and injection nodes of
     * each tile's whole 64 x AWEFD_ADJ2_ROWS extent (overlapping tiles share
     * nodes), and (batched) this shot's first block */
    int ntz2, ntx2, nblk2, blk02;
    const int *rbeg2, *sbeg2;
    const awefd_inj *rinj2, *sinj2;
} awefd_adj2d_args;

int awefd_launch_adj2d_step(const awefd_adj2d_args *a, int n, int nop, void *stream);
/* shots: device array of nshot argument records; bshot: block -> shot */
int awefd_launch_adj2d_batch(const awefd_adj2d_args *shots, const int *bshot, int nblk, int n, int nop,
                             void *stream);
/* backward steps n and n-1 of every shot in one pass (n-1 and n in one
 * checkpoint segment); bshot: block -> shot of the two-step tiles */
int awefd_launch_adj2d_batch2(const awefd_adj2d_args *shots, const int *bshot, int nblk, int n, int nop,
                              void *stream);
/* gout (model grid) = sum over the shots of dJ/ds; h: host copy of shot 0's
 * record (grid sizes), shots: device records */
int awefd_launch_adj2d_chain(const awefd_adj2d_args *h, const awefd_adj2d_args *shots, int nshot, float *gout,
                             void *stream);
/* dst_j = src_j (or 0 where src_j is NULL) for every job of a table:
 * cbeg[j] .. cbeg[j+1] are job j's 1024-float chunks */
typedef struct {
    const float *src;
    float *dst;
    long n;
} awefd_copy_job;
int awefd_launch_copy_jobs(const awefd_copy_job *jobs, const long *cbeg, int njob, long nchunk, void *stream);

/* n / d = (umulhi(n, m) + n) >> l for every n < 2^31 (awefd_magic_div). */
typedef struct {
    unsigned m;
    int l;
} awefd_udiv;
awefd_udiv awefd_magic_div(unsigned d);

/* Polar (2-D) and spherical (3-D) kernels.  Axes: 1 = r (fastest), 2 = theta,
 * 3 = phi.  Nodes within AWEFD_SPHER_SHELL layers of a face form the shell:
 * the step kernel stores their values before the one-way boundary (S) and the
 * shell kernel finishes them. */
#define AWEFD_SPHER_SHELL 6
#define AWEFD_SPHER_NOP 4         /* interior band and one-way depth (spher_utils.c:326) */
#define AWEFD_SPHER_MAXCHAIN 16   /* longest chain of one-way reads of final values */
#ifndef AWEFD_SPHER_RX
#define AWEFD_SPHER_RX 4          /* theta rows per wave of the 3-D step kernel */
#endif
#ifndef AWEFD_SPHER_NWAVE
#define AWEFD_SPHER_NWAVE 4
#endif
#define AWEFD_SPHER_TX (AWEFD_SPHER_NWAVE * AWEFD_SPHER_RX)
#ifndef AWEFD_SPHER_CRX
#define AWEFD_SPHER_CRX 2         /* theta rows per wave of the 3-D column kernel */
#endif

typedef struct {
    float       *P;          /* current field (after this step's injection)   */
    float       *Q;          /* previous field in, new field out (in place)   */
    float       *S;          /* shell nodes before the one-way boundary       */
    const float *vv;         /* (dt*dt)*(v*v), padded                         */
    int n1, n2, n3;          /* padded r, theta, phi (n3 = 1 in 2-D)          */
    int ld;                  /* row stride in floats                          */
    int ndim;
    /* derivative divisors, as the reference forms them in float */
    float tdr, dr2, tdt, dt2, dp2;  /* 2*dr, dr*dr, 2*dth, dth*dth, dph*dph      */
    /* their reciprocals in double, for x / c = (float)((double)x * (1/c));
     * fastdiv = 0: the kernels divide (awefd__exact_recip) */
    double rtdr, rdr2, rtdt, rdt2, rdp2;
    int fastdiv;
    /* 3-D metric: per r 2/r and 1/(r*r) (float); per (theta, r)
     * cos/((double)(r*r)*sin) and 1/((double)(r*r)*sin*sin) */
    const float *mr2, *mr1;
    const double *mt1, *mt2;
    /* 2-D polar coefficients per r: r-1, centre-r, r+1, theta+-1, centre-theta */
    const float *pc;         /* 5 floats per r                                */
    /* sponge: damp[i], i = depth into the strip; fsrf */
    int nb, fsrf;
    const double *damp;
    /* one-way tables: theta faces [phi][r], r faces [phi][theta] (the table
     * each rule uses), phi faces [theta][r] */
    const float *bthl, *bthh, *brl, *brh, *bphl, *bphh;
    /* step tiling; col: the 3-D column kernel (one wave per block) */
    int ntz, ntx, nty, nblk, col;
    /* shell enumeration: counts of the r-face, theta-face and phi-face
     * groups (below 2^31 in all), and the divisions that decode a node
     * number: by the r-face layer count, n2, the inner r count, the
     * theta-face layer count and the inner theta count */
    unsigned nsr, nst, nsp;
    awefd_udiv dv_rc, dv_n2, dv_rin, dv_tc, dv_tin;
    /* injection: nodes (flat offsets) with their contributions in order */
    int ninode;
    const long *inode;
    const int *ibeg;
    const float *iw;         /* weight (times reflectivity when expl)         */
    const float *wav;
    int it;
    /* extraction */
    int nr;
    const int *rj;           /* (jr, jtheta, jphi) per receiver               */
    const float *rw;
    float *trow;
} awefd_spher_args;

/* Launchers (cart_kernels.hip).  Return 0 or a HIP error code. */
int awefd_launch_cart_step(const awefd_cart_args *a, int nop, int ndim, int sse, int varden,
                           void *stream);
int awefd_launch_cart_abc_xy(const awefd_cart_args *a, int nop, void *stream);
/* hybrid boundary of the new field P (apply_abc): z lines, x lines, then in
 * 3-D y lines, each face of a line swept by one thread */
int awefd_launch_cart_hybrid(const awefd_cart_args *a, int nop, int ndim, void *stream);
int awefd_launch_fill(float *p, long n, float v, void *stream);
/* dvel (padded, row stride ld) <- v (v dt2) of the edge-replicated model
 * velocity d_vel (n[0] fastest), zero for z <= fsnb (fsnb < 0: no free
 * surface) */
int awefd_launch_set_vel(const float *d_vel, float *dvel, const int n[3], const int nbs[3], int ld, float dt2,
                         int fsnb, void *stream);
/* Batched 2-D step over many shots (cart_kernels.hip): shots = device array
 * of per-shot arguments, bshot = shot of every block, rx in {2,4,8,16}. */
/* Batched 2-D steps it (and it+1 when two) of every shot in one pass:
 * roles packs the indices into F of the previous and current field and of
 * the outputs (bits 0-1, 2-3, 4-5, 6-7); extraction blocks write trace rows
 * it (current field) and it-1 (previous field, it > 0). */
int awefd_launch_cart_batch_t2(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int roles, int two,
                               int nop, int sse, int rx, void *stream);
/* Batched 2-D steps it, it+1, it+2 of every shot in one pass (nst = 3), or
 * step it alone (nst = 1) on the same tiles: roles as above (outputs: the
 * second and third steps' fields); rx in {4, 8}, AWEFD_T3_NWAVE(rx) waves
 * per block. */
int awefd_launch_cart_batch_t3(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int roles, int nst,
                               int nop, int sse, int rx, void *stream);
/* recompute step n of every shot inside its checkpoint segment (out of
 * place, no extraction), same tables as awefd_launch_cart_batch2 */
int awefd_launch_cart_batch_rc(const awefd_cart_args *shots, const int *bshot, int nblk, int n, int nop,
                               int sse, int rx, void *stream);
/* recompute steps n, n+1, n+2 of every shot inside its checkpoint segment in
 * one pass (out of place, all three fields kept, no extraction); the tables
 * of awefd_launch_cart_batch_t3, and n + 3 stays inside the segment */
int awefd_launch_cart_batch_t3rc(const awefd_cart_args *shots, const int *bshot, int nblk, int n, int nop,
                                 int sse, int rx, void *stream);
int awefd_launch_cart_batch2(const awefd_cart_args *shots, const int *bshot, int nblk, int it, int nop,
                             int sse, int rx, void *stream);
/* resident blocks per CU of the cartesian step kernel (0 if unknown) */
int awefd_cart_step_occupancy(int nop, int ndim, int sse);

/* Launchers (spher_kernels.hip). */
int awefd_launch_spher_inject(const awefd_spher_args *a, void *stream);
int awefd_launch_spher_step(const awefd_spher_args *a, void *stream);
int awefd_launch_spher_shell(const awefd_spher_args *a, void *stream);
int awefd_launch_spher_extract(const awefd_spher_args *a, void *stream);
int awefd_spher_step_occupancy(int ndim);

#ifdef __cplusplus
}
#endif
#endif
