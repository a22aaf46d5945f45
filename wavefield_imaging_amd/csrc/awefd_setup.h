/* Host-side set-up arithmetic of libawefd (awefd_setup.c). */
#ifndef AWEFD_SETUP_H
#define AWEFD_SETUP_H

#include "awefd_internal.h"

#define AWEFD_SETUP_MAX_NOP 6

typedef struct {
    int bucket;
    awefd_inj e;
} awefd_inj_key;

int    awefd_fdcoef(int nop, int ndim, const float *d, int optfd, int order, float *out);
void   awefd_sponge_weights(int nb, float *w);
void   awefd_sponge_axis(const float *w, int nb, int n, float *f);
double awefd_spher_damp(int i, int nb);
void   awefd_pad(const float *a, float *b, const int n[3], const int nbs[3]);
void   awefd_lint(int ndim, int na, const float *coords, const int npad[3],
                  const float opad[3], const float d[3], int *j, float *w);
float  awefd_oneway_cart(float v, float dt, float h);
float  awefd_oneway_spher(float v, float dt, float h);
void   awefd_sort_inj(awefd_inj_key *keys, long n, int nbucket, awefd_inj *out, int *ibeg);
double awefd__exact_recip(float c);
void   awefd_sinc_axis(float q, int zaxis, float w[9], int *f, int *n);
int    awefd_sinc_points(int ndim, int na, const float *coords, const int npad[3], const float opad[3],
                         const float d[3], awefd_sinc_pt *out);

#endif
