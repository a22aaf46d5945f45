/*
 * Host-side set-up arithmetic of libawefd: finite-difference coefficients,
 * padding, index maps, sponge and one-way coefficient tables, injection
 * buckets.  Everything here must round exactly like the reference's C code,
 * so this file is compiled by gcc with -ffp-contract=off and no fast-math,
 * and every expression keeps the reference's float/double types.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "awefd_setup.h"

#define SETUP_PI 3.14159265358979323846264338328

/* MAWEFDcart2Dcpu.c:476-484 */
static int fact(int n)
{
    int r = 1;
    for (int i = 2; i <= n; i++) r *= i;
    return r;
}

/* Optimized coefficients of optfd=y (Liu 2013 second derivative; Zhang & Yao
 * 2013 first derivative), the values the reference tabulates at
 * MAWEFDcart2Dcpu.c:508-549, rounded to float as it stores them. */
static const float OPT_D2[7][7] = {
    {0},
    {0.f, 1.f},
    {0.f, 1.369074f, -(float)0.09266816},
    {0.f, 1.573661f, -0.1820268f, 0.01728053f},
    {0.f, 1.700010f, -0.2554615f, 0.04445392f, -0.004946851f},
    {0.f, 1.782836f, -0.3124513f, 0.07379487f, -0.01532122f, 0.001954439f},
    {0.f, 1.837023f, -0.3538895f, 0.09978343f, -0.02815486f, 0.006556587f, -0.0009405699f},
};
static const float OPT_D1[7][7] = {
    {0},
    {0.f, 0.5f},
    {0.f, (float)0.67880327, -(float)0.08962729},
    {0.f, (float)0.77793115, -(float)0.17388691, (float)0.02338713},
    {0.f, (float)0.84149635, -(float)0.24532989, (float)0.06081891, -(float)0.00839807},
    {0.f, (float)0.88414717, -(float)0.30233648, (float)0.10275057, -(float)0.02681517,
     (float)0.00398089},
    {0.f, (float)0.91067892, -(float)0.34187892, (float)0.13833962, -(float)0.04880710,
     (float)0.01302148, -(float)0.00199047},
};

/* Binomial-window weights (Chu & Stoffa 2012), MAWEFDcart2Dcpu.c:486-500:
 * c_n = -2/n^2 cos(n pi) (nop!)^2 / ((nop+n)! (nop-n)!); the first
 * derivative weight is c_n n / 2.  Types as the reference: the float
 * quotient -2.f/(n*n) times the double cosine, rounded to float. */
static void binomial_weights(int nop, int order, float *c)
{
    int hf = fact(nop);
    hf *= hf;
    c[0] = 0.f;
    for (int n = 1; n <= nop; n++) {
        const float lead = -2.f / (float)(n * n);
        c[n] = (float)((double)lead * cos(n * SETUP_PI) * (double)hf / (double)fact(nop + n)
                       / (double)fact(nop - n));
        if (order == 1) c[n] = c[n] * (float)n / 2.f;
    }
}

int awefd_fdcoef(int nop, int ndim, const float *d, int optfd, int order, float *out)
{
    float c[AWEFD_SETUP_MAX_NOP + 1];
    if (nop < 1 || nop > AWEFD_SETUP_MAX_NOP || ndim < 2 || ndim > 3) return -1;
    if (optfd) memcpy(c, order == 2 ? OPT_D2[nop] : OPT_D1[nop], sizeof(float) * (nop + 1));
    else binomial_weights(nop, order, c);
    out[0] = 0.f;
    for (int a = 0; a < ndim; a++) {
        const float h = order == 2 ? d[a] * d[a] : d[a];
        for (int k = 1; k <= nop; k++) {
            out[a * nop + k] = c[k] / h;
            if (order == 2) out[0] += out[a * nop + k];
        }
    }
    if (order == 2) out[0] *= -2.0f;
    return 0;
}

/* fdutil.c:2181-2203 */
void awefd_sponge_weights(int nb, float *w)
{
    const float sb = (float)(4.0 * nb);
    for (int ib = 0; ib < nb; ib++) {
        const float fb = (float)(ib / (sqrt(2.0) * (double)sb));
        const float arg = -fb * fb;
        w[ib] = (float)exp((double)arg);
    }
}

/* Per-axis sponge factor of padded node i on an axis of n nodes: the strip
 * value w[nb-1-ib] for the ib-th node from either face, 1 elsewhere
 * (fdutil.c:2206-2246). */
void awefd_sponge_axis(const float *w, int nb, int n, float *f)
{
    for (int i = 0; i < n; i++) f[i] = 1.0f;
    for (int ib = 0; ib < nb && ib < n; ib++) {
        f[ib] = w[nb - ib - 1];
        f[n - ib - 1] = w[nb - ib - 1];
    }
}

/* Division by d (1 <= d < 2^31) as a multiply-high, an add and a shift,
 * exact for every n < 2^31 (Granlund and Montgomery, "Division by invariant
 * integers using multiplication", PLDI 1994, the round-up multiplier):
 * l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
 * n / d = (umulhi(n, m) + n) >> l; the sum stays below 2^32. */
awefd_udiv awefd_magic_div(unsigned d)
{
    awefd_udiv r;
    int l = 0;
    while (l < 31 && (1u << l) < d) l++;
    r.l = l;
    r.m = (unsigned)(((1ULL << 32) * ((1ULL << l) - d)) / d + 1);
    return r;
}

/* 1/c in double such that (float)((double)x * (1/c)) is the IEEE float
 * quotient x / c for every float x, or 0 when that does not hold for c.
 *
 * The double product differs from x / c by at most (2u + u^2) relative
 * (u = 2^-53: one rounding of 1/c, one of the product).  A quotient q in the
 * normal range is never a rounding midpoint mu of float (x = c mu would need
 * 25 significant bits), and |q - mu| >= 2^(E-48) for the midpoints of its
 * binade [2^E, 2^(E+1)): x - c mu is a non-zero multiple of 2^(ec+E-24) for
 * c = C 2^ec, so dividing by c < 2^(ec+24) leaves at least 2^(E-48).  The
 * product's error, below 2^(E-50.9), keeps it on q's side of every midpoint,
 * and rounding it to float gives q's rounding.  Below 2^-126 the midpoints
 * are the odd multiples of 2^-150; a subnormal quotient that is not one lies
 * at least 2^min(ex, ec-150) / c from each (x = X 2^ex), more than the
 * product's error (2^-177.9 when ex > ec - 150, else 2^(ex-ec-50.9)).  It
 * can be one, though (x = 5 2^-149, c = 10), and then the
 * product must round to that midpoint exactly for the ties-to-even rule to
 * agree; those quotients are checked here one by one: every
 * mu = (2m+1) 2^-150 below 2^-126 for which c mu (exact in double) is a
 * float. */
double awefd__exact_recip(float c)
{
    if (!(c > 0.0f) || !isfinite(c) || !isnormal(c)) return 0.0;
    const double rc = 1.0 / (double)c;
    /* c = C 2^ec with C odd; x = c mu = C (2m+1) 2^(ec-150) is a float when
     * the odd C (2m+1) has at most 24 bits and, below 2^-126, when it is a
     * multiple of 2^-149 */
    int e;
    uint64_t C = (uint64_t)ldexp(frexp((double)c, &e), 24);
    int ec = e - 24;
    while (!(C & 1)) {
        C >>= 1;
        ec++;
    }
    /* c a power of two: both forms are exact scalings.  ec < 1: a float x
     * below 2^-126 is a multiple of 2^-149 and C (2m+1) 2^(ec-150) is not,
     * and one above needs C (2m+1) >= 2^(24-ec) > 2^24 bits; no candidate. */
    if (C == 1 || ec < 1) return rc;
    for (uint64_t k = 1; C * k < (1u << 24); k += 2) {
        const double xd = ldexp((double)(C * k), ec - 150);
        const float x = (float)xd;                          /* exact */
        const volatile float q = x / c;                     /* the IEEE quotient, on the host */
        const float qv = q, f = (float)((double)x * rc);
        uint32_t qb, fb;
        memcpy(&qb, &qv, 4);
        memcpy(&fb, &f, 4);
        if (qb != fb) return 0.0;
    }
    return rc;
}

/* Spherical sponge factor, spher_utils.c:665-667 (double). */
double awefd_spher_damp(int i, int nb)
{
    const float alpha = 0.90f;
    return exp(-1.0 * fabs((pow((i - 1.0), 2) * log((double)alpha)) / (pow(nb, 2))));
}

/* Edge-replicating pad (expand2d/expand3d, fdutil.c:435-571; expand_cpu_2d/3d,
 * spher_utils.c:13-103): padded node (i1,i2,i3) takes model node
 * (clamp(i1-nb1), clamp(i2-nb2), clamp(i3-nb3)). */
void awefd_pad(const float *a, float *b, const int n[3], const int nbs[3])
{
    const long p1 = n[0] + 2L * nbs[0], p2 = n[1] + 2L * nbs[1], p3 = n[2] + 2L * nbs[2];
    for (long i3 = 0; i3 < p3; i3++) {
        long k3 = i3 - nbs[2];
        k3 = k3 < 0 ? 0 : (k3 >= n[2] ? n[2] - 1 : k3);
        for (long i2 = 0; i2 < p2; i2++) {
            long k2 = i2 - nbs[1];
            k2 = k2 < 0 ? 0 : (k2 >= n[1] ? n[1] - 1 : k2);
            const float *src = a + (k3 * n[1] + k2) * n[0];
            float *dst = b + (i3 * p2 + i2) * p1;
            for (long i1 = 0; i1 < p1; i1++) {
                long k1 = i1 - nbs[0];
                k1 = k1 < 0 ? 0 : (k1 >= n[0] ? n[0] - 1 : k1);
                dst[i1] = src[k1];
            }
        }
    }
}

/* lint2d_make / lint3d_make (fdutil.c:1363-1411, 1474-1539), generalised to
 * ndim axes.  coords holds (x[,y],z) tuples as the sou=/rec= files do; the
 * grid arguments are (z, x, y).  Float arithmetic and truncation exactly as
 * the reference; out-of-grid points map to node 0 with the weight on z+1. */
void awefd_lint(int ndim, int na, const float *coords, const int npad[3], const float opad[3],
                const float d[3], int *j /* 3*na */, float *w /* 8*na */)
{
    for (int a = 0; a < na; a++) {
        float x, y = 0.f, z;
        if (ndim == 2) {
            x = coords[2 * a];
            z = coords[2 * a + 1];
        } else {
            x = coords[3 * a];
            y = coords[3 * a + 1];
            z = coords[3 * a + 2];
        }
        const float zmax = opad[0] + (float)(npad[0] - 1) * d[0];
        const float xmax = opad[1] + (float)(npad[1] - 1) * d[1];
        int inside = z >= opad[0] && z < zmax && x >= opad[1] && x < xmax;
        if (ndim == 3) {
            const float ymax = opad[2] + (float)(npad[2] - 1) * d[2];
            inside = inside && y >= opad[2] && y < ymax;
        }
        float f1 = 1.f, f2 = 0.f, f3 = 0.f;
        int jz = 0, jx = 0, jy = 0;
        if (inside) {
            jz = (int)((z - opad[0]) / d[0]);
            jx = (int)((x - opad[1]) / d[1]);
            f1 = (z - opad[0]) / d[0] - (float)jz;
            f2 = (x - opad[1]) / d[1] - (float)jx;
            if (ndim == 3) {
                jy = (int)((y - opad[2]) / d[2]);
                f3 = (y - opad[2]) / d[2] - (float)jy;
            }
        }
        j[3 * a] = jz;
        j[3 * a + 1] = jx;
        j[3 * a + 2] = jy;
        float *wa = w + 8 * a;
        if (ndim == 2) {
            wa[0] = (1 - f1) * (1 - f2);
            wa[1] = (f1) * (1 - f2);
            wa[2] = (1 - f1) * (f2);
            wa[3] = (f1) * (f2);
            wa[4] = wa[5] = wa[6] = wa[7] = 0.f;
        } else {
            wa[0] = (1 - f3) * (1 - f1) * (1 - f2);
            wa[1] = (1 - f3) * (f1) * (1 - f2);
            wa[2] = (1 - f3) * (1 - f1) * (f2);
            wa[3] = (1 - f3) * (f1) * (f2);
            wa[4] = (f3) * (1 - f1) * (1 - f2);
            wa[5] = (f3) * (f1) * (1 - f2);
            wa[6] = (f3) * (1 - f1) * (f2);
            wa[7] = (f3) * (f1) * (f2);
        }
    }
}

/* (1-d)/(1+d) with d = v dt / h in float (abcone3d_make, fdutil.c:1995-2015) */
float awefd_oneway_cart(float v, float dt, float h)
{
    const float dd = v * dt / h;
    return (1 - dd) / (1 + dd);
}

/* (1-d)/(1+d) with d = v (dt / h) in float (MAWEFDspher3Dcpu.c:245-270) */
float awefd_oneway_spher(float v, float dt, float h)
{
    const float dd = v * (dt / h);
    return (1 - dd) / (1 + dd);
}

static int inj_cmp(const void *pa, const void *pb)
{
    const awefd_inj_key *a = (const awefd_inj_key *)pa, *b = (const awefd_inj_key *)pb;
    if (a->bucket != b->bucket) return a->bucket < b->bucket ? -1 : 1;
    if (a->e.y != b->e.y) return a->e.y < b->e.y ? -1 : 1;
    if (a->e.z != b->e.z) return a->e.z < b->e.z ? -1 : 1;
    if (a->e.x != b->e.x) return a->e.x < b->e.x ? -1 : 1;
    if (a->e.src != b->e.src) return a->e.src < b->e.src ? -1 : 1;
    return 0;
}

/* Sort injection contributions by (bucket, y, z, x, source) so that, at every
 * node, contributions are added in source order as the reference's serial
 * loop adds them; ibeg[b]..ibeg[b+1] is bucket b. */
void awefd_sort_inj(awefd_inj_key *keys, long n, int nbucket, awefd_inj *out, int *ibeg)
{
    qsort(keys, (size_t)n, sizeof(*keys), inj_cmp);
    for (int b = 0; b <= nbucket; b++) ibeg[b] = 0;
    for (long i = 0; i < n; i++) {
        out[i] = keys[i].e;
        ibeg[keys[i].bucket + 1]++;
    }
    for (int b = 0; b < nbucket; b++) ibeg[b + 1] += ibeg[b];
}

/* ------------------------------------------------------------------ */
/* 8-point sinc taps of sinc2d_make / sinc3d_make (fdutil.c:784-842,
 * 1088-1145).  Their weights are the response of ints8r -- the CWP/SU
 * library's table-driven 8-point least-squares sinc interpolator (intsinc8.c,
 * mksinc.c, stoep.c), which the reference links from Madagascar -- to a unit
 * impulse.  The method: for a fractional shift d the 8 taps s solve the
 * symmetric Toeplitz normal equations
 *     sum_k sinc(fmax (j - k)) s_k = sinc(fmax (3 - j + d)),
 *     fmax = min(0.066 + 0.265 ln 8, 1),
 * by Levinson's recursion in double; shifts are tabulated at k/512 and an
 * output point uses the row nearest its fractional position. */
enum { SINC_TAPS = 8, SINC_ROWS = 513 };

static double sinc_norm(double x)
{
    if (x == 0.0) return 1.0;
    const double px = 3.141592653589793 * x;
    return sin(px) / px;
}

/* taps of one fractional shift d (Levinson, with the prediction-error
 * filter e and its power v updated in the order of stoep.c) */
static void sinc_taps(float d, float *out)
{
    double r[SINC_TAPS], g[SINC_TAPS], f[SINC_TAPS], e[SINC_TAPS];
    double fmax = 0.066 + 0.265 * log((double)SINC_TAPS);
    fmax = fmax < 1.0 ? fmax : 1.0;
    for (int j = 0; j < SINC_TAPS; j++) {
        r[j] = sinc_norm(fmax * j);
        g[j] = sinc_norm(fmax * (SINC_TAPS / 2 - j - 1 + d));
    }
    e[0] = 1.0;
    double v = r[0];
    f[0] = g[0] / r[0];
    for (int j = 1; j < SINC_TAPS; j++) {
        e[j] = 0.0;
        f[j] = 0.0;
        double dot = 0.0;
        for (int i = 0; i < j; i++) dot += e[i] * r[j - i];
        double k = dot / v;
        v -= k * dot;
        for (int i = 0; i <= j / 2; i++) {
            const double mirror = e[j - i] - k * e[i];
            e[i] -= k * e[j - i];
            e[j - i] = mirror;
        }
        double fr = 0.0;
        for (int i = 0; i < j; i++) fr += f[i] * r[j - i];
        k = (fr - g[j]) / v;
        for (int i = 0; i <= j; i++) f[i] -= k * e[j - i];
    }
    for (int j = 0; j < SINC_TAPS; j++) out[j] = (float)f[j];
}

static float sinc_rows[SINC_ROWS][SINC_TAPS];
static pthread_once_t sinc_once = PTHREAD_ONCE_INIT;

static void sinc_build(void)
{
    for (int k = 1; k < SINC_ROWS - 1; k++) sinc_taps((float)k / (float)(SINC_ROWS - 1), sinc_rows[k]);
    memset(sinc_rows[0], 0, sizeof sinc_rows[0]);
    memset(sinc_rows[SINC_ROWS - 1], 0, sizeof sinc_rows[0]);
    sinc_rows[0][SINC_TAPS / 2 - 1] = 1.f;
    sinc_rows[SINC_ROWS - 1][SINC_TAPS / 2] = 1.f;
}

/* ints8r(9, 1, -4, impulse at the centre, 0, 0, 9, xo, w): output x uses
 * the taps from int(12 + x) - 11 on, table row round(frac * 512); the impulse
 * (input 4) contributes the row's weight at tap 4 - first, when in range. */
static void sinc_impulse_response(const float *xo, float *w)
{
    pthread_once(&sinc_once, sinc_build);
    const float rows1 = (float)(SINC_ROWS - 1);
    for (int i = 0; i < 9; i++) {
        const float xn = 12.0f + xo[i] * 1.0f;
        const int ixn = (int)xn;
        const int first = ixn - 11;
        const float frac = xn - (float)ixn;
        const int row = frac >= 0.0 ? (int)(frac * rows1 + 0.5) : (int)((frac + 1.0) * rows1 - 0.5);
        const int t = 4 - first;
        w[i] = t >= 0 && t < SINC_TAPS ? 1.0f * sinc_rows[row][t] : 0.0f;
    }
}

/* one axis: q = (node - point)/h as the reference computes it; the z axis
 * forms its 9 output positions in double (-4.0 + q + i), x and y in float */
void awefd_sinc_axis(float q, int zaxis, float w[9], int *f, int *n)
{
    float xo[9];
    for (int i = 0; i < 9; i++)
        xo[i] = zaxis ? (float)(-4.0 + (double)q + (double)(i * 1.0f)) : -4.0f + q + i * 1.0f;
    sinc_impulse_response(xo, w);
    const int on_node = w[4] == 1.0f;
    *f = on_node ? 4 : 0;
    *n = on_node ? 1 : 9;
}

/* Sinc records of na points (cart2d (x,z), cart3d (x,y,z) coordinates) on
 * a padded grid (z, x, y order).  Returns -1 when a tap of a point leaves
 * the padded grid (the reference would index outside its arrays). */
int awefd_sinc_points(int ndim, int na, const float *coords, const int npad[3], const float opad[3],
                      const float d[3], awefd_sinc_pt *out)
{
    for (int a = 0; a < na; a++) {
        const float *p = coords + (long)ndim * a;
        awefd_sinc_pt *s = &out[a];
        const float pos[3] = {p[ndim - 1], p[0], ndim == 3 ? p[1] : 0.f};
        for (int ax = 0; ax < 3; ax++) {
            if (ax >= ndim) {           /* 2-D: one tap of weight 1 on y = 0 */
                s->c[ax] = 0;
                s->f[ax] = 4;
                s->n[ax] = 1;
                for (int i = 0; i < 9; i++) s->w[ax][i] = i == 4 ? 1.f : 0.f;
                continue;
            }
            /* fdutil.c:802-806,815-819,828-832 */
            const int c = (int)((pos[ax] - opad[ax]) / d[ax] + 0.499f);
            const float off = c * d[ax] + opad[ax] - pos[ax];
            s->c[ax] = c;
            awefd_sinc_axis(off / d[ax], ax == 0, s->w[ax], &s->f[ax], &s->n[ax]);
            const int lo = c - 4 + s->f[ax], hi = lo + s->n[ax] - 1;
            if (lo < 0 || hi >= npad[ax]) return -1;
        }
    }
    return 0;
}
