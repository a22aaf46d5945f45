/*
 * CPU oracle (test infrastructure only) for the acoustic FD time loop.
 * Independent C restatement of the reference CPU programs; see awefd_oracle.h.
 *
 * Shared set-up helpers live here; the propagation loops are instantiated
 * twice from oracle_prop.inc (REAL=float -> *_f32, REAL=double -> *_f64).
 *
 * Build: gcc -O2 -fPIC -ffp-contract=off -fno-fast-math -shared (oracle/Makefile).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "awefd_oracle.h"

#define ORC_PI 3.14159265358979323846264338328

#ifdef _OPENMP
#include <omp.h>
static double orc_now(void) { return omp_get_wtime(); }
int orc_threads(void) { return omp_get_max_threads(); }
#else
#include <time.h>
static double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
int orc_threads(void) { return 1; }
#endif

/* wall time of the last time loop (omp_get_wtime around the loop, as the
 * reference's timing at MAWEFDcart2Dcpu.c:109-114,384-390) */
double orc_loop_seconds = 0.0;
double orc_last_loop_seconds(void) { return orc_loop_seconds; }

/* 1 when the last cartesian 2-D adj=y loop reached its image step nt-kt */
int orc_imaged = 0;
int orc_last_imaged(void) { return orc_imaged; }

/* ------------------------------------------------------------------ */
/* FD coefficients: MAWEFDcart2Dcpu.c:423-551 (3-D: MAWEFDcart3Dcpu.c:404-531) */

static int orc_factorial(int n)          /* MAWEFDcart2Dcpu.c:476-484 */
{
    int r = 1;
    for (int i = 2; i <= n; i++) r *= i;
    return r;
}

/* Published optimized coefficient tables (Liu 2013; Zhang & Yao 2013) used by
 * optfd=y, MAWEFDcart2Dcpu.c:508-549.  Stored as the reference stores them:
 * single-precision values of the printed decimals. */
static const float orc_opt_c[11][11] = {
    {0},
    {0.f, 1.f},
    {0.f, 1.369074f, -(float)0.09266816},
    {0.f, 1.573661f, -0.1820268f, 0.01728053f},
    {0.f, 1.700010f, -0.2554615f, 0.04445392f, -0.004946851f},
    {0.f, 1.782836f, -0.3124513f, 0.07379487f, -0.01532122f, 0.001954439f},
    {0.f, 1.837023f, -0.3538895f, 0.09978343f, -0.02815486f, 0.006556587f, -0.0009405699f},
    {0.f, 1.874503f, -0.3845794f, 0.1215162f, -0.04121749f, 0.01295522f, -0.003313813f,
     0.0005310053f},
    {0.f, 1.901160f, -0.4074304f, 0.1390909f, -0.05318775f, 0.02004823f, -0.006828249f,
     0.001895771f, -0.0003369052f},
    {0.f, 1.919909f, -0.4240446f, 0.1526043f, -0.06322328f, 0.02676005f, -0.01080739f,
     0.003907747f, -0.001158024f, 0.0002240247f},
    {0.f, 1.918204f, -0.4225858f, 0.1514992f, -0.06249474f, 0.02637196f, -0.01066631f,
     0.003915625f, -0.001219872f, 0.0002863976f, -0.00003744830f},
};
static const float orc_opt_b[7][7] = {
    {0},
    {0.0f, 0.5f},
    {0.0f, (float)0.67880327, -(float)0.08962729},
    {0.0f, (float)0.77793115, -(float)0.17388691, (float)0.02338713},
    {0.0f, (float)0.84149635, -(float)0.24532989, (float)0.06081891, -(float)0.00839807},
    {0.0f, (float)0.88414717, -(float)0.30233648, (float)0.10275057, -(float)0.02681517,
     (float)0.00398089},
    {0.0f, (float)0.91067892, -(float)0.34187892, (float)0.13833962, -(float)0.04880710,
     (float)0.01302148, -(float)0.00199047},
};

/* Binomial-window coefficients, MAWEFDcart2Dcpu.c:486-500: float quotient,
 * times a double cosine and integer factorials, rounded back to float. */
static void orc_normal_coef(int nop, int order, float *cc)
{
    int hf = orc_factorial(nop);
    hf *= hf;
    cc[0] = 0.f;
    for (int n = 1; n <= nop; n++) {
        float head = -2.f / (float)(n * n);
        double v = (double)head * cos(n * ORC_PI) * (double)hf
                   / (double)orc_factorial(nop + n) / (double)orc_factorial(nop - n);
        cc[n] = (float)v;
        if (order == 1) cc[n] = cc[n] * (float)n / 2.f;
    }
}

int orc_fdcoef(int nop, int ndim, const float *d, int optfd, int order, float *out)
{
    float c[11];
    if (nop < 1 || nop > 6 || ndim < 2 || ndim > 3) return -1;
    if (order == 2) {
        if (optfd) memcpy(c, orc_opt_c[nop], sizeof(float) * (nop + 1));
        else       orc_normal_coef(nop, 2, c);
        out[0] = 0.f;
        for (int a = 0; a < ndim; a++) {
            float h2 = d[a] * d[a];
            for (int k = 1; k <= nop; k++) {
                out[a * nop + k] = c[k] / h2;
                out[0] += out[a * nop + k];
            }
        }
        out[0] *= -2.0f;
    } else {
        if (optfd) memcpy(c, orc_opt_b[nop], sizeof(float) * (nop + 1));
        else       orc_normal_coef(nop, 1, c);
        out[0] = 0.f;
        for (int a = 0; a < ndim; a++)
            for (int k = 1; k <= nop; k++) out[a * nop + k] = c[k] / d[a];
    }
    return 0;
}

/* Double-precision coefficients for the FP64 oracle (same formulas in double). */
static int orc_fdcoef_d(int nop, int ndim, const float *d, int optfd, int order, double *out)
{
    double c[11];
    if (nop < 1 || nop > 6) return -1;
    int hf = orc_factorial(nop);
    hf *= hf;
    for (int n = 1; n <= nop; n++) {
        if (optfd) c[n] = order == 2 ? orc_opt_c[nop][n] : orc_opt_b[nop][n];
        else {
            c[n] = -2.0 / (n * n) * cos(n * ORC_PI) * hf
                   / orc_factorial(nop + n) / orc_factorial(nop - n);
            if (order == 1) c[n] = c[n] * n / 2.0;
        }
    }
    out[0] = 0.0;
    for (int a = 0; a < ndim; a++)
        for (int k = 1; k <= nop; k++) {
            out[a * nop + k] = order == 2 ? c[k] / ((double)d[a] * d[a]) : c[k] / d[a];
            if (order == 2) out[0] += out[a * nop + k];
        }
    out[0] *= -2.0;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Gaussian sponge table, fdutil.c:2181-2203:
 * sb = (float)(4.0*nb); fb = (float)(ib/(sqrt(2)*sb)); w = (float)exp(-fb*fb). */
void orc_sponge_table(int nb, float *w)
{
    float sb = (float)(4.0 * nb);
    for (int ib = 0; ib < nb; ib++) {
        float fb = (float)(ib / (sqrt(2.0) * (double)sb));
        float e = -fb * fb;
        w[ib] = (float)exp((double)e);
    }
}

/* Spherical sponge factor, spher_utils.c:665-667 (double throughout). */
double orc_spher_damp(int i, int nb)
{
    float alpha = 0.90f;
    return exp(-1.0 * fabs((pow((i - 1.0), 2) * log((double)alpha)) / (pow(nb, 2))));
}

/* ------------------------------------------------------------------ */
/* Bilinear / trilinear index maps, fdutil.c:1363-1411 and 1474-1539.
 * Float arithmetic, truncation toward zero, out-of-grid points pinned to
 * node 0 with the whole weight on the second z node. */
void orc_lint2d(int na, const float *xz, int nzpad, int nxpad, float ozpad,
                float oxpad, float dz, float dx, int *jz, int *jx, float *w)
{
    for (int a = 0; a < na; a++) {
        float x = xz[2 * a], z = xz[2 * a + 1];
        float f1, f2;
        float zmax = ozpad + (float)(nzpad - 1) * dz;
        float xmax = oxpad + (float)(nxpad - 1) * dx;
        if (z >= ozpad && z < zmax && x >= oxpad && x < xmax) {
            jz[a] = (int)((z - ozpad) / dz);
            jx[a] = (int)((x - oxpad) / dx);
            f1 = (z - ozpad) / dz - (float)jz[a];
            f2 = (x - oxpad) / dx - (float)jx[a];
        } else {
            jz[a] = 0; jx[a] = 0; f1 = 1; f2 = 0;
        }
        w[4 * a + 0] = (1 - f1) * (1 - f2);
        w[4 * a + 1] = (f1) * (1 - f2);
        w[4 * a + 2] = (1 - f1) * (f2);
        w[4 * a + 3] = (f1) * (f2);
    }
}

/* ------------------------------------------------------------------ */
/* 8-point sinc weights of sinc2d_make / sinc3d_make (fdutil.c:784-842,
 * 1088-1145).  They come from ints8r, the CWP/SU library's 8-point
 * least-squares sinc interpolator (Dave Hale; intsinc8.c with mksinc.c and
 * the Levinson solver of stoep.c), which Madagascar builds from its su tree;
 * it is not part of the reference.  Restated from that published method:
 *   - mksinc(d): the 8 coefficients s minimising the error of
 *     y(i+d) ~ sum_j s_j y(i+j-3) over frequencies up to
 *     fmax = min(0.066 + 0.265 ln 8, 1) of Nyquist, i.e. the symmetric
 *     Toeplitz system sum_k sinc(fmax (j-k)) s_k = sinc(fmax (3 - j + d)),
 *     solved in double by Levinson's recursion;
 *   - a table of 513 fractional shifts d = k/512 (k = 0 and 512: the unit
 *     impulse on the 4th / 5th tap);
 *   - intt8r: output x -> position xn = 8 + (x - x0)/dx, taps
 *     yin[int(xn) - 11 ...], table row round(frac(xn) 512), samples off
 *     the input taken as yinl / yinr.
 * Parity with the reference is unpinned for these weights: the tests pin
 * the method's own properties (the Toeplitz system, symmetry, impulse
 * rows, interpolation accuracy). */
#define ORC_SINC_L 8
#define ORC_SINC_NT 513

static double orc_dsinc(double x)
{
    if (x == 0.0) return 1.0;
    double px = 3.141592653589793 * x;
    return sin(px) / px;
}

/* Levinson recursion for R f = g, R symmetric Toeplitz with top row r
 * (Claerbout, FGDP p. 57): a = prediction-error filter, v its power. */
static void orc_levinson(int n, const double *r, const double *g, double *f, double *a)
{
    if (r[0] == 0.0) return;
    a[0] = 1.0;
    double v = r[0];
    f[0] = g[0] / r[0];
    for (int j = 1; j < n; j++) {
        a[j] = 0.0;
        f[j] = 0.0;
        double e = 0.0;
        for (int i = 0; i < j; i++) e += a[i] * r[j - i];
        double c = e / v;
        v -= c * e;
        for (int i = 0; i <= j / 2; i++) {
            double hi = a[j - i] - c * a[i];
            a[i] -= c * a[j - i];
            a[j - i] = hi;
        }
        double w = 0.0;
        for (int i = 0; i < j; i++) w += f[i] * r[j - i];
        c = (w - g[j]) / v;
        for (int i = 0; i <= j; i++) f[i] -= c * a[j - i];
    }
}

void orc_mksinc(float d, float *out)
{
    double r[20], g[20], f[20], a[20];
    double fmax = 0.066 + 0.265 * log((double)ORC_SINC_L);
    if (fmax > 1.0) fmax = 1.0;
    for (int j = 0; j < ORC_SINC_L; j++) {
        r[j] = orc_dsinc(fmax * j);
        g[j] = orc_dsinc(fmax * (ORC_SINC_L / 2 - j - 1 + d));
    }
    orc_levinson(ORC_SINC_L, r, g, f, a);
    for (int j = 0; j < ORC_SINC_L; j++) out[j] = (float)f[j];
}

static float orc_stab[ORC_SINC_NT][ORC_SINC_L];
static int orc_stab_ok;

static void orc_sinc_tabulate(void)
{
    if (orc_stab_ok) return;
    for (int k = 1; k < ORC_SINC_NT - 1; k++)
        orc_mksinc((float)k / (float)(ORC_SINC_NT - 1), orc_stab[k]);
    for (int j = 0; j < ORC_SINC_L; j++) orc_stab[0][j] = orc_stab[ORC_SINC_NT - 1][j] = 0.f;
    orc_stab[0][ORC_SINC_L / 2 - 1] = 1.f;
    orc_stab[ORC_SINC_NT - 1][ORC_SINC_L / 2] = 1.f;
    orc_stab_ok = 1;
}

void orc_ints8r(int nxin, float dxin, float fxin, const float *yin, float yinl, float yinr, int nxout,
                const float *xout, float *yout)
{
    orc_sinc_tabulate();
    float xs = 1.0f / dxin, xb = 8.0f - fxin * xs, fnm1 = (float)(ORC_SINC_NT - 1);
    for (int i = 0; i < nxout; i++) {
        float xn = xb + xout[i] * xs;
        int ixn = (int)xn;
        int k = -11 + ixn;
        float frac = xn - (float)ixn;
        int kt = frac >= 0.0 ? (int)(frac * fnm1 + 0.5) : (int)((frac + 1.0) * fnm1 - 0.5);
        const float *t = orc_stab[kt];
        if (k >= 0 && k <= nxin - 8) {
            yout[i] = yin[k] * t[0] + yin[k + 1] * t[1] + yin[k + 2] * t[2] + yin[k + 3] * t[3]
                      + yin[k + 4] * t[4] + yin[k + 5] * t[5] + yin[k + 6] * t[6] + yin[k + 7] * t[7];
        } else {
            float sum = 0.0f;
            for (int j = 0; j < 8; j++, k++) {
                float y = k < 0 ? yinl : k >= nxin ? yinr : yin[k];
                sum += y * t[j];
            }
            yout[i] = sum;
        }
    }
}

/* One axis of a sinc point: q = (node - point)/h as the reference forms it
 * (in float); the z axis builds its offsets in double (fdutil.c:834,1131:
 * -4.0 + dz/dz + i), x and y in float.  w = response of ints8r to a unit
 * impulse at the centre tap; a weight of exactly 1 at the centre (the point
 * on a node) keeps that tap only: *f = 4, *n = 1. */
void orc_sinc_axis(float q, int zaxis, float *w, int *f, int *n)
{
    float inp[9] = {0, 0, 0, 0, 1, 0, 0, 0, 0}, xo[9];
    for (int i = 0; i < 9; i++)
        xo[i] = zaxis ? (float)(-4.0 + (double)q + (double)(i * 1.0f)) : -4.0f + q + i * 1.0f;
    orc_ints8r(9, 1.0f, -4.0f, inp, 0.0f, 0.0f, 9, xo, w);
    *f = 0;
    *n = 9;
    if (w[4] == 1.0f) {
        *f = 4;
        *n = 1;
    }
}

/* Centre node and taps of a point along one axis (fdutil.c:802-808,
 * 1110-1117): i = (int)((c - o)/h + 0.499), q = (i h + o - c)/h. */
static void orc_sinc_node(float c, float o, float h, int zaxis, int *i, float *w, int *f, int *n)
{
    *i = (int)((c - o) / h + 0.499f);
    float dd = *i * h + o - c;
    orc_sinc_axis(dd / h, zaxis, w, f, n);
}

/* Sinc records of na points, 2-D coordinates (x,z) or 3-D (x,y,z): per
 * point rec[0..2] = centre (z, x, y), rec[3..8] = (fz, nz, fx, nx, fy, ny),
 * w[27] = z, x, y weights.  Returns -1 when a tap leaves the padded grid. */
int orc_sinc_make(int ndim, int na, const float *c, const int np[3], const float opad[3], const float h[3],
                  int *rec, float *w)
{
    int bad = 0;
    for (int a = 0; a < na; a++) {
        const float *p = c + (long)ndim * a;
        int *r = rec + 9 * a;
        float *wa = w + 27 * a;
        /* the reference builds y, then x, then z (fdutil.c:802-839) */
        if (ndim == 3) orc_sinc_node(p[1], opad[2], h[2], 0, &r[2], wa + 18, &r[7], &r[8]);
        else { r[2] = 0; r[7] = 4; r[8] = 1; for (int i = 0; i < 9; i++) wa[18 + i] = i == 4; }
        orc_sinc_node(p[0], opad[1], h[1], 0, &r[1], wa + 9, &r[5], &r[6]);
        orc_sinc_node(p[ndim - 1], opad[0], h[0], 1, &r[0], wa, &r[3], &r[4]);
        for (int ax = 0; ax < ndim; ax++) {
            int lo = r[ax] - 4 + r[3 + 2 * ax], hi = lo + r[4 + 2 * ax] - 1;
            if (lo < 0 || hi >= np[ax]) bad = 1;
        }
    }
    return bad ? -1 : 0;
}

void orc_lint3d(int na, const float *xyz, int nzpad, int nxpad, int nypad,
                float ozpad, float oxpad, float oypad, float dz, float dx, float dy,
                int *jz, int *jx, int *jy, float *w)
{
    for (int a = 0; a < na; a++) {
        float x = xyz[3 * a], y = xyz[3 * a + 1], z = xyz[3 * a + 2];
        float f1, f2, f3;
        float zmax = ozpad + (float)(nzpad - 1) * dz;
        float xmax = oxpad + (float)(nxpad - 1) * dx;
        float ymax = oypad + (float)(nypad - 1) * dy;
        if (z >= ozpad && z < zmax && x >= oxpad && x < xmax && y >= oypad && y < ymax) {
            jz[a] = (int)((z - ozpad) / dz);
            jx[a] = (int)((x - oxpad) / dx);
            jy[a] = (int)((y - oypad) / dy);
            f1 = (z - ozpad) / dz - (float)jz[a];
            f2 = (x - oxpad) / dx - (float)jx[a];
            f3 = (y - oypad) / dy - (float)jy[a];
        } else {
            jz[a] = 0; jx[a] = 0; jy[a] = 0; f1 = 1; f2 = 0; f3 = 0;
        }
        w[8 * a + 0] = (1 - f3) * (1 - f1) * (1 - f2);
        w[8 * a + 1] = (1 - f3) * (f1) * (1 - f2);
        w[8 * a + 2] = (1 - f3) * (1 - f1) * (f2);
        w[8 * a + 3] = (1 - f3) * (f1) * (f2);
        w[8 * a + 4] = (f3) * (1 - f1) * (1 - f2);
        w[8 * a + 5] = (f3) * (f1) * (1 - f2);
        w[8 * a + 6] = (f3) * (1 - f1) * (f2);
        w[8 * a + 7] = (f3) * (f1) * (f2);
    }
}

/* ------------------------------------------------------------------ */
/* Spherical injection reflectivity, spher_utils.c:153-173 (2-D) and
 * 222-310 (3-D): |v0-vk|/(v0+vk) for the first neighbour k (fixed search
 * order) whose velocity differs from the injection node; 0 if none. */
static float orc_refl(const float *vel, long a0, const long *offs, int noff)
{
    float v0 = vel[a0];
    for (int k = 0; k < noff; k++) {
        float vk = vel[a0 + offs[k]];
        if (v0 != vk) return (float)fabs((double)((v0 - vk) / (v0 + vk)));
    }
    return 0.f;
}

static float orc_refl2d(const float *vel, int jr, int jth, int nrpad)
{
    /* index = th*nrpad + r ; order (z=th): z+1, x+1, z+1x+1, x-1, z-1x-1, z+1x-1, z-1x+1 */
    long N = nrpad;
    long offs[7] = {N, 1, N + 1, -1, -N - 1, N - 1, -N + 1};
    return orc_refl(vel, (long)jth * N + jr, offs, 7);
}

static float orc_refl3d(const float *vel, int jr, int jph, int jth, int nrpad, int nthpad)
{
    long X = 1, Z = nrpad, Y = (long)nrpad * nthpad;
    long offs[27] = {
        Z, Y, X, -Z, -Y, -X,
        Z + X, -Z - X, Z - X, -Z + X,
        Y + Z, -Y - Z, Y - Z, -Y + Z,
        Y + X, -Y - X, Y - X, -Y + X,
        Y - X,
        Y + Z + X, Y + Z - X, Y - Z + X, Y - Z - X,
        -Y + Z + X, -Y + Z - X, -Y - Z + X, -Y - Z - X,
    };
    return orc_refl(vel, (long)jph * Y + (long)jth * Z + jr, offs, 27);
}

float orc_refl_at(int ndim, const float *vel, int jr, int jth, int jph, int nrpad, int nthpad)
{
    return ndim == 3 ? orc_refl3d(vel, jr, jph, jth, nrpad, nthpad) : orc_refl2d(vel, jr, jth, nrpad);
}

/* ------------------------------------------------------------------ */
/* Spherical one-way ABC rule (spher_utils.c:583-631), serial semantics. */
#define ORC_NOP 4
int orc_spher3d_rule(int ix, int iz, int iy, int nxpad, int nzpad, int nypad)
{
    int iop, rule = 0;
    if (ix < ORC_NOP) iop = ix;
    else if (ix > nxpad - ORC_NOP) iop = nxpad - ix;
    else if (iz < ORC_NOP) iop = iz;
    else if (iz > nzpad - ORC_NOP) iop = nzpad - iz;
    else if (iy < ORC_NOP) iop = iy;
    else if (iy > nypad - ORC_NOP) iop = nypad - iy;
    else iop = 0;
    if (iz <= ORC_NOP - iop) rule = 1;
    if (iz >= nzpad - ORC_NOP + iop - 1) rule = 2;
    if (ix <= ORC_NOP - iop) rule = 3;
    if (ix >= nxpad - ORC_NOP + iop - 1) rule = 4;
    if (iy <= ORC_NOP - iop) rule = 5;
    if (iy >= nypad - ORC_NOP + iop - 1) rule = 6;
    return rule;
}

/* ------------------------------------------------------------------ */
/* Set-up of the FP64 cartesian 2-D scheme for the adjoint oracle
 * (oracle_adj.c): the same padding, coefficients, sponge, (v dt)^2 and index
 * map as orc_cart2d_f64 (MAWEFDcart2Dcpu.c:209-295). */
static int orc_fdcoef_d(int nop, int ndim, const float *d, int optfd, int order, double *out);

void orc_cart2d_free_f64(orc_cart2d_f64_t *g)
{
    free(g->G); free(g->V); free(g->dVds); free(g->snode); free(g->rnode); free(g->sw); free(g->rw);
    memset(g, 0, sizeof *g);
}

int orc_cart2d_setup_f64(const orc_par *p, const float *vel_in, const double *slow, const float *sou,
                         const float *rec, orc_cart2d_f64_t *g)
{
    memset(g, 0, sizeof *g);
    const int fdorder = p->fdorder > 0 ? p->fdorder : 4;
    const int nop = fdorder / 2;
    if (fdorder % 2 || nop < 1 || nop > 6) return -2;
    int nb = p->nb < nop ? nop : p->nb;
    g->nop = nop; g->nb = nb; g->nt = p->nt; g->ns = p->ns; g->nr = p->nr; g->expl = p->expl;
    g->mz = p->n[0]; g->mx = p->n[1];
    g->nz = g->mz + 2 * nb; g->nx = g->mx + 2 * nb;
    g->wstride = p->expl ? 1 : p->ns;
    const long N = (long)g->nz * g->nx;
    const float d[2] = {p->d[0], p->d[1]};
    if (orc_fdcoef_d(nop, 2, d, p->optfd, 2, g->c2)) return -2;
    g->G = malloc(sizeof(double) * N); g->V = malloc(sizeof(double) * N); g->dVds = malloc(sizeof(double) * N);
    g->snode = malloc(sizeof(long) * 4 * (p->ns + 1)); g->rnode = malloc(sizeof(long) * 4 * (p->nr + 1));
    g->sw = malloc(sizeof(double) * 4 * (p->ns + 1)); g->rw = malloc(sizeof(double) * 4 * (p->nr + 1));
    if (!g->G || !g->V || !g->dVds || !g->snode || !g->rnode || !g->sw || !g->rw) {
        orc_cart2d_free_f64(g);
        return -3;
    }
    double spo[4096];
    if (nb > 4096) { orc_cart2d_free_f64(g); return -3; }
    for (int i = 0; i < nb; i++) {
        double fb = i / (sqrt(2.0) * 4.0 * nb);
        spo[i] = exp(-fb * fb);
    }
    const double dt = p->dt;
    for (int ix = 0; ix < g->nx; ix++) {
        int kx = ix - nb; kx = kx < 0 ? 0 : (kx >= g->mx ? g->mx - 1 : kx);
        for (int iz = 0; iz < g->nz; iz++) {
            int kz = iz - nb; kz = kz < 0 ? 0 : (kz >= g->mz ? g->mz - 1 : kz);
            const long i = (long)ix * g->nz + iz;
            const double v = slow ? 1.0 / sqrt(slow[(long)kx * g->mz + kz]) : vel_in[(long)kx * g->mz + kz];
            /* free surface: V = 0 on the top nb+1 rows (MAWEFDcart2Dcpu.c:289-295) */
            g->V[i] = p->fsrf && iz <= nb ? 0.0 : v * v * dt * dt;
            g->dVds[i] = -g->V[i] * v * v;
            double gz = 1.0, gx = 1.0;
            if (p->dabc) {
                if (iz < nb) gz = spo[nb - 1 - iz];
                else if (iz >= g->nz - nb) gz = spo[nb - 1 - (g->nz - 1 - iz)];
                if (ix < nb) gx = spo[nb - 1 - ix];
                else if (ix >= g->nx - nb) gx = spo[nb - 1 - (g->nx - 1 - ix)];
            }
            g->G[i] = gz * gx;
        }
    }
    const float ozp = p->o[0] - (float)nb * p->d[0], oxp = p->o[1] - (float)nb * p->d[1];
    int *jz = malloc(sizeof(int) * (p->ns + p->nr + 1)), *jx = malloc(sizeof(int) * (p->ns + p->nr + 1));
    float *w = malloc(sizeof(float) * 4 * (p->ns + p->nr + 1));
    if (!jz || !jx || !w) { free(jz); free(jx); free(w); orc_cart2d_free_f64(g); return -3; }
    orc_lint2d(p->ns, sou, g->nz, g->nx, ozp, oxp, p->d[0], p->d[1], jz, jx, w);
    for (int s = 0; s < p->ns; s++)
        for (int q = 0; q < 4; q++) {
            g->snode[4 * s + q] = (long)(jx[s] + ((q >> 1) & 1)) * g->nz + jz[s] + (q & 1);
            g->sw[4 * s + q] = w[4 * s + q];
        }
    orc_lint2d(p->nr, rec, g->nz, g->nx, ozp, oxp, p->d[0], p->d[1], jz, jx, w);
    for (int r = 0; r < p->nr; r++)
        for (int q = 0; q < 4; q++) {
            g->rnode[4 * r + q] = (long)(jx[r] + ((q >> 1) & 1)) * g->nz + jz[r] + (q & 1);
            g->rw[4 * r + q] = w[4 * r + q];
        }
    free(jz); free(jx); free(w);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Propagators, instantiated for float and double. */
#ifdef ORC_REFFORM
/* Reference-form timing build (liborc_ref.so, -O3): the constant-density
 * step in ts_kernel.c:78-138 (2-D) / 205-227 (3-D) form, 4-wide SSE
 * intrinsics over z with schedule(static,1) over x columns (2-D) or y planes
 * (3-D), as the reference compiles it.  Every lane does the scalar SSE-order
 * operations of orc_step_cart (lap from c0*u, then per k the z, x[, y] pair
 * terms; u0 = (2u - u0) + vel*lap), so the bits are the checker's; the z
 * remainder of a column (when (nz - 2 nop) % 4 != 0) runs scalar instead of
 * overrunning into the halo. */
#include <xmmintrin.h>
static int orc_step_sse(float *u0, const float *u1, const float *vel, const float *rho, const float *c2,
                        int nop, int ndim, int nz, int nx, int ny, int sse)
{
    if (!sse || rho) return 0;
    const long sx = nz, sy = (long)nz * nx;
    const float *cz = c2, *cx = c2 + nop, *cy = c2 + 2 * nop;
    const float c0 = c2[0];
    const int nouter = ndim == 3 ? ny - 2 * nop : nx - 2 * nop;
#pragma omp parallel for schedule(static, 1)
    for (int o = 0; o < nouter; o++) {
        const int y0 = ndim == 3 ? nop + o : 0, y1 = ndim == 3 ? y0 + 1 : 1;
        const int x0 = ndim == 3 ? nop : nop + o, x1 = ndim == 3 ? nx - nop : x0 + 1;
        for (int iy = y0; iy < y1; iy++)
            for (int ix = x0; ix < x1; ix++) {
                const long b = iy * sy + ix * sx;
                int iz = nop;
                for (; iz + 4 <= nz - nop; iz += 4) {
                    const long i = b + iz;
                    __m128 uc = _mm_loadu_ps(u1 + i);
                    __m128 un = _mm_sub_ps(_mm_mul_ps(_mm_set1_ps(2.f), uc), _mm_loadu_ps(u0 + i));
                    __m128 lap = _mm_mul_ps(uc, _mm_set1_ps(c0));
                    for (int k = 1; k <= nop; k++) {
                        lap = _mm_add_ps(lap, _mm_mul_ps(_mm_set1_ps(cz[k]),
                                                         _mm_add_ps(_mm_loadu_ps(u1 + i + k), _mm_loadu_ps(u1 + i - k))));
                        lap = _mm_add_ps(lap, _mm_mul_ps(_mm_set1_ps(cx[k]), _mm_add_ps(_mm_loadu_ps(u1 + i + k * sx),
                                                                                        _mm_loadu_ps(u1 + i - k * sx))));
                        if (ndim == 3)
                            lap = _mm_add_ps(lap, _mm_mul_ps(_mm_set1_ps(cy[k]), _mm_add_ps(_mm_loadu_ps(u1 + i + k * sy),
                                                                                            _mm_loadu_ps(u1 + i - k * sy))));
                    }
                    _mm_storeu_ps(u0 + i, _mm_add_ps(un, _mm_mul_ps(_mm_loadu_ps(vel + i), lap)));
                }
                for (; iz < nz - nop; iz++) {
                    const long i = b + iz;
                    float lap = u1[i] * c0;
                    for (int k = 1; k <= nop; k++) {
                        lap = lap + cz[k] * (u1[i + k] + u1[i - k]);
                        lap = lap + cx[k] * (u1[i + k * sx] + u1[i - k * sx]);
                        if (ndim == 3) lap = lap + cy[k] * (u1[i + k * sy] + u1[i - k * sy]);
                    }
                    const float t = 2.f * u1[i] - u0[i];
                    u0[i] = t + vel[i] * lap;
                }
            }
    }
    return 1;
}
#define ORC_STEP_FAST orc_step_sse
#endif

#define REAL float
#define SFX(name) name##_f32
#define ORC_FDCOEF(nop, ndim, d, optfd, order, out) orc_fdcoef(nop, ndim, d, optfd, order, out)
#define ORC_SNAP orc_snap_f32
#include "oracle_prop.inc"
#undef REAL
#undef SFX
#undef ORC_FDCOEF
#undef ORC_SNAP

#undef ORC_STEP_FAST
#define REAL double
#define SFX(name) name##_f64
#define ORC_FDCOEF(nop, ndim, d, optfd, order, out) orc_fdcoef_d(nop, ndim, d, optfd, order, out)
#define ORC_SNAP orc_snap_f64
#include "oracle_prop.inc"
