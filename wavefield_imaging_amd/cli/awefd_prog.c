/*
 * Drop-in programs sfAWEFDcart2D, sfAWEFDcart3D, sfAWEFDspher2D and
 * sfAWEFDspher3D over libawefd.  Built once per program with
 * -DAWEFD_PROG_KIND=AWEFD_CART2D ... AWEFD_SPHER3D.
 *
 *   sfAWEFDcart2D < wav.rsf vel=v.rsf sou=s.rsf rec=r.rsf [den=d.rsf] [wfl=w.rsf] [key=value ...] > dat.rsf
 *
 * Parameters, defaults, inputs and outputs follow the reference programs'
 * mains: MAWEFDcart2Dcpu.c:116-211, MAWEFDcart3Dcpu.c:116-210,
 * MAWEFDspher2Dcpu.c:56-199, MAWEFDspher3Dcpu.c:57-171, their snapshot cuts
 * (fdutil.c:574-717) and their output headers.  Errors print a message and
 * exit with status 1, as sf_error does.
 *
 * Where the reference's behaviour is undefined the programs do this instead:
 *   jdata > 1     traces are decimated (rows it = 0, jdata, ...), matching the
 *                 header's n2/d2 (the reference writes every step, or
 *                 overruns its buffer in the spherical programs)
 *   cart3D bnds=y every padded y plane is written (the reference starts nb
 *                 planes in and reads past the grid)
 *   spher3D bnds=n the unpadded part of the field is written (the reference
 *                 writes an array it never fills)
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "awefd.h"
#include "rsf.h"

#ifndef AWEFD_PROG_KIND
#error "build with -DAWEFD_PROG_KIND=<kind>"
#endif

static const char *g_prog = "sfAWEFD";

static void die(const char *fmt, ...)
{
    va_list ap;
    fprintf(stderr, "%s: ", g_prog);
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    exit(1);
}

static void chk(int rc, const char *what)
{
    if (rc) die("%s: %s", what, awefd_last_error());
}

static rsf_t input(const char *key, int required)
{
    rsf_t f;
    const char *path = NULL;
    if (strcmp(key, "in") && !par_string(key, &path)) {
        if (required) die("need %s=", key);
        rsf_axes_init(&f);
        return f;
    }
    if (rsf_read(path, &f)) die("%s", rsf_error());
    return f;
}

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* snapshot writer state */
typedef struct {
    FILE *fp;
    int kind, bnds;
    int npad[3];
    int nq[3], f0[3], jq[3];  /* window: count, first padded index, stride per axis */
    float *buf;
} snap_t;

static void snap_cb(void *user, int it, const float *u, const int dims[3])
{
    (void)it;
    snap_t *s = (snap_t *)user;
    const long n1 = dims[0], n12 = (long)dims[0] * dims[1];
    long k = 0;
    if (s->kind == AWEFD_SPHER2D) {
        /* theta fastest (MAWEFDspher2Dcpu.c:302-313) */
        for (int ir = 0; ir < s->nq[0]; ir++)
            for (int ith = 0; ith < s->nq[1]; ith++)
                s->buf[k++] = u[(long)(s->f0[1] + ith * s->jq[1]) * n1 + s->f0[0] + ir * s->jq[0]];
    } else {
        for (int i3 = 0; i3 < s->nq[2]; i3++)
            for (int i2 = 0; i2 < s->nq[1]; i2++)
                for (int i1 = 0; i1 < s->nq[0]; i1++)
                    s->buf[k++] = u[(long)(s->f0[2] + i3 * s->jq[2]) * n12 + (long)(s->f0[1] + i2 * s->jq[1]) * n1
                                    + s->f0[0] + i1 * s->jq[0]];
    }
    if (rsf_append(s->fp, s->buf, k)) die("%s", rsf_error());
}

/* first index and stride of a window axis, as cut2d/cut3d_slice form them
 * (fdutil.c:582-589): floor((o_q - o_pad)/d), floor(d_q/d) in float */
static void window_axis(float oq, float dq, float opad, float d, int *f0, int *jq)
{
    *f0 = (int)floor((oq - opad) / d);
    *jq = (int)floor(dq / d);
}

static void write_traces(const float *tr, int nt, int nr, int jdata, int cart, const rsf_t *rec, float ot,
                         float dt);

int main(int argc, char **argv)
{
    const int kind = AWEFD_PROG_KIND;
    const int cart = kind == AWEFD_CART2D || kind == AWEFD_CART3D;
    const int ndim = (kind == AWEFD_CART3D || kind == AWEFD_SPHER3D) ? 3 : 2;
    static const char *names[4] = {"sfAWEFDcart2D", "sfAWEFDcart3D", "sfAWEFDspher2D", "sfAWEFDspher3D"};
    g_prog = names[kind];
    par_init(argc, argv);
    const double t0 = now();

    awefd_params p;
    chk(awefd_params_default(kind, &p), "defaults");
    int verb = 0, snap, bnds, expl, dabc = 0, cden = 0, adj = 0, fsrf = 0, optfd = 0, fdorder = 4;
    int hybrid = 0, sinc = 0, gpu = 0, jdata = 1, jsnap, nb = -1, kt = 0;
    par_bool("verb", &verb);
    if (cart) {
        /* MAWEFDcart2Dcpu.c:116-128, MAWEFDcart3Dcpu.c:116-128 */
        snap = 0; bnds = 0; expl = 0;
        par_bool("snap", &snap);
        par_bool("bnds", &bnds);
        par_bool("expl", &expl);
        par_bool("dabc", &dabc);
        par_bool("cden", &cden);
        par_bool("adj", &adj);
        if (!par_bool("free", &fsrf)) par_bool("fsrf", &fsrf);
        par_bool("optfd", &optfd);
        par_int("fdorder", &fdorder);
        par_bool("hybridbc", &hybrid);
        par_bool("sinc", &sinc);
        par_int("kt", &kt);   /* MAWEFDcart2Dcpu.c:165-167: 0 when absent */
    } else {
        /* MAWEFDspher2Dcpu.c:56-60 (snap, bnds default y), MAWEFDspher3Dcpu.c:57-61 */
        snap = ndim == 2;
        bnds = ndim == 2;
        expl = 1;
        par_bool("free", &fsrf);
        par_bool("dabc", &dabc);   /* read, unused by the reference */
        par_bool("snap", &snap);
        par_bool("bnds", &bnds);
        par_bool("expl", &expl);
    }
    par_int("gpu", &gpu);
    par_int("nb", &nb);
    par_int("jdata", &jdata);
    if (jdata < 1) die("jdata=%d", jdata);

    rsf_t wav = input("in", 1), vel = input("vel", 1), sou = input("sou", 1), rec = input("rec", 1);
    rsf_t den;
    rsf_axes_init(&den);
    if (cart && !cden) {
        const char *dp;
        if (par_string("den", &dp)) den = input("den", 1);
        else cden = 1;   /* MAWEFDcart2Dcpu.c:138-145 */
    }
    const int nt = wav.n[1];
    const float dt = wav.d[1], ot = wav.o[1];
    if (nt < 1) die("wavelet has n2=%d time samples", nt);
    if (!snap) jsnap = nt;
    else if (!par_int("jsnap", &jsnap)) jsnap = nt;
    if (jsnap < 1) die("jsnap=%d", jsnap);

    for (int a = 0; a < 3; a++) {
        p.n[a] = a < ndim ? vel.n[a] : 1;
        p.o[a] = a < ndim ? vel.o[a] : 0.f;
        p.d[a] = a < ndim ? vel.d[a] : 1.f;
    }
    p.nt = nt;
    p.dt = dt;
    p.nb = nb;
    p.fdorder = fdorder;
    p.optfd = optfd;
    p.dabc = dabc;
    p.fsrf = fsrf;
    p.expl = expl;
    p.cden = cden;
    p.jdata = jdata;
    p.device = gpu;
    p.adj = cart && adj;
    p.kt = kt;
    p.hybrid = cart && hybrid;
    p.sinc = cart && sinc;
    /* 2-D adj=y: the output is the image at step nt-kt, not traces */
    const int image = kind == AWEFD_CART2D && adj;

    /* point counts: axis 2 of sou=/rec= (the spherical 2-D program also
     * multiplies by axis 3, MAWEFDspher2Dcpu.c:90-91) */
    const int ns = kind == AWEFD_SPHER2D ? sou.n[1] * sou.n[2] : sou.n[1];
    const int nr = kind == AWEFD_SPHER2D ? rec.n[1] * rec.n[2] : rec.n[1];
    if ((long)ns * ndim > sou.size) die("sou= holds %ld values, need %d", sou.size, ns * ndim);
    if ((long)nr * ndim > rec.size) die("rec= holds %ld values, need %d", rec.size, nr * ndim);

    awefd_model *m;
    chk(awefd_create(&p, &m), "create");
    int dims[3], nbe;
    chk(awefd_padded_dims(m, dims, &nbe), "dims");
    if (verb)
        fprintf(stderr, "%s: n=%d,%d,%d padded %d,%d,%d nb=%d nt=%d ns=%d nr=%d\n", g_prog, p.n[0], p.n[1],
                p.n[2], dims[0], dims[1], dims[2], nbe, nt, ns, nr);
    if (vel.size < (long)p.n[0] * p.n[1] * p.n[2]) die("vel= is smaller than its axes");
    chk(awefd_set_model(m, vel.data, cart && !cden ? den.data : NULL), "model");
    chk(awefd_set_sources(m, ns, sou.data), "sources");
    chk(awefd_set_receivers(m, nr, rec.data), "receivers");
    chk(awefd_set_wavelet(m, wav.size, wav.data), "wavelet");

    /* snapshot file */
    snap_t sn;
    memset(&sn, 0, sizeof sn);
    sn.kind = kind;
    sn.bnds = bnds;
    rsf_t wh, win;
    rsf_axes_init(&wh);
    rsf_axes_init(&win);
    for (int a = 0; a < 3; a++) {
        /* the image axes: the model's, or the snapshot window's when snap=y
         * (MAWEFDcart2Dcpu.c:161-164,186-196: acz/acx) */
        win.n[a] = p.n[a];
        win.o[a] = p.o[a];
        win.d[a] = p.d[a];
    }
    if (snap) {
        const char *wp;
        if (!par_string("wfl", &wp)) die("snap=y needs wfl=");
        float opad[3];
        for (int a = 0; a < 3; a++) {
            sn.npad[a] = dims[a];
            opad[a] = a < ndim ? p.o[a] - (float)nbe * p.d[a] : 0.f;
        }
        /* the 2-D adj=y loop stops after step nt-kt: the header counts the
         * frames written (the reference keeps (nt-1)/jsnap+1 and writes fewer) */
        const int ntsnap = image ? (nt - kt) / jsnap + 1 : (nt - 1) / jsnap + 1;
        if (cart) {
            /* snapshot window: nqz/nqx/nqy, oqz..., dqz... (MAWEFDcart2Dcpu.c:186-207,
             * MAWEFDcart3Dcpu.c:178-205) */
            static const char *nk[3] = {"nqz", "nqx", "nqy"}, *ok[3] = {"oqz", "oqx", "oqy"}, *dk[3] = {"dqz", "dqx", "dqy"};
            for (int a = 0; a < 3; a++) {
                int nq = a < ndim ? p.n[a] : 1;
                float oq = p.o[a], dq = p.d[a];
                if (a < ndim) {
                    par_int(nk[a], &nq);
                    par_float(ok[a], &oq);
                    par_float(dk[a], &dq);
                }
                wh.n[a] = nq;
                wh.o[a] = oq;
                wh.d[a] = dq;
                win.n[a] = nq;
                win.o[a] = oq;
                win.d[a] = dq;
                if (a >= ndim) {
                    sn.nq[a] = 1;
                    sn.jq[a] = 1;
                    continue;
                }
                if (bnds) {
                    sn.nq[a] = dims[a];
                    sn.f0[a] = 0;
                    sn.jq[a] = 1;
                    wh.n[a] = dims[a];
                    if (ndim == 2) wh.o[a] = opad[a];   /* padded axes (MAWEFDcart2Dcpu.c:222-225) */
                } else {
                    sn.nq[a] = nq;
                    window_axis(oq, dq, opad[a], p.d[a], &sn.f0[a], &sn.jq[a]);
                }
            }
            static const char *lab[3] = {"z", "x", "y"};
            for (int a = 0; a < ndim; a++) snprintf(wh.label[a], sizeof wh.label[a], "%s", lab[a]);
            wh.n[ndim] = ntsnap;
            wh.o[ndim] = ot;
            wh.d[ndim] = dt * jsnap;
            snprintf(wh.label[ndim], sizeof wh.label[ndim], "t");
        } else if (kind == AWEFD_SPHER2D) {
            /* axes (theta, r, t), origins the padded ones (MAWEFDspher2Dcpu.c:176-194) */
            const int ath = 1, ara = 0;
            wh.n[0] = bnds ? dims[ath] : p.n[ath];
            wh.o[0] = opad[ath];
            wh.d[0] = p.d[ath];
            wh.n[1] = bnds ? dims[ara] : p.n[ara];
            wh.o[1] = opad[ara];
            wh.d[1] = p.d[ara];
            snprintf(wh.label[0], sizeof wh.label[0], "th");
            snprintf(wh.label[1], sizeof wh.label[1], "ra");
            wh.n[2] = ntsnap;
            wh.o[2] = ot;
            wh.d[2] = dt * jsnap;
            snprintf(wh.label[2], sizeof wh.label[2], "t");
            for (int a = 0; a < 2; a++) {
                if (bnds) {
                    sn.nq[a] = dims[a];
                    sn.f0[a] = 0;
                    sn.jq[a] = 1;
                } else {
                    /* cut2d(po, fdm, ath, ara) with the unpadded axes */
                    sn.nq[a] = p.n[a];
                    window_axis(p.o[a], p.d[a], opad[a], p.d[a], &sn.f0[a], &sn.jq[a]);
                }
            }
            sn.nq[2] = 1;
            sn.jq[2] = 1;
        } else {
            /* axes (r, theta, phi, t), unpadded origins (MAWEFDspher3Dcpu.c:140-162) */
            static const char *lab[3] = {"ra", "th", "ph"};
            for (int a = 0; a < 3; a++) {
                wh.n[a] = bnds ? dims[a] : p.n[a];
                wh.o[a] = p.o[a];
                wh.d[a] = p.d[a];
                snprintf(wh.label[a], sizeof wh.label[a], "%s", lab[a]);
                sn.nq[a] = bnds ? dims[a] : p.n[a];
                sn.f0[a] = bnds ? 0 : nbe;
                sn.jq[a] = 1;
            }
            wh.n[3] = ntsnap;
            wh.o[3] = ot;
            wh.d[3] = dt * jsnap;
            snprintf(wh.label[3], sizeof wh.label[3], "t");
        }
        for (int a = 0; a < 3; a++) {
            const long last = sn.f0[a] + (long)(sn.nq[a] - 1) * sn.jq[a];
            if (sn.nq[a] < 1 || sn.jq[a] < 1 || sn.f0[a] < 0 || last >= dims[a])
                die("snapshot window on axis %d (first %d, step %d, n %d) leaves the padded grid of %d",
                    a + 1, sn.f0[a], sn.jq[a], sn.nq[a], dims[a]);
        }
        wh.ndim = ndim + 1;
        wh.size = (long)sn.nq[0] * sn.nq[1] * sn.nq[2] * ntsnap;
        sn.buf = malloc(sizeof(float) * (size_t)sn.nq[0] * sn.nq[1] * sn.nq[2]);
        if (!sn.buf) die("out of memory");
        sn.fp = rsf_open_out(wp, &wh, par_history());
        if (!sn.fp) die("%s", rsf_error());
    }

    if (image && (win.n[0] != p.n[0] || win.n[1] != p.n[1]))
        die("adj=y with a snapshot window of %d x %d nodes: the image is the whole %d x %d model "
            "(the reference would write it into the smaller window)", win.n[0], win.n[1], p.n[0], p.n[1]);
    float *tr = malloc(sizeof(float) * (size_t)nt * (nr > 0 ? nr : 1));
    if (!tr) die("out of memory");
    chk(awefd_run(m, tr, snap ? jsnap : 0, snap ? snap_cb : NULL, &sn), "run");

    if (image) {
        /* the interior of u0 at step nt-kt (MAWEFDcart2Dcpu.c:353-370) */
        float *u = malloc(sizeof(float) * (size_t)dims[0] * dims[1]);
        float *img = malloc(sizeof(float) * (size_t)p.n[0] * p.n[1]);
        if (!u || !img) die("out of memory");
        chk(awefd_get_field(m, u), "field");
        for (int ix = 0; ix < p.n[1]; ix++)
            memcpy(img + (size_t)ix * p.n[0], u + (size_t)(nbe + ix) * dims[0] + nbe, sizeof(float) * p.n[0]);
        rsf_t out;
        rsf_axes_init(&out);
        for (int a = 0; a < 2; a++) {
            out.n[a] = p.n[a];
            out.o[a] = win.o[a];
            out.d[a] = win.d[a];
        }
        out.ndim = 2;
        FILE *fo = rsf_open_out(NULL, &out, par_history());
        if (!fo) die("%s", rsf_error());
        if (rsf_append(fo, img, (long)p.n[0] * p.n[1]) || rsf_close(fo)) die("%s", rsf_error());
        free(u);
        free(img);
    } else {
        write_traces(tr, nt, nr, jdata, cart, &rec, ot, dt);
    }
    if (snap && rsf_close(sn.fp)) die("%s", rsf_error());
    if (verb) fprintf(stderr, "\nElapsed time: %lf s\n", now() - t0);

    awefd_destroy(m);
    free(tr);
    free(sn.buf);
    rsf_free(&wav); rsf_free(&vel); rsf_free(&sou); rsf_free(&rec); rsf_free(&den);
    return 0;
}

/* traces: n1 = receivers (axis 2 of rec=), n2 = time */
static void write_traces(const float *tr, int nt, int nr, int jdata, int cart, const rsf_t *rec, float ot, float dt)
{
    rsf_t out;
    rsf_axes_init(&out);
    out.n[0] = nr;
    out.o[0] = rec->o[1];
    out.d[0] = rec->d[1];
    snprintf(out.label[0], sizeof out.label[0], "r");
    out.n[1] = cart ? (nt - 1) / jdata + 1 : nt / jdata;
    out.o[1] = ot;
    out.d[1] = dt * jdata;
    snprintf(out.label[1], sizeof out.label[1], "t");
    out.ndim = 2;
    FILE *fo = rsf_open_out(NULL, &out, par_history());
    if (!fo) die("%s", rsf_error());
    for (int k = 0; k < out.n[1]; k++)
        if (rsf_append(fo, tr + (long)k * jdata * nr, nr)) die("%s", rsf_error());
    if (rsf_close(fo)) die("%s", rsf_error());
}
