/* RSF I/O and command-line parameters for the drop-in programs (see rsf.h). */
#include "rsf.h"

#include <ctype.h>
#include <errno.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];

static int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

const char *rsf_error(void) { return g_err; }

/* ---------------------------------------------------------------- */
/* key=value token tables */

typedef struct {
    char *key, *val;
} kv_t;

typedef struct {
    kv_t *kv;
    int n, cap;
} table_t;

static void table_put(table_t *t, const char *k, size_t kl, const char *v, size_t vl)
{
    /* strip one level of double quotes */
    if (vl >= 2 && v[0] == '"' && v[vl - 1] == '"') {
        v++;
        vl -= 2;
    }
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 64;
        t->kv = realloc(t->kv, sizeof(kv_t) * t->cap);
    }
    t->kv[t->n].key = strndup(k, kl);
    t->kv[t->n].val = strndup(v, vl);
    t->n++;
}

static const char *table_get(const table_t *t, const char *key)
{
    for (int i = t->n - 1; i >= 0; i--)
        if (!strcmp(t->kv[i].key, key)) return t->kv[i].val;
    return NULL;
}

static void table_free(table_t *t)
{
    for (int i = 0; i < t->n; i++) {
        free(t->kv[i].key);
        free(t->kv[i].val);
    }
    free(t->kv);
    memset(t, 0, sizeof *t);
}

/* every whitespace-separated key=value token of a text block */
static void table_scan(table_t *t, const char *s, size_t len)
{
    size_t i = 0;
    while (i < len) {
        while (i < len && isspace((unsigned char)s[i])) i++;
        size_t b = i;
        int q = 0;
        while (i < len && (q || !isspace((unsigned char)s[i]))) {
            if (s[i] == '"') q = !q;
            i++;
        }
        const char *tok = s + b, *eq = memchr(tok, '=', i - b);
        if (eq && eq > tok) table_put(t, tok, eq - tok, eq + 1, (s + i) - (eq + 1));
    }
}

/* ---------------------------------------------------------------- */
/* command line */

static table_t g_par;
static char g_hist[4096];

void par_init(int argc, char **argv)
{
    size_t off = 0;
    for (int i = 0; i < argc; i++) {
        if (i > 0) table_scan(&g_par, argv[i], strlen(argv[i]));
        int w = snprintf(g_hist + off, sizeof g_hist - off, "%s%s", i ? " " : "", argv[i]);
        if (w > 0 && off + w < sizeof g_hist) off += w;
    }
}

const char *par_history(void) { return g_hist; }

int par_string(const char *key, const char **val)
{
    const char *v = table_get(&g_par, key);
    if (!v) return 0;
    *val = v;
    return 1;
}

int par_int(const char *key, int *val)
{
    const char *v = table_get(&g_par, key);
    if (!v) return 0;
    *val = (int)strtol(v, NULL, 10);
    return 1;
}

int par_float(const char *key, float *val)
{
    const char *v = table_get(&g_par, key);
    if (!v) return 0;
    *val = strtof(v, NULL);
    return 1;
}

int par_bool(const char *key, int *val)
{
    const char *v = table_get(&g_par, key);
    if (!v) return 0;
    *val = (v[0] == 'y' || v[0] == 'Y' || v[0] == '1' || v[0] == 't' || v[0] == 'T');
    return 1;
}

/* ---------------------------------------------------------------- */
/* files */

void rsf_axes_init(rsf_t *f)
{
    memset(f, 0, sizeof *f);
    for (int a = 0; a < RSF_MAXDIM; a++) {
        f->n[a] = 1;
        f->d[a] = 1.f;
    }
    f->ndim = 1;
    f->size = 1;
}

void rsf_free(rsf_t *f)
{
    free(f->data);
    f->data = NULL;
}

static int read_all(FILE *fp, char **buf, size_t *len)
{
    size_t cap = 1 << 16, n = 0;
    char *b = malloc(cap);
    if (!b) return fail("out of memory");
    for (;;) {
        if (n == cap) {
            cap *= 2;
            char *nb = realloc(b, cap);
            if (!nb) {
                free(b);
                return fail("out of memory");
            }
            b = nb;
        }
        size_t r = fread(b + n, 1, cap - n, fp);
        n += r;
        if (r == 0) break;
    }
    *buf = b;
    *len = n;
    return 0;
}

static void swap4(float *x, long n)
{
    uint32_t *u = (uint32_t *)x;
    for (long i = 0; i < n; i++) {
        uint32_t v = u[i];
        u[i] = (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
    }
}

int rsf_read(const char *path, rsf_t *f)
{
    rsf_axes_init(f);
    const int std = !path || !strcmp(path, "-");
    FILE *fp = std ? stdin : fopen(path, "rb");
    if (!fp) return fail("cannot open %s: %s", path, strerror(errno));
    char *buf = NULL;
    size_t len = 0;
    int rc = read_all(fp, &buf, &len);
    if (!std) fclose(fp);
    if (rc) return rc;
    /* header ends at \f\f\004 (embedded data) or at the end of the file */
    size_t hend = len, dbeg = len;
    for (size_t i = 0; i + 2 < len; i++)
        if (buf[i] == '\014' && buf[i + 1] == '\014' && buf[i + 2] == '\004') {
            hend = i;
            dbeg = i + 3;
            break;
        }
    table_t t = {0};
    table_scan(&t, buf, hend);
    char key[16];
    int ndim = 1;
    long size = 1;
    for (int a = 0; a < RSF_MAXDIM; a++) {
        const char *v;
        snprintf(key, sizeof key, "n%d", a + 1);
        if ((v = table_get(&t, key))) f->n[a] = atoi(v);
        snprintf(key, sizeof key, "o%d", a + 1);
        if ((v = table_get(&t, key))) f->o[a] = strtof(v, NULL);
        snprintf(key, sizeof key, "d%d", a + 1);
        if ((v = table_get(&t, key))) f->d[a] = strtof(v, NULL);
        snprintf(key, sizeof key, "label%d", a + 1);
        if ((v = table_get(&t, key))) snprintf(f->label[a], sizeof f->label[a], "%s", v);
        snprintf(key, sizeof key, "unit%d", a + 1);
        if ((v = table_get(&t, key))) snprintf(f->unit[a], sizeof f->unit[a], "%s", v);
        if (f->n[a] < 1) {
            table_free(&t);
            free(buf);
            return fail("%s: n%d=%d", path ? path : "stdin", a + 1, f->n[a]);
        }
        if (f->n[a] > 1) ndim = a + 1;
        size *= f->n[a];
    }
    f->ndim = ndim;
    f->size = size;
    const char *in = table_get(&t, "in"), *fmt = table_get(&t, "data_format"),
               *es = table_get(&t, "esize");
    if (es && atoi(es) != 4) {
        table_free(&t);
        free(buf);
        return fail("%s: esize=%s (only 4-byte floats)", path ? path : "stdin", es);
    }
    const int xdr = fmt && !strncmp(fmt, "xdr", 3);
    if (fmt && strstr(fmt, "float") == NULL) {
        table_free(&t);
        free(buf);
        return fail("%s: data_format=%s (only float)", path ? path : "stdin", fmt);
    }
    f->data = malloc(sizeof(float) * (size > 0 ? size : 1));
    if (!f->data) {
        table_free(&t);
        free(buf);
        return fail("out of memory");
    }
    const size_t need = sizeof(float) * (size_t)size;
    if (!in || !strcmp(in, "stdin")) {
        if (len - dbeg < need) rc = fail("%s: %zu data bytes, need %zu", path ? path : "stdin", len - dbeg, need);
        else memcpy(f->data, buf + dbeg, need);
    } else {
        FILE *dp = fopen(in, "rb");
        if (!dp) rc = fail("cannot open data file %s: %s", in, strerror(errno));
        else {
            if (fread(f->data, 1, need, dp) != need) rc = fail("%s: short data file", in);
            fclose(dp);
        }
    }
    if (!rc && xdr) swap4(f->data, size);
    table_free(&t);
    free(buf);
    if (rc) rsf_free(f);
    return rc;
}

FILE *rsf_open_out(const char *path, const rsf_t *h, const char *history)
{
    const int std = !path || !strcmp(path, "-");
    FILE *fp = std ? stdout : fopen(path, "wb");
    if (!fp) {
        fail("cannot open %s: %s", path, strerror(errno));
        return NULL;
    }
    fprintf(fp, "%s\n", history ? history : "");
    for (int a = 0; a < RSF_MAXDIM; a++) {
        if (a >= h->ndim && h->n[a] == 1) continue;
        fprintf(fp, "\tn%d=%d\n\to%d=%.9g\n\td%d=%.9g\n", a + 1, h->n[a], a + 1, h->o[a], a + 1, h->d[a]);
        if (h->label[a][0]) fprintf(fp, "\tlabel%d=\"%s\"\n", a + 1, h->label[a]);
        if (h->unit[a][0]) fprintf(fp, "\tunit%d=\"%s\"\n", a + 1, h->unit[a]);
    }
    fprintf(fp, "\tesize=4\n\tdata_format=\"native_float\"\n\tin=\"stdin\"\n\n\014\014\004");
    return fp;
}

int rsf_append(FILE *fp, const float *x, long n)
{
    if (n > 0 && fwrite(x, sizeof(float), (size_t)n, fp) != (size_t)n) return fail("write failed: %s", strerror(errno));
    return 0;
}

int rsf_close(FILE *fp)
{
    if (fp == stdout) return fflush(fp) ? fail("flush failed") : 0;
    return fclose(fp) ? fail("close failed") : 0;
}

int rsf_write(const char *path, const rsf_t *f, const char *history)
{
    FILE *fp = rsf_open_out(path, f, history);
    if (!fp) return -1;
    int rc = rsf_append(fp, f->data, f->size);
    int rc2 = rsf_close(fp);
    return rc ? rc : rc2;
}
