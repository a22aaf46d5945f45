/*
 * Minimal RSF (Madagascar regularly sampled format) I/O and command-line
 * parameters for the drop-in programs.
 *
 * A header is text holding key=value tokens (the last occurrence of a key
 * wins); the data are esize=4 floats, native or xdr byte order, stored either
 * in the file named by in= or, for in="stdin", in the same stream after the
 * three bytes \f \f \004.  Output files written here embed their data that
 * way (in="stdin"), which every RSF reader accepts.
 */
#ifndef AWEFD_RSF_H
#define AWEFD_RSF_H

#include <stdio.h>

#define RSF_MAXDIM 9

typedef struct {
    int n[RSF_MAXDIM];
    float o[RSF_MAXDIM], d[RSF_MAXDIM];
    char label[RSF_MAXDIM][64], unit[RSF_MAXDIM][64];
    int ndim;          /* highest axis with n > 1 (at least 1) */
    long size;         /* product of n */
    float *data;
} rsf_t;

/* Read header and data.  path == NULL or "-" reads standard input.  Returns
 * 0, or -1 with a message in rsf_error(). */
int rsf_read(const char *path, rsf_t *f);
/* Write header (with the given history line) and embedded data.  path ==
 * NULL or "-" writes standard output. */
int rsf_write(const char *path, const rsf_t *f, const char *history);
/* Streaming writer for outputs produced step by step: header first, then
 * rsf_append calls, then rsf_close. */
FILE *rsf_open_out(const char *path, const rsf_t *hdr, const char *history);
int rsf_append(FILE *fp, const float *x, long n);
int rsf_close(FILE *fp);
void rsf_free(rsf_t *f);
void rsf_axes_init(rsf_t *f);
const char *rsf_error(void);

/* Command line: key=value arguments (sf_getpar family). */
void par_init(int argc, char **argv);
int par_string(const char *key, const char **val);
int par_int(const char *key, int *val);
int par_float(const char *key, float *val);
int par_bool(const char *key, int *val);      /* y/n/1/0/true/false */
const char *par_history(void);

#endif
