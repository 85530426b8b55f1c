/*
 * vacio_harness.c -- drives the VAC configuration readers and writers of
 * smaug-all_amd/host/vacio.c for tests/test_vacio.py (built there with gcc,
 * once per dimensionality).
 *
 *   vacio_harness in.ini outbase n0 n1 [n2]
 *
 * reads in.ini with readasciivacconfig and writes
 *   outbase.raw   it (int32), t (double), the NVAR fields and the NDIM
 *                 position arrays as read (doubles, field-major, cell order)
 *   outbase_<it>.out   with writevacconfig, read back with readbinvacconfig
 *                 and compared byte for byte with what was written
 *   outbase.ini   with writeasciivacconfig
 * and prints "binary roundtrip identical" when the binary read gives back
 * every written value.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "smaug_types.h"
#include "vacio.h"

int main(int argc, char **argv)
{
    if (argc < 3 + NDIM) {
        fprintf(stderr, "usage: %s in.ini outbase n0 n1 [n2]\n", argv[0]);
        return 2;
    }
    struct params p;
    memset(&p, 0, sizeof p);
    long nc = 1;
    for (int d = 0; d < NDIM; d++) {
        p.n[d] = atoi(argv[3 + d]);
        nc *= p.n[d];
    }
    real *w = calloc((size_t)NVAR * nc, sizeof(real));
    real *wd = calloc((size_t)NDERV * nc, sizeof(real));
    char *hl[5] = {0};
    struct Meta md;
    memset(&md, 0, sizeof md);
    md.name = (char *)"harness";
    struct state st;
    memset(&st, 0, sizeof st);
    int rc = readasciivacconfig(argv[1], p, md, &st, w, wd, hl, 0);
    if (rc) return 10 + rc;
    /* equation parameters from the header's fourth line: gamma eta g1 .. */
    double eq[8] = {0};
    sscanf(hl[3], "%lf %lf %lf %lf %lf", &eq[0], &eq[1], &eq[2], &eq[3], &eq[4]);
    p.gamma = eq[0];
    p.eta = eq[1];
    for (int d = 0; d < NDIM; d++) p.g[d] = eq[2 + d];

    char path[4096];
    snprintf(path, sizeof path, "%s.raw", argv[2]);
    FILE *f = fopen(path, "wb");
    if (!f) return 3;
    fwrite(&st.it, sizeof(int), 1, f);
    fwrite(&st.t, sizeof(double), 1, f);
    fwrite(w, sizeof(real), (size_t)NVAR * nc, f);
    for (int d = 0; d < NDIM; d++) fwrite(wd + (long)(pos1 + d) * nc, sizeof(real), nc, f);
    fclose(f);

    if ((rc = writevacconfig(argv[2], 0, p, md, w, wd, st))) return 20 + rc;
    real *w2 = calloc((size_t)NVAR * nc, sizeof(real));
    real *wd2 = calloc((size_t)NDERV * nc, sizeof(real));
    snprintf(path, sizeof path, "%s_%d.out", argv[2], st.it);
    if ((rc = readbinvacconfig(path, p, md, w2, wd2, st))) return 30 + rc;
    int same = !memcmp(w, w2, sizeof(real) * (size_t)NVAR * nc);
    for (int d = 0; d < NDIM; d++)
        same = same && !memcmp(wd + (long)(pos1 + d) * nc, wd2 + (long)(pos1 + d) * nc,
                               sizeof(real) * nc);
    if (same) printf("binary roundtrip identical\n");

    snprintf(path, sizeof path, "%s.ini", argv[2]);
    if ((rc = writeasciivacconfig(path, p, md, w, wd, hl, st, 0))) return 40 + rc;
    return same ? 0 : 1;
}
