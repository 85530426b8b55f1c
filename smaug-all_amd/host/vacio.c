/*
 * vacio.c -- VAC/SAC configuration files for the drop-in driver.
 *
 * readasciivacconfig   replaces readwrite.c:982-1112
 * writeasciivacconfig  replaces readwrite.c:1113-1540 (single-domain case)
 * writevacconfig       replaces readwrite.c:160-368
 * readbinvacconfig     reads what writevacconfig writes
 *
 * Ascii layout (the reference's .ini files, e.g. configs/zero1_BW_asc.ini):
 *   line 1  name
 *   line 2  it t ndim neqpar nw
 *   line 3  n0 n1 [n2]          cells per dim including ghosts
 *   line 4  eqpar: gamma eta g1 g2 [g3] 0 0
 *   line 5  variable names
 *   then one row per cell, i fastest, then j, then k:
 *   x y [z] followed by the NVAR fields (rho m.. e b.. eb rhob bb..)
 * Binary layout (writevacconfig, the reference's single-domain writer):
 * the file is <name>_<it>.out; raw fwrite of name[79], it (int32),
 * t (double), ndim neqpar nw (int32 x3), dims (int32 x ndim), eqpar
 * (double x neqpar), names[79] (the reference's fixed string), then each of
 * the ndim position arrays and the NVAR field arrays over all cells in the
 * same cell order, then '\n'.  The reference's reader (readwrite.c:587-780)
 * consumes 80 bytes for the 79-byte name; readbinvacconfig here reads the
 * 79 that the writer writes.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/smaug_types.h"
#include "vacio.h"

static long ncells(const struct params *p)
{
    long n = 1;
    for (int d = 0; d < NDIM; d++) n *= p->n[d];
    return n;
}

static int read_line(FILE *f, char **out)
{
    char buf[4096];
    if (!fgets(buf, sizeof(buf), f)) return -1;
    size_t l = strlen(buf);
    while (l && (buf[l - 1] == '\n' || buf[l - 1] == '\r')) buf[--l] = 0;
    if (out) {
        *out = (char *)malloc(l + 1);
        memcpy(*out, buf, l + 1);
    }
    return 0;
}

int readasciivacconfig(char *cfgfile, params p, meta md, state *st, real *w, real *wd,
                       char **hlines, int mode)
{
    (void)md;
    (void)mode;
    FILE *f = fopen(cfgfile, "r");
    if (!f) {
        fprintf(stderr, "readasciivacconfig: cannot open %s\n", cfgfile);
        return 1;
    }
    char *tmp[5] = {0, 0, 0, 0, 0};
    char **h = hlines ? hlines : tmp;
    for (int k = 0; k < 5; k++)
        if (read_line(f, &h[k])) {
            fclose(f);
            fprintf(stderr, "readasciivacconfig: %s: short header\n", cfgfile);
            return 2;
        }
    int it = 0, ndim = 0, neqpar = 0, nw = 0;
    double t = 0;
    if (sscanf(h[1], "%d %lf %d %d %d", &it, &t, &ndim, &neqpar, &nw) != 5 || ndim != NDIM ||
        nw != NVAR) {
        fclose(f);
        fprintf(stderr, "readasciivacconfig: %s: header '%s' does not match ndim %d nw %d\n",
                cfgfile, h[1], NDIM, NVAR);
        return 3;
    }
    int dims[3] = {1, 1, 1};
    if (sscanf(h[2], "%d %d %d", &dims[0], &dims[1], &dims[2]) < NDIM) {
        fclose(f);
        return 3;
    }
    for (int d = 0; d < NDIM; d++)
        if (dims[d] != p.n[d]) {
            fclose(f);
            fprintf(stderr, "readasciivacconfig: %s: grid %d along dim %d, params say %d\n",
                    cfgfile, dims[d], d, p.n[d]);
            return 4;
        }
    if (st) {
        st->it = it;
        st->t = t;
    }
    long nc = ncells(&p);
    for (long c = 0; c < nc; c++) {
        for (int d = 0; d < NDIM; d++)
            if (fscanf(f, "%lf", &wd[(long)(pos1 + d) * nc + c]) != 1) goto bad;
        for (int v = 0; v < NVAR; v++)
            if (fscanf(f, "%lf", &w[(long)v * nc + c]) != 1) goto bad;
    }
    fclose(f);
    for (int k = 0; k < 5; k++) free(tmp[k]);
    return 0;
bad:
    fclose(f);
    fprintf(stderr, "readasciivacconfig: %s: truncated data\n", cfgfile);
    return 5;
}

static void eqpar(const params *p, double *e, int *ne)
{
    e[0] = p->gamma;
    e[1] = p->eta;
    for (int d = 0; d < NDIM; d++) e[2 + d] = p->g[d];
    e[2 + NDIM] = 0;
    e[3 + NDIM] = 0;
    *ne = 4 + NDIM;
}

#ifdef USE_SAC_3D
static const char *VARNAMES = "x y z h m1 m2 m3 e b1 b2 b3 eb rhob bg1 bg2 bg3   gamma eta   grav1 grav2 grav3";
#else
static const char *VARNAMES = "x y h m1 m2 e b1 b2 eb rhob bg1 bg2   gamma eta   grav1 grav2";
#endif

int writeasciivacconfig(char *cfgfile, params p, meta md, real *w, real *wd, char **hlines,
                        state st, int mode)
{
    (void)hlines;
    (void)mode;
    FILE *f = fopen(cfgfile, "w");
    if (!f) return 1;
    double e[8];
    int ne;
    eqpar(&p, e, &ne);
    fprintf(f, "%s\n", md.name ? md.name : "smaug");
    fprintf(f, "%7d %23.16E %2d %2d %2d\n", st.it, st.t, NDIM, ne, NVAR);
    for (int d = 0; d < NDIM; d++) fprintf(f, "%4d%s", p.n[d], d + 1 < NDIM ? " " : "\n");
    for (int k = 0; k < ne; k++) fprintf(f, " %23.16E", e[k]);
    fprintf(f, "\n%s\n", VARNAMES);
    long nc = ncells(&p);
    for (long c = 0; c < nc; c++) {
        for (int d = 0; d < NDIM; d++) fprintf(f, " %24.16E", wd[(long)(pos1 + d) * nc + c]);
        for (int v = 0; v < NVAR; v++) fprintf(f, " %24.16E", w[(long)v * nc + c]);
        fprintf(f, "\n");
    }
    return fclose(f) ? 2 : 0;
}

#ifdef USE_SAC_3D
static const char *BINNAMES = "x y z rho mx my mz e bx by bz gamma eta g1 g2 g3";
#else
static const char *BINNAMES = "x y rho mx my e bx by  gamma eta g1 g2";
#endif

int writevacconfig(char *name, int n, params p, meta md, real *w, real *wd, state st)
{
    (void)n;
    char fname[4096];
    snprintf(fname, sizeof fname, "%s_%d.out", name, st.it);
    FILE *f = fopen(fname, "wb");
    if (!f) return 1;
    char buf[79];
    memset(buf, ' ', sizeof(buf));
    if (md.name) memcpy(buf, md.name, strlen(md.name) < 79 ? strlen(md.name) : 79);
    fwrite(buf, 1, 79, f);
    int ib[3] = {st.it, 0, 0};
    fwrite(ib, sizeof(int), 1, f);
    fwrite(&st.t, sizeof(double), 1, f);
    double e[8];
    int ne;
    eqpar(&p, e, &ne);
    ib[0] = NDIM;
    ib[1] = ne;
    ib[2] = NVAR;
    fwrite(ib, sizeof(int), 3, f);
    fwrite(p.n, sizeof(int), NDIM, f);
    fwrite(e, sizeof(double), ne, f);
    memset(buf, 0, sizeof(buf));
    memcpy(buf, BINNAMES, strlen(BINNAMES) < 79 ? strlen(BINNAMES) : 79);
    fwrite(buf, 1, 79, f);
    long nc = ncells(&p);
    for (int d = 0; d < NDIM; d++) fwrite(wd + (long)(pos1 + d) * nc, sizeof(double), nc, f);
    for (int v = 0; v < NVAR; v++) fwrite(w + (long)v * nc, sizeof(double), nc, f);
    fputc('\n', f);
    return fclose(f) ? 2 : 0;
}

int readbinvacconfig(char *name, params p, meta md, real *w, real *wd, state st)
{
    (void)md;
    (void)st;
    FILE *f = fopen(name, "rb");
    if (!f) return 1;
    char buf[79];
    int ib[3], it, dims[3];
    double t, e[8];
    if (fread(buf, 1, 79, f) != 79 || fread(&it, sizeof(int), 1, f) != 1 ||
        fread(&t, sizeof(double), 1, f) != 1 || fread(ib, sizeof(int), 3, f) != 3 ||
        ib[0] != NDIM || ib[2] != NVAR || ib[1] > 8 ||
        fread(dims, sizeof(int), NDIM, f) != (size_t)NDIM ||
        fread(e, sizeof(double), ib[1], f) != (size_t)ib[1] || fread(buf, 1, 79, f) != 79) {
        fclose(f);
        return 2;
    }
    for (int d = 0; d < NDIM; d++)
        if (dims[d] != p.n[d]) {
            fclose(f);
            return 3;
        }
    long nc = ncells(&p);
    for (int d = 0; d < NDIM; d++)
        if (fread(wd + (long)(pos1 + d) * nc, sizeof(double), nc, f) != (size_t)nc) goto bad;
    for (int v = 0; v < NVAR; v++)
        if (fread(w + (long)v * nc, sizeof(double), nc, f) != (size_t)nc) goto bad;
    fclose(f);
    return 0;
bad:
    fclose(f);
    return 4;
}
