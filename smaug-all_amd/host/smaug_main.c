/*
 * smaug_main.c -- host driver of the drop-in path, built once per
 * dimensionality (-DUSE_SAC / -DUSE_SAC_3D) against libsmaug_hip2d/3d.so.
 *
 * The onestep (rkon = 0) time loop is the reference's, call for call
 * (dev/smaug/src/iosmaug.cpp:786-1168): slot copy, optional Courant and
 * hyperviscous dt control, advection per direction and field, gravity,
 * hyperdiffusion per field/direction with the (ii, ii0) and (jj, mm, sb)
 * patterns, user source and the ghost fill of every field and direction.
 * rkon = 1 advances with classical RK4 on the fused engine (cuadvancerk4),
 * or, with -rk4calls, through the same per-entry-point calls as the onestep
 * loop, one stage at a time, with cuadvance applying the RK slot algebra.
 *
 * The parameters the reference sets in iosmaugparams.h come from the
 * configuration file header (grid, gamma, gravity) and the command line:
 *
 *   smaug_main -i in.ini [-o out.ini] [-n steps] [-dt dt] [-rk4 | -rk4calls] [-mixed]
 *              [-bc b0,b1[,b2]] [-hyperdif 0|1] [-chyp rho,e,mom,b]
 *              [-courant c] [-moddton] [-driver A,P,sh,z0,sz,xc,yc] [-torsional]
 *              [-t0 t] [-tend T] [-dtlog file]
 * -tend stops the onestep loop once the elapsed time (the sum of the steps'
 * dt) reaches T, the last step shortened to land on T, as SAC does with
 * tmaxexact (dev/sac/src/vac.t:86, 606); without -n the loop then runs until
 * T.  -dtlog writes the step number and dt of every onestep step to a file,
 * one line each.  -torsional selects the reference's torsional driver (usersource.cu), with
 * the -driver values if given and the reference's constants otherwise
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/smaug_hip.h"
#include "vacio.h"

#define CALL(x)                                                                            \
    do {                                                                                   \
        int rc_ = (x);                                                                     \
        if (rc_) {                                                                         \
            fprintf(stderr, "%s failed (%d): %s\n", #x, rc_, cuerrorstring());             \
            exit(1);                                                                       \
        }                                                                                  \
    } while (0)

/* the host and device handles of cuinit, passed by address as the entry
 * points take them */
struct handles {
    struct params *p, *d_p;
    struct bparams *bp, *d_bp;
    struct state *d_state;
    real *wd, *d_wmod, *d_dwn1, *d_wd, *d_wtemp, *d_wtemp1, *d_wtemp2;
};

/* One stage of the reference's call sequence (iosmaug.cpp:913-1160):
 * velocities, advection per direction and field, gravity, hyperdiffusion per
 * field and direction with the (ii, ii0) and (jj, mm, sb) patterns and the
 * user source.  Reads slot `order` and accumulates into slot `ordero`.  Every
 * term takes p->dt except the energy hyperdiffusion, which takes edt (the
 * reference hands it the loop's initial dt). */
static void stage(struct handles *h, int order, int ordero, real edt)
{
    struct params **pp = &h->p;
    struct params *p = h->p;
    for (int dir = 0; dir < NDIM; dir++) CALL(cucomputevels(pp, &h->d_p, &h->d_wmod, &h->d_wd, order, dir));
    for (int dir = 0; dir < NDIM; dir++) {
        for (int fld = rho; fld <= (mom1 + NDIM - 1); fld++) {
            if (fld == mom1 + dir) {
                CALL(cucomputept(pp, &h->d_p, &h->d_wmod, &h->d_wd, order, dir));
                CALL(cucomputepbg(pp, &h->d_p, &h->d_wmod, &h->d_wd, order, dir));
            }
            CALL(cucentdiff1(pp, &h->d_p, &h->d_state, &h->d_wmod, &h->d_wmod, &h->d_dwn1, &h->d_wd,
                             order, ordero, p->dt, fld, dir));
        }
        for (int fld = energy; fld <= (b1 + (NDIM - 1)); fld++) {
            if (fld == energy) {
                CALL(cucomputevels(pp, &h->d_p, &h->d_wmod, &h->d_wd, order, dir));
                CALL(cucomputepbg(pp, &h->d_p, &h->d_wmod, &h->d_wd, ordero, dir));
                CALL(cucomputept(pp, &h->d_p, &h->d_wmod, &h->d_wd, order, dir));
            }
            CALL(cucentdiff2(pp, &h->d_p, &h->d_state, &h->d_wmod, &h->d_wmod, &h->d_dwn1, &h->d_wd,
                             order, ordero, p->dt, fld, dir));
        }
    }
    CALL(cugrav(pp, &h->d_p, &h->d_state, &h->d_wmod, &h->d_wmod, &h->d_dwn1, &h->d_wd, order, ordero, p->dt));
    if (p->divbon == 1)
        CALL(cudivb(pp, &h->d_p, &h->d_wmod, &h->d_wmod, &h->d_dwn1, &h->d_wd, order, ordero, p->dt));
    if (p->hyperdifmom == 1) {
        real cmax[NDIM];
        real **wd = &h->wd, **d_wmod = &h->d_wmod, **d_wd = &h->d_wd, **d_dwn1 = &h->d_dwn1;
        real **t0 = &h->d_wtemp, **t1 = &h->d_wtemp1, **t2 = &h->d_wtemp2;
        struct params **d_p = &h->d_p;
        p->maxviscoef = 0.0;
        for (int dim = 0; dim <= (NDIM - 1); dim++) {
            CALL(cucomputec(pp, d_p, d_wmod, d_wd, order, dim));
            CALL(cucomputemaxc(pp, d_p, d_wmod, d_wd, order, dim, wd, t0));
            cmax[dim] = p->cmax;
            CALL(cuhyperdifvisc1ir(pp, d_p, d_wmod, d_wd, order, t0, t1, t2, rho, dim));
            CALL(cuhyperdifvisc1r(pp, d_p, d_wmod, wd, d_wd, order, t0, t1, t2, rho, dim));
            CALL(cuhyperdifvisc1l(pp, d_p, d_wmod, wd, d_wd, order, t0, t1, t2, rho, dim));
            CALL(cuhyperdifrhosource1(pp, d_p, d_wmod, d_dwn1, d_wd, order, ordero, t0, rho, dim, p->dt));
        }
        for (int dim = 0; dim <= (NDIM - 1); dim++) {
            p->cmax = cmax[dim];
            CALL(cuhyperdifvisc1ir(pp, d_p, d_wmod, d_wd, order, t0, t1, t2, energy, dim));
            CALL(cuhyperdifvisc1r(pp, d_p, d_wmod, wd, d_wd, order, t0, t1, t2, energy, dim));
            CALL(cuhyperdifvisc1l(pp, d_p, d_wmod, wd, d_wd, order, t0, t1, t2, energy, dim));
            CALL(cuhyperdifesource1(pp, d_p, d_wmod, d_dwn1, d_wd, order, ordero, t0, energy, dim, edt));
        }
        for (int dim = 0; dim <= (NDIM - 1); dim++)
            for (int fi = 0; fi <= (NDIM - 1); fi++) {
                p->cmax = cmax[dim];
                CALL(cuhyperdifvisc1ir(pp, d_p, d_wmod, d_wd, order, t0, t1, t2, mom1 + fi, dim));
                CALL(cuhyperdifvisc1r(pp, d_p, d_wmod, wd, d_wd, order, t0, t1, t2, mom1 + fi, dim));
                CALL(cuhyperdifvisc1l(pp, d_p, d_wmod, wd, d_wd, order, t0, t1, t2, mom1 + fi, dim));
                for (int ii1 = 0; ii1 <= 1; ii1++) {
                    int ii = ii1 == 0 ? dim : fi;
                    int ii0 = ii1 == 0 ? fi : dim;
                    if (ii == dim)
                        CALL(cuhyperdifmomsource1(pp, d_p, d_wmod, d_dwn1, d_wd, order, ordero, t0, fi, dim, ii, ii0, p->dt));
                    else
                        CALL(cuhyperdifmomsourcene1(pp, d_p, d_wmod, d_dwn1, d_wd, order, ordero, t0, fi, dim, ii, ii0, p->dt));
                }
            }
        for (int dim = 0; dim <= (NDIM - 1); dim++)
            for (int fi = 0; fi <= (NDIM - 1); fi++) {
                if (fi == dim) continue;
                p->cmax = cmax[dim];
                CALL(cuhyperdifvisc1ir(pp, d_p, d_wmod, d_wd, order, t0, t1, t2, b1 + fi, dim));
                CALL(cuhyperdifvisc1r(pp, d_p, d_wmod, wd, d_wd, order, t0, t1, t2, b1 + fi, dim));
                CALL(cuhyperdifvisc1l(pp, d_p, d_wmod, wd, d_wd, order, t0, t1, t2, b1 + fi, dim));
                for (int ii1 = 0; ii1 <= 1; ii1++) {
                    int jj, mm, ii0;
                    real sb;
                    if (ii1 == 0) {
                        jj = dim;
                        mm = fi;
                        sb = -1.0;
                        ii0 = dim;
                    } else {
                        ii0 = fi;
                        mm = dim;
                        sb = 1.0;
                        jj = fi;
                    }
                    if (mm == dim)
                        CALL(cuhyperdifbsource1(pp, d_p, d_wmod, d_dwn1, d_wd, order, ordero, t0, fi, dim, jj, ii0, mm, sb, p->dt));
                    else
                        CALL(cuhyperdifbsourcene1(pp, d_p, d_wmod, d_dwn1, d_wd, order, ordero, t0, fi, dim, jj, ii0, mm, sb, p->dt));
                }
            }
    }
    CALL(cusource(pp, &h->d_p, &h->d_state, &h->d_wmod, &h->d_wmod, &h->d_dwn1, &h->d_wd, order, ordero, p->dt));
}

/* ghost fill of every field and direction of one slot */
static void fill_ghosts(struct handles *h, int slot)
{
    for (int ii = 0; ii <= (b1 + (NDIM - 1)); ii++)
        for (int idir = 0; idir < NDIM; idir++)
            CALL(cuboundary(&h->p, &h->bp, &h->d_p, &h->d_bp, &h->d_state, &h->d_wmod, slot, idir, ii));
}

static int parse_list(const char *s, double *v, int maxn)
{
    int k = 0;
    size_t l = strlen(s);
    char *dup = (char *)malloc(l + 1), *tok;
    memcpy(dup, s, l + 1);
    tok = strtok(dup, ",");
    while (tok && k < maxn) {
        v[k++] = atof(tok);
        tok = strtok(NULL, ",");
    }
    free(dup);
    return k;
}

int main(int argc, char **argv)
{
    const char *in = NULL, *out = NULL, *dtlog = NULL;
    int nt = 1, rk4 = 0, rk4calls = 0, mixed = 0, hyperdif = 1, moddton = 0, bcset = 0, driver = 0, have_drv = 0;
    double dtarg = 0, courant = 0.3, t0 = 0, tend = 0;
    int nset = 0;
    double bcv[3] = {4, 4, 4}, chv[4] = {0.02, 0.02, 0.4, 0.02}, drv[7] = {0};
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "-i") && a + 1 < argc) in = argv[++a];
        else if (!strcmp(argv[a], "-o") && a + 1 < argc) out = argv[++a];
        else if (!strcmp(argv[a], "-n") && a + 1 < argc) {
            nt = atoi(argv[++a]);
            nset = 1;
        } else if (!strcmp(argv[a], "-tend") && a + 1 < argc)
            tend = atof(argv[++a]);
        else if (!strcmp(argv[a], "-dtlog") && a + 1 < argc)
            dtlog = argv[++a];
        else if (!strcmp(argv[a], "-dt") && a + 1 < argc) dtarg = atof(argv[++a]);
        else if (!strcmp(argv[a], "-t0") && a + 1 < argc) t0 = atof(argv[++a]);
        else if (!strcmp(argv[a], "-rk4")) rk4 = 1;
        else if (!strcmp(argv[a], "-rk4calls")) rk4 = rk4calls = 1;
        else if (!strcmp(argv[a], "-mixed")) mixed = 1;
        else if (!strcmp(argv[a], "-moddton")) moddton = 1;
        else if (!strcmp(argv[a], "-courant") && a + 1 < argc) courant = atof(argv[++a]);
        else if (!strcmp(argv[a], "-hyperdif") && a + 1 < argc) hyperdif = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-bc") && a + 1 < argc) bcset = parse_list(argv[++a], bcv, 3);
        else if (!strcmp(argv[a], "-chyp") && a + 1 < argc) parse_list(argv[++a], chv, 4);
        else if (!strcmp(argv[a], "-driver") && a + 1 < argc) {
            have_drv = parse_list(argv[++a], drv, 7) == 7;
            if (driver != 2) driver = have_drv;
        } else if (!strcmp(argv[a], "-torsional"))
            driver = 2;
        else {
            fprintf(stderr, "unknown or incomplete option %s\n", argv[a]);
            return 2;
        }
    }
    if (tend > 0 && rk4) {
        fprintf(stderr, "-tend applies to the onestep loop only\n");
        return 2;
    }
    if (rk4calls && (mixed || moddton)) {
        fprintf(stderr, "-rk4calls runs in FP64 with a fixed dt (no -mixed, no -moddton)\n");
        return 2;
    }
    if (tend > 0 && !nset) nt = 1000000000;
    if (!in) {
        fprintf(stderr, "usage: %s -i config.ini [-o out.ini] [-n steps] [-dt dt] [-rk4] ...\n", argv[0]);
        return 2;
    }

    /* grid and equation parameters from the configuration header */
    FILE *f = fopen(in, "r");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", in);
        return 1;
    }
    char line[4096];
    int dims[3] = {1, 1, 1};
    double eq[8] = {0};
    if (!fgets(line, sizeof line, f) || !fgets(line, sizeof line, f) || !fgets(line, sizeof line, f)) {
        fprintf(stderr, "%s: short header\n", in);
        return 1;
    }
    sscanf(line, "%d %d %d", &dims[0], &dims[1], &dims[2]);
    if (!fgets(line, sizeof line, f)) return 1;
    sscanf(line, "%lf %lf %lf %lf %lf", &eq[0], &eq[1], &eq[2], &eq[3], &eq[4]);
    fclose(f);

    struct params *p = (struct params *)calloc(1, sizeof(struct params));
    struct state *st = (struct state *)calloc(1, sizeof(struct state));
    struct bparams *bp = (struct bparams *)calloc(1, sizeof(struct bparams));
    long dimp = 1;
    for (int d = 0; d < NDIM; d++) {
        p->n[d] = dims[d];
        p->ng[d] = 2;
        dimp *= dims[d];
        p->g[d] = eq[2 + d];
    }
    p->gamma = eq[0];
    p->eta = eq[1];
    p->mu = 1.0;
    p->rkon = rk4;
    p->moddton = moddton;
    p->courant = courant;
    p->hyperdifmom = hyperdif;
    p->divbon = 0;
    for (int i = 0; i < NVAR; i++) p->chyp[i] = 0.0;
    p->chyp[rho] = chv[0];
    p->chyp[energy] = chv[1];
    for (int d = 0; d < NDIM; d++) {
        p->chyp[mom1 + d] = chv[2];
        p->chyp[b1 + d] = chv[3];
    }
    for (int ii = 0; ii < NVAR; ii++)
        for (int idir = 0; idir < NDIM; idir++)
            for (int ib = 0; ib < 2; ib++)
                p->boundtype[ii][idir][ib] = bcset ? (int)bcv[idir < bcset ? idir : bcset - 1] : 4;

    real *wmod = (real *)calloc((size_t)NVAR * dimp, sizeof(real));
    real *wd = (real *)calloc((size_t)NDERV * dimp, sizeof(real));
    char *hlines[5] = {0};
    struct Meta meta;
    memset(&meta, 0, sizeof(meta));
    meta.name = (char *)"smaug";
    CALL(readasciivacconfig((char *)in, *p, meta, st, wmod, wd, hlines, 0));
    for (int d = 0; d < NDIM; d++) {
        long stride = d == 0 ? 1 : (d == 1 ? dims[0] : (long)dims[0] * dims[1]);
        p->dx[d] = 0.5 * (wd[(long)(pos1 + d) * dimp + 2 * stride] - wd[(long)(pos1 + d) * dimp]);
        p->xmin[d] = wd[(long)(pos1 + d) * dimp];
        p->xmax[d] = wd[(long)(pos1 + d) * dimp + (dims[d] - 1) * stride];
    }
    p->dt = dtarg > 0 ? dtarg : 1e-4;
    p->qt = t0;
    real dt = p->dt;

    struct params *d_p;
    struct bparams *d_bp;
    struct state *d_state;
    real *d_wnew, *d_wmod, *d_dwn1, *d_wd, *d_wtemp, *d_wtemp1, *d_wtemp2, *wnew = NULL;
    CALL(cusetdriver(driver, (driver == 2 && !have_drv) ? NULL : drv));
    CALL(cuinit(&p, &bp, &wmod, &wnew, &wd, &st, &d_p, &d_bp, &d_wnew, &d_wmod, &d_dwn1, &d_wd,
                &d_state, &d_wtemp, &d_wtemp1, &d_wtemp2));
    CALL(initgrid(&p, &st, &wd, &d_p, &d_dwn1, &d_wd, &d_state, &d_wtemp, &d_wtemp1, &d_wtemp2));

    /* record the values of the fixed faces (boundtype 5) before the time
     * loop (iosmaug.cpp:645-659) */
    {
        int its = p->it;
        for (int ii = 0; ii < NVAR; ii++)
            for (int idir = 0; idir < NDIM; idir++)
                if (p->boundtype[ii][idir][0] == 5) {
                    p->it = -1;
                    CALL(cuboundary(&p, &bp, &d_p, &d_bp, &d_state, &d_wmod, 0, idir, ii));
                }
        p->it = its;
    }
    struct handles h = {p, d_p, bp, d_bp, d_state, wd, d_wmod, d_dwn1, d_wd, d_wtemp, d_wtemp1, d_wtemp2};
    /* ghost fill of the initial state (iosmaug.cpp:783-791) */
    fill_ghosts(&h, 1);

    if (p->rkon == 1) {
        CALL(cucomputedervfields(&p, &d_p, &d_wmod, &d_wd, 0));
        if (!rk4calls) CALL(cuadvancerk4(&p, &d_p, &d_wmod, nt, mixed));
        p->it = nt;
    }
    /* rkon = 1 through the entry points (-rk4calls): each step is SAC's
     * fourstep (dev/sac/src/vac.t:447-464), the scheme the reference's RK
     * branch (iosmaug.cpp:1182-1395) sets out to run.  Stage `order` reads
     * slot order and accumulates into slot order + 1 (slot 0 for the last)
     * with dt/2, dt/2, dt, dt/6 at times t, t + dt/2, t + dt/2, t + dt, and
     * cuadvance(order) applies the slot algebra after it: slot 2 <- slot 0,
     * slot 3 <- slot 0, slot 1 <- (slot 1 + 2 slot 2 + slot 3 - 4 slot 0)/3,
     * slot 0 <- slot 0 + slot 1.  Where the reference's branch departs from
     * fourstep this loop follows fourstep: every direction is advected, every
     * term takes the stage's dt (through p->dt), the ghosts of every field are
     * filled in the slot the next stage reads, and slot 1 starts each step as
     * a copy of slot 0, refreshed through the host (cuupdate, cuupdatemod)
     * because the stage-0 copy of cucomputedervfields runs the other way. */
    for (int n = 1; rk4calls && n <= nt; n++) {
        const real t = p->qt;
        const real sdt[4] = {dt / 2, dt / 2, dt, dt / 6};
        const real stime[4] = {t, t + dt / 2, t + dt / 2, t + dt};
        p->it = n;
        CALL(cuupdate(&p, &wmod, &wmod, &d_wtemp2, &st, &d_p, &d_wmod, &d_wmod, &d_wtemp2, &d_state, n));
        CALL(cuupdatemod(&p, &bp, &wmod, &wnew, &wd, &st, &d_p, &d_bp, &d_wmod, &d_wnew, &d_wmod,
                         &d_dwn1, &d_wd, &d_state, &d_wtemp, &d_wtemp1, &d_wtemp2));
        for (int order = 0; order < 4; order++) {
            const int ordero = order == 3 ? 0 : order + 1;
            p->dt = sdt[order];
            p->qt = stime[order];
            stage(&h, order, ordero, p->dt);
            CALL(cuadvance(&p, &d_p, &d_wmod, &d_wmod, order));
            fill_ghosts(&h, ordero);
        }
        p->dt = dt;
        p->qt = t + dt;
    }
    double elapsed = 0.0; /* sum of the steps' dt */
    FILE *flog = NULL;
    if (dtlog && !(flog = fopen(dtlog, "w"))) {
        fprintf(stderr, "cannot open %s\n", dtlog);
        return 1;
    }
    int nrun = p->rkon == 1 ? nt : 0;
    for (int n = 1; p->rkon == 0 && n <= nt; n++) {
        int order = 0, ordero = 1;
        p->it = n;
        CALL(cucomputedervfields(&p, &d_p, &d_wmod, &d_wd, 0));
        if (p->moddton == 1) {
            /* Courant and hyperviscous dt (iosmaug.cpp:855-898) */
            p->maxcourant = 0.0;
            for (int dim = 0; dim <= (NDIM - 1); dim++) {
                CALL(cucomputec(&p, &d_p, &d_wmod, &d_wd, order, dim));
                CALL(cucomputemaxc(&p, &d_p, &d_wmod, &d_wd, order, dim, &wd, &d_wtemp));
                CALL(cucomputemaxcourant(&p, &d_p, &d_wmod, &d_wd, order, dim, &wd, &d_wtemp));
            }
            if (((p->courant) / (p->maxcourant)) > 1.0e-8) p->dt = (p->courant) / (p->maxcourant);
            if (n > 1) CALL(cugetdtvisc1(&p, &d_p, &d_wmod, &wd, &d_wd, order, &d_wtemp, &d_wtemp1, &d_wtemp2));
            if ((p->dtdiffvisc) > 1.0e-8 && (p->dt) > (p->dtdiffvisc)) p->dt = (p->dtdiffvisc);
        }
        if (tend > 0 && elapsed + p->dt > tend) p->dt = tend - elapsed;
        p->qt = (p->qt) + dt;
        stage(&h, order, ordero, dt);
        fill_ghosts(&h, ordero);
        elapsed += p->dt;
        nrun = n;
        if (flog) fprintf(flog, "%d %.17g\n", n, p->dt);
        if (tend > 0 && elapsed >= tend) break;
    }
    if (flog) fclose(flog);
    /* the newest state is in slot 1 after an onestep loop: bring it to slot 0 */
    if (p->rkon == 0) CALL(cucomputedervfields(&p, &d_p, &d_wmod, &d_wd, 0));
    CALL(cuupdate(&p, &wmod, &wmod, &d_wtemp2, &st, &d_p, &d_wmod, &d_wmod, &d_wtemp2, &d_state, nrun));
    st->it = p->it;
    st->t = p->qt;
    printf("smaug_main: %d steps, t = %.10g, dt = %.17g, elapsed = %.17g\n", nrun, p->qt, p->dt,
           p->rkon == 1 ? nt * dt : elapsed);
    if (out) CALL(writeasciivacconfig((char *)out, *p, meta, wmod, wd, hlines, *st, 0));
    CALL(cufinish(&p, &wmod, &wnew, &st, &d_p, &d_bp, &d_wmod, &d_wnew, &d_wmod, &d_dwn1, &d_wd,
                  &d_state, &d_wtemp, &d_wtemp1, &d_wtemp2));
    return 0;
}
