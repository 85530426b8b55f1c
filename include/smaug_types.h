/*
 * smaug_types.h -- parameter and state types of the drop-in C-ABI.
 *
 * Same type names, field names and enum values as the reference's
 * dev/smaug/include/iotypes.h:86-194, 213-240 and iobparams.h:9-35, so that
 * driver code written against the reference (its params files assign
 * p->n[], p->dt, p->chyp[rho], p->boundtype[f][d][b], ...) compiles
 * unchanged.  As in the reference the layout depends on the build's
 * dimensionality: compile with -DUSE_SAC (2-D, NVAR 10) or -DUSE_SAC_3D
 * (3-D, NVAR 13); libsmaug_hip2d.so / libsmaug_hip3d.so are the matching
 * libraries.  real is double (USE_REAL, single-precision builds of the
 * reference, has no counterpart: the reduced-precision path here is the
 * engine's mixed mode, selected per run).
 */
#ifndef SMAUG_TYPES_H
#define SMAUG_TYPES_H

typedef double real;

#if defined(USE_SAC_3D)
#define NDIM 3
#define NVECDIM 3
#define NVAR 13
#define NDERV 23
#elif defined(USE_SAC)
#define NDIM 2
#define NVECDIM 2
#define NVAR 10
#define NDERV 19
#else
#error "define USE_SAC (2-D) or USE_SAC_3D (3-D) as for the reference build"
#endif
#define NTEMP 8
#define NTEMP1 2
#define NTEMP2 1

#ifndef PI
#define PI 3.14159265358979
#endif

struct params {
    int n[NDIM];      /* cells per dim including 2 ghost layers per side */
    int ng[NDIM];
    int npgp[NDIM];
    int fullgridini;

    real hdmean, hdmax, pmax, pmin, pmean, emax, emin, emean;

    real xmax[NDIM];
    real xmin[NDIM];
    int nt;
    int it;
    real tmax;
    real qt;          /* simulation time seen by the user source */

    real boundu[NDIM][NVAR];
    real boundl[NDIM][NVAR];

    real cmax;        /* written by cucomputemaxc */
    real maxcourant;  /* written by cucomputemaxcourant */
    int steeringenabled;
    int finishsteering;
    real dt;
    real dx[NDIM];

    real gamma;
    real mu;
    real eta;
    real g[NDIM];

    int sodifon;
    int rkon;         /* 0: onestep (forward Euler), 1: classical RK4 */
    int moddton;      /* 1: Courant + hyperviscous dt control */
    int divbon;
    int divbfix;
    int cfgsavefrequency;
    int hyperdifmom;  /* 1: hyperdiffusion on */
    int mode;
    int readini;
    real courant;
    real maxviscoef;  /* written by cuhyperdifvisc1r/1l when moddton == 1 */
    real dtdiffvisc;  /* written by cugetdtvisc1 */
    real chyp[NVAR];
    real chyp3;
    real test;
    int boundtype[NVAR][NDIM][2]; /* 0 periodic, 3 cont, 4 contcd4, 5 fixed, ... */

    int gpid[16];
    int npe;
    int noghost;
};

struct state {
    int it;
    real t;
    real dt;
    real rho, m1, m2, m3, e, b1, b2, b3;
};

#define MSIZE1 128
#define MSIZE2 128
#define MSIZE3 128
struct bparams {
#ifdef USE_SAC_3D
    real fixed1[4 * MSIZE2 * MSIZE3 * NVAR];
    real fixed2[4 * MSIZE1 * MSIZE3 * NVAR];
    real fixed3[4 * MSIZE1 * MSIZE2 * NVAR];
#else
    real fixed1[4 * MSIZE2 * NVAR];
    real fixed2[4 * MSIZE1 * NVAR];
#endif
};

/* run metadata carried by the config-file writers (iotypes.h:58-70) */
struct Meta {
    char *directory;
    char *author;
    char *sdate;
    char *platform;
    char *desc;
    char *name;
    char *ini_file;
    char *log_file;
    char *out_file;
};

typedef enum mode { run, scatter, gather, init, redistribute } MODE;

#ifdef USE_SAC_3D
typedef enum vars { rho, mom1, mom2, mom3, energy, b1, b2, b3, energyb, rhob, b1b, b2b, b3b } CEV;
typedef enum dvars {
    vel1, vel2, vel3, flux, hdnur, hdnul, nushk1, nushk2, nushk3, soundspeed, pressuret,
    pressurek, bdotv, divb, cfast, ptb, pkb, pos1, pos2, pos3, delx1, delx2, delx3
} DEV;
#else
typedef enum vars { rho, mom1, mom2, energy, b1, b2, energyb, rhob, b1b, b2b } CEV;
typedef enum dvars {
    vel1, vel2, flux, hdnur, hdnul, nushk1, nushk2, soundspeed, pressuret, pressurek, bdotv,
    divb, cfast, ptb, pkb, pos1, pos2, delx1, delx2
} DEV;
#endif
typedef enum tempvars { tmp1, tmp2, tmp3, tmp4, tmp5, tmp6, tmp7, tmp8 } TEV;
typedef enum temp1vars { d1, d3 } TEV1;
typedef enum temp2vars { tmpnui, tmpnui1, tmpnui2 } TEV2;

typedef struct params Params;
typedef struct params params;
typedef struct state state;
typedef struct Meta meta;

#endif
