/*
 * smaug_hip.h -- drop-in C-ABI of the MI355X SMAUG step.
 *
 * Every entry point replaces the one of the same name and signature in the
 * reference's dev/smaug/include/smaugcukernels.h (cited per line below);
 * a driver written against that header (dev/smaug/src/iosmaug.cpp:802-1550)
 * links against libsmaug_hip2d.so (-DUSE_SAC) or libsmaug_hip3d.so
 * (-DUSE_SAC_3D) instead of the CUDA objects.
 *
 * Conventions kept from the reference:
 *  - everything passes by pointer-to-pointer; cuinit allocates and returns
 *    the device arrays through those arguments;
 *  - d_wmod holds slots of NVAR*dimp doubles (rkon=0: 3 slots, rkon=1: 6),
 *    field order and cell order as in iotypes.h; `order` is the slot read,
 *    `ordero` the slot updated;
 *  - scalar results (cmax, maxcourant, maxviscoef, dtdiffvisc) are written
 *    into the host struct *p;
 *  - calls are synchronous: each returns after its kernels complete.
 * Differences: every entry point returns 0 on success and a nonzero
 * SMAUG_E* code on failure (the reference's wrappers return nothing), with
 * the message in cuerrorstring(); d_wd/d_wtemp* hold this library's own
 * derived quantities (hdnur/hdnul, the ones the hyperdiffusion sources read,
 * keep the reference's slots).  Boundary types other than 0 (periodic) and
 * 4 (contcd4), divbon and shock viscosity are rejected with SMAUG_EINVAL.
 */
#ifndef SMAUG_HIP_H
#define SMAUG_HIP_H

#include "smaug_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* smaugcukernels.h:8 */
int cuinit(struct params **p, struct bparams **bp, real **wmod, real **wnew, real **wd,
           struct state **state, struct params **d_p, struct bparams **d_bp, real **d_wnew,
           real **d_wmod, real **d_dwn1, real **d_wd, struct state **d_state, real **d_wtemp,
           real **d_wtemp1, real **d_wtemp2);
/* smaugcukernels.h:9 */
int cuupdatemod(struct params **p, struct bparams **bp, real **w, real **wnew, real **wd,
                struct state **state, struct params **d_p, struct bparams **d_bp, real **d_w,
                real **d_wnew, real **d_wmod, real **d_dwn1, real **d_wd, struct state **d_state,
                real **d_wtemp, real **d_wtemp1, real **d_wtemp2);
/* smaugcukernels.h:10 */
int initgrid(struct params **p, struct state **state, real **wd, struct params **d_p,
             real **d_dwn1, real **d_wd, struct state **d_state, real **d_wtemp,
             real **d_wtemp1, real **d_wtemp2);
/* smaugcukernels.h:12 */
int cufinish(struct params **p, real **w, real **wnew, struct state **state,
             struct params **d_p, struct bparams **d_bp, real **d_w, real **d_wnew,
             real **d_wmod, real **d_dwn1, real **d_wd, struct state **d_state, real **d_wtemp,
             real **d_wtemp1, real **d_wtemp2);
/* smaugcukernels.h:14.  Ghost cells of one field along idir by its boundtype:
 * 0 periodic, 3 cont (both layers copy the face's edge cell), 4 contcd4,
 * 5 fixed.  A fixed face keeps the values that the call with p->it == -1
 * recorded (the driver makes that call once per fixed face before the time
 * loop, iosmaug.cpp:645-659); using one before it is recorded is an error. */
int cuboundary(struct params **p, struct bparams **bp, struct params **d_p,
               struct bparams **d_bp, struct state **d_s, real **d_wmod, int order, int idir,
               int field);
/* smaugcukernels.h:15 */
int cuupdate(struct params **p, real **w, real **wmod, real **wtemp2, struct state **state,
             struct params **d_p, real **d_w, real **d_wmod, real **d_wtemp2,
             struct state **d_state, int step);
/* smaugcukernels.h:17 */
int cuupdatehostwd(struct params **p, real **wd, real **wmod, real **wtemp2,
                   struct state **state, struct params **d_p, real **d_wd, real **d_wmod,
                   real **d_wtemp2, struct state **d_state, int step);
/* smaugcukernels.h:18 */
int cuupdatedevicewd(struct params **p, real **wd, real **wmod, real **wtemp2,
                     struct state **state, struct params **d_p, real **d_wd, real **d_wmod,
                     real **d_wtemp2, struct state **d_state, int step);
/* smaugcukernels.h:22 */
int cucentdiff1(struct params **p, struct params **d_p, struct state **d_s, real **d_w,
                real **d_wmod, real **d_dwn1, real **d_wd, int order, int ordero, real dt,
                int field, int dir);
/* smaugcukernels.h:23 */
int cucentdiff2(struct params **p, struct params **d_p, struct state **d_s, real **d_w,
                real **d_wmod, real **d_dwn1, real **d_wd, int order, int ordero, real dt,
                int field, int dir);
/* smaugcukernels.h:25 */
int cugrav(struct params **p, struct params **d_p, struct state **d_s, real **d_w,
           real **d_wmod, real **d_dwn1, real **d_wd, int order, int ordero, real dt);
/* smaugcukernels.h:26 */
int cusource(struct params **p, struct params **d_p, struct state **d_s, real **d_w,
             real **d_wmod, real **d_dwn1, real **d_wd, int order, int ordero, real dt);
/* smaugcukernels.h:28 */
int cuadvance(struct params **p, struct params **d_p, real **d_wmod, real **d_w, int order);
/* smaugcukernels.h:29 */
int cucomputedervfields(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                        int order);
/* smaugcukernels.h:30 */
int cucomputevels(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                  int order, int dir);
/* smaugcukernels.h:31 */
int cucomputemaxc(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                  int order, int dir, real **wd, real **d_wtemp);
/* smaugcukernels.h:32 */
int cucomputemaxcourant(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                        int order, int dir, real **wd, real **d_wtemp);
/* smaugcukernels.h:33 */
int cucomputec(struct params **p, struct params **d_p, real **d_wmod, real **d_wd, int order,
               int dir);
/* smaugcukernels.h:34 */
int cucomputept(struct params **p, struct params **d_p, real **d_wmod, real **d_wd, int order,
                int dir);
/* smaugcukernels.h:35 */
int cucomputepk(struct params **p, struct params **d_p, real **d_wmod, real **d_wd, int order,
                int dir);
/* smaugcukernels.h:36 */
int cucomputepbg(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                 int order, int dir);
/* smaugcukernels.h:37 */
int cudivb(struct params **p, struct params **d_p, real **d_w, real **d_wmod, real **d_dwn1,
           real **d_wd, int order, int ordero, real dt);
/* smaugcukernels.h:39 */
int cuhyperdifmomsource1(struct params **p, struct params **d_p, real **d_wmod,
                         real **d_dwn1, real **d_wd, int order, int ordero, real **d_wtemp,
                         int field, int dim, int ii, int ii0, real dt);
/* smaugcukernels.h:41 */
int cuhyperdifmomsourcene1(struct params **p, struct params **d_p, real **d_wmod,
                           real **d_dwn1, real **d_wd, int order, int ordero, real **d_wtemp,
                           int field, int dim, int ii, int ii0, real dt);
/* smaugcukernels.h:43 */
int cuhyperdifesource1(struct params **p, struct params **d_p, real **d_wmod, real **d_dwn1,
                       real **d_wd, int order, int ordero, real **d_wtemp, int field, int dim,
                       real dt);
/* smaugcukernels.h:45 */
int cuhyperdifbsource1(struct params **p, struct params **d_p, real **d_wmod, real **d_dwn1,
                       real **d_wd, int order, int ordero, real **d_wtemp, int field, int dim,
                       int jj, int ii0, int mm, real sb, real dt);
/* smaugcukernels.h:47 */
int cuhyperdifbsourcene1(struct params **p, struct params **d_p, real **d_wmod,
                         real **d_dwn1, real **d_wd, int order, int ordero, real **d_wtemp,
                         int field, int dim, int jj, int ii0, int mm, real sb, real dt);
/* smaugcukernels.h:50 */
int cuhyperdifrhosource1(struct params **p, struct params **d_p, real **d_wmod,
                         real **d_dwn1, real **d_wd, int order, int ordero, real **d_wtemp,
                         int field, int dim, real dt);
/* smaugcukernels.h:54 */
int cunushk1(struct params **p, struct params **d_p, real **d_wmod, real **d_wd, int order,
             real **d_wtemp, real **d_wtemp1, real **d_wtemp2);
/* smaugcukernels.h:55 */
int cugetdtvisc1(struct params **p, struct params **d_p, real **d_wmod, real **wd,
                 real **d_wd, int order, real **d_wtemp, real **d_wtemp1, real **d_wtemp2);
/* smaugcukernels.h:57 */
int cuhyperdifvisc1(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                    int order, real **d_wtemp, real **d_wtemp1, real **d_wtemp2, int field,
                    int dim, int hand);
/* smaugcukernels.h:58 */
int cuhyperdifvisc1r(struct params **p, struct params **d_p, real **d_wmod, real **wd,
                     real **d_wd, int order, real **d_wtemp, real **d_wtemp1,
                     real **d_wtemp2, int field, int dim);
/* smaugcukernels.h:59 */
int cuhyperdifvisc1l(struct params **p, struct params **d_p, real **d_wmod, real **wd,
                     real **d_wd, int order, real **d_wtemp, real **d_wtemp1,
                     real **d_wtemp2, int field, int dim);
/* smaugcukernels.h:60 */
int cuhyperdifvisc1ir(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                      int order, real **d_wtemp, real **d_wtemp1, real **d_wtemp2, int field,
                      int dim);
/* smaugcukernels.h:61 */
int cuhyperdifvisc1il(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                      int order, real **d_wtemp, real **d_wtemp1, real **d_wtemp2, int field,
                      int dim);
/* smaugcukernels.h:64 */
int cuhyperdifviscmax(struct params **p, real **w, real **wnew, struct params **d_p,
                      real **d_w, real **d_wnew, real **d_wmod, real **d_dwn1, real **d_wd,
                      int order, real **d_wtemp, int field, int dim);
/* smaugcukernels.h:65 */
int cusync(struct params **p);

/*
 * Extensions (no counterpart in the reference header).
 *
 * cuadvancerk4: nsteps of classical RK4 (the intended semantics of the
 * reference's defective rkon=1 branch, iosmaug.cpp:1182-1395) applied to
 * slot 0 with the fused stage kernels, replayed from a hipGraph.  precision
 * 0 = FP64, 1 = FP32 perturbations / FP64 background, pressure and cmax.
 * Slot 0 and slot 1 hold the result; p->qt advances by nsteps*p->dt.
 * The user source is the driver configured by cusetdriver (off by
 * default).
 */
int cuadvancerk4(struct params **p, struct params **d_p, real **d_wmod, int nsteps,
                 int precision);
/* driver parameters: amplitude, period, horizontal width, height, vertical
 * width, horizontal centre (dims 1 and 2); enable = 0 switches it off,
 * 1 selects the vertical Gaussian driver, 2 (3-D only) the reference's
 * torsional driver (usersource.cu:4-86: vx = (A yp/L) G sin(2 pi qt/P)
 * along dim 1, vy = -(A xp/L) G sin(2 pi qt/P) along dim 2, L = 4 Mm).
 * With enable = 2 and drv = NULL the reference's constants are used
 * (A = 500, P = 30, widths 0.1 Mm, height 0.1 Mm, vertical width
 * 0.05 Mm, centre 2 Mm, 2 Mm). */
int cusetdriver(int enable, const real drv[7]);
/* message of the last failure in this thread */
const char *cuerrorstring(void);

#ifdef __cplusplus
}
#endif
#endif
