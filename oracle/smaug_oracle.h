/*
 * smaug_oracle.h -- CPU restatement of SMAUG's per-step MHD update.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP path and
 * the CPU baseline timed by bench.py; it is never linked into, loaded by or
 * called from the product library.
 *
 * State layout (same as the reference's wmod slot, dev/smaug/include/iotypes.h:20-43):
 *   cell index c = i + n0*(j + n1*k), i (dim 0) fastest;
 *   evolved fields  (NVE = 2+2*ndim): rho', m_1..m_D, e', b'_1..b'_D
 *   background      (NB  = 2+ndim)  : e_b, rho_b, b_b1..b_bD
 *   field f is the contiguous block w[f*ncell .. (f+1)*ncell).
 * n[d] counts two ghost layers per side; the interior along d is [2, n[d]-3].
 */
#ifndef SMAUG_ORACLE_H
#define SMAUG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int ndim;            /* 2 or 3 */
    int n[3];            /* cells per dim incl. ghosts; n[2] == 1 when ndim == 2 */
    int bc[3];           /* per dim: 0 periodic, 3 cont, 4 contcd4 mirror, 5 fixed
                            (iosmaugparams boundtype) */
    int hyperdif;        /* p->hyperdifmom */
    int driver;          /* 0 none, 1 vertical Gaussian velocity driver,
                            2 torsional driver of usersource.cu (3-D) */
    double gamma;
    double g[3];
    double chyp[13];     /* indexed by the reference field enum (CEV) */
    double drv[8];       /* A, period, sigma_h, z0, sigma_z, xc (dim 1), yc (dim 2),
                            and for driver 2 the normalising length L */
    const double *dx[3]; /* cell metric per dim, length n[d]  (wd[delx1..3]) */
    const double *pos[3];/* cell centre per dim, length n[d]  (wd[pos1..3])  */
    const double *wfix;  /* bc 5: the evolved state whose ghost values the fixed
                            faces keep (NVE x ncell); NULL otherwise */
} sm_grid;

long sm_ncell(const sm_grid *g);
int  sm_nve(const sm_grid *g);
int  sm_nbg(const sm_grid *g);

/* ghost fill of every evolved field, dims in order 0..ndim-1 (cuboundary) */
void sm_boundary(const sm_grid *g, double *w);

/* cmax[d] = max over all cells of the fast speed along d plus |v_d| (cucomputemaxc) */
void sm_cmax(const sm_grid *g, const double *w, const double *bg, double cmax[3]);

/* Courant limit max_d max_c cfast_d/dx_d (cucomputemaxcourant) */
double sm_maxcourant(const sm_grid *g, const double *w, const double *bg);
/* the same for one direction d: max_c cfast_d/dx_d */
double sm_maxcourant_dim(const sm_grid *g, const double *w, const double *bg, int d);

/* largest hyperviscosity coefficient nu_R / nu_L over the cells and the
 * (field, dim) pairs of the hyperdiffusion (p->maxviscoef with moddton);
 * cmax may be NULL (then computed) */
double sm_numax(const sm_grid *g, const double *w, const double *bg, const double *cmax);
/* hyperviscous dt limit 0.25 min(dx^2) / numax (SAC getdt_visc); 0 if numax is 0 */
double sm_dtvisc(const sm_grid *g, double numax);

/* hyperviscosity coefficients nu_R, nu_L for one (field, dim) (cuhyperdifvisc1r/1l) */
void sm_nu(const sm_grid *g, const double *w, const double *bg, int field, int dim,
           double cmax_d, double *nuR, double *nuL);

/* 4th-order central difference of q along d with the reference's edge rule (grad3dn) */
void sm_grad4(const sm_grid *g, const double *q, int d, double *out);

/* dU/dt on interior cells, 0 on ghosts; cmax may be NULL (then computed) */
void sm_rhs(const sm_grid *g, const double *w, const double *bg, double t,
            const double *cmax, double *rhs);

/* w <- w + dt L(w), then ghost fill (the rkon=0 "onestep" path) */
void sm_step_euler(const sm_grid *g, double *w, const double *bg, double t, double dt,
                   double *scratch /* NVE*ncell */);

/* classical RK4 in the slot algebra of SAC's fourstep (vac.t:447-464) */
void sm_step_rk4(const sm_grid *g, double *w, const double *bg, double t, double dt,
                 double *scratch /* 5*NVE*ncell */);

/* run nsteps of RK4 (or Euler when rk4 == 0) with fixed dt; returns final time */
double sm_run(const sm_grid *g, double *w, const double *bg, double t0, double dt,
              int nsteps, int rk4);

int sm_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
