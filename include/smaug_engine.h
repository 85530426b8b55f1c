/*
 * smaug_engine.h -- C-ABI of the MI355X stepping engine behind the
 * reference's cu* entry points (declared in smaug_hip.h).
 *
 * The reference has no counterpart of this handle API: its state lives in
 * globals created by cuinit (dev/smaug/src/init.cu:816-1068).  The engine
 * is the same state made explicit, so that one process can hold several
 * grids (the ensemble) and so tests can drive the path with plain arrays.
 *
 * Array conventions (same as the reference's wmod slot, iotypes.h:20-43):
 *   cell c = i + n0*(j + n1*k), i fastest; n[d] counts 2 ghost layers per side.
 *   evolved w  : (2+2*ndim) fields  rho', m_1..m_D, e', b'_1..b'_D
 *   background : (2+ndim)   fields  e_b, rho_b, b_b1..b_bD
 *   field f occupies [f*ncell, (f+1)*ncell).  Host arrays are always FP64.
 * Every function returns 0 on success and a nonzero SMAUG_E* code on
 * failure; smaug_last_error() gives the message.
 */
#ifndef SMAUG_ENGINE_H
#define SMAUG_ENGINE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SMAUG_OK = 0,
    SMAUG_EINVAL = 1,   /* bad argument or unsupported configuration */
    SMAUG_EHIP = 2,     /* a HIP runtime call failed */
    SMAUG_ENODEV = 3,   /* no usable gfx950 device */
    SMAUG_ENOMEM = 4    /* device allocation failed */
};

enum { SMAUG_FP64 = 0, SMAUG_MIXED = 1 };

typedef struct {
    int ndim;             /* 2 or 3 */
    int n[3];             /* cells per dim incl. ghosts, each >= 8; n[2] ignored for ndim 2 */
    int bc[3];            /* boundtype per dim: 0 periodic, 4 contcd4 */
    int hyperdif;         /* p->hyperdifmom */
    int driver;           /* 0 none, 1 vertical Gaussian velocity driver,
                             2 the torsional driver of usersource.cu (3-D) */
    int precision;        /* SMAUG_FP64 or SMAUG_MIXED */
    double gamma;
    double g[3];
    double chyp[13];      /* indexed by the reference field enum */
    double drv[8];        /* A, period, sigma_h, z0, sigma_z, xc, yc; for driver 2
                             also L, the normalising length of the lever arm */
    const double *dx[3];  /* cell metric per dim, length n[d] (copied) */
    const double *pos[3]; /* cell centres per dim, length n[d] (copied) */
} smaug_cfg;

typedef struct smaug_engine smaug_engine;

const char *smaug_last_error(void);

int smaug_create(const smaug_cfg *cfg, smaug_engine **out);
int smaug_destroy(smaug_engine *e);

/* upload state (FP64 host arrays; rounded to FP32 for evolved fields in
 * mixed mode) and fill the ghosts.  Along a periodic dim the background of
 * every ghost cell must equal that of its periodic image (the ghost fill
 * copies the perturbation only); SMAUG_EINVAL otherwise. */
int smaug_set_state(smaug_engine *e, const double *w, const double *bg);
/* download the evolved state of the current step (FP64 host array) */
int smaug_get_state(smaug_engine *e, double *w);

/* L(w) = dU/dt of the current state on interior cells, 0 on ghosts */
int smaug_rhs(smaug_engine *e, double t, double *rhs);
/* cmax[d] over all cells of the current state */
int smaug_cmax(smaug_engine *e, double cmax[3]);
/* max over dims and cells of cfast_d/dx_d (Courant limit) */
int smaug_maxcourant(smaug_engine *e, double *out);

/* advance nsteps with fixed dt: classical RK4 / forward Euler.  Steps are
 * replayed from a captured hipGraph; the call returns after the last step
 * completes. */
int smaug_rk4(smaug_engine *e, double t0, double dt, int nsteps);
int smaug_euler(smaug_engine *e, double t0, double dt, int nsteps);

/* time nsteps of RK4: total milliseconds of the step loop (HIP events on the
 * engine stream) and the mean duration of one stage-update kernel launch
 * (HIP events around each launch, eager launches) */
int smaug_time_rk4(smaug_engine *e, double t0, double dt, int nsteps,
                   double *total_ms, double *stage_kernel_ms);

/* Launch shape.  The ratio (pass 0) and update (pass 1) kernels cover the
 * grid with 64 x 4-cell tiles grouped in XCD panels of tpx x tpy tiles, each
 * panel swept plane by plane on one XCD.  smaug_set_panel fixes a panel
 * (0, 0 = default), smaug_get_panel reports the one in use,
 * smaug_time_pass times reps launches of a pass on the current state (the
 * state is not changed), and smaug_autotune times every panel of
 * tpx in {1,2,4,8} x tpy in {1,..,32} and, for 3-D grids, the z-streamed
 * ratio kernel at depths 8..128 and the z-streamed update at depths 8..128
 * with either form of the hyperdiffusive-source kernel, and keeps the
 * fastest form per pass; then ratio form 2 (smaug_set_ratio_form) at depths
 * 4..32, kept if the ratio and update passes take less time together with
 * it.  best_ms[p] is the mean time of pass p in the forms kept. */
int smaug_set_panel(smaug_engine *e, int pass, int tpx, int tpy);
int smaug_get_panel(smaug_engine *e, int pass, int *tpx, int *tpy);
int smaug_time_pass(smaug_engine *e, int pass, int reps, double *ms);
int smaug_autotune(smaug_engine *e, int reps, int best[2][2], double best_ms[2]);
/* 3-D grids: the stage update runs z-streamed (enable = 1, the default;
 * blocks of 64 x 4 columns walk kz output planes, kz = 0 picks 32, depths
 * above 256 run as 256; the z-streamed source kernel uses the same depth)
 * with the
 * z-streamed hyperdiffusive-source kernel, z-streamed with the per-cell
 * (enable = 2) or the LDS-plane (enable = 3) source kernel, or per cell
 * (enable = 0; the update pass then uses the XCD panel) */
int smaug_set_zstream(smaug_engine *e, int enable, int kz);
/* 3-D grids: the hyperviscosity ratios z-streamed (1; kz output planes per
 * block, 0 picks 32) or per cell (0).  zstreamed = 2 (the default): the ratio
 * pass forms the seven pairs along k alone, z-streamed, and the z-streamed
 * hyperdiffusive-source kernel forms the fourteen in-plane pairs itself, in
 * LDS, from the state around its tile -- the same ratios without their trip
 * through HBM.  Form 2 takes effect with the z-streamed update and source
 * kernel (smaug_set_zstream enable = 1); with any other form of the update
 * the ratio pass forms every pair, as form 1.  2-D grids: 2 (the default)
 * the update kernel forms the ten pairs of its tile and a 1-cell halo in
 * LDS and the stage has no ratio pass, 0 the ratio pass writes them to HBM
 * and the update reads them there; same ratios either way (form 1 is
 * rejected). */
int smaug_set_ratio_form(smaug_engine *e, int zstreamed, int kz);
/* the form of the ratio pass in use (0, 1 or 2) and its depth (0 per cell) */
int smaug_get_ratio_form(smaug_engine *e, int *zstreamed, int *kz);
/* the form of the 3-D update pass in use (enable as above) and its depth */
int smaug_get_zstream(smaug_engine *e, int *enable, int *kz);
/* 3-D grids: when the uploaded background is the same on every k-plane (a
 * stratified atmosphere whose vertical axis is i), smaug_set_state also
 * keeps it as one k-plane, and the z-streamed source and update kernels
 * and the ghost fill read that plane, once per column, instead of the full
 * background (enable = 1, the default; 0 always reads the full
 * arrays).  Same arithmetic either way.  smaug_get_bg_plane reports whether
 * the plane is in use. */
int smaug_set_bg_plane(smaug_engine *e, int enable);
int smaug_get_bg_plane(smaug_engine *e, int *active);

/* conservation diagnostics: sums[f] = sum over the interior cells of evolved
 * field f of the current state, accumulated in FP64 with compensated
 * (Neumaier) summation on the device and over the per-block partials */
int smaug_sums(smaug_engine *e, double *sums);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the sums of
 * smaug_sums for an evolved state d_w that the caller holds on the device
 * (layout of smaug_set_state; double when precision is SMAUG_FP64, float when
 * SMAUG_MIXED; n[2] ignored for ndim 2), written as 2+2*ndim doubles to the
 * device array d_sums.  The work is enqueued on `stream` (a hipStream_t;
 * NULL for the default stream) and the call returns without waiting for it.
 * Allocates nothing: the per-block partials live in a static device array of
 * 64 KiB that the library holds on each device it runs on until the process
 * ends.  Every call on one device shares that array, so calls that overlap
 * on one device (other streams, other host threads) race and must be
 * ordered by the caller.  Grids above 2^29 cells are rejected. */
int smaug_sums_noalloc(int ndim, const int n[3], int precision, const void *d_w,
                       double *d_sums, void *stream);

/* one more RK4 step, with HIP events between its passes: ms[p] is the mean
 * duration over the four stages of pass p = 0 aux quantities and cmax,
 * 1 hyperviscosity ratios, 2 hyperdiffusive sources (3-D z-streamed form
 * only; 0 otherwise), 3 stage update, 4 ghost fill.  Where the stage update
 * also computes the aux quantities of its output (3-D grids with the
 * k-plane background; 2-D stages after the first), pass 0 is empty there and
 * passes 3 and 4 carry that work. */
int smaug_time_passes(smaug_engine *e, double t0, double dt, double ms[5]);

/* measured copy-bandwidth ceiling: GB/s (read + write bytes) of a streaming
 * copy of nbytes, mean of reps launches, the best over 1, 4 or 8 16-byte
 * loads in flight per thread and 8, 16 or 32 workgroups per CU */
int smaug_copy_probe(long long nbytes, int reps, double *gbs);
/* measured many-stream bandwidth: GB/s (read + write bytes) of a kernel that
 * reads one FP32 value of each of 32 arrays of ncell values and writes one
 * of each of 20 arrays per cell (the streams of a mixed 3-D stage update),
 * mean of reps launches, the best over 8, 16 or 32 workgroups per CU.  It
 * allocates 208 bytes per cell. */
int smaug_stream_probe(long long ncell, int reps, double *gbs);
/* the same streams walked as the z-streamed stage update walks them: GB/s
 * (read + write bytes of the interior) of blocks of 64 x 4 columns of an
 * n0 x n1 x n2 grid, each walking a chunk of kz planes with the next
 * plane's loads issued before the current plane's stores, the tiles of an
 * XCD contiguous; mean of reps launches.  Allocates 208 bytes per cell. */
int smaug_walk_probe(int n0, int n1, int n2, int kz, int reps, double *gbs);

/* sizes */
long long smaug_ncell(const smaug_engine *e);
long long smaug_ninterior(const smaug_engine *e);

/* ---------------------------------------------------------------------------
 * Ensemble: nrun independent grids of uneven sizes advanced together.  The
 * reference runs one grid per process (iosmaug.cpp:786-1168) and has no
 * counterpart.  Every run keeps its own n, bc, dx, pos, time and dt; ndim,
 * precision, hyperdif, driver, gamma, g, chyp and drv must agree across runs.
 * Each pass of an RK stage is one persistent launch pulling (run, tile)
 * items from a device queue; one RK4 step of all runs replays a hipGraph.
 * A member evolves bit-for-bit like the same grid in smaug_engine.
 * ------------------------------------------------------------------------- */
typedef struct smaug_ensemble smaug_ensemble;

int smaug_ens_create(const smaug_cfg *cfgs, int nrun, smaug_ensemble **out);
int smaug_ens_destroy(smaug_ensemble *e);
int smaug_ens_nrun(const smaug_ensemble *e);
/* interior cells summed over runs */
long long smaug_ens_ninterior(const smaug_ensemble *e);
/* state of one run, arrays as for smaug_set_state / smaug_get_state */
int smaug_ens_set_state(smaug_ensemble *e, int run, const double *w, const double *bg);
int smaug_ens_get_state(smaug_ensemble *e, int run, double *w);
/* conservation diagnostics of one run, as smaug_sums */
int smaug_ens_sums(smaug_ensemble *e, int run, double *sums);
/* nsteps of classical RK4; t0 (may be NULL for all zero) and dt hold one
 * value per run */
int smaug_ens_rk4(smaug_ensemble *e, const double *t0, const double *dt, int nsteps);
/* one more RK4 step of every run with HIP events between the passes: ms[p]
 * is the mean duration over the four stages of pass p = 0 aux and cmax,
 * 1 ratios, 2 stage update, 3 ghost fill */
int smaug_ens_time_passes(smaug_ensemble *e, const double *t0, const double *dt, double ms[4]);
/* as smaug_ens_rk4, and the milliseconds of the step loop (HIP events) */
int smaug_ens_time_rk4(smaug_ensemble *e, const double *t0, const double *dt, int nsteps,
                       double *total_ms);
/* 2-D runs with hyperdiffusion: 2 (the default) the stage update pulls
 * tiles of 64 x 8 cells (FP32; 64 x 4 in FP64) and forms their ratio pairs itself, in LDS, and the
 * stage has no ratio pass; 0 the ratio pass writes them to HBM.  Same
 * ratios either way.  smaug_ens_get_ratio_form reports the form in use. */
int smaug_ens_set_ratio_form(smaug_ensemble *e, int form);
int smaug_ens_get_ratio_form(smaug_ensemble *e, int *form);

#ifdef __cplusplus
}
#endif
#endif
