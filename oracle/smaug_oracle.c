/*
 * smaug_oracle.c -- CPU restatement (FP64, OpenMP) of SMAUG's per-step MHD
 * operator, its ghost fill, and the onestep / classical RK4 integrators.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (tests/, smoke())
 * and the CPU baseline leg of bench.py.  Nothing in the product links it.
 *
 * Each routine follows the reference kernel it cites.  The operator is
 * written array-at-a-time the way the reference's kernel sequence is:
 * derived arrays over the whole grid, then per-(field, direction) updates
 * restricted to the index ranges the reference kernels use.  Instead of
 * accumulating "w[slot1] -= dt*term" term by term, the terms are summed
 * into rhs = dU/dt, so L(w) is a function of one state and can drive both
 * the forward-Euler step (the reference's rkon=0 loop) and RK4.
 *
 * Deviations from the reference, each forced by a defect there:
 *  - hyperdifvisc1r reads the pad cell beyond the last ghost along dim
 *    (encode3p2 index n+1) that is never filled (bc_hyperdifr is not
 *    launched, and hyperdifvisc1r_parallel races a zeroing against a
 *    mis-strided write, hyperdifvisc1r.cu:1105-1253).  Here that pad holds
 *    the mirror value q[n-6], which is what bc_hyperdifr defines
 *    (hyperdifvisc1r.cu:17-60) and what SAC's setnu does
 *    (vacusr.viscosity.t:377-395); the left pad is q[5] as in
 *    bc_hyperdifl (hyperdifvisc1l.cu:17-60).
 *  - the user driver reads rho from the stage input state (the reference
 *    reads the partly updated output slot, usersource.cu:83-87) and takes
 *    the stage time t (the reference advances p->qt by the initial dt,
 *    iosmaug.cpp:918).
 *  - RK4 follows SAC's fourstep (vac.t:447-464); SMAUG's rkon=1 branch
 *    skips the last direction and gravity (iosmaug.cpp:1205-1395).
 */
#include "smaug_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* reference field enum (iotypes.h:213-221) as functions of ndim */
#define F_RHO 0
#define F_MOM(D, d) (1 + (d))
#define F_EN(D) (1 + (D))
#define F_B(D, d) (2 + (D) + (d))
/* background block */
#define G_EB 0
#define G_RHOB 1
#define G_BB(d) (2 + (d))

typedef struct {
    int D;
    int n[3];
    long s[3];
    long nc;
} geo;

static geo mkgeo(const sm_grid *g)
{
    geo G;
    G.D = g->ndim;
    G.n[0] = g->n[0];
    G.n[1] = g->n[1];
    G.n[2] = (g->ndim == 3) ? g->n[2] : 1;
    G.s[0] = 1;
    G.s[1] = G.n[0];
    G.s[2] = (long)G.n[0] * G.n[1];
    G.nc = (long)G.n[0] * G.n[1] * G.n[2];
    return G;
}

long sm_ncell(const sm_grid *g) { return mkgeo(g).nc; }
int sm_nve(const sm_grid *g) { return 2 + 2 * g->ndim; }
int sm_nbg(const sm_grid *g) { return 2 + g->ndim; }

int sm_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

#define FORALL(G)                                                              \
    _Pragma("omp parallel for collapse(2) schedule(static)")                   \
    for (int k_ = 0; k_ < (G).n[2]; k_++)                                      \
        for (int j_ = 0; j_ < (G).n[1]; j_++)                                  \
            for (int i_ = 0; i_ < (G).n[0]; i_++)

#define CIDX(G) ((long)i_ + (G).s[1] * j_ + (G).s[2] * k_)

static inline int coord(int d, int i, int j, int k) { return d == 0 ? i : (d == 1 ? j : k); }

static inline int interior(const geo *G, int i, int j, int k)
{
    if (i < 2 || i > G->n[0] - 3) return 0;
    if (j < 2 || j > G->n[1] - 3) return 0;
    if (G->D == 3 && (k < 2 || k > G->n[2] - 3)) return 0;
    return 1;
}

static double *alloc(long n)
{
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    if (!p) abort();
    return p;
}

/* ------------------------------------------------------------------------
 * Edge rules of the stencil operators for a non-periodic direction.
 * grad3dn (gradops.cu:437-602): zero at x == n-3, x == 2 always and at
 * x == n-4, x == 3 when the transverse indices satisfy 1 < t < nt-2.
 * grad1l3n / grad1r3n / grad13n (gradops.cu:608-830): zero at x == n-2,
 * x == 1 always and at x == n-3, x == 2 when 0 < t < nt-1, with the few
 * per-direction variations of the reference's masks written out below.
 * -------------------------------------------------------------------- */
enum { K_G4 = 0, K_G1L = 1, K_G1R = 2, K_G13 = 3 };

static int edge_zero(const geo *G, int kind, int d, int i, int j, int k)
{
    int x = coord(d, i, j, k), n = G->n[d];
    int ii[3] = {i, j, k};
    if (kind == K_G4) {
        int ok = 1;
        for (int t = 0; t < G->D; t++)
            if (t != d && !(ii[t] > 1 && ii[t] < G->n[t] - 2)) ok = 0;
        if (x == n - 3 || (x == n - 4 && ok)) return 1;
        if (x == 2 || (x == 3 && ok)) return 1;
        return 0;
    }
    /* first-order / second-order one-cell operators */
    int okhi = 1, oklo = 1;
    for (int t = 0; t < G->D; t++) {
        if (t == d) continue;
        int lo_hi = 0, lo_lo = 0; /* lower bound used by the masks: t > lo */
        if (G->D == 3) {
            if (kind == K_G1L && d == 0 && t == 2) lo_hi = 1; /* "k>1" in the upper mask */
            if (kind == K_G13 && d == 2 && t == 1) lo_lo = 1; /* "j>1" in the lower mask */
        }
        if (!(ii[t] > lo_hi && ii[t] < G->n[t] - 1)) okhi = 0;
        if (!(ii[t] > lo_lo && ii[t] < G->n[t] - 1)) oklo = 0;
    }
    if (x == n - 2 || (x == n - 3 && okhi)) return 1;
    if (x == 1 || (x == 2 && oklo)) return 1;
    return 0;
}

/* grad3dn: ((8(q+1 - q-1) + q-2 - q+2)/6)/(2 dx), defined for 1 < x < n-2 */
static void g4(const geo *G, const sm_grid *g, const double *q, int d, double *out)
{
    long s = G->s[d];
    int per = (g->bc[d] == 0);
    FORALL(*G)
    {
        long c = CIDX(*G);
        int x = coord(d, i_, j_, k_);
        double r = 0.0;
        if (x > 1 && x < G->n[d] - 2) {
            r = ((8 * q[c + s] - 8 * q[c - s] + q[c - 2 * s] - q[c + 2 * s]) / 6.0) /
                (2.0 * g->dx[d][x]);
            if (!per && edge_zero(G, K_G4, d, i_, j_, k_)) r = 0.0;
        }
        out[c] = r;
    }
}

void sm_grad4(const sm_grid *g, const double *q, int d, double *out)
{
    geo G = mkgeo(g);
    g4(&G, g, q, d, out);
}

/* grad1r3n: (q+1 - q)/dx for x < n-1 */
static void g1r(const geo *G, const sm_grid *g, const double *q, int d, double *out)
{
    long s = G->s[d];
    int per = (g->bc[d] == 0);
    FORALL(*G)
    {
        long c = CIDX(*G);
        int x = coord(d, i_, j_, k_);
        double r = 0.0;
        if (x < G->n[d] - 1) {
            r = (q[c + s] - q[c]) / g->dx[d][x];
            if (!per && edge_zero(G, K_G1R, d, i_, j_, k_)) r = 0.0;
        }
        out[c] = r;
    }
}

/* grad1l3n: (q - q-1)/dx for x > 0 */
static void g1l(const geo *G, const sm_grid *g, const double *q, int d, double *out)
{
    long s = G->s[d];
    int per = (g->bc[d] == 0);
    FORALL(*G)
    {
        long c = CIDX(*G);
        int x = coord(d, i_, j_, k_);
        double r = 0.0;
        if (x > 0) {
            r = (q[c] - q[c - s]) / g->dx[d][x];
            if (!per && edge_zero(G, K_G1L, d, i_, j_, k_)) r = 0.0;
        }
        out[c] = r;
    }
}

/* grad13n: ((q+1 - q-1)/dx)/2 for 0 < x < n-1 */
static void g13(const geo *G, const sm_grid *g, const double *q, int d, double *out)
{
    long s = G->s[d];
    int per = (g->bc[d] == 0);
    FORALL(*G)
    {
        long c = CIDX(*G);
        int x = coord(d, i_, j_, k_);
        double r = 0.0;
        if (x > 0 && x < G->n[d] - 1) {
            r = (q[c + s] - q[c - s]) / g->dx[d][x] / 2.0;
            if (!per && edge_zero(G, K_G13, d, i_, j_, k_)) r = 0.0;
        }
        out[c] = r;
    }
}

/* ------------------------------------------------------------------------
 * Ghost fill.  contcd4 (boundary.cu:16-88): 0 <- 4, 1 <- 3, n-1 <- n-5,
 * n-2 <- n-4.  periodic (gradops.cu:1271-1317, models/boundary.cu.ot):
 * 0,1 <- n-4,n-3 and n-2,n-1 <- 2,3.  cont (SAC vacgrid.t "cont"): both
 * ghost layers of a face copy the face's edge cell, 0,1 <- 2 and
 * n-2,n-1 <- n-3.  fixed: the ghosts keep the values of g->wfix, the state
 * captured when the fixed faces were initialised (bc3_setfixed_dir /
 * bc3_fixed_dir, gradops.cu:1201-1262).  Each direction is filled over the
 * whole transverse range, directions in order, as the driver loop does
 * (iosmaug.cpp:1157-1166).
 * -------------------------------------------------------------------- */
void sm_boundary(const sm_grid *g, double *w)
{
    geo G = mkgeo(g);
    int nve = sm_nve(g);
    for (int f = 0; f < nve; f++) {
        double *q = w + (long)f * G.nc;
        for (int d = 0; d < G.D; d++) {
            long s = G.s[d];
            int n = G.n[d];
            int bt = g->bc[d];
            const double *qf = g->wfix ? g->wfix + (long)f * G.nc : NULL;
            FORALL(G)
            {
                long c = CIDX(G);
                int x = coord(d, i_, j_, k_);
                int src = -1;
                int ghost = x < 2 || x > n - 3;
                if (bt == 5) {
                    if (ghost && qf) q[c] = qf[c];
                    continue;
                }
                if (bt == 0) {
                    if (x == 0 || x == 1) src = n - 4 + x;
                    else if (x == n - 1 || x == n - 2) src = 4 - n + x;
                } else if (bt == 3) {
                    if (x == 0 || x == 1) src = 2;
                    else if (x == n - 1 || x == n - 2) src = n - 3;
                } else {
                    if (x == 0 || x == 1) src = 4 - x;
                    else if (x == n - 1) src = x - 4;
                    else if (x == n - 2) src = x - 2;
                }
                if (src >= 0) q[c] = q[c + (long)(src - x) * s];
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * Fast speed along d plus |v_d| (computec3, dervfields.cu:418-486).
 * -------------------------------------------------------------------- */
static double cfast_at(const geo *G, const sm_grid *g, const double *w, const double *bg,
                       long c, int d)
{
    int D = G->D;
    long nc = G->nc;
    double gam = g->gamma;
    double rho = w[F_RHO * nc + c] + bg[G_RHOB * nc + c];
    double mm = 0, bb = 0, bbb = 0, bgb = 0;
    for (int t = 0; t < D; t++) {
        double m = w[F_MOM(D, t) * nc + c];
        double b = w[F_B(D, t) * nc + c];
        double B = bg[G_BB(t) * nc + c];
        mm += m * m;
        bb += b * b;
        bgb += B * b;
        bbb += B * B;
    }
    double e = w[F_EN(D) * nc + c];
    double pk = (gam - 1) * (e - 0.5 * (mm / rho) - 0.5 * bb - bgb);
    double cs2 = (gam * (pk + (gam - 1) * (bg[G_EB * nc + c] - 0.5 * bbb))) / rho;
    double cf2 = ((bb + bbb + 2.0 * bgb) / rho) + cs2;
    double btd = bg[G_BB(d) * nc + c] + w[F_B(D, d) * nc + c];
    double cfi = 0.5 * (cf2 + sqrt(cf2 * cf2 - 4.0 * cs2 * (btd * btd) / rho));
    return sqrt(cfi) + fabs(w[F_MOM(D, d) * nc + c] / rho);
}

/* cucomputemaxc: max over every cell (ghosts included) of cfast, from 0
 * (computedervfields.cu:1922-2116, newreduction0computemax 1054-1086). */
void sm_cmax(const sm_grid *g, const double *w, const double *bg, double cmax[3])
{
    geo G = mkgeo(g);
    for (int d = 0; d < 3; d++) cmax[d] = 0.0;
    for (int d = 0; d < G.D; d++) {
        double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
        for (long c = 0; c < G.nc; c++) {
            double v = cfast_at(&G, g, w, bg, c, d);
            if (v > m) m = v;
        }
        cmax[d] = m;
    }
}

/* cucomputemaxcourant for one direction: max over cells of cfast_d / dx_d
 * (computedervfields.cu:2118-2296, zeropadmaxcourant_parallel 1296-1340) */
double sm_maxcourant_dim(const sm_grid *g, const double *w, const double *bg, int d)
{
    geo G = mkgeo(g);
    double m = 0.0;
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
    for (int k_ = 0; k_ < G.n[2]; k_++)
        for (int j_ = 0; j_ < G.n[1]; j_++)
            for (int i_ = 0; i_ < G.n[0]; i_++) {
                long c = CIDX(G);
                double v = cfast_at(&G, g, w, bg, c, d) / g->dx[d][coord(d, i_, j_, k_)];
                if (v > m) m = v;
            }
    return m;
}

/* the Courant limit over every direction: max_d of sm_maxcourant_dim */
double sm_maxcourant(const sm_grid *g, const double *w, const double *bg)
{
    double m = 0.0;
    for (int d = 0; d < g->ndim; d++) {
        double v = sm_maxcourant_dim(g, w, bg, d);
        if (v > m) m = v;
    }
    return m;
}

/* ------------------------------------------------------------------------
 * Hyperviscosity (hyperdifvisc1r.cu:426-1300, hyperdifvisc1l.cu:180-1450;
 * SAC setnu, vacusr.viscosity.t:225-457).
 * q is rho', e' - (b'^2 + m^2/rho)/2, m_f/rho or b'_f from the stage state
 * (hyperdifvisc1r_parallel).  With the pads q[-1] = q[5], q[n] = q[n-6]:
 *   D3(f) = |3(q[f+1]-q[f]) - (q[f+2]-q[f-1])|,  D1(f) = |q[f+1]-q[f]|
 *   nu_R(x) = dx*cmax*chyp * max D3(x-1..x+1) / max D1(x-2..x+2)
 *   nu_L(x) = dx*cmax*chyp * max D3(x-2..x)   / max D1(x-3..x+1)
 * (0 where the denominator is 0) on interior cells, 0 elsewhere.
 * -------------------------------------------------------------------- */
static void hyper_q(const geo *G, const double *w, const double *bg, int field, double *q)
{
    int D = G->D;
    long nc = G->nc;
    FORALL(*G)
    {
        long c = CIDX(*G);
        double rho = w[F_RHO * nc + c] + bg[G_RHOB * nc + c];
        double v;
        if (field == F_EN(D)) {
            double bb = 0, mm = 0;
            for (int t = 0; t < D; t++) {
                double b = w[F_B(D, t) * nc + c], m = w[F_MOM(D, t) * nc + c];
                bb += b * b;
                mm += m * m;
            }
            v = w[F_EN(D) * nc + c] - 0.5 * (bb + mm / rho);
        } else if (field >= F_MOM(D, 0) && field <= F_MOM(D, D - 1)) {
            v = w[(long)field * nc + c] / rho;
        } else {
            v = w[(long)field * nc + c];
        }
        q[c] = v;
    }
}

static void nu_from_q(const geo *G, const sm_grid *g, const double *q, int field, int dim,
                      double cmax_d, double *nuR, double *nuL)
{
    long s = G->s[dim];
    int n = G->n[dim];
    double chyp = g->chyp[field];
    FORALL(*G)
    {
        long c = CIDX(*G);
        nuR[c] = 0.0;
        nuL[c] = 0.0;
        if (!interior(G, i_, j_, k_)) continue;
        int x = coord(dim, i_, j_, k_);
        long c0 = c - (long)x * s; /* cell 0 of this line */
        /* q along the line with the two pads, index -1..n */
#define QX(xx) ((xx) < 0 ? q[c0 + 5 * s] : ((xx) >= n ? q[c0 + (long)(n - 6) * s] : q[c0 + (long)(xx) * s]))
#define D3(f) fabs(3.0 * (QX((f) + 1) - QX(f)) - (QX((f) + 2) - QX((f) - 1)))
#define D1(f) fabs(QX((f) + 1) - QX(f))
        for (int side = 0; side < 2; side++) {
            int a = (side == 0) ? x : x - 1; /* R window centre, L window centre */
            double m3 = 0.0, m1 = 0.0;
            for (int t = -1; t <= 1; t++) {
                double v = D3(a + t);
                if (v > m3) m3 = v;
            }
            for (int t = -2; t <= 2; t++) {
                double v = D1(a + t);
                if (v > m1) m1 = v;
            }
            double nu = 0.0;
            if (m1 > 0) nu = g->dx[dim][x] * cmax_d * chyp * m3 / m1;
            if (side == 0) nuR[c] = nu;
            else nuL[c] = nu;
        }
#undef D1
#undef D3
#undef QX
    }
}

void sm_nu(const sm_grid *g, const double *w, const double *bg, int field, int dim,
           double cmax_d, double *nuR, double *nuL)
{
    geo G = mkgeo(g);
    double *q = alloc(G.nc);
    hyper_q(&G, w, bg, field, q);
    nu_from_q(&G, g, q, field, dim, cmax_d, nuR, nuL);
    free(q);
}

/* Largest hyperviscosity coefficient of a state: max of nu_R and nu_L over
 * the cells and over the (field, dim) pairs the hyperdiffusion visits (rho
 * and e along every dim, v_f along every dim, b_f along every dim but f),
 * each with cmax[dim].  With moddton the reference keeps this maximum in
 * p->maxviscoef while it forms the coefficients of a step
 * (hyperdifvisc1r.cu:1303-1330, hyperdifvisc1l.cu:1373-1411, reset at
 * iosmaug.cpp:973). */
double sm_numax(const sm_grid *g, const double *w, const double *bg, const double *cmax_in)
{
    geo G = mkgeo(g);
    int D = G.D;
    double cmax[3];
    if (cmax_in) {
        for (int d = 0; d < 3; d++) cmax[d] = cmax_in[d];
    } else {
        sm_cmax(g, w, bg, cmax);
    }
    double *q = alloc(G.nc), *nuR = alloc(G.nc), *nuL = alloc(G.nc);
    double m = 0.0;
    int fields[8], nf = 0;
    fields[nf++] = F_RHO;
    fields[nf++] = F_EN(D);
    for (int f = 0; f < D; f++) fields[nf++] = F_MOM(D, f);
    for (int f = 0; f < D; f++) fields[nf++] = F_B(D, f);
    for (int a = 0; a < nf; a++) {
        int fld = fields[a];
        hyper_q(&G, w, bg, fld, q);
        for (int dim = 0; dim < D; dim++) {
            if (fld >= F_B(D, 0) && fld - F_B(D, 0) == dim) continue;
            nu_from_q(&G, g, q, fld, dim, cmax[dim], nuR, nuL);
            for (long c = 0; c < G.nc; c++) {
                if (nuR[c] > m) m = nuR[c];
                if (nuL[c] > m) m = nuL[c];
            }
        }
    }
    free(q);
    free(nuR);
    free(nuL);
    return m;
}

/* Hyperviscous time-step limit, SAC getdt_visc (vacusr.viscosity.t:512-546,
 * dtdiffpar = 0.25): 0.25 dx^2 / numax, minimised over the dims and the
 * cells' own dx; 0 when numax is 0 (no limit).  The reference's cugetdtvisc1
 * keeps this formula in a comment (nushk1.cu:688-694); its reduction
 * (nushk1.cu:700-761) carries each direction's result into the next
 * direction's maximum, which SAC does not. */
double sm_dtvisc(const sm_grid *g, double numax)
{
    if (!(numax > 0)) return 0.0;
    geo G = mkgeo(g);
    double best = -1.0;
    for (int d = 0; d < G.D; d++)
        for (int x = 0; x < G.n[d]; x++) {
            double t = 0.25 * g->dx[d][x] * g->dx[d][x] / numax;
            if (best < 0 || t < best) best = t;
        }
    return best;
}

/* ------------------------------------------------------------------------
 * The operator L(w) = dU/dt.
 * -------------------------------------------------------------------- */
void sm_rhs(const sm_grid *g, const double *w, const double *bg, double t,
            const double *cmax_in, double *rhs)
{
    geo G = mkgeo(g);
    int D = G.D;
    long nc = G.nc;
    int nve = sm_nve(g);
    double gam = g->gamma;

    const double *rp = w + F_RHO * nc;
    const double *ep = w + F_EN(D) * nc;
    const double *mp[3], *bp[3], *bbp[3];
    for (int d = 0; d < D; d++) {
        mp[d] = w + F_MOM(D, d) * nc;
        bp[d] = w + F_B(D, d) * nc;
        bbp[d] = bg + G_BB(d) * nc;
    }
    const double *ebp = bg + G_EB * nc, *rbp = bg + G_RHOB * nc;

    memset(rhs, 0, sizeof(double) * (size_t)nve * nc);

    double *rho = alloc(nc), *pt = alloc(nc), *ptb = alloc(nc), *flux = alloc(nc),
           *grd = alloc(nc), *vel[3], *gv = alloc(nc);
    for (int d = 0; d < D; d++) vel[d] = alloc(nc);

    /* derived fields: computevel3, computept3, computepbg3 (dervfields.cu:256-368) */
    FORALL(G)
    {
        long c = CIDX(G);
        double r = rp[c] + rbp[c];
        double mm = 0, bgb = 0, bb = 0, bbb = 0;
        for (int d = 0; d < D; d++) {
            vel[d][c] = mp[d][c] / r;
            mm += mp[d][c] * mp[d][c];
            bgb += bp[d][c] * bbp[d][c];
            bb += bp[d][c] * bp[d][c];
            bbb += bbp[d][c] * bbp[d][c];
        }
        rho[c] = r;
        double ptmp = (gam - 2.0) * (bgb + 0.5 * bb);
        pt[c] = (gam - 1.0) * (ep[c] - 0.5 * mm / r) - ptmp;
        ptb[c] = (gam - 1) * ebp[c] - 0.5 * (gam - 2) * bbb;
    }

    /* advection: centdiff1 (rho, m) and centdiff2 (e, b), each direction */
    for (int dir = 0; dir < D; dir++) {
        for (int f = 0; f < nve; f++) {
            FORALL(G)
            {
                long c = CIDX(G);
                double r = rho[c], md = mp[dir][c];
                double F;
                if (f == F_RHO) {
                    /* computefluxrho (centdiff1.cu:140-154) */
                    F = md * rp[c] / r + (rbp[c] * md) / (rbp[c] + rp[c]);
                } else if (f <= F_MOM(D, D - 1)) {
                    /* transportflux + fluxmom1x (centdiff1.cu:51-120) */
                    int ii = f - 1;
                    double x1 = bp[ii][c] * bbp[dir][c] + bbp[ii][c] * bp[dir][c];
                    double x2 = bp[ii][c] * bp[dir][c];
                    F = md * mp[ii][c] / r + (dir == ii ? pt[c] - (x1) - x2 : -(x1) - x2);
                } else if (f == F_EN(D)) {
                    /* computefluxe: transportflux_cd2 + fluxe1 (centdiff2.cu:203-300) */
                    double bbm = 0, bm = 0;
                    for (int t = 0; t < D; t++) {
                        bbm += bbp[t][c] * mp[t][c];
                        bm += bp[t][c] * mp[t][c];
                    }
                    double ft = md * pt[c];
                    ft -= bp[dir][c] * bbm;
                    ft -= bbp[dir][c] * bm;
                    ft /= r;
                    ft += md * ebp[c] / r;
                    ft -= bp[dir][c] * bm / r;
                    F = md * ep[c] / r + ft;
                } else {
                    /* computefluxb1/2/3 (centdiff2.cu:302-356) */
                    int ii = f - F_B(D, 0);
                    if (ii == dir)
                        F = 0.0;
                    else
                        F = md * bp[ii][c] / r - (bp[dir][c] + bbp[dir][c]) * mp[ii][c] / r +
                            (bbp[ii][c]) * md / r;
                }
                flux[c] = F;
            }
            g4(&G, g, flux, dir, grd);
            double *out = rhs + (long)f * nc;
            FORALL(G)
            {
                if (!interior(&G, i_, j_, k_)) continue;
                long c = CIDX(G);
                out[c] -= grd[c];
            }
            if (f == F_EN(D)) {
                /* addenergyterms_cd2 (centdiff2.cu:80-115) */
                g4(&G, g, vel[dir], dir, grd);
                FORALL(G)
                {
                    if (!interior(&G, i_, j_, k_)) continue;
                    long c = CIDX(G);
                    out[c] -= ptb[c] * grd[c];
                }
                for (int idim = 0; idim < D; idim++) {
                    g4(&G, g, vel[dir], idim, gv);
                    FORALL(G)
                    {
                        if (!interior(&G, i_, j_, k_)) continue;
                        long c = CIDX(G);
                        out[c] += bbp[idim][c] * bbp[dir][c] * gv[c];
                    }
                }
            }
        }
    }

    /* gravity: addgrav_cd2 (centdiff2.cu:118-154) */
    FORALL(G)
    {
        if (!interior(&G, i_, j_, k_)) continue;
        long c = CIDX(G);
        for (int d = 0; d < D; d++) rhs[F_MOM(D, d) * nc + c] += g->g[d] * rp[c];
        for (int d = 0; d < D; d++)
            rhs[F_EN(D) * nc + c] += rp[c] * (g->g[d] * mp[d][c]) / (rp[c] + rbp[c]);
    }

    /* hyperdiffusion (iosmaug.cpp:969-1146) */
    if (g->hyperdif) {
        double cmax[3];
        if (cmax_in) {
            for (int d = 0; d < 3; d++) cmax[d] = cmax_in[d];
        } else {
            sm_cmax(g, w, bg, cmax);
        }
        double *q = alloc(nc), *nuR = alloc(nc), *nuL = alloc(nc), *tA = alloc(nc),
               *tB = alloc(nc), *tC = alloc(nc);
        double *re = rhs + F_EN(D) * nc;

        /* density, hyperdifrhosource1.cu:15-271 */
        for (int dim = 0; dim < D; dim++) {
            hyper_q(&G, w, bg, F_RHO, q);
            nu_from_q(&G, g, q, F_RHO, dim, cmax[dim], nuR, nuL);
            g1r(&G, g, rp, dim, tA);
            g1l(&G, g, rp, dim, tB);
            double *out = rhs + F_RHO * nc;
            FORALL(G)
            {
                long c = CIDX(G);
                out[c] += (nuR[c] * tA[c] - nuL[c] * tB[c]) / g->dx[dim][coord(dim, i_, j_, k_)];
            }
        }
        /* energy, hyperdifesource1.cu:18-545 */
        for (int dim = 0; dim < D; dim++) {
            hyper_q(&G, w, bg, F_EN(D), q);
            nu_from_q(&G, g, q, F_EN(D), dim, cmax[dim], nuR, nuL);
            g1l(&G, g, q, dim, tB);
            g1r(&G, g, q, dim, tA);
            FORALL(G)
            {
                long c = CIDX(G);
                re[c] += (nuR[c] * tA[c] - nuL[c] * tB[c]) / g->dx[dim][coord(dim, i_, j_, k_)];
            }
        }
        /* momentum, hyperdifmomsource1.cu:17-533 and hyperdifmomsourcene1.cu:16-740,
         * with the (ii, ii0) pattern of iosmaug.cpp:1046-1090 */
        for (int dim = 0; dim < D; dim++)
            for (int f = 0; f < D; f++) {
                int fld = F_MOM(D, f);
                hyper_q(&G, w, bg, fld, q); /* v_f */
                nu_from_q(&G, g, q, fld, dim, cmax[dim], nuR, nuL);
                for (int ii1 = 0; ii1 <= 1; ii1++) {
                    int ii = (ii1 == 0) ? dim : f;
                    int ii0 = (ii1 == 0) ? f : dim;
                    if (ii == dim) {
                        /* diagonal: ii0 == f in both calls */
                        g1r(&G, g, q, dim, tA);
                        g1l(&G, g, q, dim, tB);
                        long sd = G.s[dim];
                        double *om = rhs + F_MOM(D, ii0) * nc;
                        const double *m0 = mp[ii0];
                        FORALL(G)
                        {
                            long c = CIDX(G);
                            int x = coord(dim, i_, j_, k_);
                            if (!interior(&G, i_, j_, k_)) continue;
                            double dxl = g->dx[dim][x];
                            double rl = (rp[c] + rbp[c] + rp[c - sd] + rbp[c - sd]) / 2;
                            double rr = (rp[c] + rbp[c] + rp[c + sd] + rbp[c + sd]) / 2;
                            double ml = (m0[c] + m0[c - sd]) / 2;
                            double mr = (m0[c] + m0[c + sd]) / 2;
                            om[c] += (rr * nuR[c] * tA[c] - rl * nuL[c] * tB[c]) / dxl / 2;
                            re[c] += (mr * nuR[c] * tA[c] - ml * nuL[c] * tB[c]) / dxl / 2;
                        }
                    } else {
                        /* off-diagonal: d/d(ii) [ rho (nuR+nuL)/4 d/d(dim) v_f ] into m_dim */
                        g13(&G, g, q, dim, tA);
                        FORALL(G)
                        {
                            long c = CIDX(G);
                            double g6 = tA[c] * (nuR[c] + nuL[c]) / 4.0;
                            tB[c] = rho[c] * g6;
                            tC[c] = g6 * mp[ii0][c];
                        }
                        g13(&G, g, tB, ii, tA);
                        double *om = rhs + F_MOM(D, ii0) * nc;
                        FORALL(G)
                        {
                            if (!interior(&G, i_, j_, k_)) continue;
                            long c = CIDX(G);
                            om[c] += tA[c];
                        }
                        g13(&G, g, tC, ii, tA);
                        FORALL(G)
                        {
                            if (!interior(&G, i_, j_, k_)) continue;
                            long c = CIDX(G);
                            re[c] += tA[c];
                        }
                    }
                }
            }
        /* magnetic field, hyperdifbsource1.cu:21-440 and hyperdifbsourcene1.cu:18-820,
         * with the (jj, ii0, mm, sb) pattern of iosmaug.cpp:1096-1144 */
        for (int dim = 0; dim < D; dim++)
            for (int f = 0; f < D; f++) {
                if (f == dim) continue;
                int fld = F_B(D, f);
                hyper_q(&G, w, bg, fld, q); /* b'_f */
                nu_from_q(&G, g, q, fld, dim, cmax[dim], nuR, nuL);
                /* ii1 = 0: jj=dim, mm=f, sb=-1, ii0=dim -> cuhyperdifbsourcene1 */
                {
                    g13(&G, g, q, dim, tA); /* tmp2 */
                    FORALL(G)
                    {
                        long c = CIDX(G);
                        double t3 = tA[c] * (nuL[c] + nuR[c]) / 2;
                        tB[c] = t3;                  /* tmp3 */
                        tC[c] = t3 * bp[dim][c];     /* tmp5 */
                    }
                    g13(&G, g, tB, f, tA); /* tmp4 */
                    double *ob = rhs + F_B(D, dim) * nc;
                    FORALL(G)
                    {
                        if (!interior(&G, i_, j_, k_)) continue;
                        long c = CIDX(G);
                        ob[c] += -1.0 * tA[c];
                    }
                    g13(&G, g, tC, f, tA); /* tmp6 */
                    FORALL(G)
                    {
                        if (!interior(&G, i_, j_, k_)) continue;
                        long c = CIDX(G);
                        re[c] += -1.0 * tA[c];
                    }
                }
                /* ii1 = 1: ii0=f, mm=dim, sb=+1, jj=f -> cuhyperdifbsource1 */
                {
                    g1l(&G, g, q, dim, tB); /* tmp4 */
                    g1r(&G, g, q, dim, tA); /* tmp5 */
                    long sd = G.s[dim];
                    double *ob = rhs + F_B(D, f) * nc;
                    FORALL(G)
                    {
                        long c = CIDX(G);
                        if (!interior(&G, i_, j_, k_)) continue;
                        int x = coord(dim, i_, j_, k_);
                        double dxl = g->dx[dim][x];
                        double bl = (bp[f][c] + bp[f][c - sd]) / 2; /* tmp2 */
                        double br = (bp[f][c] + bp[f][c + sd]) / 2; /* tmp3 */
                        ob[c] += 1.0 * (tA[c] * nuR[c] - tB[c] * nuL[c]) / dxl;
                        re[c] += 1.0 * (br * tA[c] * nuR[c] - bl * tB[c] * nuL[c]) / dxl;
                    }
                }
            }
        free(q);
        free(nuR);
        free(nuL);
        free(tA);
        free(tB);
        free(tC);
    }

    /* user source: vertical velocity driver (usersource.cu:4-92 pattern;
     * profile per the stratified-atmosphere workload) */
    if (g->driver == 1) {
        const double PI2 = 6.283185307179586;
        double A = g->drv[0], P = g->drv[1], sh = g->drv[2], z0 = g->drv[3], sz = g->drv[4],
               xc = g->drv[5], yc = g->drv[6];
        double tdep = sin(t * PI2 / P);
        FORALL(G)
        {
            if (!interior(&G, i_, j_, k_)) continue;
            long c = CIDX(G);
            double zp = g->pos[0][i_] - z0;
            double xp = g->pos[1][j_] - xc;
            double r2 = xp * xp;
            if (D == 3) {
                double yp = g->pos[2][k_] - yc;
                r2 += yp * yp;
            }
            double v = A * tdep * exp(-zp * zp / (sz * sz)) * exp(-r2 / (sh * sh));
            double r = rp[c] + rbp[c];
            rhs[F_MOM(D, 0) * nc + c] += v * r;
            rhs[F_EN(D) * nc + c] += (v * v) * r / 2.0;
        }
    }
    /* user source: the torsional driver of the reference
     * (dev/smaug/src/usersource.cu:4-86): horizontal velocities
     * vx = (A yp/L) G tdep along dim 1 and vy = -(A xp/L) G tdep along dim 2,
     * G = exp(-xp^2/sh^2) exp(-yp^2/sh^2) exp(-zp^2/sz^2), with xp, yp, zp the
     * offsets from (xc, yc, z0) along dims 1, 2, 0; m_1 += dt vx rho,
     * m_2 += dt vy rho, e += dt (vx^2 + vy^2) rho / 2 */
    if (g->driver == 2 && D == 3) {
        const double PI2 = 6.283185307179586;
        double A = g->drv[0], P = g->drv[1], sh = g->drv[2], z0 = g->drv[3], sz = g->drv[4],
               xc = g->drv[5], yc = g->drv[6], L = g->drv[7];
        double tdep = sin(t * PI2 / P);
        FORALL(G)
        {
            if (!interior(&G, i_, j_, k_)) continue;
            long c = CIDX(G);
            double xp = g->pos[1][j_] - xc;
            double zp = g->pos[0][i_] - z0;
            double yp = g->pos[2][k_] - yc;
            double ez = exp(-zp * zp / (sz * sz));
            double ex = exp(-xp * xp / (sh * sh));
            double ey = exp(-yp * yp / (sh * sh));
            double exyz = ex * ey * ez;
            double vx = (A * yp / L) * exyz * tdep;
            double vy = -(A * xp / L) * exyz * tdep;
            double r = rp[c] + rbp[c];
            rhs[F_MOM(D, 1) * nc + c] += vx * r;
            rhs[F_MOM(D, 2) * nc + c] += vy * r;
            rhs[F_EN(D) * nc + c] += (vx * vx + vy * vy) * r / 2.0;
        }
    }

    free(rho);
    free(pt);
    free(ptb);
    free(flux);
    free(grd);
    free(gv);
    for (int d = 0; d < D; d++) free(vel[d]);
}

/* ------------------------------------------------------------------------
 * Integrators.
 * -------------------------------------------------------------------- */
static void axpy_all(long n, double *y, const double *x, double a, const double *b)
{
    /* y = x + a*b */
#pragma omp parallel for schedule(static)
    for (long c = 0; c < n; c++) y[c] = x[c] + a * b[c];
}

void sm_step_euler(const sm_grid *g, double *w, const double *bg, double t, double dt,
                   double *scratch)
{
    long n = (long)sm_nve(g) * sm_ncell(g);
    sm_rhs(g, w, bg, t, NULL, scratch);
    axpy_all(n, w, w, dt, scratch);
    sm_boundary(g, w);
}

void sm_step_rk4(const sm_grid *g, double *w, const double *bg, double t, double dt,
                 double *scratch)
{
    long n = (long)sm_nve(g) * sm_ncell(g);
    double *k = scratch, *w1 = scratch + n, *w2 = scratch + 2 * n, *w3 = scratch + 3 * n;
    /* w1 = w + dt/2 k1 */
    sm_rhs(g, w, bg, t, NULL, k);
    axpy_all(n, w1, w, dt / 2, k);
    sm_boundary(g, w1);
    /* w2 = w + dt/2 k2(w1) */
    sm_rhs(g, w1, bg, t + dt / 2, NULL, k);
    axpy_all(n, w2, w, dt / 2, k);
    sm_boundary(g, w2);
    /* w3 = w + dt k3(w2) */
    sm_rhs(g, w2, bg, t + dt / 2, NULL, k);
    axpy_all(n, w3, w, dt, k);
    sm_boundary(g, w3);
    /* w1 = (w1 + 2 w2 + w3 - 4 w)/3 */
#pragma omp parallel for schedule(static)
    for (long c = 0; c < n; c++) w1[c] = (w1[c] + 2 * w2[c] + w3[c] - 4 * w[c]) / 3;
    /* w = w + dt/6 k4(w3); w = w + w1 */
    sm_rhs(g, w3, bg, t + dt, NULL, k);
#pragma omp parallel for schedule(static)
    for (long c = 0; c < n; c++) {
        double v = w[c] + dt / 6 * k[c];
        w[c] = v + w1[c];
    }
    sm_boundary(g, w);
}

double sm_run(const sm_grid *g, double *w, const double *bg, double t0, double dt, int nsteps,
              int rk4)
{
    long n = (long)sm_nve(g) * sm_ncell(g);
    double *scratch = alloc((rk4 ? 5 : 1) * n);
    sm_boundary(g, w);
    double t = t0;
    for (int s = 0; s < nsteps; s++) {
        if (rk4)
            sm_step_rk4(g, w, bg, t, dt, scratch);
        else
            sm_step_euler(g, w, bg, t, dt, scratch);
        t += dt;
    }
    free(scratch);
    return t;
}
