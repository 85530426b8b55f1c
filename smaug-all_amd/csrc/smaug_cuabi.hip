// smaug_cuabi.hip -- the reference's cu* entry points (include/smaug_hip.h)
// implemented with gfx950 kernels.  Built twice: -DUSE_SAC (2-D, NVAR 10)
// into libsmaug_hip2d.so and -DUSE_SAC_3D (3-D, NVAR 13) into
// libsmaug_hip3d.so, because struct params changes layout with NDIM as in
// the reference.
//
// Each entry point performs its own term of the step on the FP64 slots of
// d_wmod (per-entry-point semantics, used by the drop-in driver loop and
// the per-term parity tests).  cuadvancerk4 runs whole RK4 steps on the
// fused engine instead.
#include "smaug_phys.hpp"
#include "../../include/smaug_engine.h"
#include "../../include/smaug_hip.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>

using namespace smaug;

#define D NDIM
typedef Lay<D> LY;

static thread_local char c_err[512] = "";

static int cfail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c_err, sizeof(c_err), fmt, ap);
    va_end(ap);
    return code;
}

#define CHK(x)                                                                             \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return cfail(SMAUG_EHIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                        \
    } while (0)

extern "C" const char *cuerrorstring(void) { return c_err; }

// ---------------------------------------------------------------------------
// library state created by cuinit (the reference keeps the same state in
// globals and in the caller's pointer arguments, init.cu:816-1068)
// ---------------------------------------------------------------------------
namespace {
struct Ctx {
    bool ready = false;
    int n[3] = {1, 1, 1};
    idx_t dimp = 0;
    int nslots = 0;
    double *wmod = nullptr, *wd = nullptr, *dwn1 = nullptr, *wtemp = nullptr,
           *wtemp1 = nullptr, *wtemp2 = nullptr, *wnew = nullptr;
    struct params *dp = nullptr;
    struct state *ds = nullptr;
    struct bparams *dbp = nullptr;
    std::vector<double> dx[3], pos[3];
    double *d_dx[3] = {nullptr, nullptr, nullptr};
    double *d_idx[3] = {nullptr, nullptr, nullptr};
    double *d_prof[3] = {nullptr, nullptr, nullptr};
    double *d_profq[3] = {nullptr, nullptr, nullptr};
    unsigned long long *d_red = nullptr;
    smaug_engine *eng[2] = {nullptr, nullptr};
    // fixed faces (boundtype 5): the values captured by the cuboundary call
    // with p->it == -1, NVAR fields of dimp cells; fixed_set per field
    double *fixed = nullptr;
    bool fixed_set[NVAR] = {};
    int driver = 0;
    double drv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
Ctx g_ctx;
} // namespace

static Geo make_geo(const struct params *p)
{
    Geo g;
    for (int d = 0; d < 3; d++) {
        g.n[d] = g_ctx.n[d];
        g.per[d] = (d < D) ? (p->boundtype[0][d][0] == 0) : 1;
        g.dx[d] = g_ctx.d_dx[d];
        g.idx[d] = g_ctx.d_idx[d];
        g.prof[d] = g_ctx.d_prof[d];
        g.profq[d] = g_ctx.d_profq[d];
    }
    g.s1 = g_ctx.n[0];
    g.s2 = (idx_t)g_ctx.n[0] * g_ctx.n[1];
    g.nc = g_ctx.dimp;
    g.bgf = g.nc;
    g.bgm = ~0u;
    return g;
}

static Phys make_phys(const struct params *p)
{
    Phys ph;
    ph.gamma = p->gamma;
    for (int d = 0; d < 3; d++) ph.g[d] = d < D ? p->g[d] : 0.0;
    for (int f = 0; f < 13; f++) ph.chyp[f] = f < NVAR ? p->chyp[f] : 0.0;
    for (int k = 0; k < 8; k++) ph.drv[k] = g_ctx.drv[k];
    ph.hyperdif = p->hyperdifmom;
    ph.driver = g_ctx.driver;
    return ph;
}

static int check_params(const struct params *p)
{
    if (!g_ctx.ready) return cfail(SMAUG_EINVAL, "cuinit has not been called");
    for (int d = 0; d < D; d++) {
        int bt = p->boundtype[0][d][0];
        if (bt != 0 && bt != 3 && bt != 4 && bt != 5)
            return cfail(SMAUG_EINVAL,
                         "boundtype %d along dim %d unsupported (0 periodic, 3 cont, 4 contcd4, "
                         "5 fixed)", bt, d);
        for (int f = 0; f < NVAR; f++)
            for (int b = 0; b < 2; b++)
                if (p->boundtype[f][d][b] != bt)
                    return cfail(SMAUG_EINVAL,
                                 "boundtype must be the same for every field and side along a dim");
    }
    return 0;
}

static inline double *slot(int s) { return g_ctx.wmod + (idx_t)s * NVAR * g_ctx.dimp; }

static dim3 interior_grid(dim3 blk)
{
    int i0 = g_ctx.n[0] - 4, i1 = g_ctx.n[1] - 4;
    return dim3((i0 + blk.x - 1) / blk.x, (i1 + blk.y - 1) / blk.y, D == 3 ? g_ctx.n[2] - 4 : 1);
}

static dim3 all_grid(dim3 blk)
{
    return dim3((g_ctx.n[0] + blk.x - 1) / blk.x, (g_ctx.n[1] + blk.y - 1) / blk.y,
                D == 3 ? g_ctx.n[2] : 1);
}

#define CELL_ALL()                                                                         \
    const int x0 = blockIdx.x * blockDim.x + threadIdx.x;                                  \
    const int x1 = blockIdx.y * blockDim.y + threadIdx.y;                                  \
    const int x2 = blockIdx.z;                                                             \
    if (x0 >= g.n[0] || x1 >= g.n[1]) return;                                              \
    const int xs[3] = {x0, x1, x2};                                                        \
    const idx_t c = (idx_t)x0 + g.s1 * x1 + g.s2 * x2;

#define CELL_INT()                                                                         \
    const int x0 = 2 + blockIdx.x * blockDim.x + threadIdx.x;                              \
    const int x1 = 2 + blockIdx.y * blockDim.y + threadIdx.y;                              \
    const int x2 = (D == 3) ? 2 + (int)blockIdx.z : 0;                                     \
    if (x0 > g.n[0] - 3 || x1 > g.n[1] - 3) return;                                        \
    const int xs[3] = {x0, x1, x2};                                                        \
    const idx_t c = (idx_t)x0 + g.s1 * x1 + g.s2 * x2;

__device__ __forceinline__ bool is_interior(const Geo &g, const int *xs)
{
    for (int d = 0; d < D; d++)
        if (xs[d] < 2 || xs[d] > g.n[d] - 3) return false;
    return true;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void kc_copy(double *__restrict__ dst, const double *__restrict__ src, idx_t n)
{
    for (idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (idx_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ghost fill of one field along one direction over the whole transverse
// range (boundary0/1/2_contcd4_parallel, boundary.cu:16-88; periodic as in
// gradops.cu:1271-1317)
// ghost cells of one field along dir by boundtype bt: 0 periodic, 3 cont
// (both layers copy the face's edge cell, SAC vacgrid.t "cont"), 4 contcd4,
// 5 fixed (the values wf captured when the faces were initialised)
__global__ void kc_bc(Geo g, double *__restrict__ w, int dir, int bt,
                      const double *__restrict__ wf)
{
    CELL_ALL();
    const int x = xs[dir], n = g.n[dir];
    int src;
    if (x >= 2 && x <= n - 3) return;
    if (bt == 5) {
        w[c] = wf[c];
        return;
    }
    if (bt == 0)
        src = x < 2 ? n - 4 + x : 4 - n + x;
    else if (bt == 3)
        src = x < 2 ? 2 : n - 3;
    else
        src = x < 2 ? 4 - x : (x == n - 1 ? x - 4 : x - 2);
    w[c] = w[c + (idx_t)(src - x) * stride<D>(g, dir)];
}

// derived fields into wd (computevel3, computept3, computepbg3, computepk3,
// computec3; dervfields.cu:196-486)
__global__ void kc_derived(Geo g, Phys ph, const double *__restrict__ w, double *__restrict__ wd,
                           int kind, int dir)
{
    CELL_ALL();
    (void)xs;
    const idx_t nc = g.nc;
    St<D, double> s;
    load_st<D, double>(s, w, w + (idx_t)LY::NVE * nc, nc, c);
    const double rho = rho_full(s);
    if (kind == 0) {
        for (int d = 0; d < D; d++) wd[(idx_t)(vel1 + d) * nc + c] = s.m[d] / rho;
    } else if (kind == 1) {
        wd[(idx_t)pressuret * nc + c] = ptot_pert<D, double>(s, rho, ph.gamma);
    } else if (kind == 2) {
        wd[(idx_t)ptb * nc + c] = ptot_bg<D, double>(s, ph.gamma);
    } else if (kind == 3) {
        double mm = 0, bb = 0, bgb = 0;
        for (int d = 0; d < D; d++) {
            mm += s.m[d] * s.m[d];
            bb += s.b[d] * s.b[d];
            bgb += s.b[d] * s.bb[d];
        }
        wd[(idx_t)pressurek * nc + c] = (ph.gamma - 1) * s.e - (0.5 * mm / rho + (0.5 * bb + bgb));
    } else {
        wd[(idx_t)::cfast * nc + c] = smaug::cfast<D, double>(s, dir, ph.gamma);
    }
}

// max over every cell of cfast_dir (mode 0) or cfast_dir/dx_dir (mode 1)
__global__ __launch_bounds__(256) void kc_maxc(Geo g, Phys ph, const double *__restrict__ w,
                                               int dir, int mode, unsigned long long *red)
{
    __shared__ double part[4];
    const idx_t nc = g.nc;
    double mx = 0.0;
    for (idx_t c = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (idx_t)gridDim.x * blockDim.x) {
        St<D, double> s;
        load_st<D, double>(s, w, w + (idx_t)LY::NVE * nc, nc, c);
        double v = smaug::cfast<D, double>(s, dir, ph.gamma);
        if (mode == 1) {
            idx_t r = c;
            int x0 = (int)(r % g.n[0]);
            r /= g.n[0];
            int x1 = (int)(r % g.n[1]);
            int x2 = (int)(r / g.n[1]);
            int x = dir == 0 ? x0 : (dir == 1 ? x1 : x2);
            v = v / g.dx[dir][x];
        }
        mx = v > mx ? v : mx;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = part[0];
        for (int k = 1; k < 4; k++) v = part[k] > v ? part[k] : v;
        atomicMax(red, (unsigned long long)__double_as_longlong(v));
    }
}

__global__ void kc_zero_red(unsigned long long *red)
{
    if (threadIdx.x == 0) red[0] = 0ull;
}

// flux of one evolved field along a (centdiff1.cu:19-256, centdiff2.cu:160-356)
__device__ __forceinline__ double flux_one(const St<D, double> &s, int f, int a, double gam)
{
    const double rho = rho_full(s);
    const double ma = s.m[a];
    if (f == LY::RHO) return ma * s.r / rho + (s.rb * ma) / (s.rb + s.r);
    if (f >= LY::M0 && f < LY::M0 + D) {
        int i = f - LY::M0;
        double x1 = s.b[i] * s.bb[a] + s.bb[i] * s.b[a];
        double x2 = s.b[i] * s.b[a];
        double pt = ptot_pert<D, double>(s, rho, gam);
        return ma * s.m[i] / rho + (a == i ? pt - (x1) - x2 : -(x1) - x2);
    }
    if (f == LY::EN) {
        double bbm = 0, bm = 0;
        for (int t = 0; t < D; t++) {
            bbm += s.bb[t] * s.m[t];
            bm += s.b[t] * s.m[t];
        }
        double pt = ptot_pert<D, double>(s, rho, gam);
        double ft = ma * pt;
        ft -= s.b[a] * bbm;
        ft -= s.bb[a] * bm;
        ft /= rho;
        ft += ma * s.eb / rho;
        ft -= s.b[a] * bm / rho;
        return ma * s.e / rho + ft;
    }
    int i = f - LY::B0;
    if (i == a) return 0.0;
    return ma * s.b[i] / rho - (s.b[a] + s.bb[a]) * s.m[i] / rho + (s.bb[i]) * ma / rho;
}

__device__ __forceinline__ double g4(double qm2, double qm1, double q1, double q2, double dx)
{
    return ((8 * q1 - 8 * qm1 + qm2 - q2) / 6.0) / (2.0 * dx);
}

// w_out[f] -= dt * d/dx_dir F_f,dir(w_in); for the energy also the background
// work terms with p->dt (addenergyterms_cd2, centdiff2.cu:80-115)
__global__ void kc_centdiff(Geo g, Phys ph, const double *__restrict__ win,
                            double *__restrict__ wout, int f, int dir, double dt, double pdt)
{
    CELL_INT();
    const idx_t nc = g.nc;
    const double *bgp = win + (idx_t)LY::NVE * nc;
    const idx_t sd = stride<D>(g, dir);
    const double dxd = g.dx[dir][xs[dir]];
    double F[4];
    for (int p = 0; p < 4; p++) {
        const int off = p < 2 ? p - 2 : p - 1;
        St<D, double> s;
        load_st<D, double>(s, win, bgp, nc, c + off * sd);
        F[p] = flux_one(s, f, dir, ph.gamma);
    }
    double gdiv = zero4(g.per[dir], xs[dir], g.n[dir]) ? 0.0 : g4(F[0], F[1], F[2], F[3], dxd);
    wout[(idx_t)f * nc + c] -= dt * gdiv;
    if (f == LY::EN) {
        St<D, double> sc;
        load_st<D, double>(sc, win, bgp, nc, c);
        const double ptbv = ptot_bg<D, double>(sc, ph.gamma);
        double acc = 0;
        for (int idim = 0; idim < D; idim++) {
            const idx_t si = stride<D>(g, idim);
            double v[4];
            for (int p = 0; p < 4; p++) {
                const int off = p < 2 ? p - 2 : p - 1;
                const idx_t cn = c + off * si;
                v[p] = win[(idx_t)(LY::M0 + dir) * nc + cn] / (win[cn] + bgp[nc + cn]);
            }
            double gv = zero4(g.per[idim], xs[idim], g.n[idim])
                            ? 0.0
                            : g4(v[0], v[1], v[2], v[3], g.dx[idim][xs[idim]]);
            if (idim == dir) acc -= pdt * ptbv * gv;
            acc += pdt * sc.bb[idim] * sc.bb[dir] * gv;
        }
        wout[(idx_t)LY::EN * nc + c] += acc;
    }
}

// gravity (addgrav_cd2, centdiff2.cu:118-154) with p->dt
__global__ void kc_grav(Geo g, Phys ph, const double *__restrict__ win,
                        double *__restrict__ wout, double pdt)
{
    CELL_INT();
    (void)xs;
    const idx_t nc = g.nc;
    const double r = win[c], rb = win[(idx_t)(LY::NVE + 1) * nc + c];
    for (int d = 0; d < D; d++) wout[(idx_t)(LY::M0 + d) * nc + c] += pdt * ph.g[d] * r;
    for (int d = 0; d < D; d++)
        wout[(idx_t)LY::EN * nc + c] += pdt * r * (ph.g[d] * win[(idx_t)(LY::M0 + d) * nc + c]) / (r + rb);
}

// user driver at time qt (source_parallel, centdiff2.cu:785-867): the
// vertical Gaussian driver (1) or the torsional one of usersource.cu:4-86
// (2: vx = (A yp/L) G tdep along dim 1, vy = -(A xp/L) G tdep along dim 2)
__global__ void kc_source(Geo g, Phys ph, const double *__restrict__ win,
                          double *__restrict__ wout, double pdt, double qt)
{
    CELL_INT();
    const idx_t nc = g.nc;
    (void)xs;
    const double rho = win[c] + win[(idx_t)(LY::NVE + 1) * nc + c];
    const double tdep = sin(qt * 6.283185307179586 / ph.drv[1]);
    if (ph.driver == 2) {
        if (D == 3) {
            const double a = ph.drv[0] * tdep / ph.drv[7] * g.prof[0][x0];
            const double vx = a * g.prof[1][x1] * g.profq[2][x2];
            const double vy = -(a * g.profq[1][x1] * g.prof[2][x2]);
            wout[(idx_t)(LY::M0 + 1) * nc + c] += pdt * vx * rho;
            wout[(idx_t)(LY::M0 + D - 1) * nc + c] += pdt * vy * rho;
            wout[(idx_t)LY::EN * nc + c] += pdt * (vx * vx + vy * vy) * rho / 2.0;
        }
        return;
    }
    double pr = ph.drv[0] * tdep * g.prof[0][x0] * g.prof[1][x1];
    if (D == 3) pr *= g.prof[2][x2];
    wout[(idx_t)LY::M0 * nc + c] += pdt * pr * rho;
    wout[(idx_t)LY::EN * nc + c] += pdt * (pr * pr) * rho / 2.0;
}

// hyperdiffusion variable q of a field at cell c
__device__ __forceinline__ double hq(const double *__restrict__ w, idx_t nc, idx_t c, int field)
{
    const double *bgp = w + (idx_t)LY::NVE * nc;
    if (field == LY::EN) {
        St<D, double> s;
        load_st<D, double>(s, w, bgp, nc, c);
        return e_int<D, double>(s, rho_full(s));
    }
    if (field >= LY::M0 && field < LY::M0 + D) return w[(idx_t)field * nc + c] / (w[c] + bgp[nc + c]);
    return w[(idx_t)field * nc + c];
}

// nu_R (side 0) or nu_L (side 1) of (field, dim) into wd[hdnur/hdnul]
// (hyperdifvisc1r/1l kernel chains), 0 off the interior
__global__ void kc_nu(Geo g, Phys ph, const double *__restrict__ w, double *__restrict__ wd,
                      int field, int dim, int side, double cmax)
{
    CELL_ALL();
    const idx_t nc = g.nc;
    double *out = wd + (idx_t)(side == 0 ? hdnur : hdnul) * nc;
    if (!is_interior(g, xs)) {
        out[c] = 0.0;
        return;
    }
    const int x = xs[dim], n = g.n[dim];
    const idx_t sd = stride<D>(g, dim);
    const idx_t c0 = c - (idx_t)x * sd;
    const int a = side == 0 ? x : x - 1;
    double q[6];
    for (int p = 0; p < 6; p++) {
        int xx = a - 2 + p;
        if (xx < 0) xx = 5;
        if (xx >= n) xx = n - 6;
        q[p] = hq(w, nc, c0 + (idx_t)xx * sd, field);
    }
    double m1 = 0, m3 = 0;
    for (int f = 0; f < 5; f++) {
        double v = fabs(q[f + 1] - q[f]);
        m1 = v > m1 ? v : m1;
    }
    for (int f = 1; f < 4; f++) {
        double v = fabs(3.0 * (q[f + 1] - q[f]) - (q[f + 2] - q[f - 1]));
        m3 = v > m3 ? v : m3;
    }
    out[c] = m1 > 0 ? g.dx[dim][x] * cmax * ph.chyp[field] * m3 / m1 : 0.0;
}

// diagonal hyperdiffusive source (hyperdifrhosource1, hyperdifesource1,
// hyperdifmomsource1, hyperdifbsource1): kind 0 rho, 1 energy, 2 momentum,
// 3 magnetic field.  qf: field whose q is diffused, tgt: updated field,
// af: field averaged for the energy term (m_ii0 or b_jj), coef: sb or 1/2
__global__ void kc_hd_diag(Geo g, const double *__restrict__ win, double *__restrict__ wout,
                           const double *__restrict__ wd, int kind, int qf, int tgt, int af,
                           int dim, double coef, double dt)
{
    CELL_INT();
    const idx_t nc = g.nc;
    const idx_t sd = stride<D>(g, dim);
    const int x = xs[dim];
    const double dxd = g.dx[dim][x];
    const bool z1 = zero1(g.per[dim], x, g.n[dim]);
    const double qc = hq(win, nc, c, qf), ql = hq(win, nc, c - sd, qf), qr = hq(win, nc, c + sd, qf);
    const double gR = z1 ? 0.0 : (qr - qc) / dxd;
    const double gL = z1 ? 0.0 : (qc - ql) / dxd;
    const double nR = wd[(idx_t)hdnur * nc + c], nL = wd[(idx_t)hdnul * nc + c];
    if (kind <= 1) {
        wout[(idx_t)tgt * nc + c] += dt * (nR * gR - nL * gL) / dxd;
        return;
    }
    const double *bgp = win + (idx_t)LY::NVE * nc;
    const double ac = win[(idx_t)af * nc + c];
    const double aL = (ac + win[(idx_t)af * nc + c - sd]) / 2;
    const double aR = (ac + win[(idx_t)af * nc + c + sd]) / 2;
    if (kind == 2) {
        const double rL = (win[c] + bgp[nc + c] + win[c - sd] + bgp[nc + c - sd]) / 2;
        const double rR = (win[c] + bgp[nc + c] + win[c + sd] + bgp[nc + c + sd]) / 2;
        wout[(idx_t)tgt * nc + c] += dt * ((rR * nR * gR - rL * nL * gL) / dxd * coef);
        wout[(idx_t)LY::EN * nc + c] += dt * ((aR * nR * gR - aL * nL * gL) / dxd * coef);
    } else {
        wout[(idx_t)tgt * nc + c] += dt * (coef * (gR * nR - gL * nL) / dxd);
        wout[(idx_t)LY::EN * nc + c] += dt * (coef * (aR * gR * nR - aL * gL * nL) / dxd);
    }
}

// off-diagonal hyperdiffusive source, inner product at every cell:
// G = d/dx_dim q * (nu_R + nu_L) * fac, tA = W G (W = rho or 1), tB = X G
__global__ void kc_off_inner(Geo g, const double *__restrict__ win, const double *__restrict__ wd,
                             double *__restrict__ tA, double *__restrict__ tB, int qf, int dim,
                             int wrho, int xf, double fac)
{
    CELL_ALL();
    const idx_t nc = g.nc;
    const int x = xs[dim];
    double gq = 0.0;
    bool inside = true;
    for (int d = 0; d < D; d++)
        if (xs[d] < 1 || xs[d] > g.n[d] - 2) inside = false;
    if (inside && !zero1(g.per[dim], x, g.n[dim])) {
        const idx_t sd = stride<D>(g, dim);
        gq = (hq(win, nc, c + sd, qf) - hq(win, nc, c - sd, qf)) / g.dx[dim][x] / 2.0;
    }
    const double G = gq * (wd[(idx_t)hdnur * nc + c] + wd[(idx_t)hdnul * nc + c]) * fac;
    const double W = wrho ? win[c] + win[(idx_t)(LY::NVE + 1) * nc + c] : 1.0;
    tA[c] = W * G;
    tB[c] = G * win[(idx_t)xf * nc + c];
}

// outer derivative: wout[tgt] += sg dt d/dx_od tA, wout[e] += sg dt d/dx_od tB
__global__ void kc_off_outer(Geo g, const double *__restrict__ tA, const double *__restrict__ tB,
                             double *__restrict__ wout, int od, int tgt, double sg, double dt)
{
    CELL_INT();
    const idx_t nc = g.nc;
    if (zero1(g.per[od], xs[od], g.n[od])) return;
    const idx_t so = stride<D>(g, od);
    const double dxo = g.dx[od][xs[od]];
    wout[(idx_t)tgt * nc + c] += sg * dt * ((tA[c + so] - tA[c - so]) / dxo / 2.0);
    wout[(idx_t)LY::EN * nc + c] += sg * dt * ((tB[c + so] - tB[c - so]) / dxo / 2.0);
}

// RK4 slot algebra of SAC's fourstep on the evolved fields (advance.cu:18-133)
__global__ void kc_advance(double *__restrict__ wm, idx_t dimp, int order)
{
    const idx_t n = (idx_t)LY::NVE * dimp;
    double *s0 = wm, *s1 = wm + (idx_t)NVAR * dimp, *s2 = wm + 2 * (idx_t)NVAR * dimp,
           *s3 = wm + 3 * (idx_t)NVAR * dimp;
    for (idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (idx_t)gridDim.x * blockDim.x) {
        if (order == 0)
            s2[i] = s0[i];
        else if (order == 1)
            s3[i] = s0[i];
        else if (order == 2)
            s1[i] = (s1[i] + 2.0 * s2[i] + s3[i] - 4.0 * s0[i]) / 3;
        else
            s0[i] = s0[i] + s1[i];
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static int sync_check()
{
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    return 0;
}

static int reduce_max(const Geo &g, const Phys &ph, const double *w, int dir, int mode, double *out)
{
    hipLaunchKernelGGL(kc_zero_red, dim3(1), dim3(64), 0, 0, g_ctx.d_red);
    idx_t nb = (g.nc + 255) / 256;
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(kc_maxc, dim3((unsigned)nb), dim3(256), 0, 0, g, ph, w, dir, mode, g_ctx.d_red);
    unsigned long long b;
    CHK(hipMemcpy(&b, g_ctx.d_red, sizeof(b), hipMemcpyDeviceToHost));
    memcpy(out, &b, sizeof(double));
    return 0;
}

static int upload_axes(const struct params *p)
{
    for (int d = 0; d < 3; d++) {
        int nd = g_ctx.n[d];
        std::vector<double> dx(nd, 1.0), idx(nd, 1.0), prof(nd, 1.0), profq(nd, 0.0);
        if (d < D) {
            for (int i = 0; i < nd; i++) {
                dx[i] = g_ctx.dx[d][i];
                idx[i] = 1.0 / dx[i];
                if (g_ctx.driver) {
                    double pp = g_ctx.pos[d][i];
                    double c0 = d == 0 ? g_ctx.drv[3] : (d == 1 ? g_ctx.drv[5] : g_ctx.drv[6]);
                    double s = d == 0 ? g_ctx.drv[4] : g_ctx.drv[2];
                    prof[i] = exp(-(pp - c0) * (pp - c0) / (s * s));
                    if (d > 0) profq[i] = (pp - c0) * prof[i];
                }
            }
        }
        if (!g_ctx.d_dx[d]) {
            CHK(hipMalloc((void **)&g_ctx.d_dx[d], sizeof(double) * nd));
            CHK(hipMalloc((void **)&g_ctx.d_idx[d], sizeof(double) * nd));
            CHK(hipMalloc((void **)&g_ctx.d_prof[d], sizeof(double) * nd));
            CHK(hipMalloc((void **)&g_ctx.d_profq[d], sizeof(double) * nd));
        }
        CHK(hipMemcpy(g_ctx.d_dx[d], dx.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
        CHK(hipMemcpy(g_ctx.d_idx[d], idx.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
        CHK(hipMemcpy(g_ctx.d_prof[d], prof.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
        CHK(hipMemcpy(g_ctx.d_profq[d], profq.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
    }
    (void)p;
    return 0;
}

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------
extern "C" int cuinit(struct params **p, struct bparams **bp, real **wmod, real **wnew,
                      real **wd, struct state **state, struct params **d_p,
                      struct bparams **d_bp, real **d_wnew, real **d_wmod, real **d_dwn1,
                      real **d_wd, struct state **d_state, real **d_wtemp, real **d_wtemp1,
                      real **d_wtemp2)
{
    (void)bp;
    (void)wnew;
    if (!p || !*p || !wmod || !*wmod || !wd || !*wd) return cfail(SMAUG_EINVAL, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return cfail(SMAUG_ENODEV, "no HIP device");
    const struct params *pp = *p;
    for (int d = 0; d < 3; d++) g_ctx.n[d] = d < D ? pp->n[d] : 1;
    for (int d = 0; d < D; d++)
        if (g_ctx.n[d] < 8) return cfail(SMAUG_EINVAL, "n[%d] = %d: need >= 8 incl. ghosts", d, g_ctx.n[d]);
    g_ctx.dimp = (idx_t)g_ctx.n[0] * g_ctx.n[1] * g_ctx.n[2];
    g_ctx.nslots = pp->rkon == 1 ? 6 : 3;
    idx_t dimp = g_ctx.dimp;
    idx_t n1p = (idx_t)(g_ctx.n[0] + 1) * (g_ctx.n[1] + 1) * (D == 3 ? g_ctx.n[2] + 1 : 1);
    idx_t n2p = (idx_t)(g_ctx.n[0] + 2) * (g_ctx.n[1] + 2) * (D == 3 ? g_ctx.n[2] + 2 : 1);
    CHK(hipMalloc((void **)&g_ctx.wmod, sizeof(double) * g_ctx.nslots * NVAR * dimp));
    CHK(hipMalloc((void **)&g_ctx.wnew, sizeof(double) * NVAR * dimp));
    CHK(hipMalloc((void **)&g_ctx.dwn1, sizeof(double) * NVAR * dimp));
    CHK(hipMalloc((void **)&g_ctx.wd, sizeof(double) * NDERV * dimp));
    CHK(hipMalloc((void **)&g_ctx.wtemp, sizeof(double) * NTEMP * dimp));
    CHK(hipMalloc((void **)&g_ctx.wtemp1, sizeof(double) * NTEMP1 * n1p));
    CHK(hipMalloc((void **)&g_ctx.wtemp2, sizeof(double) * NTEMP2 * n2p));
    CHK(hipMalloc((void **)&g_ctx.dp, sizeof(struct params)));
    CHK(hipMalloc((void **)&g_ctx.ds, sizeof(struct state)));
    CHK(hipMalloc((void **)&g_ctx.dbp, sizeof(struct bparams)));
    CHK(hipMalloc((void **)&g_ctx.d_red, sizeof(unsigned long long) * 2));
    CHK(hipMemcpy(g_ctx.wmod, *wmod, sizeof(double) * NVAR * dimp, hipMemcpyHostToDevice));
    for (int s = 1; s < g_ctx.nslots; s++)
        CHK(hipMemcpy(slot(s), g_ctx.wmod, sizeof(double) * NVAR * dimp, hipMemcpyDeviceToDevice));
    CHK(hipMemcpy(g_ctx.wd, *wd, sizeof(double) * NDERV * dimp, hipMemcpyHostToDevice));
    CHK(hipMemset(g_ctx.dwn1, 0, sizeof(double) * NVAR * dimp));
    CHK(hipMemset(g_ctx.wtemp, 0, sizeof(double) * NTEMP * dimp));
    CHK(hipMemcpy(g_ctx.dp, pp, sizeof(struct params), hipMemcpyHostToDevice));
    if (state && *state) CHK(hipMemcpy(g_ctx.ds, *state, sizeof(struct state), hipMemcpyHostToDevice));
    // grid metrics default to the uniform spacing of p->dx until initgrid
    for (int d = 0; d < 3; d++) {
        g_ctx.dx[d].assign(g_ctx.n[d], d < D ? pp->dx[d] : 1.0);
        g_ctx.pos[d].assign(g_ctx.n[d], 0.0);
        for (int i = 0; i < g_ctx.n[d]; i++)
            g_ctx.pos[d][i] = (d < D ? pp->xmin[d] : 0.0) + (i - 2 + 0.5) * g_ctx.dx[d][i];
    }
    g_ctx.ready = true;
    int rc = upload_axes(pp);
    if (rc) return rc;
    if (d_p) *d_p = g_ctx.dp;
    if (d_bp) *d_bp = g_ctx.dbp;
    if (d_wnew) *d_wnew = g_ctx.wnew;
    if (d_wmod) *d_wmod = g_ctx.wmod;
    if (d_dwn1) *d_dwn1 = g_ctx.dwn1;
    if (d_wd) *d_wd = g_ctx.wd;
    if (d_state) *d_state = g_ctx.ds;
    if (d_wtemp) *d_wtemp = g_ctx.wtemp;
    if (d_wtemp1) *d_wtemp1 = g_ctx.wtemp1;
    if (d_wtemp2) *d_wtemp2 = g_ctx.wtemp2;
    return sync_check();
}

// grid metrics from the cell positions in wd[pos1..posD]: dx = (x[i+1] - x[i-1])/2,
// edge cells take their neighbour's value (setupdx_parallel, init.cu:482-600)
extern "C" int initgrid(struct params **p, struct state **state, real **wd, struct params **d_p,
                        real **d_dwn1, real **d_wd, struct state **d_state, real **d_wtemp,
                        real **d_wtemp1, real **d_wtemp2)
{
    (void)state;
    (void)d_p;
    (void)d_dwn1;
    (void)d_state;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    if (!g_ctx.ready) return cfail(SMAUG_EINVAL, "cuinit has not been called");
    if (!p || !*p || !wd || !*wd) return cfail(SMAUG_EINVAL, "null argument");
    double *h = *wd;
    idx_t dimp = g_ctx.dimp;
    idx_t st[3] = {1, g_ctx.n[0], (idx_t)g_ctx.n[0] * g_ctx.n[1]};
    for (int d = 0; d < D; d++) {
        int nd = g_ctx.n[d];
        std::vector<double> pos(nd), dx(nd);
        for (int i = 0; i < nd; i++) pos[i] = h[(idx_t)(pos1 + d) * dimp + i * st[d]];
        bool ok = true;
        for (int i = 1; i < nd - 1; i++) {
            dx[i] = 0.5 * (pos[i + 1] - pos[i - 1]);
            if (!(dx[i] > 0)) ok = false;
        }
        if (ok) {
            dx[0] = dx[1];
            dx[nd - 1] = dx[nd - 2];
            g_ctx.dx[d] = dx;
            g_ctx.pos[d] = pos;
        }
        // tensor-product grid: the metric of a cell depends only on its index along d
        for (idx_t c = 0; c < dimp; c++) {
            int x = (int)((c / st[d]) % nd);
            h[(idx_t)(delx1 + d) * dimp + c] = g_ctx.dx[d][x];
        }
    }
    CHK(hipMemcpy(g_ctx.wd, h, sizeof(double) * NDERV * dimp, hipMemcpyHostToDevice));
    int rc = upload_axes(*p);
    if (rc) return rc;
    return sync_check();
}

extern "C" int cuboundary(struct params **p, struct bparams **bp, struct params **d_p,
                          struct bparams **d_bp, struct state **d_s, real **d_wmod, int order,
                          int idir, int field)
{
    (void)bp;
    (void)d_p;
    (void)d_bp;
    (void)d_s;
    (void)d_wmod;
    int rc = check_params(*p);
    if (rc) return rc;
    if (idir < 0 || idir >= D || field < 0 || field >= NVAR || order < 0 || order >= g_ctx.nslots)
        return cfail(SMAUG_EINVAL, "bad arguments");
    Geo g = make_geo(*p);
    const int bt = (*p)->boundtype[field][idir][0];
    double *wf = slot(order) + (idx_t)field * g.nc;
    if (bt == 5) {
        // the reference's driver calls cuboundary once per fixed face with
        // p->it == -1 before the time loop (iosmaug.cpp:645-659) to record the
        // face values (bc3_setfixed_dir, gradops.cu:1201-1230); later calls
        // restore them (bc3_fixed_dir, 1234-1262)
        if (!g_ctx.fixed) CHK(hipMalloc((void **)&g_ctx.fixed, sizeof(double) * NVAR * g.nc));
        double *fx = g_ctx.fixed + (idx_t)field * g.nc;
        if ((*p)->it == -1) {
            CHK(hipMemcpy(fx, wf, sizeof(double) * g.nc, hipMemcpyDeviceToDevice));
            g_ctx.fixed_set[field] = true;
            return 0;
        }
        if (!g_ctx.fixed_set[field])
            return cfail(SMAUG_EINVAL,
                         "fixed boundary of field %d used before its values were recorded "
                         "(cuboundary with p->it == -1 records them)", field);
        dim3 blk(64, 4, 1);
        hipLaunchKernelGGL(kc_bc, all_grid(blk), blk, 0, 0, g, wf, idir, bt, (const double *)fx);
        return sync_check();
    }
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(kc_bc, all_grid(blk), blk, 0, 0, g, wf, idir, bt, (const double *)nullptr);
    return sync_check();
}

extern "C" int cucomputedervfields(struct params **p, struct params **d_p, real **d_wmod,
                                   real **d_wd, int order)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    int rc = check_params(*p);
    if (rc) return rc;
    // order 0: slot 0 <- slot 1 for the evolved fields (computedervfields.cu:1729-1761)
    if (order == 0) {
        hipLaunchKernelGGL(kc_copy, dim3(2048), dim3(256), 0, 0, slot(0), slot(1),
                           (idx_t)LY::NVE * g_ctx.dimp);
    }
    return sync_check();
}

static int derived(struct params **p, int order, int kind, int dir)
{
    int rc = check_params(*p);
    if (rc) return rc;
    Geo g = make_geo(*p);
    Phys ph = make_phys(*p);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(kc_derived, all_grid(blk), blk, 0, 0, g, ph, (const double *)slot(order),
                       g_ctx.wd, kind, dir);
    return sync_check();
}

extern "C" int cucomputevels(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                             int order, int dir)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    return derived(p, order, 0, dir);
}

extern "C" int cucomputept(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                           int order, int dir)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    return derived(p, order, 1, dir);
}

extern "C" int cucomputepbg(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                            int order, int dir)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    return derived(p, order, 2, dir);
}

extern "C" int cucomputepk(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                           int order, int dir)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    return derived(p, order, 3, dir);
}

extern "C" int cucomputec(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                          int order, int dir)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    return derived(p, order, 4, dir);
}

// p->cmax = max over every cell of cfast_dir (cucomputemaxc,
// computedervfields.cu:1922-2116)
extern "C" int cucomputemaxc(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                             int order, int dir, real **wd, real **d_wtemp)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    (void)wd;
    (void)d_wtemp;
    int rc = check_params(*p);
    if (rc) return rc;
    double v;
    if ((rc = reduce_max(make_geo(*p), make_phys(*p), slot(order), dir, 0, &v))) return rc;
    (*p)->cmax = v;
    return sync_check();
}

// p->maxcourant = max over cells of cfast_dir/dx_dir for this dir only: the
// reference resets it on every call (computedervfields.cu:2118-2296)
extern "C" int cucomputemaxcourant(struct params **p, struct params **d_p, real **d_wmod,
                                   real **d_wd, int order, int dir, real **wd, real **d_wtemp)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    (void)wd;
    (void)d_wtemp;
    int rc = check_params(*p);
    if (rc) return rc;
    double v;
    if ((rc = reduce_max(make_geo(*p), make_phys(*p), slot(order), dir, 1, &v))) return rc;
    (*p)->maxcourant = v;
    return sync_check();
}

extern "C" int cucentdiff1(struct params **p, struct params **d_p, struct state **d_s,
                           real **d_w, real **d_wmod, real **d_dwn1, real **d_wd, int order,
                           int ordero, real dt, int field, int dir)
{
    (void)d_p;
    (void)d_s;
    (void)d_w;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    int rc = check_params(*p);
    if (rc) return rc;
    if (field < LY::RHO || field > LY::M0 + D - 1 || dir < 0 || dir >= D)
        return cfail(SMAUG_EINVAL, "cucentdiff1: field %d dir %d", field, dir);
    Geo g = make_geo(*p);
    Phys ph = make_phys(*p);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(kc_centdiff, interior_grid(blk), blk, 0, 0, g, ph, (const double *)slot(order),
                       slot(ordero), field, dir, dt, (*p)->dt);
    return sync_check();
}

extern "C" int cucentdiff2(struct params **p, struct params **d_p, struct state **d_s,
                           real **d_w, real **d_wmod, real **d_dwn1, real **d_wd, int order,
                           int ordero, real dt, int field, int dir)
{
    (void)d_p;
    (void)d_s;
    (void)d_w;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    int rc = check_params(*p);
    if (rc) return rc;
    if (field < LY::EN || field > LY::B0 + D - 1 || dir < 0 || dir >= D)
        return cfail(SMAUG_EINVAL, "cucentdiff2: field %d dir %d", field, dir);
    Geo g = make_geo(*p);
    Phys ph = make_phys(*p);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(kc_centdiff, interior_grid(blk), blk, 0, 0, g, ph, (const double *)slot(order),
                       slot(ordero), field, dir, dt, (*p)->dt);
    return sync_check();
}

extern "C" int cugrav(struct params **p, struct params **d_p, struct state **d_s, real **d_w,
                      real **d_wmod, real **d_dwn1, real **d_wd, int order, int ordero, real dt)
{
    (void)d_p;
    (void)d_s;
    (void)d_w;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)dt;
    int rc = check_params(*p);
    if (rc) return rc;
    Geo g = make_geo(*p);
    Phys ph = make_phys(*p);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(kc_grav, interior_grid(blk), blk, 0, 0, g, ph, (const double *)slot(order),
                       slot(ordero), (*p)->dt);
    return sync_check();
}

extern "C" int cusource(struct params **p, struct params **d_p, struct state **d_s, real **d_w,
                        real **d_wmod, real **d_dwn1, real **d_wd, int order, int ordero, real dt)
{
    (void)d_p;
    (void)d_s;
    (void)d_w;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)dt;
    int rc = check_params(*p);
    if (rc) return rc;
    if (!g_ctx.driver) return 0;
    Geo g = make_geo(*p);
    Phys ph = make_phys(*p);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(kc_source, interior_grid(blk), blk, 0, 0, g, ph, (const double *)slot(order),
                       slot(ordero), (*p)->dt, (*p)->qt);
    return sync_check();
}

extern "C" int cusetdriver(int enable, const real drv[7])
{
    if (enable < 0 || enable > 2) return cfail(SMAUG_EINVAL, "cusetdriver: enable must be 0, 1 or 2");
    if (enable == 2 && D != 3)
        return cfail(SMAUG_EINVAL, "cusetdriver: the torsional driver drives m_y and m_z (3-D only)");
    g_ctx.driver = enable;
    if (enable && drv)
        for (int k = 0; k < 7; k++) g_ctx.drv[k] = drv[k];
    if (enable == 2) {
        // the reference's constants (usersource.cu:26-57)
        static const double ref[8] = {500.0, 30.0, 0.1e6, 0.1e6, 0.05e6, 2.0e6, 2.0e6, 4.0e6};
        if (!drv)
            for (int k = 0; k < 7; k++) g_ctx.drv[k] = ref[k];
        g_ctx.drv[7] = ref[7];
    }
    if (g_ctx.ready) {
        struct params dummy;
        memset(&dummy, 0, sizeof(dummy));
        return upload_axes(&dummy);
    }
    return 0;
}

extern "C" int cudivb(struct params **p, struct params **d_p, real **d_w, real **d_wmod,
                      real **d_dwn1, real **d_wd, int order, int ordero, real dt)
{
    (void)d_p;
    (void)d_w;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)order;
    (void)ordero;
    (void)dt;
    if ((*p)->divbon == 1) return cfail(SMAUG_EINVAL, "divbon (Powell div-B source) is not supported");
    return 0;
}

extern "C" int cunushk1(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                        int order, real **d_wtemp, real **d_wtemp1, real **d_wtemp2)
{
    (void)p;
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    (void)order;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    return cfail(SMAUG_EINVAL, "shock viscosity (USE_SHOCKVISC) is not supported");
}

static int nu_side(struct params **p, int order, int field, int dim, int side)
{
    int rc = check_params(*p);
    if (rc) return rc;
    if (field < 0 || field >= LY::NVE || dim < 0 || dim >= D)
        return cfail(SMAUG_EINVAL, "hyperviscosity: field %d dim %d", field, dim);
    Geo g = make_geo(*p);
    Phys ph = make_phys(*p);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(kc_nu, all_grid(blk), blk, 0, 0, g, ph, (const double *)slot(order), g_ctx.wd,
                       field, dim, side, (*p)->cmax);
    if ((rc = sync_check())) return rc;
    if ((*p)->moddton == 1) {
        // max of the coefficient for the hyperviscous dt limit
        std::vector<double> h(g.nc);
        CHK(hipMemcpy(h.data(), g_ctx.wd + (idx_t)(side == 0 ? hdnur : hdnul) * g.nc,
                      sizeof(double) * g.nc, hipMemcpyDeviceToHost));
        double m = (*p)->maxviscoef;
        for (idx_t c = 0; c < g.nc; c++) m = h[c] > m ? h[c] : m;
        (*p)->maxviscoef = m;
    }
    return 0;
}

extern "C" int cuhyperdifvisc1r(struct params **p, struct params **d_p, real **d_wmod, real **wd,
                                real **d_wd, int order, real **d_wtemp, real **d_wtemp1,
                                real **d_wtemp2, int field, int dim)
{
    (void)d_p;
    (void)d_wmod;
    (void)wd;
    (void)d_wd;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    return nu_side(p, order, field, dim, 0);
}

extern "C" int cuhyperdifvisc1l(struct params **p, struct params **d_p, real **d_wmod, real **wd,
                                real **d_wd, int order, real **d_wtemp, real **d_wtemp1,
                                real **d_wtemp2, int field, int dim)
{
    (void)d_p;
    (void)d_wmod;
    (void)wd;
    (void)d_wd;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    return nu_side(p, order, field, dim, 1);
}

// the reference's 1ir/1il only prepare its temporaries; the coefficients are
// formed in cuhyperdifvisc1r/1l here
extern "C" int cuhyperdifvisc1ir(struct params **p, struct params **d_p, real **d_wmod,
                                 real **d_wd, int order, real **d_wtemp, real **d_wtemp1,
                                 real **d_wtemp2, int field, int dim)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    (void)order;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    (void)field;
    (void)dim;
    return check_params(*p);
}

extern "C" int cuhyperdifvisc1il(struct params **p, struct params **d_p, real **d_wmod,
                                 real **d_wd, int order, real **d_wtemp, real **d_wtemp1,
                                 real **d_wtemp2, int field, int dim)
{
    return cuhyperdifvisc1ir(p, d_p, d_wmod, d_wd, order, d_wtemp, d_wtemp1, d_wtemp2, field, dim);
}

extern "C" int cuhyperdifvisc1(struct params **p, struct params **d_p, real **d_wmod, real **d_wd,
                               int order, real **d_wtemp, real **d_wtemp1, real **d_wtemp2,
                               int field, int dim, int hand)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_wd;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    return nu_side(p, order, field, dim, hand == 0 ? 0 : 1);
}

extern "C" int cuhyperdifviscmax(struct params **p, real **w, real **wnew, struct params **d_p,
                                 real **d_w, real **d_wnew, real **d_wmod, real **d_dwn1,
                                 real **d_wd, int order, real **d_wtemp, int field, int dim)
{
    (void)w;
    (void)wnew;
    (void)d_p;
    (void)d_w;
    (void)d_wnew;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)order;
    (void)d_wtemp;
    (void)field;
    (void)dim;
    int rc = check_params(*p);
    if (rc) return rc;
    std::vector<double> h(g_ctx.dimp);
    double m = (*p)->maxviscoef;
    for (int side = 0; side < 2; side++) {
        CHK(hipMemcpy(h.data(), g_ctx.wd + (idx_t)(side == 0 ? hdnur : hdnul) * g_ctx.dimp,
                      sizeof(double) * g_ctx.dimp, hipMemcpyDeviceToHost));
        for (idx_t c = 0; c < g_ctx.dimp; c++) m = h[c] > m ? h[c] : m;
    }
    (*p)->maxviscoef = m;
    return 0;
}

// dt limit of the hyperviscosity, 0.25 dx^2 / max nu over the dims (SAC
// getdt_visc, vacusr.viscosity.t:512-546; the reference keeps the same
// formula in a comment of nushk1.cu:688-694)
extern "C" int cugetdtvisc1(struct params **p, struct params **d_p, real **d_wmod, real **wd,
                            real **d_wd, int order, real **d_wtemp, real **d_wtemp1,
                            real **d_wtemp2)
{
    (void)d_p;
    (void)d_wmod;
    (void)wd;
    (void)d_wd;
    (void)order;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    int rc = check_params(*p);
    if (rc) return rc;
    double best = 0.0;
    if ((*p)->maxviscoef > 0) {
        best = 1e300;
        for (int d = 0; d < D; d++) {
            double mn = 1e300;
            for (double v : g_ctx.dx[d]) mn = v < mn ? v : mn;
            double t = 0.25 * mn * mn / (*p)->maxviscoef;
            best = t < best ? t : best;
        }
    }
    (*p)->dtdiffvisc = best;
    return 0;
}

static int hd_diag(struct params **p, int order, int ordero, int kind, int qf, int tgt, int af,
                   int dim, double coef, double dt)
{
    int rc = check_params(*p);
    if (rc) return rc;
    Geo g = make_geo(*p);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(kc_hd_diag, interior_grid(blk), blk, 0, 0, g, (const double *)slot(order),
                       slot(ordero), (const double *)g_ctx.wd, kind, qf, tgt, af, dim, coef, dt);
    return sync_check();
}

static int hd_off(struct params **p, int order, int ordero, int qf, int dim, int od, int tgt,
                  int wrho, int xf, double fac, double sg, double dt)
{
    int rc = check_params(*p);
    if (rc) return rc;
    Geo g = make_geo(*p);
    dim3 blk(64, 4, 1);
    double *tA = g_ctx.wtemp, *tB = g_ctx.wtemp + g.nc;
    hipLaunchKernelGGL(kc_off_inner, all_grid(blk), blk, 0, 0, g, (const double *)slot(order),
                       (const double *)g_ctx.wd, tA, tB, qf, dim, wrho, xf, fac);
    hipLaunchKernelGGL(kc_off_outer, interior_grid(blk), blk, 0, 0, g, (const double *)tA,
                       (const double *)tB, slot(ordero), od, tgt, sg, dt);
    return sync_check();
}

extern "C" int cuhyperdifrhosource1(struct params **p, struct params **d_p, real **d_wmod,
                                    real **d_dwn1, real **d_wd, int order, int ordero,
                                    real **d_wtemp, int field, int dim, real dt)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)d_wtemp;
    (void)field;
    return hd_diag(p, order, ordero, 0, LY::RHO, LY::RHO, 0, dim, 1.0, dt);
}

extern "C" int cuhyperdifesource1(struct params **p, struct params **d_p, real **d_wmod,
                                  real **d_dwn1, real **d_wd, int order, int ordero,
                                  real **d_wtemp, int field, int dim, real dt)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)d_wtemp;
    (void)field;
    return hd_diag(p, order, ordero, 1, LY::EN, LY::EN, 0, dim, 1.0, dt);
}

// field: velocity component whose gradient diffuses; ii0: momentum updated
extern "C" int cuhyperdifmomsource1(struct params **p, struct params **d_p, real **d_wmod,
                                    real **d_dwn1, real **d_wd, int order, int ordero,
                                    real **d_wtemp, int field, int dim, int ii, int ii0, real dt)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)d_wtemp;
    (void)ii;
    return hd_diag(p, order, ordero, 2, LY::M0 + field, LY::M0 + ii0, LY::M0 + ii0, dim, 0.5, dt);
}

extern "C" int cuhyperdifmomsourcene1(struct params **p, struct params **d_p, real **d_wmod,
                                      real **d_dwn1, real **d_wd, int order, int ordero,
                                      real **d_wtemp, int field, int dim, int ii, int ii0,
                                      real dt)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)d_wtemp;
    return hd_off(p, order, ordero, LY::M0 + field, dim, ii, LY::M0 + ii0, 1, LY::M0 + ii0, 0.25,
                  1.0, dt);
}

extern "C" int cuhyperdifbsource1(struct params **p, struct params **d_p, real **d_wmod,
                                  real **d_dwn1, real **d_wd, int order, int ordero,
                                  real **d_wtemp, int field, int dim, int jj, int ii0, int mm,
                                  real sb, real dt)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)d_wtemp;
    (void)mm;
    return hd_diag(p, order, ordero, 3, LY::B0 + field, LY::B0 + ii0, LY::B0 + jj, dim, sb, dt);
}

extern "C" int cuhyperdifbsourcene1(struct params **p, struct params **d_p, real **d_wmod,
                                    real **d_dwn1, real **d_wd, int order, int ordero,
                                    real **d_wtemp, int field, int dim, int jj, int ii0, int mm,
                                    real sb, real dt)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)d_wtemp;
    return hd_off(p, order, ordero, LY::B0 + field, dim, mm, LY::B0 + ii0, 0, LY::B0 + jj, 0.5, sb,
                  dt);
}

extern "C" int cuadvance(struct params **p, struct params **d_p, real **d_wmod, real **d_w,
                         int order)
{
    (void)d_p;
    (void)d_wmod;
    (void)d_w;
    int rc = check_params(*p);
    if (rc) return rc;
    if ((*p)->rkon == 1) {
        if (g_ctx.nslots < 4) return cfail(SMAUG_EINVAL, "rkon=1 needs cuinit with rkon=1");
        hipLaunchKernelGGL(kc_advance, dim3(2048), dim3(256), 0, 0, g_ctx.wmod, g_ctx.dimp, order);
    } else {
        hipLaunchKernelGGL(kc_copy, dim3(2048), dim3(256), 0, 0, slot(0), slot(1),
                           (idx_t)LY::NVE * g_ctx.dimp);
    }
    return sync_check();
}

extern "C" int cuupdate(struct params **p, real **w, real **wmod, real **wtemp2,
                        struct state **state, struct params **d_p, real **d_w, real **d_wmod,
                        real **d_wtemp2, struct state **d_state, int step)
{
    (void)p;
    (void)w;
    (void)wtemp2;
    (void)state;
    (void)d_p;
    (void)d_w;
    (void)d_wmod;
    (void)d_wtemp2;
    (void)d_state;
    (void)step;
    if (!g_ctx.ready) return cfail(SMAUG_EINVAL, "cuinit has not been called");
    if (!wmod || !*wmod) return cfail(SMAUG_EINVAL, "null argument");
    CHK(hipMemcpy(*wmod, g_ctx.wmod, sizeof(double) * NVAR * g_ctx.dimp, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int cuupdatehostwd(struct params **p, real **wd, real **wmod, real **wtemp2,
                              struct state **state, struct params **d_p, real **d_wd,
                              real **d_wmod, real **d_wtemp2, struct state **d_state, int step)
{
    (void)p;
    (void)wmod;
    (void)wtemp2;
    (void)state;
    (void)d_p;
    (void)d_wd;
    (void)d_wmod;
    (void)d_wtemp2;
    (void)d_state;
    (void)step;
    if (!g_ctx.ready) return cfail(SMAUG_EINVAL, "cuinit has not been called");
    CHK(hipMemcpy(*wd, g_ctx.wd, sizeof(double) * NDERV * g_ctx.dimp, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int cuupdatedevicewd(struct params **p, real **wd, real **wmod, real **wtemp2,
                                struct state **state, struct params **d_p, real **d_wd,
                                real **d_wmod, real **d_wtemp2, struct state **d_state, int step)
{
    (void)p;
    (void)wmod;
    (void)wtemp2;
    (void)state;
    (void)d_p;
    (void)d_wd;
    (void)d_wmod;
    (void)d_wtemp2;
    (void)d_state;
    (void)step;
    if (!g_ctx.ready) return cfail(SMAUG_EINVAL, "cuinit has not been called");
    CHK(hipMemcpy(g_ctx.wd, *wd, sizeof(double) * NDERV * g_ctx.dimp, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int cuupdatemod(struct params **p, struct bparams **bp, real **w, real **wnew,
                           real **wd, struct state **state, struct params **d_p,
                           struct bparams **d_bp, real **d_w, real **d_wnew, real **d_wmod,
                           real **d_dwn1, real **d_wd, struct state **d_state, real **d_wtemp,
                           real **d_wtemp1, real **d_wtemp2)
{
    (void)p;
    (void)bp;
    (void)wnew;
    (void)wd;
    (void)state;
    (void)d_p;
    (void)d_bp;
    (void)d_w;
    (void)d_wnew;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)d_state;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    if (!g_ctx.ready) return cfail(SMAUG_EINVAL, "cuinit has not been called");
    // host slot 0 -> every device slot (init.cu updatemod)
    CHK(hipMemcpy(g_ctx.wmod, *w, sizeof(double) * NVAR * g_ctx.dimp, hipMemcpyHostToDevice));
    for (int s = 1; s < g_ctx.nslots; s++)
        CHK(hipMemcpy(slot(s), g_ctx.wmod, sizeof(double) * NVAR * g_ctx.dimp, hipMemcpyDeviceToDevice));
    return 0;
}

extern "C" int cufinish(struct params **p, real **w, real **wnew, struct state **state,
                        struct params **d_p, struct bparams **d_bp, real **d_w, real **d_wnew,
                        real **d_wmod, real **d_dwn1, real **d_wd, struct state **d_state,
                        real **d_wtemp, real **d_wtemp1, real **d_wtemp2)
{
    (void)p;
    (void)w;
    (void)wnew;
    (void)state;
    (void)d_p;
    (void)d_bp;
    (void)d_w;
    (void)d_wnew;
    (void)d_wmod;
    (void)d_dwn1;
    (void)d_wd;
    (void)d_state;
    (void)d_wtemp;
    (void)d_wtemp1;
    (void)d_wtemp2;
    if (!g_ctx.ready) return 0;
    for (int k = 0; k < 2; k++)
        if (g_ctx.eng[k]) smaug_destroy(g_ctx.eng[k]);
    double *bufs[] = {g_ctx.wmod, g_ctx.wd, g_ctx.dwn1, g_ctx.wtemp, g_ctx.wtemp1, g_ctx.wtemp2,
                      g_ctx.wnew};
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    for (int d = 0; d < 3; d++) {
        if (g_ctx.d_dx[d]) (void)hipFree(g_ctx.d_dx[d]);
        if (g_ctx.d_idx[d]) (void)hipFree(g_ctx.d_idx[d]);
        if (g_ctx.d_prof[d]) (void)hipFree(g_ctx.d_prof[d]);
        if (g_ctx.d_profq[d]) (void)hipFree(g_ctx.d_profq[d]);
    }
    (void)hipFree(g_ctx.dp);
    (void)hipFree(g_ctx.ds);
    (void)hipFree(g_ctx.dbp);
    (void)hipFree(g_ctx.d_red);
    if (g_ctx.fixed) (void)hipFree(g_ctx.fixed);
    g_ctx.fixed = nullptr;
    for (int f = 0; f < NVAR; f++) g_ctx.fixed_set[f] = false;
    int drv = g_ctx.driver;
    double dv[8];
    memcpy(dv, g_ctx.drv, sizeof(dv));
    g_ctx = Ctx();
    g_ctx.driver = drv;
    memcpy(g_ctx.drv, dv, sizeof(dv));
    return 0;
}

extern "C" int cusync(struct params **p)
{
    (void)p;
    CHK(hipDeviceSynchronize());
    return 0;
}

// ---------------------------------------------------------------------------
// fused RK4 on the engine
// ---------------------------------------------------------------------------
extern "C" int cuadvancerk4(struct params **p, struct params **d_p, real **d_wmod, int nsteps,
                            int precision)
{
    (void)d_p;
    (void)d_wmod;
    int rc = check_params(*p);
    if (rc) return rc;
    if (precision != SMAUG_FP64 && precision != SMAUG_MIXED)
        return cfail(SMAUG_EINVAL, "precision must be 0 (FP64) or 1 (mixed)");
    const struct params *pp = *p;
    idx_t dimp = g_ctx.dimp;
    if (!g_ctx.eng[precision]) {
        smaug_cfg cfg;
        memset(&cfg, 0, sizeof(cfg));
        cfg.ndim = D;
        for (int d = 0; d < 3; d++) {
            cfg.n[d] = g_ctx.n[d];
            cfg.bc[d] = d < D ? pp->boundtype[0][d][0] : 0;
            cfg.g[d] = d < D ? pp->g[d] : 0.0;
            cfg.dx[d] = g_ctx.dx[d].data();
            cfg.pos[d] = g_ctx.pos[d].data();
        }
        cfg.hyperdif = pp->hyperdifmom;
        cfg.driver = g_ctx.driver;
        cfg.precision = precision;
        cfg.gamma = pp->gamma;
        for (int f = 0; f < NVAR; f++) cfg.chyp[f] = pp->chyp[f];
        for (int k = 0; k < 8; k++) cfg.drv[k] = g_ctx.drv[k];
        if ((rc = smaug_create(&cfg, &g_ctx.eng[precision])))
            return cfail(rc, "%s", smaug_last_error());
    }
    smaug_engine *e = g_ctx.eng[precision];
    std::vector<double> h((size_t)NVAR * dimp);
    CHK(hipMemcpy(h.data(), g_ctx.wmod, sizeof(double) * NVAR * dimp, hipMemcpyDeviceToHost));
    const double *w = h.data(), *bg = h.data() + (idx_t)LY::NVE * dimp;
    if ((rc = smaug_set_state(e, w, bg))) return cfail(rc, "%s", smaug_last_error());
    if ((rc = smaug_rk4(e, pp->qt, pp->dt, nsteps))) return cfail(rc, "%s", smaug_last_error());
    if ((rc = smaug_get_state(e, h.data()))) return cfail(rc, "%s", smaug_last_error());
    CHK(hipMemcpy(g_ctx.wmod, h.data(), sizeof(double) * LY::NVE * dimp, hipMemcpyHostToDevice));
    CHK(hipMemcpy(slot(1), h.data(), sizeof(double) * LY::NVE * dimp, hipMemcpyHostToDevice));
    (*p)->qt += nsteps * pp->dt;
    return sync_check();
}
