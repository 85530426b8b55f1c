// smaug_phys.hpp -- device-side layout and pointwise physics shared by the
// fused stage kernels and the per-entry-point kernels.
//
// Field layout follows the reference's enum order (dev/smaug/include/iotypes.h:213-221):
//   evolved    0: rho'   1..D: m_d   1+D: e'   2+D..1+2D: b'_d
//   background 0: e_b    1: rho_b    2..1+D: b_bd
// T is the storage/arithmetic type of the evolved (perturbation) fields:
// double for the plain FP64 path, float for the mixed path.  Background
// fields are always double; pressure recovery and internal energy are
// evaluated in double in both paths.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smaug {

typedef long long idx_t;

struct Geo {
    int n[3];         // cells incl. ghosts per dim (n[2] == 1 in 2-D)
    int per[3];       // 1 if periodic along d (boundtype 0)
    idx_t s1, s2, nc; // strides of dims 1 and 2, cell count
    // background as the per-cell kernels read it: field f of cell c at
    // bg[f * bgf + (c & bgm)]; a full array has bgf = nc, bgm = ~0, a
    // background that is the same in every cell is passed as its nb
    // values with bgf = 1, bgm = 0
    idx_t bgf;
    unsigned bgm;
    const double *dx[3];
    const double *idx[3];  // 1/dx per cell
    const double *prof[3]; // driver profile factors per dim
    const double *profq[3]; // torsional driver: (x_d - c_d) prof[d] along dims 1, 2
};

struct Phys {
    double gamma;
    double g[3];
    double chyp[13];
    double drv[8];
    int hyperdif;
    int driver;
};

template <int D>
struct Lay {
    static constexpr int NVE = 2 + 2 * D;
    static constexpr int NB = 2 + D;
    static constexpr int RHO = 0;
    static constexpr int EN = 1 + D;
    static constexpr int M0 = 1;
    static constexpr int B0 = 2 + D;
    // (q, dim) pairs for which a hyperviscosity ratio is formed:
    // rho' x D, e_int x D, v_f x D*D, b'_f (f != dim) x D*(D-1)
    static constexpr int NPAIR = 2 * D + D * D + D * (D - 1);
    __host__ __device__ static constexpr int prho(int d) { return d; }
    __host__ __device__ static constexpr int pe(int d) { return D + d; }
    __host__ __device__ static constexpr int pv(int f, int d) { return 2 * D + f * D + d; }
    __host__ __device__ static constexpr int pb(int f, int d)
    {
        return 2 * D + D * D + f * (D - 1) + (d < f ? d : d - 1);
    }
};

template <int D>
__device__ __forceinline__ idx_t stride(const Geo &g, int d)
{
    return d == 0 ? (idx_t)1 : (d == 1 ? g.s1 : g.s2);
}

// Cell index within one field for the fused kernels.  A grid is limited to
// 2^29 cells so that the byte offset of a cell in a double field fits in 32
// bits: each access is then one 32-bit VGPR offset against a field base
// address held in SGPRs (the global_load saddr form), shared by every field.
typedef unsigned int ci_t;
constexpr long long MAX_CELLS = 1ll << 29;

template <typename T>
__device__ __forceinline__ T ld(const T *__restrict__ base, ci_t c)
{
    return *(const T *)((const char *)base + (ci_t)(c * (ci_t)sizeof(T)));
}

template <typename T>
__device__ __forceinline__ void st(T *__restrict__ base, ci_t c, T v)
{
    *(T *)((char *)base + (ci_t)(c * (ci_t)sizeof(T))) = v;
}

template <int D>
__device__ __forceinline__ ci_t stride32(const Geo &g, int d)
{
    return d == 0 ? 1u : (d == 1 ? (ci_t)g.s1 : (ci_t)g.s2);
}

// one cell's state
template <int D, typename T>
struct St {
    T r, e, m[D], b[D];   // perturbations
    double rb, eb, bb[D]; // background
};

template <int D, typename T>
__device__ __forceinline__ void load_st(St<D, T> &s, const T *__restrict__ u,
                                        const double *__restrict__ bg, idx_t nc, idx_t c)
{
    s.r = u[c];
#pragma unroll
    for (int d = 0; d < D; d++) s.m[d] = u[(idx_t)(1 + d) * nc + c];
    s.e = u[(idx_t)(1 + D) * nc + c];
#pragma unroll
    for (int d = 0; d < D; d++) s.b[d] = u[(idx_t)(2 + D + d) * nc + c];
    s.eb = bg[c];
    s.rb = bg[nc + c];
#pragma unroll
    for (int d = 0; d < D; d++) s.bb[d] = bg[(idx_t)(2 + d) * nc + c];
}

// the perturbation fields of cell c only (the background is left as it is)
template <int D, typename T>
__device__ __forceinline__ void load_pert32(St<D, T> &s, const T *__restrict__ u, size_t nc, ci_t c)
{
    s.r = ld(u, c);
#pragma unroll
    for (int d = 0; d < D; d++) s.m[d] = ld(u + (1 + d) * nc, c);
    s.e = ld(u + (1 + D) * nc, c);
#pragma unroll
    for (int d = 0; d < D; d++) s.b[d] = ld(u + (2 + D + d) * nc, c);
}

// the background fields of a cell from an array of field stride nb (nc for
// a full background, the plane size for a background stored as one k-plane)
template <int D, typename T>
__device__ __forceinline__ void load_bg32(St<D, T> &s, const double *__restrict__ bg, size_t nb,
                                          ci_t c)
{
    s.eb = ld(bg, c);
    s.rb = ld(bg + nb, c);
#pragma unroll
    for (int d = 0; d < D; d++) s.bb[d] = ld(bg + (2 + d) * nb, c);
}

template <int D, typename T>
__device__ __forceinline__ void load_st32(St<D, T> &s, const T *__restrict__ u,
                                          const double *__restrict__ bg, size_t nc, ci_t c)
{
    s.r = ld(u, c);
#pragma unroll
    for (int d = 0; d < D; d++) s.m[d] = ld(u + (1 + d) * nc, c);
    s.e = ld(u + (1 + D) * nc, c);
#pragma unroll
    for (int d = 0; d < D; d++) s.b[d] = ld(u + (2 + D + d) * nc, c);
    s.eb = ld(bg, c);
    s.rb = ld(bg + nc, c);
#pragma unroll
    for (int d = 0; d < D; d++) s.bb[d] = ld(bg + (2 + d) * nc, c);
}

// the perturbations and rho_b only (what the hyperdiffusion variables need)
template <int D, typename T>
__device__ __forceinline__ void load_light(St<D, T> &s, const T *__restrict__ u,
                                           const double *__restrict__ bg, idx_t nc, idx_t c)
{
    s.r = u[c];
#pragma unroll
    for (int d = 0; d < D; d++) s.m[d] = u[(idx_t)(1 + d) * nc + c];
    s.e = u[(idx_t)(1 + D) * nc + c];
#pragma unroll
    for (int d = 0; d < D; d++) s.b[d] = u[(idx_t)(2 + D + d) * nc + c];
    s.rb = bg[nc + c];
    s.eb = 0;
#pragma unroll
    for (int d = 0; d < D; d++) s.bb[d] = 0;
}

// Block-to-tile map that keeps the L2 of each XCD on one sub-domain.
// Workgroups are handed to the 8 XCDs round-robin in launch order, so
// block b runs on XCD b % 8.  The (tile_x, tile_y) tiles of a plane are
// grouped into panels of tpx x tpy tiles; panel p belongs to XCD p % 8,
// and an XCD walks each of its panels plane by plane (k innermost over
// panels), so the k-1..k+2 planes a stencil touches are still in its L2.
struct TileMap {
    int ntx, nty, nk;  // tiles along x, y and planes
    int tpx, tpy;      // panel size in tiles
    int npx, npanels;  // panels along x, total panels
};

__device__ __forceinline__ bool map_block(const TileMap &m, int &tx, int &ty, int &k)
{
    const int b = blockIdx.x;
    const int xcd = b & 7;
    const int local = b >> 3;
    const int tpp = m.tpx * m.tpy;
    const int per_panel = tpp * m.nk;
    const int pidx = local / per_panel;
    const int panel = xcd + 8 * pidx;
    if (panel >= m.npanels) return false;
    const int rem = local - pidx * per_panel;
    k = rem / tpp;
    const int t = rem - k * tpp;
    tx = (panel % m.npx) * m.tpx + t % m.tpx;
    ty = (panel / m.npx) * m.tpy + t / m.tpx;
    return tx < m.ntx && ty < m.nty;
}

// full density rho' + rho_b (formed in double)
template <int D, typename T>
__device__ __forceinline__ double rho_full(const St<D, T> &s)
{
    return (double)s.r + s.rb;
}

// 1/x for a positive normal x (a density) in FP64 without the full IEEE
// division: the hardware estimate and two Newton steps, within an ulp or so
// of the correctly rounded quotient
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// total-pressure perturbation, computept3 (dervfields.cu:284-330):
// (g-1)(e' - m^2/2rho) - (g-2)(b'.b_b + b'^2/2), in double.  m^2/rho is
// taken as a product with 1/rho (rcp64, which the callers form anyway and
// the compiler shares), and every fused multiply-add is written out with
// contraction off, so that every kernel that inlines this gets the same bits.
template <int D, typename T>
__device__ __forceinline__ double ptot_pert_ir(const St<D, T> &s, double ir, double gam)
{
#pragma clang fp contract(off)
    double mm = 0, bgb = 0, bb = 0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const double m = s.m[d], b = s.b[d];
        mm = fma(m, m, mm);
        bgb = fma(b, (double)s.bb[d], bgb);
        bb = fma(b, b, bb);
    }
    const double ptmp = (gam - 2.0) * fma(0.5, bb, bgb);
    return (gam - 1.0) * fma(-0.5 * mm, ir, (double)s.e) - ptmp;
}

template <int D, typename T>
__device__ __forceinline__ double ptot_pert(const St<D, T> &s, double rho, double gam)
{
    return ptot_pert_ir<D, T>(s, rcp64(rho), gam);
}

// background total pressure, computepbg3 (dervfields.cu:332-368)
template <int D, typename T>
__device__ __forceinline__ double ptot_bg(const St<D, T> &s, double gam)
{
    double bbb = 0;
#pragma unroll
    for (int d = 0; d < D; d++) bbb += s.bb[d] * s.bb[d];
    return (gam - 1) * s.eb - 0.5 * (gam - 2) * bbb;
}

// internal-energy perturbation used by the energy hyperdiffusion,
// e' - (b'^2 + m^2/rho)/2 (hyperdifesource1.cu:300-330), in double; 1/rho
// and contraction as in ptot_pert
template <int D, typename T>
__device__ __forceinline__ double e_int(const St<D, T> &s, double rho)
{
#pragma clang fp contract(off)
    double bb = 0, mm = 0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const double b = s.b[d], m = s.m[d];
        bb = fma(b, b, bb);
        mm = fma(m, m, mm);
    }
    return fma(-0.5, fma(mm, rcp64(rho), bb), (double)s.e);
}

// fast magnetosonic speed along d plus |v_d|, computec3 (dervfields.cu:418-486), in double.
// The five divisions by rho of that formula are products with 1/rho here
// (rcp64, one reciprocal per cell instead of nine divisions once the
// d-independent terms are shared): the result differs from the divided form
// by rounding only.
// Contraction into FMAs is off, so that every kernel that inlines this gets
// the same bits whatever else the products feed.
template <int D, typename T>
__device__ __forceinline__ double cfast(const St<D, T> &s, int d, double gam)
{
#pragma clang fp contract(off)
    const double rho = rho_full(s);
    const double ir = rcp64(rho);
    double mm = 0, bb = 0, bgb = 0, bbb = 0;
#pragma unroll
    for (int t = 0; t < D; t++) {
        double m = s.m[t], b = s.b[t], B = s.bb[t];
        mm += m * m;
        bb += b * b;
        bgb += B * b;
        bbb += B * B;
    }
    double pk = (gam - 1) * ((double)s.e - 0.5 * (mm * ir) - 0.5 * bb - bgb);
    double cs2 = (gam * (pk + (gam - 1) * (s.eb - 0.5 * bbb))) * ir;
    double cf2 = ((bb + bbb + 2.0 * bgb) * ir) + cs2;
    double btd = 0, md = 0;
#pragma unroll
    for (int t = 0; t < D; t++)
        if (t == d) {
            btd = s.bb[t] + (double)s.b[t];
            md = s.m[t];
        }
    double cfi = 0.5 * (cf2 + sqrt(cf2 * cf2 - 4.0 * cs2 * (btd * btd) * ir));
    return sqrt(cfi) + fabs(md * ir);
}

// edge rules of the reference stencils along a non-periodic direction:
// the 4th-order divergence is zero at x in {2,3,n-4,n-3} (grad3dn,
// gradops.cu:437-602); the one-cell differences are zero at
// x in {1,2,n-3,n-2} (grad1l3n/grad1r3n/grad13n, gradops.cu:608-830)
__device__ __forceinline__ bool zero4(int per, int x, int n)
{
    return !per && (x == 2 || x == 3 || x == n - 4 || x == n - 3);
}
__device__ __forceinline__ bool zero1(int per, int x, int n)
{
    return !per && (x == 1 || x == 2 || x == n - 3 || x == n - 2);
}

} // namespace smaug
