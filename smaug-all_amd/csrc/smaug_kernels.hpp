// smaug_kernels.hpp -- fused per-stage kernels of the MI355X path.
//
// One RK stage of the semi-discrete operator L(w) is four launches:
//   k_prep    : per-cell 1/rho, pressure perturbation, internal energy and
//               velocity (double arithmetic) into an aux buffer, and cmax_d =
//               max over all cells of cfast_d (wavefront-shuffle max, one
//               64-bit atomic max per block); replaces cucomputec +
//               cucomputemaxc (computedervfields.cu:1922-2116)
//   k_ratio   : per cell the hyperviscosity ratio max|D3|/max|D1| of every
//               (q, dim) pair; replaces the hyperdifvisc1r/1l kernel chains
//   k_update  : per interior cell the whole of L(w) -- 4th-order flux
//               divergence of every field in every direction (centdiff1/2),
//               background work terms, gravity, all hyperdiffusive sources
//               and the user driver -- fused with the RK slot algebra
//   k_bc      : ghost fill of the output state (cuboundary)
#pragma once
#include "smaug_phys.hpp"

namespace smaug {

// ---------------------------------------------------------------------------
// cmax
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

static __global__ void k_reset_cmax(unsigned long long *cm)
{
    if (threadIdx.x < 3) cm[threadIdx.x] = 0ull;
}

// Per-cell auxiliary quantities of a state, formed once per stage and read
// by the ratio and update kernels instead of being recomputed at every
// neighbour visit: aux[0] = 1/rho, aux[1] = total-pressure perturbation
// (double arithmetic, computept3), aux[2] = e' - (b'^2 + m^2/rho)/2,
// aux[3+d] = m_d/rho.  The same pass forms cmax_d = max over every cell of
// cfast_d (cucomputec + cucomputemaxc, computedervfields.cu:1922-2116) with a
// wavefront-shuffle max and one 64-bit atomic max per block.
template <int D>
struct Aux {
    static constexpr int IRHO = 0, PT = 1, EINT = 2, V0 = 3, N = 3 + D;
};

// aux of cell c (state s loaded) and the running cfast maxima; full = 0
// leaves out 1/rho and the pressure perturbation, which only the per-cell
// update reads
template <int D, typename T>
__device__ __forceinline__ void prep_st(const Geo &g, const Phys &ph, const St<D, T> &s,
                                        T *__restrict__ aux, ci_t c, double *mx, int full = 1)
{
    typedef Aux<D> AX;
    const size_t nc = (size_t)g.nc;
    const double rho = rho_full(s);
    const double ir = rcp64(rho);
    if (full) {
        st(aux + AX::IRHO * nc, c, (T)ir);
        st(aux + AX::PT * nc, c, (T)ptot_pert<D, T>(s, rho, ph.gamma));
    }
    st(aux + AX::EINT * nc, c, (T)e_int<D, T>(s, rho));
#pragma unroll
    for (int d = 0; d < D; d++) st(aux + (AX::V0 + d) * nc, c, (T)((double)s.m[d] * ir));
#pragma unroll
    for (int d = 0; d < D; d++) {
        double v = cfast<D, T>(s, d, ph.gamma);
        mx[d] = v > mx[d] ? v : mx[d];
    }
}

template <int D, typename T>
__device__ __forceinline__ void prep_cell(const Geo &g, const Phys &ph, const T *__restrict__ u,
                                          const double *__restrict__ bg, T *__restrict__ aux,
                                          ci_t c, double *mx)
{
    St<D, T> s;
    load_pert32<D, T>(s, u, (size_t)g.nc, c);
    load_bg32<D, T>(s, bg, (size_t)g.bgf, c & g.bgm);
    prep_st<D, T>(g, ph, s, aux, c, mx);
}

// block-wide max of mx[] (blockDim.x == 256) folded into cm[] with a 64-bit
// atomic max; non-negative doubles order like their bit patterns
template <int D>
__device__ __forceinline__ void block_max_to(double *mx, double (*red)[4], unsigned long long *cm)
{
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < D; d++) {
        double v = wave_max(mx[d]);
        if (lane == 0) red[d][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < D) {
        int d = threadIdx.x;
        double v = red[d][0];
        for (int k = 1; k < 4; k++) v = red[d][k] > v ? red[d][k] : v;
        atomicMax(&cm[d], (unsigned long long)__double_as_longlong(v));
    }
}

// block-wide max of mx[] (blockDim.x == 64 NW) written to part[0..D-1]:
// the single-grid passes leave one partial maximum per block, and
// k_cmax_reduce folds them.  (One 64-bit atomic max per block on the same
// few addresses serialises: thousands of short blocks then spend more time
// in the atomics than in their cells.)
template <int D, int NW = 4>
__device__ __forceinline__ void block_max_part(double *mx, double (*red)[NW], double *part)
{
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < D; d++) {
        double v = wave_max(mx[d]);
        if (lane == 0) red[d][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < D) {
        int d = threadIdx.x;
        double v = red[d][0];
        for (int k = 1; k < NW; k++) v = red[d][k] > v ? red[d][k] : v;
        part[d] = v;
    }
}

// cm[d] = max over the nb partial maxima part[4 b + d] (one block of 1024
// threads), as the bits of a non-negative double
template <int D>
__global__ __launch_bounds__(1024) void k_cmax_reduce(const double *__restrict__ part, int nb,
                                                      unsigned long long *cm)
{
    __shared__ double red[D][16];
    double mx[D];
#pragma unroll
    for (int d = 0; d < D; d++) mx[d] = 0.0;
    for (int b = threadIdx.x; b < nb; b += 1024)
#pragma unroll
        for (int d = 0; d < D; d++) {
            const double v = part[4 * b + d];
            mx[d] = v > mx[d] ? v : mx[d];
        }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const double v = wave_max(mx[d]);
        if (lane == 0) red[d][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < D) {
        const int d = threadIdx.x;
        double v = red[d][0];
        for (int k = 1; k < 16; k++) v = red[d][k] > v ? red[d][k] : v;
        cm[d] = (unsigned long long)__double_as_longlong(v);
    }
}

// Courant limit: max over every cell and dim of cfast_d / dx_d at the cell
// (cucomputemaxcourant, computedervfields.cu:2118-2296), the block's
// maximum into part[4 blockIdx.x]; blockDim.x == 256, grid-stride
template <int D, typename T>
__global__ __launch_bounds__(256) void k_courant(Geo g, Phys ph, const T *__restrict__ u,
                                                  const double *__restrict__ bg, double *part)
{
    __shared__ double red[1][4];
    double mx[1] = {0.0};
    const ci_t nc = (ci_t)g.nc, s1 = (ci_t)g.s1, s2 = (ci_t)g.s2;
    for (ci_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        St<D, T> s;
        load_pert32<D, T>(s, u, (size_t)g.nc, c);
        load_bg32<D, T>(s, bg, (size_t)g.bgf, c & g.bgm);
        const int x[3] = {(int)(c % s1), (int)((c / s1) % (ci_t)g.n[1]), (int)(c / s2)};
#pragma unroll
        for (int d = 0; d < D; d++) {
            const double v = cfast<D, T>(s, d, ph.gamma) * g.idx[d][x[d]];
            mx[0] = v > mx[0] ? v : mx[0];
        }
    }
    block_max_part<1>(mx, red, part + 4 * blockIdx.x);
}

// blockDim.x == 256; grid-stride over every cell (ghosts included); the
// block's cfast maxima go to part[4 blockIdx.x + d]
template <int D, typename T>
__global__ __launch_bounds__(256) void k_prep(Geo g, Phys ph, const T *__restrict__ u,
                                               const double *__restrict__ bg,
                                               T *__restrict__ aux, double *part)
{
    __shared__ double red[D][4];
    double mx[D];
#pragma unroll
    for (int d = 0; d < D; d++) mx[d] = 0.0;
    const ci_t nc = (ci_t)g.nc;
    for (ci_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
        prep_cell<D, T>(g, ph, u, bg, aux, c, mx);
    block_max_part<D>(mx, red, part + 4 * blockIdx.x);
}

// ---------------------------------------------------------------------------
// hyperviscosity ratio.  For pair (q, dim) at cell x (along dim) with the
// pads q[-1] = q[5], q[n] = q[n-6]:
//   R(x) = max(D3(x-1), D3(x), D3(x+1)) / max(D1(x-2), .., D1(x+2)), 0 if 0/0
//   D3(f) = |3(q[f+1]-q[f]) - (q[f+2]-q[f-1])|, D1(f) = |q[f+1]-q[f]|
// nu_R(x) = dx cmax chyp R(x), nu_L(x) = dx cmax chyp R(x-1)
// (hyperdifvisc1r.cu:426-1300, hyperdifvisc1l.cu:180-1450).
// Computed on the box [1, n-3] in every dim; R is stored as T.
// ---------------------------------------------------------------------------
// R from the two maxima m1 = max D1, m3 = max D3 of a window
template <typename T>
__device__ __forceinline__ T ratio_fin(T m1, T m3)
{
    if constexpr (sizeof(T) == 4) {
        // FP32: the hardware reciprocal (1 ulp) and one multiply, within
        // 2 ulp of m3/m1.  The denominator is clamped to the smallest normal
        // number, so the reciprocal's argument and result stay normal: where
        // m1 is 0 every difference is 0, so m3 is too and the ratio is 0, as
        // the reference's guard gives; a denormal m1 (differences below
        // 1.2e-38, too small for any term of L(w) to carry) gives
        // m3 / 2^-126, below the exact ratio.  m1 above 2^126 does not arise
        // in data that the FP32 flux arithmetic can carry.
        return m3 * __builtin_amdgcn_rcpf(fmax(m1, 0x1p-126f));
    } else {
        if (!(m1 > (T)0)) return (T)0;
        return m3 / m1;
    }
}

// D3 at the centre difference d = q[f+1]-q[f] of a window: formed with one
// fused multiply-add from the two first differences, so it carries a
// single rounding even where it is a small remainder of nearly equal terms
// (smooth data); in FP32 the differences of the FP32 state are then exact
// to one rounding, as in FP64.  qa = q[f+2], qb = q[f-1].
template <typename T>
__device__ __forceinline__ T d3abs(T d, T qa, T qb)
{
    return fabs(fma((T)3, d, -(qa - qb)));
}

// R of one window
template <typename T>
__device__ __forceinline__ T ratio6(const T q[6])
{
    T d[5];
#pragma unroll
    for (int f = 0; f < 5; f++) d[f] = q[f + 1] - q[f];
    // running maxima of non-negative values: fmax(v, m) equals
    // (v > m ? v : m) here, a NaN v included (both keep m)
    T m1 = 0, m3 = 0;
#pragma unroll
    for (int f = 0; f < 5; f++) m1 = fmax(fabs(d[f]), m1);
#pragma unroll
    for (int f = 1; f < 4; f++) m3 = fmax(d3abs(d[f], q[f + 2], q[f - 1]), m3);
    return ratio_fin(m1, m3);
}

// index of window position x along an axis of n cells: the pads q[-1] =
// q[5] and q[n] = q[n-6] of the ratio windows
__device__ __forceinline__ int pad_map(int x, int n) { return x < 0 ? 5 : (x >= n ? n - 6 : x); }

// R of the L windows q[c..c+5] (c = 0..L-1) of the L + 5 values q[x st],
// sharing their differences and partial maxima; each equals ratio6 of its
// window.  The maxima are over the same values as ratio6's, in another
// order: fmax ignores a NaN operand and the values are non-negative, so the
// order does not change the result.  The D3 maxima keep ratio6's starting
// 0 (as fmax(t, 0) per difference); for the D1 maxima ratio_fin's clamp
// (FP32) or its test m1 > 0 (FP64) gives what the starting 0 does.
template <int L, typename T>
__device__ __forceinline__ void ratio_run(const T *q, int st, T r[L])
{
    T v[L + 5], d[L + 4], t[L + 3], a[L + 2];
#pragma unroll
    for (int x = 0; x < L + 5; x++) v[x] = q[x * st];
#pragma unroll
    for (int x = 0; x < L + 4; x++) d[x] = v[x + 1] - v[x];
    // D3 at the centre differences 1..L+2
#pragma unroll
    for (int x = 1; x < L + 3; x++) t[x] = fmax(d3abs(d[x], v[x + 2], v[x - 1]), (T)0);
    // max |d[x..x+2]|
#pragma unroll
    for (int x = 0; x < L + 2; x++) a[x] = fmax(fmax(fabs(d[x]), fabs(d[x + 1])), fabs(d[x + 2]));
#pragma unroll
    for (int c = 0; c < L; c++) {
        const T m1 = fmax(fmax(a[c], fabs(d[c + 3])), fabs(d[c + 4]));
        const T m3 = fmax(fmax(t[c + 1], t[c + 2]), t[c + 3]);
        r[c] = ratio_fin(m1, m3);
    }
}

// r[c] to o[c st] for c <= lim, 0 beyond
template <int L, typename T>
__device__ __forceinline__ void ratio_store(T *o, int st, const T r[L], int lim)
{
    if (lim >= L - 1) {
#pragma unroll
        for (int c = 0; c < L; c++) o[c * st] = r[c];
    } else {
#pragma unroll
        for (int c = 0; c < L; c++) o[c * st] = c <= lim ? r[c] : (T)0;
    }
}

template <int D, typename T>
__device__ __forceinline__ void ratio_cell(const Geo &g, const T *__restrict__ u,
                                           const T *__restrict__ aux, T *__restrict__ R,
                                           const int x0, const int x1, const int x2)
{
    typedef Lay<D> LY;
    typedef Aux<D> AX;
    int xs[3] = {x0, x1, x2};
    const size_t nc = (size_t)g.nc;
    const ci_t c = (ci_t)x0 + (ci_t)g.s1 * (ci_t)x1 + (ci_t)g.s2 * (ci_t)x2;
#pragma unroll
    for (int d = 0; d < D; d++) {
        int x = xs[d], n = g.n[d];
        const ci_t sd = stride32<D>(g, d);
        const ci_t c0 = c - (ci_t)x * sd;
        T qr[6], qe[6], qv[D][6], qb[D][6];
#pragma unroll
        for (int p = 0; p < 6; p++) {
            int xx = x - 2 + p;
            if (xx < 0) xx = 5;
            if (xx >= n) xx = n - 6;
            const ci_t cp = c0 + (ci_t)xx * sd;
            qr[p] = ld(u, cp);
            qe[p] = ld(aux + AX::EINT * nc, cp);
#pragma unroll
            for (int f = 0; f < D; f++) {
                qv[f][p] = ld(aux + (AX::V0 + f) * nc, cp);
                qb[f][p] = ld(u + (LY::B0 + f) * nc, cp);
            }
        }
        st(R + LY::prho(d) * nc, c, ratio6<T>(qr));
        st(R + LY::pe(d) * nc, c, ratio6<T>(qe));
#pragma unroll
        for (int f = 0; f < D; f++) st(R + LY::pv(f, d) * nc, c, ratio6<T>(qv[f]));
#pragma unroll
        for (int f = 0; f < D; f++)
            if (f != d) st(R + LY::pb(f, d) * nc, c, ratio6<T>(qb[f]));
    }
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_ratio(Geo g, TileMap tm, const T *__restrict__ u,
                                                const T *__restrict__ aux, T *__restrict__ R)
{
    int tx, ty, tk;
    if (!map_block(tm, tx, ty, tk)) return;
    int x0 = 1 + tx * 64 + (threadIdx.x & 63);
    int x1 = 1 + ty * 4 + (threadIdx.x >> 6);
    int x2 = (D == 3) ? 1 + tk : 0;
    if (x0 > g.n[0] - 3 || x1 > g.n[1] - 3) return;
    ratio_cell<D, T>(g, u, aux, R, x0, x1, x2);
}

// ---------------------------------------------------------------------------
// ghost fill: every ghost cell takes the value at the image of each of its
// coordinates (contcd4: 0<-4, 1<-3, n-1<-n-5, n-2<-n-4; periodic: 0,1 <-
// n-4,n-3, n-2,n-1 <- 2,3), which equals filling dim 0, 1, 2 in turn over
// the full transverse range as the driver does (iosmaug.cpp:1157-1166).
// One thread per cell of the three ghost slabs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bc_map(int x, int n, int per)
{
    if (x >= 2 && x <= n - 3) return x;
    if (per) return (x < 2) ? n - 4 + x : 4 - n + x;
    if (x < 2) return 4 - x;
    return (x == n - 1) ? x - 4 : x - 2;
}

// The aux/cmax pass for a 3-D background that does not vary along k, stored
// as one plane (bgp, field stride s2).  Thread = one column (i, j) of the
// plane, block row y = a chunk of kc planes: the column's background is
// loaded once and only the evolved fields are streamed.  full as in
// prep_st.  fill = 1 also fills the ghost cells of u first: a ghost cell
// takes the state of its image (bc_map in every dim, as k_bc does) and its
// aux and cfast are formed with its own background.  Only ghost cells are
// written and the images are interior cells, so no thread reads a cell that
// another writes.  The block's cfast maxima go to part[4 b + d], b = the
// linear block index.
template <typename T>
__global__ __launch_bounds__(256) void k_prep_col(Geo g, Phys ph, T *__restrict__ u,
                                                  const double *__restrict__ bgp,
                                                  T *__restrict__ aux, double *part,
                                                  int kc, int full, int fill)
{
    __shared__ double red[3][4];
    double mx[3] = {0.0, 0.0, 0.0};
    const ci_t s2 = (ci_t)g.s2, s1 = (ci_t)g.s1;
    const size_t nc = (size_t)g.nc;
    const ci_t p = blockIdx.x * 256 + threadIdx.x;
    const int ka = blockIdx.y * kc;
    const int kb = min(ka + kc, g.n[2]);
    if (p < s2) {
        St<3, T> s;
        load_bg32<3, T>(s, bgp, (size_t)s2, p);
        const int i = (int)(p % s1), j = (int)(p / s1);
        const ci_t pimg = (ci_t)bc_map(i, g.n[0], g.per[0]) + s1 * (ci_t)bc_map(j, g.n[1], g.per[1]);
        for (int k = ka; k < kb; k++) {
            const ci_t c = p + s2 * (ci_t)k;
            const int yk = bc_map(k, g.n[2], g.per[2]);
            // one load site and one arithmetic path for ghost and interior
            // cells, so that both get the same instruction sequence
            const bool gh = fill && (pimg != p || yk != k);
            load_pert32<3, T>(s, u, nc, gh ? pimg + s2 * (ci_t)yk : c);
            if (gh) {
                st(u, c, s.r);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    st(u + (1 + d) * nc, c, s.m[d]);
                    st(u + (5 + d) * nc, c, s.b[d]);
                }
                st(u + 4 * nc, c, s.e);
            }
            prep_st<3, T>(g, ph, s, aux, c, mx, full);
        }
    }
    block_max_part<3>(mx, red, part + 4 * (blockIdx.x + gridDim.x * blockIdx.y));
}

// Ghost cells of a 3-D grid, each once: the four k-ghost planes in full,
// then per interior plane its ring (rows j = 0, 1, n1-2, n1-1 in full, then
// the cells i = 0, 1, n0-2, n0-1 of the rows between)
__host__ __device__ __forceinline__ idx_t ghost_count(int n0, int n1, int n2)
{
    return (idx_t)4 * n0 * n1 + (idx_t)(n2 - 4) * (4 * n0 + 4 * (n1 - 4));
}

__device__ __forceinline__ void ghost_cell(idx_t t, int n0, int n1, int n2, int &i, int &j, int &k)
{
    const idx_t pl = (idx_t)n0 * n1;
    if (t < 4 * pl) {
        const int q = (int)(t / pl);
        const int p = (int)(t - q * pl);
        k = q < 2 ? q : n2 - 4 + q;
        j = p / n0;
        i = p - j * n0;
        return;
    }
    t -= 4 * pl;
    const int ring = 4 * n0 + 4 * (n1 - 4);
    const int kk = (int)(t / ring);
    int r = (int)(t - (idx_t)kk * ring);
    k = 2 + kk;
    if (r < 4 * n0) {
        const int q = r / n0;
        j = q < 2 ? q : n1 - 4 + q;
        i = r - q * n0;
    } else {
        r -= 4 * n0;
        j = 2 + (r >> 2);
        const int q = r & 3;
        i = q < 2 ? q : n0 - 4 + q;
    }
}

// The ghost-cell share of k_prep_col with fill = 1, for a state whose
// interior aux and cfast maxima the stage update that wrote it has formed
// already: every ghost cell takes the state of its image, and its aux and
// cfast are formed with its own background (the k-plane bgp) by the same
// load site and arithmetic.  The block's maxima go to part[4 b + d].
template <typename T>
__global__ __launch_bounds__(256) void k_ghost_prep(Geo g, Phys ph, T *__restrict__ u,
                                                    const double *__restrict__ bgp,
                                                    T *__restrict__ aux, double *part, int full)
{
    __shared__ double red[3][4];
    double mx[3] = {0.0, 0.0, 0.0};
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const idx_t t = (idx_t)blockIdx.x * 256 + threadIdx.x;
    if (t < ghost_count(n0, n1, n2)) {
        const ci_t s2 = (ci_t)g.s2, s1 = (ci_t)g.s1;
        const size_t nc = (size_t)g.nc;
        int i, j, k;
        ghost_cell(t, n0, n1, n2, i, j, k);
        const ci_t p = (ci_t)i + s1 * (ci_t)j;
        const ci_t c = p + s2 * (ci_t)k;
        const ci_t img = (ci_t)bc_map(i, n0, g.per[0]) + s1 * (ci_t)bc_map(j, n1, g.per[1]) +
                         s2 * (ci_t)bc_map(k, n2, g.per[2]);
        St<3, T> s;
        load_bg32<3, T>(s, bgp, (size_t)s2, p);
        load_pert32<3, T>(s, u, nc, img);
        st(u, c, s.r);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            st(u + (1 + d) * nc, c, s.m[d]);
            st(u + (5 + d) * nc, c, s.b[d]);
        }
        st(u + 4 * nc, c, s.e);
        prep_st<3, T>(g, ph, s, aux, c, mx, full);
    }
    block_max_part<3>(mx, red, part + 4 * blockIdx.x);
}

// Ghost cells of a 2-D grid, each once: rows j = 0, 1, n1-2, n1-1 in full,
// then the cells i = 0, 1, n0-2, n0-1 of the rows between
__host__ __device__ __forceinline__ idx_t ghost_count2(int n0, int n1)
{
    return (idx_t)4 * n0 + (idx_t)4 * (n1 - 4);
}

// The ghost-cell share of the aux pass of a 2-D state whose interior aux and
// cfast maxima the stage update that wrote it has formed (k_update with
// FOLD): every ghost cell takes the state of its image, and its aux (all of
// them) and cfast are formed with its own background, by the same load and
// arithmetic as the aux pass.  The block's maxima go to part[4 b + d].
template <typename T>
__global__ __launch_bounds__(256) void k_ghost_prep2(Geo g, Phys ph, T *__restrict__ u,
                                                     const double *__restrict__ bg,
                                                     T *__restrict__ aux, double *part)
{
    typedef Lay<2> LY;
    __shared__ double red[3][4];
    double mx[3] = {0.0, 0.0, 0.0};
    const int n0 = g.n[0], n1 = g.n[1];
    const int t = (int)(blockIdx.x * 256 + threadIdx.x);
    if (t < (int)ghost_count2(n0, n1)) {
        int i, j;
        if (t < 4 * n0) {
            const int q = t / n0;
            j = q < 2 ? q : n1 - 4 + q;
            i = t - q * n0;
        } else {
            const int r = t - 4 * n0;
            j = 2 + (r >> 2);
            const int q = r & 3;
            i = q < 2 ? q : n0 - 4 + q;
        }
        const ci_t s1 = (ci_t)g.s1;
        const size_t nc = (size_t)g.nc;
        const ci_t c = (ci_t)i + s1 * (ci_t)j;
        const ci_t img = (ci_t)bc_map(i, n0, g.per[0]) + s1 * (ci_t)bc_map(j, n1, g.per[1]);
        St<2, T> s;
        load_pert32<2, T>(s, u, nc, img);
        load_bg32<2, T>(s, bg, (size_t)g.bgf, c & g.bgm);
        st(u + LY::RHO * nc, c, s.r);
        st(u + LY::EN * nc, c, s.e);
#pragma unroll
        for (int d = 0; d < 2; d++) {
            st(u + (LY::M0 + d) * nc, c, s.m[d]);
            st(u + (LY::B0 + d) * nc, c, s.b[d]);
        }
        prep_st<2, T>(g, ph, s, aux, c, mx, 1);
    }
    block_max_part<2>(mx, red, part + 4 * blockIdx.x);
}

// number of ghost-slab elements of a grid (4 layers per slab, slabs may overlap)
template <int D>
__host__ __device__ __forceinline__ idx_t bc_count(int n0, int n1, int n2)
{
    return (idx_t)4 * n1 * n2 + (idx_t)4 * n0 * n2 + (D == 3 ? (idx_t)4 * n0 * n1 : 0);
}

// fill ghost-slab element t
template <int D, typename T>
__device__ __forceinline__ void bc_elem(const Geo &g, T *__restrict__ u, idx_t t)
{
    typedef Lay<D> LY;
    int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    idx_t a0 = (idx_t)4 * n1 * n2, a1 = (idx_t)4 * n0 * n2, a2 = (D == 3) ? (idx_t)4 * n0 * n1 : 0;
    int x[3];
    if (t < a0) {
        int r = (int)(t % 4);
        idx_t q = t / 4;
        x[0] = r < 2 ? r : n0 - 4 + r;
        x[1] = (int)(q % n1);
        x[2] = (int)(q / n1);
    } else if (t < a0 + a1) {
        t -= a0;
        int r = (int)(t % 4);
        idx_t q = t / 4;
        x[1] = r < 2 ? r : n1 - 4 + r;
        x[0] = (int)(q % n0);
        x[2] = (int)(q / n0);
    } else if (t < a0 + a1 + a2) {
        t -= a0 + a1;
        int r = (int)(t % 4);
        idx_t q = t / 4;
        x[2] = r < 2 ? r : n2 - 4 + r;
        x[0] = (int)(q % n0);
        x[1] = (int)(q / n0);
    } else {
        return;
    }
    int y0 = bc_map(x[0], n0, g.per[0]);
    int y1 = bc_map(x[1], n1, g.per[1]);
    int y2 = (D == 3) ? bc_map(x[2], n2, g.per[2]) : 0;
    idx_t dst = (idx_t)x[0] + g.s1 * x[1] + g.s2 * x[2];
    idx_t src = (idx_t)y0 + g.s1 * y1 + g.s2 * y2;
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) u[(idx_t)f * g.nc + dst] = u[(idx_t)f * g.nc + src];
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_bc(Geo g, T *__restrict__ u)
{
    bc_elem<D, T>(g, u, (idx_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// conservation diagnostics: compensated FP64 sums of every evolved field over
// the interior cells.  Each thread accumulates its cells with Neumaier's
// compensated summation; the block combines the per-thread (sum,
// compensation) pairs with error-free two-sums and writes one pair per field;
// the host adds the per-block pairs, again compensated.
// ---------------------------------------------------------------------------
struct CSum {
    double s, c;
};

__device__ __forceinline__ void csum_add(CSum &a, double x)
{
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x))
        a.c += (a.s - t) + x;
    else
        a.c += (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ void csum_merge(CSum &a, const CSum &b)
{
    csum_add(a, b.s);
    a.c += b.c;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_sums(Geo g, const T *__restrict__ u, double *part)
{
    typedef Lay<D> LY;
    __shared__ double red[2][LY::NVE][256];
    const int e0 = g.n[0] - 4, e1 = g.n[1] - 4, e2 = (D == 3) ? g.n[2] - 4 : 1;
    const long long ni = (long long)e0 * e1 * e2;
    const size_t nc = (size_t)g.nc;
    CSum acc[LY::NVE];
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) acc[f].s = acc[f].c = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ni;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % e0);
        const long long r = t / e0;
        const int j = (int)(r % e1), k = (int)(r / e1);
        const ci_t c = (ci_t)(2 + i) + (ci_t)g.s1 * (ci_t)(2 + j) +
                       (D == 3 ? (ci_t)g.s2 * (ci_t)(2 + k) : 0u);
#pragma unroll
        for (int f = 0; f < LY::NVE; f++) csum_add(acc[f], (double)ld(u + f * nc, c));
    }
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) {
        red[0][f][threadIdx.x] = acc[f].s;
        red[1][f][threadIdx.x] = acc[f].c;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int f = 0; f < LY::NVE; f++) {
                CSum a = {red[0][f][threadIdx.x], red[1][f][threadIdx.x]};
                const CSum b = {red[0][f][threadIdx.x + w], red[1][f][threadIdx.x + w]};
                csum_merge(a, b);
                red[0][f][threadIdx.x] = a.s;
                red[1][f][threadIdx.x] = a.c;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < LY::NVE) {
        part[(size_t)blockIdx.x * 2 * LY::NVE + 2 * threadIdx.x] = red[0][threadIdx.x][0];
        part[(size_t)blockIdx.x * 2 * LY::NVE + 2 * threadIdx.x + 1] = red[1][threadIdx.x][0];
    }
}

// ---------------------------------------------------------------------------
// stage update
// ---------------------------------------------------------------------------
struct StageArgs {
    int mode;      // 0: RK stage, 1: forward Euler, 2: write raw L(w)
    int stage;     // RK stage 0..3
    double ct;     // stage time offset in units of dt (0, 1/2, 1/2, 1)
    const double *tm; // device [t, dt]
};

// Global-memory view of the hyperdiffusion inputs around cell c: field f of
// the state, aux quantity a, rho_b and ratio pair p at c + (di, dj, dk).
// UNI: a background that is the same in every cell (Geo::bgm == 0), whose
// rho_b is read once into rb0
template <int D, typename T, bool UNI = false>
struct HyGlobal {
    const T *__restrict__ uin;
    const T *__restrict__ aux;
    const double *__restrict__ bg;
    const T *__restrict__ R;
    size_t nc;
    size_t bgf; // background field stride and cell mask (Geo::bgf, bgm)
    ci_t bgm;
    double rb0;
    ci_t c, s1, s2;
    __device__ __forceinline__ ci_t at(int di, int dj, int dk) const
    {
        return c + (ci_t)di + s1 * (ci_t)dj + s2 * (ci_t)dk;
    }
    __device__ __forceinline__ T U(int f, int di, int dj, int dk) const
    {
        return ld(uin + f * nc, at(di, dj, dk));
    }
    __device__ __forceinline__ T A(int a, int di, int dj, int dk) const
    {
        return ld(aux + a * nc, at(di, dj, dk));
    }
    __device__ __forceinline__ double rhob(int di, int dj, int dk) const
    {
        if constexpr (UNI)
            return rb0;
        else
            return ld(bg + bgf, at(di, dj, dk) & bgm);
    }
    __device__ __forceinline__ T Rp(int p, int di, int dj, int dk) const
    {
        return ld(R + p * nc, at(di, dj, dk));
    }
};

// unit displacement along dim d, scaled by s
template <int D>
__device__ __forceinline__ void unit(int d, int s, int &di, int &dj, int &dk)
{
    di = d == 0 ? s : 0;
    dj = d == 1 ? s : 0;
    dk = d == 2 ? s : 0;
}

// the hyperdiffusive sources of interior cell (x0, x1, x2) added to dU
// (iosmaug.cpp:969-1146); inputs through the view V (HyGlobal or an LDS
// staged view with the same members)
template <int D, typename T, typename V>
__device__ __forceinline__ void hyper_terms(const Geo &g, const Phys &ph, const V &vw,
                                            const unsigned long long *__restrict__ cmb,
                                            const int x0, const int x1, const int x2, T *dU)
{
    typedef Lay<D> LY;
    typedef Aux<D> AX;
    const int xs[3] = {x0, x1, x2};
    const T rc = vw.U(LY::RHO, 0, 0, 0);
    const double rbc = vw.rhob(0, 0, 0);
    double cmax[D];
#pragma unroll
    for (int d = 0; d < D; d++) cmax[d] = __longlong_as_double((long long)cmb[d]);
    const T eic = vw.A(AX::EINT, 0, 0, 0);
    // full density rho' + rho_b of a cell, formed in double and rounded to T
    auto F = [&](int di, int dj, int dk) -> T {
        return (T)((double)vw.U(LY::RHO, di, dj, dk) + vw.rhob(di, dj, dk));
    };
    const T fc = (T)((double)rc + rbc);

#pragma unroll
    for (int d = 0; d < D; d++) {
        int ri, rj, rk, li, lj, lk;
        unit<D>(d, 1, ri, rj, rk);
        unit<D>(d, -1, li, lj, lk);
        const int xd = xs[d], nd = g.n[d];
        const double dcd = g.dx[d][xd] * cmax[d];
        const T idxd = (T)g.idx[d][xd];
        // 1/dx of the gradients, 0 where the reference zeroes them (zero1);
        // every term below is (nu/R) (R_R g_R - R_L g_L) times this and the
        // 1/dx of the divergence, with nu/R = dx cmax chyp rounded to T
        const T gxd = zero1(g.per[d], xd, nd) ? (T)0 : idxd;
        const T sxd = gxd * idxd, hxd = gxd / (T)2;
        auto C = [&](int field) -> T { return (T)(dcd * ph.chyp[field]) * sxd; };
        T pR, pL;
        // density (hyperdifrhosource1.cu:15-120)
        pR = vw.Rp(LY::prho(d), 0, 0, 0) * (vw.U(LY::RHO, ri, rj, rk) - rc);
        pL = vw.Rp(LY::prho(d), li, lj, lk) * (rc - vw.U(LY::RHO, li, lj, lk));
        dU[LY::RHO] += (pR - pL) * C(LY::RHO);
        // energy (hyperdifesource1.cu:18-330)
        pR = vw.Rp(LY::pe(d), 0, 0, 0) * (vw.A(AX::EINT, ri, rj, rk) - eic);
        pL = vw.Rp(LY::pe(d), li, lj, lk) * (eic - vw.A(AX::EINT, li, lj, lk));
        dU[LY::EN] += (pR - pL) * C(LY::EN);
        // momentum, diagonal part for every f; twice when f == d
        // (hyperdifmomsource1.cu:17-533 called for ii1 = 0 and, if f == d, ii1 = 1);
        // rR, rL, mR, mL, bR, bL are twice the face means, the 1/2 is in the factor
        const T rR = fc + F(ri, rj, rk);
        const T rL = fc + F(li, lj, lk);
#pragma unroll
        for (int f = 0; f < D; f++) {
            const T vc = vw.A(AX::V0 + f, 0, 0, 0);
            pR = vw.Rp(LY::pv(f, d), 0, 0, 0) * (vw.A(AX::V0 + f, ri, rj, rk) - vc);
            pL = vw.Rp(LY::pv(f, d), li, lj, lk) * (vc - vw.A(AX::V0 + f, li, lj, lk));
            const T mc = vw.U(LY::M0 + f, 0, 0, 0);
            const T mR = mc + vw.U(LY::M0 + f, ri, rj, rk);
            const T mL = mc + vw.U(LY::M0 + f, li, lj, lk);
            const T ck = C(LY::M0 + f) * ((f == d) ? (T)0.5 : (T)0.25);
            dU[LY::M0 + f] += (rR * pR - rL * pL) * ck;
            dU[LY::EN] += (mR * pR - mL * pL) * ck;
        }
        // magnetic field, diagonal part for f != d (hyperdifbsource1.cu:21-440)
#pragma unroll
        for (int f = 0; f < D; f++) {
            if (f == d) continue;
            const T bc_ = vw.U(LY::B0 + f, 0, 0, 0);
            const T br = vw.U(LY::B0 + f, ri, rj, rk), bl = vw.U(LY::B0 + f, li, lj, lk);
            pR = vw.Rp(LY::pb(f, d), 0, 0, 0) * (br - bc_);
            pL = vw.Rp(LY::pb(f, d), li, lj, lk) * (bc_ - bl);
            const T cb = C(LY::B0 + f);
            dU[LY::B0 + f] += (pR - pL) * cb;
            dU[LY::EN] += ((bc_ + br) * pR - (bc_ + bl) * pL) * (cb * (T)0.5);
        }

        // off-diagonal terms for f != d: d/dx_f of products evaluated at
        // c -/+ e_f (hyperdifmomsourcene1.cu:16-740, hyperdifbsourcene1.cu:18-820)
#pragma unroll
        for (int f = 0; f < D; f++) {
            if (f == d) continue;
            const int xf = xs[f], nf = g.n[f];
            if (zero1(g.per[f], xf, nf)) continue;
            T Xr[2], Xm[2], Hb[2], He[2];
#pragma unroll
            for (int side = 0; side < 2; side++) {
                const int sg = side ? 1 : -1;
                int ni, nj, nk;
                unit<D>(f, sg, ni, nj, nk); // cn = c + sg e_f
                Xr[side] = Xm[side] = Hb[side] = He[side] = 0;
                const int xfn = xf + sg;
                if (xfn < 2 || xfn > nf - 3) continue; // nu vanishes off the interior
                // ca = cn - e_d, cb = cn + e_d
                const int ai = ni + li, aj = nj + lj, ak = nk + lk;
                const int bi = ni + ri, bj = nj + rj, bk = nk + rk;
                const T dv = (vw.A(AX::V0 + f, bi, bj, bk) - vw.A(AX::V0 + f, ai, aj, ak)) * hxd;
                const T db = (vw.U(LY::B0 + f, bi, bj, bk) - vw.U(LY::B0 + f, ai, aj, ak)) * hxd;
                {
                    const T K = (T)(dcd * ph.chyp[LY::M0 + f]);
                    const T s2 = K * (vw.Rp(LY::pv(f, d), ni, nj, nk) +
                                      vw.Rp(LY::pv(f, d), ai, aj, ak));
                    const T G = dv * s2 / (T)4;
                    const T rhon = F(ni, nj, nk);
                    Xr[side] = rhon * G;
                    Xm[side] = G * vw.U(LY::M0 + d, ni, nj, nk);
                }
                {
                    const T K = (T)(dcd * ph.chyp[LY::B0 + f]);
                    const T s2 = K * (vw.Rp(LY::pb(f, d), ni, nj, nk) +
                                      vw.Rp(LY::pb(f, d), ai, aj, ak));
                    const T H = db * s2 / (T)2;
                    Hb[side] = H;
                    He[side] = H * vw.U(LY::B0 + d, ni, nj, nk);
                }
            }
            const T h = (T)g.idx[f][xf] / (T)2;
            dU[LY::M0 + d] += (Xr[1] - Xr[0]) * h;
            dU[LY::EN] += (Xm[1] - Xm[0]) * h;
            dU[LY::B0 + d] -= (Hb[1] - Hb[0]) * h;
            dU[LY::EN] -= (He[1] - He[0]) * h;
        }
    }
}

// the global-memory view of cell (x0, x1, x2)
template <int D, typename T, bool UNI>
__device__ __forceinline__ void hy_global_init(HyGlobal<D, T, UNI> &vw, const Geo &g,
                                               const T *__restrict__ uin,
                                               const T *__restrict__ R,
                                               const T *__restrict__ aux,
                                               const double *__restrict__ bg, const int x0,
                                               const int x1, const int x2)
{
    vw.rb0 = UNI ? ld(bg + (size_t)g.bgf, 0u) : 0.0;
    vw.uin = uin;
    vw.aux = aux;
    vw.bg = bg;
    vw.R = R;
    vw.nc = (size_t)g.nc;
    vw.bgf = (size_t)g.bgf;
    vw.bgm = g.bgm;
    vw.s1 = (ci_t)g.s1;
    vw.s2 = (ci_t)g.s2;
    vw.c = (ci_t)x0 + vw.s1 * (ci_t)x1 + vw.s2 * (ci_t)x2;
}

template <int D, typename T, bool UNI = false>
__device__ __forceinline__ void hyper_cell(const Geo &g, const Phys &ph, const T *__restrict__ uin,
                                           const T *__restrict__ R, const T *__restrict__ aux,
                                           const double *__restrict__ bg,
                                           const unsigned long long *__restrict__ cmb,
                                           const int x0, const int x1, const int x2, T *dU)
{
    HyGlobal<D, T, UNI> vw;
    hy_global_init<D, T, UNI>(vw, g, uin, R, aux, bg, x0, x1, x2);
    hyper_terms<D, T>(g, ph, vw, cmb, x0, x1, x2, dU);
}

// HyGlobal with the ratio pairs read from a box in LDS: pair p of the cell
// at (di, dj) from the box cell own is sR[p nb + own + di + bi dj] (2-D)
template <int D, typename T, bool UNI = false>
struct HyLdsR : HyGlobal<D, T, UNI> {
    const T *sR;
    int nb, own, bi;
    __device__ __forceinline__ T Rp(int p, int di, int dj, int dk) const
    {
        return sR[p * nb + own + di + bi * dj];
    }
};

// LR (update_cell): the hyperdiffusive sources with the ratio pairs from LDS
struct RBox {
    const void *sR; // T [NPAIR][nb]
    int nb, own, bi;
};

// the driver's time factor drv_0 sin(2 pi (t + ct dt) / drv_1) at the stage
// time (usersource.cu:4-92)
__device__ __forceinline__ double driver_amp(const Phys &ph, const StageArgs &sa, double t,
                                             double dt)
{
    const double ts = t + sa.ct * dt;
    return ph.drv[0] * sin(ts * 6.283185307179586 / ph.drv[1]);
}

// the driver's sources (vertical velocity v = pr: m_0 += rho v,
// e += rho v^2/2) of a cell with total density rc + rbc, added to dU
template <int D, typename T>
__device__ __forceinline__ void driver_add(double pr, const T rc, const double rbc, T *dU)
{
    typedef Lay<D> LY;
    const T v = (T)pr;
    const T rfull = (T)((double)rc + rbc);
    dU[LY::M0] += v * rfull;
    dU[LY::EN] += v * v * rfull / (T)2;
}

// the torsional driver's sources (horizontal velocities vx along dim 1,
// vy along dim 2: m_1 += rho vx, m_2 += rho vy, e += rho (vx^2 + vy^2)/2)
// of a cell with total density rc + rbc, added to dU
template <typename T>
__device__ __forceinline__ void driver_add_tor(double vx_, double vy_, const T rc,
                                               const double rbc, T *dU)
{
    typedef Lay<3> LY;
    const T vx = (T)vx_, vy = (T)vy_;
    const T rfull = (T)((double)rc + rbc);
    dU[LY::M0 + 1] += vx * rfull;
    dU[LY::M0 + 2] += vy * rfull;
    dU[LY::EN] += (vx * vx + vy * vy) * rfull / (T)2;
}

// the user driver at interior cell (x0, x1, x2) added to dU; its spatial
// factor is the product of the per-axis profiles (usersource.cu:4-92).
// Driver 2 is the reference's torsional one: vx = (A/L) sin(..) P0 P1 Q2,
// vy = -(A/L) sin(..) P0 Q1 P2 with P the Gaussian profiles and Q the lever
// profiles (x - xc) P (usersource.cu:26-73).
template <int D, typename T>
__device__ __forceinline__ void driver_cell(const Geo &g, const Phys &ph, const StageArgs &sa,
                                            const T rc, const double rbc, const int x0,
                                            const int x1, const int x2, T *dU)
{
    if (ph.driver == 1) {
        double pr = driver_amp(ph, sa, sa.tm[0], sa.tm[1]) * g.prof[0][x0] * g.prof[1][x1];
        if (D == 3) pr *= g.prof[2][x2];
        driver_add<D, T>(pr, rc, rbc, dU);
    } else if constexpr (D == 3) {
        if (ph.driver == 2) {
            const double a = driver_amp(ph, sa, sa.tm[0], sa.tm[1]) / ph.drv[7] * g.prof[0][x0];
            driver_add_tor<T>(a * g.prof[1][x1] * g.profq[2][x2],
                              -(a * g.profq[1][x1] * g.prof[2][x2]), rc, rbc, dU);
        }
    }
}

// the stage result of cell c from its L(w) = dU (mode 0: low-storage RK4
// slot algebra, 1: forward Euler, 2: raw L(w))
template <int D, typename T>
__device__ __forceinline__ void rk_write(const StageArgs &sa, const T *__restrict__ un,
                                         T *__restrict__ uout, T *__restrict__ acc, size_t nc,
                                         ci_t c, const T *dU, T *ov = nullptr)
{
    typedef Lay<D> LY;
    const double dt = sa.tm[1];
    if (sa.mode == 2) {
#pragma unroll
        for (int f = 0; f < LY::NVE; f++) st(uout + f * nc, c, dU[f]);
        return;
    }
    if (sa.mode == 1) {
        const T tdt = (T)dt;
#pragma unroll
        for (int f = 0; f < LY::NVE; f++)
            st(uout + f * nc, c, ld(un + f * nc, c) + tdt * dU[f]);
        return;
    }
    // low-storage classical RK4:
    //  s0: out = un + dt/2 k, acc = un + dt/6 k
    //  s1: out = un + dt/2 k, acc += dt/3 k
    //  s2: out = un + dt k,   acc += dt/3 k
    //  s3: out = acc + dt/6 k
    // dt/6 and dt/3 as products: no FP64 division per cell
    const int s = sa.stage;
    const T a = (T)(s == 2 ? dt : (s == 3 ? dt * (1.0 / 6.0) : dt * 0.5));
    const T b = (T)(s == 0 ? dt * (1.0 / 6.0) : dt * (1.0 / 3.0));
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) {
        const size_t o = f * nc;
        if (s == 3) {
            const T v = ld(acc + o, c) + a * dU[f];
            st(uout + o, c, v);
            if (ov) ov[f] = v;
        } else {
            const T base = ld(un + o, c);
            const T v = base + a * dU[f];
            st(uout + o, c, v);
            if (ov) ov[f] = v;
            st(acc + o, c, (s == 0 ? base : ld(acc + o, c)) + b * dU[f]);
        }
    }
}


// the whole stage update of interior cell (x0, x1, x2)
// UNI: the background is the same in every cell (Geo::bgm == 0); its values
// are read once instead of at every neighbour.  FOLD (RK stages 0-2): the
// aux quantities of the stage result (all of them, as the aux pass forms
// them) go to aux_out and its cfast into the running maxima mx, so that the
// next stage needs no aux/cmax pass over the interior.
template <int D, typename T, bool UNI = false, bool FOLD = false, bool LR = false>
__device__ __forceinline__ void update_cell(const Geo &g, const Phys &ph, const StageArgs &sa,
                                            const T *__restrict__ uin,
                                            const T *__restrict__ un,
                                            T *__restrict__ uout, T *__restrict__ acc,
                                            const T *__restrict__ R,
                                            const T *__restrict__ aux,
                                            const double *__restrict__ bg,
                                            const unsigned long long *__restrict__ cmb,
                                            const int x0, const int x1, const int x2,
                                            T *__restrict__ aux_out = nullptr,
                                            double *mx = nullptr, RBox rbox = RBox())
{
    typedef Lay<D> LY;
    typedef Aux<D> AX;
    const int xs[3] = {x0, x1, x2};
    const size_t nc = (size_t)g.nc;
    const ci_t c = (ci_t)x0 + (ci_t)g.s1 * (ci_t)x1 + (ci_t)g.s2 * (ci_t)x2;
    const double gam = ph.gamma;
    // field and aux accessors
    auto U = [&](int f, ci_t cc) -> T { return ld(uin + f * nc, cc); };
    auto A = [&](int a, ci_t cc) -> T { return ld(aux + a * nc, cc); };
    double bgv[2 + D];
#pragma unroll
    for (int f = 0; f < 2 + D; f++) bgv[f] = UNI ? ld(bg + f * (size_t)g.bgf, 0u) : 0.0;
    auto BG = [&](int f, ci_t cc) -> double {
        if constexpr (UNI)
            return bgv[f];
        else
            return ld(bg + f * (size_t)g.bgf, cc & g.bgm);
    };

    T dU[LY::NVE];
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) dU[f] = 0;

    const double rbc = BG(1, c);
    T bbc[D];
#pragma unroll
    for (int d = 0; d < D; d++) bbc[d] = (T)BG(2 + d, c);
    double bbb = 0;
#pragma unroll
    for (int d = 0; d < D; d++) bbb += BG(2 + d, c) * BG(2 + d, c);
    // background total pressure (computepbg3)
    const T ptb = (T)((gam - 1) * BG(0, c) - 0.5 * (gam - 2) * bbb);

    // ---- advection: 4th-order divergence of the fluxes along each axis a,
    // accumulated neighbour by neighbour, plus the background work terms of
    // the energy equation (centdiff1.cu:19-256, centdiff2.cu:20-356)
#pragma unroll
    for (int a = 0; a < D; a++) {
        const ci_t sa_ = stride32<D>(g, a);
        const int xa = xs[a], na = g.n[a];
        if (zero4(g.per[a], xa, na)) continue;
        // d/dx_a q = (q[-2] - 8 q[-1] + 8 q[+1] - q[+2]) / (12 dx)
        const T i12 = (T)(g.idx[a][xa] * (1.0 / 12.0)); // 1/(12 dx), without a division per cell
        T gv[D];
#pragma unroll
        for (int d = 0; d < D; d++) gv[d] = 0;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int off = (p < 2) ? p - 2 : p - 1; // -2,-1,+1,+2
            const T cw = (p == 0 ? (T)1 : (p == 1 ? (T)-8 : (p == 2 ? (T)8 : (T)-1))) * i12;
            const ci_t cn = c + (ci_t)off * sa_;
            T m[D], b[D], bb[D];
#pragma unroll
            for (int d = 0; d < D; d++) {
                m[d] = U(LY::M0 + d, cn);
                b[d] = U(LY::B0 + d, cn);
                bb[d] = (T)BG(2 + d, cn);
            }
            const T e = U(LY::EN, cn);
            const T eb = (T)BG(0, cn);
            const T ir = A(AX::IRHO, cn);
            const T pt = A(AX::PT, cn);
            const T ma = m[a];
            T bbm = 0, bm = 0;
#pragma unroll
            for (int t = 0; t < D; t++) {
                bbm += bb[t] * m[t];
                bm += b[t] * m[t];
            }
            // F_rho = m_a rho'/rho + rho_b m_a/rho = m_a
            dU[LY::RHO] -= cw * ma;
#pragma unroll
            for (int i = 0; i < D; i++) {
                const T mag = -(b[i] * bb[a] + bb[i] * b[a]) - b[i] * b[a];
                dU[LY::M0 + i] -= cw * (ma * m[i] * ir + ((i == a) ? pt + mag : mag));
            }
            dU[LY::EN] -= cw * ((ma * (e + eb + pt) - b[a] * bbm - bb[a] * bm - b[a] * bm) * ir);
#pragma unroll
            for (int i = 0; i < D; i++) {
                if (i == a) continue;
                dU[LY::B0 + i] -= cw * ((ma * b[i] - (b[a] + bb[a]) * m[i] + bb[i] * ma) * ir);
            }
#pragma unroll
            for (int d = 0; d < D; d++) gv[d] += cw * A(AX::V0 + d, cn);
        }
        // -p_Tb dv_a/dx_a + sum_d b_ba b_bd dv_d/dx_a  (addenergyterms_cd2)
        T w = -ptb * gv[a];
#pragma unroll
        for (int d = 0; d < D; d++) w += bbc[a] * bbc[d] * gv[d];
        dU[LY::EN] += w;
    }

    const T rc = U(LY::RHO, c);
    // ---- gravity (addgrav_cd2, centdiff2.cu:118-154)
    {
        T ge = 0;
#pragma unroll
        for (int d = 0; d < D; d++) {
            dU[LY::M0 + d] += (T)ph.g[d] * rc;
            ge += (T)ph.g[d] * U(LY::M0 + d, c);
        }
        dU[LY::EN] += rc * ge * A(AX::IRHO, c);
    }

    if (ph.hyperdif) {
        if constexpr (LR) {
            HyLdsR<D, T, UNI> vw;
            hy_global_init<D, T, UNI>(vw, g, uin, nullptr, aux, bg, x0, x1, x2);
            vw.sR = (const T *)rbox.sR;
            vw.nb = rbox.nb;
            vw.own = rbox.own;
            vw.bi = rbox.bi;
            hyper_terms<D, T>(g, ph, vw, cmb, x0, x1, x2, dU);
        } else {
            hyper_cell<D, T, UNI>(g, ph, uin, R, aux, bg, cmb, x0, x1, x2, dU);
        }
    }

    driver_cell<D, T>(g, ph, sa, rc, rbc, x0, x1, x2, dU);
    if constexpr (FOLD) {
        T ov[LY::NVE];
        rk_write<D, T>(sa, un, uout, acc, nc, c, dU, ov);
        St<D, T> so;
        so.r = ov[LY::RHO];
        so.e = ov[LY::EN];
#pragma unroll
        for (int d = 0; d < D; d++) {
            so.m[d] = ov[LY::M0 + d];
            so.b[d] = ov[LY::B0 + d];
            so.bb[d] = BG(2 + d, c);
        }
        so.rb = rbc;
        so.eb = BG(0, c);
        prep_st<D, T>(g, ph, so, aux_out, c, mx, 1);
    } else {
        rk_write<D, T>(sa, un, uout, acc, nc, c, dU);
    }
}

// FOLD: as update_cell; the block's cfast maxima go to part[4 b + d]
template <int D, typename T, bool UNI = false, bool FOLD = false>
__global__ __launch_bounds__(256) void k_update(Geo g, TileMap tm, Phys ph, StageArgs sa,
                                                 const T *__restrict__ uin,
                                                 const T *__restrict__ un,
                                                 T *__restrict__ uout, T *__restrict__ acc,
                                                 const T *__restrict__ R,
                                                 const T *__restrict__ aux,
                                                 const double *__restrict__ bg,
                                                 const unsigned long long *__restrict__ cmb,
                                                 T *__restrict__ aux_out = nullptr,
                                                 double *part = nullptr)
{
    int tx, ty, tk;
    if (!map_block(tm, tx, ty, tk)) {
        if (FOLD && threadIdx.x < D) part[4 * blockIdx.x + threadIdx.x] = 0.0;
        return;
    }
    const int x0 = 2 + tx * 64 + (threadIdx.x & 63);
    const int x1 = 2 + ty * 4 + (threadIdx.x >> 6);
    const int x2 = (D == 3) ? 2 + tk : 0;
    if constexpr (FOLD) {
        __shared__ double red[3][4];
        double mx[3] = {0.0, 0.0, 0.0};
        if (x0 <= g.n[0] - 3 && x1 <= g.n[1] - 3)
            update_cell<D, T, UNI, true>(g, ph, sa, uin, un, uout, acc, R, aux, bg, cmb, x0, x1,
                                         x2, aux_out, mx);
        block_max_part<D>(mx, red, part + 4 * blockIdx.x);
    } else {
        if (x0 > g.n[0] - 3 || x1 > g.n[1] - 3) return;
        update_cell<D, T, UNI>(g, ph, sa, uin, un, uout, acc, R, aux, bg, cmb, x0, x1, x2);
    }
}

// ---------------------------------------------------------------------------
// 2-D stage update with the hyperviscosity ratios formed in the kernel
// (ratio form 2 of 2-D grids).  A block stages rho', e_int, v_0, v_1, b'_0
// and b'_1 of its 64 x TJ tile and a 3-cell halo in LDS (pads mirrored as in
// ratio_cell), forms the ten ratio pairs of the tile and a 1-cell halo from
// them (0 outside the ratio box [1, n-3], where R, which k_ratio fills only
// there, holds the zeros it was cleared to), and runs update_cell with the
// pairs read from LDS: bit for bit k_ratio followed by k_update, without R
// in HBM.  The pairs along i are formed in runs of 11 cells of a row,
// those along j per column of the box.
// ---------------------------------------------------------------------------
// tiles of 64 x TJ cells, 64 TJ threads
template <int TJ>
struct U2Shape {
    static constexpr int TI = 64, NT = TI * TJ;
    static constexpr int RI = TI + 2, RJ = TJ + 2, RN = RI * RJ; // ratio box
    static constexpr int QI = TI + 7, QJ = TJ + 7, QN = QI * QJ; // window box
    static constexpr int RL = 11, NRUN = RI / RL;
    static_assert(RL * NRUN == RI, "the runs cover the rows of the ratio box");
};
// the tile height of k_update2r and ke_update2r: 8 rows in FP32 (77 VGPRs,
// 52 KB of LDS, six waves per SIMD; 4 rows: four), 4 in FP64 (8 rows would
// take 104 KB of LDS, one block per CU)
template <typename T>
__host__ __device__ constexpr int u2_tj()
{
    return sizeof(T) == 4 ? 8 : 4;
}
template <typename T>
using U2 = U2Shape<u2_tj<T>()>;

// pair a (0..4) of the pairs along d: rho', e_int, v_0, v_1, b'_f (f != d):
// its window quantity in the staged box and its ratio pair
__device__ __forceinline__ int u2_field(int d, int a) { return a < 4 ? a : 5 - d; }
__device__ __forceinline__ int u2_pair(int d, int a)
{
    typedef Lay<2> LY;
    return a == 0 ? LY::prho(d)
                  : (a == 1 ? LY::pe(d) : (a < 4 ? LY::pv(a - 2, d) : LY::pb(1 - d, d)));
}

// one 64 x TJ tile (tx, ty) of the interior: the window quantities staged,
// the ratio pairs formed, update_cell run (FOLD: the cfast maxima into mx);
// every thread of the block calls it (two block barriers)
template <typename T, bool UNI, bool FOLD, int TJ = u2_tj<T>()>
__device__ __forceinline__ void update2r_tile(const Geo &g, const Phys &ph, const StageArgs &sa,
                                              const T *__restrict__ uin,
                                              const T *__restrict__ un, T *__restrict__ uout,
                                              T *__restrict__ acc, const T *__restrict__ aux,
                                              const double *__restrict__ bg,
                                              const unsigned long long *__restrict__ cmb,
                                              const int tx, const int ty,
                                              T (*sQ)[U2Shape<TJ>::QN], T *sR,
                                              T *__restrict__ aux_out, double *mx)
{
    typedef Lay<2> LY;
    typedef Aux<2> AX;
    typedef U2Shape<TJ> SH;
    const int n0 = g.n[0], n1 = g.n[1];
    const size_t nc = (size_t)g.nc;
    const int i0 = 2 + tx * SH::TI, j0 = 2 + ty * TJ;
    if (ph.hyperdif) {
        // window box cell (qi, qj) = cell (i0 - 3 + qi, j0 - 3 + qj)
        for (int q = threadIdx.x; q < SH::QN; q += SH::NT) {
            const int qj = q / SH::QI, qi = q - qj * SH::QI;
            const ci_t c = (ci_t)pad_map(i0 - 3 + qi, n0) + (ci_t)g.s1 * (ci_t)pad_map(j0 - 3 + qj, n1);
            sQ[0][q] = ld(uin, c);
            sQ[1][q] = ld(aux + AX::EINT * nc, c);
            sQ[2][q] = ld(aux + AX::V0 * nc, c);
            sQ[3][q] = ld(aux + (AX::V0 + 1) * nc, c);
            sQ[4][q] = ld(uin + LY::B0 * nc, c);
            sQ[5][q] = ld(uin + (LY::B0 + 1) * nc, c);
        }
        __syncthreads();
        // ratio box cell (bi, bj) = cell (i0 - 1 + bi, j0 - 1 + bj); its
        // window along i starts at window-box column bi of row bj + 2, along
        // j at row bj of column bi + 2
        constexpr int NI = 5 * SH::RJ * SH::NRUN, NJ = 5 * SH::RI;
        for (int u = threadIdx.x; u < NI + NJ; u += SH::NT) {
            if (u < NI) {
                const int a = u / (SH::RJ * SH::NRUN);
                const int y = u - a * (SH::RJ * SH::NRUN);
                const int bj = y / SH::NRUN, bi = SH::RL * (y - bj * SH::NRUN);
                T r[SH::RL];
                ratio_run<SH::RL, T>(sQ[u2_field(0, a)] + (bj + 2) * SH::QI + bi, 1, r);
                const int gj = j0 - 1 + bj;
                const int lim = gj > n1 - 3 ? -1 : n0 - 3 - (i0 - 1 + bi);
                ratio_store<SH::RL, T>(sR + u2_pair(0, a) * SH::RN + bj * SH::RI + bi, 1, r, lim);
            } else {
                const int x = u - NI;
                const int a = x / SH::RI, bi = x - a * SH::RI;
                T r[SH::RJ];
                ratio_run<SH::RJ, T>(sQ[u2_field(1, a)] + bi + 2, SH::QI, r);
                const int gi = i0 - 1 + bi;
                const int lim = gi > n0 - 3 ? -1 : n1 - 3 - (j0 - 1);
                ratio_store<SH::RJ, T>(sR + u2_pair(1, a) * SH::RN + bi, SH::RI, r, lim);
            }
        }
        __syncthreads();
    }
    const int ti = threadIdx.x & 63, tj = threadIdx.x >> 6;
    const int x0 = i0 + ti, x1 = j0 + tj;
    if (x0 > n0 - 3 || x1 > n1 - 3) return;
    RBox rb;
    rb.sR = sR;
    rb.nb = SH::RN;
    rb.own = (tj + 1) * SH::RI + ti + 1;
    rb.bi = SH::RI;
    update_cell<2, T, UNI, FOLD, true>(g, ph, sa, uin, un, uout, acc, nullptr, aux, bg, cmb, x0,
                                       x1, 0, aux_out, mx, rb);
}

template <typename T, bool UNI = false, bool FOLD = false>
__global__ __launch_bounds__(U2<T>::NT) void k_update2r(Geo g, TileMap tm, Phys ph, StageArgs sa,
                                                   const T *__restrict__ uin,
                                                   const T *__restrict__ un,
                                                   T *__restrict__ uout, T *__restrict__ acc,
                                                   const T *__restrict__ aux,
                                                   const double *__restrict__ bg,
                                                   const unsigned long long *__restrict__ cmb,
                                                   T *__restrict__ aux_out = nullptr,
                                                   double *part = nullptr)
{
    __shared__ T sQ[6][U2<T>::QN];
    __shared__ T sR[Lay<2>::NPAIR * U2<T>::RN];
    int tx, ty, tk;
    if (!map_block(tm, tx, ty, tk)) {
        if (FOLD && threadIdx.x < 2) part[4 * blockIdx.x + threadIdx.x] = 0.0;
        return;
    }
    if constexpr (FOLD) {
        __shared__ double red[3][U2<T>::NT / 64];
        double mx[3] = {0.0, 0.0, 0.0};
        update2r_tile<T, UNI, true>(g, ph, sa, uin, un, uout, acc, aux, bg, cmb, tx, ty, sQ, sR,
                                    aux_out, mx);
        block_max_part<2, U2<T>::NT / 64>(mx, red, part + 4 * blockIdx.x);
    } else {
        update2r_tile<T, UNI, false>(g, ph, sa, uin, un, uout, acc, aux, bg, cmb, tx, ty, sQ, sR,
                                     nullptr, nullptr);
    }
}

} // namespace smaug
