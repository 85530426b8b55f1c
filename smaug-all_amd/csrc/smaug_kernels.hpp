// smaug_kernels.hpp -- fused per-stage kernels of the MI355X path.
//
// One RK stage of the semi-discrete operator L(w) is four launches:
//   k_cmax    : cmax_d = max over all cells of cfast_d (wavefront-shuffle
//               max, one 64-bit atomic max per block); replaces
//               cucomputec + cucomputemaxc (computedervfields.cu:1922-2116)
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

__global__ void k_reset_cmax(unsigned long long *cm)
{
    if (threadIdx.x < 3) cm[threadIdx.x] = 0ull;
}

// blockDim.x == 256; grid-stride over every cell (ghosts included)
template <int D, typename T>
__global__ __launch_bounds__(256) void k_cmax(Geo g, Phys ph, const T *__restrict__ u,
                                               const double *__restrict__ bg,
                                               unsigned long long *cm)
{
    __shared__ double red[D][4];
    double mx[D];
#pragma unroll
    for (int d = 0; d < D; d++) mx[d] = 0.0;
    for (idx_t c = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; c < g.nc;
         c += (idx_t)gridDim.x * blockDim.x) {
        St<D, T> s;
        load_st<D, T>(s, u, bg, g.nc, c);
#pragma unroll
        for (int d = 0; d < D; d++) {
            double v = cfast<D, T>(s, d, ph.gamma);
            mx[d] = v > mx[d] ? v : mx[d];
        }
    }
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
        // non-negative doubles order like their bit patterns
        atomicMax(&cm[d], (unsigned long long)__double_as_longlong(v));
    }
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
template <typename T>
__device__ __forceinline__ T ratio6(const T q[6])
{
    T m1 = 0, m3 = 0;
#pragma unroll
    for (int f = 0; f < 5; f++) {
        T v = fabs(q[f + 1] - q[f]);
        m1 = v > m1 ? v : m1;
    }
#pragma unroll
    for (int f = 1; f < 4; f++) {
        T v = fabs((T)3.0 * (q[f + 1] - q[f]) - (q[f + 2] - q[f - 1]));
        m3 = v > m3 ? v : m3;
    }
    return m1 > (T)0 ? m3 / m1 : (T)0;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_ratio(Geo g, TileMap tm, const T *__restrict__ u,
                                                const double *__restrict__ bg, T *__restrict__ R)
{
    typedef Lay<D> LY;
    int tx, ty, tk;
    if (!map_block(tm, tx, ty, tk)) return;
    int x0 = 1 + tx * 64 + (threadIdx.x & 63);
    int x1 = 1 + ty * 4 + (threadIdx.x >> 6);
    int x2 = (D == 3) ? 1 + tk : 0;
    if (x0 > g.n[0] - 3 || x1 > g.n[1] - 3) return;
    int xs[3] = {x0, x1, x2};
    idx_t c = (idx_t)x0 + g.s1 * x1 + g.s2 * x2;
#pragma unroll
    for (int d = 0; d < D; d++) {
        int x = xs[d], n = g.n[d];
        idx_t sd = stride<D>(g, d);
        idx_t c0 = c - (idx_t)x * sd;
        // the differences and the ratio are formed in double in both paths:
        // third differences of nearly smooth FP32 data sit near the FP32
        // rounding level, and the ratio is the quantity that amplifies it
        double qr[6], qe[6], qv[D][6], qb[D][6];
#pragma unroll
        for (int p = 0; p < 6; p++) {
            int xx = x - 2 + p;
            if (xx < 0) xx = 5;
            if (xx >= n) xx = n - 6;
            St<D, T> s;
            load_light<D, T>(s, u, bg, g.nc, c0 + (idx_t)xx * sd);
            double rho = rho_full(s);
            qr[p] = s.r;
            qe[p] = e_int<D, T>(s, rho);
            double ir = 1.0 / rho;
#pragma unroll
            for (int f = 0; f < D; f++) {
                qv[f][p] = (double)s.m[f] * ir;
                qb[f][p] = s.b[f];
            }
        }
        R[(idx_t)LY::prho(d) * g.nc + c] = (T)ratio6<double>(qr);
        R[(idx_t)LY::pe(d) * g.nc + c] = (T)ratio6<double>(qe);
#pragma unroll
        for (int f = 0; f < D; f++) R[(idx_t)LY::pv(f, d) * g.nc + c] = (T)ratio6<double>(qv[f]);
#pragma unroll
        for (int f = 0; f < D; f++)
            if (f != d) R[(idx_t)LY::pb(f, d) * g.nc + c] = (T)ratio6<double>(qb[f]);
    }
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

template <int D, typename T>
__global__ __launch_bounds__(256) void k_bc(Geo g, T *__restrict__ u)
{
    typedef Lay<D> LY;
    idx_t t = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
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

// ---------------------------------------------------------------------------
// stage update
// ---------------------------------------------------------------------------
struct StageArgs {
    int mode;      // 0: RK stage, 1: forward Euler, 2: write raw L(w)
    int stage;     // RK stage 0..3
    double ct;     // stage time offset in units of dt (0, 1/2, 1/2, 1)
    const double *tm; // device [t, dt]
};

// 4th-order divergence, grad3dn form ((8(q1-qm1) + qm2 - q2)/6)/(2 dx)
template <typename T>
__device__ __forceinline__ T div4(T qm2, T qm1, T q1, T q2, T i12dx)
{
    return ((T)8 * (q1 - qm1) + (qm2 - q2)) * i12dx;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_update(Geo g, TileMap tm, Phys ph, StageArgs sa,
                                                 const T *__restrict__ uin,
                                                 const T *__restrict__ un,
                                                 T *__restrict__ uout, T *__restrict__ acc,
                                                 const T *__restrict__ R,
                                                 const double *__restrict__ bg,
                                                 const unsigned long long *__restrict__ cmb)
{
    typedef Lay<D> LY;
    int tx, ty, tk;
    if (!map_block(tm, tx, ty, tk)) return;
    const int x0 = 2 + tx * 64 + (threadIdx.x & 63);
    const int x1 = 2 + ty * 4 + (threadIdx.x >> 6);
    const int x2 = (D == 3) ? 2 + tk : 0;
    if (x0 > g.n[0] - 3 || x1 > g.n[1] - 3) return;
    const int xs[3] = {x0, x1, x2};
    const idx_t nc = g.nc;
    const idx_t c = (idx_t)x0 + g.s1 * x1 + g.s2 * x2;
    const double gam = ph.gamma;

    T dU[LY::NVE];
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) dU[f] = 0;

    St<D, T> sc;
    load_st<D, T>(sc, uin, bg, nc, c);
    const double rhoc = rho_full(sc);
    const T ptb = (T)ptot_bg<D, T>(sc, gam);

    // ---- advection: 4th-order divergence of the fluxes along each axis a,
    // plus the background work terms of the energy equation
    // (centdiff1.cu:19-256, centdiff2.cu:20-356)
#pragma unroll
    for (int a = 0; a < D; a++) {
        const idx_t sa_ = stride<D>(g, a);
        const int xa = xs[a], na = g.n[a];
        if (zero4(g.per[a], xa, na)) continue;
        // d/dx_a q = (q[-2] - 8 q[-1] + 8 q[+1] - q[+2]) / (12 dx), accumulated
        // neighbour by neighbour
        const T i12 = (T)(1.0 / (12.0 * g.dx[a][xa]));
        T gv[D];
#pragma unroll
        for (int d = 0; d < D; d++) gv[d] = 0;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int off = (p < 2) ? p - 2 : p - 1; // -2,-1,+1,+2
            const T cw = (p == 0 ? (T)1 : (p == 1 ? (T)-8 : (p == 2 ? (T)8 : (T)-1))) * i12;
            St<D, T> s;
            load_st<D, T>(s, uin, bg, nc, c + (idx_t)off * sa_);
            const double rho = rho_full(s);
            const T pt = (T)ptot_pert<D, T>(s, rho, gam);
            const T ir = (T)(1.0 / rho);
            const T ma = s.m[a];
            T bbm = 0, bm = 0;
#pragma unroll
            for (int t = 0; t < D; t++) {
                bbm += (T)s.bb[t] * s.m[t];
                bm += s.b[t] * s.m[t];
            }
            // F_rho = m_a rho'/rho + rho_b m_a/rho = m_a
            dU[LY::RHO] -= cw * ma;
#pragma unroll
            for (int i = 0; i < D; i++) {
                T mag = -(s.b[i] * (T)s.bb[a] + (T)s.bb[i] * s.b[a]) - s.b[i] * s.b[a];
                dU[LY::M0 + i] -= cw * (ma * s.m[i] * ir + ((i == a) ? pt + mag : mag));
            }
            dU[LY::EN] -= cw * ((ma * (s.e + (T)s.eb + pt) - s.b[a] * bbm - (T)s.bb[a] * bm -
                                 s.b[a] * bm) * ir);
#pragma unroll
            for (int i = 0; i < D; i++) {
                if (i == a) continue;
                dU[LY::B0 + i] -= cw * ((ma * s.b[i] - (s.b[a] + (T)s.bb[a]) * s.m[i] +
                                         (T)s.bb[i] * ma) * ir);
            }
#pragma unroll
            for (int d = 0; d < D; d++) gv[d] += cw * (s.m[d] * ir);
        }
        // -p_Tb dv_a/dx_a + sum_d b_ba b_bd dv_d/dx_a  (addenergyterms_cd2)
        T w = -ptb * gv[a];
#pragma unroll
        for (int d = 0; d < D; d++) w += (T)sc.bb[a] * (T)sc.bb[d] * gv[d];
        dU[LY::EN] += w;
    }

    // ---- gravity (addgrav_cd2, centdiff2.cu:118-154)
    {
        T ge = 0;
#pragma unroll
        for (int d = 0; d < D; d++) {
            dU[LY::M0 + d] += (T)ph.g[d] * sc.r;
            ge += (T)ph.g[d] * sc.m[d];
        }
        dU[LY::EN] += sc.r * ge * (T)(1.0 / rhoc);
    }

    // ---- hyperdiffusion (iosmaug.cpp:969-1146)
    if (ph.hyperdif) {
        double cmax[D];
#pragma unroll
        for (int d = 0; d < D; d++) cmax[d] = __longlong_as_double((long long)cmb[d]);
        const T eic = (T)e_int<D, T>(sc, rhoc);
        const T irc = (T)(1.0 / rhoc);
        T vc[D];
#pragma unroll
        for (int f = 0; f < D; f++) vc[f] = sc.m[f] * irc;

#pragma unroll
        for (int d = 0; d < D; d++) {
            const idx_t sd = stride<D>(g, d);
            const int xd = xs[d], nd = g.n[d];
            const double dxd = g.dx[d][xd];
            const T idxd = (T)g.idx[d][xd];
            const bool z1 = zero1(g.per[d], xd, nd);
            St<D, T> sl, sr;
            load_light<D, T>(sl, uin, bg, nc, c - sd);
            load_light<D, T>(sr, uin, bg, nc, c + sd);
            const double rhol = rho_full(sl), rhor = rho_full(sr);
            const T irl = (T)(1.0 / rhol), irr = (T)(1.0 / rhor);

            // diagonal terms: q += (nu_R dq_R - nu_L dq_L)/dx
            auto nuRL = [&](int pair, int field, T &nR, T &nL) {
                const T K = (T)(dxd * cmax[d] * ph.chyp[field]);
                nR = K * R[(idx_t)pair * nc + c];
                nL = K * R[(idx_t)pair * nc + c - sd];
            };
            T nR, nL, gR, gL;
            // density (hyperdifrhosource1.cu:15-120)
            nuRL(LY::prho(d), LY::RHO, nR, nL);
            gR = z1 ? (T)0 : (sr.r - sc.r) * idxd;
            gL = z1 ? (T)0 : (sc.r - sl.r) * idxd;
            dU[LY::RHO] += (nR * gR - nL * gL) * idxd;
            // energy (hyperdifesource1.cu:18-330)
            nuRL(LY::pe(d), LY::EN, nR, nL);
            {
                const T el = (T)e_int<D, T>(sl, rhol), er = (T)e_int<D, T>(sr, rhor);
                gR = z1 ? (T)0 : (er - eic) * idxd;
                gL = z1 ? (T)0 : (eic - el) * idxd;
                dU[LY::EN] += (nR * gR - nL * gL) * idxd;
            }
            // momentum, diagonal part for every f; twice when f == d
            // (hyperdifmomsource1.cu:17-533 called for ii1 = 0 and, if f == d, ii1 = 1)
#pragma unroll
            for (int f = 0; f < D; f++) {
                nuRL(LY::pv(f, d), LY::M0 + f, nR, nL);
                const T vl = sl.m[f] * irl, vr = sr.m[f] * irr;
                gR = z1 ? (T)0 : (vr - vc[f]) * idxd;
                gL = z1 ? (T)0 : (vc[f] - vl) * idxd;
                const T rR = (T)((rhoc + rhor) / 2), rL = (T)((rhoc + rhol) / 2);
                const T mR = (sc.m[f] + sr.m[f]) / (T)2, mL = (sc.m[f] + sl.m[f]) / (T)2;
                const T k2 = (f == d) ? (T)1 : (T)0.5;
                dU[LY::M0 + f] += k2 * (rR * nR * gR - rL * nL * gL) * idxd;
                dU[LY::EN] += k2 * (mR * nR * gR - mL * nL * gL) * idxd;
            }
            // magnetic field, diagonal part for f != d (hyperdifbsource1.cu:21-440)
#pragma unroll
            for (int f = 0; f < D; f++) {
                if (f == d) continue;
                nuRL(LY::pb(f, d), LY::B0 + f, nR, nL);
                gR = z1 ? (T)0 : (sr.b[f] - sc.b[f]) * idxd;
                gL = z1 ? (T)0 : (sc.b[f] - sl.b[f]) * idxd;
                const T bR = (sc.b[f] + sr.b[f]) / (T)2, bL = (sc.b[f] + sl.b[f]) / (T)2;
                dU[LY::B0 + f] += (gR * nR - gL * nL) * idxd;
                dU[LY::EN] += (bR * gR * nR - bL * gL * nL) * idxd;
            }

            // off-diagonal terms for f != d: d/dx_f of a product evaluated at
            // c -/+ e_f (hyperdifmomsourcene1.cu:16-740, hyperdifbsourcene1.cu:18-820)
#pragma unroll
            for (int f = 0; f < D; f++) {
                if (f == d) continue;
                const idx_t sf = stride<D>(g, f);
                const int xf = xs[f], nf = g.n[f];
                T Xr[2], Xm[2], Hb[2], He[2];
#pragma unroll
                for (int side = 0; side < 2; side++) {
                    const int sg = side ? 1 : -1;
                    const idx_t cn = c + sg * sf;
                    Xr[side] = Xm[side] = Hb[side] = He[side] = 0;
                    const int xfn = xf + sg;
                    if (xfn < 2 || xfn > nf - 3) continue; // nu vanishes off the interior
                    St<D, T> sn;
                    load_light<D, T>(sn, uin, bg, nc, cn);
                    const double rhon = rho_full(sn);
                    // v_f and b_f at cn -/+ e_d
                    const idx_t ca = cn - sd, cb = cn + sd;
                    const double rha = (double)uin[ca] + bg[nc + ca];
                    const double rhb = (double)uin[cb] + bg[nc + cb];
                    const T vfa = uin[(idx_t)(LY::M0 + f) * nc + ca] * (T)(1.0 / rha);
                    const T vfb = uin[(idx_t)(LY::M0 + f) * nc + cb] * (T)(1.0 / rhb);
                    const T bfa = uin[(idx_t)(LY::B0 + f) * nc + ca];
                    const T bfb = uin[(idx_t)(LY::B0 + f) * nc + cb];
                    const T dv = z1 ? (T)0 : (vfb - vfa) * idxd / (T)2;
                    const T db = z1 ? (T)0 : (bfb - bfa) * idxd / (T)2;
                    {
                        const T K = (T)(dxd * cmax[d] * ph.chyp[LY::M0 + f]);
                        const T s2 = K * (R[(idx_t)LY::pv(f, d) * nc + cn] +
                                          R[(idx_t)LY::pv(f, d) * nc + cn - sd]);
                        const T G = dv * s2 / (T)4;
                        Xr[side] = (T)rhon * G;
                        Xm[side] = G * sn.m[d];
                    }
                    {
                        const T K = (T)(dxd * cmax[d] * ph.chyp[LY::B0 + f]);
                        const T s2 = K * (R[(idx_t)LY::pb(f, d) * nc + cn] +
                                          R[(idx_t)LY::pb(f, d) * nc + cn - sd]);
                        const T H = db * s2 / (T)2;
                        Hb[side] = H;
                        He[side] = H * sn.b[d];
                    }
                }
                if (!zero1(g.per[f], xf, nf)) {
                    const T h = (T)g.idx[f][xf] / (T)2;
                    dU[LY::M0 + d] += (Xr[1] - Xr[0]) * h;
                    dU[LY::EN] += (Xm[1] - Xm[0]) * h;
                    dU[LY::B0 + d] -= (Hb[1] - Hb[0]) * h;
                    dU[LY::EN] -= (He[1] - He[0]) * h;
                }
            }
        }
    }

    // ---- user driver: vertical velocity, m_0 += rho v, e += rho v^2/2
    const double t = sa.tm[0], dt = sa.tm[1];
    if (ph.driver == 1) {
        const double ts = t + sa.ct * dt;
        const double tdep = sin(ts * 6.283185307179586 / ph.drv[1]);
        double pr = ph.drv[0] * tdep * g.prof[0][x0] * g.prof[1][x1];
        if (D == 3) pr *= g.prof[2][x2];
        const T v = (T)pr;
        dU[LY::M0] += v * (T)rhoc;
        dU[LY::EN] += v * v * (T)rhoc / (T)2;
    }

    // ---- write
    if (sa.mode == 2) {
#pragma unroll
        for (int f = 0; f < LY::NVE; f++) uout[(idx_t)f * nc + c] = dU[f];
        return;
    }
    if (sa.mode == 1) {
        const T tdt = (T)dt;
#pragma unroll
        for (int f = 0; f < LY::NVE; f++)
            uout[(idx_t)f * nc + c] = un[(idx_t)f * nc + c] + tdt * dU[f];
        return;
    }
    // low-storage classical RK4:
    //  s0: out = un + dt/2 k, acc = un + dt/6 k
    //  s1: out = un + dt/2 k, acc += dt/3 k
    //  s2: out = un + dt k,   acc += dt/3 k
    //  s3: out = acc + dt/6 k
    const int s = sa.stage;
    const T a = (T)(s == 2 ? dt : (s == 3 ? dt / 6 : dt / 2));
    const T b = (T)(s == 0 ? dt / 6 : dt / 3);
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) {
        const idx_t o = (idx_t)f * nc + c;
        if (s == 3) {
            uout[o] = acc[o] + a * dU[f];
        } else {
            const T base = un[o];
            uout[o] = base + a * dU[f];
            acc[o] = (s == 0 ? base : acc[o]) + b * dU[f];
        }
    }
}

} // namespace smaug
