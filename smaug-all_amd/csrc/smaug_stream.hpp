// smaug_stream.hpp -- the 2.5-D z-streamed stage update for 3-D grids.
//
// A block owns a tile of 64 x 4 columns (i fastest, then j) and walks its
// columns upward in k over a chunk of kz output planes.  For every plane it
// stages, in LDS, the advective fluxes along i and j and the velocity of
// the tile plus a 2-cell halo (one global read of each cell's state per
// block and plane); the 4th-order differences along i and j then read LDS.
// Along k each thread keeps a ring of five pending outputs in registers: the
// k-flux of plane P is formed once and scattered into outputs P-2..P+2, so
// no plane is read twice along the stream.  An output plane is complete two
// planes behind the front, where the hyperdiffusive sources (formed by
// k_hyper into their own buffer), the driver and the RK slot algebra are
// applied and the result written.
//
// The per-cell arithmetic is the reference operator of centdiff1.cu:19-256
// and centdiff2.cu:20-356 (fluxes and background work terms), addgrav_cd2
// (centdiff2.cu:118-154), and the hyperdiffusion, driver and RK combination
// of smaug_kernels.hpp.  Only the order in which the twelve neighbour
// contributions are summed differs from the per-cell kernel.
#pragma once
#include "smaug_kernels.hpp"

namespace smaug {

constexpr int ZS_TI = 64, ZS_TJ = 4;                // output columns per block
constexpr int ZS_BI = ZS_TI + 4, ZS_BJ = ZS_TJ + 4; // LDS box incl. 2-cell halo
constexpr int ZS_NB = ZS_BI * ZS_BJ;
constexpr int ZS_NT = ZS_TI * ZS_TJ;                // threads per block
constexpr int ZS_NH = ZS_NB - ZS_NT;                // halo cells of the box
constexpr int ZS_NHQ = (ZS_NH + ZS_NT - 1) / ZS_NT; // halo cells per thread
constexpr int ZS_KZMAX = 256; // deepest z-stream chunk (per-plane factors staged in LDS)

// block -> (tile i, tile j, z-chunk).  The tiles are cut into 8 contiguous
// groups, one per XCD (block b runs on XCD b % 8); consecutive blocks of an
// XCD take neighbouring tiles of the same z-chunk, whose halos then meet in
// that XCD's L2.
struct ZMap {
    int ntx, nty;   // tiles along i and j
    int kz, nkz;    // output planes per chunk, chunks along k
    int per_xcd;    // tiles per XCD group
};

__device__ __forceinline__ bool zmap_block(const ZMap &m, int &tx, int &ty, int &kc)
{
    const int b = blockIdx.x;
    const int xcd = b & 7, local = b >> 3;
    kc = local / m.per_xcd;
    const int t = xcd * m.per_xcd + (local - kc * m.per_xcd);
    if (kc >= m.nkz || t >= m.ntx * m.nty) return false;
    tx = t % m.ntx;
    ty = t / m.ntx;
    return true;
}

// box coordinates of halo cell h of the box: rows 0,1 and the last two in
// full, then columns 0,1 and the last two of the rows between
__device__ __forceinline__ void zs_halo(int h, int &bi, int &bj)
{
    if (h < 2 * ZS_BI) {
        bj = h / ZS_BI;
        bi = h - bj * ZS_BI;
    } else if (h < 4 * ZS_BI) {
        const int q = h - 2 * ZS_BI;
        bj = ZS_BJ - 2 + q / ZS_BI;
        bi = q % ZS_BI;
    } else {
        const int q = h - 4 * ZS_BI;
        bj = 2 + (q >> 2);
        const int c = q & 3;
        bi = c < 2 ? c : ZS_BI - 4 + c;
    }
}

// one cell's state in the form the flux needs (T arithmetic, as the
// per-cell kernel uses it: background rounded to T, 1/rho and the pressure
// perturbation formed in double and rounded)
template <typename T>
struct ZCell {
    T r, e, eb, ir, pt, m[3], b[3], bb[3], v[3];
};

// a cell's raw state as loaded (the loads of the next plane are issued
// into it while the current plane is being processed).  BGP: the background
// does not vary along k and was loaded into s once for the column, so only
// the evolved fields are loaded.
template <bool BGP, typename T>
__device__ __forceinline__ void zs_raw(St<3, T> &s, const T *__restrict__ u,
                                       const double *__restrict__ bg, size_t nc, ci_t c)
{
    if constexpr (BGP)
        load_pert32<3, T>(s, u, nc, c);
    else
        load_st32<3, T>(s, u, bg, nc, c);
}

template <typename T>
__device__ __forceinline__ void zs_derive(ZCell<T> &z, const St<3, T> &s, double gam,
                                          double *rb_out = nullptr, double *ptb_out = nullptr)
{
    const double rho = rho_full(s);
    const double ird = rcp64(rho);
    z.r = s.r;
    z.e = s.e;
    z.eb = (T)s.eb;
    z.ir = (T)ird;
    z.pt = (T)ptot_pert_ir<3, T>(s, ird, gam);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        z.m[d] = s.m[d];
        z.b[d] = s.b[d];
        z.bb[d] = (T)s.bb[d];
        z.v[d] = (T)((double)s.m[d] * ird);
    }
    if (rb_out) *rb_out = s.rb;
    if (ptb_out) {
        double bbb = 0;
#pragma unroll
        for (int d = 0; d < 3; d++) bbb += s.bb[d] * s.bb[d];
        *ptb_out = (gam - 1) * s.eb - 0.5 * (gam - 2) * bbb;
    }
}

// the seven non-trivial components of the flux along a: rho, m_0..2, e,
// and b_i for the two i != a (increasing i)
template <typename T>
__device__ __forceinline__ void zs_flux(int a, const ZCell<T> &z, T F[7])
{
    T bbm = 0, bm = 0;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        bbm += z.bb[t] * z.m[t];
        bm += z.b[t] * z.m[t];
    }
    const T ma = z.m[a];
    F[0] = ma;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T mag = -(z.b[i] * z.bb[a] + z.bb[i] * z.b[a]) - z.b[i] * z.b[a];
        F[1 + i] = ma * z.m[i] * z.ir + ((i == a) ? z.pt + mag : mag);
    }
    F[4] = (ma * (z.e + z.eb + z.pt) - z.b[a] * bbm - z.bb[a] * bm - z.b[a] * bm) * z.ir;
    int q = 5;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i == a) continue;
        F[q++] = (ma * z.b[i] - (z.b[a] + z.bb[a]) * z.m[i] + z.bb[i] * ma) * z.ir;
    }
}

// evolved-field index of flux component q along a
__device__ __forceinline__ int zs_field(int a, int q)
{
    if (q < 5) return q; // rho, m_0..2, e
    const int first = a == 0 ? 1 : 0;
    const int second = a == 2 ? 1 : 2;
    return 5 + (q == 5 ? first : second);
}

// BGP: bg is a background that does not vary along k, stored as one plane
// (field stride s2); its values are loaded once per column.  FOLD (with
// BGP): the aux quantities (e_int, v) of every output cell are written to
// aux_out and the block's cfast maxima of the output to part[4 b + d], so
// that the next stage needs no aux/cmax pass over the interior
template <bool BGP, bool FOLD, typename T>
__device__ __forceinline__ void zstream_body(const Geo &g, const Phys &ph, const StageArgs &sa,
                                             const ZMap &zm, const T *__restrict__ uin,
                                             const T *__restrict__ un, T *__restrict__ uout,
                                             T *__restrict__ acc, const T *__restrict__ dUh,
                                             const double *__restrict__ bg,
                                             T (*sF)[7][ZS_NB], T (*sV)[ZS_NB], T *sZ,
                                             double *sP, double *sQ, T *__restrict__ aux_out,
                                             double *part, double (*red)[4])
{
    typedef Lay<3> LY;

    int tx, ty, kc;
    if (!zmap_block(zm, tx, ty, kc)) {
        if (FOLD && threadIdx.x < 3) part[4 * blockIdx.x + threadIdx.x] = 0.0;
        return;
    }
    double mxf[3] = {0.0, 0.0, 0.0};
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const size_t nc = (size_t)g.nc;
    const ci_t s1 = (ci_t)g.s1, s2 = (ci_t)g.s2;
    const double gam = ph.gamma;
    const int ti = threadIdx.x % ZS_TI, tj = threadIdx.x / ZS_TI;
    const int i0 = 2 + tx * ZS_TI, j0 = 2 + ty * ZS_TJ;
    const int i = i0 + ti, j = j0 + tj;
    const bool act = i <= n0 - 3 && j <= n1 - 3;
    const int k0 = 2 + kc * zm.kz;
    const int k1 = min(k0 + zm.kz, n2 - 2);
    const int own = (tj + 2) * ZS_BI + ti + 2;
    // this thread's halo cells of the box (one or two).  A cell of the two
    // rows above or below the tile is read only by the differences along j,
    // one of the two columns either side only by those along i, so each
    // forms its flux along that axis alone (hax); the corner cells of the
    // box are read by neither and are skipped.
    int hb[ZS_NHQ], hc[ZS_NHQ], hax[ZS_NHQ];
    bool hv[ZS_NHQ];
#pragma unroll
    for (int q = 0; q < ZS_NHQ; q++) {
        const int h = threadIdx.x + q * ZS_NT;
        hv[q] = h < ZS_NH;
        int bi = 0, bj = 0;
        if (hv[q]) zs_halo(h, bi, bj);
        const int gi = i0 - 2 + bi, gj = j0 - 2 + bj;
        const bool rowh = bj < 2 || bj >= ZS_BJ - 2, colh = bi < 2 || bi >= ZS_BI - 2;
        hv[q] = hv[q] && gi <= n0 - 1 && gj <= n1 - 1 && !(rowh && colh);
        hax[q] = rowh ? 1 : 0;
        hb[q] = bj * ZS_BI + bi;
        hc[q] = gi + (int)s1 * gj;
    }
    const bool own_in = i <= n0 - 1 && j <= n1 - 1; // own cell lies in the grid
    const ci_t colc = (ci_t)i + s1 * (ci_t)j;

    // Metric and driver factors, read once per block: the in-plane ones per
    // thread into registers, those of the chunk's output planes into LDS (a
    // global load inside the plane loop would be waited for together with
    // the prefetched next plane, as vector memory operations retire in
    // order).  Same arithmetic as the per-cell kernel.
    const double dt = sa.tm[1];
    // stage result = old + rk_a L(w) (old = the accumulator in RK stage 3),
    // accumulator = (u_n or accumulator) + rk_b L(w) in RK stages 0-2
    const int rs = sa.stage;
    const T rk_a = sa.mode == 1 ? (T)dt : (T)(rs == 2 ? dt : (rs == 3 ? dt * (1.0 / 6.0) : dt * 0.5));
    const T rk_b = (T)(rs == 0 ? dt * (1.0 / 6.0) : dt * (1.0 / 3.0));
    const bool rk_old_acc = sa.mode == 0 && rs == 3;
    T i12x[2] = {0, 0};
    // driver 1: v = pxy P2(k); driver 2 (torsional): vx = pxy Q2(k),
    // vy = pxq P2(k), as driver_cell forms them
    double pxy = 0, pxq = 0;
    if (act) {
        i12x[0] = (T)(1.0 / (12.0 * g.dx[0][i]));
        i12x[1] = (T)(1.0 / (12.0 * g.dx[1][j]));
        if (ph.driver == 1) {
            pxy = driver_amp(ph, sa, sa.tm[0], dt) * g.prof[0][i] * g.prof[1][j];
        } else if (ph.driver == 2) {
            const double a = driver_amp(ph, sa, sa.tm[0], dt) / ph.drv[7] * g.prof[0][i];
            pxy = a * g.prof[1][j];
            pxq = -(a * g.profq[1][j]);
        }
    }
    for (int t = threadIdx.x; t < k1 - k0; t += ZS_NT) {
        sZ[t] = (T)(1.0 / (12.0 * g.dx[2][k0 + t]));
        sP[t] = g.prof[2][k0 + t];
        sQ[t] = g.profq[2][k0 + t];
    }
    __syncthreads();

    // Ring of pending outputs, slots 0..3 = planes P-2..P+1 at the start of
    // front P: rU holds dU; rG the k-differences of v still waiting for
    // the background-work coefficients of planes P and P+1 (slots 2, 3);
    // rC those coefficients (-p_Tb, b_b0..2) of planes P-2 and P-1.
    T rU[4][8], rG[2][3], rC[2][4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int f = 0; f < 8; f++) rU[q][f] = 0;
#pragma unroll
    for (int q = 0; q < 2; q++) {
#pragma unroll
        for (int d = 0; d < 3; d++) rG[q][d] = 0;
#pragma unroll
        for (int f = 0; f < 4; f++) rC[q][f] = 0;
    }

    const bool zper = g.per[2];
    const bool need_un = sa.mode == 1 || (sa.mode == 0 && sa.stage != 3);
    const bool need_acc = sa.mode == 0 && sa.stage != 0;
    // raw state of the own cell and the halo cells of the current plane
    St<3, T> ro, rh[ZS_NHQ];
    if constexpr (BGP) {
        load_bg32<3, T>(ro, bg, (size_t)s2, own_in ? colc : 0u);
#pragma unroll
        for (int q = 0; q < ZS_NHQ; q++) load_bg32<3, T>(rh[q], bg, (size_t)s2, hv[q] ? (ci_t)hc[q] : 0u);
    }
    zs_raw<BGP, T>(ro, uin, bg, nc, own_in ? colc + s2 * (ci_t)(k0 - 2) : 0u);
    // The output completed in one iteration (stage result ov, accumulator
    // av) is stored at the head of the next one, behind the loads of the own
    // cell's plane, which the head then waits for without waiting for the
    // stores.  Every lane stores on every path: without an output it writes
    // 0 into cell 0, a corner ghost cell that no kernel of the stage reads
    // and the ghost fill rewrites.
    const bool wacc = sa.mode == 0 && sa.stage != 3;
    T ov[8], av[8];
#pragma unroll
    for (int f = 0; f < 8; f++) ov[f] = av[f] = 0;
    ci_t cs = 0, ca = 0;
    auto store_out = [&]() {
#pragma unroll
        for (int f = 0; f < 8; f++) st(uout + f * nc, cs, ov[f]);
#pragma unroll
        for (int f = 0; f < 8; f++) st(acc + f * nc, ca, av[f]);
    };
    for (int P = k0 - 2; P < k1 + 2; P++) {
        store_out();
        const bool outp = P >= k0 && P < k1;
        ZCell<T> z;
        double rbo = 0, ptbo = 0;
        if (own_in) zs_derive<T>(z, ro, gam, &rbo, &ptbo);
        T Fz[7];
        if (act) zs_flux<T>(2, z, Fz);
        if (outp) {
            __syncthreads(); // the previous plane's LDS reads are done
            if (own_in) {
                T F[7];
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    zs_flux<T>(a, z, F);
#pragma unroll
                    for (int q = 0; q < 7; q++) sF[a][q][own] = F[q];
                }
#pragma unroll
                for (int d = 0; d < 3; d++) sV[d][own] = z.v[d];
            }
#pragma unroll
            for (int q = 0; q < ZS_NHQ; q++)
                if (hv[q]) {
                    ZCell<T> h;
                    zs_derive<T>(h, rh[q], gam);
                    T F[7];
                    if (hax[q]) {
                        zs_flux<T>(1, h, F);
#pragma unroll
                        for (int f = 0; f < 7; f++) sF[1][f][hb[q]] = F[f];
                    } else {
                        zs_flux<T>(0, h, F);
#pragma unroll
                        for (int f = 0; f < 7; f++) sF[0][f][hb[q]] = F[f];
                    }
#pragma unroll
                    for (int d = 0; d < 3; d++) sV[d][hb[q]] = h.v[d];
                }
        }
        // Loads for later, issued before the in-plane work so they are in
        // flight while it runs: first the inputs of the output completed in
        // this iteration, then the raw state of plane P+1 (vector memory
        // operations retire in order, so waiting for the former does not
        // wait for the latter).
        const int O = P - 2;
        const bool fin = act && O >= k0 && O < k1;
        const ci_t co = colc + s2 * (ci_t)O;
        // (issued by every lane: an input the stage does not use, or a lane
        // without an output, reads cell 0)
        T fu[8], fa[8], fh[8], fr;
        double frb;
        {
            const ci_t cu = fin && need_un ? co : 0u, cc = fin && need_acc ? co : 0u;
            const ci_t chy = fin && ph.hyperdif ? co : 0u, cd = fin && ph.driver ? co : 0u;
#pragma unroll
            for (int f = 0; f < 8; f++) {
                fu[f] = ld(un + f * nc, cu);
                fa[f] = ld(acc + f * nc, cc);
                fh[f] = ld(dUh + f * nc, chy);
            }
            fr = ld(uin, cd);
            if constexpr (BGP)
                frb = ro.rb;
            else
                frb = ld(bg + nc, cd);
        }
        // The own cell's loads come last and are issued on every path (the
        // last iteration reloads its plane, lanes off the grid read cell 0):
        // the compiler's waits for the loads above then count past them.
        if (P + 1 >= k0 && P + 1 < k1)
#pragma unroll
            for (int q = 0; q < ZS_NHQ; q++)
                if (hv[q]) zs_raw<BGP, T>(rh[q], uin, bg, nc, (ci_t)hc[q] + s2 * (ci_t)(P + 1));
        zs_raw<BGP, T>(ro, uin, bg, nc, own_in ? colc + s2 * (ci_t)min(P + 1, k1 + 1) : 0u);
        if (outp) {
            __syncthreads();
            if (act) {
                // 4th-order differences along i and j from LDS, background
                // work terms, gravity
                T bbc[3];
#pragma unroll
                for (int d = 0; d < 3; d++) bbc[d] = z.bb[d];
                const T ptb = (T)ptbo;
                const int xs[2] = {i, j};
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    // the divergence along a is left out where zero4 holds:
                    // a zero factor (the box cells read lie on the grid)
                    const int st_ = a == 0 ? 1 : ZS_BI;
                    const T i12 = zero4(g.per[a], xs[a], g.n[a]) ? (T)0 : i12x[a];
                    T gv[3] = {0, 0, 0};
#pragma unroll
                    for (int p = 0; p < 4; p++) {
                        const int off = (p < 2) ? p - 2 : p - 1;
                        const T cw = (p == 0 ? (T)1 : (p == 1 ? (T)-8 : (p == 2 ? (T)8 : (T)-1))) * i12;
                        const int nb = own + off * st_;
#pragma unroll
                        for (int q = 0; q < 7; q++) rU[2][zs_field(a, q)] -= cw * sF[a][q][nb];
#pragma unroll
                        for (int d = 0; d < 3; d++) gv[d] += cw * sV[d][nb];
                    }
                    T w = -ptb * gv[a];
#pragma unroll
                    for (int d = 0; d < 3; d++) w += bbc[a] * bbc[d] * gv[d];
                    rU[2][LY::EN] += w;
                }
                T ge = 0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    rU[2][LY::M0 + d] += (T)ph.g[d] * z.r;
                    ge += (T)ph.g[d] * z.m[d];
                }
                rU[2][LY::EN] += z.r * ge * z.ir;
            }
        }
#pragma unroll
        for (int f = 0; f < 8; f++) ov[f] = av[f] = 0;
        cs = fin ? co : 0u;
        ca = fin && wacc ? co : 0u;
        if (act) {
            // background-work coefficients of plane P; the k-differences of
            // v gathered for output P so far become its energy term
            T cP[4];
            cP[0] = -(T)ptbo;
#pragma unroll
            for (int d = 0; d < 3; d++) cP[1 + d] = z.bb[d];
            auto bgwork = [&](const T *c, const T *gz) -> T {
                T w = c[0] * gz[2];
#pragma unroll
                for (int d = 0; d < 3; d++) w += c[3] * c[1 + d] * gz[d];
                return w;
            };
            if (outp) rU[2][LY::EN] += bgwork(cP, rG[0]);
            // scatter the k-flux of plane P into the outputs P-2..P+2
            const T *F = Fz;
            T nU[8], nG[3]; // the new slot of output P+2
#pragma unroll
            for (int f = 0; f < 8; f++) nU[f] = 0;
#pragma unroll
            for (int d = 0; d < 3; d++) nG[d] = 0;
#pragma unroll
            for (int q = 0; q < 5; q++) {
                if (q == 2) continue;
                const int O = P - 2 + q;
                if (O < k0 || O >= k1 || zero4(zper, O, n2)) continue;
                // neighbour offset P - O = 2 - q; weights 1, -8, 8, -1 for
                // offsets -2, -1, +1, +2
                const int off = 2 - q;
                const T base = off == -2 ? (T)1 : (off == -1 ? (T)-8 : (off == 1 ? (T)8 : (T)-1));
                const T cw = base * sZ[O - k0];
                T *dU = q < 4 ? rU[q] : nU;
#pragma unroll
                for (int f = 0; f < 7; f++) dU[zs_field(2, f)] -= cw * F[f];
                T gz[3];
#pragma unroll
                for (int d = 0; d < 3; d++) gz[d] = cw * z.v[d];
                if (q < 2)
                    dU[LY::EN] += bgwork(rC[q], gz);
                else if (q == 3)
#pragma unroll
                    for (int d = 0; d < 3; d++) rG[1][d] += gz[d];
                else
#pragma unroll
                    for (int d = 0; d < 3; d++) nG[d] = gz[d];
            }
            // output P-2 is complete
            if (fin) {
                T *dU = rU[0];
                if (ph.hyperdif) {
#pragma unroll
                    for (int f = 0; f < 8; f++) dU[f] += fh[f];
                }
                if (ph.driver == 1)
                    driver_add<3, T>(pxy * sP[O - k0], fr, frb, dU);
                else if (ph.driver == 2)
                    driver_add_tor<T>(pxy * sQ[O - k0], pxq * sP[O - k0], fr, frb, dU);
                // the stage result and accumulator as rk_write forms them
#pragma unroll
                for (int f = 0; f < 8; f++) {
                    ov[f] = sa.mode == 2 ? dU[f] : (rk_old_acc ? fa[f] : fu[f]) + rk_a * dU[f];
                    av[f] = wacc ? (sa.stage == 0 ? fu[f] : fa[f]) + rk_b * dU[f] : (T)0;
                }
                if constexpr (FOLD) {
                    // the output's aux and cfast, as the aux pass forms them
                    St<3, T> so = ro; // the column's background
                    so.r = ov[0];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        so.m[d] = ov[1 + d];
                        so.b[d] = ov[5 + d];
                    }
                    so.e = ov[4];
                    prep_st<3, T>(g, ph, so, aux_out, co, mxf, 0);
                }
            }
            // advance the ring
#pragma unroll
            for (int q = 0; q < 3; q++)
#pragma unroll
                for (int f = 0; f < 8; f++) rU[q][f] = rU[q + 1][f];
#pragma unroll
            for (int f = 0; f < 8; f++) rU[3][f] = nU[f];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                rG[0][d] = rG[1][d];
                rG[1][d] = nG[d];
            }
#pragma unroll
            for (int f = 0; f < 4; f++) {
                rC[0][f] = rC[1][f];
                rC[1][f] = cP[f];
            }
        }
    }
    store_out();
    if constexpr (FOLD) block_max_part<3>(mxf, red, part + 4 * blockIdx.x);
}

// FP32 state: up to 256 VGPRs (the prefetched plane and the output ring
// stay in registers without spills), two 256-thread blocks of 37 KB of LDS
// per CU.  BGP: bg is a background stored as one k-plane.
template <bool BGP, bool FOLD = false>
__global__ __launch_bounds__(ZS_NT) __attribute__((amdgpu_waves_per_eu(2))) void
k_zstream_f(Geo g, Phys ph, StageArgs sa, ZMap zm, const float *__restrict__ uin,
            const float *__restrict__ un, float *__restrict__ uout, float *__restrict__ acc,
            const float *__restrict__ dUh, const double *__restrict__ bg,
            float *__restrict__ aux_out = nullptr, double *part = nullptr)
{
    __shared__ double red[3][4];
    __shared__ float sF[2][7][ZS_NB]; // fluxes along i (0) and j (1)
    __shared__ float sV[3][ZS_NB];    // velocity
    __shared__ float sZ[ZS_KZMAX];    // 1/(12 dz) of the output planes
    __shared__ double sP[ZS_KZMAX];   // driver profile along k
    __shared__ double sQ[ZS_KZMAX];   // torsional driver's lever profile along k
    zstream_body<BGP, FOLD, float>(g, ph, sa, zm, uin, un, uout, acc, dUh, bg, sF, sV, sZ, sP, sQ,
                                   aux_out, part, red);
}

// FP64 state: one wave per SIMD, so that nothing spills; 74 KB of LDS
template <bool BGP, bool FOLD = false>
__global__ __launch_bounds__(ZS_NT) __attribute__((amdgpu_waves_per_eu(1))) void
k_zstream_d(Geo g, Phys ph, StageArgs sa, ZMap zm, const double *__restrict__ uin,
            const double *__restrict__ un, double *__restrict__ uout, double *__restrict__ acc,
            const double *__restrict__ dUh, const double *__restrict__ bg,
            double *__restrict__ aux_out = nullptr, double *part = nullptr)
{
    __shared__ double red[3][4];
    __shared__ double sF[2][7][ZS_NB];
    __shared__ double sV[3][ZS_NB];
    __shared__ double sZ[ZS_KZMAX];
    __shared__ double sP[ZS_KZMAX];
    __shared__ double sQ[ZS_KZMAX];
    zstream_body<BGP, FOLD, double>(g, ph, sa, zm, uin, un, uout, acc, dUh, bg, sF, sV, sZ, sP, sQ,
                                    aux_out, part, red);
}

// ---------------------------------------------------------------------------
// Hyperviscosity ratios, z-streamed.  A block owns 64 x 8 columns of the box
// [1, n-3] and walks a chunk of its planes.  Along i the blocks cover whole
// rows [0, n-1]: the ratios at i = 0, n-2, n-1 are never read, but writing
// them keeps the 21 output rows free of partly written cache lines.  Each plane's window quantities
// (rho', e_int, v, b') over the tile plus the window halo (2 cells below, 3
// above, pads mirrored as in ratio_cell) are staged in LDS once; the ratios
// along i and j read their 6-point windows there.  Along k every thread
// keeps the last six planes of its own column in registers and forms the
// k-ratios of the plane three behind the front.  The next plane's loads are
// issued into registers before the current plane is processed.  Same
// arithmetic (ratio6) as k_ratio, so the two give identical ratios.
// ---------------------------------------------------------------------------
constexpr int RZ_TI = 64, RZ_TJ = 8, RZ_NT = RZ_TI * RZ_TJ; // columns (threads) per block
constexpr int RB_BI = RZ_TI + 5, RB_BJ = RZ_TJ + 5, RB_NB = RB_BI * RB_BJ;
constexpr int RB_NQ = 8;                           // rho', e_int, v_0..2, b'_0..2
constexpr int RB_PER = (RB_NB + RZ_NT - 1) / RZ_NT; // box cells per thread


template <typename T>
__device__ __forceinline__ void ratio_zs_body(const Geo &g, const ZMap &zm,
                                              const T *__restrict__ u,
                                              const T *__restrict__ aux, T *__restrict__ R,
                                              T (*sQ)[RB_NB])
{
    typedef Lay<3> LY;
    typedef Aux<3> AX;
    int tx, ty, kc;
    if (!zmap_block(zm, tx, ty, kc)) return;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const size_t nc = (size_t)g.nc;
    const ci_t s1 = (ci_t)g.s1, s2 = (ci_t)g.s2;
    const int ti = threadIdx.x % RZ_TI, tj = threadIdx.x / RZ_TI;
    const int i0 = tx * RZ_TI, j0 = 1 + ty * RZ_TJ;
    const int i = i0 + ti, j = j0 + tj;
    const bool act = i <= n0 - 1 && j <= n1 - 3;
    const int k0 = 1 + kc * zm.kz;
    const int k1 = min(k0 + zm.kz, n2 - 2);
    const int own = (tj + 2) * RB_BI + ti + 2;
    // in-plane offsets of this thread's box cells (pads mirrored)
    ci_t bc[RB_PER];
    bool bv[RB_PER];
#pragma unroll
    for (int q = 0; q < RB_PER; q++) {
        const int b = threadIdx.x + q * RZ_NT;
        bv[q] = b < RB_NB;
        const int bi = b % RB_BI, bj = b / RB_BI;
        const int gi = pad_map(i0 - 2 + bi, n0), gj = pad_map(j0 - 2 + bj, n1);
        bc[q] = (ci_t)gi + s1 * (ci_t)gj;
    }
    const ci_t cell = (ci_t)i + s1 * (ci_t)j;
    auto load = [&](T (*v)[RB_NQ], int P) {
        const ci_t pc = s2 * (ci_t)pad_map(P, n2);
#pragma unroll
        for (int q = 0; q < RB_PER; q++)
            if (bv[q]) {
                const ci_t c = bc[q] + pc;
                v[q][0] = ld(u, c);
                v[q][1] = ld(aux + AX::EINT * nc, c);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    v[q][2 + d] = ld(aux + (AX::V0 + d) * nc, c);
                    v[q][5 + d] = ld(u + (LY::B0 + d) * nc, c);
                }
            }
    };
    // own-column window along k: planes P-5..P of rho', e_int, v_0..2, b'_0, b'_1
    T zw[6][7];
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int q = 0; q < 7; q++) zw[p][q] = 0;
    T nx[RB_PER][RB_NQ];
    load(nx, k0 - 2);
    for (int P = k0 - 2; P <= k1 + 2; P++) {
        __syncthreads(); // the previous plane's LDS reads are done
#pragma unroll
        for (int q = 0; q < RB_PER; q++)
            if (bv[q]) {
                const int b = threadIdx.x + q * RZ_NT;
#pragma unroll
                for (int f = 0; f < RB_NQ; f++) sQ[f][b] = nx[q][f];
            }
        if (P + 1 <= k1 + 2) load(nx, P + 1);
        __syncthreads();
        if (!act) continue;
#pragma unroll
        for (int p = 0; p < 5; p++)
#pragma unroll
            for (int q = 0; q < 7; q++) zw[p][q] = zw[p + 1][q];
#pragma unroll
        for (int q = 0; q < 7; q++) zw[5][q] = sQ[q][own];
        if (P >= k0 && P < k1) {
            const ci_t c = cell + s2 * (ci_t)P;
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int st_ = d == 0 ? 1 : RB_BI;
                auto win = [&](int f, T w[6]) {
#pragma unroll
                    for (int p = 0; p < 6; p++) w[p] = sQ[f][own + (p - 2) * st_];
                };
                T w[6];
                win(0, w);
                st(R + LY::prho(d) * nc, c, ratio6<T>(w));
                win(1, w);
                st(R + LY::pe(d) * nc, c, ratio6<T>(w));
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    win(2 + f, w);
                    st(R + LY::pv(f, d) * nc, c, ratio6<T>(w));
                }
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    if (f == d) continue;
                    win(5 + f, w);
                    st(R + LY::pb(f, d) * nc, c, ratio6<T>(w));
                }
            }
        }
        const int O = P - 3; // the window planes O-2..O+3 are in zw
        if (O >= k0 && O < k1) {
            const ci_t c = cell + s2 * (ci_t)O;
            T w[6];
            auto zwin = [&](int q) {
#pragma unroll
                for (int p = 0; p < 6; p++) w[p] = zw[p][q];
            };
            zwin(0);
            st(R + LY::prho(2) * nc, c, ratio6<T>(w));
            zwin(1);
            st(R + LY::pe(2) * nc, c, ratio6<T>(w));
#pragma unroll
            for (int f = 0; f < 3; f++) {
                zwin(2 + f);
                st(R + LY::pv(f, 2) * nc, c, ratio6<T>(w));
            }
#pragma unroll
            for (int f = 0; f < 2; f++) {
                zwin(5 + f);
                st(R + LY::pb(f, 2) * nc, c, ratio6<T>(w));
            }
        }
    }
}

#ifndef SMAUG_ILP_TU // built in smaug_engine.hip
__global__ __launch_bounds__(RZ_NT) void k_ratio_zs_f(Geo g, ZMap zm, const float *__restrict__ u,
                                                       const float *__restrict__ aux,
                                                       float *__restrict__ R)
{
    __shared__ float sQ[RB_NQ][RB_NB];
    ratio_zs_body<float>(g, zm, u, aux, R, sQ);
}

__global__ __launch_bounds__(RZ_NT) void k_ratio_zs_d(Geo g, ZMap zm, const double *__restrict__ u,
                                                       const double *__restrict__ aux,
                                                       double *__restrict__ R)
{
    __shared__ double sQ[RB_NQ][RB_NB];
    ratio_zs_body<double>(g, zm, u, aux, R, sQ);
}
#endif

// ---------------------------------------------------------------------------
// The seven ratio pairs along k alone (prho, pe, pv(0..2), pb(0..1) of
// dim 2), for the source kernel that forms the in-plane pairs itself.  The
// cells, blocks and arithmetic are those of k_ratio_zs: each thread streams
// its own column through planes k0-2 .. k1+2 (mirrored pads along k) with
// the last six planes in registers, and writes the ratios of the plane three
// behind the front.  No LDS: the column is the only neighbourhood, so the
// blocks are 64 x KZ_TJ columns, smaller than k_ratio_zs's, for more blocks
// at a given depth.
// ---------------------------------------------------------------------------
constexpr int KZ_TJ = 4, KZ_NT = RZ_TI * KZ_TJ;
template <typename T>
__device__ __forceinline__ void ratio_kz_body(const Geo &g, const ZMap &zm,
                                              const T *__restrict__ u,
                                              const T *__restrict__ aux, T *__restrict__ R)
{
    typedef Lay<3> LY;
    typedef Aux<3> AX;
    int tx, ty, kc;
    if (!zmap_block(zm, tx, ty, kc)) return;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const size_t nc = (size_t)g.nc;
    const ci_t s1 = (ci_t)g.s1, s2 = (ci_t)g.s2;
    const int i = tx * RZ_TI + threadIdx.x % RZ_TI, j = 1 + ty * KZ_TJ + threadIdx.x / RZ_TI;
    if (i > n0 - 1 || j > n1 - 3) return;
    const int k0 = 1 + kc * zm.kz;
    const int k1 = min(k0 + zm.kz, n2 - 2);
    const ci_t cell = (ci_t)i + s1 * (ci_t)j;
    // rho', e_int, v_0..2, b'_0, b'_1 of one plane
    auto load = [&](T v[7], int P) {
        const ci_t c = cell + s2 * (ci_t)pad_map(P, n2);
        v[0] = ld(u, c);
        v[1] = ld(aux + AX::EINT * nc, c);
#pragma unroll
        for (int d = 0; d < 3; d++) v[2 + d] = ld(aux + (AX::V0 + d) * nc, c);
        v[5] = ld(u + LY::B0 * nc, c);
        v[6] = ld(u + (LY::B0 + 1) * nc, c);
    };
    T zw[6][7];
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int q = 0; q < 7; q++) zw[p][q] = 0;
    T nx[7];
    load(nx, k0 - 2);
    for (int P = k0 - 2; P <= k1 + 2; P++) {
#pragma unroll
        for (int p = 0; p < 5; p++)
#pragma unroll
            for (int q = 0; q < 7; q++) zw[p][q] = zw[p + 1][q];
#pragma unroll
        for (int q = 0; q < 7; q++) zw[5][q] = nx[q];
        if (P + 1 <= k1 + 2) load(nx, P + 1);
        const int O = P - 3; // the window planes O-2..O+3 are in zw
        if (O >= k0 && O < k1) {
            const ci_t c = cell + s2 * (ci_t)O;
            T w[6];
            auto zwin = [&](int q) {
#pragma unroll
                for (int p = 0; p < 6; p++) w[p] = zw[p][q];
            };
            zwin(0);
            st(R + LY::prho(2) * nc, c, ratio6<T>(w));
            zwin(1);
            st(R + LY::pe(2) * nc, c, ratio6<T>(w));
#pragma unroll
            for (int f = 0; f < 3; f++) {
                zwin(2 + f);
                st(R + LY::pv(f, 2) * nc, c, ratio6<T>(w));
            }
#pragma unroll
            for (int f = 0; f < 2; f++) {
                zwin(5 + f);
                st(R + LY::pb(f, 2) * nc, c, ratio6<T>(w));
            }
        }
    }
}

// built in smaug_ilp.hip alone (under its scheduler), which defines SMAUG_ILP_TU
#ifdef SMAUG_ILP_TU
__global__ __launch_bounds__(KZ_NT) void k_ratio_kz_f(Geo g, ZMap zm, const float *__restrict__ u,
                                                       const float *__restrict__ aux,
                                                       float *__restrict__ R)
{
    ratio_kz_body<float>(g, zm, u, aux, R);
}

__global__ __launch_bounds__(KZ_NT) void k_ratio_kz_d(Geo g, ZMap zm, const double *__restrict__ u,
                                                       const double *__restrict__ aux,
                                                       double *__restrict__ R)
{
    ratio_kz_body<double>(g, zm, u, aux, R);
}
#endif

// ---------------------------------------------------------------------------
// Hyperdiffusive sources with the plane staged in LDS.  A block takes a
// 64 x 4 tile of one plane; the state, aux quantities, rho_b and 18 of the
// 21 ratio pairs of the tile and a 1-cell halo are loaded once, coalesced,
// into LDS, and every in-plane neighbour access of hyper_terms reads LDS.
// Accesses to the planes above and below (and the three ratio pairs that
// are only read at the cell itself or below it) go to global memory.
// ---------------------------------------------------------------------------
constexpr int HB_TI = 64, HB_TJ = 4; // the 64 x 4 tiles of make_map / map_block
constexpr int HB_BI = HB_TI + 2, HB_BJ = HB_TJ + 2, HB_NB = HB_BI * HB_BJ;
constexpr int HB_NH = HB_NB - HB_TI * HB_TJ;
constexpr int HB_NR = 18; // ratio pairs staged: all but prho(2), pe(2), pv(2,2)

__host__ __device__ constexpr int hb_rslot(int p)
{
    return (p == 2 || p == 5 || p == 14) ? -1 : p - (p > 2) - (p > 5) - (p > 14);
}
// state field -> LDS slot (rho, m_0..2, b_0..2; e is not read)
__host__ __device__ constexpr int hb_uslot(int f) { return f < 4 ? f : f - 1; }

template <typename T>
struct HyLds {
    HyGlobal<3, T> gm;
    const T (*sU)[HB_NB];
    const T (*sA)[HB_NB];
    const double *sRb;
    const T (*sR)[HB_NB];
    int own;
    __device__ __forceinline__ int bx(int di, int dj) const { return own + di + HB_BI * dj; }
    __device__ __forceinline__ T U(int f, int di, int dj, int dk) const
    {
        return dk == 0 ? sU[hb_uslot(f)][bx(di, dj)] : gm.U(f, di, dj, dk);
    }
    __device__ __forceinline__ T A(int a, int di, int dj, int dk) const
    {
        return dk == 0 ? sA[a - Aux<3>::EINT][bx(di, dj)] : gm.A(a, di, dj, dk);
    }
    __device__ __forceinline__ double rhob(int di, int dj, int dk) const
    {
        return dk == 0 ? sRb[bx(di, dj)] : gm.rhob(di, dj, dk);
    }
    __device__ __forceinline__ T Rp(int p, int di, int dj, int dk) const
    {
        return (dk == 0 && hb_rslot(p) >= 0) ? sR[hb_rslot(p)][bx(di, dj)] : gm.Rp(p, di, dj, dk);
    }
};

// box coordinates of halo cell h of the 66 x 6 box: rows 0 and 5 in full,
// then columns 0 and 65 of rows 1..4
__device__ __forceinline__ void hb_halo(int h, int &bi, int &bj)
{
    if (h < 2 * HB_BI) {
        bj = h < HB_BI ? 0 : HB_BJ - 1;
        bi = h < HB_BI ? h : h - HB_BI;
    } else {
        const int q = h - 2 * HB_BI;
        bj = 1 + (q >> 1);
        bi = (q & 1) ? HB_BI - 1 : 0;
    }
}

template <typename T>
__device__ __forceinline__ void hyper_lds_body(const Geo &g, const TileMap &tm, const Phys &ph,
                                               const T *__restrict__ uin,
                                               const T *__restrict__ R,
                                               const T *__restrict__ aux,
                                               const double *__restrict__ bg,
                                               const unsigned long long *__restrict__ cmb,
                                               T *__restrict__ dUh, T (*sU)[HB_NB],
                                               T (*sA)[HB_NB], double *sRb, T (*sR)[HB_NB])
{
    typedef Lay<3> LY;
    int tx, ty, tk;
    if (!map_block(tm, tx, ty, tk)) return;
    const int n0 = g.n[0], n1 = g.n[1];
    const size_t nc = (size_t)g.nc;
    const ci_t s1 = (ci_t)g.s1, s2 = (ci_t)g.s2;
    const int ti = threadIdx.x & 63, tj = threadIdx.x >> 6;
    const int i0 = 2 + tx * HB_TI, j0 = 2 + ty * HB_TJ, k = 2 + tk;
    const ci_t pc = s2 * (ci_t)k;
    // stage the tile and its 1-cell halo
    auto stage = [&](int bi, int bj) {
        const int gi = i0 - 1 + bi, gj = j0 - 1 + bj;
        if (gi > n0 - 1 || gj > n1 - 1) return;
        const ci_t c = (ci_t)gi + s1 * (ci_t)gj + pc;
        const int b = bj * HB_BI + bi;
        sU[0][b] = ld(uin, c);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            sU[1 + d][b] = ld(uin + (LY::M0 + d) * nc, c);
            sU[4 + d][b] = ld(uin + (LY::B0 + d) * nc, c);
        }
#pragma unroll
        for (int a = 0; a < 4; a++) sA[a][b] = ld(aux + (Aux<3>::EINT + a) * nc, c);
        sRb[b] = ld(bg + nc, c);
#pragma unroll
        for (int p = 0; p < LY::NPAIR; p++)
            if (hb_rslot(p) >= 0) sR[hb_rslot(p)][b] = ld(R + p * nc, c);
    };
    stage(ti + 1, tj + 1);
    if (threadIdx.x < HB_NH) {
        int bi, bj;
        hb_halo(threadIdx.x, bi, bj);
        stage(bi, bj);
    }
    __syncthreads();
    const int i = i0 + ti, j = j0 + tj;
    if (i > n0 - 3 || j > n1 - 3) return;
    HyLds<T> vw;
    vw.gm.uin = uin;
    vw.gm.aux = aux;
    vw.gm.bg = bg;
    vw.gm.R = R;
    vw.gm.nc = nc;
    vw.gm.bgf = nc; // this kernel reads the full background
    vw.gm.bgm = ~0u;
    vw.gm.s1 = s1;
    vw.gm.s2 = s2;
    vw.gm.c = (ci_t)i + s1 * (ci_t)j + pc;
    vw.sU = sU;
    vw.sA = sA;
    vw.sRb = sRb;
    vw.sR = sR;
    vw.own = (tj + 1) * HB_BI + ti + 1;
    T dU[8];
#pragma unroll
    for (int f = 0; f < 8; f++) dU[f] = 0;
    hyper_terms<3, T>(g, ph, vw, cmb, i, j, k, dU);
#pragma unroll
    for (int f = 0; f < 8; f++) st(dUh + f * nc, vw.gm.c, dU[f]);
}

#ifndef SMAUG_ILP_TU // built in smaug_engine.hip
__global__ __launch_bounds__(256) void k_hyper_lds_f(Geo g, TileMap tm, Phys ph,
                                                      const float *__restrict__ uin,
                                                      const float *__restrict__ R,
                                                      const float *__restrict__ aux,
                                                      const double *__restrict__ bg,
                                                      const unsigned long long *__restrict__ cmb,
                                                      float *__restrict__ dUh)
{
    __shared__ float sU[7][HB_NB], sA[4][HB_NB], sR[HB_NR][HB_NB];
    __shared__ double sRb[HB_NB];
    hyper_lds_body<float>(g, tm, ph, uin, R, aux, bg, cmb, dUh, sU, sA, sRb, sR);
}

__global__ __launch_bounds__(256) void k_hyper_lds_d(Geo g, TileMap tm, Phys ph,
                                                      const double *__restrict__ uin,
                                                      const double *__restrict__ R,
                                                      const double *__restrict__ aux,
                                                      const double *__restrict__ bg,
                                                      const unsigned long long *__restrict__ cmb,
                                                      double *__restrict__ dUh)
{
    __shared__ double sU[7][HB_NB], sA[4][HB_NB], sR[HB_NR][HB_NB];
    __shared__ double sRb[HB_NB];
    hyper_lds_body<double>(g, tm, ph, uin, R, aux, bg, cmb, dUh, sU, sA, sRb, sR);
}
#endif

// ---------------------------------------------------------------------------
// Hyperdiffusive sources, z-streamed.  A block owns 32 x 8 columns and walks
// a chunk of output planes.  Per plane P (the "centre") the state, AUX,
// rho_b and the 14 in-plane ratio pairs of the tile + 1-cell halo are staged
// in LDS, and:
//  - the terms along i and j of output P (diagonal, and the off-diagonal
//    pair (i, j)) read LDS;
//  - the k-coupled off-diagonal products X_{k,d}(P) (d in-plane) of the own
//    column are formed from LDS and scattered into outputs P-1 and P+1;
//  - the diagonal k-terms of output P-1 use the own column's values of
//    planes P-2..P, kept in registers;
//  - the products X_{f,k}(P-1) (f in-plane) are formed by the owner of each
//    box column from its planes P-2..P (registers) and shared through LDS;
// and output P-1 is written.  The box cells each thread loads stay fixed
// along the stream, so every history is private to its thread.  The next
// plane's loads are issued before the current plane is processed.  The
// arithmetic of each term is that of hyper_terms; only the summation order
// differs.  k_hyper_zsr forms the 14 in-plane ratio pairs of each plane in
// the kernel, from the state of a wider box staged in LDS, instead of
// loading them from R (which then holds the 7 pairs along k); the ratios
// are bit for bit those of the ratio kernels.
// ---------------------------------------------------------------------------
constexpr int HZ_NIP = 14; // in-plane ratio pairs staged in LDS (every pair with d < 2)

// tiles of 32 x 8 columns; the box of the tile and its 1-cell halo is 34 x 10
constexpr int HZ_TI = 32, HZ_TJ = 8;
constexpr int HZ_BI = HZ_TI + 2, HZ_BJ = HZ_TJ + 2, HZ_NB = HZ_BI * HZ_BJ;
constexpr int HZ_NH = HZ_NB - HZ_TI * HZ_TJ;
static_assert(HZ_TI * HZ_TJ == 256 && HZ_NH <= 256, "one own column and at most one halo column per thread");

// box coordinates of halo cell h of the 34 x 10 box: rows 0 and 9 in full,
// then columns 0 and 33 of rows 1..8
__device__ __forceinline__ void hz_halo(int h, int &bi, int &bj)
{
    if (h < 2 * HZ_BI) {
        bj = h < HZ_BI ? 0 : HZ_BJ - 1;
        bi = h < HZ_BI ? h : h - HZ_BI;
    } else {
        const int q = h - 2 * HZ_BI;
        bj = 1 + (q >> 1);
        bi = (q & 1) ? HZ_BI - 1 : 0;
    }
}

__host__ __device__ constexpr int hz_islot(int p)
{
    // pairs along d = 2 (3-D numbering): 2, 5, 8, 11, 14, 16, 18
    return (p == 2 || p == 5 || p == 8 || p == 11 || p == 14 || p == 16 || p == 18)
               ? -1
               : p - (p > 2) - (p > 5) - (p > 8) - (p > 11) - (p > 14) - (p > 16) - (p > 18);
}

template <typename T>
struct HzCell {
    T r, e, m[3], b[3], v[3];
    double rb;
    T R[21];
};

// In-plane ratios formed in the source kernel (RIN): the state that ratio6
// reads -- rho', e_int, v_0..2, b'_0..2 -- staged on the 38 x 14 box of
// cells i0-3 .. i0+34, j0-3 .. j0+10 (pads mirrored as in k_ratio_zs), which
// holds the 34 x 10 box at offset (2, 2).  The pairs along i are formed on the
// 33 x 10 box cells that read them (box columns 0..32), those along j on the
// 34 x 9 cells (box rows 0..8).
constexpr int QB_BI = HZ_BI + 4, QB_BJ = HZ_BJ + 4, QB_NB = QB_BI * QB_BJ;
constexpr int QB_NX = QB_NB - HZ_NB;       // ratio-box cells outside the 34 x 10 box
constexpr int QB_PX = (QB_NX + 255) / 256; // of them per thread

// ratio-box coordinates of the cells outside the 34 x 10 box: rows 0, 1, 12,
// 13 in full, then columns 0, 1, 36, 37 of rows 2..11
__device__ __forceinline__ void qb_extra(int e, int &qi, int &qj)
{
    if (e < 4 * QB_BI) {
        const int r = e / QB_BI;
        qj = r < 2 ? r : QB_BJ - 4 + r;
        qi = e - r * QB_BI;
    } else {
        const int x = e - 4 * QB_BI;
        qj = 2 + (x >> 2);
        qi = (x & 3) < 2 ? (x & 3) : QB_BI - 4 + (x & 3);
    }
}

// pair a (0..6) of the in-plane pairs along d: rho', e_int, v_0..2, then the
// two b'_f with f != d -- its ratio-box field and its sR slot (hz_islot)
__device__ __forceinline__ int rin_field(int d, int a)
{
    return a < 5 ? a : (a == 6 ? 7 : (d == 0 ? 6 : 5));
}
__device__ __forceinline__ int rin_slot(int d, int a)
{
    return a < 5 ? 2 * a + d : (d == 0 ? 6 + a : (a == 5 ? 10 : 13));
}

// runs of the items along i: the 33 box cells of a row in 3 runs of 11
constexpr int HZ_RL = 11, HZ_NRUN = (HZ_BI - 1) / HZ_RL;
static_assert(HZ_RL * HZ_NRUN == HZ_BI - 1, "the runs cover box columns 0..HZ_BI-2");

// all 21 ratio pairs for the own column (full = true); a halo column needs
// only the in-plane pairs and pv/pb(f, 2), f in-plane.  BGP: the background
// does not vary along k, and h.rb was set once for the column.  RIN: the
// in-plane pairs are formed in the kernel and not loaded; rho', e_int, v and
// b' come from cq (the cell with pads mirrored in-plane, which is c itself
// for a cell on the grid), the rest from c.
template <bool BGP, bool RIN, typename T>
__device__ __forceinline__ void hz_load(HzCell<T> &h, const T *__restrict__ u,
                                        const T *__restrict__ aux, const double *__restrict__ bg,
                                        const T *__restrict__ R, size_t nc, ci_t c, bool full,
                                        ci_t cq = 0)
{
    typedef Lay<3> LY;
    typedef Aux<3> AX;
    if constexpr (!RIN) cq = c;
    h.r = ld(u, cq);
    h.e = ld(aux + AX::EINT * nc, cq);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        h.m[d] = ld(u + (LY::M0 + d) * nc, c);
        h.b[d] = ld(u + (LY::B0 + d) * nc, cq);
        h.v[d] = ld(aux + (AX::V0 + d) * nc, cq);
    }
    if constexpr (!BGP) h.rb = ld(bg + nc, c);
#pragma unroll
    for (int p = 0; p < LY::NPAIR; p++) {
        const bool kh = p == LY::pv(0, 2) || p == LY::pv(1, 2) || p == LY::pb(0, 2) ||
                        p == LY::pb(1, 2);
        const bool need = RIN ? (kh || (full && hz_islot(p) < 0))
                              : (full || hz_islot(p) >= 0 || kh);
        h.R[p] = need ? ld(R + p * nc, c) : (T)0;
    }
}

// LDS view of the centre plane for the in-plane terms (accesses with dk = 0).
// RIN: rho', e_int, v and b' are read from the ratio box sQ, m from sU.
template <typename T, bool RIN>
struct HzView {
    const T (*sU)[HZ_NB]; // rho', m_0..2, b'_0..2 (RIN: m_0..2)
    const T (*sA)[HZ_NB]; // e_int, v_0..2 (not with RIN)
    const T (*sQ)[QB_NB]; // RIN: rho', e_int, v_0..2, b'_0..2
    const T *sRf; // full density rho' + rho_b (formed in double, stored in T)
    const T (*sR)[HZ_NB];
    int own, ownq;
    __device__ __forceinline__ int bx(int di, int dj) const { return own + di + HZ_BI * dj; }
    __device__ __forceinline__ int qx(int di, int dj) const { return ownq + di + QB_BI * dj; }
    __device__ __forceinline__ T U(int f, int di, int dj) const
    {
        if constexpr (RIN) {
            if (f == Lay<3>::RHO) return sQ[0][qx(di, dj)];
            if (f >= Lay<3>::B0) return sQ[5 + f - Lay<3>::B0][qx(di, dj)];
            return sU[f - Lay<3>::M0][bx(di, dj)];
        } else {
            return sU[hb_uslot(f)][bx(di, dj)];
        }
    }
    __device__ __forceinline__ T A(int a, int di, int dj) const
    {
        if constexpr (RIN)
            return sQ[1 + a - Aux<3>::EINT][qx(di, dj)];
        else
            return sA[a - Aux<3>::EINT][bx(di, dj)];
    }
    __device__ __forceinline__ T rhof(int di, int dj) const { return sRf[bx(di, dj)]; }
    __device__ __forceinline__ T Rp(int p, int di, int dj) const { return sR[hz_islot(p)][bx(di, dj)]; }
};

// BGP: bg is a background stored as one k-plane (field stride s2).  RIN:
// the in-plane ratio pairs are formed here from the ratio box sQ (R holds
// the pairs along k only); sU then holds m_0..2 and sA is not used.
template <bool BGP, bool RIN, typename T>
__device__ __forceinline__ void hyper_zs_body(const Geo &g, const ZMap &zm, const Phys &ph,
                                              const T *__restrict__ uin,
                                              const T *__restrict__ R,
                                              const T *__restrict__ aux,
                                              const double *__restrict__ bg,
                                              const unsigned long long *__restrict__ cmb,
                                              T *__restrict__ dUh, T (*sU)[HZ_NB],
                                              T (*sA)[HZ_NB], T *sRf, T (*sR)[HZ_NB],
                                              T (*sX)[HZ_NB], T *sIz, T (*sCx)[HZ_TI],
                                              T (*sCy)[HZ_TJ], T (*sCz)[ZS_KZMAX],
                                              T (*sKx)[HZ_TI], T (*sKy)[HZ_TJ], T (*sKz)[ZS_KZMAX],
                                              T (*sQ)[QB_NB])
{
    typedef Lay<3> LY;
    typedef Aux<3> AX;
    int tx, ty, kc;
    if (!zmap_block(zm, tx, ty, kc)) return;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const size_t nc = (size_t)g.nc;
    const ci_t s1 = (ci_t)g.s1, s2 = (ci_t)g.s2;
    const int ti = threadIdx.x % HZ_TI, tj = threadIdx.x / HZ_TI;
    const int i0 = 2 + tx * HZ_TI, j0 = 2 + ty * HZ_TJ;
    const int i = i0 + ti, j = j0 + tj;
    const bool act = i <= n0 - 3 && j <= n1 - 3;
    const int k0 = 2 + kc * zm.kz;
    const int k1 = min(k0 + zm.kz, n2 - 2);
    double cmax[3];
#pragma unroll
    for (int d = 0; d < 3; d++) cmax[d] = __longlong_as_double((long long)cmb[d]);
    // Metric factors, read once per block: 1/dx of the own cell in-plane
    // into registers, the rest into LDS tables per column, row and output
    // plane (a global load inside the plane loop would be waited for
    // together with the prefetched next plane, as vector memory operations
    // retire in order).  nu = (dx cmax) chyp R as in hyper_terms.
    T id[2] = {0, 0};
    // 1/dx, or 0 where the gradients of a cell along that axis are zeroed
    // (zero1): the gradients below take their zeroing from this factor
    T gx[2] = {0, 0};
    if (act) {
        id[0] = (T)g.idx[0][i];
        id[1] = (T)g.idx[1][j];
        gx[0] = zero1(g.per[0], i, n0) ? (T)0 : id[0];
        gx[1] = zero1(g.per[1], j, n1) ? (T)0 : id[1];
    }
    // 1/(2 dx) of the outer derivative of the off-diagonal products, 0 where
    // the reference leaves that derivative out (zero1 along its axis)
    const T hf[2] = {gx[0] / (T)2, gx[1] / (T)2};
    // Per field slot q (rho, e, m_0..2, b_0..2) the factor of the diagonal
    // terms of a cell along an axis: (dx cmax chyp_q), rounded to T, times
    // 1/dx of the gradients (0 where zero1) and 1/dx of the divergence --
    // sCx per column along i, sCy per row along j, sCz per output plane
    // along k (slots up to b_1).  The off-diagonal terms take nu/R alone,
    // (dx cmax chyp) rounded to T: sKx / sKy for m and b of the other
    // in-plane axis and of k (slots m_o, b_o, m_2, b_2), sKz for m_0, m_1,
    // b_0, b_1 per output plane.
    constexpr int CF[8] = {LY::RHO, LY::EN, LY::M0, LY::M0 + 1, LY::M0 + 2, LY::B0, LY::B0 + 1, LY::B0 + 2};
    for (int t = threadIdx.x; t < k1 - k0; t += 256) {
        const int kk = k0 + t;
        const double dz = g.dx[2][kk] * cmax[2];
        const T izt = (T)g.idx[2][kk];
        const T s = (zero1(g.per[2], kk, n2) ? (T)0 : izt) * izt;
        sIz[t] = izt;
#pragma unroll
        for (int q = 0; q < 7; q++) sCz[q][t] = (T)(dz * ph.chyp[CF[q]]) * s;
#pragma unroll
        for (int f = 0; f < 2; f++) {
            sKz[f][t] = (T)(dz * ph.chyp[LY::M0 + f]);
            sKz[2 + f][t] = (T)(dz * ph.chyp[LY::B0 + f]);
        }
    }
    if (threadIdx.x < HZ_TI + HZ_TJ) {
        const bool alongi = threadIdx.x < HZ_TI;
        const int a = alongi ? 0 : 1;
        const int x = alongi ? i0 + threadIdx.x : j0 + (threadIdx.x - HZ_TI);
        const bool in = x <= g.n[a] - 3;
        const double dd = in ? g.dx[a][x] * cmax[a] : 0.0;
        const T ix = in ? (T)g.idx[a][x] : (T)0;
        const T s = (in && !zero1(g.per[a], x, g.n[a]) ? ix : (T)0) * ix;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const T c = (T)(dd * ph.chyp[CF[q]]) * s;
            if (alongi)
                sCx[q][threadIdx.x] = c;
            else
                sCy[q][threadIdx.x - HZ_TI] = c;
        }
        const int o = 1 - a; // the other in-plane axis
        const int KF[4] = {LY::M0 + o, LY::B0 + o, LY::M0 + 2, LY::B0 + 2};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const T kq = (T)(dd * ph.chyp[KF[q]]);
            if (alongi)
                sKx[q][threadIdx.x] = kq;
            else
                sKy[q][threadIdx.x - HZ_TI] = kq;
        }
    }
    __syncthreads();
    // box cells of this thread: 0 = its own column, 1 = a halo column (if any)
    int bb[2], bi_[2], bj_[2];
    bool bv[2];
    ci_t bc[2];
    // RIN: whether the box position exists (on the grid or not), its index
    // in the ratio box and its cell with pads mirrored in-plane
    bool pq[2];
    int bq[2];
    ci_t cqa[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        int bi, bj;
        if (q == 0) {
            bi = ti + 1;
            bj = tj + 1;
            bv[q] = true;
        } else {
            bv[q] = threadIdx.x < HZ_NH;
            bi = bj = 0;
            if (bv[q]) hz_halo(threadIdx.x, bi, bj);
        }
        pq[q] = bv[q];
        const int gi = i0 - 1 + bi, gj = j0 - 1 + bj;
        bv[q] = bv[q] && gi <= n0 - 1 && gj <= n1 - 1;
        bi_[q] = gi;
        bj_[q] = gj;
        bb[q] = bj * HZ_BI + bi;
        bc[q] = (ci_t)gi + s1 * (ci_t)gj;
        bq[q] = (bj + 2) * QB_BI + bi + 2;
        cqa[q] = (ci_t)pad_map(gi, n0) + s1 * (ci_t)pad_map(gj, n1);
    }
    // RIN: the ratio-box cells outside the 34 x 10 box that this thread loads
    bool ev[QB_PX];
    int eq[QB_PX];
    ci_t ce[QB_PX];
    T xq[QB_PX][8];
#pragma unroll
    for (int x = 0; x < QB_PX; x++) {
        const int e = threadIdx.x + 256 * x;
        ev[x] = RIN && e < QB_NX;
        int qi = 0, qj = 0;
        if (ev[x]) qb_extra(e, qi, qj);
        eq[x] = qj * QB_BI + qi;
        ce[x] = (ci_t)pad_map(i0 - 3 + qi, n0) + s1 * (ci_t)pad_map(j0 - 3 + qj, n1);
#pragma unroll
        for (int f = 0; f < 8; f++) xq[x][f] = 0;
    }
    auto xload = [&](int P) {
        typedef Aux<3> AX;
#pragma unroll
        for (int x = 0; x < QB_PX; x++)
            if (ev[x]) {
                const ci_t c = ce[x] + s2 * (ci_t)P;
                xq[x][0] = ld(uin, c);
                xq[x][1] = ld(aux + AX::EINT * nc, c);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    xq[x][2 + d] = ld(aux + (AX::V0 + d) * nc, c);
                    xq[x][5 + d] = ld(uin + (LY::B0 + d) * nc, c);
                }
            }
    };
    const int own = bb[0];
    HzView<T, RIN> vw;
    vw.sU = sU;
    vw.sA = sA;
    vw.sQ = sQ;
    vw.sRf = sRf;
    vw.sR = sR;
    vw.own = own;
    vw.ownq = bq[0];

    // per box column, planes P-1 and P-2 (index 0 and 1): v_0, v_1, b_0, b_1,
    // and the k-pairs pv(0,2), pv(1,2), pb(0,2), pb(1,2); rho, m_2, b_2 of P-1
    T hv[2][2][4], hr[2][2][4], hn[2][3];
    // own column, planes P-1 and P-2: the diagonal k-term inputs not already
    // in hv/hr (rho', e_int, v_2, m_0..2; prho(2), pe(2), pv(2,2); rho_b)
    T ow[2][6], oz[2][3];
    T of2 = 0; // full density of the own column, plane P-2
    // pending outputs P-1, P, P+1
    T rA[8], rB[8], rC[8];
#pragma unroll
    for (int f = 0; f < 8; f++) rA[f] = rB[f] = rC[f] = 0;
#pragma unroll
    for (int q = 0; q < 2; q++) {
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int f = 0; f < 4; f++) hv[q][t][f] = hr[q][t][f] = 0;
#pragma unroll
        for (int f = 0; f < 3; f++) hn[q][f] = 0;
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
#pragma unroll
        for (int f = 0; f < 6; f++) ow[t][f] = 0;
#pragma unroll
        for (int f = 0; f < 3; f++) oz[t][f] = 0;
    }
    const bool zper = g.per[2];

    HzCell<T> cur[2];
    if constexpr (BGP) {
#pragma unroll
        for (int q = 0; q < 2; q++) cur[q].rb = ld(bg + (size_t)s2, bv[q] ? bc[q] : 0u);
    }
    if constexpr (RIN) {
        xload(k0 - 1);
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (pq[q])
                hz_load<BGP, RIN, T>(cur[q], uin, aux, bg, R, nc,
                                     (bv[q] ? bc[q] : 0u) + s2 * (ci_t)(k0 - 1), q == 0,
                                     cqa[q] + s2 * (ci_t)(k0 - 1));
    } else {
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (bv[q])
                hz_load<BGP, RIN, T>(cur[q], uin, aux, bg, R, nc, bc[q] + s2 * (ci_t)(k0 - 1),
                                     q == 0);
    }

    // The output completed in one iteration is stored at the head of the
    // next one, behind the loads of the own column's plane, which the head
    // then waits for without waiting for the stores.  Every lane stores on
    // every path: without an output it writes 0 into cell 0, a corner ghost
    // cell that the stage update does not read.
    T hs[8];
#pragma unroll
    for (int f = 0; f < 8; f++) hs[f] = 0;
    ci_t hco = 0;
    for (int P = k0 - 1; P <= k1; P++) {
#pragma unroll
        for (int f = 0; f < 8; f++) st(dUh + f * nc, hco, hs[f]);
        __syncthreads(); // the previous plane's LDS reads are done
        if constexpr (RIN) {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (pq[q]) {
                    const int b = bq[q];
                    sQ[0][b] = cur[q].r;
                    sQ[1][b] = cur[q].e;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        sQ[2 + d][b] = cur[q].v[d];
                        sQ[5 + d][b] = cur[q].b[d];
                    }
                }
                if (bv[q]) {
                    const int b = bb[q];
#pragma unroll
                    for (int d = 0; d < 3; d++) sU[d][b] = cur[q].m[d];
                    sRf[b] = (T)((double)cur[q].r + cur[q].rb);
                }
            }
#pragma unroll
            for (int x = 0; x < QB_PX; x++)
                if (ev[x])
#pragma unroll
                    for (int f = 0; f < 8; f++) sQ[f][eq[x]] = xq[x][f];
        } else {
#pragma unroll
            for (int q = 0; q < 2; q++)
                if (bv[q]) {
                    const int b = bb[q];
                    sU[0][b] = cur[q].r;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        sU[1 + d][b] = cur[q].m[d];
                        sU[4 + d][b] = cur[q].b[d];
                        sA[1 + d][b] = cur[q].v[d];
                    }
                    sA[0][b] = cur[q].e;
                    sRf[b] = (T)((double)cur[q].r + cur[q].rb);
#pragma unroll
                    for (int p = 0; p < LY::NPAIR; p++)
                        if (hz_islot(p) >= 0) sR[hz_islot(p)][b] = cur[q].R[p];
                }
        }
        // X_{f,k}(P-1) of this thread's box columns, from planes P-2, P-1
        // (registers) and P (cur)
        const int O = P - 1;
        const bool outO = O >= k0 && O < k1;
        if (outO) {
            // idz / 2, or 0 where the k-gradients of plane O are zeroed
            const T hz = zero1(zper, O, n2) ? (T)0 : sIz[O - k0] / (T)2;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (!bv[q]) continue;
                const int xq[2] = {bi_[q], bj_[q]};
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    // the products vanish off the interior along f (the
                    // factor is 0 there; every input is finite: loaded
                    // cells, and ratios that were written or are the zeros
                    // the buffer was cleared to)
                    const T hzf = (xq[f] >= 2 && xq[f] <= g.n[f] - 3) ? hz : (T)0;
                    // ca = p - e_k (plane P-2), cb = p + e_k (plane P)
                    const T dv = (cur[q].v[f] - hv[q][1][f]) * hzf;
                    const T db = (cur[q].b[f] - hv[q][1][2 + f]) * hzf;
                    const T Km = sKz[f][O - k0];
                    const T sv = Km * (hr[q][0][f] + hr[q][1][f]);
                    const T G = dv * sv / (T)4;
                    const T Xr = hn[q][0] * G;
                    const T Xm = G * hn[q][1];
                    const T Kb = sKz[2 + f][O - k0];
                    const T sb = Kb * (hr[q][0][2 + f] + hr[q][1][2 + f]);
                    const T Hb = db * sb / (T)2;
                    const T He = Hb * hn[q][2];
                    sX[4 * f + 0][bb[q]] = Xr;
                    sX[4 * f + 1][bb[q]] = Xm;
                    sX[4 * f + 2][bb[q]] = Hb;
                    sX[4 * f + 3][bb[q]] = He;
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // ---- diagonal k-terms of output P-1: own column planes P-2, P-1
        // (registers) and P (cur)
        if (act && outO) {
            // plane t (0: P-1, 1: P-2) or P (t = 2) of quantity k: rho', e_int,
            // v_0..2, m_0..2, b'_0, b'_1
            auto W = [&](int t, int k) -> T {
                if (t == 2) {
                    if (k == 0) return cur[0].r;
                    if (k == 1) return cur[0].e;
                    if (k < 5) return cur[0].v[k - 2];
                    if (k < 8) return cur[0].m[k - 5];
                    return cur[0].b[k - 8];
                }
                if (k == 0) return ow[t][0];
                if (k == 1) return ow[t][1];
                if (k == 2 || k == 3) return hv[0][t][k - 2];
                if (k == 4) return ow[t][2];
                if (k < 8) return ow[t][3 + k - 5];
                return hv[0][t][2 + k - 8];
            };
            // ratio pair t of the k-set (prho, pe, pv(0..2), pb(0..1) along k)
            auto Z = [&](int t, int k) -> T {
                if (k == 0 || k == 1) return oz[t][k];
                if (k == 2 || k == 3) return hr[0][t][k - 2];
                if (k == 4) return oz[t][2];
                return hr[0][t][2 + k - 5];
            };
            // the factors of the plane (as sCx along i)
            const T *cz = &sCz[0][O - k0];
            auto C = [&](int q) -> T { return cz[q * ZS_KZMAX]; };
            T pR, pL;
            const T rc = W(0, 0);
            pR = Z(0, 0) * (W(2, 0) - rc);
            pL = Z(1, 0) * (rc - W(1, 0));
            rA[LY::RHO] += (pR - pL) * C(0);
            pR = Z(0, 1) * (W(2, 1) - W(0, 1));
            pL = Z(1, 1) * (W(0, 1) - W(1, 1));
            rA[LY::EN] += (pR - pL) * C(1);
            // twice the mean full density over the faces (planes P-1, P and
            // P-1, P-2)
            const T fc = hn[0][0];
            const T rR = fc + (T)((double)cur[0].r + cur[0].rb);
            const T rL = fc + of2;
#pragma unroll
            for (int f = 0; f < 3; f++) {
                const T vc = W(0, 2 + f);
                pR = Z(0, 2 + f) * (W(2, 2 + f) - vc);
                pL = Z(1, 2 + f) * (vc - W(1, 2 + f));
                const T mc = W(0, 5 + f);
                const T mR = mc + W(2, 5 + f), mL = mc + W(1, 5 + f);
                const T ck = C(2 + f) * ((f == 2) ? (T)0.5 : (T)0.25);
                rA[LY::M0 + f] += (rR * pR - rL * pL) * ck;
                rA[LY::EN] += (mR * pR - mL * pL) * ck;
            }
#pragma unroll
            for (int f = 0; f < 2; f++) {
                const T bc_ = W(0, 8 + f), br = W(2, 8 + f), bl = W(1, 8 + f);
                pR = Z(0, 5 + f) * (br - bc_);
                pL = Z(1, 5 + f) * (bc_ - bl);
                const T cb = C(5 + f);
                rA[LY::B0 + f] += (pR - pL) * cb;
                rA[LY::EN] += ((bc_ + br) * pR - (bc_ + bl) * pL) * (cb * (T)0.5);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        // the histories move up one plane
        of2 = hn[0][0];
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (bv[q]) {
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    hv[q][1][f] = hv[q][0][f];
                    hv[q][1][2 + f] = hv[q][0][2 + f];
                    hv[q][0][f] = cur[q].v[f];
                    hv[q][0][2 + f] = cur[q].b[f];
                    hr[q][1][f] = hr[q][0][f];
                    hr[q][1][2 + f] = hr[q][0][2 + f];
                    hr[q][0][f] = cur[q].R[LY::pv(f, 2)];
                    hr[q][0][2 + f] = cur[q].R[LY::pb(f, 2)];
                }
                hn[q][0] = (T)((double)cur[q].r + cur[q].rb);
                hn[q][1] = cur[q].m[2];
                hn[q][2] = cur[q].b[2];
            }
        if (act) {
#pragma unroll
            for (int f = 0; f < 6; f++) ow[1][f] = ow[0][f];
#pragma unroll
            for (int f = 0; f < 3; f++) oz[1][f] = oz[0][f];
            ow[0][0] = cur[0].r;
            ow[0][1] = cur[0].e;
            ow[0][2] = cur[0].v[2];
#pragma unroll
            for (int f = 0; f < 3; f++) ow[0][3 + f] = cur[0].m[f];
            oz[0][0] = cur[0].R[LY::prho(2)];
            oz[0][1] = cur[0].R[LY::pe(2)];
            oz[0][2] = cur[0].R[LY::pv(2, 2)];
        }
        // The next plane's loads, in flight through the rest of this plane.
        // The own column's come last and are issued on every path (the last
        // iteration reloads its plane, a column off the grid reads cell 0),
        // so that the compiler's waits for earlier loads count past them.
        if constexpr (RIN) {
            if (P + 1 <= k1) {
                xload(P + 1);
                if (pq[1])
                    hz_load<BGP, RIN, T>(cur[1], uin, aux, bg, R, nc,
                                         (bv[1] ? bc[1] : 0u) + s2 * (ci_t)(P + 1), false,
                                         cqa[1] + s2 * (ci_t)(P + 1));
            }
            hz_load<BGP, RIN, T>(cur[0], uin, aux, bg, R, nc,
                                 bv[0] ? bc[0] + s2 * (ci_t)min(P + 1, k1) : 0u, true,
                                 cqa[0] + s2 * (ci_t)min(P + 1, k1));
        } else {
            if (P + 1 <= k1 && bv[1])
                hz_load<BGP, RIN, T>(cur[1], uin, aux, bg, R, nc, bc[1] + s2 * (ci_t)(P + 1),
                                     false);
            hz_load<BGP, RIN, T>(cur[0], uin, aux, bg, R, nc,
                                 bv[0] ? bc[0] + s2 * (ci_t)min(P + 1, k1) : 0u, true);
        }
        __syncthreads();
        if constexpr (RIN) {
            // the in-plane ratio pairs of plane P on the box cells that read
            // them, from the ratio box; 0 where k_ratio_zs writes none (rows
            // above n1-3, planes above n2-3, cells off the grid).  Items:
            // pair a of each of the 34 box columns along j (the column's 14
            // values give the pairs of its cells in box rows 0..8), then
            // pair a of each of the 10 box rows along i, in runs of HZ_RL
            // cells (box columns 0..32).  Iteration 0 of the 256 threads
            // takes the 238 items along j and the first 18 along i, so one
            // wave runs both kinds; iteration 1 the other 192.
            const bool zk = P > n2 - 3;
            constexpr int NJ = 7 * HZ_BI, NI = 7 * HZ_BJ * HZ_NRUN;
            for (int u = threadIdx.x; u < NJ + NI; u += 256) {
                if (u < NJ) {
                    const int a = u / HZ_BI, bi = u - a * HZ_BI;
                    T r[HZ_BJ - 1];
                    ratio_run<HZ_BJ - 1, T>(sQ[rin_field(1, a)] + bi + 2, QB_BI, r);
                    // cell c at gj = j0 - 1 + c of the column gi = i0 - 1 + bi
                    const int lim = zk || i0 - 1 + bi > n0 - 1 ? -1 : n1 - 2 - j0;
                    ratio_store<HZ_BJ - 1, T>(sR[rin_slot(1, a)] + bi, HZ_BI, r, lim);
                } else {
                    const int x = u - NJ;
                    const int a = x / (HZ_BJ * HZ_NRUN);
                    const int y = x - a * (HZ_BJ * HZ_NRUN);
                    const int bj = y / HZ_NRUN, bi = HZ_RL * (y - bj * HZ_NRUN);
                    T r[HZ_RL];
                    ratio_run<HZ_RL, T>(sQ[rin_field(0, a)] + (bj + 2) * QB_BI + bi, 1, r);
                    // cell c at gi = i0 - 1 + bi + c of the row gj = j0 - 1 + bj
                    const int lim = zk || j0 - 1 + bj > n1 - 3 ? -1 : n0 - i0 - bi;
                    ratio_store<HZ_RL, T>(sR[rin_slot(0, a)] + bj * HZ_BI + bi, 1, r, lim);
                }
            }
            __syncthreads();
        }
        if (act) {
            const int xs[2] = {i, j};
            // ---- in-plane terms of output P
            if (P >= k0 && P < k1) {
                const T rc = vw.U(LY::RHO, 0, 0);
                const T fc = vw.rhof(0, 0);
                const T eic = vw.A(AX::EINT, 0, 0);
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    const int ri = d == 0, rj = d == 1;
                    const T hxd = gx[d] / (T)2;
                    auto K = [&](int q) -> T { return d == 0 ? sKx[q][ti] : sKy[q][tj]; };
                    // nu/R of field slot q times the gradient and divergence
                    // factors of this cell along d (0 where zero1)
                    auto C = [&](int q) -> T { return d == 0 ? sCx[q][ti] : sCy[q][tj]; };
                    T pR, pL;
                    pR = vw.Rp(LY::prho(d), 0, 0) * (vw.U(LY::RHO, ri, rj) - rc);
                    pL = vw.Rp(LY::prho(d), -ri, -rj) * (rc - vw.U(LY::RHO, -ri, -rj));
                    rB[LY::RHO] += (pR - pL) * C(0);
                    pR = vw.Rp(LY::pe(d), 0, 0) * (vw.A(AX::EINT, ri, rj) - eic);
                    pL = vw.Rp(LY::pe(d), -ri, -rj) * (eic - vw.A(AX::EINT, -ri, -rj));
                    rB[LY::EN] += (pR - pL) * C(1);
                    // twice the mean full density over the two faces
                    const T rR = fc + vw.rhof(ri, rj);
                    const T rL = fc + vw.rhof(-ri, -rj);
#pragma unroll
                    for (int f = 0; f < 3; f++) {
                        const T vc = vw.A(AX::V0 + f, 0, 0);
                        pR = vw.Rp(LY::pv(f, d), 0, 0) * (vw.A(AX::V0 + f, ri, rj) - vc);
                        pL = vw.Rp(LY::pv(f, d), -ri, -rj) * (vc - vw.A(AX::V0 + f, -ri, -rj));
                        const T mc = vw.U(LY::M0 + f, 0, 0);
                        const T mR = mc + vw.U(LY::M0 + f, ri, rj);
                        const T mL = mc + vw.U(LY::M0 + f, -ri, -rj);
                        // the face means' 1/2 times 1 (f == d) or 1/2
                        const T ck = C(2 + f) * ((f == d) ? (T)0.5 : (T)0.25);
                        rB[LY::M0 + f] += (rR * pR - rL * pL) * ck;
                        rB[LY::EN] += (mR * pR - mL * pL) * ck;
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int f = 0; f < 3; f++) {
                        if (f == d) continue;
                        const T bc_ = vw.U(LY::B0 + f, 0, 0);
                        const T br = vw.U(LY::B0 + f, ri, rj), bl = vw.U(LY::B0 + f, -ri, -rj);
                        pR = vw.Rp(LY::pb(f, d), 0, 0) * (br - bc_);
                        pL = vw.Rp(LY::pb(f, d), -ri, -rj) * (bc_ - bl);
                        const T cb = C(5 + f);
                        rB[LY::B0 + f] += (pR - pL) * cb;
                        rB[LY::EN] += ((bc_ + br) * pR - (bc_ + bl) * pL) * (cb * (T)0.5);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    // off-diagonal pair (f, d) with f the other in-plane dim
                    const int f = 1 - d;
                    const int xf = xs[f], nf = g.n[f];
                    {
                        T Xr[2], Xm[2], Hb[2], He[2];
#pragma unroll
                        for (int side = 0; side < 2; side++) {
                            const int sg = side ? 1 : -1;
                            const int ni = f == 0 ? sg : 0, nj = f == 1 ? sg : 0;
                            // nu vanishes off the interior: factor 0 there
                            // (inputs finite, as for the k-products)
                            const int xfn = xf + sg;
                            const T hs = (xfn < 2 || xfn > nf - 3) ? (T)0 : hxd;
                            const int ai = ni - ri, aj = nj - rj, bi = ni + ri, bj = nj + rj;
                            const T dv = (vw.A(AX::V0 + f, bi, bj) - vw.A(AX::V0 + f, ai, aj)) * hs;
                            const T db = (vw.U(LY::B0 + f, bi, bj) - vw.U(LY::B0 + f, ai, aj)) * hs;
                            {
                                const T s2_ = K(0) * (vw.Rp(LY::pv(f, d), ni, nj) + vw.Rp(LY::pv(f, d), ai, aj));
                                const T G = dv * s2_ / (T)4;
                                const T rhon = vw.rhof(ni, nj);
                                Xr[side] = rhon * G;
                                Xm[side] = G * vw.U(LY::M0 + d, ni, nj);
                            }
                            {
                                const T s2_ = K(1) * (vw.Rp(LY::pb(f, d), ni, nj) + vw.Rp(LY::pb(f, d), ai, aj));
                                const T H = db * s2_ / (T)2;
                                Hb[side] = H;
                                He[side] = H * vw.U(LY::B0 + d, ni, nj);
                            }
                        }
                        const T h = hf[f];
                        rB[LY::M0 + d] += (Xr[1] - Xr[0]) * h;
                        rB[LY::EN] += (Xm[1] - Xm[0]) * h;
                        rB[LY::B0 + d] -= (Hb[1] - Hb[0]) * h;
                        rB[LY::EN] -= (He[1] - He[0]) * h;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        // ---- X_{k,d}(P) of the own column into outputs P-1 (+) and P+1 (-)
            if (P >= 2 && P <= n2 - 3) {
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    const int ri = d == 0, rj = d == 1;
                    const T hxd = gx[d] / (T)2;
                    const T dv = (vw.A(AX::V0 + 2, ri, rj) - vw.A(AX::V0 + 2, -ri, -rj)) * hxd;
                    const T db = (vw.U(LY::B0 + 2, ri, rj) - vw.U(LY::B0 + 2, -ri, -rj)) * hxd;
                    const T Km = d == 0 ? sKx[2][ti] : sKy[2][tj];
                    const T sv = Km * (vw.Rp(LY::pv(2, d), 0, 0) + vw.Rp(LY::pv(2, d), -ri, -rj));
                    const T G = dv * sv / (T)4;
                    const T Xr = vw.rhof(0, 0) * G, Xm = G * vw.U(LY::M0 + d, 0, 0);
                    const T Kb = d == 0 ? sKx[3][ti] : sKy[3][tj];
                    const T sb = Kb * (vw.Rp(LY::pb(2, d), 0, 0) + vw.Rp(LY::pb(2, d), -ri, -rj));
                    const T H = db * sb / (T)2;
                    const T He = H * vw.U(LY::B0 + d, 0, 0);
                    // output P-1 sees this column as its c + e_k side
                    if (P - 1 >= k0 && P - 1 < k1 && !zero1(zper, P - 1, n2)) {
                        const T h = sIz[P - 1 - k0] / (T)2;
                        rA[LY::M0 + d] += Xr * h;
                        rA[LY::EN] += Xm * h;
                        rA[LY::B0 + d] -= H * h;
                        rA[LY::EN] -= He * h;
                    }
                    // output P+1 sees it as its c - e_k side
                    if (P + 1 >= k0 && P + 1 < k1 && !zero1(zper, P + 1, n2)) {
                        const T h = sIz[P + 1 - k0] / (T)2;
                        rC[LY::M0 + d] -= Xr * h;
                        rC[LY::EN] -= Xm * h;
                        rC[LY::B0 + d] += H * h;
                        rC[LY::EN] += He * h;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        // the products X_{f,k}(P-1) were written to LDS before the barrier
        // that follows the staging, and are next rewritten after the first
        // barrier of the next plane, so no barrier is needed here
        if (act && outO) {
            // every box cell read here was written this plane: the
            // neighbours of an active column lie on the grid
#pragma unroll
            for (int f = 0; f < 2; f++) {
                const int sf = f == 0 ? 1 : HZ_BI;
                const T h = hf[f];
                const T Xr = sX[4 * f + 0][own + sf] - sX[4 * f + 0][own - sf];
                const T Xm = sX[4 * f + 1][own + sf] - sX[4 * f + 1][own - sf];
                const T Hb = sX[4 * f + 2][own + sf] - sX[4 * f + 2][own - sf];
                const T He = sX[4 * f + 3][own + sf] - sX[4 * f + 3][own - sf];
                rA[LY::M0 + 2] += Xr * h;
                rA[LY::EN] += Xm * h;
                rA[LY::B0 + 2] -= Hb * h;
                rA[LY::EN] -= He * h;
            }
        }
        const bool wr = act && outO;
        hco = wr ? bc[0] + s2 * (ci_t)O : 0u;
#pragma unroll
        for (int f = 0; f < 8; f++) hs[f] = wr ? rA[f] : (T)0;
        // advance the output ring and the own-column history
#pragma unroll
        for (int f = 0; f < 8; f++) {
            rA[f] = rB[f];
            rB[f] = rC[f];
            rC[f] = 0;
        }
    }
#pragma unroll
    for (int f = 0; f < 8; f++) st(dUh + f * nc, hco, hs[f]);
}

template <bool BGP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
k_hyper_zs_f(Geo g, ZMap zm, Phys ph, const float *__restrict__ uin, const float *__restrict__ R,
             const float *__restrict__ aux, const double *__restrict__ bg,
             const unsigned long long *__restrict__ cmb, float *__restrict__ dUh)
{
    __shared__ float sU[7][HZ_NB], sA[4][HZ_NB], sR[HZ_NIP][HZ_NB], sX[8][HZ_NB];
    __shared__ float sRf[HZ_NB];
    __shared__ float sIz[ZS_KZMAX]; // 1/dz of the output planes
    __shared__ float sCx[8][HZ_TI], sCy[8][HZ_TJ], sCz[7][ZS_KZMAX];
    __shared__ float sKx[4][HZ_TI], sKy[4][HZ_TJ], sKz[4][ZS_KZMAX];
    hyper_zs_body<BGP, false, float>(g, zm, ph, uin, R, aux, bg, cmb, dUh, sU, sA, sRf, sR, sX,
                                     sIz, sCx, sCy, sCz, sKx, sKy, sKz, nullptr);
}

// the same sources with the in-plane ratio pairs formed in the kernel from
// the ratio box; R holds the pairs along k (k_ratio_kz_f)
template <bool BGP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
k_hyper_zsr_f(Geo g, ZMap zm, Phys ph, const float *__restrict__ uin, const float *__restrict__ R,
              const float *__restrict__ aux, const double *__restrict__ bg,
              const unsigned long long *__restrict__ cmb, float *__restrict__ dUh)
{
    __shared__ float sM[3][HZ_NB], sR[HZ_NIP][HZ_NB], sX[8][HZ_NB], sQ[8][QB_NB];
    __shared__ float sRf[HZ_NB];
    __shared__ float sIz[ZS_KZMAX];
    __shared__ float sCx[8][HZ_TI], sCy[8][HZ_TJ], sCz[7][ZS_KZMAX];
    __shared__ float sKx[4][HZ_TI], sKy[4][HZ_TJ], sKz[4][ZS_KZMAX];
    hyper_zs_body<BGP, true, float>(g, zm, ph, uin, R, aux, bg, cmb, dUh, sM, nullptr, sRf, sR,
                                    sX, sIz, sCx, sCy, sCz, sKx, sKy, sKz, sQ);
}

template <bool BGP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void
k_hyper_zs_d(Geo g, ZMap zm, Phys ph, const double *__restrict__ uin,
             const double *__restrict__ R, const double *__restrict__ aux,
             const double *__restrict__ bg, const unsigned long long *__restrict__ cmb,
             double *__restrict__ dUh)
{
    __shared__ double sU[7][HZ_NB], sA[4][HZ_NB], sR[HZ_NIP][HZ_NB], sX[8][HZ_NB];
    __shared__ double sRf[HZ_NB];
    __shared__ double sIz[ZS_KZMAX];
    __shared__ double sCx[8][HZ_TI], sCy[8][HZ_TJ], sCz[7][ZS_KZMAX];
    __shared__ double sKx[4][HZ_TI], sKy[4][HZ_TJ], sKz[4][ZS_KZMAX];
    hyper_zs_body<BGP, false, double>(g, zm, ph, uin, R, aux, bg, cmb, dUh, sU, sA, sRf, sR, sX,
                                      sIz, sCx, sCy, sCz, sKx, sKy, sKz, nullptr);
}

template <bool BGP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void
k_hyper_zsr_d(Geo g, ZMap zm, Phys ph, const double *__restrict__ uin,
              const double *__restrict__ R, const double *__restrict__ aux,
              const double *__restrict__ bg, const unsigned long long *__restrict__ cmb,
              double *__restrict__ dUh)
{
    __shared__ double sM[3][HZ_NB], sR[HZ_NIP][HZ_NB], sX[8][HZ_NB], sQ[8][QB_NB];
    __shared__ double sRf[HZ_NB];
    __shared__ double sIz[ZS_KZMAX];
    __shared__ double sCx[8][HZ_TI], sCy[8][HZ_TJ], sCz[7][ZS_KZMAX];
    __shared__ double sKx[4][HZ_TI], sKy[4][HZ_TJ], sKz[4][ZS_KZMAX];
    hyper_zs_body<BGP, true, double>(g, zm, ph, uin, R, aux, bg, cmb, dUh, sM, nullptr, sRf, sR,
                                     sX, sIz, sCx, sCy, sCz, sKx, sKy, sKz, sQ);
}

// the hyperdiffusive sources of every interior cell into dUh (the stage
// update adds them); same tiles and XCD panels as k_update
template <int D, typename T>
__global__ __launch_bounds__(256) void k_hyper(Geo g, TileMap tm, Phys ph,
                                                const T *__restrict__ uin,
                                                const T *__restrict__ R,
                                                const T *__restrict__ aux,
                                                const double *__restrict__ bg,
                                                const unsigned long long *__restrict__ cmb,
                                                T *__restrict__ dUh)
{
    typedef Lay<D> LY;
    int tx, ty, tk;
    if (!map_block(tm, tx, ty, tk)) return;
    const int x0 = 2 + tx * 64 + (threadIdx.x & 63);
    const int x1 = 2 + ty * 4 + (threadIdx.x >> 6);
    const int x2 = (D == 3) ? 2 + tk : 0;
    if (x0 > g.n[0] - 3 || x1 > g.n[1] - 3) return;
    T dU[LY::NVE];
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) dU[f] = 0;
    hyper_cell<D, T>(g, ph, uin, R, aux, bg, cmb, x0, x1, x2, dU);
    const size_t nc = (size_t)g.nc;
    const ci_t c = (ci_t)x0 + (ci_t)g.s1 * (ci_t)x1 + (ci_t)g.s2 * (ci_t)x2;
#pragma unroll
    for (int f = 0; f < LY::NVE; f++) st(dUh + f * nc, c, dU[f]);
}

// Launchers of the kernels that smaug_ilp.hip builds under the max-ILP
// machine scheduler (defined there; the other translation units only call
// them, so that these kernels have one build): the FP32 z-streamed sources
// with the in-plane ratios, and the ratio pairs along k in both precisions.
hipError_t launch_hyper_zsr_f(bool bgp, unsigned nb, hipStream_t st, const Geo &g, const ZMap &zm,
                              const Phys &ph, const float *uin, const float *R, const float *aux,
                              const double *bg, const unsigned long long *cmb, float *dUh);
hipError_t launch_ratio_kz_f(unsigned nb, hipStream_t st, const Geo &g, const ZMap &zm,
                             const float *u, const float *aux, float *R);
hipError_t launch_ratio_kz_d(unsigned nb, hipStream_t st, const Geo &g, const ZMap &zm,
                             const double *u, const double *aux, double *R);

} // namespace smaug
