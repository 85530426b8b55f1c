// smaug_engine.hip -- host side of the MI355X stepping engine: device
// arena, launch geometry, the captured RK4 step graph, and the C-ABI of
// smaug_engine.h.
#include "smaug_kernels.hpp"
#include "smaug_stream.hpp"
#include "smaug_host.hpp"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>

using namespace smaug;

static thread_local char g_err[512] = "";

int smaug::fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int smaug::check_cfg(const smaug_cfg *cfg)
{
    int D = cfg->ndim;
    if (D != 2 && D != 3) return fail(SMAUG_EINVAL, "ndim must be 2 or 3 (got %d)", D);
    for (int d = 0; d < D; d++) {
        if (cfg->n[d] < 8) return fail(SMAUG_EINVAL, "n[%d]=%d: need >= 8 cells incl. ghosts", d, cfg->n[d]);
        if (cfg->bc[d] != 0 && cfg->bc[d] != 4)
            return fail(SMAUG_EINVAL, "boundtype %d along dim %d unsupported (0 periodic, 4 contcd4)",
                        cfg->bc[d], d);
        if (!cfg->dx[d]) return fail(SMAUG_EINVAL, "dx[%d] missing", d);
        if (cfg->driver && !cfg->pos[d]) return fail(SMAUG_EINVAL, "pos[%d] missing", d);
    }
    if (cfg->driver < 0 || cfg->driver > 2)
        return fail(SMAUG_EINVAL, "driver must be 0 (none), 1 (vertical) or 2 (torsional)");
    if (cfg->driver == 2) {
        if (D != 3) return fail(SMAUG_EINVAL, "the torsional driver drives m_y and m_z: 3-D only");
        if (!(cfg->drv[7] != 0.0)) return fail(SMAUG_EINVAL, "torsional driver: drv[7] (L) is 0");
    }
    if (cfg->precision != SMAUG_FP64 && cfg->precision != SMAUG_MIXED)
        return fail(SMAUG_EINVAL, "precision must be SMAUG_FP64 or SMAUG_MIXED");
    long long nc = (long long)cfg->n[0] * cfg->n[1] * (D == 3 ? cfg->n[2] : 1);
    if (nc > MAX_CELLS)
        return fail(SMAUG_EINVAL, "%lld cells exceed the per-grid limit of %lld", nc, MAX_CELLS);
    return 0;
}

int smaug::check_periodic_bg(int D, const int n[3], const int bc[3], const double *bg)
{
    const int n2 = D == 3 ? n[2] : 1;
    const size_t nc = (size_t)n[0] * n[1] * n2;
    const size_t st[3] = {1, (size_t)n[0], (size_t)n[0] * n[1]};
    for (int d = 0; d < D; d++) {
        if (bc[d] != 0) continue;
        const int nd = n[d];
        for (int f = 0; f < 2 + D; f++) {
            const double *b = bg + (size_t)f * nc;
            for (size_t c = 0; c < nc; c++) {
                const int x = (int)((c / st[d]) % (size_t)nd);
                int img = -1;
                if (x < 2) img = nd - 4 + x;
                else if (x >= nd - 2) img = x - (nd - 4);
                if (img < 0) continue;
                const size_t ci = c + ((long long)img - x) * (long long)st[d];
                if (memcmp(&b[c], &b[ci], sizeof(double)))
                    return fail(SMAUG_EINVAL,
                                "dim %d is periodic but background field %d is not: ghost cell "
                                "%d differs from its image %d (the ghost fill copies only the "
                                "perturbation)", d, f, x, img);
            }
        }
    }
    return 0;
}

void smaug::axis_arrays(const smaug_cfg *cfg, int d, int nd, std::vector<double> &dx,
                        std::vector<double> &idx, std::vector<double> &prof,
                        std::vector<double> &profq)
{
    dx.assign(nd, 1.0);
    idx.assign(nd, 1.0);
    prof.assign(nd, 1.0);
    profq.assign(nd, 0.0);
    if (d >= cfg->ndim) return;
    for (int i = 0; i < nd; i++) {
        dx[i] = cfg->dx[d][i];
        idx[i] = 1.0 / dx[i];
    }
    if (cfg->driver == 1 || cfg->driver == 2) {
        for (int i = 0; i < nd; i++) {
            double p = cfg->pos[d][i];
            double c0 = d == 0 ? cfg->drv[3] : (d == 1 ? cfg->drv[5] : cfg->drv[6]);
            double s = d == 0 ? cfg->drv[4] : cfg->drv[2];
            prof[i] = exp(-(p - c0) * (p - c0) / (s * s));
            if (d > 0) profq[i] = (p - c0) * prof[i];
        }
    }
}

extern "C" const char *smaug_last_error(void) { return g_err; }

struct smaug_engine {
    smaug_cfg cfg;
    int D;
    int n[3];
    idx_t nc, nint;
    int prec; // SMAUG_FP64 / SMAUG_MIXED
    size_t tsz;
    hipStream_t stream;
    void *U[3];   // U[0] = current state, U[1], U[2] stage buffers
    void *ACC;
    void *R;      // hyperviscosity ratios
    void *AUX;    // per-cell 1/rho, pressure, internal energy, velocity
    void *HYP;    // 3-D: hyperdiffusive part of L(w) for the z-streamed update
    double *d_part; // per-block partial sums of the conservation diagnostics
    double *bg;
    double *bgp;  // 3-D: the background as one k-plane (field stride n0*n1)
    int bgp_ok;   // the uploaded background does not vary along k
    int bgp_on;   // use bgp when bgp_ok (default on)
    double *bgu;  // the background's nb values when it is the same in every cell
    int bgu_ok;   // the per-cell kernels then read bgu (Geo::bgf = 1, bgm = 0)
    double *d_dx[3], *d_idx[3], *d_prof[3], *d_profq[3];
    unsigned long long *d_cmax;
    double *d_cpart; // per-block partial cfast maxima of the aux/cmax pass
    int ncpart;      // its capacity in blocks
    void *AUXB;      // 2-D: the aux quantities of the next stage's input, formed by
                     // the update of RK stages 0-2 while it reads AUX (swapped per stage)
    void *aux_in;    // the aux quantities of the current stage's input (AUX or AUXB)
    int aux0_ok;     // AUX and cmax hold the aux pass of U0 (left by the last folded
                     // RK4 step); any other change of the state clears it
    double *d_tm; // [t, dt]
    Geo geo;
    Phys phys;
    hipGraphExec_t gexec[2]; // 0: RK4 step, 1: Euler step
    bool gvalid[2];
    int panel[2][2];         // XCD panel (tpx, tpy) of the ratio and update passes; 0 = default
    int zstream;             // 3-D: 1 = z-streamed stage update, 0 = per-cell stage update
    int kz;                  // z-stream depth (output planes per block)
    int hyper_lds;           // 3-D: hyperdiffusive sources with the plane staged in LDS
    int ratio_zs;            // 3-D: z-streamed hyperviscosity ratios (2: the pairs along k
                             // alone, the in-plane pairs formed by the source kernel)
    int kzr;                 // their z-stream depth (0: 32)
    int kzh;                 // z-stream depth of the z-streamed hyperdiffusive sources (0: 32)
};

// ---------------------------------------------------------------------------
// launch helpers (templated on D and T)
// ---------------------------------------------------------------------------
// the background of a 3-D grid is read from its k-plane copy
static bool bg_planar(const smaug_engine *e) { return e->D == 3 && e->bgp_ok && e->bgp_on; }

constexpr int PREP_KC = 16; // planes per block of k_prep_col

// geometry and background pointer for the per-cell kernels: a background
// that is the same in every cell is read as its nb values
static Geo pc_geo(const smaug_engine *e, const double **bg)
{
    Geo g = e->geo;
    if (e->bgu_ok) {
        g.bgf = 1;
        g.bgm = 0u;
        *bg = e->bgu;
    } else {
        *bg = e->bg;
    }
    return g;
}

// blocks of the grid-stride aux/cmax pass
static idx_t prep_blocks(idx_t nc)
{
    idx_t nb = (nc + 255) / 256;
    return nb > 4096 ? 4096 : nb;
}

// aux quantities and cmax of u; fill = 1 (k-plane background only) also
// fills the ghost cells of u, which the previous stage then left unfilled
template <int D, typename T>
static int launch_cmax(smaug_engine *e, const void *u, int fill = 0)
{
    unsigned nb;
    if (D == 3 && bg_planar(e)) {
        const unsigned nx = (unsigned)((e->geo.s2 + 255) / 256);
        const unsigned ny = (unsigned)((e->n[2] + PREP_KC - 1) / PREP_KC);
        nb = nx * ny;
        hipLaunchKernelGGL((k_prep_col<T>), dim3(nx, ny), dim3(256), 0, e->stream, e->geo,
                           e->phys, (T *)u, (const double *)e->bgp, (T *)e->aux_in, e->d_cpart,
                           PREP_KC, e->zstream ? 0 : 1, fill);
    } else {
        nb = (unsigned)prep_blocks(e->nc);
        const double *bg;
        const Geo g = pc_geo(e, &bg);
        hipLaunchKernelGGL((k_prep<D, T>), dim3(nb), dim3(256), 0, e->stream, g, e->phys,
                           (const T *)u, bg, (T *)e->aux_in, e->d_cpart);
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_cmax_reduce<D>, dim3(1), dim3(1024), 0, e->stream,
                       (const double *)e->d_cpart, (int)nb, e->d_cmax);
    HIPCHK(hipGetLastError());
    return 0;
}

// tiles of 64 x th cells over an extent of e0 x e1 x e2 cells, grouped in
// panels of tpx x tpy tiles so that each XCD sweeps its own sub-domain (see
// map_block); tpx, tpy <= 0 pick the default panel
static TileMap make_map(int e0, int e1, int e2, int tpx, int tpy, unsigned *nblocks, int th = 4)
{
    TileMap m;
    m.ntx = (e0 + 63) / 64;
    m.nty = (e1 + th - 1) / th;
    m.nk = e2;
    if (tpx <= 0 || tpy <= 0) {
        tpx = m.ntx < 8 ? m.ntx : 8;
        tpy = 16 / tpx > 0 ? 16 / tpx : 1;
    }
    m.tpx = tpx < m.ntx ? tpx : m.ntx;
    m.tpy = tpy < m.nty ? tpy : m.nty;
    m.npx = (m.ntx + m.tpx - 1) / m.tpx;
    int npy = (m.nty + m.tpy - 1) / m.tpy;
    m.npanels = m.npx * npy;
    int per_xcd = (m.npanels + 7) / 8;
    *nblocks = (unsigned)(8 * per_xcd * m.tpx * m.tpy * m.nk);
    return m;
}

// tiles and z-chunks of the ratio box [1, n-3] for the z-streamed ratio
// kernels: tiles of RZ_TI x tj columns (RZ_TJ for k_ratio_zs, KZ_TJ for
// k_ratio_kz)
static ZMap make_rmap(const smaug_engine *e, unsigned *nblocks, int tj = RZ_TJ)
{
    ZMap m;
    m.ntx = (e->n[0] + RZ_TI - 1) / RZ_TI;
    m.nty = (e->n[1] - 3 + tj - 1) / tj;
    const int nk = e->n[2] - 3;
    m.kz = e->kzr > 0 ? (e->kzr < nk ? e->kzr : nk) : (nk < 32 ? nk : 32);
    m.nkz = (nk + m.kz - 1) / m.kz;
    m.per_xcd = (m.ntx * m.nty + 7) / 8;
    *nblocks = (unsigned)(8 * m.per_xcd * m.nkz);
    return m;
}

// 3-D, ratio form 2 with the z-streamed update and source kernel: the ratio
// pass forms the pairs along k alone and the source kernel the in-plane
// pairs (the per-cell forms read every pair from R)
static bool ratio_in_sources(const smaug_engine *e)
{
    return e->D == 3 && e->ratio_zs == 2 && e->zstream && e->hyper_lds == 2;
}

// 2-D, ratio form 2: the update kernel forms the ratios itself
// (k_update2r) and the stage has no ratio pass
static bool ratio_in_update(const smaug_engine *e)
{
    return e->D == 2 && e->ratio_zs == 2 && e->phys.hyperdif;
}

// the tiles of the per-cell update pass over the interior: 64 x 4 cells,
// or 64 x u2_tj for the 2-D update that forms its ratios.  The ghost fill
// that follows a folded update places its cfast partials after the
// update's, so both take the block count from here.
static TileMap update_map(const smaug_engine *e, unsigned *nb)
{
    const int th = !ratio_in_update(e) ? 4 : (e->prec == SMAUG_FP64 ? u2_tj<double>() : u2_tj<float>());
    return make_map(e->n[0] - 4, e->n[1] - 4, e->D == 3 ? e->n[2] - 4 : 1, e->panel[1][0],
                    e->panel[1][1], nb, th);
}

template <int D, typename T>
static int launch_ratio(smaug_engine *e, const void *u)
{
    if constexpr (D == 3) {
        if (e->ratio_zs) {
            unsigned nb;
            const bool kz = ratio_in_sources(e);
            ZMap m = make_rmap(e, &nb, kz ? KZ_TJ : RZ_TJ);
            const int nt = kz ? KZ_NT : RZ_NT;
            if (kz) {
                hipError_t he;
                if constexpr (sizeof(T) == 4)
                    he = launch_ratio_kz_f(nb, e->stream, e->geo, m, (const T *)u,
                                           (const T *)e->aux_in, (T *)e->R);
                else
                    he = launch_ratio_kz_d(nb, e->stream, e->geo, m, (const T *)u,
                                           (const T *)e->aux_in, (T *)e->R);
                HIPCHK(he);
                return 0;
            }
            if constexpr (sizeof(T) == 4)
                hipLaunchKernelGGL(k_ratio_zs_f, dim3(nb), dim3(nt), 0, e->stream, e->geo, m,
                                   (const T *)u, (const T *)e->aux_in, (T *)e->R);
            else
                hipLaunchKernelGGL(k_ratio_zs_d, dim3(nb), dim3(nt), 0, e->stream, e->geo, m,
                                   (const T *)u, (const T *)e->aux_in, (T *)e->R);
            HIPCHK(hipGetLastError());
            return 0;
        }
    }
    unsigned nb;
    TileMap m = make_map(e->n[0] - 3, e->n[1] - 3, D == 3 ? e->n[2] - 3 : 1, e->panel[0][0],
                         e->panel[0][1], &nb); // box [1, n-3]
    hipLaunchKernelGGL((k_ratio<D, T>), dim3(nb), dim3(256), 0, e->stream, e->geo, m, (const T *)u,
                       (const T *)e->aux_in, (T *)e->R);
    HIPCHK(hipGetLastError());
    return 0;
}

// aux_out (2-D, RK stages 0-2): the update also forms the aux quantities
// and cfast partials of its output there (launch_ghost_fold then does the
// ghost cells and the reduction) -- the next stage's aux pass without a
// pass over the interior
template <int D, typename T>
static int launch_update(smaug_engine *e, StageArgs sa, const void *uin, const void *un,
                         void *uout, void *aux_out = nullptr)
{
    unsigned nb;
    TileMap m = update_map(e, &nb);
    const double *bg;
    const Geo g = pc_geo(e, &bg);
    // a background that is the same in every cell (bgm == 0) is read once
    const bool uni = g.bgm == 0u;
    if constexpr (D == 2) {
        if (ratio_in_update(e)) {
            if (aux_out) {
                const idx_t ng = ghost_count2(e->n[0], e->n[1]);
                const unsigned gb = (unsigned)((ng + 255) / 256);
                if ((idx_t)nb + gb > (idx_t)e->ncpart)
                    return fail(SMAUG_EINVAL, "cfast partials overflow");
                auto kern = uni ? k_update2r<T, true, true> : k_update2r<T, false, true>;
                hipLaunchKernelGGL(kern, dim3(nb), dim3(U2<T>::NT), 0, e->stream, g, m, e->phys, sa,
                                   (const T *)uin, (const T *)un, (T *)uout, (T *)e->ACC,
                                   (const T *)e->aux_in, bg,
                                   (const unsigned long long *)e->d_cmax, (T *)aux_out,
                                   e->d_cpart);
            } else {
                auto kern = uni ? k_update2r<T, true> : k_update2r<T, false>;
                hipLaunchKernelGGL(kern, dim3(nb), dim3(U2<T>::NT), 0, e->stream, g, m, e->phys, sa,
                                   (const T *)uin, (const T *)un, (T *)uout, (T *)e->ACC,
                                   (const T *)e->aux_in, bg,
                                   (const unsigned long long *)e->d_cmax, nullptr, nullptr);
            }
            HIPCHK(hipGetLastError());
            return 0;
        }
    }
    if (aux_out) {
        if constexpr (D == 2) {
            const idx_t ng = ghost_count2(e->n[0], e->n[1]);
            const unsigned gb = (unsigned)((ng + 255) / 256);
            if ((idx_t)nb + gb > (idx_t)e->ncpart)
                return fail(SMAUG_EINVAL, "cfast partials overflow");
            auto kern = uni ? k_update<D, T, true, true> : k_update<D, T, false, true>;
            hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, e->stream, g, m, e->phys, sa,
                               (const T *)uin, (const T *)un, (T *)uout, (T *)e->ACC,
                               (const T *)e->R, (const T *)e->aux_in, bg,
                               (const unsigned long long *)e->d_cmax, (T *)aux_out, e->d_cpart);
            HIPCHK(hipGetLastError());
            return 0;
        } else {
            return fail(SMAUG_EINVAL, "the folded per-cell aux pass is for 2-D grids");
        }
    }
    auto kern = uni ? k_update<D, T, true> : k_update<D, T, false>;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, e->stream, g, m, e->phys, sa,
                       (const T *)uin, (const T *)un, (T *)uout, (T *)e->ACC,
                       (const T *)e->R, (const T *)e->aux_in, bg,
                       (const unsigned long long *)e->d_cmax, nullptr, nullptr);
    HIPCHK(hipGetLastError());
    return 0;
}

static ZMap make_zmap(const smaug_engine *e, unsigned *nblocks)
{
    ZMap m;
    m.ntx = (e->n[0] - 4 + ZS_TI - 1) / ZS_TI;
    m.nty = (e->n[1] - 4 + ZS_TJ - 1) / ZS_TJ;
    const int nk = e->n[2] - 4;
    m.kz = e->kz > 0 ? (e->kz < nk ? e->kz : nk) : (nk < 32 ? nk : 32);
    if (m.kz > ZS_KZMAX) m.kz = ZS_KZMAX;
    m.nkz = (nk + m.kz - 1) / m.kz;
    m.per_xcd = (m.ntx * m.nty + 7) / 8;
    *nblocks = (unsigned)(8 * m.per_xcd * m.nkz);
    return m;
}

// 3-D z-streamed form: the hyperdiffusive sources of uin into HYP (the
// stage update adds them)
static ZMap make_hmap(const smaug_engine *e, unsigned *nblocks)
{
    ZMap m;
    m.ntx = (e->n[0] - 4 + HZ_TI - 1) / HZ_TI;
    m.nty = (e->n[1] - 4 + HZ_TJ - 1) / HZ_TJ;
    const int nk = e->n[2] - 4;
    m.kz = e->kzh > 0 ? (e->kzh < nk ? e->kzh : nk) : (nk < 32 ? nk : 32);
    if (m.kz > ZS_KZMAX) m.kz = ZS_KZMAX;
    m.nkz = (nk + m.kz - 1) / m.kz;
    m.per_xcd = (m.ntx * m.nty + 7) / 8;
    *nblocks = (unsigned)(8 * m.per_xcd * m.nkz);
    return m;
}

template <typename T>
static int launch_hyper_sources(smaug_engine *e, const void *uin)
{
    if (e->phys.hyperdif && e->hyper_lds == 2) {
        unsigned nb;
        ZMap m = make_hmap(e, &nb);
        const bool bp = bg_planar(e);
        const double *bg = bp ? e->bgp : e->bg;
        const bool rin = ratio_in_sources(e);
        if constexpr (sizeof(T) == 4) {
            if (rin) {
                HIPCHK(launch_hyper_zsr_f(bp, nb, e->stream, e->geo, m, e->phys, (const T *)uin,
                                          (const T *)e->R, (const T *)e->aux_in, bg,
                                          (const unsigned long long *)e->d_cmax, (T *)e->HYP));
                return 0;
            }
            auto kern = bp ? k_hyper_zs_f<true> : k_hyper_zs_f<false>;
            hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, e->stream, e->geo, m, e->phys,
                               (const T *)uin, (const T *)e->R, (const T *)e->aux_in, bg,
                               (const unsigned long long *)e->d_cmax, (T *)e->HYP);
        } else {
            auto kern = rin ? (bp ? k_hyper_zsr_d<true> : k_hyper_zsr_d<false>)
                            : (bp ? k_hyper_zs_d<true> : k_hyper_zs_d<false>);
            hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, e->stream, e->geo, m, e->phys,
                               (const T *)uin, (const T *)e->R, (const T *)e->aux_in, bg,
                               (const unsigned long long *)e->d_cmax, (T *)e->HYP);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }
    if (e->phys.hyperdif) {
        unsigned nh;
        TileMap tm = make_map(e->n[0] - 4, e->n[1] - 4, e->n[2] - 4, e->panel[1][0],
                              e->panel[1][1], &nh);
        if (e->hyper_lds) {
            if constexpr (sizeof(T) == 4)
                hipLaunchKernelGGL(k_hyper_lds_f, dim3(nh), dim3(256), 0, e->stream, e->geo, tm,
                                   e->phys, (const T *)uin, (const T *)e->R, (const T *)e->aux_in,
                                   (const double *)e->bg, (const unsigned long long *)e->d_cmax,
                                   (T *)e->HYP);
            else
                hipLaunchKernelGGL(k_hyper_lds_d, dim3(nh), dim3(256), 0, e->stream, e->geo, tm,
                                   e->phys, (const T *)uin, (const T *)e->R, (const T *)e->aux_in,
                                   (const double *)e->bg, (const unsigned long long *)e->d_cmax,
                                   (T *)e->HYP);
        } else {
            const double *bg;
            const Geo g = pc_geo(e, &bg);
            hipLaunchKernelGGL((k_hyper<3, T>), dim3(nh), dim3(256), 0, e->stream, g, tm,
                               e->phys, (const T *)uin, (const T *)e->R, (const T *)e->aux_in, bg,
                               (const unsigned long long *)e->d_cmax, (T *)e->HYP);
        }
        HIPCHK(hipGetLastError());
    }
    return 0;
}

// fold (k-plane background only): the update also forms the aux
// quantities and cfast partials of its output, k_ghost_prep fills the
// output's ghost cells and forms theirs, and the partials are reduced into
// cmax -- the aux/cmax pass of the next stage, done without reading the
// interior of the output again
template <typename T>
static int launch_zstream(smaug_engine *e, StageArgs sa, const void *uin, const void *un,
                          void *uout, bool fold = false)
{
    unsigned nb;
    ZMap m = make_zmap(e, &nb);
    const bool bp = bg_planar(e);
    const double *bg = bp ? e->bgp : e->bg;
    if (fold && !bp) return fail(SMAUG_EINVAL, "folded aux pass needs the k-plane background");
    // Every lane of the update issues its loads of the sources, at cell 0
    // where it has no use for them (no output, or no hyperdiffusion).  HYP
    // is only allocated with hyperdiffusion; without it the update is given
    // uin, 8 fields of nc cells like HYP, so that those loads stay in bounds.
    const void *hyp = e->HYP ? e->HYP : uin;
    if constexpr (sizeof(T) == 4) {
        if (fold)
            hipLaunchKernelGGL((k_zstream_f<true, true>), dim3(nb), dim3(ZS_NT), 0, e->stream,
                               e->geo, e->phys, sa, m, (const T *)uin, (const T *)un, (T *)uout,
                               (T *)e->ACC, (const T *)hyp, bg, (T *)e->AUX, e->d_cpart);
        else
            hipLaunchKernelGGL(bp ? k_zstream_f<true> : k_zstream_f<false>, dim3(nb), dim3(ZS_NT),
                               0, e->stream, e->geo, e->phys, sa, m, (const T *)uin,
                               (const T *)un, (T *)uout, (T *)e->ACC, (const T *)hyp, bg,
                               nullptr, nullptr);
    } else {
        if (fold)
            hipLaunchKernelGGL((k_zstream_d<true, true>), dim3(nb), dim3(ZS_NT), 0, e->stream,
                               e->geo, e->phys, sa, m, (const T *)uin, (const T *)un, (T *)uout,
                               (T *)e->ACC, (const T *)hyp, bg, (T *)e->AUX, e->d_cpart);
        else
            hipLaunchKernelGGL(bp ? k_zstream_d<true> : k_zstream_d<false>, dim3(nb), dim3(ZS_NT),
                               0, e->stream, e->geo, e->phys, sa, m, (const T *)uin,
                               (const T *)un, (T *)uout, (T *)e->ACC, (const T *)hyp, bg,
                               nullptr, nullptr);
    }
    HIPCHK(hipGetLastError());
    if (fold) {
        const idx_t ng = ghost_count(e->n[0], e->n[1], e->n[2]);
        if ((idx_t)nb + (ng + 255) / 256 > (idx_t)e->ncpart)
            return fail(SMAUG_EINVAL, "cfast partials overflow");
    }
    return 0;
}

// The ghost fill of an update output whose interior aux pass the update
// folded in: the ghost cells take their images' state and get their aux
// quantities and cfast partials (after the update's partials), and the
// partials are reduced into cmax.  3-D: k_ghost_prep (k-plane background,
// aux in AUX); 2-D: k_ghost_prep2 (aux in aux_out).
template <int D, typename T>
static int launch_ghost_fold(smaug_engine *e, void *uout, void *aux_out)
{
    unsigned nb, gb;
    if constexpr (D == 3) {
        make_zmap(e, &nb);
        gb = (unsigned)((ghost_count(e->n[0], e->n[1], e->n[2]) + 255) / 256);
        if ((idx_t)nb + gb > (idx_t)e->ncpart) return fail(SMAUG_EINVAL, "cfast partials overflow");
        hipLaunchKernelGGL((k_ghost_prep<T>), dim3(gb), dim3(256), 0, e->stream, e->geo, e->phys,
                           (T *)uout, (const double *)e->bgp, (T *)e->AUX,
                           e->d_cpart + (size_t)4 * nb, 0);
    } else {
        update_map(e, &nb);
        gb = (unsigned)((ghost_count2(e->n[0], e->n[1]) + 255) / 256);
        if ((idx_t)nb + gb > (idx_t)e->ncpart) return fail(SMAUG_EINVAL, "cfast partials overflow");
        const double *bg;
        const Geo g = pc_geo(e, &bg);
        hipLaunchKernelGGL((k_ghost_prep2<T>), dim3(gb), dim3(256), 0, e->stream, g, e->phys,
                           (T *)uout, bg, (T *)aux_out, e->d_cpart + (size_t)4 * nb);
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_cmax_reduce<D>, dim3(1), dim3(1024), 0, e->stream,
                       (const double *)e->d_cpart, (int)(nb + gb), e->d_cmax);
    HIPCHK(hipGetLastError());
    return 0;
}

template <int D, typename T>
static bool zstreamed(const smaug_engine *e)
{
    return D == 3 && e->zstream;
}

// the hyperdiffusive-source pass that precedes the z-streamed update
template <int D, typename T>
static int launch_pre_update(smaug_engine *e, const void *uin)
{
    if constexpr (D == 3)
        if (e->zstream && e->phys.hyperdif) return launch_hyper_sources<T>(e, uin);
    return 0;
}

template <int D, typename T>
static int launch_stage_update(smaug_engine *e, StageArgs sa, const void *uin, const void *un,
                               void *uout, bool fold = false, void *aux_out = nullptr)
{
    if constexpr (D == 3)
        if (e->zstream) return launch_zstream<T>(e, sa, uin, un, uout, fold);
    return launch_update<D, T>(e, sa, uin, un, uout, aux_out);
}

template <int D, typename T>
static int launch_bc(smaug_engine *e, void *u)
{
    idx_t tot = (idx_t)4 * e->n[1] * e->n[2] + (idx_t)4 * e->n[0] * e->n[2] +
                (D == 3 ? (idx_t)4 * e->n[0] * e->n[1] : 0);
    hipLaunchKernelGGL((k_bc<D, T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       e->stream, e->geo, (T *)u);
    HIPCHK(hipGetLastError());
    return 0;
}

__global__ void k_advance_time(double *tm)
{
    if (threadIdx.x == 0) tm[0] += tm[1];
}

// one stage: cmax(uin), ratios(uin), update -> uout, ghost fill of uout.
// With the k-plane background, the outputs of RK stages 0-2 are left
// unfilled and the aux/cmax pass of the next stage fills their ghosts; with
// the z-streamed update as well, that pass is folded into the update of
// every stage (launch_zstream), and every stage starts at the ratios: the
// step's stage 3 does the aux pass of the next step's stage 0.
// 2-D grids fold it into the per-cell update of RK stages 0-2 likewise
// (launch_update), alternating two aux buffers.
template <int D, typename T>
static int stage_launches(smaug_engine *e, StageArgs sa, const void *uin, const void *un,
                          void *uout, hipEvent_t ev0, hipEvent_t ev1, const hipEvent_t *evk)
{
    int rc;
    const bool defer = D == 3 && bg_planar(e) && sa.mode == 0;
    const bool fold = defer && e->zstream;
    // 2-D: the updates of RK stages 0-2 form the next stage's aux quantities
    // in the other aux buffer (the update reads the current one at the
    // neighbours) and fill the output's ghosts; the stages alternate buffers
    const bool fold2 = D == 2 && sa.mode == 0 && e->AUXB;
    e->aux_in = fold2 && (sa.stage & 1) ? e->AUXB : e->AUX;
    void *aux_out = fold2 && sa.stage < 3 ? ((sa.stage & 1) ? e->AUX : e->AUXB) : nullptr;
    if (evk) HIPCHK(hipEventRecord(evk[0], e->stream));
    // 3-D fold: the aux pass of every stage is done by the update before it,
    // that of stage 0 by the previous step's stage 3 (or ensure_aux0)
    if (!fold && !(fold2 && sa.stage > 0))
        if ((rc = launch_cmax<D, T>(e, uin, defer && sa.stage > 0))) return rc;
    if (evk) HIPCHK(hipEventRecord(evk[1], e->stream));
    if (e->phys.hyperdif && !ratio_in_update(e))
        if ((rc = launch_ratio<D, T>(e, uin))) return rc;
    if (evk) HIPCHK(hipEventRecord(evk[2], e->stream));
    if ((rc = launch_pre_update<D, T>(e, uin))) return rc;
    if (ev0) HIPCHK(hipEventRecord(ev0, e->stream));
    if (evk) HIPCHK(hipEventRecord(evk[3], e->stream));
    if ((rc = launch_stage_update<D, T>(e, sa, uin, un, uout, fold, aux_out))) return rc;
    if (ev1) HIPCHK(hipEventRecord(ev1, e->stream));
    if (evk) HIPCHK(hipEventRecord(evk[4], e->stream));
    if (fold || aux_out) {
        if ((rc = launch_ghost_fold<D, T>(e, uout, aux_out))) return rc;
    } else if (sa.mode != 2 && !((defer || fold2) && sa.stage < 3)) {
        if ((rc = launch_bc<D, T>(e, uout))) return rc;
    }
    if (evk) HIPCHK(hipEventRecord(evk[5], e->stream));
    return 0;
}

template <int D, typename T>
static int stage(smaug_engine *e, StageArgs sa, const void *uin, const void *un, void *uout,
                 hipEvent_t ev0, hipEvent_t ev1, const hipEvent_t *evk = nullptr)
{
    const int rc = stage_launches<D, T>(e, sa, uin, un, uout, ev0, ev1, evk);
    e->aux_in = e->AUX; // every other launch reads the aux quantities from AUX
    return rc;
}

// RK4 steps whose stages all fold the aux pass into the update before them
template <int D, typename T>
static bool folds_all(const smaug_engine *e)
{
    return D == 3 && e->zstream && bg_planar(e);
}

// before RK4 steps that fold: the aux pass of U0 unless the last folded
// step left it (eager, ahead of the captured steps)
template <int D, typename T>
static int ensure_aux0(smaug_engine *e)
{
    if (!folds_all<D, T>(e) || e->aux0_ok) return 0;
    int rc = launch_cmax<D, T>(e, e->U[0]);
    if (rc) return rc;
    e->aux0_ok = 1;
    return 0;
}

template <int D, typename T>
static int rk4_step(smaug_engine *e, std::vector<hipEvent_t> *evs,
                    std::vector<hipEvent_t> *evk = nullptr)
{
    // buffers: U0 = u_n (and u_{n+1}), U1/U2 stage outputs, ACC accumulator
    const double ct[4] = {0.0, 0.5, 0.5, 1.0};
    void *in[4] = {e->U[0], e->U[1], e->U[2], e->U[1]};
    void *out[4] = {e->U[1], e->U[2], e->U[1], e->U[0]};
    for (int s = 0; s < 4; s++) {
        StageArgs sa;
        sa.mode = 0;
        sa.stage = s;
        sa.ct = ct[s];
        sa.tm = e->d_tm;
        hipEvent_t a = 0, b = 0;
        if (evs) {
            a = (*evs)[2 * s];
            b = (*evs)[2 * s + 1];
        }
        int rc = stage<D, T>(e, sa, in[s], e->U[0], out[s], a, b,
                             evk ? evk->data() + 6 * s : nullptr);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(64), 0, e->stream, e->d_tm);
    HIPCHK(hipGetLastError());
    return 0;
}

template <int D, typename T>
static int euler_step(smaug_engine *e)
{
    StageArgs sa;
    sa.mode = 1;
    sa.stage = 0;
    sa.ct = 0.0;
    sa.tm = e->d_tm;
    int rc = stage<D, T>(e, sa, e->U[0], e->U[0], e->U[1], 0, 0);
    if (rc) return rc;
    // the step result lives in U1; copy back into U0 with a kernel-free swap
    void *t = e->U[0];
    e->U[0] = e->U[1];
    e->U[1] = t;
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(64), 0, e->stream, e->d_tm);
    HIPCHK(hipGetLastError());
    return 0;
}

// dispatch over (D, precision)
// DISPATCH_KEEP runs the instance of FN for the engine's dimensionality and
// precision.  DISPATCH does the same after clearing aux0_ok: every entry
// point that reaches the state goes through it, so a new one invalidates the
// aux quantities of U0 by default; smaug_rk4 alone keeps them, as its
// folded steps leave them valid for the next call.
#define DISPATCH_KEEP(e, FN, ...)                                                          \
    ((e)->D == 2 ? ((e)->prec == SMAUG_FP64 ? FN<2, double>(__VA_ARGS__)                   \
                                            : FN<2, float>(__VA_ARGS__))                   \
                 : ((e)->prec == SMAUG_FP64 ? FN<3, double>(__VA_ARGS__)                   \
                                            : FN<3, float>(__VA_ARGS__)))
#define DISPATCH(e, FN, ...) ((e)->aux0_ok = 0, DISPATCH_KEEP(e, FN, __VA_ARGS__))

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
static void free_engine(smaug_engine *e)
{
    if (!e) return;
    for (int k = 0; k < 2; k++)
        if (e->gvalid[k]) hipGraphExecDestroy(e->gexec[k]);
    for (int k = 0; k < 3; k++) {
        if (e->U[k]) hipFree(e->U[k]);
        if (e->d_dx[k]) hipFree(e->d_dx[k]);
        if (e->d_idx[k]) hipFree(e->d_idx[k]);
        if (e->d_prof[k]) hipFree(e->d_prof[k]);
        if (e->d_profq[k]) hipFree(e->d_profq[k]);
    }
    if (e->ACC) hipFree(e->ACC);
    if (e->R) hipFree(e->R);
    if (e->AUX) hipFree(e->AUX);
    if (e->AUXB) hipFree(e->AUXB);
    if (e->HYP) hipFree(e->HYP);
    if (e->d_part) hipFree(e->d_part);
    if (e->bg) hipFree(e->bg);
    if (e->bgp) hipFree(e->bgp);
    if (e->bgu) hipFree(e->bgu);
    if (e->d_cmax) hipFree(e->d_cmax);
    if (e->d_cpart) hipFree(e->d_cpart);
    if (e->d_tm) hipFree(e->d_tm);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

#define TRYALLOC(ptr, bytes)                                                               \
    do {                                                                                   \
        hipError_t e_ = hipMalloc((void **)&(ptr), (bytes));                               \
        if (e_ != hipSuccess) {                                                            \
            free_engine(e);                                                                \
            return fail(SMAUG_ENOMEM, "hipMalloc of %zu bytes failed: %s", (size_t)(bytes), \
                        hipGetErrorString(e_));                                            \
        }                                                                                  \
    } while (0)

extern "C" int smaug_create(const smaug_cfg *cfg, smaug_engine **out)
{
    if (!cfg || !out) return fail(SMAUG_EINVAL, "null argument");
    *out = nullptr;
    int rc0 = check_cfg(cfg);
    if (rc0) return rc0;
    int D = cfg->ndim;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(SMAUG_ENODEV, "no HIP device");

    smaug_engine *e = new smaug_engine();
    memset((void *)e, 0, sizeof(*e));
    e->cfg = *cfg;
    e->D = D;
    e->n[0] = cfg->n[0];
    e->n[1] = cfg->n[1];
    e->n[2] = D == 3 ? cfg->n[2] : 1;
    e->nc = (idx_t)e->n[0] * e->n[1] * e->n[2];
    e->nint = (idx_t)(e->n[0] - 4) * (e->n[1] - 4) * (D == 3 ? e->n[2] - 4 : 1);
    e->prec = cfg->precision;
    e->tsz = e->prec == SMAUG_FP64 ? sizeof(double) : sizeof(float);
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        free_engine(e);
        return fail(SMAUG_EHIP, "hipStreamCreate: %s", hipGetErrorString(he));
    }
    int nve = 2 + 2 * D, nbg = 2 + D;
    int npair = 2 * D + D * D + D * (D - 1);
    for (int k = 0; k < 3; k++) TRYALLOC(e->U[k], e->tsz * nve * e->nc);
    TRYALLOC(e->ACC, e->tsz * nve * e->nc);
    if (cfg->hyperdif) {
        TRYALLOC(e->R, e->tsz * npair * e->nc);
        // ratios outside the boxes the ratio kernels write stay 0: the 3-D
        // source kernel multiplies some of them by a zero factor
        if (hipMemset(e->R, 0, e->tsz * npair * e->nc) != hipSuccess) {
            free_engine(e);
            return fail(SMAUG_EHIP, "hipMemset of the ratio buffer failed");
        }
    }
    TRYALLOC(e->AUX, e->tsz * (3 + D) * e->nc);
    if (D == 2) TRYALLOC(e->AUXB, e->tsz * (3 + D) * e->nc);
    e->aux_in = e->AUX;
    if (D == 3 && cfg->hyperdif) TRYALLOC(e->HYP, e->tsz * nve * e->nc);
    TRYALLOC(e->bg, sizeof(double) * nbg * e->nc);
    if (D == 3) TRYALLOC(e->bgp, sizeof(double) * nbg * (size_t)e->n[0] * e->n[1]);
    TRYALLOC(e->bgu, sizeof(double) * 8);
    TRYALLOC(e->d_cmax, sizeof(unsigned long long) * 4);
    {
        idx_t nb = prep_blocks(e->nc);
        if (D == 3) {
            const idx_t ncol = ((idx_t)e->n[0] * e->n[1] + 255) / 256 *
                               ((e->n[2] + PREP_KC - 1) / PREP_KC);
            if (ncol > nb) nb = ncol;
            // the folded aux pass: z-stream blocks at depth 1 and the ghost cells
            const idx_t tiles = ((idx_t)(e->n[0] - 4 + ZS_TI - 1) / ZS_TI) *
                                ((e->n[1] - 4 + ZS_TJ - 1) / ZS_TJ);
            const idx_t nzs = 8 * ((tiles + 7) / 8) * (e->n[2] - 4) +
                              (ghost_count(e->n[0], e->n[1], e->n[2]) + 255) / 256;
            if (nzs > nb) nb = nzs;
        } else {
            // the folded aux pass of the 2-D update: its blocks for any XCD
            // panel (8 ceil(np / 8) tpx tpy <= 11 tiles) and the ghost cells
            const idx_t tiles = ((idx_t)(e->n[0] - 4 + 63) / 64) * ((e->n[1] - 4 + 3) / 4);
            const idx_t nup = 11 * tiles + (ghost_count2(e->n[0], e->n[1]) + 255) / 256;
            if (nup > nb) nb = nup;
        }
        e->ncpart = (int)nb;
        TRYALLOC(e->d_cpart, sizeof(double) * 4 * nb);
    }
    TRYALLOC(e->d_tm, sizeof(double) * 2);
    for (int d = 0; d < 3; d++) {
        int nd = e->n[d];
        std::vector<double> dx, idx, prof, profq;
        axis_arrays(cfg, d, nd, dx, idx, prof, profq);
        TRYALLOC(e->d_dx[d], sizeof(double) * nd);
        TRYALLOC(e->d_idx[d], sizeof(double) * nd);
        TRYALLOC(e->d_prof[d], sizeof(double) * nd);
        TRYALLOC(e->d_profq[d], sizeof(double) * nd);
        hipMemcpy(e->d_profq[d], profq.data(), sizeof(double) * nd, hipMemcpyHostToDevice);
        hipMemcpy(e->d_dx[d], dx.data(), sizeof(double) * nd, hipMemcpyHostToDevice);
        hipMemcpy(e->d_idx[d], idx.data(), sizeof(double) * nd, hipMemcpyHostToDevice);
        hipMemcpy(e->d_prof[d], prof.data(), sizeof(double) * nd, hipMemcpyHostToDevice);
    }
    Geo &g = e->geo;
    for (int d = 0; d < 3; d++) {
        g.n[d] = e->n[d];
        g.per[d] = (d < D) ? (cfg->bc[d] == 0) : 1;
        g.dx[d] = e->d_dx[d];
        g.idx[d] = e->d_idx[d];
        g.prof[d] = e->d_prof[d];
        g.profq[d] = e->d_profq[d];
    }
    g.s1 = e->n[0];
    g.s2 = (idx_t)e->n[0] * e->n[1];
    g.nc = e->nc;
    g.bgf = g.nc;
    g.bgm = ~0u;
    Phys &p = e->phys;
    p.gamma = cfg->gamma;
    for (int d = 0; d < 3; d++) p.g[d] = cfg->g[d];
    for (int f = 0; f < 13; f++) p.chyp[f] = cfg->chyp[f];
    for (int k = 0; k < 8; k++) p.drv[k] = cfg->drv[k];
    p.hyperdif = cfg->hyperdif;
    p.driver = cfg->driver;
    e->zstream = D == 3;
    e->kz = 0;
    e->hyper_lds = 2;
    e->kzh = 0;
    e->ratio_zs = 2; // 3-D: in-plane pairs in the source kernel; 2-D: every pair in the update
    e->kzr = 0;
    e->bgp_on = 1;
    *out = e;
    return 0;
}

extern "C" int smaug_destroy(smaug_engine *e)
{
    if (e) {
        hipStreamSynchronize(e->stream);
        free_engine(e);
    }
    return 0;
}

extern "C" long long smaug_ncell(const smaug_engine *e) { return e ? e->nc : 0; }
extern "C" long long smaug_ninterior(const smaug_engine *e) { return e ? e->nint : 0; }

template <typename T>
static void to_t(std::vector<char> &buf, const double *w, idx_t n)
{
    buf.resize(sizeof(T) * n);
    T *b = (T *)buf.data();
    for (idx_t i = 0; i < n; i++) b[i] = (T)w[i];
}

template <int D, typename T>
static int bc_only(smaug_engine *e, void *u)
{
    return launch_bc<D, T>(e, u);
}

static void drop_graphs(smaug_engine *e);

// 1 if every background field of a 3-D grid is the same on every k-plane
// (bit for bit), as for a stratified atmosphere with k horizontal
static int bg_k_invariant(const smaug_engine *e, const double *bg)
{
    if (e->D != 3) return 0;
    const size_t s2 = (size_t)e->n[0] * e->n[1];
    for (int f = 0; f < 2 + e->D; f++) {
        const double *b = bg + (size_t)f * e->nc;
        for (int k = 1; k < e->n[2]; k++)
            if (memcmp(b, b + k * s2, s2 * sizeof(double))) return 0;
    }
    return 1;
}

// 1 if every background field has the same value (bit for bit) in every cell
static int bg_uniform(int nbg, idx_t nc, const double *bg)
{
    for (int f = 0; f < nbg; f++) {
        const double *b = bg + (size_t)f * nc;
        for (idx_t c = 1; c < nc; c++)
            if (memcmp(&b[c], &b[0], sizeof(double))) return 0;
    }
    return 1;
}

extern "C" int smaug_set_state(smaug_engine *e, const double *w, const double *bg)
{
    if (!e || !w || !bg) return fail(SMAUG_EINVAL, "null argument");
    {
        const int rc = check_periodic_bg(e->D, e->n, e->cfg.bc, bg);
        if (rc) return rc;
    }
    e->aux0_ok = 0; // U0 is written below, ahead of the dispatched ghost fill
    int nve = 2 + 2 * e->D, nbg = 2 + e->D;
    {
        const int ok = bg_uniform(nbg, e->nc, bg);
        if (ok) {
            double v[8];
            for (int f = 0; f < nbg; f++) v[f] = bg[(size_t)f * e->nc];
            HIPCHK(hipMemcpy(e->bgu, v, sizeof(double) * nbg, hipMemcpyHostToDevice));
        }
        if (ok != e->bgu_ok) drop_graphs(e); // the captured launches read the other copy
        e->bgu_ok = ok;
    }
    if (e->D == 3) {
        const int ok = bg_k_invariant(e, bg);
        if (ok) {
            const size_t s2 = (size_t)e->n[0] * e->n[1];
            for (int f = 0; f < nbg; f++)
                HIPCHK(hipMemcpyAsync(e->bgp + f * s2, bg + (size_t)f * e->nc, s2 * sizeof(double),
                                      hipMemcpyHostToDevice, e->stream));
        }
        if (ok != e->bgp_ok) drop_graphs(e); // the captured launches read the other copy
        e->bgp_ok = ok;
    }
    idx_t n = (idx_t)nve * e->nc;
    std::vector<char> buf;
    if (e->prec == SMAUG_FP64)
        to_t<double>(buf, w, n);
    else
        to_t<float>(buf, w, n);
    HIPCHK(hipMemcpyAsync(e->U[0], buf.data(), buf.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->bg, bg, sizeof(double) * nbg * e->nc, hipMemcpyHostToDevice,
                          e->stream));
    int rc = DISPATCH(e, bc_only, e, e->U[0]);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

extern "C" int smaug_get_state(smaug_engine *e, double *w)
{
    if (!e || !w) return fail(SMAUG_EINVAL, "null argument");
    int nve = 2 + 2 * e->D;
    idx_t n = (idx_t)nve * e->nc;
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->prec == SMAUG_FP64) {
        HIPCHK(hipMemcpy(w, e->U[0], sizeof(double) * n, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> b(n);
        HIPCHK(hipMemcpy(b.data(), e->U[0], sizeof(float) * n, hipMemcpyDeviceToHost));
        for (idx_t i = 0; i < n; i++) w[i] = b[i];
    }
    return 0;
}

static int set_time(smaug_engine *e, double t, double dt)
{
    double tm[2] = {t, dt};
    HIPCHK(hipMemcpyAsync(e->d_tm, tm, sizeof(tm), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

template <int D, typename T>
static int do_rhs(smaug_engine *e, double t, double *rhs)
{
    int rc = set_time(e, t, 0.0);
    if (rc) return rc;
    StageArgs sa;
    sa.mode = 2;
    sa.stage = 0;
    sa.ct = 0.0;
    sa.tm = e->d_tm;
    int nve = 2 + 2 * D;
    HIPCHK(hipMemsetAsync(e->U[1], 0, sizeof(T) * nve * e->nc, e->stream));
    rc = stage<D, T>(e, sa, e->U[0], e->U[0], e->U[1], 0, 0);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    idx_t n = (idx_t)nve * e->nc;
    std::vector<T> b(n);
    HIPCHK(hipMemcpy(b.data(), e->U[1], sizeof(T) * n, hipMemcpyDeviceToHost));
    for (idx_t i = 0; i < n; i++) rhs[i] = (double)b[i];
    return 0;
}

extern "C" int smaug_rhs(smaug_engine *e, double t, double *rhs)
{
    if (!e || !rhs) return fail(SMAUG_EINVAL, "null argument");
    return DISPATCH(e, do_rhs, e, t, rhs);
}

template <int D, typename T>
static int do_cmax(smaug_engine *e, double *cm)
{
    int rc = launch_cmax<D, T>(e, e->U[0]);
    if (rc) return rc;
    unsigned long long b[4];
    HIPCHK(hipMemcpyAsync(b, e->d_cmax, sizeof(b), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int d = 0; d < 3; d++) {
        double v = 0.0;
        if (d < D) memcpy(&v, &b[d], sizeof(double));
        cm[d] = v;
    }
    return 0;
}

extern "C" int smaug_cmax(smaug_engine *e, double cmax[3])
{
    if (!e || !cmax) return fail(SMAUG_EINVAL, "null argument");
    return DISPATCH(e, do_cmax, e, cmax);
}

template <int D, typename T>
static int do_courant(smaug_engine *e, double *out)
{
    const unsigned nb = (unsigned)prep_blocks(e->nc);
    const double *bg;
    const Geo g = pc_geo(e, &bg);
    hipLaunchKernelGGL((k_courant<D, T>), dim3(nb), dim3(256), 0, e->stream, g, e->phys,
                       (const T *)e->U[0], bg, e->d_cpart);
    HIPCHK(hipGetLastError());
    // one value, reduced into the spare fourth cmax slot
    hipLaunchKernelGGL(k_cmax_reduce<1>, dim3(1), dim3(1024), 0, e->stream,
                       (const double *)e->d_cpart, (int)nb, e->d_cmax + 3);
    HIPCHK(hipGetLastError());
    unsigned long long b;
    HIPCHK(hipMemcpyAsync(&b, e->d_cmax + 3, sizeof(b), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    memcpy(out, &b, sizeof(double));
    return 0;
}

// max over cells and dims of cfast_d / dx_d at the cell (the Courant limit
// of cucomputemaxcourant over every dim), exact on non-uniform grids
extern "C" int smaug_maxcourant(smaug_engine *e, double *out)
{
    if (!e || !out) return fail(SMAUG_EINVAL, "null argument");
    return DISPATCH(e, do_courant, e, out);
}

template <int D, typename T>
static int capture(smaug_engine *e, int which)
{
    if (e->gvalid[which]) return 0;
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = which == 0 ? rk4_step<D, T>(e, nullptr) : euler_step<D, T>(e);
    hipError_t ce = hipStreamEndCapture(e->stream, &graph);
    if (rc) return rc;
    if (ce != hipSuccess) return fail(SMAUG_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(&e->gexec[which], graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(SMAUG_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    e->gvalid[which] = true;
    return 0;
}

template <int D, typename T>
static int run_steps(smaug_engine *e, int which, double t0, double dt, int nsteps)
{
    if (which == 1) {
        // the Euler step swaps buffers on the host, so it runs eagerly
        e->aux0_ok = 0;
        int rc = set_time(e, t0, dt);
        if (rc) return rc;
        for (int s = 0; s < nsteps; s++)
            if ((rc = euler_step<D, T>(e))) return rc;
        HIPCHK(hipStreamSynchronize(e->stream));
        return 0;
    }
    int rc = capture<D, T>(e, which);
    if (rc) return rc;
    if ((rc = set_time(e, t0, dt))) return rc;
    if ((rc = ensure_aux0<D, T>(e))) return rc;
    for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(e->gexec[which], e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

extern "C" int smaug_rk4(smaug_engine *e, double t0, double dt, int nsteps)
{
    if (!e || nsteps < 0) return fail(SMAUG_EINVAL, "bad argument");
    return DISPATCH_KEEP(e, run_steps, e, 0, t0, dt, nsteps);
}

extern "C" int smaug_euler(smaug_engine *e, double t0, double dt, int nsteps)
{
    if (!e || nsteps < 0) return fail(SMAUG_EINVAL, "bad argument");
    return DISPATCH(e, run_steps, e, 1, t0, dt, nsteps);
}

template <int D, typename T>
static int time_rk4(smaug_engine *e, double t0, double dt, int nsteps, double *total_ms,
                    double *stage_ms)
{
    int rc = capture<D, T>(e, 0);
    if (rc) return rc;
    if ((rc = set_time(e, t0, dt))) return rc;
    if ((rc = ensure_aux0<D, T>(e))) return rc;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, e->stream));
    for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(e->gexec[0], e->stream));
    HIPCHK(hipEventRecord(b, e->stream));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    *total_ms = ms;
    // per-launch duration of the stage-update kernel: one more eager step
    // with events around each k_update launch on the engine stream
    std::vector<hipEvent_t> evs(8);
    for (auto &v : evs) HIPCHK(hipEventCreate(&v));
    if ((rc = rk4_step<D, T>(e, &evs))) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    double sum = 0;
    for (int s = 0; s < 4; s++) {
        float m = 0;
        HIPCHK(hipEventElapsedTime(&m, evs[2 * s], evs[2 * s + 1]));
        sum += m;
    }
    *stage_ms = sum / 4;
    for (auto &v : evs) hipEventDestroy(v);
    hipEventDestroy(a);
    hipEventDestroy(b);
    return 0;
}

// one eager RK4 step with events between the passes of every stage: the
// mean duration of each pass over the four stages
template <int D, typename T>
static int time_passes(smaug_engine *e, double t0, double dt, double *ms)
{
    int rc = set_time(e, t0, dt);
    if (rc) return rc;
    if ((rc = ensure_aux0<D, T>(e))) return rc;
    std::vector<hipEvent_t> evk(24);
    for (auto &v : evk) HIPCHK(hipEventCreate(&v));
    if ((rc = rk4_step<D, T>(e, nullptr, &evk))) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int p = 0; p < 5; p++) {
        double sum = 0;
        for (int s = 0; s < 4; s++) {
            float m = 0;
            HIPCHK(hipEventElapsedTime(&m, evk[6 * s + p], evk[6 * s + p + 1]));
            sum += m;
        }
        ms[p] = sum / 4;
    }
    for (auto &v : evk) hipEventDestroy(v);
    return 0;
}

extern "C" int smaug_time_passes(smaug_engine *e, double t0, double dt, double ms[5])
{
    if (!e || !ms) return fail(SMAUG_EINVAL, "bad argument");
    return DISPATCH(e, time_passes, e, t0, dt, ms);
}

extern "C" int smaug_time_rk4(smaug_engine *e, double t0, double dt, int nsteps,
                              double *total_ms, double *stage_kernel_ms)
{
    if (!e || nsteps < 1 || !total_ms || !stage_kernel_ms) return fail(SMAUG_EINVAL, "bad argument");
    return DISPATCH(e, time_rk4, e, t0, dt, nsteps, total_ms, stage_kernel_ms);
}

// ---------------------------------------------------------------------------
// launch-shape tuning
// ---------------------------------------------------------------------------
static void drop_graphs(smaug_engine *e)
{
    e->aux0_ok = 0;
    for (int k = 0; k < 2; k++)
        if (e->gvalid[k]) {
            hipGraphExecDestroy(e->gexec[k]);
            e->gvalid[k] = false;
        }
}

extern "C" int smaug_set_panel(smaug_engine *e, int pass, int tpx, int tpy)
{
    if (!e || pass < 0 || pass > 1) return fail(SMAUG_EINVAL, "bad argument");
    e->panel[pass][0] = tpx;
    e->panel[pass][1] = tpy;
    drop_graphs(e);
    return 0;
}

extern "C" int smaug_get_panel(smaug_engine *e, int pass, int *tpx, int *tpy)
{
    if (!e || pass < 0 || pass > 1 || !tpx || !tpy) return fail(SMAUG_EINVAL, "bad argument");
    unsigned nb;
    const int D = e->D;
    TileMap m = pass == 1 ? update_map(e, &nb)
                          : make_map(e->n[0] - 3, e->n[1] - 3, D == 3 ? e->n[2] - 3 : 1,
                                     e->panel[0][0], e->panel[0][1], &nb);
    *tpx = m.tpx;
    *tpy = m.tpy;
    return 0;
}

// mean duration (ms) of reps launches of one pass with the current panel;
// the inputs are the current state, the outputs go to the stage buffers,
// which every RK4 step rewrites before reading
template <int D, typename T>
static int time_pass(smaug_engine *e, int pass, int reps, double *ms)
{
    int rc = launch_cmax<D, T>(e, e->U[0]);
    if (!rc && e->phys.hyperdif) rc = launch_ratio<D, T>(e, e->U[0]);
    if (rc) return rc;
    StageArgs sa;
    sa.mode = 0;
    sa.stage = 1;
    sa.ct = 0.5;
    sa.tm = e->d_tm;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    for (int r = -1; r < reps; r++) {
        if (r == 0) HIPCHK(hipEventRecord(a, e->stream));
        if (pass == 0) {
            rc = launch_ratio<D, T>(e, e->U[0]);
        } else {
            rc = launch_pre_update<D, T>(e, e->U[0]);
            if (!rc) rc = launch_stage_update<D, T>(e, sa, e->U[0], e->U[0], e->U[1]);
        }
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(b, e->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    *ms = f / reps;
    hipEventDestroy(a);
    hipEventDestroy(b);
    return 0;
}

extern "C" int smaug_set_zstream(smaug_engine *e, int enable, int kz)
{
    if (!e || kz < 0) return fail(SMAUG_EINVAL, "bad argument");
    if (enable && e->D != 3) return fail(SMAUG_EINVAL, "the z-streamed update is for 3-D grids");
    e->zstream = enable ? 1 : 0;
    // hyperdiffusive sources: 1 z-streamed (the default), 2 per cell, 3 plane
    // in LDS
    e->hyper_lds = enable == 2 ? 0 : (enable == 3 ? 1 : 2);
    e->kz = kz;
    e->kzh = kz;
    drop_graphs(e);
    return 0;
}

extern "C" int smaug_time_pass(smaug_engine *e, int pass, int reps, double *ms)
{
    if (!e || pass < 0 || pass > 1 || reps < 1 || !ms) return fail(SMAUG_EINVAL, "bad argument");
    if (pass == 0 && !e->phys.hyperdif) return fail(SMAUG_EINVAL, "no ratio pass without hyperdiffusion");
    return DISPATCH(e, time_pass, e, pass, reps, ms);
}

// picks, for the ratio and update passes, the fastest XCD panel among
// tpx in {1, 2, 4, 8} and tpy in {1, 2, 4, 8, 16, 32}; for 3-D grids the
// ratio pass also tries the z-streamed ratio kernel at depths 8..128 and
// the update pass the z-streamed update at depths 8..128 with either
// z-streamed source kernel, and each keeps whichever form is fastest; then
// ratio form 2 at depths 4..32 is kept if the two passes together take less
// time with it
extern "C" int smaug_autotune(smaug_engine *e, int reps, int best[2][2], double best_ms[2])
{
    if (!e || reps < 1) return fail(SMAUG_EINVAL, "bad argument");
    static const int TX[] = {1, 2, 4, 8}, TY[] = {1, 2, 4, 8, 16, 32};
    const int zs0 = e->zstream;
    e->zstream = 0;
    e->ratio_zs = 0;
    double bms[2] = {0, 0};
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 0 && !e->phys.hyperdif) {
            if (best) best[0][0] = best[0][1] = 0;
            if (best_ms) best_ms[0] = 0;
            continue;
        }
        double bm = 1e30;
        int bx = 0, by = 0;
        for (int tx : TX)
            for (int ty : TY) {
                int gx, gy;
                e->panel[pass][0] = tx;
                e->panel[pass][1] = ty;
                smaug_get_panel(e, pass, &gx, &gy);
                if (gx != tx || gy != ty) continue; // clamped duplicate
                double ms;
                int rc = DISPATCH(e, time_pass, e, pass, reps, &ms);
                if (rc) return rc;
                if (ms < bm) {
                    bm = ms;
                    bx = tx;
                    by = ty;
                }
            }
        e->panel[pass][0] = bx;
        e->panel[pass][1] = by;
        if (pass == 0 && e->D == 3) {
            static const int KZ[] = {8, 16, 32, 64, 128};
            int bkz = -1;
            e->ratio_zs = 1;
            for (int kz : KZ) {
                e->kzr = kz;
                double ms;
                int rc = DISPATCH(e, time_pass, e, pass, reps, &ms);
                if (rc) return rc;
                if (ms < bm) {
                    bm = ms;
                    bkz = kz;
                }
            }
            e->ratio_zs = bkz > 0;
            e->kzr = bkz > 0 ? bkz : 0;
        }
        if (pass == 1 && e->D == 3) {
            // z-streamed update with the z-streamed (hyper_lds 2) or the
            // LDS-plane (1) hyperdiffusive-source kernel
            static const int KZ[] = {8, 16, 32, 64, 128};
            int bkz = -1, bhy = 2;
            e->zstream = 1;
            for (int hy = 2; hy >= 1; hy--) {
                e->hyper_lds = hy;
                for (int kz : KZ) {
                    e->kz = kz;
                    e->kzh = kz;
                    double ms;
                    int rc = DISPATCH(e, time_pass, e, pass, reps, &ms);
                    if (rc) return rc;
                    if (ms < bm) {
                        bm = ms;
                        bkz = kz;
                        bhy = hy;
                    }
                }
            }
            e->zstream = bkz > 0;
            e->kz = e->kzh = bkz > 0 ? bkz : 0;
            e->hyper_lds = bhy;
        }
        if (best) {
            best[pass][0] = bx;
            best[pass][1] = by;
        }
        bms[pass] = bm;
    }
    // ratio form 2 (the pairs along k in the ratio pass, the in-plane pairs
    // formed by the z-streamed source kernel) against the tuned form: kept
    // where the ratio pass and the update pass with its sources take less
    // time together
    if (e->D == 3 && e->phys.hyperdif && e->zstream && e->hyper_lds == 2) {
        static const int KZ[] = {4, 6, 8, 16, 32};
        const int rz = e->ratio_zs, kzr = e->kzr;
        double bt = bms[0] + bms[1], b0 = 0, b1 = 0;
        int bk = -1;
        e->ratio_zs = 2;
        for (int kz : KZ) {
            e->kzr = kz;
            double m0, m1;
            int rc = DISPATCH(e, time_pass, e, 0, reps, &m0);
            if (!rc) rc = DISPATCH(e, time_pass, e, 1, reps, &m1);
            if (rc) return rc;
            if (m0 + m1 < bt) {
                bt = m0 + m1;
                bk = kz;
                b0 = m0;
                b1 = m1;
            }
        }
        if (bk > 0) {
            e->kzr = bk;
            bms[0] = b0;
            bms[1] = b1;
        } else {
            e->ratio_zs = rz;
            e->kzr = kzr;
        }
    }
    // 2-D: the ratios formed by the update kernel (form 2) against the ratio
    // pass and the update, each at its tuned panel
    if (e->D != 3) {
        e->zstream = zs0;
        e->ratio_zs = 0;
        if (e->phys.hyperdif) {
            e->ratio_zs = 2;
            double m1;
            int rc = DISPATCH(e, time_pass, e, 1, reps, &m1);
            if (rc) return rc;
            if (m1 < bms[0] + bms[1]) {
                bms[0] = 0;
                bms[1] = m1;
            } else {
                e->ratio_zs = 0;
            }
        }
    }
    if (best_ms) {
        best_ms[0] = bms[0];
        best_ms[1] = bms[1];
    }
    drop_graphs(e);
    return 0;
}

extern "C" int smaug_set_ratio_form(smaug_engine *e, int zstreamed, int kz)
{
    if (!e || kz < 0 || zstreamed < 0 || zstreamed > 2) return fail(SMAUG_EINVAL, "bad argument");
    if (zstreamed == 1 && e->D != 3) return fail(SMAUG_EINVAL, "the z-streamed ratios are for 3-D grids");
    e->ratio_zs = zstreamed;
    e->kzr = kz;
    drop_graphs(e);
    return 0;
}

extern "C" int smaug_get_ratio_form(smaug_engine *e, int *zstreamed, int *kz)
{
    if (!e || !zstreamed || !kz) return fail(SMAUG_EINVAL, "null argument");
    unsigned nb;
    *zstreamed = e->ratio_zs;
    *kz = e->D == 3 && *zstreamed ? make_rmap(e, &nb).kz : 0;
    return 0;
}

extern "C" int smaug_set_bg_plane(smaug_engine *e, int enable)
{
    if (!e) return fail(SMAUG_EINVAL, "null argument");
    e->bgp_on = enable ? 1 : 0;
    drop_graphs(e);
    return 0;
}

extern "C" int smaug_get_bg_plane(smaug_engine *e, int *active)
{
    if (!e || !active) return fail(SMAUG_EINVAL, "null argument");
    *active = bg_planar(e) ? 1 : 0;
    return 0;
}

extern "C" int smaug_get_zstream(smaug_engine *e, int *enable, int *kz)
{
    if (!e || !enable || !kz) return fail(SMAUG_EINVAL, "bad argument");
    *enable = e->zstream ? (e->hyper_lds == 2 ? 1 : (e->hyper_lds == 1 ? 3 : 2)) : 0;
    unsigned nb;
    *kz = e->D == 3 ? make_zmap(e, &nb).kz : 0;
    return 0;
}

// ---------------------------------------------------------------------------
// conservation diagnostics
// ---------------------------------------------------------------------------
int smaug::field_sums(int D, int prec, const Geo &g, const void *u, hipStream_t st, double *part,
                      double *sums)
{
    const int nve = 2 + 2 * D, nb = sums_blocks();
    if (D == 2) {
        if (prec == SMAUG_FP64)
            hipLaunchKernelGGL((k_sums<2, double>), dim3(nb), dim3(256), 0, st, g, (const double *)u, part);
        else
            hipLaunchKernelGGL((k_sums<2, float>), dim3(nb), dim3(256), 0, st, g, (const float *)u, part);
    } else {
        if (prec == SMAUG_FP64)
            hipLaunchKernelGGL((k_sums<3, double>), dim3(nb), dim3(256), 0, st, g, (const double *)u, part);
        else
            hipLaunchKernelGGL((k_sums<3, float>), dim3(nb), dim3(256), 0, st, g, (const float *)u, part);
    }
    HIPCHK(hipGetLastError());
    std::vector<double> h((size_t)2 * nve * nb);
    HIPCHK(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int f = 0; f < nve; f++) {
        // Neumaier over the per-block (sum, compensation) pairs
        double s = 0, c = 0;
        auto add = [&](double x) {
            const double t = s + x;
            c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
            s = t;
        };
        for (int b = 0; b < nb; b++) {
            add(h[(size_t)b * 2 * nve + 2 * f]);
            c += h[(size_t)b * 2 * nve + 2 * f + 1];
        }
        sums[f] = s + c;
    }
    return 0;
}

extern "C" int smaug_sums(smaug_engine *e, double *sums)
{
    if (!e || !sums) return fail(SMAUG_EINVAL, "null argument");
    if (!e->d_part) HIPCHK(hipMalloc((void **)&e->d_part, sizeof(double) * 2 * 8 * sums_blocks()));
    return field_sums(e->D, e->prec, e->geo, e->U[0], e->stream, e->d_part, sums);
}

// ---------------------------------------------------------------------------
// conservation diagnostics without allocation (smaug_sums_noalloc)
//
// The per-block partials of k_sums live in `scratch`, a static device array
// that every call on a device shares (2 * 8 * sums_blocks() doubles, 64 KiB).
// Host code never passes `scratch` itself to a kernel: its device address is
// looked up once per device with hipGetSymbolAddress and kept in
// g_scratch_ptr.  k_sums writes every partial of every block, so nothing in
// the array needs clearing between calls.
// ---------------------------------------------------------------------------
static constexpr int SCRATCH_DOUBLES = 2 * 8 * sums_blocks();
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

static constexpr int MAX_DEVICES = 64;
static double *g_scratch_ptr[MAX_DEVICES];
static std::mutex g_scratch_mu;

// the device address of `scratch` on the current device (looked up on the
// first call there)
static int scratch_on_current_device(double **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(SMAUG_ENODEV, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// the per-block (sum, compensation) pairs folded field by field in block
// order with Neumaier's step, the same operations in the same order as the
// host pass of field_sums, so both give the same bits
__global__ __launch_bounds__(64) void k_sums_merge(const double *__restrict__ part, int nb,
                                                   int nve, double *__restrict__ out)
{
    const int f = threadIdx.x;
    if (f >= nve) return;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < nb; b++) {
        const double x = part[(size_t)b * 2 * nve + 2 * f];
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
        c += part[(size_t)b * 2 * nve + 2 * f + 1];
    }
    out[f] = s + c;
}

extern "C" int smaug_sums_noalloc(int ndim, const int n[3], int precision, const void *d_w,
                                  double *d_sums, void *stream)
{
    if (!n || !d_w || !d_sums) return fail(SMAUG_EINVAL, "null argument");
    if (ndim != 2 && ndim != 3) return fail(SMAUG_EINVAL, "ndim must be 2 or 3 (got %d)", ndim);
    if (precision != SMAUG_FP64 && precision != SMAUG_MIXED)
        return fail(SMAUG_EINVAL, "precision must be SMAUG_FP64 or SMAUG_MIXED");
    for (int d = 0; d < ndim; d++)
        if (n[d] < 5) return fail(SMAUG_EINVAL, "n[%d]=%d: need >= 5 cells incl. ghosts", d, n[d]);
    const long long nc = (long long)n[0] * n[1] * (ndim == 3 ? n[2] : 1);
    if (nc > MAX_CELLS)
        return fail(SMAUG_EINVAL,
                    "%lld cells exceed the limit of %lld cells per grid of smaug_sums_noalloc", nc,
                    MAX_CELLS);
    const int nve = 2 + 2 * ndim, nb = sums_blocks();
    if (2 * nve * nb > SCRATCH_DOUBLES)
        return fail(SMAUG_EINVAL, "scratch of %d doubles is smaller than the %d partials", SCRATCH_DOUBLES,
                    2 * nve * nb);
    double *part = nullptr;
    int rc = scratch_on_current_device(&part);
    if (rc) return rc;
    Geo g;
    memset((void *)&g, 0, sizeof(g));
    g.n[0] = n[0];
    g.n[1] = n[1];
    g.n[2] = ndim == 3 ? n[2] : 1;
    g.s1 = n[0];
    g.s2 = (idx_t)n[0] * n[1];
    g.nc = nc;
    hipStream_t st = (hipStream_t)stream;
    if (ndim == 2) {
        if (precision == SMAUG_FP64)
            hipLaunchKernelGGL((k_sums<2, double>), dim3(nb), dim3(256), 0, st, g, (const double *)d_w, part);
        else
            hipLaunchKernelGGL((k_sums<2, float>), dim3(nb), dim3(256), 0, st, g, (const float *)d_w, part);
    } else {
        if (precision == SMAUG_FP64)
            hipLaunchKernelGGL((k_sums<3, double>), dim3(nb), dim3(256), 0, st, g, (const double *)d_w, part);
        else
            hipLaunchKernelGGL((k_sums<3, float>), dim3(nb), dim3(256), 0, st, g, (const float *)d_w, part);
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_sums_merge, dim3(1), dim3(64), 0, st, (const double *)part, nb, nve, d_sums);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// measured copy-bandwidth ceiling: a streaming copy of nbytes in which every
// thread keeps U 16-byte loads in flight before it stores them; the probe
// sweeps U and the grid size and reports the fastest
// ---------------------------------------------------------------------------
template <int U>
__global__ __launch_bounds__(256) void k_copy_probe(const float4 *__restrict__ a,
                                                     float4 *__restrict__ b, idx_t n)
{
    const idx_t stride = (idx_t)gridDim.x * 256 * U;
    for (idx_t base = (idx_t)blockIdx.x * 256 * U + threadIdx.x; base < n; base += stride) {
        float4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const idx_t i = base + (idx_t)u * 256;
            if (i < n) v[u] = a[i];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const idx_t i = base + (idx_t)u * 256;
            if (i < n) b[i] = v[u];
        }
    }
}

// many-stream bandwidth: every thread reads one 4-byte value of each of NR
// arrays at its cell and writes one of each of NW arrays there -- the
// access pattern of a stage update that reads the 8 fields of four arrays
// (state, u_n, accumulator, sources) and writes those of two (stage
// result, accumulator) and 4 aux quantities, one cell per thread
template <int NR, int NW>
__global__ __launch_bounds__(256) void k_stream_probe(const float *__restrict__ a,
                                                      float *__restrict__ b, idx_t n)
{
    const idx_t stride = (idx_t)gridDim.x * 256;
    for (idx_t c = (idx_t)blockIdx.x * 256 + threadIdx.x; c < n; c += stride) {
        float v[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) v[r] = a[(idx_t)r * n + c];
        float sum = 0.0f;
#pragma unroll
        for (int r = 0; r < NR; r++) sum += v[r];
#pragma unroll
        for (int w = 0; w < NW; w++) b[(idx_t)w * n + c] = sum + (float)w;
    }
}

constexpr int SP_NR = 32, SP_NW = 20;

extern "C" int smaug_stream_probe(long long ncell, int reps, double *gbs)
{
    if (ncell < 256 || reps < 1 || !gbs) return fail(SMAUG_EINVAL, "bad argument");
    const idx_t n = ncell;
    float *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void **)&a, (size_t)SP_NR * n * sizeof(float)));
    hipError_t he = hipMalloc((void **)&b, (size_t)SP_NW * n * sizeof(float));
    if (he != hipSuccess) {
        hipFree(a);
        return fail(SMAUG_ENOMEM, "hipMalloc: %s", hipGetErrorString(he));
    }
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipMemsetAsync(a, 0, (size_t)SP_NR * n * sizeof(float), st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 0.0;
    const unsigned grids[3] = {256 * 8, 256 * 16, 256 * 32}; // workgroups: 8..32 per CU
    for (int gi = 0; gi < 3; gi++) {
        const unsigned nb = grids[gi];
        hipLaunchKernelGGL((k_stream_probe<SP_NR, SP_NW>), dim3(nb), dim3(256), 0, st, a, b, n);
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL((k_stream_probe<SP_NR, SP_NW>), dim3(nb), dim3(256), 0, st, a, b, n);
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipGetLastError());
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        const double g = (double)(SP_NR + SP_NW) * n * sizeof(float) * reps / (ms / 1e3) / 1e9;
        if (g > best) best = g;
    }
    *gbs = best;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    hipFree(a);
    hipFree(b);
    return 0;
}

// the same 32 read and 20 written FP32 streams walked as the z-streamed
// stage update walks them: blocks of 64 x 4 columns of an n0 x n1 x n2 grid
// (the interior, 2 cells in from each side), each walking a chunk of kz
// planes with the next plane's loads issued before the current plane's
// stores, the tiles of an XCD contiguous (as ZMap's)
template <int NR, int NW>
__global__ __launch_bounds__(256) void k_walk_probe(const float *__restrict__ a,
                                                    float *__restrict__ b, ZMap m, int n0, int n1,
                                                    int n2, idx_t nc)
{
    int tx, ty, kc;
    if (!zmap_block(m, tx, ty, kc)) return;
    const int i = 2 + tx * 64 + (threadIdx.x & 63), j = 2 + ty * 4 + (threadIdx.x >> 6);
    const int k0 = 2 + kc * m.kz, k1 = min(k0 + m.kz, n2 - 2);
    if (i > n0 - 3 || j > n1 - 3 || k0 >= k1) return;
    const idx_t col = (idx_t)i + (idx_t)n0 * j, s2 = (idx_t)n0 * n1;
    float v[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) v[r] = a[(idx_t)r * nc + col + s2 * k0];
    for (int k = k0; k < k1; k++) {
        float sum = 0.0f;
#pragma unroll
        for (int r = 0; r < NR; r++) sum += v[r];
        const int kn = min(k + 1, k1 - 1);
#pragma unroll
        for (int r = 0; r < NR; r++) v[r] = a[(idx_t)r * nc + col + s2 * kn];
#pragma unroll
        for (int w = 0; w < NW; w++) b[(idx_t)w * nc + col + s2 * k] = sum + (float)w;
    }
}

extern "C" int smaug_walk_probe(int n0, int n1, int n2, int kz, int reps, double *gbs)
{
    if (n0 < 8 || n1 < 8 || n2 < 8 || kz < 1 || reps < 1 || !gbs) return fail(SMAUG_EINVAL, "bad argument");
    const idx_t n = (idx_t)n0 * n1 * n2;
    float *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void **)&a, (size_t)SP_NR * n * sizeof(float)));
    hipError_t he = hipMalloc((void **)&b, (size_t)SP_NW * n * sizeof(float));
    if (he != hipSuccess) {
        hipFree(a);
        return fail(SMAUG_ENOMEM, "hipMalloc: %s", hipGetErrorString(he));
    }
    ZMap m;
    m.ntx = (n0 - 4 + 63) / 64;
    m.nty = (n1 - 4 + 3) / 4;
    m.kz = kz < n2 - 4 ? kz : n2 - 4;
    m.nkz = (n2 - 4 + m.kz - 1) / m.kz;
    m.per_xcd = (m.ntx * m.nty + 7) / 8;
    const unsigned nb = (unsigned)(8 * m.per_xcd * m.nkz);
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipMemsetAsync(a, 0, (size_t)SP_NR * n * sizeof(float), st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipLaunchKernelGGL((k_walk_probe<SP_NR, SP_NW>), dim3(nb), dim3(256), 0, st, a, b, m, n0, n1, n2, n);
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL((k_walk_probe<SP_NR, SP_NW>), dim3(nb), dim3(256), 0, st, a, b, m, n0, n1,
                           n2, n);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipGetLastError());
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    const double interior = (double)(n0 - 4) * (n1 - 4) * (n2 - 4);
    *gbs = (double)(SP_NR + SP_NW) * interior * sizeof(float) * reps / (ms / 1e3) / 1e9;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    hipFree(a);
    hipFree(b);
    return 0;
}

extern "C" int smaug_copy_probe(long long nbytes, int reps, double *gbs)
{
    if (nbytes < 16 || reps < 1 || !gbs) return fail(SMAUG_EINVAL, "bad argument");
    idx_t n = nbytes / 16;
    float4 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void **)&a, n * 16));
    hipError_t he = hipMalloc((void **)&b, n * 16);
    if (he != hipSuccess) {
        hipFree(a);
        return fail(SMAUG_ENOMEM, "hipMalloc: %s", hipGetErrorString(he));
    }
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipMemsetAsync(a, 0, n * 16, st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 0.0;
    const unsigned grids[3] = {256 * 8, 256 * 16, 256 * 32}; // workgroups: 8..32 per CU
    for (int ui = 0; ui < 3; ui++)
        for (int gi = 0; gi < 3; gi++) {
            const unsigned nb = grids[gi];
            auto launch = [&]() {
                if (ui == 0)
                    hipLaunchKernelGGL((k_copy_probe<1>), dim3(nb), dim3(256), 0, st, a, b, n);
                else if (ui == 1)
                    hipLaunchKernelGGL((k_copy_probe<4>), dim3(nb), dim3(256), 0, st, a, b, n);
                else
                    hipLaunchKernelGGL((k_copy_probe<8>), dim3(nb), dim3(256), 0, st, a, b, n);
            };
            launch(); // warm-up
            HIPCHK(hipEventRecord(e0, st));
            for (int r = 0; r < reps; r++) launch();
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipGetLastError());
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            const double g = 2.0 * (double)n * 16 * reps / (ms / 1e3) / 1e9;
            if (g > best) best = g;
        }
    *gbs = best;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    hipFree(a);
    hipFree(b);
    return 0;
}
