// smaug_ensemble.hip -- many independent grids of uneven sizes advanced
// together by persistent, queue-driven kernels.
//
// Every pass of an RK stage (prep, ratio, update, ghost fill) is one
// persistent launch over the whole ensemble: its work items are runs of
// consecutive cells of every grid, numbered grid by grid, and the blocks
// pull them from a device queue.  The item range is cut into 8 contiguous
// slices, one per XCD; a block first drains the slice of its own XCD
// (blocks are dispatched to XCDs round-robin, so block b runs on XCD b % 8)
// and then steals from the others.  Neighbouring tiles of one run are thus
// mostly processed on one XCD and share its L2.  Each run keeps its own
// geometry, cmax, time and dt; all runs share the physics parameters.
// One RK4 step (a counter reset, 4 stages x 4 passes and the time advance)
// is captured once into a hipGraph and replayed.
//
// The per-cell arithmetic is the single-grid engine's (prep_cell,
// ratio_cell, update_cell, bc_elem), so an ensemble member evolves
// bit-for-bit like the same grid stepped alone.
#include "smaug_kernels.hpp"
#include "smaug_host.hpp"

#include <stdio.h>
#include <string.h>
#include <vector>

using namespace smaug;

namespace {

// Work-item sizes.  An item costs a queue atomic, a block-wide pull and the
// run's descriptor loads, so items are several cells per thread; the update
// items stay smaller, as the heaviest pass needs the finer balance.
constexpr int PREP_CELLS = 4096;   // cells per prep item (16 per thread)
constexpr int RATIO_CELLS = 1024;  // ratio-box cells per ratio item (4 per thread)
constexpr int UPDATE_CELLS = 512;  // interior cells per update item (2 per thread)
constexpr int BC_CELLS = 1024;     // ghost-slab elements per item (4 per thread)

// t / d for t < 2^31 as (t * m) >> s with s = 31 + ceil(log2 d) and
// m = floor(2^s / d) + 1: the error of the reciprocal, m d - 2^s <= d <=
// 2^(s-31), keeps t (m d - 2^s) < 2^s, so the quotient is exact
struct FastDiv {
    unsigned long long m;
    int s;
};

static FastDiv make_fastdiv(unsigned d)
{
    int c = 0;
    while ((1ull << c) < d) c++;
    FastDiv f;
    f.s = 31 + c;
    f.m = (unsigned long long)((((unsigned __int128)1) << f.s) / d) + 1;
    return f;
}

__device__ __forceinline__ unsigned fdiv(unsigned t, const FastDiv &f)
{
    return (unsigned)(((unsigned long long)t * f.m) >> f.s);
}
constexpr int NPASS = 4;         // prep, ratio, update, ghost fill
constexpr int NLIST = NPASS + 1; // their item lists, and the 64 x u2_tj tiles of the 2-D
                                 // update that forms the ratios itself (ke_update2r,
                                 // one block per item, no queue)
constexpr int NQ = 8;            // queue slices, one per XCD
// slice counters lie 4 KiB apart: atomics on one cache line (or one memory
// channel) are served one after another, and every pull of a pass, as well
// as every block's probes of drained slices at its end, is such an atomic
constexpr int CSTRIDE = 1024;

template <typename T>
struct RunDesc {
    Geo geo;
    T *U[3];
    T *ACC, *R, *AUX;
    const double *bg;
    unsigned long long *cm; // [4 stages][4] cmax bit patterns
    double *tm;             // [t, dt]
    int ext[2][3];          // extents of the ratio box [1, n-3] (0) and the interior (1)
    FastDiv dv[2][2];       // division by ext[b][0] and ext[b][1]
};

struct Queue {
    const int *base;   // nrun + 1 prefix sums of the items of each run
    const int *run;    // the run of every item
    int nrun, nitems;
    unsigned *ctr;     // NQ slice counters, CSTRIDE apart
};

// next item for the calling block (thread 0 only), -1 when every slice is
// drained; v is the number of slices this block has found empty so far
__device__ __forceinline__ int pull(const Queue &q, int xcd, int &v)
{
    while (v < NQ) {
        const int s = (xcd + v) & (NQ - 1);
        const long long lo = ((long long)q.nitems * s) / NQ;
        const long long hi = ((long long)q.nitems * (s + 1)) / NQ;
        const unsigned k = atomicAdd(&q.ctr[(size_t)s * CSTRIDE], 1u);
        if (lo + (long long)k < hi) return (int)(lo + k);
        v++;
    }
    return -1;
}

// The block-uniform next item and its run.  Thread 0 keeps one item (and
// the load of its run) pulled ahead, so the latency of the queue atomic and
// of the run lookup overlaps the processing of the current item; s_it[1]
// starts at -2 ("nothing pulled yet").  Two barriers so that s_it is not
// rewritten before every thread has read it.
__device__ __forceinline__ int block_pull(const Queue &q, int &v, int *s_it, int &run)
{
    if (threadIdx.x == 0) {
        const int xcd = blockIdx.x & (NQ - 1);
        int cur = s_it[1], cr = s_it[3];
        if (cur == -2) {
            cur = pull(q, xcd, v);
            cr = cur >= 0 ? q.run[cur] : 0;
        }
        s_it[0] = cur;
        s_it[2] = cr;
        const int nx = cur >= 0 ? pull(q, xcd, v) : -1;
        s_it[1] = nx;
        s_it[3] = nx >= 0 ? q.run[nx] : 0;
    }
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_it[0]);
    run = __builtin_amdgcn_readfirstlane(s_it[2]);
    __syncthreads();
    return it;
}

// cell t (linear, i fastest) of a box of extents ext starting at lo; a run
// has fewer than 2^31 cells, so the quotients fit the fast division
__device__ __forceinline__ void box_cell(unsigned t, const int *ext, const FastDiv *dv, int lo,
                                         int &x0, int &x1, int &x2)
{
    const unsigned e0 = (unsigned)ext[0], e1 = (unsigned)ext[1];
    const unsigned r = fdiv(t, dv[0]);
    x0 = lo + (int)(t - r * e0);
    const unsigned k = fdiv(r, dv[1]);
    x1 = lo + (int)(r - k * e1);
    x2 = lo + (int)k;
}

__global__ void ke_reset(unsigned *ctr, int nctr, unsigned long long *cm, int ncm)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nctr + ncm;
         i += gridDim.x * blockDim.x) {
        if (i < nctr)
            ctr[(size_t)i * CSTRIDE] = 0u;
        else
            cm[i - nctr] = 0ull;
    }
}

__global__ void ke_advance_time(double *tm, int nrun)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrun) tm[2 * r] += tm[2 * r + 1];
}

template <int D, typename T>
__global__ __launch_bounds__(256) void ke_prep(const RunDesc<T> *__restrict__ runs, Queue q,
                                                Phys ph, int stage, int src)
{
    __shared__ int s_it[4];
    __shared__ double red[3][4];
    int v = 0;
    if (threadIdx.x == 0) s_it[1] = -2;
    // the block's cfast maxima are kept over consecutive items of one run
    // and folded into the run's cmax (one atomic per dim) when the run
    // changes and at the end
    int prev = -1;
    double mx[3] = {0.0, 0.0, 0.0};
    for (;;) {
        int r;
        const int it = block_pull(q, v, s_it, r);
        if (it < 0) break;
        if (r != prev) {
            if (prev >= 0) block_max_to<D>(mx, red, runs[prev].cm + 4 * stage);
#pragma unroll
            for (int d = 0; d < 3; d++) mx[d] = 0.0;
            prev = r;
        }
        const RunDesc<T> &rd = runs[r];
        const idx_t c0 = (idx_t)(it - q.base[r]) * PREP_CELLS;
#pragma unroll 4
        for (int k = 0; k < PREP_CELLS / 256; k++) {
            const idx_t c = c0 + k * 256 + threadIdx.x;
            if (c < rd.geo.nc) prep_cell<D, T>(rd.geo, ph, rd.U[src], rd.bg, rd.AUX, c, mx);
        }
    }
    if (prev >= 0) block_max_to<D>(mx, red, runs[prev].cm + 4 * stage);
}

template <int D, typename T>
__global__ __launch_bounds__(256) void ke_ratio(const RunDesc<T> *__restrict__ runs, Queue q,
                                                 int src)
{
    __shared__ int s_it[4];
    int v = 0;
    if (threadIdx.x == 0) s_it[1] = -2;
    for (;;) {
        int r;
        const int it = block_pull(q, v, s_it, r);
        if (it < 0) break;
        const RunDesc<T> &rd = runs[r];
        const unsigned nb = (unsigned)rd.ext[0][0] * rd.ext[0][1] * rd.ext[0][2];
#pragma unroll 2
        for (int k = 0; k < RATIO_CELLS / 256; k++) {
            const unsigned t = (unsigned)(it - q.base[r]) * RATIO_CELLS + k * 256 + threadIdx.x;
            if (t >= nb) break;
            int x0, x1, x2;
            box_cell(t, rd.ext[0], rd.dv[0], 1, x0, x1, x2);
            ratio_cell<D, T>(rd.geo, rd.U[src], rd.AUX, rd.R, x0, x1, D == 3 ? x2 : 0);
        }
    }
}

template <int D, typename T>
__global__ __launch_bounds__(256) void ke_update(const RunDesc<T> *__restrict__ runs, Queue q,
                                                  Phys ph, int stage, double ct, int src, int dst)
{
    __shared__ int s_it[4];
    int v = 0;
    if (threadIdx.x == 0) s_it[1] = -2;
    for (;;) {
        int r;
        const int it = block_pull(q, v, s_it, r);
        if (it < 0) break;
        const RunDesc<T> &rd = runs[r];
        const unsigned nb = (unsigned)rd.ext[1][0] * rd.ext[1][1] * rd.ext[1][2];
#pragma unroll
        for (int k = 0; k < UPDATE_CELLS / 256; k++) {
            const unsigned t = (unsigned)(it - q.base[r]) * UPDATE_CELLS + k * 256 + threadIdx.x;
            if (t >= nb) break;
            int x0, x1, x2;
            box_cell(t, rd.ext[1], rd.dv[1], 2, x0, x1, x2);
            if (D != 3) x2 = 0;
            StageArgs sa;
            sa.mode = 0;
            sa.stage = stage;
            sa.ct = ct;
            sa.tm = rd.tm;
            // the run's background layout is wave-uniform: read once when
            // it is the same in every cell
            if (rd.geo.bgm == 0u)
                update_cell<D, T, true>(rd.geo, ph, sa, rd.U[src], rd.U[0], rd.U[dst], rd.ACC,
                                        rd.R, rd.AUX, rd.bg, rd.cm + 4 * stage, x0, x1, x2);
            else
                update_cell<D, T, false>(rd.geo, ph, sa, rd.U[src], rd.U[0], rd.U[dst], rd.ACC,
                                         rd.R, rd.AUX, rd.bg, rd.cm + 4 * stage, x0, x1, x2);
        }
    }
}

// 2-D ratio form 2: the update of the 64 x u2_tj-cell tiles of every run, each
// forming the ratio pairs of its tile and a 1-cell halo in LDS as
// k_update2r does (the stage then has no ratio pass).  One block per tile
// item rather than a persistent queue (1-2% faster here: the items are
// equal, so the hardware dispatcher balances them): the items are cut into 8
// contiguous slices and block b takes item b / 8 of slice b mod 8 (its
// XCD's), so the tiles of a run mostly share one L2
template <typename T>
__global__ __launch_bounds__(U2<T>::NT) void ke_update2r(const RunDesc<T> *__restrict__ runs,
                                                      const int *__restrict__ base,
                                                      const int *__restrict__ irun, int nitems,
                                                      Phys ph, int stage, double ct, int src,
                                                      int dst)
{
    __shared__ T sQ[6][U2<T>::QN];
    __shared__ T sR[Lay<2>::NPAIR * U2<T>::RN];
    const int per = (nitems + 7) >> 3;
    const int it = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if ((int)(blockIdx.x >> 3) >= per || it >= nitems) return;
    const int r = irun[it];
    const RunDesc<T> &rd = runs[r];
    const int ntx = (rd.geo.n[0] - 4 + U2<T>::TI - 1) / U2<T>::TI;
    const int t = it - base[r];
    const int ty = t / ntx, tx = t - ty * ntx;
    StageArgs sa;
    sa.mode = 0;
    sa.stage = stage;
    sa.ct = ct;
    sa.tm = rd.tm;
    if (rd.geo.bgm == 0u)
        update2r_tile<T, true, false>(rd.geo, ph, sa, rd.U[src], rd.U[0], rd.U[dst], rd.ACC,
                                      rd.AUX, rd.bg, rd.cm + 4 * stage, tx, ty, sQ, sR, nullptr,
                                      nullptr);
    else
        update2r_tile<T, false, false>(rd.geo, ph, sa, rd.U[src], rd.U[0], rd.U[dst], rd.ACC,
                                       rd.AUX, rd.bg, rd.cm + 4 * stage, tx, ty, sQ, sR, nullptr,
                                       nullptr);
}

template <int D, typename T>
__global__ __launch_bounds__(256) void ke_bc(const RunDesc<T> *__restrict__ runs, Queue q, int dst)
{
    __shared__ int s_it[4];
    int v = 0;
    if (threadIdx.x == 0) s_it[1] = -2;
    for (;;) {
        int r;
        const int it = block_pull(q, v, s_it, r);
        if (it < 0) break;
        const RunDesc<T> &rd = runs[r];
        const idx_t nbc = bc_count<D>(rd.geo.n[0], rd.geo.n[1], rd.geo.n[2]);
#pragma unroll 4
        for (int k = 0; k < BC_CELLS / 256; k++) {
            const idx_t t = (idx_t)(it - q.base[r]) * BC_CELLS + k * 256 + threadIdx.x;
            if (t < nbc) bc_elem<D, T>(rd.geo, rd.U[dst], t);
        }
    }
}

} // namespace

struct smaug_ensemble {
    int nrun, D, prec;
    size_t tsz;
    hipStream_t stream;
    std::vector<smaug_cfg> cfg;
    std::vector<Geo> geo;          // host copies (device pointers inside)
    std::vector<idx_t> nc, nint;
    std::vector<size_t> off_u, off_r, off_aux, off_bg; // byte offsets into the arenas
    char *state;                   // U0, U1, U2, ACC of every run
    char *ratio, *aux;
    double *bg, *axes, *tm;
    double *bgu;                   // nrun x 8: a run's background values when uniform
    unsigned long long *cm;        // nrun x 16
    unsigned *ctr;                 // 16 launches x NQ
    void *descs;                   // RunDesc<T>[nrun]
    int *base[NLIST];              // per-list item prefix sums (device)
    int *irun[NLIST];              // per-list run of every item (device)
    int nitems[NLIST];
    unsigned grid[NLIST];          // persistent grid per list's kernel
    int rform;                     // 2-D: 2 the update forms the ratios (ke_update2r), 0 not
    double *d_part;                // conservation-diagnostic partial sums
    Phys phys;
    hipGraphExec_t gexec;
    bool gvalid;
};

static void free_ens(smaug_ensemble *e)
{
    if (!e) return;
    if (e->gvalid) hipGraphExecDestroy(e->gexec);
    void *ptrs[] = {e->state, e->ratio, e->aux, e->bg, e->axes, e->tm, e->cm, e->ctr, e->descs,
                    e->d_part, e->bgu};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (int k = 0; k < NLIST; k++) {
        if (e->base[k]) hipFree(e->base[k]);
        if (e->irun[k]) hipFree(e->irun[k]);
    }
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static bool same_phys(const smaug_cfg &a, const smaug_cfg &b)
{
    if (a.ndim != b.ndim || a.precision != b.precision || a.hyperdif != b.hyperdif ||
        a.driver != b.driver || a.gamma != b.gamma)
        return false;
    for (int k = 0; k < 3; k++)
        if (a.g[k] != b.g[k]) return false;
    for (int k = 0; k < 13; k++)
        if (a.chyp[k] != b.chyp[k]) return false;
    for (int k = 0; k < 8; k++)
        if (a.drv[k] != b.drv[k]) return false;
    return true;
}

// persistent grid of a pass: resident blocks per CU x CUs, a multiple of 8
template <typename K>
static int pass_grid(K kernel, unsigned *grid)
{
    int dev = 0, ncu = 0, per = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0));
    if (per < 1) per = 1;
    unsigned g = (unsigned)(per * ncu);
    *grid = (g + 7) & ~7u;
    return 0;
}

template <int D, typename T>
static int setup_typed(smaug_ensemble *e)
{
    std::vector<RunDesc<T>> d(e->nrun);
    std::vector<int> base[NLIST];
    for (int k = 0; k < NLIST; k++) base[k].assign(e->nrun + 1, 0);
    const int nve = 2 + 2 * D;
    for (int r = 0; r < e->nrun; r++) {
        RunDesc<T> &rd = d[r];
        const Geo &g = e->geo[r];
        rd.geo = g;
        T *u = (T *)(e->state + e->off_u[r]);
        for (int k = 0; k < 3; k++) rd.U[k] = u + (idx_t)k * nve * g.nc;
        rd.ACC = u + (idx_t)3 * nve * g.nc;
        rd.R = e->phys.hyperdif ? (T *)(e->ratio + e->off_r[r]) : nullptr;
        rd.AUX = (T *)(e->aux + e->off_aux[r]);
        rd.bg = (const double *)((char *)e->bg + e->off_bg[r]);
        rd.cm = e->cm + 16 * r;
        rd.tm = e->tm + 2 * r;
        const int n2 = D == 3 ? g.n[2] : 1;
        // ratio box [1, n-3], interior [2, n-3]
        for (int d = 0; d < 3; d++) {
            rd.ext[0][d] = d < D ? g.n[d] - 3 : 1;
            rd.ext[1][d] = d < D ? g.n[d] - 4 : 1;
        }
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) rd.dv[b][d] = make_fastdiv((unsigned)rd.ext[b][d]);
        const long long nbox[2] = {(long long)rd.ext[0][0] * rd.ext[0][1] * rd.ext[0][2],
                                   (long long)rd.ext[1][0] * rd.ext[1][1] * rd.ext[1][2]};
        if (g.nc >= (1ll << 31)) return fail(SMAUG_EINVAL, "run %d: too many cells for a member", r);
        long long items[NLIST] = {
            (long long)((g.nc + PREP_CELLS - 1) / PREP_CELLS),
            e->phys.hyperdif ? (nbox[0] + RATIO_CELLS - 1) / RATIO_CELLS : 0,
            (nbox[1] + UPDATE_CELLS - 1) / UPDATE_CELLS,
            (long long)((bc_count<D>(g.n[0], g.n[1], n2) + BC_CELLS - 1) / BC_CELLS),
            D == 2 ? (long long)((g.n[0] - 4 + U2<T>::TI - 1) / U2<T>::TI) *
                         ((g.n[1] - 4 + u2_tj<T>() - 1) / u2_tj<T>())
                   : 0};
        for (int k = 0; k < NLIST; k++) {
            long long nb = (long long)base[k][r] + items[k];
            if (nb > 0x7fffffffLL) return fail(SMAUG_EINVAL, "ensemble too large for the work queue");
            base[k][r + 1] = (int)nb;
        }
    }
    HIPCHK(hipMalloc(&e->descs, sizeof(RunDesc<T>) * e->nrun));
    HIPCHK(hipMemcpy(e->descs, d.data(), sizeof(RunDesc<T>) * e->nrun, hipMemcpyHostToDevice));
    for (int k = 0; k < NLIST; k++) {
        e->nitems[k] = base[k][e->nrun];
        HIPCHK(hipMalloc((void **)&e->base[k], sizeof(int) * (e->nrun + 1)));
        HIPCHK(hipMemcpy(e->base[k], base[k].data(), sizeof(int) * (e->nrun + 1),
                         hipMemcpyHostToDevice));
        std::vector<int> ir(e->nitems[k] > 0 ? e->nitems[k] : 1, 0);
        for (int r = 0; r < e->nrun; r++)
            for (int it = base[k][r]; it < base[k][r + 1]; it++) ir[it] = r;
        HIPCHK(hipMalloc((void **)&e->irun[k], sizeof(int) * ir.size()));
        HIPCHK(hipMemcpy(e->irun[k], ir.data(), sizeof(int) * ir.size(), hipMemcpyHostToDevice));
    }
    int rc;
    if ((rc = pass_grid(ke_prep<D, T>, &e->grid[0]))) return rc;
    if ((rc = pass_grid(ke_ratio<D, T>, &e->grid[1]))) return rc;
    if ((rc = pass_grid(ke_update<D, T>, &e->grid[2]))) return rc;
    if ((rc = pass_grid(ke_bc<D, T>, &e->grid[3]))) return rc;
    e->grid[4] = 0; // ke_update2r: one block per item
    // no more blocks than items (a multiple of 8 so every XCD slice has blocks)
    for (int k = 0; k < NLIST; k++) {
        const unsigned need = ((unsigned)e->nitems[k] + 7) & ~7u;
        if (need > 0 && need < e->grid[k]) e->grid[k] = need;
    }
    return 0;
}

#define EDISPATCH(e, FN, ...)                                                              \
    ((e)->D == 2 ? ((e)->prec == SMAUG_FP64 ? FN<2, double>(__VA_ARGS__)                   \
                                            : FN<2, float>(__VA_ARGS__))                   \
                 : ((e)->prec == SMAUG_FP64 ? FN<3, double>(__VA_ARGS__)                   \
                                            : FN<3, float>(__VA_ARGS__)))

extern "C" int smaug_ens_create(const smaug_cfg *cfgs, int nrun, smaug_ensemble **out)
{
    if (!cfgs || !out || nrun < 1) return fail(SMAUG_EINVAL, "bad argument");
    *out = nullptr;
    for (int r = 0; r < nrun; r++) {
        int rc = check_cfg(&cfgs[r]);
        if (rc) return rc;
        if (!same_phys(cfgs[0], cfgs[r]))
            return fail(SMAUG_EINVAL,
                        "run %d: ndim, precision, hyperdif, driver, gamma, g, chyp and drv must "
                        "be the same for every run of an ensemble", r);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SMAUG_ENODEV, "no HIP device");

    smaug_ensemble *e = new smaug_ensemble();
    e->nrun = nrun;
    e->D = cfgs[0].ndim;
    e->prec = cfgs[0].precision;
    e->tsz = e->prec == SMAUG_FP64 ? sizeof(double) : sizeof(float);
    e->cfg.assign(cfgs, cfgs + nrun);
    e->state = e->ratio = e->aux = nullptr;
    e->bg = e->axes = e->tm = e->bgu = nullptr;
    e->cm = nullptr;
    e->ctr = nullptr;
    e->descs = nullptr;
    for (int k = 0; k < NLIST; k++) e->base[k] = e->irun[k] = nullptr;
    // 2-D with hyperdiffusion: the update forms the ratios (no ratio pass)
    e->rform = e->D == 2 && cfgs[0].hyperdif ? 2 : 0;
    e->stream = nullptr;
    e->d_part = nullptr;
    e->gvalid = false;
#define ECHK(x)                                                                            \
    do {                                                                                   \
        hipError_t he_ = (x);                                                              \
        if (he_ != hipSuccess) {                                                           \
            free_ens(e);                                                                   \
            return fail(he_ == hipErrorOutOfMemory ? SMAUG_ENOMEM : SMAUG_EHIP,            \
                        "%s failed: %s", #x, hipGetErrorString(he_));                       \
        }                                                                                  \
    } while (0)
    ECHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));

    const int D = e->D, nve = 2 + 2 * D, nbg = 2 + D;
    const int npair = 2 * D + D * D + D * (D - 1);
    size_t su = 0, sr = 0, sa = 0, sb = 0, sx = 0;
    for (int r = 0; r < nrun; r++) {
        const smaug_cfg &c = cfgs[r];
        const idx_t n2 = D == 3 ? c.n[2] : 1;
        const idx_t nc = (idx_t)c.n[0] * c.n[1] * n2;
        e->nc.push_back(nc);
        e->nint.push_back((idx_t)(c.n[0] - 4) * (c.n[1] - 4) * (D == 3 ? n2 - 4 : 1));
        e->off_u.push_back(su);
        e->off_r.push_back(sr);
        e->off_aux.push_back(sa);
        e->off_bg.push_back(sb);
        su += align256(e->tsz * 4 * nve * nc);
        if (c.hyperdif) sr += align256(e->tsz * npair * nc);
        sa += align256(e->tsz * (3 + D) * nc);
        sb += align256(sizeof(double) * nbg * nc);
        sx += 4 * (size_t)(c.n[0] + c.n[1] + n2);
    }
    ECHK(hipMalloc((void **)&e->state, su));
    if (sr) ECHK(hipMalloc((void **)&e->ratio, sr));
    ECHK(hipMalloc((void **)&e->aux, sa));
    ECHK(hipMalloc((void **)&e->bg, sb));
    ECHK(hipMalloc((void **)&e->axes, sizeof(double) * sx));
    ECHK(hipMalloc((void **)&e->tm, sizeof(double) * 2 * nrun));
    ECHK(hipMalloc((void **)&e->cm, sizeof(unsigned long long) * 16 * nrun));
    ECHK(hipMalloc((void **)&e->ctr, sizeof(unsigned) * 4 * NPASS * NQ * CSTRIDE));
    ECHK(hipMalloc((void **)&e->bgu, sizeof(double) * 8 * nrun));
    ECHK(hipMemset(e->state, 0, su));
    ECHK(hipMemset(e->tm, 0, sizeof(double) * 2 * nrun));

    // per-run metric arrays, packed [dx | 1/dx | profile | lever profile] per dim
    std::vector<double> host(sx);
    size_t pos = 0;
    for (int r = 0; r < nrun; r++) {
        const smaug_cfg &c = cfgs[r];
        Geo g;
        memset(&g, 0, sizeof(g));
        for (int d = 0; d < 3; d++) {
            const int nd = d < D ? c.n[d] : 1;
            std::vector<double> dx, idx, prof, profq;
            axis_arrays(&c, d, nd, dx, idx, prof, profq);
            memcpy(&host[pos], dx.data(), sizeof(double) * nd);
            memcpy(&host[pos + nd], idx.data(), sizeof(double) * nd);
            memcpy(&host[pos + 2 * nd], prof.data(), sizeof(double) * nd);
            memcpy(&host[pos + 3 * nd], profq.data(), sizeof(double) * nd);
            g.n[d] = nd;
            g.per[d] = d < D ? (c.bc[d] == 0) : 1;
            g.dx[d] = e->axes + pos;
            g.idx[d] = e->axes + pos + nd;
            g.prof[d] = e->axes + pos + 2 * nd;
            g.profq[d] = e->axes + pos + 3 * nd;
            pos += 4 * (size_t)nd;
        }
        g.s1 = g.n[0];
        g.s2 = (idx_t)g.n[0] * g.n[1];
        g.nc = e->nc[r];
        g.bgf = g.nc;
        g.bgm = ~0u;
        e->geo.push_back(g);
    }
    ECHK(hipMemcpy(e->axes, host.data(), sizeof(double) * pos, hipMemcpyHostToDevice));

    Phys &p = e->phys;
    const smaug_cfg &c0 = cfgs[0];
    p.gamma = c0.gamma;
    for (int d = 0; d < 3; d++) p.g[d] = c0.g[d];
    for (int f = 0; f < 13; f++) p.chyp[f] = c0.chyp[f];
    for (int k = 0; k < 8; k++) p.drv[k] = c0.drv[k];
    p.hyperdif = c0.hyperdif;
    p.driver = c0.driver;
#undef ECHK
    int rc = EDISPATCH(e, setup_typed, e);
    if (rc) {
        free_ens(e);
        return rc;
    }
    *out = e;
    return 0;
}

extern "C" int smaug_ens_destroy(smaug_ensemble *e)
{
    if (e) {
        hipStreamSynchronize(e->stream);
        free_ens(e);
    }
    return 0;
}

extern "C" int smaug_ens_nrun(const smaug_ensemble *e) { return e ? e->nrun : 0; }

extern "C" long long smaug_ens_ninterior(const smaug_ensemble *e)
{
    long long s = 0;
    if (e)
        for (idx_t v : e->nint) s += v;
    return s;
}

template <int D, typename T>
static int ens_set_state(smaug_ensemble *e, int r, const double *w, const double *bg)
{
    const int nve = 2 + 2 * D, nbg = 2 + D;
    const idx_t nc = e->nc[r];
    std::vector<T> b((size_t)nve * nc);
    for (size_t i = 0; i < b.size(); i++) b[i] = (T)w[i];
    T *u0 = (T *)(e->state + e->off_u[r]);
    HIPCHK(hipMemcpyAsync(u0, b.data(), sizeof(T) * b.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync((char *)e->bg + e->off_bg[r], bg, sizeof(double) * nbg * nc,
                          hipMemcpyHostToDevice, e->stream));
    // a background that is the same in every cell is read by the per-cell
    // kernels as its nbg values (Geo::bgf = 1, bgm = 0), else in full; the
    // run's descriptor is rewritten (the captured step reads it at run time)
    {
        int uni = 1;
        for (int f = 0; f < nbg && uni; f++)
            for (idx_t c = 1; c < nc; c++)
                if (memcmp(&bg[(size_t)f * nc + c], &bg[(size_t)f * nc], sizeof(double))) {
                    uni = 0;
                    break;
                }
        RunDesc<T> rd;
        RunDesc<T> *dd = (RunDesc<T> *)e->descs + r;
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipMemcpy(&rd, dd, sizeof(rd), hipMemcpyDeviceToHost));
        if (uni) {
            double v[8];
            for (int f = 0; f < nbg; f++) v[f] = bg[(size_t)f * nc];
            HIPCHK(hipMemcpy(e->bgu + 8 * r, v, sizeof(double) * nbg, hipMemcpyHostToDevice));
            rd.bg = e->bgu + 8 * r;
            rd.geo.bgf = 1;
            rd.geo.bgm = 0u;
        } else {
            rd.bg = (const double *)((char *)e->bg + e->off_bg[r]);
            rd.geo.bgf = rd.geo.nc;
            rd.geo.bgm = ~0u;
        }
        HIPCHK(hipMemcpy(dd, &rd, sizeof(rd), hipMemcpyHostToDevice));
    }
    const Geo &g = e->geo[r];
    const idx_t tot = bc_count<D>(g.n[0], g.n[1], g.n[2]);
    hipLaunchKernelGGL((k_bc<D, T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, e->stream,
                       g, u0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

extern "C" int smaug_ens_set_state(smaug_ensemble *e, int run, const double *w, const double *bg)
{
    if (!e || !w || !bg || run < 0 || run >= e->nrun) return fail(SMAUG_EINVAL, "bad argument");
    const smaug_cfg &c = e->cfg[run];
    const int rc = check_periodic_bg(e->D, c.n, c.bc, bg);
    if (rc) return rc;
    return EDISPATCH(e, ens_set_state, e, run, w, bg);
}

extern "C" int smaug_ens_get_state(smaug_ensemble *e, int run, double *w)
{
    if (!e || !w || run < 0 || run >= e->nrun) return fail(SMAUG_EINVAL, "bad argument");
    const idx_t n = (idx_t)(2 + 2 * e->D) * e->nc[run];
    HIPCHK(hipStreamSynchronize(e->stream));
    const char *u0 = e->state + e->off_u[run];
    if (e->prec == SMAUG_FP64) {
        HIPCHK(hipMemcpy(w, u0, sizeof(double) * n, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> b(n);
        HIPCHK(hipMemcpy(b.data(), u0, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (idx_t i = 0; i < n; i++) w[i] = b[i];
    }
    return 0;
}

extern "C" int smaug_ens_sums(smaug_ensemble *e, int run, double *sums)
{
    if (!e || !sums || run < 0 || run >= e->nrun) return fail(SMAUG_EINVAL, "bad argument");
    if (!e->d_part) HIPCHK(hipMalloc((void **)&e->d_part, sizeof(double) * 2 * 8 * sums_blocks()));
    return field_sums(e->D, e->prec, e->geo[run], e->state + e->off_u[run], e->stream, e->d_part,
                      sums);
}

static Queue queue(smaug_ensemble *e, int pass, int launch)
{
    Queue q;
    q.base = e->base[pass];
    q.run = e->irun[pass];
    q.nrun = e->nrun;
    q.nitems = e->nitems[pass];
    q.ctr = e->ctr + (size_t)launch * NQ * CSTRIDE;
    return q;
}

// one RK4 step of every run (same slot algebra as the single-grid engine)
// evk (optional): 5 events per stage, recorded between its passes
template <int D, typename T>
static int ens_step(smaug_ensemble *e, const hipEvent_t *evk = nullptr)
{
    const double ct[4] = {0.0, 0.5, 0.5, 1.0};
    const int in[4] = {0, 1, 2, 1}, out[4] = {1, 2, 1, 0};
    const RunDesc<T> *rd = (const RunDesc<T> *)e->descs;
    hipLaunchKernelGGL(ke_reset, dim3(8), dim3(256), 0, e->stream, e->ctr, 4 * NPASS * NQ, e->cm,
                       16 * e->nrun);
    for (int s = 0; s < 4; s++) {
        if (evk) HIPCHK(hipEventRecord(evk[5 * s + 0], e->stream));
        hipLaunchKernelGGL((ke_prep<D, T>), dim3(e->grid[0]), dim3(256), 0, e->stream, rd,
                           queue(e, 0, 4 * s + 0), e->phys, s, in[s]);
        if (evk) HIPCHK(hipEventRecord(evk[5 * s + 1], e->stream));
        const bool r2 = D == 2 && e->rform == 2 && e->phys.hyperdif;
        if (e->phys.hyperdif && !r2)
            hipLaunchKernelGGL((ke_ratio<D, T>), dim3(e->grid[1]), dim3(256), 0, e->stream, rd,
                               queue(e, 1, 4 * s + 1), in[s]);
        if (evk) HIPCHK(hipEventRecord(evk[5 * s + 2], e->stream));
        if constexpr (D == 2) {
            if (r2)
                hipLaunchKernelGGL(ke_update2r<T>, dim3(8 * ((e->nitems[4] + 7) / 8)), dim3(U2<T>::NT),
                                   0, e->stream, rd, e->base[4], e->irun[4], e->nitems[4],
                                   e->phys, s, ct[s], in[s], out[s]);
        }
        if (!r2)
            hipLaunchKernelGGL((ke_update<D, T>), dim3(e->grid[2]), dim3(256), 0, e->stream, rd,
                               queue(e, 2, 4 * s + 2), e->phys, s, ct[s], in[s], out[s]);
        if (evk) HIPCHK(hipEventRecord(evk[5 * s + 3], e->stream));
        hipLaunchKernelGGL((ke_bc<D, T>), dim3(e->grid[3]), dim3(256), 0, e->stream, rd,
                           queue(e, 3, 4 * s + 3), out[s]);
        if (evk) HIPCHK(hipEventRecord(evk[5 * s + 4], e->stream));
    }
    hipLaunchKernelGGL(ke_advance_time, dim3((e->nrun + 63) / 64), dim3(64), 0, e->stream, e->tm,
                       e->nrun);
    HIPCHK(hipGetLastError());
    return 0;
}

template <int D, typename T>
static int ens_capture(smaug_ensemble *e)
{
    if (e->gvalid) return 0;
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = ens_step<D, T>(e);
    hipError_t ce = hipStreamEndCapture(e->stream, &graph);
    if (rc) return rc;
    if (ce != hipSuccess) return fail(SMAUG_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(SMAUG_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    e->gvalid = true;
    return 0;
}

static int ens_set_time(smaug_ensemble *e, const double *t0, const double *dt)
{
    std::vector<double> tm(2 * (size_t)e->nrun);
    for (int r = 0; r < e->nrun; r++) {
        tm[2 * r] = t0 ? t0[r] : 0.0;
        tm[2 * r + 1] = dt[r];
    }
    HIPCHK(hipMemcpyAsync(e->tm, tm.data(), sizeof(double) * tm.size(), hipMemcpyHostToDevice,
                          e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

template <int D, typename T>
static int ens_run(smaug_ensemble *e, const double *t0, const double *dt, int nsteps, double *ms)
{
    int rc = ens_capture<D, T>(e);
    if (rc) return rc;
    if ((rc = ens_set_time(e, t0, dt))) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    if (ms) {
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        HIPCHK(hipEventRecord(a, e->stream));
    }
    for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(e->gexec, e->stream));
    if (ms) {
        HIPCHK(hipEventRecord(b, e->stream));
        HIPCHK(hipEventSynchronize(b));
        float f = 0;
        HIPCHK(hipEventElapsedTime(&f, a, b));
        *ms = f;
        hipEventDestroy(a);
        hipEventDestroy(b);
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

extern "C" int smaug_ens_rk4(smaug_ensemble *e, const double *t0, const double *dt, int nsteps)
{
    if (!e || !dt || nsteps < 0) return fail(SMAUG_EINVAL, "bad argument");
    return EDISPATCH(e, ens_run, e, t0, dt, nsteps, (double *)nullptr);
}

template <int D, typename T>
static int ens_time_passes(smaug_ensemble *e, const double *t0, const double *dt, double *ms)
{
    int rc = ens_set_time(e, t0, dt);
    if (rc) return rc;
    std::vector<hipEvent_t> evk(20);
    for (auto &v : evk) HIPCHK(hipEventCreate(&v));
    if ((rc = ens_step<D, T>(e, evk.data()))) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int p = 0; p < 4; p++) {
        double sum = 0;
        for (int s = 0; s < 4; s++) {
            float m = 0;
            HIPCHK(hipEventElapsedTime(&m, evk[5 * s + p], evk[5 * s + p + 1]));
            sum += m;
        }
        ms[p] = sum / 4;
    }
    for (auto &v : evk) hipEventDestroy(v);
    return 0;
}

extern "C" int smaug_ens_time_passes(smaug_ensemble *e, const double *t0, const double *dt,
                                     double ms[4])
{
    if (!e || !dt || !ms) return fail(SMAUG_EINVAL, "bad argument");
    return EDISPATCH(e, ens_time_passes, e, t0, dt, ms);
}

extern "C" int smaug_ens_time_rk4(smaug_ensemble *e, const double *t0, const double *dt,
                                  int nsteps, double *total_ms)
{
    if (!e || !dt || nsteps < 1 || !total_ms) return fail(SMAUG_EINVAL, "bad argument");
    return EDISPATCH(e, ens_run, e, t0, dt, nsteps, total_ms);
}

extern "C" int smaug_ens_set_ratio_form(smaug_ensemble *e, int form)
{
    if (!e || (form != 0 && form != 2)) return fail(SMAUG_EINVAL, "bad argument");
    if (form == 2 && e->D != 2)
        return fail(SMAUG_EINVAL, "the ratios formed by the update are for 2-D runs");
    e->rform = form;
    if (e->gvalid) {
        hipGraphExecDestroy(e->gexec);
        e->gvalid = false;
    }
    return 0;
}

extern "C" int smaug_ens_get_ratio_form(smaug_ensemble *e, int *form)
{
    if (!e || !form) return fail(SMAUG_EINVAL, "null argument");
    *form = e->D == 2 && e->phys.hyperdif ? e->rform : 0;
    return 0;
}
