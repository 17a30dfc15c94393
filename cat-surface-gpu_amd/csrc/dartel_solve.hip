// Host-side orchestration of the DARTEL hot path on one HIP stream:
// the full-multigrid solve, one Gauss-Newton step (dartel_step) and the
// multi-step solve (solve_dartel_maps), all over a caller-provided
// workspace so that nothing here allocates and every launch can be captured
// into a hipGraph.
//
// Reference semantics: src/cat_surface_gpu/dartel_grid.py:1211-1616 and
// src/cat_surface_gpu/surface_warp.py:76-107, :239-367.
#include "dartel_solve.h"

#include <atomic>
#include <vector>

namespace csg {

// ---------------------------------------------------------------- workspace

namespace {

struct Carver {
    char *base;
    size_t used = 0;
    size_t cap;
    bool dry;
    template <typename U>
    U *take(size_t count)
    {
        size_t bytes = (count * sizeof(U) + 255) & ~(size_t)255;
        U *p = dry ? nullptr : reinterpret_cast<U *>(base + used);
        used += bytes;
        return p;
    }
};

template <typename T>
__global__ void k_copy(const T *__restrict__ in, T *__restrict__ out, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = in[i];
}

template <typename T>
__global__ void k_fill(T *__restrict__ out, T value, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = value;
}

template <typename T>
int copy_async(const T *in, T *out, size_t n, hipStream_t s)
{
    if (n == 0)
        return kOk;
    hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out, n);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int fill_async(T *out, T value, size_t n, hipStream_t s)
{
    if (n == 0)
        return kOk;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, out, value, n);
    CSG_LAUNCH_CHECK();
    return kOk;
}

}  // namespace

int fmg_levels(int nx, int ny, int *lnx, int *lny)
{
    int L = 1;
    lnx[0] = nx;
    lny[0] = ny;
    while (lnx[L - 1] >= 2 || lny[L - 1] >= 2) {
        lnx[L] = (lnx[L - 1] + 1) / 2;
        lny[L] = (lny[L - 1] + 1) / 2;
        ++L;
        if (lnx[L - 1] < 2 && lny[L - 1] < 2)
            break;
        if (L >= kMaxLevels)
            break;
    }
    return L;
}

// ---------------------------------------------------------------- FMG

template <typename T>
struct FmgPlan {
    int L;
    int nx[kMaxLevels], ny[kMaxLevels];
    T *A[kMaxLevels], *B0[kMaxLevels], *B[kMaxLevels], *U[kMaxLevels], *R[kMaxLevels];
    T *U2[kMaxLevels];  // second correction buffer of the fused schedule (aliases R)
    unsigned char *meta;  // the fused schedule's coarse state (coarse_meta_bytes)
    size_t P(int l) const { return (size_t)nx[l] * ny[l]; }
};
static_assert(kMaxLevels == kFmgMaxLevels, "one level limit");

template <typename T>
static void carve_fmg(Carver &c, int nx, int ny, int batch, FmgPlan<T> &plan)
{
    plan.L = fmg_levels(nx, ny, plan.nx, plan.ny);
    for (int l = 0; l < plan.L; ++l) {
        size_t P = plan.P(l);
        plan.R[l] = c.template take<T>(2 * P * batch);
        plan.U2[l] = plan.R[l];  // the fused schedule needs no residual buffer
        if (l == 0) {
            plan.A[0] = plan.B0[0] = plan.B[0] = plan.U[0] = nullptr;
            continue;
        }
        plan.A[l] = c.template take<T>(3 * P * batch);
        plan.B0[l] = c.template take<T>(2 * P * batch);
        plan.B[l] = c.template take<T>(2 * P * batch);
        plan.U[l] = c.template take<T>(2 * P * batch);
    }
    plan.meta = c.template take<unsigned char>(coarse_meta_bytes<T>(plan.nx, plan.ny, plan.L));
}

template <typename T>
size_t fmg_workspace_bytes(int nx, int ny, int batch)
{
    Carver c{nullptr, 0, 0, true};
    FmgPlan<T> plan;
    carve_fmg<T>(c, nx, ny, batch, plan);
    return c.used;
}

namespace {

template <typename T>
struct FmgRun {
    FmgPlan<T> &p;
    const double *params;
    int nit, batch;
    hipStream_t s;

    MembraneWeights weights(int l) const
    {
        double lp[5] = {params[0] * p.nx[l] / p.nx[0], params[1] * p.ny[l] / p.ny[0], params[2],
                        params[3], params[4]};
        return membrane_weights(lp);
    }
    int relax(int l)
    {
        size_t P = p.P(l);
        return launch_relax<T>(p.A[l], p.B[l], p.U[l], weights(l), p.nx[l], p.ny[l], batch,
                               3 * P, 2 * P, nit, s);
    }
    // B[l+1] = restrict(B[l] - (L + A) U[l]); U[l+1] = 0.  dartel_grid.py:1356-1391
    int down(int l)
    {
        size_t P = p.P(l), Pc = p.P(l + 1);
        CSG_TRY(launch_apply_membrane<T>(p.A[l], p.U[l], weights(l), p.nx[l], p.ny[l], batch,
                                         3 * P, 2 * P, p.R[l], p.B[l], s));
        CSG_TRY(launch_resize<T>(p.R[l], 2, p.nx[l], p.ny[l], p.nx[l + 1], p.ny[l + 1], batch,
                                 2 * P, 2 * Pc, p.B[l + 1], 0, s));
        return fill_async<T>(p.U[l + 1], (T)0, 2 * Pc * batch, s);
    }
    // U[l] += prolong(U[l+1]).  dartel_grid.py:1392-1400
    int up(int l)
    {
        size_t P = p.P(l), Pc = p.P(l + 1);
        return launch_resize<T>(p.U[l + 1], 2, p.nx[l + 1], p.ny[l + 1], p.nx[l], p.ny[l], batch,
                                2 * Pc, 2 * P, p.U[l], 1, s);
    }
    int cycle(int start)
    {
        int last = p.L - 1;
        for (int l = start; l < last; ++l) {
            CSG_TRY(relax(l));
            CSG_TRY(down(l));
        }
        CSG_TRY(relax(last));
        for (int l = last - 1; l >= start; --l) {
            CSG_TRY(up(l));
            CSG_TRY(relax(l));
        }
        return kOk;
    }
};

// The fused schedule: fine levels on LDS tiles, levels c0.. in one
// workgroup per job (fmg_fused.hip).  Same operations, same order.
template <typename T>
struct FusedFmg {
    FmgPlan<T> &p;
    const double *params;
    int nit, batch, c0;
    hipStream_t s;
    int cur[kMaxLevels] = {};
    CoarseLevels<T> lv;
    size_t lds = 0;
    bool built = false;  // the coarse state of this solve is stored

    T *ubuf(int l, int which) { return which ? p.U2[l] : p.U[l]; }
    T *u(int l) { return ubuf(l, cur[l]); }
    MembraneWeights weights(int l) const
    {
        double lp[5] = {params[0] * p.nx[l] / p.nx[0], params[1] * p.ny[l] / p.ny[0], params[2],
                        params[3], params[4]};
        return membrane_weights(lp);
    }
    void setup()
    {
        lv.L = p.L;
        lv.c0 = c0;
        for (int l = 0; l < p.L; ++l) {
            MembraneWeights w = weights(l);
            lv.nx[l] = p.nx[l];
            lv.ny[l] = p.ny[l];
            lv.w00[l] = (T)w.w00;
            lv.w01[l] = (T)w.w01;
            lv.w10[l] = (T)w.w10;
            // levels below c0 are made by the coarse kernel, which keeps A there
            lv.apyr[l] = l > c0 ? p.A[l] : nullptr;
        }
        lv.meta = p.meta;
        lds = coarse_lds_bytes<T>(p.nx, p.ny, c0, p.L);
    }
    // relax_membrane at level l, optionally after u[l] += prolong(u[l+1])
    int relax(int l, bool prolong)
    {
        size_t P = p.P(l), Pc = prolong ? p.P(l + 1) : 0;
        T *in = u(l), *out = ubuf(l, cur[l] ^ 1);
        CSG_TRY(launch_relax_tile<T>(p.A[l], p.B[l], in, out, prolong ? u(l + 1) : nullptr,
                                     prolong ? p.nx[l + 1] : 0, prolong ? p.ny[l + 1] : 0,
                                     weights(l), p.nx[l], p.ny[l], batch, 3 * P, 2 * P, 2 * Pc,
                                     2 * nit, s));
        cur[l] ^= 1;
        return kOk;
    }
    int down(int l)
    {
        size_t P = p.P(l), Pc = p.P(l + 1);
        cur[l + 1] = 0;
        return launch_restrict_residual<T>(p.A[l], p.B[l], u(l), weights(l), p.nx[l], p.ny[l],
                                           p.nx[l + 1], p.ny[l + 1], batch, 3 * P, 2 * P, 2 * Pc,
                                           p.B[l + 1], u(l + 1), s);
    }
    int coarse(int init, int ncycles)
    {
        size_t P = p.P(c0);
        const int build = built ? 0 : 1;
        built = true;
        return launch_fmg_coarse<T>(lv, p.A[c0], init ? p.B0[c0] : p.B[c0], u(c0), 3 * P, 2 * P,
                                    params[4], nit, ncycles, init, build, batch, lds, s);
    }
    int cycle(int start)
    {
        for (int l = start; l < c0; ++l) {
            CSG_TRY(relax(l, false));
            CSG_TRY(down(l));
        }
        CSG_TRY(coarse(0, 1));
        for (int l = c0 - 1; l >= start; --l)
            CSG_TRY(relax(l, true));
        return kOk;
    }
};

std::atomic<int> g_fused_schedule{1};

// First level whose whole sub-pyramid fits in LDS, or -1 when the fused
// schedule does not apply (a fine level of odd width, or tiles too large).
template <typename T>
int fused_first_coarse_level(const FmgPlan<T> &p, int nit)
{
    if (!g_fused_schedule.load(std::memory_order_relaxed))
        return -1;
    const size_t budget = fmg_lds_budget();
    int c0 = -1;
    for (int l = 0; l < p.L; ++l)
        if (coarse_lds_bytes<T>(p.nx, p.ny, l, p.L) <= budget) {
            c0 = l;
            break;
        }
    if (c0 < 0)
        return -1;
    for (int l = 0; l < c0; ++l) {
        if ((p.nx[l] & 1) || p.nx[l] < 2)
            return -1;
        const TileShape tile = relax_tile_shape(p.nx[l], p.ny[l], (int)sizeof(T));
        if (tile_lds_bytes<T>(2 * nit, tile) > budget || !tile_fits_threads(2 * nit, tile))
            return -1;
    }
    return c0;
}

}  // namespace

int fused_tiled_levels(int nx, int ny, int nit, int elem)
{
    FmgPlan<float> pf{};
    FmgPlan<double> pd{};
    int L = fmg_levels(nx, ny, pf.nx, pf.ny);
    pf.L = L;
    for (int l = 0; l < L; ++l) {
        pd.nx[l] = pf.nx[l];
        pd.ny[l] = pf.ny[l];
    }
    pd.L = L;
    const int c0 = elem == 8 ? fused_first_coarse_level<double>(pd, nit)
                             : fused_first_coarse_level<float>(pf, nit);
    return c0 < 0 ? 0 : c0;
}

int set_fused_fmg_schedule(int enable)
{
    return g_fused_schedule.exchange(enable ? 1 : 0);
}

template <typename T>
int fmg_run(const T *hes, const T *rhs, const double *params, int cycles, int nit,
            const T *initial, int coarse_init, int nx, int ny, int batch, T *out, void *ws,
            size_t ws_bytes, hipStream_t s)
{
    if (nx < 1 || ny < 1 || batch < 1 || cycles < 0 || nit < 0)
        return fail(kInvalidArgument, "invalid FMG arguments");
    Carver c{(char *)ws, 0, ws_bytes, false};
    FmgPlan<T> plan;
    carve_fmg<T>(c, nx, ny, batch, plan);
    if (c.used > ws_bytes)
        return fail(kInvalidArgument, "FMG workspace too small: need %zu bytes, got %zu", c.used,
                    ws_bytes);
    size_t P0 = plan.P(0);
    plan.A[0] = const_cast<T *>(hes);
    plan.B0[0] = plan.B[0] = const_cast<T *>(rhs);
    plan.U[0] = out;
    const int c0 = fused_first_coarse_level<T>(plan, nit);
    // level pyramids of A and the original right-hand side.  dartel_grid.py:1247-1272
    // (the fused schedule's coarse kernel builds the levels below c0 itself from
    // level c0, so only the fine levels and c0 are made here)
    const int top = c0 >= 0 ? c0 : plan.L - 1;
    for (int l = 1; l <= top; ++l) {
        size_t P = plan.P(l - 1), Pc = plan.P(l);
        CSG_TRY(launch_resize_two<T>(plan.A[l - 1], 3, 3 * P, 3 * Pc, plan.A[l], plan.B0[l - 1],
                                     2, 2 * P, 2 * Pc, plan.B0[l], plan.nx[l - 1],
                                     plan.ny[l - 1], plan.nx[l], plan.ny[l], batch, s));
    }
    if (initial) {
        if (initial != out)
            CSG_TRY(copy_async<T>(initial, out, 2 * P0 * batch, s));
    } else {
        CSG_TRY(fill_async<T>(out, (T)0, 2 * P0 * batch, s));
    }
    if (c0 >= 0) {
        FusedFmg<T> run{plan, params, nit, batch, c0, s};
        run.setup();
        if (coarse_init) {
            CSG_TRY(run.coarse(1, cycles));
            for (int level = c0 - 1; level >= 0; --level) {
                size_t P = plan.P(level), Pc = plan.P(level + 1);
                CSG_TRY(launch_resize<T>(run.u(level + 1), 2, plan.nx[level + 1],
                                         plan.ny[level + 1], plan.nx[level], plan.ny[level],
                                         batch, 2 * Pc, 2 * P, run.u(level), 0, s));
                if (level > 0)
                    CSG_TRY(copy_async<T>(plan.B0[level], plan.B[level], 2 * P * batch, s));
                for (int cyc = 0; cyc < cycles; ++cyc)
                    CSG_TRY(run.cycle(level));
            }
        } else if (c0 == 0) {
            CSG_TRY(run.coarse(0, cycles));
        } else {
            for (int cyc = 0; cyc < cycles; ++cyc)
                CSG_TRY(run.cycle(0));
        }
        if (run.cur[0] != 0)
            CSG_TRY(copy_async<T>(plan.U2[0], out, 2 * P0 * batch, s));
        return kOk;
    }
    FmgRun<T> run{plan, params, nit, batch, s};
    int last = plan.L - 1;
    if (coarse_init) {
        // dartel_grid.py:1301-1326, :1402-1413
        size_t Pl = plan.P(last);
        CSG_TRY(fill_async<T>(plan.B[last], (T)0, 2 * Pl * batch, s));
        CSG_TRY(launch_coarse_solve<T>(plan.A[last], plan.B0[last], params[4], (int)Pl, batch,
                                       3 * Pl, 2 * Pl, plan.U[last], s));
        CSG_TRY(run.relax(last));
        for (int level = last - 1; level >= 0; --level) {
            size_t P = plan.P(level), Pc = plan.P(level + 1);
            CSG_TRY(launch_resize<T>(plan.U[level + 1], 2, plan.nx[level + 1], plan.ny[level + 1],
                                     plan.nx[level], plan.ny[level], batch, 2 * Pc, 2 * P,
                                     plan.U[level], 0, s));
            if (level > 0)
                CSG_TRY(copy_async<T>(plan.B0[level], plan.B[level], 2 * P * batch, s));
            for (int cyc = 0; cyc < cycles; ++cyc)
                CSG_TRY(run.cycle(level));
        }
    } else {
        for (int cyc = 0; cyc < cycles; ++cyc)
            CSG_TRY(run.cycle(0));
    }
    return kOk;
}

// ---------------------------------------------------------------- dartel_step

template <typename T>
struct StepPlan {
    size_t P;
    T *em[2], *ej[2];          // exponentiation ping-pong
    T *fm[2], *fj[2];          // forward small-deformation chain
    T *gm[2], *gj[2];          // reverse small-deformation chain
    T *fb[2], *fa[2];          // forward b / A ping-pong
    T *rb[2], *ra[2];          // reverse b / A ping-pong
    T *crhs, *chess, *u;       // combined system and its solution
    double *partial[3];        // objective / regulariser / norm block partials
    int *flags;                // per-element "initial guess is non-zero"
    int nblocks;
    void *fmg_ws;
    size_t fmg_bytes;
};

template <typename T>
static void carve_step(Carver &c, int nx, int ny, int batch, StepPlan<T> &plan)
{
    size_t P = (size_t)nx * ny;
    plan.P = P;
    for (int i = 0; i < 2; ++i) {
        plan.em[i] = c.template take<T>(2 * P * batch);
        plan.ej[i] = c.template take<T>(4 * P * batch);
        plan.fm[i] = c.template take<T>(2 * P * batch);
        plan.fj[i] = c.template take<T>(4 * P * batch);
        plan.gm[i] = c.template take<T>(2 * P * batch);
        plan.gj[i] = c.template take<T>(4 * P * batch);
        plan.fb[i] = c.template take<T>(2 * P * batch);
        plan.fa[i] = c.template take<T>(3 * P * batch);
        plan.rb[i] = c.template take<T>(2 * P * batch);
        plan.ra[i] = c.template take<T>(3 * P * batch);
    }
    plan.crhs = c.template take<T>(2 * P * batch);
    plan.chess = c.template take<T>(3 * P * batch);
    plan.u = c.template take<T>(2 * P * batch);
    plan.nblocks = objective_blocks(nx, ny);
    for (int i = 0; i < 3; ++i)
        plan.partial[i] = c.template take<double>((size_t)plan.nblocks * batch);
    plan.flags = c.template take<int>(batch);
    size_t before = c.used;
    plan.fmg_ws = c.template take<char>(fmg_workspace_bytes<T>(nx, ny, batch));
    plan.fmg_bytes = c.used - before;
}

template <typename T>
size_t step_workspace_bytes(int nx, int ny, int batch)
{
    Carver c{nullptr, 0, 0, true};
    StepPlan<T> plan;
    carve_step<T>(c, nx, ny, batch, plan);
    return c.used;
}

namespace {

// k squaring updates starting from (b0, a0) at map slot 0.  Returns pointers
// to the final b / A, the final map / Jacobian slot, and the scratch buffer
// (the reference's map_buffers[0]).  dartel_grid.py:778-914
template <typename T>
struct ChainResult {
    T *b, *a, *map, *jac, *scratch;
};

template <typename T>
int squaring_chain(T *b0, T *a0, T **maps, T **jacs, T **bbuf, T **abuf, int k, bool save,
                   int nx, int ny, int batch, ChainResult<T> &res, hipStream_t s, bool disp)
{
    size_t P = (size_t)nx * ny;
    int cur = 0, oth = 1;
    T *bc = b0, *ac = a0;
    int pp = 0;
    for (int st = 0; st < k; ++st) {
        bool compose = save || st < k - 1;
        T *bn = bbuf[pp], *an = abuf[pp];
        CSG_TRY(launch_squaring<T>(bc, ac, maps[cur], jacs[cur], nx, ny, batch, 2 * P, 3 * P,
                                   2 * P, 4 * P, bn, an, compose ? nullptr : maps[oth],
                                   compose ? maps[oth] : nullptr, compose ? jacs[oth] : nullptr,
                                   s, disp));
        bc = bn;
        ac = an;
        pp ^= 1;
        if (compose) {
            int t = cur;
            cur = oth;
            oth = t;
        }
    }
    if (save && cur != 0) {
        CSG_TRY(copy_async<T>(maps[cur], maps[0], 2 * P * batch, s));
        CSG_TRY(copy_async<T>(jacs[cur], jacs[0], 4 * P * batch, s));
    }
    res.b = bc;
    res.a = ac;
    res.map = maps[cur];
    res.jac = jacs[cur];
    res.scratch = maps[0];
    return kOk;
}

}  // namespace

template <typename T>
int dartel_step_run(const StepArgs<T> &a, void *ws, size_t ws_bytes, hipStream_t s)
{
    const int nx = a.nx, ny = a.ny, batch = a.batch, k = a.k;
    if (nx < 1 || ny < 1 || batch < 1 || k < 0 || k > 30)
        return fail(kInvalidArgument, "invalid dartel_step arguments");
    if (a.code != 0 && a.code != 1)
        return fail(kUnsupported, "dartel_step supports only code 0 and 1");
    Carver c{(char *)ws, 0, ws_bytes, false};
    StepPlan<T> pl;
    carve_step<T>(c, nx, ny, batch, pl);
    if (c.used > ws_bytes)
        return fail(kInvalidArgument, "dartel_step workspace too small: need %zu, got %zu",
                    c.used, ws_bytes);
    const size_t P = pl.P;
    const double scale = 1.0 / (double)(1LL << k);
    // float32 keeps its internal maps as displacements (see csg_common.h)
    const bool disp = sizeof(T) == 4;

    // expdef(v, k) with Jacobian.  dartel_grid.py:1493-1501
    CSG_TRY(launch_expdef_init<T>(a.v, scale, nx, ny, batch, 2 * P, pl.em[0], pl.ej[0], 2 * P,
                                  4 * P, s, disp));
    int e = 0;
    for (int i = 0; i < k; ++i) {
        CSG_TRY(launch_compose_self<T>(pl.em[e], pl.ej[e], nx, ny, batch, 2 * P, 4 * P,
                                       pl.em[e ^ 1], pl.ej[e ^ 1], s, disp));
        e ^= 1;
    }
    // jac_div_smalldef + initialise_objfun + smalldef_jac(v, -scale).  :1502-1523
    ObjectiveArgs<T> o{};
    o.src = a.src;
    o.tgt = a.tgt;
    o.v = a.v;
    o.map = pl.em[e];
    o.jac = pl.ej[e];
    o.dist = a.dist;
    o.s_div = scale;
    o.s_small = -scale;
    o.nx = nx;
    o.ny = ny;
    o.batch = batch;
    o.istride = P;
    o.vstride = 2 * P;
    o.mstride = 2 * P;
    o.jstride = 4 * P;
    o.dstride = a.dist_batch_stride;
    o.rhs = pl.rb[0];  // forward objective output lives in the reverse ping-pong until the
    o.hes = pl.ra[0];  // forward chain consumes it
    o.smap = pl.fm[0];
    o.sjac = pl.fj[0];
    o.rstride = 2 * P;
    o.hstride = 3 * P;
    o.partial = pl.partial[0];
    o.disp = disp;
    CSG_TRY(launch_objective<T>(o, s));
    CSG_TRY(launch_reduce_partials(pl.partial[0], pl.nblocks, batch, 0.5, a.metrics,
                                   a.metrics_stride, 0, s));
    // forward squaring.  :1524-1542
    ChainResult<T> fwd;
    T *fm[2] = {pl.fm[0], pl.fm[1]}, *fj[2] = {pl.fj[0], pl.fj[1]};
    CSG_TRY(squaring_chain<T>(pl.rb[0], pl.ra[0], fm, fj, pl.fb, pl.fa, k, a.code == 1, nx, ny,
                              batch, fwd, s, disp));
    T *bf = fwd.b, *af = fwd.a;
    if (k == 0) {
        // no squaring step ran: the forward result is the objective output itself;
        // move it out of the reverse buffers before they are reused
        CSG_TRY(copy_async<T>(pl.rb[0], pl.fb[0], 2 * P * batch, s));
        CSG_TRY(copy_async<T>(pl.ra[0], pl.fa[0], 3 * P * batch, s));
        bf = pl.fb[0];
        af = pl.fa[0];
    }
    T *scratch = fwd.scratch;
    // the scratch holds a map unless the chain ended on an increment (k even >= 2, no save)
    bool scratch_is_map = (a.code == 1) || k == 0 || (k % 2 == 1);
    T *br = nullptr, *ar = nullptr;
    if (a.code == 1) {
        // reverse objective at the saved forward map.  :1544-1570
        ObjectiveArgs<T> r = o;
        r.src = a.tgt;
        r.tgt = a.src;
        r.map = fwd.map;
        r.jac = fwd.jac;
        r.s_div = -scale;
        r.s_small = scale;
        r.rhs = pl.rb[0];
        r.hes = pl.ra[0];
        r.smap = pl.gm[0];
        r.sjac = pl.gj[0];
        CSG_TRY(launch_objective<T>(r, s));
        CSG_TRY(launch_reduce_partials(pl.partial[0], pl.nblocks, batch, 0.5, a.metrics,
                                       a.metrics_stride, 1, s));
        ChainResult<T> rev;
        T *gm[2] = {pl.gm[0], pl.gm[1]}, *gj[2] = {pl.gj[0], pl.gj[1]};
        T *rbb[2] = {pl.rb[1], pl.crhs}, *rab[2] = {pl.ra[1], pl.chess};
        CSG_TRY(squaring_chain<T>(pl.rb[0], pl.ra[0], gm, gj, rbb, rab, k, false, nx, ny, batch,
                                  rev, s, disp));
        br = rev.b;
        ar = rev.a;
        scratch = rev.scratch;
        scratch_is_map = k == 0 || (k % 2 == 1);
        if (br == pl.crhs) {
            // the combine kernel writes crhs/chess: keep the reverse result elsewhere
            CSG_TRY(copy_async<T>(pl.crhs, pl.rb[1], 2 * P * batch, s));
            CSG_TRY(copy_async<T>(pl.chess, pl.ra[1], 3 * P * batch, s));
            br = pl.rb[1];
            ar = pl.ra[1];
        }
    }
    // rhs = (bf - br)·s + Lv, H = (Af + Ar)·s, metrics.  :1584-1595
    double rp[5] = {a.params[0], a.params[1], a.params[2], a.params[3], a.params[4]};
    MembraneWeights w = membrane_weights(rp);
    int nb = 0;
    CSG_TRY(launch_combine<T>(a.v, bf, af, br, ar, scale, w.w00, w.w01, w.w10, nx, ny, batch,
                              2 * P, 2 * P, 3 * P, pl.crhs, pl.chess, pl.partial[1],
                              pl.partial[2], &nb, s));
    CSG_TRY(launch_reduce_partials(pl.partial[1], nb, batch, 0.5, a.metrics + 1,
                                   a.metrics_stride, 0, s));
    CSG_TRY(launch_reduce_partials(pl.partial[2], nb, batch, 1.0, a.metrics + 2,
                                   a.metrics_stride, 0, s));
    CSG_TRY(launch_sqrt_inplace(a.metrics + 2, batch, a.metrics_stride, s));
    // FMG from the scratch guess.  :1596-1609, :1392-1415
    double sp[5] = {rp[0], rp[1], rp[2], rp[3], rp[4] + a.lmreg};
    if (disp && scratch_is_map) {
        // the reference starts the solver from the map's coordinates
        CSG_TRY(launch_disp_to_coord<T>(scratch, nx, ny, batch, 2 * P, pl.u, s));
        scratch = pl.u;
    }
    if (a.assume_nonzero) {
        CSG_TRY(fmg_run<T>(pl.chess, pl.crhs, sp, a.cycles, a.nit, scratch, 0, nx, ny, batch,
                           pl.u, pl.fmg_ws, pl.fmg_bytes, s));
    } else {
        // the reference checks torch.any(initial) on the host; so do we
        CSG_TRY(launch_any_nonzero<T>(scratch, 2 * P, batch, 2 * P, pl.flags, s));
        std::vector<int> flags(batch);
        CSG_HIP_CHECK(hipMemcpyAsync(flags.data(), pl.flags, sizeof(int) * batch,
                                     hipMemcpyDeviceToHost, s));
        CSG_HIP_CHECK(hipStreamSynchronize(s));
        bool all_same = true;
        for (int i = 1; i < batch; ++i)
            all_same = all_same && (flags[i] == flags[0]);
        if (all_same) {
            CSG_TRY(fmg_run<T>(pl.chess, pl.crhs, sp, a.cycles, a.nit, scratch, !flags[0], nx,
                               ny, batch, pl.u, pl.fmg_ws, pl.fmg_bytes, s));
        } else {
            for (int i = 0; i < batch; ++i)
                CSG_TRY(fmg_run<T>(pl.chess + 3 * P * i, pl.crhs + 2 * P * i, sp, a.cycles,
                                   a.nit, scratch + 2 * P * i, !flags[i], nx, ny, 1,
                                   pl.u + 2 * P * i, pl.fmg_ws, pl.fmg_bytes, s));
        }
    }
    // v' = v - u·w.  :1610-1614
    CSG_TRY(launch_update_velocity<T>(a.v, pl.u, a.dist, nx, ny, batch, 2 * P,
                                      a.dist_batch_stride, a.v_out, s));
    return kOk;
}

// ---------------------------------------------------------------- solve

template <typename T>
size_t solve_workspace_bytes(int nx, int ny, int batch)
{
    Carver c{nullptr, 0, 0, true};
    StepPlan<T> plan;
    carve_step<T>(c, nx, ny, batch, plan);
    size_t P = (size_t)nx * ny;
    c.template take<T>(2 * P * batch);  // velocity A
    c.template take<T>(2 * P * batch);  // velocity B
    c.template take<double>(3 * (size_t)batch);  // metrics when the caller keeps none
    return c.used;
}

template <typename T>
int solve_run(const SolveArgs<T> &a, void *ws, size_t ws_bytes, hipStream_t s)
{
    const int nx = a.nx, ny = a.ny, batch = a.batch;
    if (a.loop < 1 || a.its < 0 || a.steps_per_job < 1 || batch % a.steps_per_job != 0)
        return fail(kInvalidArgument, "invalid solve arguments");
    size_t need = solve_workspace_bytes<T>(nx, ny, batch);
    if (need > ws_bytes)
        return fail(kInvalidArgument, "solve workspace too small: need %zu, got %zu", need,
                    ws_bytes);
    size_t step_bytes = step_workspace_bytes<T>(nx, ny, batch);
    const size_t P = (size_t)nx * ny;
    const bool disp = sizeof(T) == 4;
    Carver c{(char *)ws, step_bytes, ws_bytes, false};
    T *v[2] = {c.template take<T>(2 * P * batch), c.template take<T>(2 * P * batch)};
    double *mscratch = c.template take<double>(3 * (size_t)batch);
    CSG_TRY(fill_async<T>(v[0], (T)0, 2 * P * batch, s));
    // loop parameters.  surface_warp.py:76-107
    std::vector<double> params(5 * a.loop);
    double mu = 0.125, lam = 0.0;
    for (int i = 0; i < a.loop; ++i) {
        params[5 * i + 0] = 1.0;
        params[5 * i + 1] = 1.0;
        params[5 * i + 2] = mu;
        params[5 * i + 3] = lam;
        params[5 * i + 4] = lam / 2.0;
        if ((i + 1) % 4 == 0)
            mu /= 1.25;
        lam /= 5.0;
    }
    int cur = 0, m = 0;
    const int per_elem = a.loop * a.its * 3;
    for (int l = 0; l < a.loop; ++l) {
        for (int it = 0; it < a.its; ++it) {
            StepArgs<T> st{};
            st.src = a.src;
            st.tgt = a.tgt;
            st.v = v[cur];
            st.v_out = v[cur ^ 1];
            st.dist = nullptr;
            st.dist_batch_stride = 0;
            st.k = l;
            for (int q = 0; q < 5; ++q)
                st.params[q] = params[5 * l + q];
            st.lmreg = a.lmreg;
            st.cycles = a.cycles;
            st.nit = a.nit;
            st.code = a.code;
            st.assume_nonzero = a.assume_nonzero;
            st.nx = nx;
            st.ny = ny;
            st.batch = batch;
            st.metrics = a.metrics ? a.metrics + 3 * m : mscratch;
            st.metrics_stride = a.metrics ? per_elem : 3;
            CSG_TRY(dartel_step_run<T>(st, ws, step_bytes, s));
            cur ^= 1;
            ++m;
        }
    }
    if (a.velocity)
        CSG_TRY(copy_async<T>(v[cur], a.velocity, 2 * P * batch, s));
    // flow = expdef(v_last, 10) for the last step of every job.  surface_warp.py:354-361
    if (a.flow) {
        StepPlan<T> pl;
        Carver c2{(char *)ws, 0, ws_bytes, false};
        carve_step<T>(c2, nx, ny, batch, pl);
        int jobs = batch / a.steps_per_job;
        for (int j = 0; j < jobs; ++j) {
            const T *vl = v[cur] + 2 * P * ((size_t)j * a.steps_per_job + a.steps_per_job - 1);
            T *fm[2] = {pl.em[0] + 2 * P * j, pl.em[1] + 2 * P * j};
            CSG_TRY(launch_expdef_init<T>(vl, 1.0 / 1024.0, nx, ny, 1, 0, fm[0], pl.ej[0], 0, 0,
                                          s, disp));
        }
        int e = 0;
        for (int i = 0; i < 10; ++i) {
            CSG_TRY(launch_compose_self<T>(pl.em[e], nullptr, nx, ny, jobs, 2 * P, 0,
                                           pl.em[e ^ 1], nullptr, s, disp));
            e ^= 1;
        }
        if (disp)
            CSG_TRY(launch_disp_to_coord<T>(pl.em[e], nx, ny, jobs, 2 * P, a.flow, s));
        else
            CSG_TRY(copy_async<T>(pl.em[e], a.flow, 2 * P * jobs, s));
    }
    return kOk;
}

#define CSG_INSTANTIATE_SOLVE(T)                                                               \
    template size_t fmg_workspace_bytes<T>(int, int, int);                                     \
    template int fmg_run<T>(const T *, const T *, const double *, int, int, const T *, int,     \
                            int, int, int, T *, void *, size_t, hipStream_t);                  \
    template size_t step_workspace_bytes<T>(int, int, int);                                    \
    template int dartel_step_run<T>(const StepArgs<T> &, void *, size_t, hipStream_t);          \
    template size_t solve_workspace_bytes<T>(int, int, int);                                   \
    template int solve_run<T>(const SolveArgs<T> &, void *, size_t, hipStream_t);

CSG_INSTANTIATE_SOLVE(float)
CSG_INSTANTIATE_SOLVE(double)

}  // namespace csg
