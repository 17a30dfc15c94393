// Fused schedule of the membrane full multigrid (dartel_grid.py:1211-1417).
//
// The reference relaxes, restricts and prolongs level by level, one operator
// call (here: one launch) at a time; at 512x256 a 3-cycle FMG is ~450 small
// dependent launches.  This schedule keeps the arithmetic of every pixel and
// its order exactly (the per-pixel formulas come from fmg_device.h, as in the
// one-launch-per-sweep kernels) and only regroups the work:
//
//   * fine levels: all 2·nit half-sweeps of one relax_membrane call run in one
//     launch on LDS tiles with a halo of 2·nit pixels (temporal blocking: a
//     pixel at distance >= s from the tile edge is exact after s sweeps); the
//     prolongation that precedes the upward relaxation is added while the tile
//     is loaded; residual and restriction are one launch.
//   * coarse levels (from the first level whose whole sub-pyramid fits in LDS
//     down to 1x1): one workgroup per job runs every relaxation, residual,
//     restriction, prolongation and the coarse 2x2 solve in LDS, with only
//     workgroup barriers between steps.
#include "dartel_kernels.h"
#include "fmg_device.h"
#include "kernel_timing.h"

#include <map>
#include <mutex>
#include <type_traits>

namespace csg {

// ---------------------------------------------------------------- fine levels

namespace {

// One relax_membrane call (`sweeps` = 2·nit half-sweeps, first colour
// (x + y) & 1 == 1) on a TX x TY interior (TX even).  The tile holds columns
// x0 - H .. x0 + TX + H - 1 in unwrapped coordinates (nx is even, so a
// pixel's colour is the same unwrapped or wrapped) and the rows of
// y0 - H .. y0 + TY + H - 1 that exist (the y boundary reflects onto rows
// inside the grid).  Half-sweep s updates the pixels at distance >= s + 1
// from the tile's cut edges (a row range that reaches the grid's y boundary
// is not cut): after s sweeps every pixel at distance >= s holds the value of
// the global sweep, so the interior (distance >= H) is exact after H.
//
// u is staged in LDS with the two channels of a pixel side by side (one
// 8- or 16-byte LDS access per neighbour).  Each thread owns up to kTilePix
// pixels of each colour for the whole call (colour c's pixels of the tile in
// row-major order, entries tid + k·kTileThreads), with their LDS offsets,
// right-hand side and 2x2 block (relax_coef) in registers.  The kernel is
// bound by VALU issue, so the index arithmetic around the sweeps is kept
// short: tile columns are wrapped by one compare when the tiles stay within
// one period of the grid, and the store loop divides through a reciprocal.
//
// With a coarse correction Uc the tile loads u + prolong(Uc).  `colsum` (set
// by the host for exact 2:1 levels, nx = 2·cnx and ny = 2·cny) evaluates the
// prolongation's y pass once per (tile row, coarse column) into LDS and the x
// pass per pixel from those sums, with the taps of every tile row and column
// resolved once into LDS tables: the same operations in the same order as
// resize_from, with 3 coarse reads per column sum instead of 18 per pixel.  At
// 2:1 resize_taps' origin for fine column g is exactly g >> 1, so unwrapped
// tile column x0 + i reads the column sums of unwrapped coarse columns
// ((x0 + i) >> 1) - 1 .. + 1, which are wrapped when the coarse values are
// fetched.
//
// HC > 0 fixes the halo (= sweeps) at compile time; HC == 0 takes it from H.
constexpr int kTilePix = 4;
// tile pixels one thread loads (tile_fits_threads: at most 2·kTilePix each)
constexpr int kTileLoad = 2 * kTilePix;

template <typename T>
struct alignas(2 * sizeof(T)) U2 {
    T x, y;
};

// Resolved y taps of one tile row (coarse row offsets premultiplied by cnx) and
// x taps of one tile column (first column-sum slot m) of the 2:1 prolongation.
template <typename T>
struct RowTaps {
    int ym, y0, yp;
    T wm, w0, wp;
};

template <typename T>
struct ColTaps {
    int m;
    T wm, w0, wp;
};

// LDS bytes of the column-sum prolongation for a tile of ex x rows pixels.
template <typename T>
constexpr size_t colsum_lds_bytes(size_t ex, size_t rows)
{
    return rows * (ex / 2 + 3) * sizeof(U2<T>) + rows * sizeof(RowTaps<T>) +
           ex * sizeof(ColTaps<T>);
}

// ix wrapped into [0, n) for -n <= ix < 2n.
__device__ __forceinline__ int wrap_once(int ix, int n)
{
    return ix < 0 ? ix + n : (ix >= n ? ix - n : ix);
}

template <typename T, int HC>
__global__ void __launch_bounds__(kTileThreads)
k_relax_tile(const T *__restrict__ A, const T *__restrict__ B, const T *__restrict__ Uin,
             T *__restrict__ Uout, const T *__restrict__ Uc, int cnx, int cny, int nx, int ny,
             T w00, T w01, T w10, int H_arg, int TX, int TY, int tiles_x, int colsum,
             size_t astride, size_t bstride, size_t cstride)
{
    const int H = HC > 0 ? HC : H_arg;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    U2<T> *uu = reinterpret_cast<U2<T> *>(lds_raw);
    const int P = nx * ny, CP = cnx * cny;
    // tiles relabelled XCD by XCD: a tile's halo lies in neighbouring tiles
    const Block2 blk = xcd_block();
    const int job = blk.y;
    A += job * astride;
    B += job * bstride;
    Uin += job * bstride;
    Uout += job * bstride;
    if (Uc)
        Uc += job * cstride;
    const int tyi = blk.x / tiles_x, txi = blk.x - tyi * tiles_x;
    const int ex = TX + 2 * H;
    const int x0 = txi * TX - H;
    const int ylo = max(0, tyi * TY - H), yhi = min(ny, tyi * TY + TY + H);
    const int rows = yhi - ylo;
    const int npx = ex * rows;
    // every unwrapped tile column lies in [-nx + 2, 2nx - 4), so one compare
    // wraps it, and the coarse columns of its prolongation taps (2:1) lie in
    // [-cnx, 2cnx)
    const bool once = H + 2 <= nx && tiles_x * TX + H + 4 <= 2 * nx;
    auto wrapc = [&](int ix) { return once ? wrap_once(ix, nx) : wrap_x(ix, nx); };
    const float inv_ex = 1.0f / (float)ex;
    // this thread's pixels, and their A and b loads issued before the tile's
    // (both sets of global loads in flight at once; the relaxation
    // coefficients are formed after the tile is in LDS)
    const int hex = ex / 2;  // pixels of one colour per tile row (ex is even)
    const float inv_hex = 1.0f / (float)hex;
    int pq[2][kTilePix], pym[2][kTilePix], pyp[2][kTilePix], plim[2][kTilePix];
    RelaxCoef<T> pk[2][kTilePix];
    T pa0[2][kTilePix], pa1[2][kTilePix], pa2[2][kTilePix];
    T pbx[2][kTilePix], pby[2][kTilePix];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int k = 0; k < kTilePix; ++k) {
            const int j = threadIdx.x + k * kTileThreads;
            plim[c][k] = 0;
            pq[c][k] = pym[c][k] = pyp[c][k] = 0;
            pa0[c][k] = pa1[c][k] = (T)1;
            pa2[c][k] = (T)0;
            pbx[c][k] = pby[c][k] = (T)0;
            if (j >= hex * rows)
                continue;
            const DivMod rk = divmod_small(j, hex, inv_hex);
            const int r = rk.q, gy = ylo + r;
            // the column of this row with (x0 + i + gy) & 1 == c
            const int i = 2 * rk.r + ((c - x0 - gy) & 1);
            int d = min(i, ex - 1 - i);
            if (ylo > 0)
                d = min(d, r);
            if (yhi < ny)
                d = min(d, rows - 1 - r);
            if (d < 1)
                continue;
            const int q = r * ex + i;
            const int p = gy * nx + wrapc(x0 + i);
            plim[c][k] = d;
            pq[c][k] = q;
            // the y neighbours; the grid's first and last rows reflect onto themselves
            pym[c][k] = gy == 0 ? q : q - ex;
            pyp[c][k] = gy == ny - 1 ? q : q + ex;
            pa0[c][k] = A[p];
            pa1[c][k] = A[P + p];
            pa2[c][k] = A[2 * P + p];
            pbx[c][k] = B[p];
            pby[c][k] = B[P + p];
        }
    }
    if (!Uc || colsum) {
        // u of the tile (every load issued before the first use)
        U2<T> v[kTileLoad];
#pragma unroll
        for (int j = 0; j < kTileLoad; ++j) {
            const int q = threadIdx.x + j * kTileThreads;
            v[j] = U2<T>{(T)0, (T)0};
            if (q < npx) {
                const DivMod ri = divmod_small(q, ex, inv_ex);
                const int p = (ylo + ri.q) * nx + wrapc(x0 + ri.r);
                v[j] = U2<T>{Uin[p], Uin[P + p]};
            }
        }
        if (Uc) {
            // y pass of the prolongation: column sums of the unwrapped coarse
            // columns cbase .. cbase + W - 1 for every tile row, both channels
            const int cbase = (x0 >> 1) - 1;
            const int W = ((x0 + ex - 1) >> 1) + 2 - cbase;
            U2<T> *cs = uu + npx;
            RowTaps<T> *rt = reinterpret_cast<RowTaps<T> *>(cs + rows * W);
            ColTaps<T> *ct = reinterpret_cast<ColTaps<T> *>(rt + rows);
            for (int t = threadIdx.x; t < rows + ex; t += kTileThreads) {
                if (t < rows) {
                    const Taps<T> ty = resize_taps<T>(ylo + t, cny, ny);
                    rt[t] = RowTaps<T>{reflect_y(ty.o - 1, cny) * cnx, reflect_y(ty.o, cny) * cnx,
                                       reflect_y(ty.o + 1, cny) * cnx, ty.wm, ty.w0, ty.wp};
                } else {
                    const int xu = x0 + (t - rows), gx = wrapc(xu);
                    const Taps<T> tx = resize_taps<T>(gx, cnx, nx);
                    // unwrapped origin: tx.o plus the whole coarse periods of xu - gx
                    ct[t - rows] = ColTaps<T>{tx.o + ((xu - gx) >> 1) - 1 - cbase, tx.wm, tx.w0,
                                              tx.wp};
                }
            }
            __syncthreads();
            const float inv_w = 1.0f / (float)W;
            for (int q = threadIdx.x; q < rows * W; q += kTileThreads) {
                const DivMod rk = divmod_small(q, W, inv_w);
                const RowTaps<T> r = rt[rk.q];
                const int cx = once ? wrap_once(cbase + rk.r, cnx) : wrap_x(cbase + rk.r, cnx);
                cs[q] = U2<T>{r.wm * Uc[r.ym + cx] + r.w0 * Uc[r.y0 + cx] + r.wp * Uc[r.yp + cx],
                              r.wm * Uc[CP + r.ym + cx] + r.w0 * Uc[CP + r.y0 + cx] +
                                  r.wp * Uc[CP + r.yp + cx]};
            }
            __syncthreads();
            // x pass per tile pixel, added to u
#pragma unroll
            for (int j = 0; j < kTileLoad; ++j) {
                const int q = threadIdx.x + j * kTileThreads;
                if (q < npx) {
                    const DivMod ri = divmod_small(q, ex, inv_ex);
                    const ColTaps<T> c = ct[ri.r];
                    const U2<T> *sm = cs + ri.q * W + c.m;
                    const U2<T> a = sm[0], b = sm[1], d = sm[2];
                    v[j].x = v[j].x + (c.wm * a.x + c.w0 * b.x + c.wp * d.x);
                    v[j].y = v[j].y + (c.wm * a.y + c.w0 * b.y + c.wp * d.y);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kTileLoad; ++j) {
            const int q = threadIdx.x + j * kTileThreads;
            if (q < npx)
                uu[q] = v[j];
        }
    } else {
        // u + prolong(Uc) per pixel from the coarse level (any size ratio)
        for (int q = threadIdx.x; q < npx; q += kTileThreads) {
            const DivMod ri = divmod_small(q, ex, inv_ex);
            const int gx = wrapc(x0 + ri.r), gy = ylo + ri.q, p = gy * nx + gx;
            const Taps<T> tx = resize_taps<T>(gx, cnx, nx), ty = resize_taps<T>(gy, cny, ny);
            const T rx = resize_at<T>(tx, ty, cnx, cny,
                                      [&](int x, int y) { return Uc[y * cnx + x]; });
            const T ry = resize_at<T>(tx, ty, cnx, cny,
                                      [&](int x, int y) { return Uc[CP + y * cnx + x]; });
            uu[q] = U2<T>{Uin[p] + rx, Uin[P + p] + ry};
        }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < kTilePix; ++k)
            pk[c][k] = plim[c][k] > 0 ? relax_coef<T>(pa0[c][k], pa1[c][k], pa2[c][k], w00)
                                      : RelaxCoef<T>{(T)1, (T)1, (T)0, (T)1};
    __syncthreads();
    // half-sweep s relaxes colour 1 - (s & 1); the colour is a compile-time
    // index into the register arrays (a run-time one would put them in
    // scratch memory), so the sweeps run in (1, 0) pairs
    auto sweep = [&](auto cc, int s) {
        constexpr int c = decltype(cc)::value;
#pragma unroll
        for (int k = 0; k < kTilePix; ++k) {
            if (s < plim[c][k]) {
                const int q = pq[c][k];
                const U2<T> a = uu[pym[c][k]], b = uu[pyp[c][k]], l = uu[q - 1], r = uu[q + 1];
                const Stencil<T> n{(T)0, (T)0, a.x + b.x, l.x + r.x, a.y + b.y, l.y + r.y};
                const Pair<T> v = relax_solve<T>(pbx[c][k], pby[c][k], n, pk[c][k], w01, w10);
                uu[q] = U2<T>{v.x, v.y};
            }
        }
        __syncthreads();
    };
    for (int s = 0; s + 1 < H; s += 2) {
        sweep(std::integral_constant<int, 1>{}, s);
        sweep(std::integral_constant<int, 0>{}, s + 1);
    }
    if (H & 1)
        sweep(std::integral_constant<int, 1>{}, H - 1);
    const float inv_tx = 1.0f / (float)TX;
    for (int q = threadIdx.x; q < TX * TY; q += kTileThreads) {
        const DivMod yx = divmod_small(q, TX, inv_tx);
        const int gx = txi * TX + yx.r, gy = tyi * TY + yx.q;
        if (gx >= nx || gy >= ny)
            continue;
        const U2<T> w = uu[(gy - ylo) * ex + H + yx.r];
        Uout[gy * nx + gx] = w.x;
        Uout[P + gy * nx + gx] = w.y;
    }
}

// Residual of the finer level restricted to the coarser one; the coarser
// level's correction starts from zero.  dartel_grid.py:1356-1368
//
// Four lanes per coarse pixel: lane k < 3 evaluates the three residual pairs
// of x tap k and their y-weighted column sum, and lane 0 combines the three
// columns (resize_from's order per channel, so the values are the same).  Splitting the
// 9 stencil evaluations over lanes gives each thread a third of the loads and
// four times as many waves to hide their latency.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_restrict_residual(const T *__restrict__ A, const T *__restrict__ B, const T *__restrict__ U,
                    int nx, int ny, int cnx, int cny, T w00, T w01, T w10, size_t astride,
                    size_t bstride, size_t cstride, T *__restrict__ Bc, T *__restrict__ Uc)
{
    const int CP = cnx * cny, P = nx * ny;
    const Block2 blk = xcd_block();
    const int gid = blk.x * blockDim.x + threadIdx.x;
    const int q = gid >> 2, k = gid & 3;
    const bool live = q < CP;
    const int job = blk.y;
    A += job * astride;
    B += job * bstride;
    U += job * bstride;
    Bc += job * cstride;
    Uc += job * cstride;
    const int cy = live ? q / cnx : 0, cx = live ? q - cy * cnx : 0;
    const Taps<T> tx = resize_taps<T>(cx, nx, cnx), ty = resize_taps<T>(cy, ny, cny);
    Pair<T> col{(T)0, (T)0};
    if (live && k < 3) {
        auto res = [&](int x, int y) {
            const int p = y * nx + x;
            const Pair<T> m = membrane_pixel<T>(stencil_at<T>(U, P, nx, ny, x, y), A[p],
                                                A[P + p], A[2 * P + p], w00, w01, w10);
            return Pair<T>{B[p] - m.x, B[P + p] - m.y};
        };
        const int x = wrap_x(tx.o - 1 + k, nx);
        const Pair<T> a = res(x, reflect_y(ty.o - 1, ny)), b = res(x, reflect_y(ty.o, ny)),
                      c = res(x, reflect_y(ty.o + 1, ny));
        col = {ty.wm * a.x + ty.w0 * b.x + ty.wp * c.x, ty.wm * a.y + ty.w0 * b.y + ty.wp * c.y};
    }
    const T c1x = __shfl_down(col.x, 1, 4), c1y = __shfl_down(col.y, 1, 4);
    const T c2x = __shfl_down(col.x, 2, 4), c2y = __shfl_down(col.y, 2, 4);
    if (!live || k != 0)
        return;
    Bc[q] = tx.wm * col.x + tx.w0 * c1x + tx.wp * c2x;
    Bc[CP + q] = tx.wm * col.y + tx.w0 * c1y + tx.wp * c2y;
    Uc[q] = (T)0;
    Uc[CP + q] = (T)0;
}

// ---------------------------------------------------------------- coarse levels
//
// One workgroup per job holds levels c0 .. L-1 in LDS (per level: A 3 planes,
// b 2, u 2) plus a two-plane residual buffer of level c0's size.  The schedule
// is the reference's (dartel_grid.py:1341-1413).  Every step is coarse_op
// with plain scalar arguments (level sizes, LDS offsets), inlined at each of
// its few call sites: inside it every access is an LDS offset and every loop
// strides by the fixed block size, so a step costs no global memory access.
// (As a called function its frame kept callee-saved registers in scratch
// memory, and a run-time index into the relaxation's register arrays put them
// there too: 78.5 -> 55.5 us per launch in float32, 106 -> 45 us in float64.
// Running the levels of <= 128 pixels on one wave without workgroup barriers
// measured slower again: the steps wait on their dependent LDS and arithmetic
// chains, not on the barrier.)

extern __shared__ __align__(16) unsigned char g_lds[];

// The level table, copied to LDS at kernel start.
template <typename T>
struct LevelTable {
    int nx[kFmgMaxLevels], ny[kFmgMaxLevels];
    unsigned off[kFmgMaxLevels];
    unsigned roff;  // residual buffer
    // resolved resize taps (TapIdx entries from the tap region): down[l] maps
    // level l to l + 1 (one x entry per column of l + 1, one y entry per row),
    // up[l] maps l + 1 to l
    unsigned down_x[kFmgMaxLevels], down_y[kFmgMaxLevels], up_x[kFmgMaxLevels],
        up_y[kFmgMaxLevels];
    T w00[kFmgMaxLevels], w01[kFmgMaxLevels], w10[kFmgMaxLevels];
    // reciprocals for the float divmods of each level: 1 / nx, 1 / (nx·ny) and
    // 1 / ceil(nx / 2) (the same expressions the steps evaluated themselves)
    float inv_nx[kFmgMaxLevels], inv_p[kFmgMaxLevels], inv_half[kFmgMaxLevels];
    unsigned tap0, tap1;  // the tap entries' range
};

// CoarseLevels::meta: the table, then the tap entries from a 256-byte boundary.
template <typename T>
constexpr size_t meta_taps_offset()
{
    return (sizeof(LevelTable<T>) + 255) / 256 * 256;
}

enum CoarseOp : int { kRelax, kDown, kUpAdd, kProlong, kPyramidA, kPyramidB };

// LDS resize of `channels` planes from resolved taps: out = resize(in) [+ out].
// Channels are spread over threads as well as pixels.
template <typename T>
__device__ __forceinline__ void lds_resize(const T *in, int snx, int sny, T *out, int tnx,
                                           int tny, int channels, bool accumulate,
                                           const TapIdx<T> *tabx, const TapIdx<T> *taby,
                                           float inv_tp, float inv_tnx)
{
    const int SP = snx * sny, TP = tnx * tny;
    for (int q = threadIdx.x; q < channels * TP; q += kCoarseThreads) {
        const DivMod cp = divmod_small(q, TP, inv_tp);
        const DivMod yx = divmod_small(cp.r, tnx, inv_tnx);
        const T *f = in + cp.q * SP;
        T r = resize_from<T>(tabx[yx.r], taby[yx.q],
                             [&](int x, int y) { return f[y * snx + x]; });
        T *o = out + q;
        *o = accumulate ? *o + r : r;
    }
}

// One step of the schedule on level (nx, ny) at LDS offset `off`; the other
// level (onx, ony, ooff) is l+1 for kDown / kUpAdd / kProlong and l-1 for the
// pyramid steps.  inv_* are the two levels' reciprocals from the table.  Ends
// with a workgroup barrier.
template <typename T>
__device__ __forceinline__ void coarse_op(int op, int nx, int ny, unsigned off, T w00, T w01, T w10,
                                       int onx, int ony, unsigned ooff, unsigned roff,
                                       unsigned tx_off, unsigned ty_off, int nit, float inv_nx,
                                       float inv_p, float inv_half, float oinv_nx, float oinv_p)
{
    T *base = reinterpret_cast<T *>(g_lds);
    const TapIdx<T> *tabx = reinterpret_cast<const TapIdx<T> *>(g_lds) + tx_off;
    const TapIdx<T> *taby = reinterpret_cast<const TapIdx<T> *>(g_lds) + ty_off;
    const int P = nx * ny, OP = onx * ony;
    T *a = base + off, *b = a + 3 * P, *u = a + 5 * P;
    T *oa = base + ooff, *ob = oa + 3 * OP, *ou = oa + 5 * OP;
    if (op == kRelax) {
        // relax_membrane: 2·nit half-sweeps, two-phase order for odd widths
        const int half = (nx + 1) >> 1;
        const bool odd = (nx & 1) && nx > 1;
        // one first-phase pixel per thread and colour: its stencil offsets, 2x2
        // block and right-hand side stay in registers for the whole call (A and
        // b do not change during a relaxation), the same values relax_pixel
        // reads.  The odd-width last column keeps the general loop below.
        const bool in_regs = half * ny <= kCoarseThreads;
        int rp[2] = {0, 0}, rxm[2] = {0, 0}, rxp[2] = {0, 0}, rym[2] = {0, 0}, ryp[2] = {0, 0};
        bool rok[2] = {false, false};
        RelaxCoef<T> rk[2] = {};
        T rbx[2] = {(T)0, (T)0}, rby[2] = {(T)0, (T)0};
        if (in_regs && (int)threadIdx.x < half * ny) {
            const DivMod yk = divmod_small((int)threadIdx.x, half, inv_half);
            const int y = yk.q, row = y * nx;
            // both colours without branches, so that their loads and divisions
            // interleave: a pixel that does not exist (rok false) is replaced by
            // column 2·r of the row, which does, and is never updated
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int xc = 2 * yk.r + ((c + y) & 1);
                rok[c] = xc < nx && !(odd && xc == nx - 1);
                const int x = xc < nx ? xc : 2 * yk.r;
                const int p = row + x;
                rp[c] = p;
                rxm[c] = row + (x == 0 ? nx - 1 : x - 1);
                rxp[c] = row + (x == nx - 1 ? 0 : x + 1);
                rym[c] = (y == 0 ? 0 : y - 1) * nx + x;
                ryp[c] = (y == ny - 1 ? y : y + 1) * nx + x;
            }
            T a0[2], a1[2], a2[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int p = rp[c];
                a0[c] = a[p];
                a1[c] = a[P + p];
                a2[c] = a[2 * P + p];
                rbx[c] = b[p];
                rby[c] = b[P + p];
            }
#pragma unroll
            for (int c = 0; c < 2; ++c)
                rk[c] = relax_coef<T>(a0[c], a1[c], a2[c], w00);
        }
        // the half-sweeps in colour pairs (1, 0), so that every index into the
        // register arrays is a compile-time constant (a run-time index would
        // place them in scratch memory)
        auto in_regs_sweep = [&](auto cc) {
            constexpr int c = decltype(cc)::value;
            if (rok[c]) {
                const int p = rp[c], xm = rxm[c], xp = rxp[c], ym = rym[c], yp = ryp[c];
                const Stencil<T> st{u[p],           u[P + p],           u[ym] + u[yp],
                                    u[xm] + u[xp], u[P + ym] + u[P + yp], u[P + xm] + u[P + xp]};
                const Pair<T> v = relax_solve<T>(rbx[c], rby[c], st, rk[c], w01, w10);
                u[p] = v.x;
                u[P + p] = v.y;
            }
            __syncthreads();
        };
        if (in_regs && !odd) {
            for (int it = 0; it < nit; ++it) {
                in_regs_sweep(std::integral_constant<int, 1>{});
                in_regs_sweep(std::integral_constant<int, 0>{});
            }
            return;
        }
        for (int it = 0; it < 2 * nit; ++it) {
            const int colour = (it % 2 == 0) ? 1 : 0;
            if (in_regs) {
                if (colour == 1)
                    in_regs_sweep(std::integral_constant<int, 1>{});
                else
                    in_regs_sweep(std::integral_constant<int, 0>{});
            }
            for (int phase = in_regs ? 1 : 0; phase < (odd ? 2 : 1); ++phase) {
                const int count = phase == 0 ? half * ny : ny;
                for (int q = threadIdx.x; q < count; q += kCoarseThreads) {
                    int x, y;
                    if (phase == 0) {
                        const DivMod yk = divmod_small(q, half, inv_half);
                        y = yk.q;
                        x = 2 * yk.r + ((colour + y) & 1);
                        if (x >= nx || (odd && x == nx - 1))
                            continue;
                    } else {
                        y = q;
                        x = nx - 1;
                        if (((x + y) & 1) != colour)
                            continue;
                    }
                    const int p = y * nx + x;
                    const Pair<T> v =
                        relax_pixel<T>(b[p], b[P + p], stencil_at<T>(u, P, nx, ny, x, y), a[p],
                                       a[P + p], a[2 * P + p], w00, w01, w10);
                    u[p] = v.x;
                    u[P + p] = v.y;
                }
                __syncthreads();
            }
        }
    } else if (op == kDown) {
        // r = b - (L + A) u; b[l+1] = restrict(r); u[l+1] = 0.  dartel_grid.py:1327-1368
        T *r = base + roff;
        for (int p = threadIdx.x; p < P; p += kCoarseThreads) {
            const DivMod yx = divmod_small(p, nx, inv_nx);
            const int y = yx.q, x = yx.r;
            const Pair<T> m = membrane_pixel<T>(stencil_at<T>(u, P, nx, ny, x, y), a[p],
                                                a[P + p], a[2 * P + p], w00, w01, w10);
            r[p] = b[p] - m.x;
            r[P + p] = b[P + p] - m.y;
        }
        __syncthreads();
        lds_resize<T>(r, nx, ny, ob, onx, ony, 2, false, tabx, taby, oinv_p, oinv_nx);
        for (int p = threadIdx.x; p < 2 * OP; p += kCoarseThreads)
            ou[p] = (T)0;
        __syncthreads();
    } else if (op == kUpAdd || op == kProlong) {
        lds_resize<T>(ou, onx, ony, u, nx, ny, 2, op == kUpAdd, tabx, taby, inv_p, inv_nx);
        __syncthreads();
    } else {
        // pyramid: A (or the original b) of this level from the finer one
        if (op == kPyramidA)
            lds_resize<T>(oa, onx, ony, a, nx, ny, 3, false, tabx, taby, inv_p, inv_nx);
        else
            lds_resize<T>(ob, onx, ony, b, nx, ny, 2, false, tabx, taby, inv_p, inv_nx);
        __syncthreads();
    }
}

// coarse_op at level l of the table (the table is in LDS, the call takes scalars).
template <typename T>
__device__ __forceinline__ void run_op(int op, int l, const LevelTable<T> &t, int nit)
{
    // the other level: none for a relaxation (pass l itself), l - 1 for the
    // pyramid steps, l + 1 otherwise
    const bool pyramid = op == kPyramidA || op == kPyramidB;
    const int o = op == kRelax ? l : (pyramid ? l - 1 : l + 1);
    // taps: restriction l -> l + 1 (kDown), prolongation l + 1 -> l (kUpAdd,
    // kProlong), restriction l - 1 -> l (pyramid)
    unsigned tx = 0, ty = 0;
    if (op == kDown) {
        tx = t.down_x[l];
        ty = t.down_y[l];
    } else if (op == kUpAdd || op == kProlong) {
        tx = t.up_x[l];
        ty = t.up_y[l];
    } else if (pyramid) {
        tx = t.down_x[l - 1];
        ty = t.down_y[l - 1];
    }
    coarse_op<T>(op, t.nx[l], t.ny[l], t.off[l], t.w00[l], t.w01[l], t.w10[l], t.nx[o], t.ny[o],
                 t.off[o], t.roff, tx, ty, nit, t.inv_nx[l], t.inv_p[l], t.inv_half[l],
                 t.inv_nx[o], t.inv_p[o]);
}

// `build`: the first launch of a multigrid solve, which makes the level table,
// the resolved taps and the A pyramid and stores them in lv.meta / lv.apyr;
// the solve's later launches load them (the same values: A does not change
// during a solve).
template <typename T>
__global__ void __launch_bounds__(kCoarseThreads)
k_fmg_coarse(CoarseLevels<T> lv, const T *__restrict__ A, const T *__restrict__ B,
             T *__restrict__ U, size_t astride, size_t bstride, double id, int nit, int cycles,
             int init, int build, int P0)
{
    __shared__ LevelTable<T> tab;
    T *base = reinterpret_cast<T *>(g_lds);
    const int job = blockIdx.x;
    A += job * astride;
    B += job * bstride;
    U += job * bstride;
    const int c0 = lv.c0, L = lv.L, last = L - 1;
    TapIdx<T> *taps = reinterpret_cast<TapIdx<T> *>(g_lds);
    const TapIdx<T> *gtaps = reinterpret_cast<const TapIdx<T> *>(lv.meta + meta_taps_offset<T>());
    // level c0 (at LDS offset 0 in every layout) does not depend on the table:
    // its loads go out with the table's
    T *ta = base, *tb = ta + 3 * P0, *tu = ta + 5 * P0;
    for (int p = threadIdx.x; p < P0; p += kCoarseThreads) {
        ta[p] = A[p];
        ta[P0 + p] = A[P0 + p];
        ta[2 * P0 + p] = A[2 * P0 + p];
        tb[p] = B[p];
        tb[P0 + p] = B[P0 + p];
        if (!init) {
            tu[p] = U[p];
            tu[P0 + p] = U[P0 + p];
        }
    }
    if (!build) {
        const unsigned *src = reinterpret_cast<const unsigned *>(lv.meta);
        unsigned *dst = reinterpret_cast<unsigned *>(&tab);
        for (int i = threadIdx.x; i < (int)(sizeof(LevelTable<T>) / 4); i += kCoarseThreads)
            dst[i] = src[i];
        __syncthreads();
        for (int t = threadIdx.x; t < (int)(tab.tap1 - tab.tap0); t += kCoarseThreads)
            taps[tab.tap0 + t] = gtaps[t];
        for (int l = c0 + 1; l <= last; ++l) {
            const int P = tab.nx[l] * tab.ny[l];
            const T *ga = lv.apyr[l] + (size_t)job * 3 * P;
            T *la = base + tab.off[l];
            for (int p = threadIdx.x; p < 3 * P; p += kCoarseThreads)
                la[p] = ga[p];
        }
    } else if (threadIdx.x == 0) {
        // constant indices only: a run-time index into the argument would
        // copy it to scratch memory
        unsigned off = 0;
#pragma unroll
        for (int l = 0; l < kFmgMaxLevels; ++l) {
            if (l >= c0 && l < L) {
                tab.nx[l] = lv.nx[l];
                tab.ny[l] = lv.ny[l];
                tab.w00[l] = lv.w00[l];
                tab.w01[l] = lv.w01[l];
                tab.w10[l] = lv.w10[l];
                tab.inv_nx[l] = 1.0f / (float)lv.nx[l];
                tab.inv_p[l] = 1.0f / (float)(lv.nx[l] * lv.ny[l]);
                tab.inv_half[l] = 1.0f / (float)((lv.nx[l] + 1) >> 1);
                tab.off[l] = off;
                off += 7u * (unsigned)(lv.nx[l] * lv.ny[l]);
            }
        }
        tab.roff = off;
        off += 2u * (unsigned)(lv.nx[c0] * lv.ny[c0]);
        // tap entries start at the first TapIdx-aligned position after the planes
        unsigned e = (unsigned)((off * sizeof(T) + sizeof(TapIdx<T>) - 1) / sizeof(TapIdx<T>));
        tab.tap0 = e;
#pragma unroll
        for (int l = 0; l < kFmgMaxLevels - 1; ++l) {
            if (l >= c0 && l < L - 1) {
                tab.down_x[l] = e;
                e += (unsigned)lv.nx[l + 1];
                tab.down_y[l] = e;
                e += (unsigned)lv.ny[l + 1];
                tab.up_x[l] = e;
                e += (unsigned)lv.nx[l];
                tab.up_y[l] = e;
                e += (unsigned)lv.ny[l];
            }
        }
        tab.tap1 = e;
    }
    __syncthreads();
    // resolve every resize tap of the sub-pyramid once
    for (int l = c0; l < last && build; ++l) {
        const int fx = tab.nx[l], fy = tab.ny[l], cx = tab.nx[l + 1], cy = tab.ny[l + 1];
        for (int t = threadIdx.x; t < cx + cy + fx + fy; t += kCoarseThreads) {
            if (t < cx)
                taps[tab.down_x[l] + t] = taps_x<T>(t, fx, cx);
            else if (t < cx + cy)
                taps[tab.down_y[l] + t - cx] = taps_y<T>(t - cx, fy, cy);
            else if (t < cx + cy + fx)
                taps[tab.up_x[l] + t - cx - cy] = taps_x<T>(t - cx - cy, cx, fx);
            else
                taps[tab.up_y[l] + t - cx - cy - fx] = taps_y<T>(t - cx - cy - fx, cy, fy);
        }
    }
    __syncthreads();
    // pyramids of A (and, for the coarse initialisation, of the original b)
    for (int l = c0 + 1; l <= last; ++l) {
        if (build)
            run_op<T>(kPyramidA, l, tab, nit);
        if (init)
            run_op<T>(kPyramidB, l, tab, nit);
    }
    if (build) {
        // kept for the solve's later launches
        for (int l = c0 + 1; l <= last; ++l) {
            const int P = tab.nx[l] * tab.ny[l];
            T *ga = lv.apyr[l] + (size_t)job * 3 * P;
            const T *la = base + tab.off[l];
            for (int p = threadIdx.x; p < 3 * P; p += kCoarseThreads)
                ga[p] = la[p];
        }
        if (job == 0) {
            const unsigned *src = reinterpret_cast<const unsigned *>(&tab);
            unsigned *dst = reinterpret_cast<unsigned *>(lv.meta);
            for (int i = threadIdx.x; i < (int)(sizeof(LevelTable<T>) / 4); i += kCoarseThreads)
                dst[i] = src[i];
            TapIdx<T> *out = reinterpret_cast<TapIdx<T> *>(lv.meta + meta_taps_offset<T>());
            for (int t = threadIdx.x; t < (int)(tab.tap1 - tab.tap0); t += kCoarseThreads)
                out[t] = taps[tab.tap0 + t];
        }
    }
    if (init) {
        // dartel_grid.py:1301-1326: direct 2x2 solve against the original b of
        // the coarsest level, then relaxation against that level's (zero) b
        const int PC = tab.nx[last] * tab.ny[last];
        T *ca = base + tab.off[last], *cb = ca + 3 * PC, *cu = ca + 5 * PC;
        for (int p = threadIdx.x; p < PC; p += kCoarseThreads) {
            double dx = (double)ca[p] + id, dy = (double)ca[PC + p] + id,
                   cr = (double)ca[2 * PC + p];
            double det = dx * dy * 1.0000000001 - cr * cr;
            double bx = (double)cb[p], by = (double)cb[PC + p];
            cu[p] = (T)((dy * bx - cr * by) / det);
            cu[PC + p] = (T)((-cr * bx + dx * by) / det);
        }
        __syncthreads();
        for (int p = threadIdx.x; p < 2 * PC; p += kCoarseThreads)
            cb[p] = (T)0;
        __syncthreads();
        run_op<T>(kRelax, last, tab, nit);
        // dartel_grid.py:1402-1413; b of each level still holds its original
        // pyramid value when the level is reached (only coarser levels have
        // been overwritten by then)
        for (int level = last - 1; level >= c0; --level) {
            run_op<T>(kProlong, level, tab, nit);
            for (int cyc = 0; cyc < cycles; ++cyc) {
                // cycle(level), dartel_grid.py:1341-1391
                for (int l = level; l < last; ++l) {
                    run_op<T>(kRelax, l, tab, nit);
                    run_op<T>(kDown, l, tab, nit);
                }
                run_op<T>(kRelax, last, tab, nit);
                for (int l = last - 1; l >= level; --l) {
                    run_op<T>(kUpAdd, l, tab, nit);
                    run_op<T>(kRelax, l, tab, nit);
                }
            }
        }
    } else {
        for (int cyc = 0; cyc < cycles; ++cyc) {
            for (int l = c0; l < last; ++l) {
                run_op<T>(kRelax, l, tab, nit);
                run_op<T>(kDown, l, tab, nit);
            }
            run_op<T>(kRelax, last, tab, nit);
            for (int l = last - 1; l >= c0; --l) {
                run_op<T>(kUpAdd, l, tab, nit);
                run_op<T>(kRelax, l, tab, nit);
            }
        }
    }
    for (int p = threadIdx.x; p < P0; p += kCoarseThreads) {
        U[p] = tu[p];
        U[P0 + p] = tu[P0 + p];
    }
}

}  // namespace

// ---------------------------------------------------------------- LDS limits

namespace {

std::once_flag g_lds_once;
size_t g_lds_budget = 65536;

template <typename F>
void raise_limit(F kernel, int bytes)
{
    (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              bytes);
}

void init_lds_budget()
{
    int dev = 0, per_block = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess ||
        per_block <= 0) {
        (void)hipGetLastError();
        return;
    }
    // leave room for the kernels' static shared variables
    int budget = per_block - 4096;
    raise_limit(k_relax_tile<float, 0>, budget);
    raise_limit(k_relax_tile<float, 6>, budget);
    raise_limit(k_relax_tile<double, 0>, budget);
    raise_limit(k_relax_tile<double, 6>, budget);
    raise_limit(k_fmg_coarse<float>, budget);
    raise_limit(k_fmg_coarse<double>, budget);
    if (hipGetLastError() == hipSuccess)
        g_lds_budget = (size_t)budget;
}

}  // namespace

size_t fmg_lds_budget()
{
    std::call_once(g_lds_once, init_lds_budget);
    return g_lds_budget;
}

template <typename T>
size_t tile_lds_bytes(int sweeps, TileShape tile)
{
    return (size_t)(tile.x + 2 * sweeps) * (tile.y + 2 * sweeps) * 2 * sizeof(T);
}

// Pixels of one colour in a tile with its halo: each thread owns at most
// kTilePix of them.
bool tile_fits_threads(int sweeps, TileShape tile)
{
    return (size_t)(tile.x + 2 * sweeps) / 2 * (tile.y + 2 * sweeps) <=
           (size_t)kTilePix * kTileThreads;
}

namespace {

struct TileKey {
    int nx, ny, elem;
    bool operator<(const TileKey &o) const
    {
        return nx != o.nx ? nx < o.nx : (ny != o.ny ? ny < o.ny : elem < o.elem);
    }
};
std::mutex g_tile_mutex;
std::map<TileKey, TileShape> g_tile_table;

}  // namespace

TileShape relax_tile_shape(int nx, int ny, int elem)
{
    std::lock_guard<std::mutex> lock(g_tile_mutex);
    auto it = g_tile_table.find(TileKey{nx, ny, elem});
    return it == g_tile_table.end() ? kTileDefault : it->second;
}

int set_relax_tile_shape(int nx, int ny, int elem, int tx, int ty)
{
    if (nx < 1 || ny < 1 || (elem != 4 && elem != 8))
        return fail(kInvalidArgument, "relaxation tile shapes are set per level size and "
                                      "element size 4 or 8");
    if (tx == 0 && ty == 0) {
        std::lock_guard<std::mutex> lock(g_tile_mutex);
        g_tile_table.erase(TileKey{nx, ny, elem});
        return kOk;
    }
    if (tx < 2 || (tx & 1) || ty < 1)
        return fail(kInvalidArgument, "tile width must be even and >= 2, height >= 1 (got %d x %d)",
                    tx, ty);
    const size_t need = (size_t)(tx + 12) * (ty + 12) * 2 * elem;
    if (need > fmg_lds_budget())
        return fail(kInvalidArgument, "a %d x %d tile needs %zu bytes of LDS, over the budget",
                    tx, ty, need);
    if (!tile_fits_threads(6, TileShape{tx, ty}))
        return fail(kInvalidArgument, "a %d x %d tile has more pixels of one colour than its "
                                      "threads hold", tx, ty);
    std::lock_guard<std::mutex> lock(g_tile_mutex);
    g_tile_table[TileKey{nx, ny, elem}] = TileShape{tx, ty};
    return kOk;
}

template <typename T>
size_t coarse_lds_bytes(const int *nx, const int *ny, int c0, int L)
{
    size_t px = 0;
    for (int l = c0; l < L; ++l)
        px += (size_t)nx[l] * ny[l];
    // A, b, u of every level, a residual buffer of level c0's size, then the
    // resolved resize taps (four tables per level pair)
    size_t planes = (px * 7 + (size_t)2 * nx[c0] * ny[c0]) * sizeof(T);
    size_t entries = 0;
    for (int l = c0; l + 1 < L; ++l)
        entries += (size_t)nx[l + 1] + ny[l + 1] + nx[l] + ny[l];
    size_t tap = sizeof(TapIdx<T>);
    return (planes + tap - 1) / tap * tap + entries * tap;
}

// Values the reference's operator sequence reads and writes once for the
// coarse levels of one call (per job): level c0 in and out, the pyramids,
// the coarse solve, then per operator: relax_membrane 9·nit, residual and
// restriction 12, prolongation-add 4.5, prolongation 2.5 (per fine pixel).
template <typename T>
static double coarse_reference_values(const CoarseLevels<T> &lv, int nit, int cycles, int init)
{
    auto M = [&](int l) { return (double)lv.nx[l] * lv.ny[l]; };
    const int c0 = lv.c0, last = lv.L - 1;
    double v = 7.0 * M(c0) + 2.0 * M(c0);
    for (int l = c0 + 1; l <= last; ++l)
        v += (init ? 5.0 : 3.0) * (M(l - 1) + M(l));
    auto cycle = [&](int start) {
        double c = 9.0 * nit * M(last);
        for (int l = start; l < last; ++l)
            c += 2.0 * 9.0 * nit * M(l) + 12.0 * M(l) + 4.5 * M(l);
        return c;
    };
    if (init) {
        v += 7.0 * M(last) + 9.0 * nit * M(last);
        for (int level = last - 1; level >= c0; --level)
            v += 2.5 * M(level) + cycles * cycle(level);
    } else {
        v += cycles * cycle(c0);
    }
    return v;
}

// ---------------------------------------------------------------- launchers

template <typename T>
int launch_relax_tile(const T *a, const T *b, const T *uin, T *uout, const T *uc, int cnx,
                      int cny, MembraneWeights w, int nx, int ny, int batch, size_t astride,
                      size_t bstride, size_t cstride, int sweeps, hipStream_t s)
{
    if ((nx & 1) || nx < 2 || sweeps < 0)
        return fail(kInvalidArgument, "tiled relaxation needs an even width, got %d", nx);
    const TileShape tile = relax_tile_shape(nx, ny, (int)sizeof(T));
    size_t lds = tile_lds_bytes<T>(sweeps, tile);
    if (lds > fmg_lds_budget() || !tile_fits_threads(sweeps, tile))
        return fail(kInvalidArgument, "tiled relaxation of %d x %d tiles with %d sweeps does not "
                                      "fit (%zu bytes of LDS)", tile.x, tile.y, sweeps, lds);
    // the prolongation's column sums and tap tables in LDS (exact 2:1 levels,
    // when they fit) after the tile
    int colsum = 0;
    if (uc && nx == 2 * cnx && ny == 2 * cny) {
        const size_t ex = (size_t)tile.x + 2 * sweeps, rows = (size_t)tile.y + 2 * sweeps;
        const size_t extra = colsum_lds_bytes<T>(ex, rows);
        if (lds + extra <= fmg_lds_budget()) {
            lds += extra;
            colsum = 1;
        }
    }
    const int tiles_x = (nx + tile.x - 1) / tile.x, tiles_y = (ny + tile.y - 1) / tile.y;
    dim3 grid(tiles_x * tiles_y, batch);
    // reference bytes: 9 values per pixel per red+black iteration, plus the
    // prolongation u += resize(u_c) (reads u_c and u, writes u)
    const double P = (double)nx * ny;
    TimedLaunch timer(kTimedRelaxTile, s,
                      (9.0 * (sweeps / 2.0) + (uc ? 4.5 : 0.0)) * P * sizeof(T) * batch);
    if (sweeps == 6)  // nit = 3, the reference default
        hipExtLaunchKernelGGL((k_relax_tile<T, 6>), grid, dim3(kTileThreads), lds, s, timer.start(), timer.stop(), 0, a, b, uin, uout,
                           uc, cnx, cny, nx, ny, (T)w.w00, (T)w.w01, (T)w.w10, sweeps, tile.x,
                           tile.y, tiles_x, colsum, astride, bstride, cstride);
    else
        hipExtLaunchKernelGGL((k_relax_tile<T, 0>), grid, dim3(kTileThreads), lds, s, timer.start(), timer.stop(), 0, a, b, uin, uout,
                           uc, cnx, cny, nx, ny, (T)w.w00, (T)w.w01, (T)w.w10, sweeps, tile.x,
                           tile.y, tiles_x, colsum, astride, bstride, cstride);
    CSG_LAUNCH_CHECK();
    timer.done();
    return kOk;
}

template <typename T>
int launch_restrict_residual(const T *a, const T *b, const T *u, MembraneWeights w, int nx,
                             int ny, int cnx, int cny, int batch, size_t astride, size_t bstride,
                             size_t cstride, T *bc, T *uc, hipStream_t s)
{
    dim3 g(grid_for((size_t)cnx * cny * 4), batch);
    // reference bytes: apply_membrane (A, u in, R out: 7), b (2), restriction
    // (R in: 2, b_c out: 0.5) and the zeroed u_c (0.5) per fine pixel
    TimedLaunch timer(kTimedRestrict, s, 12.0 * nx * ny * sizeof(T) * batch);
    hipExtLaunchKernelGGL(k_restrict_residual<T>, g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, a, b, u, nx, ny, cnx, cny,
                       (T)w.w00, (T)w.w01, (T)w.w10, astride, bstride, cstride, bc, uc);
    CSG_LAUNCH_CHECK();
    timer.done();
    return kOk;
}

template <typename T>
size_t coarse_meta_bytes(const int *nx, const int *ny, int L)
{
    size_t entries = 0;
    for (int l = 0; l + 1 < L; ++l)
        entries += (size_t)nx[l + 1] + ny[l + 1] + nx[l] + ny[l];
    return meta_taps_offset<T>() + entries * sizeof(TapIdx<T>);
}

template <typename T>
int launch_fmg_coarse(const CoarseLevels<T> &lv, const T *a, const T *b, T *u, size_t astride,
                      size_t bstride, double id, int nit, int cycles, int init, int build,
                      int batch, size_t lds_bytes, hipStream_t s)
{
    if (lds_bytes > fmg_lds_budget())
        return fail(kInvalidArgument, "coarse multigrid levels need %zu bytes of LDS", lds_bytes);
    if (!lv.meta || (init && !build))
        return fail(kInvalidArgument, "coarse multigrid launch without its kept state");
    TimedLaunch timer(kTimedFmgCoarse, s,
                      coarse_reference_values(lv, nit, cycles, init) * sizeof(T) * batch);
    hipExtLaunchKernelGGL(k_fmg_coarse<T>, dim3(batch), dim3(kCoarseThreads), lds_bytes, s, timer.start(), timer.stop(), 0, lv, a,
                       b, u, astride, bstride, id, nit, cycles, init, build,
                       lv.nx[lv.c0] * lv.ny[lv.c0]);
    CSG_LAUNCH_CHECK();
    timer.done();
    return kOk;
}

#define CSG_INSTANTIATE_FUSED(T)                                                                 \
    template size_t tile_lds_bytes<T>(int, TileShape);                                                      \
    template size_t coarse_lds_bytes<T>(const int *, const int *, int, int);                     \
    template int launch_relax_tile<T>(const T *, const T *, const T *, T *, const T *, int, int, \
                                      MembraneWeights, int, int, int, size_t, size_t, size_t,    \
                                      int, hipStream_t);                                         \
    template int launch_restrict_residual<T>(const T *, const T *, const T *, MembraneWeights,   \
                                             int, int, int, int, int, size_t, size_t, size_t,    \
                                             T *, T *, hipStream_t);                             \
    template size_t coarse_meta_bytes<T>(const int *, const int *, int);                        \
    template int launch_fmg_coarse<T>(const CoarseLevels<T> &, const T *, const T *, T *,       \
                                      size_t, size_t, double, int, int, int, int, int, size_t,   \
                                      hipStream_t);

CSG_INSTANTIATE_FUSED(float)
CSG_INSTANTIATE_FUSED(double)

}  // namespace csg
