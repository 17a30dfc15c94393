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

#include <mutex>

namespace csg {

// ---------------------------------------------------------------- fine levels

namespace {

// One relax_membrane call (`sweeps` = 2·nit half-sweeps, first colour
// (x + y) & 1 == 1) on a kTileX x kTileY interior.  The tile holds columns
// x0 - H .. x0 + kTileX + H - 1 in unwrapped coordinates (nx is even, so a
// pixel's colour is the same unwrapped or wrapped) and the rows of
// y0 - H .. y0 + kTileY + H - 1 that exist (the y boundary reflects onto rows
// inside the grid).  Pixels whose neighbours lie outside the tile are not
// updated; after s sweeps every pixel at distance >= s from the tile edge
// holds the value of the global sweep, so the interior is exact after H.
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
k_relax_tile(const T *__restrict__ A, const T *__restrict__ B, const T *__restrict__ Uin,
             T *__restrict__ Uout, const T *__restrict__ Uc, int cnx, int cny, int nx, int ny,
             T w00, T w01, T w10, int H, int tiles_x, size_t astride, size_t bstride,
             size_t cstride)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    const int P = nx * ny, CP = cnx * cny;
    const int job = blockIdx.y;
    A += job * astride;
    B += job * bstride;
    Uin += job * bstride;
    Uout += job * bstride;
    if (Uc)
        Uc += job * cstride;
    const int tyi = blockIdx.x / tiles_x, txi = blockIdx.x - tyi * tiles_x;
    const int ex = kTileX + 2 * H;
    const int x0 = txi * kTileX - H;
    const int ylo = max(0, tyi * kTileY - H), yhi = min(ny, tyi * kTileY + kTileY + H);
    const int npx = ex * (yhi - ylo);
    T *u0 = lds, *u1 = lds + npx, *a0 = lds + 2 * npx, *a1 = lds + 3 * npx, *a2 = lds + 4 * npx,
      *b0 = lds + 5 * npx, *b1 = lds + 6 * npx;
    for (int q = threadIdx.x; q < npx; q += blockDim.x) {
        const int r = q / ex, i = q - r * ex;
        const int gx = wrap_x(x0 + i, nx), gy = ylo + r, p = gy * nx + gx;
        T ux = Uin[p], uy = Uin[P + p];
        if (Uc) {
            const Taps<T> tx = resize_taps<T>(gx, cnx, nx), ty = resize_taps<T>(gy, cny, ny);
            T rx = resize_at<T>(tx, ty, cnx, cny, [&](int x, int y) { return Uc[y * cnx + x]; });
            T ry = resize_at<T>(tx, ty, cnx, cny,
                                [&](int x, int y) { return Uc[CP + y * cnx + x]; });
            ux = ux + rx;
            uy = uy + ry;
        }
        u0[q] = ux;
        u1[q] = uy;
        a0[q] = A[p];
        a1[q] = A[P + p];
        a2[q] = A[2 * P + p];
        b0[q] = B[p];
        b1[q] = B[P + p];
    }
    __syncthreads();
    for (int s = 0; s < H; ++s) {
        const int colour = (s & 1) ? 0 : 1;
        for (int q = threadIdx.x; q < npx; q += blockDim.x) {
            const int r = q / ex, i = q - r * ex;
            const int gy = ylo + r;
            if (i == 0 || i == ex - 1 || (((x0 + i + gy) & 1) != colour))
                continue;
            const int ym = reflect_y(gy - 1, ny), yp = reflect_y(gy + 1, ny);
            if (ym < ylo || ym >= yhi || yp < ylo || yp >= yhi)
                continue;
            const int qym = (ym - ylo) * ex + i, qyp = (yp - ylo) * ex + i;
            const Stencil<T> n{u0[q],           u1[q],           u0[qym] + u0[qyp],
                               u0[q - 1] + u0[q + 1], u1[qym] + u1[qyp], u1[q - 1] + u1[q + 1]};
            const Pair<T> v = relax_pixel<T>(b0[q], b1[q], n, a0[q], a1[q], a2[q], w00, w01, w10);
            u0[q] = v.x;
            u1[q] = v.y;
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < kTileX * kTileY; q += blockDim.x) {
        const int ly = q / kTileX, lx = q - ly * kTileX;
        const int gx = txi * kTileX + lx, gy = tyi * kTileY + ly;
        if (gx >= nx || gy >= ny)
            continue;
        const int t = (gy - ylo) * ex + H + lx;
        Uout[gy * nx + gx] = u0[t];
        Uout[P + gy * nx + gx] = u1[t];
    }
}

// Residual of the finer level restricted to the coarser one; the coarser
// level's correction starts from zero.  dartel_grid.py:1356-1368
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_restrict_residual(const T *__restrict__ A, const T *__restrict__ B, const T *__restrict__ U,
                    int nx, int ny, int cnx, int cny, T w00, T w01, T w10, size_t astride,
                    size_t bstride, size_t cstride, T *__restrict__ Bc, T *__restrict__ Uc)
{
    const int CP = cnx * cny, P = nx * ny;
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= CP)
        return;
    const int job = blockIdx.y;
    A += job * astride;
    B += job * bstride;
    U += job * bstride;
    Bc += job * cstride;
    Uc += job * cstride;
    const int cy = q / cnx, cx = q - cy * cnx;
    const Taps<T> tx = resize_taps<T>(cx, nx, cnx), ty = resize_taps<T>(cy, ny, cny);
    auto res = [&](int x, int y, int ch) {
        const int p = y * nx + x;
        const Pair<T> m = membrane_pixel<T>(stencil_at<T>(U, P, nx, ny, x, y), A[p], A[P + p],
                                            A[2 * P + p], w00, w01, w10);
        return ch ? B[P + p] - m.y : B[p] - m.x;
    };
    Bc[q] = resize_at<T>(tx, ty, nx, ny, [&](int x, int y) { return res(x, y, 0); });
    Bc[CP + q] = resize_at<T>(tx, ty, nx, ny, [&](int x, int y) { return res(x, y, 1); });
    Uc[q] = (T)0;
    Uc[CP + q] = (T)0;
}

// ---------------------------------------------------------------- coarse levels

// LDS view of one level: A (3 planes), b (2), u (2).
template <typename T>
struct LdsLevel {
    T *a, *b, *u;
    int nx, ny;
    T w00, w01, w10;
    __device__ int P() const { return nx * ny; }
};

// The level table, copied to LDS at kernel start so that levels can be
// indexed at run time without spilling the kernel argument to scratch.
template <typename T>
struct LevelTable {
    int nx[kFmgMaxLevels], ny[kFmgMaxLevels];
    unsigned off[kFmgMaxLevels];
    T w00[kFmgMaxLevels], w01[kFmgMaxLevels], w10[kFmgMaxLevels];
};

template <typename T>
__device__ __forceinline__ LdsLevel<T> lds_level(const LevelTable<T> &t, T *base, int l)
{
    const int P = t.nx[l] * t.ny[l];
    T *a = base + t.off[l];
    return {a, a + 3 * P, a + 5 * P, t.nx[l], t.ny[l], t.w00[l], t.w01[l], t.w10[l]};
}

// relax_membrane on one LDS level (including the two-phase odd-width order).
template <typename T>
__device__ __forceinline__ void lds_relax(const LdsLevel<T> &L, int nit)
{
    const int nx = L.nx, ny = L.ny, P = L.P();
    const int half = (nx + 1) >> 1;
    const bool odd = (nx & 1) && nx > 1;
    auto update = [&](int x, int y) {
        const int p = y * nx + x;
        const Pair<T> v = relax_pixel<T>(L.b[p], L.b[P + p], stencil_at<T>(L.u, P, nx, ny, x, y),
                                         L.a[p], L.a[P + p], L.a[2 * P + p], L.w00, L.w01, L.w10);
        L.u[p] = v.x;
        L.u[P + p] = v.y;
    };
    for (int it = 0; it < 2 * nit; ++it) {
        const int colour = (it % 2 == 0) ? 1 : 0;
        for (int q = threadIdx.x; q < half * ny; q += blockDim.x) {
            const int y = q / half;
            const int x = 2 * (q - y * half) + ((colour + y) & 1);
            if (x >= nx || (odd && x == nx - 1))
                continue;
            update(x, y);
        }
        __syncthreads();
        if (odd) {
            for (int y = threadIdx.x; y < ny; y += blockDim.x)
                if (((nx - 1 + y) & 1) == colour)
                    update(nx - 1, y);
            __syncthreads();
        }
    }
}

// Resize `channels` planes of level `src` into level `dst` (src → dst sizes).
template <typename T>
__device__ __forceinline__ void lds_resize(const T *in, int snx, int sny, T *out, int tnx, int tny, int channels,
                           bool accumulate)
{
    const int SP = snx * sny, TP = tnx * tny;
    for (int p = threadIdx.x; p < TP; p += blockDim.x) {
        const int ty = p / tnx, tx = p - ty * tnx;
        const Taps<T> txp = resize_taps<T>(tx, snx, tnx), typ = resize_taps<T>(ty, sny, tny);
        for (int c = 0; c < channels; ++c) {
            const T *f = in + c * SP;
            T r = resize_at<T>(txp, typ, snx, sny, [&](int x, int y) { return f[y * snx + x]; });
            T *o = out + c * TP + p;
            *o = accumulate ? *o + r : r;
        }
    }
    __syncthreads();
}

// b[l+1] = restrict(b[l] - (L + A) u[l]); u[l+1] = 0.
template <typename T>
__device__ __forceinline__ void lds_down(const LdsLevel<T> &F, const LdsLevel<T> &C)
{
    const int nx = F.nx, ny = F.ny, P = F.P(), CP = C.P();
    for (int q = threadIdx.x; q < CP; q += blockDim.x) {
        const int cy = q / C.nx, cx = q - cy * C.nx;
        const Taps<T> tx = resize_taps<T>(cx, nx, C.nx), ty = resize_taps<T>(cy, ny, C.ny);
        auto res = [&](int x, int y, int ch) {
            const int p = y * nx + x;
            const Pair<T> m = membrane_pixel<T>(stencil_at<T>(F.u, P, nx, ny, x, y), F.a[p],
                                                F.a[P + p], F.a[2 * P + p], F.w00, F.w01, F.w10);
            return ch ? F.b[P + p] - m.y : F.b[p] - m.x;
        };
        C.b[q] = resize_at<T>(tx, ty, nx, ny, [&](int x, int y) { return res(x, y, 0); });
        C.b[CP + q] = resize_at<T>(tx, ty, nx, ny, [&](int x, int y) { return res(x, y, 1); });
        C.u[q] = (T)0;
        C.u[CP + q] = (T)0;
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void lds_cycle(const LevelTable<T> &lv, int L, T *base, int start, int nit)
{
    const int last = L - 1;
    for (int l = start; l < last; ++l) {
        LdsLevel<T> F = lds_level(lv, base, l);
        lds_relax(F, nit);
        lds_down(F, lds_level(lv, base, l + 1));
    }
    lds_relax(lds_level(lv, base, last), nit);
    for (int l = last - 1; l >= start; --l) {
        LdsLevel<T> F = lds_level(lv, base, l), C = lds_level(lv, base, l + 1);
        lds_resize(C.u, C.nx, C.ny, F.u, F.nx, F.ny, 2, true);
        lds_relax(F, nit);
    }
}

template <typename T>
__global__ void __launch_bounds__(kCoarseThreads)
k_fmg_coarse(CoarseLevels<T> lv, const T *__restrict__ A, const T *__restrict__ B,
             T *__restrict__ U, size_t astride, size_t bstride, double id, int nit, int cycles,
             int init)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ LevelTable<T> tab;
    T *base = reinterpret_cast<T *>(lds_raw);
    const int job = blockIdx.x;
    A += job * astride;
    B += job * bstride;
    U += job * bstride;
    const int c0 = lv.c0, L = lv.L, last = L - 1;
    if (threadIdx.x == 0) {
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
                tab.off[l] = off;
                off += 7u * (unsigned)(lv.nx[l] * lv.ny[l]);
            }
        }
    }
    __syncthreads();
    LdsLevel<T> top = lds_level(tab, base, c0);
    const int P0 = top.P();
    for (int p = threadIdx.x; p < P0; p += blockDim.x) {
        top.a[p] = A[p];
        top.a[P0 + p] = A[P0 + p];
        top.a[2 * P0 + p] = A[2 * P0 + p];
        top.b[p] = B[p];
        top.b[P0 + p] = B[P0 + p];
        if (!init) {
            top.u[p] = U[p];
            top.u[P0 + p] = U[P0 + p];
        }
    }
    __syncthreads();
    // pyramids of A (and, for the coarse initialisation, of the original b)
    for (int l = c0 + 1; l <= last; ++l) {
        LdsLevel<T> F = lds_level(tab, base, l - 1), C = lds_level(tab, base, l);
        lds_resize(F.a, F.nx, F.ny, C.a, C.nx, C.ny, 3, false);
        if (init)
            lds_resize(F.b, F.nx, F.ny, C.b, C.nx, C.ny, 2, false);
    }
    if (init) {
        // dartel_grid.py:1301-1326: direct 2x2 solve against the original b of
        // the coarsest level, then relaxation against that level's (zero) b
        LdsLevel<T> C = lds_level(tab, base, last);
        const int PC = C.P();
        for (int p = threadIdx.x; p < PC; p += blockDim.x) {
            double dx = (double)C.a[p] + id, dy = (double)C.a[PC + p] + id,
                   cr = (double)C.a[2 * PC + p];
            double det = dx * dy * 1.0000000001 - cr * cr;
            double bx = (double)C.b[p], by = (double)C.b[PC + p];
            C.u[p] = (T)((dy * bx - cr * by) / det);
            C.u[PC + p] = (T)((-cr * bx + dx * by) / det);
        }
        __syncthreads();
        for (int p = threadIdx.x; p < 2 * PC; p += blockDim.x)
            C.b[p] = (T)0;
        __syncthreads();
        lds_relax(C, nit);
        // dartel_grid.py:1402-1413; b of each level still holds its original
        // pyramid value when the level is reached (only coarser levels have
        // been overwritten by then)
        for (int level = last - 1; level >= c0; --level) {
            LdsLevel<T> F = lds_level(tab, base, level), Cl = lds_level(tab, base, level + 1);
            lds_resize(Cl.u, Cl.nx, Cl.ny, F.u, F.nx, F.ny, 2, false);
            for (int cyc = 0; cyc < cycles; ++cyc)
                lds_cycle(tab, L, base, level, nit);
        }
    } else {
        for (int cyc = 0; cyc < cycles; ++cyc)
            lds_cycle(tab, L, base, c0, nit);
    }
    for (int p = threadIdx.x; p < P0; p += blockDim.x) {
        U[p] = top.u[p];
        U[P0 + p] = top.u[P0 + p];
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
    raise_limit(k_relax_tile<float>, budget);
    raise_limit(k_relax_tile<double>, budget);
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
size_t tile_lds_bytes(int sweeps)
{
    return (size_t)(kTileX + 2 * sweeps) * (kTileY + 2 * sweeps) * 7 * sizeof(T);
}

template <typename T>
size_t coarse_lds_bytes(const int *nx, const int *ny, int c0, int L)
{
    size_t px = 0;
    for (int l = c0; l < L; ++l)
        px += (size_t)nx[l] * ny[l];
    return px * 7 * sizeof(T);
}

// ---------------------------------------------------------------- launchers

template <typename T>
int launch_relax_tile(const T *a, const T *b, const T *uin, T *uout, const T *uc, int cnx,
                      int cny, MembraneWeights w, int nx, int ny, int batch, size_t astride,
                      size_t bstride, size_t cstride, int sweeps, hipStream_t s)
{
    if ((nx & 1) || nx < 2 || sweeps < 0)
        return fail(kInvalidArgument, "tiled relaxation needs an even width, got %d", nx);
    size_t lds = tile_lds_bytes<T>(sweeps);
    if (lds > fmg_lds_budget())
        return fail(kInvalidArgument, "tiled relaxation needs %zu bytes of LDS", lds);
    const int tiles_x = (nx + kTileX - 1) / kTileX, tiles_y = (ny + kTileY - 1) / kTileY;
    hipLaunchKernelGGL(k_relax_tile<T>, dim3(tiles_x * tiles_y, batch), dim3(kTileThreads), lds,
                       s, a, b, uin, uout, uc, cnx, cny, nx, ny, (T)w.w00, (T)w.w01, (T)w.w10,
                       sweeps, tiles_x, astride, bstride, cstride);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_restrict_residual(const T *a, const T *b, const T *u, MembraneWeights w, int nx,
                             int ny, int cnx, int cny, int batch, size_t astride, size_t bstride,
                             size_t cstride, T *bc, T *uc, hipStream_t s)
{
    dim3 g(grid_for((size_t)cnx * cny), batch);
    hipLaunchKernelGGL(k_restrict_residual<T>, g, dim3(kBlock), 0, s, a, b, u, nx, ny, cnx, cny,
                       (T)w.w00, (T)w.w01, (T)w.w10, astride, bstride, cstride, bc, uc);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_fmg_coarse(const CoarseLevels<T> &lv, const T *a, const T *b, T *u, size_t astride,
                      size_t bstride, double id, int nit, int cycles, int init, int batch,
                      size_t lds_bytes, hipStream_t s)
{
    if (lds_bytes > fmg_lds_budget())
        return fail(kInvalidArgument, "coarse multigrid levels need %zu bytes of LDS", lds_bytes);
    hipLaunchKernelGGL(k_fmg_coarse<T>, dim3(batch), dim3(kCoarseThreads), lds_bytes, s, lv, a,
                       b, u, astride, bstride, id, nit, cycles, init);
    CSG_LAUNCH_CHECK();
    return kOk;
}

#define CSG_INSTANTIATE_FUSED(T)                                                                 \
    template size_t tile_lds_bytes<T>(int);                                                      \
    template size_t coarse_lds_bytes<T>(const int *, const int *, int, int);                     \
    template int launch_relax_tile<T>(const T *, const T *, const T *, T *, const T *, int, int, \
                                      MembraneWeights, int, int, int, size_t, size_t, size_t,    \
                                      int, hipStream_t);                                         \
    template int launch_restrict_residual<T>(const T *, const T *, const T *, MembraneWeights,   \
                                             int, int, int, int, int, size_t, size_t, size_t,    \
                                             T *, T *, hipStream_t);                             \
    template int launch_fmg_coarse<T>(const CoarseLevels<T> &, const T *, const T *, T *,       \
                                      size_t, size_t, double, int, int, int, int, size_t,        \
                                      hipStream_t);

CSG_INSTANTIATE_FUSED(float)
CSG_INSTANTIATE_FUSED(double)

}  // namespace csg
