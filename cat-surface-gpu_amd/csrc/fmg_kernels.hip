// Full-multigrid pieces for the membrane system (L_membrane + A) u = b:
// quadratic B-spline resize (restriction and prolongation), red-black block
// Gauss-Seidel, residual, and the 2x2 direct solve at the coarsest level.
//
// Reference semantics: src/cat_surface_gpu/dartel_grid.py:239-305 (resize),
// :1018-1070 (membrane system), :1073-1208 (relaxation), :1211-1417 (FMG).
#include "dartel_kernels.h"
#include "fmg_device.h"
#include "kernel_timing.h"

namespace csg {

// Separable 3-tap resize; one thread per target pixel, all channels.
// Source location loc = (t + 0.5)·S/T - 0.5, origin floor(loc + 0.5).
// The y pass is evaluated first for each of the three x taps, as in the
// reference's two-stage torch evaluation.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_resize(const T *__restrict__ in, int channels, int snx, int sny, int tnx, int tny,
         size_t istride, size_t ostride, T *__restrict__ out, int accumulate)
{
    const int TP = tnx * tny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= TP)
        return;
    in += blockIdx.y * istride;
    out += blockIdx.y * ostride;
    int tx = p % tnx, ty = p / tnx;
    const Taps<T> txp = resize_taps<T>(tx, snx, tnx), typ = resize_taps<T>(ty, sny, tny);
    const int SP = snx * sny;
    for (int c = 0; c < channels; ++c) {
        const T *f = in + (size_t)c * SP;
        T r = resize_at<T>(txp, typ, snx, sny, [&](int x, int y) { return f[y * snx + x]; });
        size_t o = (size_t)c * TP + p;
        out[o] = accumulate ? out[o] + r : r;
    }
}

// Two k_resize problems of one grid geometry in one launch (blockIdx.z picks
// the problem): the level pyramids of A and of the right-hand side.
template <typename T>
struct ResizeTwo {
    const T *in[2];
    T *out[2];
    int channels[2];
    size_t istride[2], ostride[2];
};

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_resize_two(ResizeTwo<T> r, int snx, int sny, int tnx, int tny)
{
    const int TP = tnx * tny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= TP)
        return;
    const int z = blockIdx.z;
    const T *in = r.in[z] + blockIdx.y * r.istride[z];
    T *out = r.out[z] + blockIdx.y * r.ostride[z];
    int tx = p % tnx, ty = p / tnx;
    const Taps<T> txp = resize_taps<T>(tx, snx, tnx), typ = resize_taps<T>(ty, sny, tny);
    const int SP = snx * sny;
    for (int c = 0; c < r.channels[z]; ++c) {
        const T *f = in + (size_t)c * SP;
        out[(size_t)c * TP + p] =
            resize_at<T>(txp, typ, snx, sny, [&](int x, int y) { return f[y * snx + x]; });
    }
}

// One colour of the red-black sweep, in place.  Pixels with (x + y) & 1 ==
// color are updated from their four neighbours (opposite colour for even nx,
// or the pixel itself through the boundaries).  dartel_triton.py:241-291
//
// For odd nx > 1 the reference's CPU path updates row-major in place
// (dartel_grid.py:1139-1185); the only same-colour neighbours are then
// x = 0 and x = nx - 1 of a row whose sweep starts at x = 0, and the
// sequential order updates x = 0 first.  phase 0 updates every active pixel
// except x = nx - 1 of odd-width rows, phase 1 updates those last pixels, so
// the result equals the sequential sweep exactly.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_relax_color(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ u, T w00, T w01,
              T w10, int nx, int ny, size_t astride, size_t bstride, int color, int phase)
{
    const int P = nx * ny;
    int half = (nx + 1) >> 1;
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    int x, y;
    const bool odd = (nx & 1) && nx > 1;
    if (phase == 0) {
        if (q >= half * ny)
            return;
        y = q / half;
        x = 2 * (q - y * half) + ((color + y) & 1);
        if (x >= nx || (odd && x == nx - 1))
            return;
    } else {
        if (q >= ny)
            return;
        y = q;
        x = nx - 1;
        if (((x + y) & 1) != color)
            return;
    }
    a += blockIdx.y * astride;
    b += blockIdx.y * bstride;
    u += blockIdx.y * bstride;
    int p = y * nx + x;
    Pair<T> r = relax_pixel<T>(b[p], b[P + p], stencil_at<T>(u, P, nx, ny, x, y), a[p], a[P + p],
                               a[2 * P + p], w00, w01, w10);
    u[p] = r.x;
    u[P + p] = r.y;
}

// out = (L + A) u, or out = b - (L + A) u when b_minus is given.
// dartel_triton.py:189-239, dartel_grid.py:1356-1368
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_apply_membrane(const T *__restrict__ a, const T *__restrict__ u, T w00, T w01, T w10, int nx,
                 int ny, size_t astride, size_t ustride, T *__restrict__ out,
                 const T *__restrict__ b_minus)
{
    const int P = nx * ny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    a += blockIdx.y * astride;
    u += blockIdx.y * ustride;
    out += blockIdx.y * ustride;
    int x = p % nx, y = p / nx;
    Pair<T> m = membrane_pixel<T>(stencil_at<T>(u, P, nx, ny, x, y), a[p], a[P + p], a[2 * P + p],
                                  w00, w01, w10);
    T ox = m.x, oy = m.y;
    if (b_minus) {
        b_minus += blockIdx.y * ustride;
        ox = b_minus[p] - ox;
        oy = b_minus[P + p] - oy;
    }
    out[p] = ox;
    out[P + p] = oy;
}

// Coarsest level: direct 2x2 block solve in double precision.
// dartel_grid.py:1301-1315
template <typename T>
__global__ void k_coarse_solve(const T *__restrict__ a, const T *__restrict__ b, double id, int P,
                               size_t astride, size_t bstride, T *__restrict__ u)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    a += blockIdx.y * astride;
    b += blockIdx.y * bstride;
    u += blockIdx.y * bstride;
    double dx = (double)a[p] + id, dy = (double)a[P + p] + id, cr = (double)a[2 * P + p];
    double det = dx * dy * 1.0000000001 - cr * cr;
    double bx = (double)b[p], by = (double)b[P + p];
    u[p] = (T)((dy * bx - cr * by) / det);
    u[P + p] = (T)((-cr * bx + dx * by) / det);
}

// flags[b] = any(u[b] != 0): flags are cleared first, then every block that
// finds a non-zero value in its chunk stores 1 (all writers store the same
// value).
__global__ void k_clear_flags(int *__restrict__ flags, int batch)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch)
        flags[b] = 0;
}

constexpr int kNonzeroChunk = kBlock * 8;

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_any_nonzero(const T *__restrict__ u, size_t n, size_t stride, int *__restrict__ flags)
{
    __shared__ int found;
    if (threadIdx.x == 0)
        found = 0;
    __syncthreads();
    u += blockIdx.y * stride;
    const size_t begin = (size_t)blockIdx.x * kNonzeroChunk;
    const size_t end = begin + kNonzeroChunk < n ? begin + kNonzeroChunk : n;
    int local = 0;
    for (size_t i = begin + threadIdx.x; i < end; i += blockDim.x)
        local |= (u[i] != (T)0);
    if (local)
        found = 1;
    __syncthreads();
    if (threadIdx.x == 0 && found)
        flags[blockIdx.y] = 1;
}

template <typename T>
int launch_resize(const T *in, int channels, int snx, int sny, int tnx, int tny, int batch,
                  size_t istride, size_t ostride, T *out, int accumulate, hipStream_t s)
{
    dim3 g(grid_for((size_t)tnx * tny), batch);
    hipLaunchKernelGGL(k_resize<T>, g, dim3(kBlock), 0, s, in, channels, snx, sny, tnx, tny,
                       istride, ostride, out, accumulate);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_resize_two(const T *in0, int channels0, size_t istride0, size_t ostride0, T *out0,
                      const T *in1, int channels1, size_t istride1, size_t ostride1, T *out1,
                      int snx, int sny, int tnx, int tny, int batch, hipStream_t s)
{
    ResizeTwo<T> r{{in0, in1}, {out0, out1}, {channels0, channels1}, {istride0, istride1},
                   {ostride0, ostride1}};
    dim3 g(grid_for((size_t)tnx * tny), batch, 2);
    hipLaunchKernelGGL(k_resize_two<T>, g, dim3(kBlock), 0, s, r, snx, sny, tnx, tny);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_relax_sweep(const T *a, const T *b, T *u, MembraneWeights w, int nx, int ny,
                       int batch, size_t astride, size_t bstride, int color, hipStream_t s)
{
    int half = (nx + 1) / 2;
    dim3 g(grid_for((size_t)half * ny), batch);
    // reference bytes: half of the 9 values per pixel of a red+black iteration
    TimedLaunch timer(kTimedRelaxSweep, s, 4.5 * nx * ny * sizeof(T) * batch);
    hipExtLaunchKernelGGL(k_relax_color<T>, g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, a, b, u, (T)w.w00, (T)w.w01,
                       (T)w.w10, nx, ny, astride, bstride, color, 0);
    CSG_LAUNCH_CHECK();
    if ((nx & 1) && nx > 1) {
        dim3 g1(grid_for((size_t)ny), batch);
        hipExtLaunchKernelGGL(k_relax_color<T>, g1, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, a, b, u, (T)w.w00, (T)w.w01,
                           (T)w.w10, nx, ny, astride, bstride, color, 1);
        CSG_LAUNCH_CHECK();
    }
    timer.done();
    return kOk;
}

template <typename T>
int launch_relax(const T *a, const T *b, T *u, MembraneWeights w, int nx, int ny, int batch,
                 size_t astride, size_t bstride, int nit, hipStream_t s)
{
    for (int it = 0; it < 2 * nit; ++it) {
        int rc = launch_relax_sweep<T>(a, b, u, w, nx, ny, batch, astride, bstride,
                                       (it % 2 == 0) ? 1 : 0, s);
        if (rc)
            return rc;
    }
    return kOk;
}

template <typename T>
int launch_apply_membrane(const T *a, const T *u, MembraneWeights w, int nx, int ny, int batch,
                          size_t astride, size_t ustride, T *out, const T *b_minus,
                          hipStream_t s)
{
    dim3 g(grid_for((size_t)nx * ny), batch);
    hipLaunchKernelGGL(k_apply_membrane<T>, g, dim3(kBlock), 0, s, a, u, (T)w.w00, (T)w.w01,
                       (T)w.w10, nx, ny, astride, ustride, out, b_minus);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_coarse_solve(const T *a, const T *b, double id, int P, int batch, size_t astride,
                        size_t bstride, T *u, hipStream_t s)
{
    dim3 g(grid_for(P, 64), batch);
    hipLaunchKernelGGL(k_coarse_solve<T>, g, dim3(64), 0, s, a, b, id, P, astride, bstride, u);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_any_nonzero(const T *u, size_t n, int batch, size_t stride, int *flags,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_clear_flags, dim3(grid_for((size_t)batch)), dim3(kBlock), 0, s, flags,
                       batch);
    CSG_LAUNCH_CHECK();
    const unsigned chunks = (unsigned)((n + kNonzeroChunk - 1) / kNonzeroChunk);
    hipLaunchKernelGGL(k_any_nonzero<T>, dim3(chunks > 0 ? chunks : 1, batch), dim3(kBlock), 0,
                       s, u, n, stride, flags);
    CSG_LAUNCH_CHECK();
    return kOk;
}

#define CSG_INSTANTIATE_FMG(T)                                                                 \
    template int launch_resize<T>(const T *, int, int, int, int, int, int, size_t, size_t, T *, \
                                  int, hipStream_t);                                           \
    template int launch_resize_two<T>(const T *, int, size_t, size_t, T *, const T *, int,     \
                                      size_t, size_t, T *, int, int, int, int, int,            \
                                      hipStream_t);                                            \
    template int launch_relax_sweep<T>(const T *, const T *, T *, MembraneWeights, int, int,   \
                                       int, size_t, size_t, int, hipStream_t);                 \
    template int launch_relax<T>(const T *, const T *, T *, MembraneWeights, int, int, int,    \
                                 size_t, size_t, int, hipStream_t);                            \
    template int launch_apply_membrane<T>(const T *, const T *, MembraneWeights, int, int,     \
                                          int, size_t, size_t, T *, const T *, hipStream_t);   \
    template int launch_coarse_solve<T>(const T *, const T *, double, int, int, size_t,        \
                                        size_t, T *, hipStream_t);                             \
    template int launch_any_nonzero<T>(const T *, size_t, int, size_t, int *, hipStream_t);

CSG_INSTANTIATE_FMG(float)
CSG_INSTANTIATE_FMG(double)

}  // namespace csg
