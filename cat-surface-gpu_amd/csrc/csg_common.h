// Shared device helpers and host error plumbing for libcsg_hip (gfx950).
//
// Grid conventions follow the reference DARTEL operators
// (src/cat_surface_gpu/dartel_grid.py): planes are channel-major
// [C][ny][nx], coordinate maps are one-based, x is circular and y is
// reflected with period 2*ny.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>

namespace csg {

// ---------------------------------------------------------------- errors

enum Status : int {
    kOk = 0,
    kInvalidArgument = -1,   // mapped to ValueError on the Python side
    kRuntimeError = -2,      // mapped to RuntimeError
    kUnsupported = -3,       // mapped to NotImplementedError
};

void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define CSG_HIP_CHECK(expr)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return ::csg::fail(::csg::kRuntimeError, "%s failed: %s (%s:%d)", #expr, \
                               hipGetErrorString(_e), __FILE__, __LINE__);         \
    } while (0)

#define CSG_LAUNCH_CHECK() CSG_HIP_CHECK(hipGetLastError())

// Propagate a non-zero status code of a library call.
#define CSG_TRY(expr)                      \
    do {                                   \
        int _rc = (expr);                  \
        if (_rc != ::csg::kOk)             \
            return _rc;                    \
    } while (0)

inline int dtype_size(int dtype) { return dtype == 1 ? 8 : 4; }

// ---------------------------------------------------------------- grid maths

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n, int block = kBlock)
{
    return (unsigned)((n + (size_t)block - 1) / (size_t)block);
}

// x circular.
// x / d for a divisor whose reciprocal y = 1 / d (correctly rounded) was taken
// in advance: q = RN(x·y) is within one ulp of x / d, r = x - q·d is exact
// (FMA), and RN(q + r·y) is the correctly rounded quotient (Markstein's
// theorem; IEEE binary64, round to nearest, no underflow or overflow): the
// bits of x / d for three FP64 operations instead of the ~15 of a division.
__device__ __forceinline__ double div_known(double x, double d, double y)
{
    const double q = x * y;
    return fma(fma(-q, d, x), y, q);
}

// The calling workgroup's (x, y) block with the blocks relabelled XCD by XCD.
// The hardware places workgroup i (flattened x first, then y) on XCD i mod 8;
// here XCD k takes the k-th contiguous share of the flattened blocks, so that
// neighbouring blocks, whose bilinear gathers and stencil reads overlap, share
// that XCD's L2.  A bijection: every block's work and results are unchanged.
struct Block2 {
    int x, y;
};

__device__ __forceinline__ Block2 xcd_block()
{
    const unsigned gx = gridDim.x, total = gx * gridDim.y;
    const unsigned l = blockIdx.x + gx * blockIdx.y;
    const unsigned k = l & 7u, j = l >> 3, per = total >> 3, extra = total & 7u;
    const unsigned m = k * per + (k < extra ? k : extra) + j;
    return {(int)(m % gx), (int)(m / gx)};
}

__device__ __forceinline__ int wrap_x(int ix, int nx)
{
    if ((unsigned)ix < (unsigned)nx)
        return ix;
    int r = ix % nx;
    return r < 0 ? r + nx : r;
}

// y reflected with period 2*ny; ny == 1 collapses to row 0.
__device__ __forceinline__ int reflect_y(int iy, int ny)
{
    if ((unsigned)iy < (unsigned)ny)
        return iy;
    if (ny == 1)
        return 0;
    int period = 2 * ny;
    int r = iy % period;
    if (r < 0)
        r += period;
    return r >= ny ? period - r - 1 : r;
}

template <typename T>
struct Corners {
    int i22, i12, i21, i11;
    T dx1, dx2, dy1, dy2;
};

// Bilinear corners at zero-based (x, y): k22 (ix,iy), k12 (ix+1,iy),
// k21 (ix,iy+1), k11 (ix+1,iy+1).  dartel_grid.py:286-343
template <typename T>
__device__ __forceinline__ Corners<T> corners_at(T x, T y, int nx, int ny)
{
    Corners<T> c;
    T fx = floor(x), fy = floor(y);
    int ix = (int)fx, iy = (int)fy;
    c.dx1 = x - (T)ix;
    c.dy1 = y - (T)iy;
    c.dx2 = (T)1 - c.dx1;
    c.dy2 = (T)1 - c.dy1;
    int x0 = wrap_x(ix, nx), x1 = wrap_x(ix + 1, nx);
    int y0 = reflect_y(iy, ny), y1 = reflect_y(iy + 1, ny);
    c.i22 = y0 * nx + x0;
    c.i12 = y0 * nx + x1;
    c.i21 = y1 * nx + x0;
    c.i11 = y1 * nx + x1;
    return c;
}

template <typename T>
__device__ __forceinline__ T bilinear(const T *__restrict__ f, const Corners<T> &c)
{
    return (f[c.i22] * c.dx2 + f[c.i12] * c.dx1) * c.dy2 +
           (f[c.i21] * c.dx2 + f[c.i11] * c.dx1) * c.dy1;
}

// Coordinate sample with the corners unwrapped against k22 by `period`.
// dartel_grid.py:429-455
template <typename T>
__device__ __forceinline__ T wrapped_coordinate(const T *__restrict__ b, const Corners<T> &c,
                                                T period)
{
    T k22 = b[c.i22] - (T)1;
    T k12 = b[c.i12] - (T)1;
    T k21 = b[c.i21] - (T)1;
    T k11 = b[c.i11] - (T)1;
    k12 = k12 - floor((k12 - k22) / period + (T)0.5) * period;
    k21 = k21 - floor((k21 - k22) / period + (T)0.5) * period;
    k11 = k11 - floor((k11 - k22) / period + (T)0.5) * period;
    return (k22 * c.dx2 + k12 * c.dx1) * c.dy2 + (k21 * c.dx2 + k11 * c.dx1) * c.dy1 + (T)1;
}

// ---- displacement form (float32 internal maps)
//
// In float32 the internal maps hold d = φ - id instead of the one-based
// coordinate φ: a coordinate near 512 carries a float32 rounding of ~3e-5
// pixels, the displacement itself ~1e-7.  The arithmetic below is the
// reference's coordinate arithmetic rewritten exactly (in real numbers) with
// the integer parts kept as integers.

// Corners at zero-based position (px + dx, py + dy); also returns the grid
// columns x0/x1 and rows y0/y1 of the corners (after wrap / reflection).
template <typename T>
__device__ __forceinline__ Corners<T> corners_disp(int px, int py, T dx, T dy, int nx, int ny,
                                                   int &x0, int &x1, int &y0, int &y1)
{
    Corners<T> c;
    T fx = floor(dx), fy = floor(dy);
    int ix = px + (int)fx, iy = py + (int)fy;
    c.dx1 = dx - fx;
    c.dy1 = dy - fy;
    c.dx2 = (T)1 - c.dx1;
    c.dy2 = (T)1 - c.dy1;
    x0 = wrap_x(ix, nx);
    x1 = wrap_x(ix + 1, nx);
    y0 = reflect_y(iy, ny);
    y1 = reflect_y(iy + 1, ny);
    c.i22 = y0 * nx + x0;
    c.i12 = y0 * nx + x1;
    c.i21 = y1 * nx + x0;
    c.i11 = y1 * nx + x1;
    return c;
}

// Displacement of the composed map for one component: corner values are
// X_c + d_c (X_c the corner's grid index), unwrapped against k22 by `period`
// exactly as wrapped_coordinate does, interpolated, minus the pixel index p.
template <typename T>
__device__ __forceinline__ T wrapped_displacement(const T *__restrict__ d, const Corners<T> &c,
                                                  int X22, int X12, int X21, int X11, int period,
                                                  int p)
{
    T d22 = d[c.i22], d12 = d[c.i12], d21 = d[c.i21], d11 = d[c.i11];
    T P = (T)period;
    int n12 = (int)floor(((T)(X12 - X22) + (d12 - d22)) / P + (T)0.5);
    int n21 = (int)floor(((T)(X21 - X22) + (d21 - d22)) / P + (T)0.5);
    int n11 = (int)floor(((T)(X11 - X22) + (d11 - d22)) / P + (T)0.5);
    T i22 = (T)(X22 - p), i12 = (T)(X12 - n12 * period - p);
    T i21 = (T)(X21 - n21 * period - p), i11 = (T)(X11 - n11 * period - p);
    T base = (i22 * c.dx2 + i12 * c.dx1) * c.dy2 + (i21 * c.dx2 + i11 * c.dx1) * c.dy1;
    T disp = (d22 * c.dx2 + d12 * c.dx1) * c.dy2 + (d21 * c.dx2 + d11 * c.dx1) * c.dy1;
    return base + disp;
}

// Deterministic block reduction of doubles (blockDim.x == kBlock, wave64).
__device__ __forceinline__ double block_sum(double v, double *smem)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        smem[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            r += smem[w];
    }
    __syncthreads();
    return r;
}

}  // namespace csg
