// Per-pixel arithmetic of the membrane multigrid, shared by every kernel that
// relaxes, applies or resizes (one-launch-per-sweep kernels, the LDS tiles of
// the fine levels and the single-workgroup coarse-level kernel).  Keeping one
// definition makes all schedules produce bit-identical values.
//
// Reference semantics: src/cat_surface_gpu/dartel_grid.py:229-305 (resize),
// :1018-1070 (membrane system), :1073-1208 (relaxation).
#pragma once

#include "csg_common.h"

#include <type_traits>

namespace csg {

// Quadratic B-spline weight.  dartel_grid.py:229-236
template <typename T>
__device__ __forceinline__ T wt2(T d)
{
    T a = fabs(d);
    if (a < (T)0.5)
        return (T)0.75 - a * a;
    if (a < (T)1.5) {
        T o = (T)1.5 - a;
        return (T)0.5 * o * o;
    }
    return (T)0;
}

// The three source taps of target index t when a dimension is resized from S
// to Tn samples: loc = (t + 0.5)·S/Tn - 0.5, origin o = floor(loc + 0.5),
// weights of o-1, o, o+1.
template <typename T>
struct Taps {
    int o;
    T wm, w0, wp;
};

template <typename T>
__device__ __forceinline__ Taps<T> resize_taps(int t, int S, int Tn)
{
    T l = ((T)t + (T)0.5) * (T)S / (T)Tn - (T)0.5;
    int o = (int)floor(l + (T)0.5);
    return {o, wt2<T>((T)o - (T)1 - l), wt2<T>((T)o - l), wt2<T>((T)o + (T)1 - l)};
}

// Source indices (wrapped / reflected) and weights of one target index.
template <typename T>
struct TapIdx {
    int im, i0, ip;
    T wm, w0, wp;
};

template <typename T>
__device__ __forceinline__ TapIdx<T> taps_x(int t, int S, int Tn)
{
    const Taps<T> k = resize_taps<T>(t, S, Tn);
    return {wrap_x(k.o - 1, S), wrap_x(k.o, S), wrap_x(k.o + 1, S), k.wm, k.w0, k.wp};
}

template <typename T>
__device__ __forceinline__ TapIdx<T> taps_y(int t, int S, int Tn)
{
    const Taps<T> k = resize_taps<T>(t, S, Tn);
    return {reflect_y(k.o - 1, S), reflect_y(k.o, S), reflect_y(k.o + 1, S), k.wm, k.w0, k.wp};
}

// Separable resize at one target pixel from resolved taps: the y pass is
// evaluated first for each of the three x taps, then the x pass, as in the
// reference's two-stage evaluation.  f(x, y) returns the source value of one
// channel at source pixel (x, y).
template <typename T, typename F>
__device__ __forceinline__ T resize_from(const TapIdx<T> &tx, const TapIdx<T> &ty, F f)
{
    T im = ty.wm * f(tx.im, ty.im) + ty.w0 * f(tx.im, ty.i0) + ty.wp * f(tx.im, ty.ip);
    T i0 = ty.wm * f(tx.i0, ty.im) + ty.w0 * f(tx.i0, ty.i0) + ty.wp * f(tx.i0, ty.ip);
    T ip = ty.wm * f(tx.ip, ty.im) + ty.w0 * f(tx.ip, ty.i0) + ty.wp * f(tx.ip, ty.ip);
    return tx.wm * im + tx.w0 * i0 + tx.wp * ip;
}

// The same from unresolved taps (indices wrapped and reflected here).
template <typename T, typename F>
__device__ __forceinline__ T resize_at(const Taps<T> &tx, const Taps<T> &ty, int snx, int sny,
                                       F f)
{
    const TapIdx<T> x{wrap_x(tx.o - 1, snx), wrap_x(tx.o, snx), wrap_x(tx.o + 1, snx),
                      tx.wm, tx.w0, tx.wp};
    const TapIdx<T> y{reflect_y(ty.o - 1, sny), reflect_y(ty.o, sny), reflect_y(ty.o + 1, sny),
                      ty.wm, ty.w0, ty.wp};
    return resize_from<T>(x, y, f);
}

// n / d and n % d for 0 <= n < 2^22, d >= 1, from inv = 1.0f / d (computed
// once by the caller): the float quotient is off by at most one and is
// corrected exactly.  Cheaper than an integer division by a run-time value.
struct DivMod {
    int q, r;
};

__device__ __forceinline__ DivMod divmod_small(int n, int d, float inv)
{
    int q = (int)((float)n * inv);
    int r = n - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
    return {q, r};
}

template <typename T>
struct Pair {
    T x, y;
};

// Membrane stencil values around one pixel, per channel: u at the pixel and
// the sums of its two y and two x neighbours.
template <typename T>
struct Stencil {
    T vx, vy;      // u at the pixel
    T syx, sxx;    // channel 0: u[ym] + u[yp], u[xm] + u[xp]
    T syy, sxy;    // channel 1
};

// The pixel's 2x2 block A + w00 (diagonal dx, dy, off-diagonal cr) and its
// regularised determinant; constant over a relax_membrane call.  In float64
// `det` is the determinant and the update divides by it, with the reference's
// operations in the reference's order.  In float32 the block is stored already
// scaled by the reciprocal determinant (one IEEE division when it is formed)
// and the update is written with fused multiply-adds: the kernels that relax
// on LDS tiles are bound by VALU issue, and this is 12 instructions per pixel
// and half-sweep instead of 24.  Every schedule forms and applies the block
// through these two functions, so all of them still produce the same bits.
template <typename T>
struct RelaxCoef {
    T dx, dy, cr, det;
};

template <typename T>
__device__ __forceinline__ RelaxCoef<T> relax_coef(T a0, T a1, T a2, T w00)
{
    const T dx = a0 + w00, dy = a1 + w00, cr = a2;
    const T det = dx * dy * (T)1.0000000001 - cr * cr;
    if constexpr (std::is_same<T, float>::value) {
        const float inv = 1.0f / det;
        return {dx * inv, dy * inv, cr * inv, inv};
    } else {
        return {dx, dy, cr, det};
    }
}

// Gauss-Seidel block update of one pixel from its block.
template <typename T>
__device__ __forceinline__ Pair<T> relax_solve(T bx, T by, const Stencil<T> &n,
                                               const RelaxCoef<T> &k, T w01, T w10)
{
    if constexpr (std::is_same<T, float>::value) {
        const float rx = fmaf(-w10, n.sxx, fmaf(-w01, n.syx, bx));
        const float ry = fmaf(-w10, n.sxy, fmaf(-w01, n.syy, by));
        return {fmaf(k.dy, rx, -(k.cr * ry)), fmaf(k.dx, ry, -(k.cr * rx))};
    } else {
        T rx = bx - w01 * n.syx - w10 * n.sxx;
        T ry = by - w01 * n.syy - w10 * n.sxy;
        return {(k.dy * rx - k.cr * ry) / k.det, (-k.cr * rx + k.dx * ry) / k.det};
    }
}

// Gauss-Seidel block update of one pixel: solve the 2x2 system
// (A + w00) u = b - w01·Σ_y u - w10·Σ_x u.  dartel_triton.py:241-291
template <typename T>
__device__ __forceinline__ Pair<T> relax_pixel(T bx, T by, const Stencil<T> &n, T a0, T a1,
                                               T a2, T w00, T w01, T w10)
{
    return relax_solve<T>(bx, by, n, relax_coef<T>(a0, a1, a2, w00), w01, w10);
}

// (L + A) u at one pixel.  dartel_triton.py:189-239
template <typename T>
__device__ __forceinline__ Pair<T> membrane_pixel(const Stencil<T> &n, T a0, T a1, T a2, T w00,
                                                  T w01, T w10)
{
    T rx = w00 * n.vx + w01 * n.syx + w10 * n.sxx;
    T ry = w00 * n.vy + w01 * n.syy + w10 * n.sxy;
    return {rx + a0 * n.vx + a2 * n.vy, ry + a2 * n.vx + a1 * n.vy};
}

// Planar [2][P] field accessor with the mixed boundary: the stencil of pixel
// (x, y) of an nx × ny grid.  For offsets of one pixel the circular x and the
// reflected y boundaries reduce to a comparison each (nx or ny == 1 included).
template <typename T>
__device__ __forceinline__ Stencil<T> stencil_at(const T *u, int P, int nx, int ny, int x, int y)
{
    const int row = y * nx, p = row + x;
    const int xm = row + (x == 0 ? nx - 1 : x - 1), xp = row + (x == nx - 1 ? 0 : x + 1);
    const int ym = (y == 0 ? 0 : y - 1) * nx + x, yp = (y == ny - 1 ? y : y + 1) * nx + x;
    return {u[p], u[P + p], u[ym] + u[yp], u[xm] + u[xp], u[P + ym] + u[P + yp],
            u[P + xm] + u[P + xp]};
}

}  // namespace csg
