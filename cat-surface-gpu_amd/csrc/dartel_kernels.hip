// DARTEL 2-D grid kernels for gfx950: exponentiation (scaling and squaring),
// objective/gradient/Hessian, the fused squaring update, the membrane
// regulariser and the reductions that feed the step metrics.
//
// Every kernel is one thread per pixel over a 1-D range of the plane, with
// the batch index (independent registrations of one shape) in blockIdx.y.
// Planes are channel-major [C][ny][nx]; a batch element's planes are
// contiguous and elements are `stride` elements apart for each array.
//
// Reference semantics: src/cat_surface_gpu/dartel_grid.py (CPU torch path)
// and src/cat_surface_gpu/dartel_triton.py (fused squaring, lines 344-477).
#include "dartel_kernels.h"
#include "kernel_timing.h"

namespace csg {

// φ0 = id + s·v, J0 = I + s·∇v (central differences).  With D the map holds
// the displacement s·v.  dartel_grid.py:1619-1646, :624-652
template <typename T, bool D>
__global__ void __launch_bounds__(kBlock)
k_expdef_init(const T *__restrict__ v, T scale, int nx, int ny, size_t vstride,
              T *__restrict__ map, T *__restrict__ jac, size_t mstride, size_t jstride)
{
    const int P = nx * ny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    v += blockIdx.y * vstride;
    if (map)
        map += blockIdx.y * mstride;
    jac += blockIdx.y * jstride;
    int x = p % nx, y = p / nx;
    int xm = y * nx + wrap_x(x - 1, nx), xp = y * nx + wrap_x(x + 1, nx);
    int ym = reflect_y(y - 1, ny) * nx + x, yp = reflect_y(y + 1, ny) * nx + x;
    T s = scale * (T)0.5;
    if (map) {
        if (D) {
            map[p] = v[p] * scale;
            map[P + p] = v[P + p] * scale;
        } else {
            map[p] = (T)(x + 1) + v[p] * scale;
            map[P + p] = (T)(y + 1) + v[P + p] * scale;
        }
    }
    jac[p] = (v[xm] * (T)-1 + v[xp]) * s + (T)1;
    jac[P + p] = (v[P + xm] * (T)-1 + v[P + xp]) * s;
    jac[2 * P + p] = (v[ym] * (T)-1 + v[yp]) * s;
    jac[3 * P + p] = (v[P + ym] * (T)-1 + v[P + yp]) * s + (T)1;
}

// Corners of the map at pixel p, in either representation; for D also the
// corner grid indices needed by wrapped_displacement.
template <typename T, bool D>
struct MapCorners {
    Corners<T> c;
    int x0, x1, y0, y1;
};

template <typename T, bool D>
__device__ __forceinline__ MapCorners<T, D> map_corners(const T *__restrict__ map, int p, int P,
                                                        int nx, int ny)
{
    MapCorners<T, D> m;
    if (D) {
        m.c = corners_disp<T>(p % nx, p / nx, map[p], map[P + p], nx, ny, m.x0, m.x1, m.y0,
                              m.y1);
    } else {
        m.c = corners_at<T>(map[p] - (T)1, map[P + p] - (T)1, nx, ny);
    }
    return m;
}

// The composed map's two components at pixel p (coordinates, or displacements for D).
template <typename T, bool D>
__device__ __forceinline__ void composed_map(const T *__restrict__ b, const MapCorners<T, D> &m,
                                             int p, int nx, int ny, T &cx, T &cy)
{
    const int P = nx * ny;
    if (D) {
        cx = wrapped_displacement<T>(b, m.c, m.x0, m.x1, m.x0, m.x1, nx, p % nx);
        cy = wrapped_displacement<T>(b + P, m.c, m.y0, m.y0, m.y1, m.y1, ny, p / nx);
    } else {
        cx = wrapped_coordinate<T>(b, m.c, (T)nx);
        cy = wrapped_coordinate<T>(b + P, m.c, (T)ny);
    }
}

// φ' = φ∘φ and J' = (J∘φ)·J.  dartel_grid.py:458-507
template <typename T, bool D>
__global__ void __launch_bounds__(kBlock)
k_compose_self(const T *__restrict__ map, const T *__restrict__ jac, int nx, int ny,
               size_t mstride, size_t jstride, T *__restrict__ omap, T *__restrict__ ojac)
{
    const int P = nx * ny;
    const Block2 blk = xcd_block();
    int p = blk.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    map += blk.y * mstride;
    omap += blk.y * mstride;
    if (jac) {
        jac += blk.y * jstride;
        ojac += blk.y * jstride;
    }
    MapCorners<T, D> m = map_corners<T, D>(map, p, P, nx, ny);
    const Corners<T> &c = m.c;
    T cx, cy;
    composed_map<T, D>(map, m, p, nx, ny, cx, cy);
    if (jac) {
        T jb00 = bilinear<T>(jac, c), jb01 = bilinear<T>(jac + P, c);
        T jb10 = bilinear<T>(jac + 2 * P, c), jb11 = bilinear<T>(jac + 3 * P, c);
        T a0 = jac[p], a1 = jac[P + p], a2 = jac[2 * P + p], a3 = jac[3 * P + p];
        ojac[p] = jb00 * a0 + jb10 * a1;
        ojac[P + p] = jb01 * a0 + jb11 * a1;
        ojac[2 * P + p] = jb00 * a2 + jb10 * a3;
        ojac[3 * P + p] = jb01 * a2 + jb11 * a3;
    }
    omap[p] = cx;
    omap[P + p] = cy;
}

// General composition c = b∘a (+ Jacobian chain rule).  dartel_grid.py:552-621
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_compose(const T *__restrict__ a, const T *__restrict__ ja, const T *__restrict__ b,
          const T *__restrict__ jb, int nx, int ny, T *__restrict__ omap, T *__restrict__ ojac)
{
    const int P = nx * ny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    Corners<T> c = corners_at<T>(a[p] - (T)1, a[P + p] - (T)1, nx, ny);
    T cx = wrapped_coordinate<T>(b, c, (T)nx);
    T cy = wrapped_coordinate<T>(b + P, c, (T)ny);
    if (ja) {
        T jb00 = bilinear<T>(jb, c), jb01 = bilinear<T>(jb + P, c);
        T jb10 = bilinear<T>(jb + 2 * P, c), jb11 = bilinear<T>(jb + 3 * P, c);
        T a0 = ja[p], a1 = ja[P + p], a2 = ja[2 * P + p], a3 = ja[3 * P + p];
        ojac[p] = jb00 * a0 + jb10 * a1;
        ojac[P + p] = jb01 * a0 + jb11 * a1;
        ojac[2 * P + p] = jb00 * a2 + jb10 * a3;
        ojac[3 * P + p] = jb01 * a2 + jb11 * a3;
    }
    omap[p] = cx;
    omap[P + p] = cy;
}

// Fused: Jd = J·(I + s_div·∇v)^-1 (jac_div_smalldef), the objective terms at
// φ (initialise_objfun), and the small-deformation start of the squaring
// chain, smap = id + s_small·v with its Jacobian (smalldef_jac).
// dartel_grid.py:679-775, :655-676.  Per-block FP64 partials of Σ d²w.
template <typename T, bool D>
__global__ void __launch_bounds__(kBlock)
k_objective(const T *__restrict__ src, const T *__restrict__ tgt, const T *__restrict__ v,
            const T *__restrict__ map, const T *__restrict__ jac, const T *__restrict__ dist,
            T s_div, T s_small, int nx, int ny, size_t istride, size_t vstride, size_t mstride,
            size_t jstride, size_t dstride, T *__restrict__ rhs, T *__restrict__ hes,
            T *__restrict__ smap, T *__restrict__ sjac, size_t rstride, size_t hstride,
            double *__restrict__ partial, int nblocks)
{
    __shared__ double red[kBlock / 64];
    const int P = nx * ny;
    const Block2 blk = xcd_block();
    int p = blk.x * blockDim.x + threadIdx.x;
    const int b = blk.y;
    src += b * istride;
    tgt += b * istride;
    map += b * mstride;
    jac += b * jstride;
    rhs += b * rstride;
    hes += b * hstride;
    if (v) {
        v += b * vstride;
        smap += b * mstride;
        sjac += b * jstride;
    }
    if (dist)
        dist += b * dstride;
    double e = 0.0;
    if (p < P) {
        int x = p % nx, y = p / nx;
        T dvx_dx = (T)0, dvy_dx = (T)0, dvx_dy = (T)0, dvy_dy = (T)0;
        T o00 = jac[p], o10 = jac[P + p], o01 = jac[2 * P + p], o11 = jac[3 * P + p];
        T j0 = o00, j1 = o10, j2 = o01, j3 = o11;
        if (v) {
            int xm = y * nx + wrap_x(x - 1, nx), xp = y * nx + wrap_x(x + 1, nx);
            int ym = reflect_y(y - 1, ny) * nx + x, yp = reflect_y(y + 1, ny) * nx + x;
            dvx_dx = v[xm] * (T)-1 + v[xp];
            dvy_dx = v[P + xm] * (T)-1 + v[P + xp];
            dvx_dy = v[ym] * (T)-1 + v[yp];
            dvy_dy = v[P + ym] * (T)-1 + v[P + yp];
            // jac_div_smalldef
            T sd = s_div * (T)0.5;
            T d00 = dvx_dx * sd + (T)1, d10 = dvy_dx * sd, d01 = dvx_dy * sd,
              d11 = dvy_dy * sd + (T)1;
            T det = d00 * d11 - d01 * d10;
            T i00 = d11 / det, i10 = -d10 / det, i01 = -d01 / det, i11 = d00 / det;
            j0 = o00 * i00 + o01 * i10;
            j1 = o10 * i00 + o11 * i10;
            j2 = o00 * i01 + o01 * i11;
            j3 = o10 * i01 + o11 * i11;
        }
        // initialise_objfun
        MapCorners<T, D> m = map_corners<T, D>(map, p, P, nx, ny);
        const Corners<T> &c = m.c;
        T k22 = src[c.i22], k12 = src[c.i12], k21 = src[c.i21], k11 = src[c.i11];
        T sampled = (k22 * c.dx2 + k12 * c.dx1) * c.dy2 + (k21 * c.dx2 + k11 * c.dx1) * c.dy1;
        T diff = sampled - tgt[p];
        T ddx = (k11 - k21) * c.dy1 + (k12 - k22) * c.dy2;
        T ddy = (k11 * c.dx1 + k21 * c.dx2) - (k12 * c.dx1 + k22 * c.dx2);
        T gx = j0 * ddx + j1 * ddy;
        T gy = j2 * ddx + j3 * ddy;
        T w = dist ? dist[p] : (T)1;
        hes[p] = gx * gx * w;
        hes[P + p] = gy * gy * w;
        hes[2 * P + p] = gx * gy * w;
        rhs[p] = gx * diff * w;
        rhs[P + p] = gy * diff * w;
        e = (double)(diff * diff * w);
        if (v) {
            // smalldef_jac(v, s_small)
            T ss = s_small * (T)0.5;
            if (D) {
                smap[p] = v[p] * s_small;
                smap[P + p] = v[P + p] * s_small;
            } else {
                smap[p] = (T)(x + 1) + v[p] * s_small;
                smap[P + p] = (T)(y + 1) + v[P + p] * s_small;
            }
            sjac[p] = dvx_dx * ss + (T)1;
            sjac[P + p] = dvy_dx * ss;
            sjac[2 * P + p] = dvx_dy * ss;
            sjac[3 * P + p] = dvy_dy * ss + (T)1;
        }
    }
    double s = block_sum(e, red);
    if (threadIdx.x == 0 && partial)
        partial[(size_t)b * nblocks + blk.x] = s;
}

// One squaring step: b += det·Jᵀ(b∘φ), A += det·Jᵀ(A∘φ)J with the increment
// written out; optionally φ'=φ∘φ, J'=(J∘φ)J from the same corners.
// dartel_grid.py:846-913, dartel_triton.py:344-477
template <typename T, bool D>
__global__ void __launch_bounds__(kBlock)
k_squaring(const T *__restrict__ bin, const T *__restrict__ ain, const T *__restrict__ map,
           const T *__restrict__ jac, int nx, int ny, size_t bstride, size_t astride,
           size_t mstride, size_t jstride, T *__restrict__ bout, T *__restrict__ aout,
           T *__restrict__ inc, T *__restrict__ omap, T *__restrict__ ojac)
{
    const int P = nx * ny;
    const Block2 blk = xcd_block();
    int p = blk.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    const int bi = blk.y;
    bin += bi * bstride;
    ain += bi * astride;
    bout += bi * bstride;
    aout += bi * astride;
    map += bi * mstride;
    jac += bi * jstride;
    MapCorners<T, D> m = map_corners<T, D>(map, p, P, nx, ny);
    const Corners<T> &c = m.c;
    T j00 = jac[p], j10 = jac[P + p], j01 = jac[2 * P + p], j11 = jac[3 * P + p];
    T det = j00 * j11 - j01 * j10;
    T sbx = bilinear<T>(bin, c), sby = bilinear<T>(bin + P, c);
    T s00 = bilinear<T>(ain, c), s11 = bilinear<T>(ain + P, c), s01 = bilinear<T>(ain + 2 * P, c);
    T ibx = det * (sbx * j00 + sby * j10);
    T iby = det * (sbx * j01 + sby * j11);
    T t1 = s00 * j00 + s01 * j10;
    T t2 = s01 * j00 + s11 * j10;
    T t3 = s00 * j01 + s01 * j11;
    T t4 = s01 * j01 + s11 * j11;
    bout[p] = bin[p] + ibx;
    bout[P + p] = bin[P + p] + iby;
    aout[p] = ain[p] + det * (t1 * j00 + t2 * j10);
    aout[P + p] = ain[P + p] + det * (t3 * j01 + t4 * j11);
    aout[2 * P + p] = ain[2 * P + p] + det * (t1 * j01 + t2 * j11);
    if (inc) {
        inc += bi * mstride;
        inc[p] = ibx;
        inc[P + p] = iby;
    }
    if (omap) {
        omap += bi * mstride;
        ojac += bi * jstride;
        T cx, cy;
        composed_map<T, D>(map, m, p, nx, ny, cx, cy);
        T jb00 = bilinear<T>(jac, c), jb01 = bilinear<T>(jac + P, c);
        T jb10 = bilinear<T>(jac + 2 * P, c), jb11 = bilinear<T>(jac + 3 * P, c);
        ojac[p] = jb00 * j00 + jb10 * j10;
        ojac[P + p] = jb01 * j00 + jb11 * j10;
        ojac[2 * P + p] = jb00 * j01 + jb10 * j11;
        ojac[3 * P + p] = jb01 * j01 + jb11 * j11;
        omap[p] = cx;
        omap[P + p] = cy;
    }
}

// Membrane regulariser L v (rtype 1) and the combination of the forward and
// reverse chains: rhs = (bf - br)·s + Lv, H = (Af + Ar)·s.  Per-block FP64
// partials of Σ Lv·v and Σ rhs².  dartel_grid.py:917-955, :1581-1595
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_combine(const T *__restrict__ v, const T *__restrict__ bf, const T *__restrict__ af,
          const T *__restrict__ br, const T *__restrict__ ar, T scale, T w00, T w01, T w10,
          int nx, int ny, size_t vstride, size_t bstride, size_t astride, T *__restrict__ rhs,
          T *__restrict__ hes, double *__restrict__ partial_reg, double *__restrict__ partial_norm,
          int nblocks)
{
    __shared__ double red[kBlock / 64];
    const int P = nx * ny;
    const Block2 blk = xcd_block();
    int p = blk.x * blockDim.x + threadIdx.x;
    const int bi = blk.y;
    v += bi * vstride;
    bf += bi * bstride;
    af += bi * astride;
    rhs += bi * bstride;
    hes += bi * astride;
    if (br) {
        br += bi * bstride;
        ar += bi * astride;
    }
    double sreg = 0.0, snorm = 0.0;
    if (p < P) {
        int x = p % nx, y = p / nx;
        int xm = y * nx + wrap_x(x - 1, nx), xp = y * nx + wrap_x(x + 1, nx);
        int ym = reflect_y(y - 1, ny) * nx + x, yp = reflect_y(y + 1, ny) * nx + x;
        T vx = v[p], vy = v[P + p];
        T lx = w00 * vx + w01 * (v[ym] + v[yp]) + w10 * (v[xm] + v[xp]);
        T ly = w00 * vy + w01 * (v[P + ym] + v[P + yp]) + w10 * (v[P + xm] + v[P + xp]);
        T bx = bf[p], by = bf[P + p];
        T a0 = af[p], a1 = af[P + p], a2 = af[2 * P + p];
        if (br) {
            bx = bx - br[p];
            by = by - br[P + p];
            a0 = a0 + ar[p];
            a1 = a1 + ar[P + p];
            a2 = a2 + ar[2 * P + p];
        }
        bx = bx * scale + lx;
        by = by * scale + ly;
        rhs[p] = bx;
        rhs[P + p] = by;
        hes[p] = a0 * scale;
        hes[P + p] = a1 * scale;
        hes[2 * P + p] = a2 * scale;
        sreg = (double)(lx * vx) + (double)(ly * vy);
        snorm = (double)(bx * bx) + (double)(by * by);
    }
    double r1 = block_sum(sreg, red);
    double r2 = block_sum(snorm, red);
    if (threadIdx.x == 0) {
        partial_reg[(size_t)bi * nblocks + blk.x] = r1;
        partial_norm[(size_t)bi * nblocks + blk.x] = r2;
    }
}

// Regulariser alone, rtype 0/1/2.  dartel_grid.py:917-1015
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_regularization(const T *__restrict__ v, int rtype, T sx, T sy, T mu, T lam, T id, int nx,
                 int ny, T *__restrict__ out)
{
    const int P = nx * ny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    int x = p % nx, y = p / nx;
    auto at = [&](const T *f, int dx, int dy) {
        return f[reflect_y(y + dy, ny) * nx + wrap_x(x + dx, nx)];
    };
    T sx2 = sx * sx, sy2 = sy * sy;
    for (int c = 0; c < 2; ++c) {
        const T *f = v + c * P;
        T xm = at(f, -1, 0), xp = at(f, 1, 0), ym = at(f, 0, -1), yp = at(f, 0, 1);
        T r;
        if (rtype == 1) {
            T w00 = mu * ((T)2 * sx2 + (T)2 * sy2) + id;
            T w01 = -mu * sy2;
            T w10 = -mu * sx2;
            r = w00 * f[p] + w01 * (ym + yp) + w10 * (xm + xp);
        } else if (rtype == 0) {
            T wx0 = mu * ((T)2 * sy2 + (T)4 * sx2) + (T)2 * lam * sx2 + id;
            T wy0 = mu * ((T)2 * sx2 + (T)4 * sy2) + (T)2 * lam * sy2 + id;
            T wx1 = -((T)2 * mu + lam) * sx2;
            T wy1 = -((T)2 * mu + lam) * sy2;
            T wx2 = -mu * sx2;
            T wy2 = -mu * sy2;
            T wxy = (T)0.25 * (lam + mu) * sx * sy;
            const T *g = v + (1 - c) * P;
            T cross = at(g, -1, -1) - at(g, -1, 1) - at(g, 1, -1) + at(g, 1, 1);
            if (c == 0)
                r = wx0 * f[p] + wy2 * (ym + yp) + wx1 * (xm + xp) + wxy * cross;
            else
                r = wy0 * f[p] + wy1 * (ym + yp) + wx2 * (xm + xp) + wxy * cross;
        } else {
            T sx4 = sx2 * sx2, sy4 = sy2 * sy2;
            T w00 = mu * ((T)6 * sx4 + (T)6 * sy4 + (T)8 * sx2 * sy2) + id;
            T w01 = mu * ((T)-4 * sx2 * sy2 - (T)4 * sy4);
            T w02 = mu * sy4;
            T w10 = mu * ((T)-4 * sx4 - (T)4 * sx2 * sy2);
            T w11 = mu * ((T)2 * sx2 * sy2);
            T w20 = mu * sx4;
            r = w00 * f[p] + w01 * (ym + yp) + w02 * (at(f, 0, -2) + at(f, 0, 2)) +
                w10 * (xm + xp) +
                w11 * (at(f, -1, -1) + at(f, 1, -1) + at(f, -1, 1) + at(f, 1, 1)) +
                w20 * (at(f, -2, 0) + at(f, 2, 0));
        }
        out[c * P + p] = r;
    }
}

// Final sum of per-block partials: one block per batch element, fixed
// strided order per thread then a fixed-shape tree, so the result is
// deterministic.
__global__ void __launch_bounds__(kBlock)
k_reduce_partials(const double *__restrict__ partial, int nblocks, int batch, double scale,
                  double *__restrict__ out, int out_stride, int accumulate)
{
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += blockDim.x)
        s += partial[(size_t)b * nblocks + i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        s *= scale;
        if (accumulate)
            out[(size_t)b * out_stride] += s;
        else
            out[(size_t)b * out_stride] = s;
    }
}

__global__ void k_sqrt_inplace(double *__restrict__ v, int batch, int stride)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch)
        v[(size_t)b * stride] = sqrt(v[(size_t)b * stride]);
}

// v' = v - u·w (w = distortion[y, 0] or 1).  dartel_grid.py:1610-1614
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_update_velocity(const T *__restrict__ v, const T *__restrict__ u, const T *__restrict__ dist,
                  int nx, int ny, size_t vstride, size_t dstride, T *__restrict__ out)
{
    const int P = nx * ny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    const int bi = blockIdx.y;
    v += bi * vstride;
    u += bi * vstride;
    out += bi * vstride;
    T w = (T)1;
    if (dist)
        w = dist[bi * dstride + (size_t)(p / nx) * nx];
    out[p] = v[p] - u[p] * w;
    out[P + p] = v[P + p] - u[P + p] * w;
}

// J·(I + s·∇v)^-1 alone.  dartel_grid.py:679-720
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_jac_div(const T *__restrict__ jac, const T *__restrict__ v, T scale, int nx, int ny,
          T *__restrict__ out)
{
    const int P = nx * ny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    int x = p % nx, y = p / nx;
    int xm = y * nx + wrap_x(x - 1, nx), xp = y * nx + wrap_x(x + 1, nx);
    int ym = reflect_y(y - 1, ny) * nx + x, yp = reflect_y(y + 1, ny) * nx + x;
    T sd = scale * (T)0.5;
    T d00 = (v[xm] * (T)-1 + v[xp]) * sd + (T)1;
    T d10 = (v[P + xm] * (T)-1 + v[P + xp]) * sd;
    T d01 = (v[ym] * (T)-1 + v[yp]) * sd;
    T d11 = (v[P + ym] * (T)-1 + v[P + yp]) * sd + (T)1;
    T det = d00 * d11 - d01 * d10;
    T i00 = d11 / det, i10 = -d10 / det, i01 = -d01 / det, i11 = d00 / det;
    T o00 = jac[p], o10 = jac[P + p], o01 = jac[2 * P + p], o11 = jac[3 * P + p];
    out[p] = o00 * i00 + o01 * i10;
    out[P + p] = o10 * i00 + o11 * i10;
    out[2 * P + p] = o00 * i01 + o01 * i11;
    out[3 * P + p] = o10 * i01 + o11 * i11;
}

template <typename T>
int launch_jac_div(const T *jac, const T *v, double scale, int nx, int ny, T *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_jac_div<T>, dim3(grid_for((size_t)nx * ny)), dim3(kBlock), 0, s, jac, v,
                       (T)scale, nx, ny, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}
template int launch_jac_div<float>(const float *, const float *, double, int, int, float *,
                                   hipStream_t);
template int launch_jac_div<double>(const double *, const double *, double, int, int, double *,
                                    hipStream_t);

// Bilinear sample of a scalar field at zero-based coordinates.
// dartel_grid.py:510-531
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_sample_field(const T *__restrict__ field, int nx, int ny, const T *__restrict__ x,
               const T *__restrict__ y, size_t n, T *__restrict__ out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Corners<T> c = corners_at<T>(x[i], y[i], nx, ny);
    out[i] = bilinear<T>(field, c);
}

// φ = id + d: one-based coordinates of a displacement-form map.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_disp_to_coord(const T *__restrict__ d, int nx, int ny, size_t stride, T *__restrict__ out)
{
    const int P = nx * ny;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    d += blockIdx.y * stride;
    out += blockIdx.y * stride;
    out[p] = (T)(p % nx + 1) + d[p];
    out[P + p] = (T)(p / nx + 1) + d[P + p];
}

// ------------------------------------------------------------- launchers

template <typename T>
int launch_disp_to_coord(const T *d, int nx, int ny, int batch, size_t stride, T *out,
                         hipStream_t s)
{
    dim3 g(grid_for((size_t)nx * ny), batch);
    hipLaunchKernelGGL(k_disp_to_coord<T>, g, dim3(kBlock), 0, s, d, nx, ny, stride, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_expdef_init(const T *v, double scale, int nx, int ny, int batch, size_t vstride,
                       T *map, T *jac, size_t mstride, size_t jstride, hipStream_t s, bool disp)
{
    dim3 g(grid_for((size_t)nx * ny), batch);
    if (disp)
        hipLaunchKernelGGL((k_expdef_init<T, true>), g, dim3(kBlock), 0, s, v, (T)scale, nx, ny,
                           vstride, map, jac, mstride, jstride);
    else
        hipLaunchKernelGGL((k_expdef_init<T, false>), g, dim3(kBlock), 0, s, v, (T)scale, nx, ny,
                           vstride, map, jac, mstride, jstride);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_compose_self(const T *map, const T *jac, int nx, int ny, int batch, size_t mstride,
                        size_t jstride, T *omap, T *ojac, hipStream_t s, bool disp)
{
    dim3 g(grid_for((size_t)nx * ny), batch);
    // reference bytes: composition with its Jacobian, 12 values per pixel (4 without)
    TimedLaunch timer(kTimedComposeSelf, s,
                      (jac ? 12.0 : 4.0) * nx * ny * sizeof(T) * batch);
    if (disp)
        hipExtLaunchKernelGGL((k_compose_self<T, true>), g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, map, jac, nx, ny,
                           mstride, jstride, omap, ojac);
    else
        hipExtLaunchKernelGGL((k_compose_self<T, false>), g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, map, jac, nx, ny,
                           mstride, jstride, omap, ojac);
    CSG_LAUNCH_CHECK();
    timer.done();
    return kOk;
}

template <typename T>
int launch_compose(const T *a, const T *ja, const T *b, const T *jb, int nx, int ny, T *omap,
                   T *ojac, hipStream_t s)
{
    dim3 g(grid_for((size_t)nx * ny), 1);
    hipLaunchKernelGGL(k_compose<T>, g, dim3(kBlock), 0, s, a, ja, b, jb, nx, ny, omap, ojac);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_objective(const ObjectiveArgs<T> &a, hipStream_t s)
{
    dim3 g(grid_for((size_t)a.nx * a.ny), a.batch);
    // reference bytes: jac_div_smalldef 10, initialise_objfun 13, smalldef_jac 8 values
    TimedLaunch timer(kTimedObjective, s, 31.0 * a.nx * a.ny * sizeof(T) * a.batch);
    if (a.disp)
        hipExtLaunchKernelGGL((k_objective<T, true>), g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, a.src, a.tgt, a.v,
                           a.map, a.jac, a.dist, (T)a.s_div, (T)a.s_small, a.nx, a.ny, a.istride,
                           a.vstride, a.mstride, a.jstride, a.dstride, a.rhs, a.hes, a.smap,
                           a.sjac, a.rstride, a.hstride, a.partial, (int)g.x);
    else
        hipExtLaunchKernelGGL((k_objective<T, false>), g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, a.src, a.tgt, a.v,
                           a.map, a.jac, a.dist, (T)a.s_div, (T)a.s_small, a.nx, a.ny, a.istride,
                           a.vstride, a.mstride, a.jstride, a.dstride, a.rhs, a.hes, a.smap,
                           a.sjac, a.rstride, a.hstride, a.partial, (int)g.x);
    CSG_LAUNCH_CHECK();
    timer.done();
    return kOk;
}

template <typename T>
int launch_squaring(const T *bin, const T *ain, const T *map, const T *jac, int nx, int ny,
                    int batch, size_t bstride, size_t astride, size_t mstride, size_t jstride,
                    T *bout, T *aout, T *inc, T *omap, T *ojac, hipStream_t s, bool disp)
{
    dim3 g(grid_for((size_t)nx * ny), batch);
    // reference bytes: squaring step 18 values per pixel, the composition of phi
    // and J between steps 12 more
    TimedLaunch timer(kTimedSquaring, s,
                      (18.0 + (omap ? 12.0 : 0.0)) * nx * ny * sizeof(T) * batch);
    if (disp)
        hipExtLaunchKernelGGL((k_squaring<T, true>), g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, bin, ain, map, jac, nx,
                           ny, bstride, astride, mstride, jstride, bout, aout, inc, omap, ojac);
    else
        hipExtLaunchKernelGGL((k_squaring<T, false>), g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, bin, ain, map, jac, nx,
                           ny, bstride, astride, mstride, jstride, bout, aout, inc, omap, ojac);
    CSG_LAUNCH_CHECK();
    timer.done();
    return kOk;
}

template <typename T>
int launch_combine(const T *v, const T *bf, const T *af, const T *br, const T *ar, double scale,
                   double w00, double w01, double w10, int nx, int ny, int batch, size_t vstride,
                   size_t bstride, size_t astride, T *rhs, T *hes, double *partial_reg,
                   double *partial_norm, int *nblocks_out, hipStream_t s)
{
    dim3 g(grid_for((size_t)nx * ny), batch);
    // reference bytes: the combination and the regulariser, 17 values per pixel
    TimedLaunch timer(kTimedCombine, s, 17.0 * nx * ny * sizeof(T) * batch);
    hipExtLaunchKernelGGL(k_combine<T>, g, dim3(kBlock), 0, s, timer.start(), timer.stop(), 0, v, bf, af, br, ar, (T)scale, (T)w00,
                       (T)w01, (T)w10, nx, ny, vstride, bstride, astride, rhs, hes, partial_reg,
                       partial_norm, (int)g.x);
    CSG_LAUNCH_CHECK();
    timer.done();
    if (nblocks_out)
        *nblocks_out = (int)g.x;
    return kOk;
}

template <typename T>
int launch_regularization(const T *v, int rtype, const double *params, int nx, int ny, T *out,
                          hipStream_t s)
{
    dim3 g(grid_for((size_t)nx * ny));
    hipLaunchKernelGGL(k_regularization<T>, g, dim3(kBlock), 0, s, v, rtype, (T)params[0],
                       (T)params[1], (T)params[2], (T)params[3], (T)params[4], nx, ny, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_reduce_partials(const double *partial, int nblocks, int batch, double scale,
                           double *out, int out_stride, int accumulate, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(batch), dim3(kBlock), 0, s, partial, nblocks,
                       batch, scale, out, out_stride, accumulate);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_sqrt_inplace(double *v, int batch, int stride, hipStream_t s)
{
    hipLaunchKernelGGL(k_sqrt_inplace, dim3(grid_for(batch, 64)), dim3(64), 0, s, v, batch,
                       stride);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_update_velocity(const T *v, const T *u, const T *dist, int nx, int ny, int batch,
                           size_t vstride, size_t dstride, T *out, hipStream_t s)
{
    dim3 g(grid_for((size_t)nx * ny), batch);
    hipLaunchKernelGGL(k_update_velocity<T>, g, dim3(kBlock), 0, s, v, u, dist, nx, ny, vstride,
                       dstride, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_sample_field(const T *field, int nx, int ny, const T *x, const T *y, size_t n, T *out,
                        hipStream_t s)
{
    if (n == 0)
        return kOk;
    hipLaunchKernelGGL(k_sample_field<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, field, nx, ny,
                       x, y, n, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

#define CSG_INSTANTIATE(T)                                                                     \
    template int launch_expdef_init<T>(const T *, double, int, int, int, size_t, T *, T *,     \
                                       size_t, size_t, hipStream_t, bool);                     \
    template int launch_compose_self<T>(const T *, const T *, int, int, int, size_t, size_t,   \
                                        T *, T *, hipStream_t, bool);                          \
    template int launch_disp_to_coord<T>(const T *, int, int, int, size_t, T *, hipStream_t);  \
    template int launch_compose<T>(const T *, const T *, const T *, const T *, int, int, T *,  \
                                   T *, hipStream_t);                                          \
    template int launch_objective<T>(const ObjectiveArgs<T> &, hipStream_t);                   \
    template int launch_squaring<T>(const T *, const T *, const T *, const T *, int, int, int, \
                                    size_t, size_t, size_t, size_t, T *, T *, T *, T *, T *,   \
                                    hipStream_t, bool);                                        \
    template int launch_combine<T>(const T *, const T *, const T *, const T *, const T *,      \
                                   double, double, double, double, int, int, int, size_t,      \
                                   size_t, size_t, T *, T *, double *, double *, int *,        \
                                   hipStream_t);                                               \
    template int launch_regularization<T>(const T *, int, const double *, int, int, T *,       \
                                          hipStream_t);                                        \
    template int launch_update_velocity<T>(const T *, const T *, const T *, int, int, int,     \
                                           size_t, size_t, T *, hipStream_t);                  \
    template int launch_sample_field<T>(const T *, int, int, const T *, const T *, size_t,     \
                                        T *, hipStream_t);

CSG_INSTANTIATE(float)
CSG_INSTANTIATE(double)

}  // namespace csg
