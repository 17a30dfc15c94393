// Element-wise and reduction steps of the in-process stencil builder
// (native/cat_surface_stencil_builder.c: the sphere normalisation, the output
// sphere at 0.975 of the mean radius around the bound sums, the resampled
// sphere and the unit points), each a kernel instead of a chain of small
// tensor operations around the closest-triangle search.
#include "../../include/csg.h"
#include "csg_common.h"

namespace csg {
namespace {

constexpr int kSphereThreads = 1024;

// Sum, minimum and maximum over a workgroup of kSphereThreads threads, in a
// fixed order (wave shuffles, then the waves in order by thread 0).
struct Red {
    double *sum, *lo, *hi;  // [kSphereThreads / 64] each
};

__device__ void reduce3(double s, double lo, double hi, Red r, double &so, double &loo,
                        double &hio)
{
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off, 64);
        lo = fmin(lo, __shfl_down(lo, off, 64));
        hi = fmax(hi, __shfl_down(hi, off, 64));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        r.sum[wave] = s;
        r.lo[wave] = lo;
        r.hi[wave] = hi;
    }
    __syncthreads();
    __shared__ double out[3];
    if (threadIdx.x == 0) {
        double a = 0.0, b = INFINITY, c = -INFINITY;
        for (int w = 0; w < kSphereThreads / 64; ++w) {
            a += r.sum[w];
            b = fmin(b, r.lo[w]);
            c = fmax(c, r.hi[w]);
        }
        out[0] = a;
        out[1] = b;
        out[2] = c;
    }
    __syncthreads();
    so = out[0];
    loo = out[1];
    hio = out[2];
    __syncthreads();
}

// One workgroup: p = f64(v) [- mean, then / |p| when normalise]; params =
// (lo x, y, z, hi x, y, z, 0.975 · mean |p|).
__global__ void __launch_bounds__(kSphereThreads)
k_stencil_sphere(const float *__restrict__ v, int n, int normalise, double *__restrict__ pts,
                 double *__restrict__ params)
{
    __shared__ double rs[kSphereThreads / 64], rl[kSphereThreads / 64], rh[kSphereThreads / 64];
    const Red red{rs, rl, rh};
    double c[3] = {0.0, 0.0, 0.0};
    if (normalise) {
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
            for (int i = threadIdx.x; i < n; i += kSphereThreads)
                s += (double)v[3 * i + d];
            double t, lo, hi;
            reduce3(s, 0.0, 0.0, red, t, lo, hi);
            c[d] = t / (double)n;
        }
    }
    double len = 0.0, lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY,
                                                                       -INFINITY};
    for (int i = threadIdx.x; i < n; i += kSphereThreads) {
        double p[3];
        for (int d = 0; d < 3; ++d)
            p[d] = (double)v[3 * i + d] - c[d];
        if (normalise) {
            const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            for (int d = 0; d < 3; ++d)
                p[d] = p[d] / r;
        }
        for (int d = 0; d < 3; ++d) {
            pts[3 * i + d] = p[d];
            lo[d] = fmin(lo[d], p[d]);
            hi[d] = fmax(hi[d], p[d]);
        }
        len += sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    }
    double total, l0, h0, l1, h1, l2, h2, zero;
    reduce3(len, lo[0], hi[0], red, total, l0, h0);
    reduce3(0.0, lo[1], hi[1], red, zero, l1, h1);
    reduce3(0.0, lo[2], hi[2], red, zero, l2, h2);
    if (threadIdx.x == 0) {
        params[0] = l0;
        params[1] = l1;
        params[2] = l2;
        params[3] = h0;
        params[4] = h1;
        params[5] = h2;
        params[6] = 0.975 * (total / (double)n);
    }
}

// The same over many workgroups: each block reduces a fixed chunk of
// kSphereChunk vertices (so the partials do not depend on the launch) and one
// block adds the chunks' partials in order.
constexpr int kSphereChunk = 1024;
constexpr int kSphereBlock = 256;

__device__ __forceinline__ void block_reduce3(double &s, double &lo, double &hi, double *ws,
                                              double *wl, double *wh)
{
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off, 64);
        lo = fmin(lo, __shfl_down(lo, off, 64));
        hi = fmax(hi, __shfl_down(hi, off, 64));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        ws[wave] = s;
        wl[wave] = lo;
        wh[wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = INFINITY, c = -INFINITY;
        for (int w = 0; w < kSphereBlock / 64; ++w) {
            a += ws[w];
            b = fmin(b, wl[w]);
            c = fmax(c, wh[w]);
        }
        s = a;
        lo = b;
        hi = c;
    }
}

// part[b][0..2] = the coordinate sums of chunk b.
__global__ void __launch_bounds__(kSphereBlock)
k_sphere_sums(const float *__restrict__ v, int n, double *__restrict__ part)
{
    __shared__ double ws[kSphereBlock / 64], wl[kSphereBlock / 64], wh[kSphereBlock / 64];
    const int i0 = blockIdx.x * kSphereChunk, i1 = min(n, i0 + kSphereChunk);
    for (int d = 0; d < 3; ++d) {
        double s = 0.0, lo = 0.0, hi = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += kSphereBlock)
            s += (double)v[3 * i + d];
        block_reduce3(s, lo, hi, ws, wl, wh);
        if (threadIdx.x == 0)
            part[3 * blockIdx.x + d] = s;
        __syncthreads();
    }
}

// centre[d] = (sum of the chunks' sums in order) / n.
__global__ void k_sphere_centre(const double *__restrict__ part, int nb, int n,
                                double *__restrict__ centre)
{
    if (threadIdx.x >= 3 || blockIdx.x != 0)
        return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b)
        s += part[3 * b + threadIdx.x];
    centre[threadIdx.x] = s / (double)n;
}

// pts of chunk b (centred and normalised when centre is given); part2[b] =
// (sum of |p|, lo x, y, z, hi x, y, z).
__global__ void __launch_bounds__(kSphereBlock)
k_sphere_points(const float *__restrict__ v, int n, const double *__restrict__ centre,
                double *__restrict__ pts, double *__restrict__ part2)
{
    __shared__ double ws[kSphereBlock / 64], wl[kSphereBlock / 64], wh[kSphereBlock / 64];
    const int i0 = blockIdx.x * kSphereChunk, i1 = min(n, i0 + kSphereChunk);
    double c[3] = {0.0, 0.0, 0.0};
    if (centre)
        for (int d = 0; d < 3; ++d)
            c[d] = centre[d];
    double len = 0.0, lo[3] = {INFINITY, INFINITY, INFINITY},
           hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = i0 + threadIdx.x; i < i1; i += kSphereBlock) {
        double p[3];
        for (int d = 0; d < 3; ++d)
            p[d] = (double)v[3 * i + d] - c[d];
        if (centre) {
            const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            for (int d = 0; d < 3; ++d)
                p[d] = p[d] / r;
        }
        for (int d = 0; d < 3; ++d) {
            pts[3 * i + d] = p[d];
            lo[d] = fmin(lo[d], p[d]);
            hi[d] = fmax(hi[d], p[d]);
        }
        len += sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    }
    for (int d = 0; d < 3; ++d) {
        double s = d == 0 ? len : 0.0;
        block_reduce3(s, lo[d], hi[d], ws, wl, wh);
        if (threadIdx.x == 0) {
            if (d == 0)
                part2[7 * blockIdx.x] = s;
            part2[7 * blockIdx.x + 1 + d] = lo[d];
            part2[7 * blockIdx.x + 4 + d] = hi[d];
        }
        __syncthreads();
    }
}

// params = (lo x, y, z, hi x, y, z, 0.975 · mean |p|) from the chunks in order.
__global__ void k_sphere_params(const double *__restrict__ part2, int nb, int n,
                                double *__restrict__ params)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double len = 0.0, lo[3] = {INFINITY, INFINITY, INFINITY},
           hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nb; ++b) {
        len += part2[7 * b];
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], part2[7 * b + 1 + d]);
            hi[d] = fmax(hi[d], part2[7 * b + 4 + d]);
        }
    }
    for (int d = 0; d < 3; ++d) {
        params[d] = lo[d];
        params[3 + d] = hi[d];
    }
    params[6] = 0.975 * (len / (double)n);
}

// out[j] = (lo + hi) + radius · ico[j] (the helper passes the bound sums as the centre).
__global__ void __launch_bounds__(kBlock)
k_out_sphere(const double *__restrict__ ico, int m, const double *__restrict__ params,
             double *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m)
        return;
    const double r = params[6];
    for (int d = 0; d < 3; ++d)
        out[3 * j + d] = (params[d] + params[3 + d]) + r * ico[3 * j + d];
}

// res[j] = Σ_k w[j][k] · pts[idx[j][k]] (k in order), unit[j] = res[j] / |res[j]|.
__global__ void __launch_bounds__(kBlock)
k_resample_unit(const double *__restrict__ pts, const int *__restrict__ idx,
                const double *__restrict__ w, int m, double *__restrict__ res,
                double *__restrict__ unit)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m)
        return;
    double r[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k) {
        const double *p = pts + 3 * idx[3 * j + k];
        const double wk = w[3 * j + k];
        for (int d = 0; d < 3; ++d)
            r[d] += p[d] * wk;
    }
    const double len = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int d = 0; d < 3; ++d) {
        res[3 * j + d] = r[d];
        unit[3 * j + d] = r[d] / len;
    }
}

// out = f32(p / |p|).
__global__ void __launch_bounds__(kBlock)
k_unit_points64(const double *__restrict__ pts, int n, float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double *p = pts + 3 * i;
    const double len = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    for (int d = 0; d < 3; ++d)
        out[3 * i + d] = (float)(p[d] / len);
}

inline hipStream_t S(void *s) { return (hipStream_t)s; }

}  // namespace
}  // namespace csg

using namespace csg;

extern "C" {

int csg_stencil_sphere(const float *v, int n, int normalise, const double *ico, int m,
                       double *pts, double *out_sphere, double *params, void *stream)
{
    if (n < 1 || m < 1 || !v || !ico || !pts || !out_sphere || !params)
        return fail(kInvalidArgument, "invalid stencil sphere arguments");
    const int nb = (n + kSphereChunk - 1) / kSphereChunk;
    if ((size_t)10 * nb + 3 <= (size_t)3 * m) {
        // the chunk partials live in out_sphere until k_out_sphere overwrites it
        double *part = out_sphere, *centre = out_sphere + 3 * (size_t)nb,
               *part2 = centre + 3;
        if (normalise) {
            hipLaunchKernelGGL(k_sphere_sums, dim3(nb), dim3(kSphereBlock), 0, S(stream), v, n,
                               part);
            CSG_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_sphere_centre, dim3(1), dim3(64), 0, S(stream), part, nb, n,
                               centre);
            CSG_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_sphere_points, dim3(nb), dim3(kSphereBlock), 0, S(stream), v, n,
                           normalise ? centre : nullptr, pts, part2);
        CSG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sphere_params, dim3(1), dim3(64), 0, S(stream), part2, nb, n,
                           params);
        CSG_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(k_stencil_sphere, dim3(1), dim3(kSphereThreads), 0, S(stream), v, n,
                           normalise, pts, params);
        CSG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_out_sphere, dim3(grid_for(m)), dim3(kBlock), 0, S(stream), ico, m,
                       params, out_sphere);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int csg_resample_unit(const double *pts, const int *idx, const double *w, int m, double *res,
                      double *unit, void *stream)
{
    if (m < 0 || (m > 0 && (!pts || !idx || !w || !res || !unit)))
        return fail(kInvalidArgument, "invalid resample arguments");
    if (m == 0)
        return kOk;
    hipLaunchKernelGGL(k_resample_unit, dim3(grid_for(m)), dim3(kBlock), 0, S(stream), pts, idx,
                       w, m, res, unit);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int csg_unit_points64(const double *pts, int n, float *out, void *stream)
{
    if (n < 0 || (n > 0 && (!pts || !out)))
        return fail(kInvalidArgument, "invalid unit point arguments");
    if (n == 0)
        return kOk;
    hipLaunchKernelGGL(k_unit_points64, dim3(grid_for(n)), dim3(kBlock), 0, S(stream), pts, n,
                       out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

}  // extern "C"
