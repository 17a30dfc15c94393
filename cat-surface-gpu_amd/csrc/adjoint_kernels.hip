// Hand-written adjoint of  E(v) = ½ Σ_p (f(φ(p)) - t(p))² + ½ Σ v·(L v),
// φ = exp(v) by scaling and squaring, on gfx950.  No counterpart in the
// reference, whose DARTEL step uses the push-forward gradient of
// src/cat_surface_gpu/dartel_grid.py:1509-1592 instead; the forward pass is
// the reference's own arithmetic: expdef (:1619-1666, K squarings of
// φ0 = id + v·2^-K), the bilinear sample and half sum of squares of
// initialise_objfun (:723-775) and the rtype-1 regulariser (:917-955) with
// parameters (1, 1, membrane, 0, 0).
//
// Maps are held as displacements d = φ - id (the float32 form of the DARTEL
// kernels, exact in real numbers).  One squaring is d' = d(id + d) composed
// with the reference's corner unwrap; its adjoint, for an upstream λ' on d':
//   λ(p)         += Σ_c λ'_c(p) ∂I_c/∂x   (the sample position p + d(p))
//   λ_c(corner k) += w_k(p) λ'_c(p)        (the four bilinear weights)
// with ∂I_c/∂x from the unwrapped corner differences (the unwrap shifts are
// piecewise constant).  The value term scatters; it is accumulated without
// atomics and in a fixed order: every pixel writes its position term to its
// own λ entry and its four corner terms, tagged with their destination, to a
// list in source order (entry 4p + k); a stable radix sort by destination
// groups them, and each destination adds its group in source order.  The
// gradient is therefore the same bits on every run and in both schedules.
// The SSD adjoint is λ_K = (f(φ) - t) ∇f(φ) with the bilinear derivative of
// initialise_objfun; ∂E/∂v = 2^-K λ_0 + L v.
//
// Two schedules: "store" keeps d_0 .. d_K from the forward pass; "recompute"
// keeps none and rebuilds d_i from v by i squarings for each backward step
// (K(K+1)/2 extra squarings, two map buffers instead of K + 1).
#include <hipcub/hipcub.hpp>

#include "../../include/csg.h"
#include "dartel_kernels.h"

using namespace csg;

namespace {

inline hipStream_t S(void *s) { return (hipStream_t)s; }
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr int kMaxSquarings = 30;

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_scale_velocity(const T *__restrict__ v, T scale, int P, T *__restrict__ d)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    d[p] = v[p] * scale;
    d[P + p] = v[P + p] * scale;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_zero(T *__restrict__ x, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        x[i] = (T)0;
}

// Sample f at p + d(p); per-block float64 partial of (f(φ) - t)²; with lam,
// also λ_K = (f(φ) - t)·∇f(φ).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_warp_ssd(const T *__restrict__ f, const T *__restrict__ t, const T *__restrict__ d, int nx,
           int ny, T *__restrict__ lam, double *__restrict__ partial)
{
    __shared__ double red[kBlock / 64];
    const int P = nx * ny;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (p < P) {
        int x0, x1, y0, y1;
        Corners<T> c = corners_disp<T>(p % nx, p / nx, d[p], d[P + p], nx, ny, x0, x1, y0, y1);
        T k22 = f[c.i22], k12 = f[c.i12], k21 = f[c.i21], k11 = f[c.i11];
        T sampled = (k22 * c.dx2 + k12 * c.dx1) * c.dy2 + (k21 * c.dx2 + k11 * c.dx1) * c.dy1;
        T diff = sampled - t[p];
        e = (double)(diff * diff);
        if (lam) {
            T ddx = (k11 - k21) * c.dy1 + (k12 - k22) * c.dy2;
            T ddy = (k11 * c.dx1 + k21 * c.dx2) - (k12 * c.dx1 + k22 * c.dx2);
            lam[p] = diff * ddx;
            lam[P + p] = diff * ddy;
        }
    }
    double s = block_sum(e, red);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = s;
}

// Unwrapped corner differences k12 - k22, k21 - k22, k11 - k22 of one map
// component, as wrapped_displacement forms them.
template <typename T>
__device__ __forceinline__ void corner_differences(const T *__restrict__ dc, const Corners<T> &c,
                                                   int X22, int X12, int X21, int X11, int period,
                                                   T &e12, T &e21, T &e11)
{
    T d22 = dc[c.i22], d12 = dc[c.i12], d21 = dc[c.i21], d11 = dc[c.i11];
    T Pp = (T)period;
    int n12 = (int)floor(((T)(X12 - X22) + (d12 - d22)) / Pp + (T)0.5);
    int n21 = (int)floor(((T)(X21 - X22) + (d21 - d22)) / Pp + (T)0.5);
    int n11 = (int)floor(((T)(X11 - X22) + (d11 - d22)) / Pp + (T)0.5);
    e12 = (T)(X12 - n12 * period - X22) + (d12 - d22);
    e21 = (T)(X21 - n21 * period - X22) + (d21 - d22);
    e11 = (T)(X11 - n11 * period - X22) + (d11 - d22);
}

// Adjoint of d' = d(id + d) for the upstream λ_out: the position term goes to
// lam_in[p]; corner k of pixel p (22, 12, 21, 11) becomes entry e = 4p + k of
// (dest, entry, cx, cy) = (corner pixel, e, w_k λx, w_k λy).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_square_adjoint_emit(const T *__restrict__ d, const T *__restrict__ lam_out, int nx, int ny,
                      T *__restrict__ lam_in, int *__restrict__ dest, int *__restrict__ entry,
                      T *__restrict__ cx, T *__restrict__ cy)
{
    const int P = nx * ny;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    int x0, x1, y0, y1;
    Corners<T> c = corners_disp<T>(p % nx, p / nx, d[p], d[P + p], nx, ny, x0, x1, y0, y1);
    const T lx = lam_out[p], ly = lam_out[P + p];
    T ex12, ex21, ex11, ey12, ey21, ey11;
    corner_differences<T>(d, c, x0, x1, x0, x1, nx, ex12, ex21, ex11);
    corner_differences<T>(d + P, c, y0, y0, y1, y1, ny, ey12, ey21, ey11);
    // ∂I/∂x = e12 dy2 + (e11 - e21) dy1,  ∂I/∂y = e21 dx2 + (e11 - e12) dx1
    const T gxx = ex12 * c.dy2 + (ex11 - ex21) * c.dy1;
    const T gxy = ex21 * c.dx2 + (ex11 - ex12) * c.dx1;
    const T gyx = ey12 * c.dy2 + (ey11 - ey21) * c.dy1;
    const T gyy = ey21 * c.dx2 + (ey11 - ey12) * c.dx1;
    lam_in[p] = lx * gxx + ly * gyx;
    lam_in[P + p] = lx * gxy + ly * gyy;
    const int to[4] = {c.i22, c.i12, c.i21, c.i11};
    const T w[4] = {c.dx2 * c.dy2, c.dx1 * c.dy2, c.dx2 * c.dy1, c.dx1 * c.dy1};
    const int e0 = 4 * p;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        dest[e0 + k] = to[k];
        entry[e0 + k] = e0 + k;
        cx[e0 + k] = w[k] * lx;
        cy[e0 + k] = w[k] * ly;
    }
}

// lam_in[q] += the corner terms whose destination is q, in source order (the
// run of q in the stably sorted destinations, found by binary search).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_square_adjoint_gather(const int *__restrict__ dest_sorted, const int *__restrict__ entry_sorted,
                        const T *__restrict__ cx, const T *__restrict__ cy, int P,
                        T *__restrict__ lam_in)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P)
        return;
    const int n = 4 * P;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (dest_sorted[mid] < q)
            lo = mid + 1;
        else
            hi = mid;
    }
    T ax = lam_in[q], ay = lam_in[P + q];
    for (int i = lo; i < n && dest_sorted[i] == q; ++i) {
        const int e = entry_sorted[i];
        ax = ax + cx[e];
        ay = ay + cy[e];
    }
    lam_in[q] = ax;
    lam_in[P + q] = ay;
}

int sort_bits(int P)
{
    int bits = 1;
    while ((1LL << bits) < (long long)P)
        ++bits;
    return bits;
}

size_t sort_temp_bytes(int P)
{
    size_t need = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, need, (const int *)nullptr, (int *)nullptr,
                                             (const int *)nullptr, (int *)nullptr, 4 * P, 0,
                                             sort_bits(P), (hipStream_t)0);
    return need;
}

// Per-block float64 partial of Σ v·(L v).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dot(const T *__restrict__ a, const T *__restrict__ b, size_t n, double *__restrict__ partial)
{
    __shared__ double red[kBlock / 64];
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e = i < n ? (double)(a[i] * b[i]) : 0.0;
    double s = block_sum(e, red);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = s;
}

// grad = g·(scale·λ_0 + L v)  (L v optional, g = *grad_scale or 1).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_adjoint_gradient(const T *__restrict__ lam0, const T *__restrict__ lv, T scale,
                   const double *__restrict__ grad_scale, size_t n, T *__restrict__ grad)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    T g = lam0[i] * scale;
    if (lv)
        g = g + lv[i];
    if (grad_scale)
        g = g * (T)grad_scale[0];
    grad[i] = g;
}

struct AdjLayout {
    size_t maps, lam, lv, partial, lists, sort_tmp, sort_tmp_bytes, total;
    int nmaps;
};

// lists: dest, entry, dest sorted, entry sorted (4P int each), cx, cy (4P T each)
AdjLayout adj_layout(int nx, int ny, int K, int dtype, int store)
{
    AdjLayout l{};
    const int P = nx * ny;
    const size_t plane = (size_t)2 * P * dtype_size(dtype);
    l.nmaps = store ? K + 1 : 2;
    l.maps = 0;
    l.lam = l.maps + align256(plane) * l.nmaps;
    l.lv = l.lam + 2 * align256(plane);
    l.partial = l.lv + align256(plane);
    l.lists = l.partial + align256(sizeof(double) * (grid_for((size_t)2 * P) + 8));
    l.sort_tmp = l.lists + 4 * align256((size_t)4 * P * sizeof(int)) +
                 2 * align256((size_t)4 * P * dtype_size(dtype));
    l.sort_tmp_bytes = K > 0 ? sort_temp_bytes(P) : 0;
    l.total = l.sort_tmp + align256(l.sort_tmp_bytes);
    return l;
}

template <typename T>
struct Adj {
    const T *v, *f, *t;
    int nx, ny, K, store;
    double membrane;
    T *maps[kMaxSquarings + 1];
    T *lam[2];
    T *lv;
    double *partial;
    int *dest, *entry, *dest_sorted, *entry_sorted;
    T *cx, *cy;
    void *sort_tmp;
    size_t sort_tmp_bytes;
    hipStream_t s;

    int P() const { return nx * ny; }
};

template <typename T>
int run_square(const T *in, T *out, int nx, int ny, hipStream_t s)
{
    return launch_compose_self<T>(in, nullptr, nx, ny, 1, (size_t)2 * nx * ny, 0, out, nullptr,
                                  s, true);
}

// d_i from v: scale, then i squarings ping-ponging maps[0] / maps[1].
template <typename T>
int rebuild(Adj<T> &a, int i, T **out)
{
    const int P = a.P();
    const T scale = (T)(1.0 / (double)(1LL << a.K));
    hipLaunchKernelGGL(k_scale_velocity<T>, dim3(grid_for(P)), dim3(kBlock), 0, a.s, a.v, scale,
                       P, a.maps[0]);
    CSG_LAUNCH_CHECK();
    int cur = 0;
    for (int k = 0; k < i; ++k) {
        CSG_TRY(run_square<T>(a.maps[cur], a.maps[cur ^ 1], a.nx, a.ny, a.s));
        cur ^= 1;
    }
    *out = a.maps[cur];
    return kOk;
}

template <typename T>
int membrane_lv(Adj<T> &a)
{
    const double params[5] = {1.0, 1.0, a.membrane, 0.0, 0.0};
    return launch_regularization<T>(a.v, 1, params, a.nx, a.ny, a.lv, a.s);
}

template <typename T>
int forward(Adj<T> &a, double *total, double *parts)
{
    const int P = a.P();
    T *dK = nullptr;
    if (a.store) {
        const T scale = (T)(1.0 / (double)(1LL << a.K));
        hipLaunchKernelGGL(k_scale_velocity<T>, dim3(grid_for(P)), dim3(kBlock), 0, a.s, a.v,
                           scale, P, a.maps[0]);
        CSG_LAUNCH_CHECK();
        for (int k = 0; k < a.K; ++k)
            CSG_TRY(run_square<T>(a.maps[k], a.maps[k + 1], a.nx, a.ny, a.s));
        dK = a.maps[a.K];
    } else {
        CSG_TRY(rebuild<T>(a, a.K, &dK));
    }
    const int nb = (int)grid_for(P);
    hipLaunchKernelGGL(k_warp_ssd<T>, dim3(nb), dim3(kBlock), 0, a.s, a.f, a.t, dK, a.nx, a.ny,
                       (T *)nullptr, a.partial);
    CSG_LAUNCH_CHECK();
    CSG_TRY(launch_reduce_partials(a.partial, nb, 1, 0.5, total, 1, 0, a.s));
    if (parts)
        CSG_TRY(launch_reduce_partials(a.partial, nb, 1, 0.5, parts, 1, 0, a.s));
    if (a.membrane != 0.0) {
        CSG_TRY(membrane_lv<T>(a));
        const int nb2 = (int)grid_for((size_t)2 * P);
        hipLaunchKernelGGL(k_dot<T>, dim3(nb2), dim3(kBlock), 0, a.s, a.v, a.lv, (size_t)2 * P,
                           a.partial);
        CSG_LAUNCH_CHECK();
        CSG_TRY(launch_reduce_partials(a.partial, nb2, 1, 0.5, total, 1, 1, a.s));
        if (parts)
            CSG_TRY(launch_reduce_partials(a.partial, nb2, 1, 0.5, parts + 1, 1, 0, a.s));
    } else if (parts) {
        hipLaunchKernelGGL(k_zero<double>, dim3(1), dim3(kBlock), 0, a.s, parts + 1, (size_t)1);
        CSG_LAUNCH_CHECK();
    }
    return kOk;
}

template <typename T>
int backward(Adj<T> &a, const double *grad_scale, T *grad)
{
    const int P = a.P();
    const size_t n2 = (size_t)2 * P;
    T *d = nullptr;
    if (a.store)
        d = a.maps[a.K];
    else
        CSG_TRY(rebuild<T>(a, a.K, &d));
    const int nb = (int)grid_for(P);
    int cur = 0;  // lam[cur] holds λ_{i+1}
    hipLaunchKernelGGL(k_warp_ssd<T>, dim3(nb), dim3(kBlock), 0, a.s, a.f, a.t, d, a.nx, a.ny,
                       a.lam[cur], a.partial);
    CSG_LAUNCH_CHECK();
    for (int i = a.K - 1; i >= 0; --i) {
        if (a.store)
            d = a.maps[i];
        else
            CSG_TRY(rebuild<T>(a, i, &d));
        hipLaunchKernelGGL(k_square_adjoint_emit<T>, dim3(nb), dim3(kBlock), 0, a.s, d,
                           a.lam[cur], a.nx, a.ny, a.lam[cur ^ 1], a.dest, a.entry, a.cx, a.cy);
        CSG_LAUNCH_CHECK();
        size_t tmp_bytes = a.sort_tmp_bytes;
        CSG_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(a.sort_tmp, tmp_bytes,
                                                         (const int *)a.dest, a.dest_sorted,
                                                         (const int *)a.entry, a.entry_sorted,
                                                         4 * P, 0, sort_bits(P), a.s));
        hipLaunchKernelGGL(k_square_adjoint_gather<T>, dim3(nb), dim3(kBlock), 0, a.s,
                           a.dest_sorted, a.entry_sorted, a.cx, a.cy, P, a.lam[cur ^ 1]);
        CSG_LAUNCH_CHECK();
        cur ^= 1;
    }
    const T *lv = nullptr;
    if (a.membrane != 0.0) {
        CSG_TRY(membrane_lv<T>(a));
        lv = a.lv;
    }
    const T scale = (T)(1.0 / (double)(1LL << a.K));
    hipLaunchKernelGGL(k_adjoint_gradient<T>, dim3(grid_for(n2)), dim3(kBlock), 0, a.s,
                       a.lam[cur], lv, scale, grad_scale, n2, grad);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int make_adj(Adj<T> &a, const void *v, const void *f, const void *t, int nx, int ny, int K,
             double membrane, int dtype, int store, void *ws, size_t ws_bytes, void *stream)
{
    if (nx < 2 || ny < 1 || K < 0 || K > kMaxSquarings || v == nullptr || f == nullptr ||
        t == nullptr)
        return fail(kInvalidArgument,
                    "exp-warp-SSD needs nx >= 2, ny >= 1, 0 <= K <= %d and all inputs",
                    kMaxSquarings);
    if ((long long)nx * ny > (1LL << 28))
        return fail(kInvalidArgument, "exp-warp-SSD supports at most 2^28 pixels");
    const AdjLayout l = adj_layout(nx, ny, K, dtype, store);
    if (ws_bytes < l.total)
        return fail(kInvalidArgument, "exp-warp-SSD workspace too small");
    char *b = (char *)ws;
    const size_t plane = align256((size_t)2 * nx * ny * sizeof(T));
    a.v = (const T *)v;
    a.f = (const T *)f;
    a.t = (const T *)t;
    a.nx = nx;
    a.ny = ny;
    a.K = K;
    a.store = store;
    a.membrane = membrane;
    for (int i = 0; i < l.nmaps; ++i)
        a.maps[i] = (T *)(b + l.maps + plane * i);
    a.lam[0] = (T *)(b + l.lam);
    a.lam[1] = (T *)(b + l.lam + plane);
    a.lv = (T *)(b + l.lv);
    a.partial = (double *)(b + l.partial);
    const size_t ilist = align256((size_t)4 * nx * ny * sizeof(int));
    const size_t tlist = align256((size_t)4 * nx * ny * sizeof(T));
    a.dest = (int *)(b + l.lists);
    a.entry = (int *)(b + l.lists + ilist);
    a.dest_sorted = (int *)(b + l.lists + 2 * ilist);
    a.entry_sorted = (int *)(b + l.lists + 3 * ilist);
    a.cx = (T *)(b + l.lists + 4 * ilist);
    a.cy = (T *)(b + l.lists + 4 * ilist + tlist);
    a.sort_tmp = b + l.sort_tmp;
    a.sort_tmp_bytes = l.sort_tmp_bytes;
    a.s = S(stream);
    return kOk;
}

}  // namespace

extern "C" {

size_t csg_exp_warp_ssd_workspace_bytes(int nx, int ny, int K, int dtype, int store)
{
    if (nx < 1 || ny < 1 || K < 0 || K > kMaxSquarings || (long long)nx * ny > (1LL << 28))
        return 0;
    return adj_layout(nx, ny, K, dtype, store ? 1 : 0).total + 256;
}

int csg_exp_warp_ssd_forward(const void *v, const void *feature, const void *target, int nx,
                             int ny, int K, double membrane, int dtype, int store,
                             double *energy, double *parts, void *ws, size_t ws_bytes,
                             void *stream)
{
    if (energy == nullptr)
        return fail(kInvalidArgument, "exp-warp-SSD needs an energy output");
    if (dtype == 0) {
        Adj<float> a{};
        CSG_TRY(make_adj<float>(a, v, feature, target, nx, ny, K, membrane, 0, store ? 1 : 0, ws,
                                ws_bytes, stream));
        return forward<float>(a, energy, parts);
    }
    if (dtype == 1) {
        Adj<double> a{};
        CSG_TRY(make_adj<double>(a, v, feature, target, nx, ny, K, membrane, 1, store ? 1 : 0,
                                 ws, ws_bytes, stream));
        return forward<double>(a, energy, parts);
    }
    return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}

int csg_exp_warp_ssd_backward(const void *v, const void *feature, const void *target, int nx,
                              int ny, int K, double membrane, int dtype, int store,
                              const double *grad_scale, void *grad_v, void *ws,
                              size_t ws_bytes, void *stream)
{
    if (grad_v == nullptr)
        return fail(kInvalidArgument, "exp-warp-SSD needs a gradient output");
    if (dtype == 0) {
        Adj<float> a{};
        CSG_TRY(make_adj<float>(a, v, feature, target, nx, ny, K, membrane, 0, store ? 1 : 0, ws,
                                ws_bytes, stream));
        return backward<float>(a, grad_scale, (float *)grad_v);
    }
    if (dtype == 1) {
        Adj<double> a{};
        CSG_TRY(make_adj<double>(a, v, feature, target, nx, ny, K, membrane, 1, store ? 1 : 0,
                                 ws, ws_bytes, stream));
        return backward<double>(a, grad_scale, (double *)grad_v);
    }
    return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}

}  // extern "C"
