// Surface-side kernels for gfx950: resampling onto the stencil sphere,
// heat-kernel smoothing over a CSR neighbour table, atomic-free
// angle-weighted normals, type-5 curvature, sheet gather, min/max
// normalisation, the depth-potential solve for the rotation feature, the
// final sphere warp and the pole-frame average.
//
// Reference semantics: src/cat_surface_gpu/surface_stencil.py:282-550,
// src/cat_surface_gpu/rotation_feature.py:24-317, :320-399,
// src/cat_surface_gpu/surface_warp.py:491-584,
// src/cat_surface_gpu/gpu_pipeline.py:248-271.
//
// Scatter-adds of the reference (index_add_) are replaced by gathers over a
// per-vertex list of incident (face, corner) entries sorted in the order the
// reference adds them, so every sum is deterministic and atomic-free.
#include "surface_kernels.h"
#include "kernel_timing.h"

#include <algorithm>

#include <mutex>

namespace csg {

__device__ __forceinline__ float norm3f(float x, float y, float z)
{
    // torch's float32 2-norm: sqrt(fma(z, z, fma(y, y, x * x)))
    return sqrtf(fmaf(z, z, fmaf(y, y, x * x)));
}

// out[r] = Σ_c f32(f64(x[idx[r,c]])·w[r,c]) in corner order (float32 input) or
// the float64 sum (float64 input).  surface_stencil.py:282-306
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_resample(const T *__restrict__ x, const int *__restrict__ idx, const double *__restrict__ w,
           int n, T *__restrict__ out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n)
        return;
    int r = t / 3, d = t - 3 * r;
    const int *ix = idx + 3 * r;
    const double *wr = w + 3 * r;
    if (sizeof(T) == 4) {
        float acc = 0.0f;
        for (int c = 0; c < 3; ++c)
            acc = acc + (float)((double)x[3 * ix[c] + d] * wr[c]);
        out[t] = (T)acc;
    } else {
        double acc = 0.0;
        for (int c = 0; c < 3; ++c)
            acc = acc + (double)x[3 * ix[c] + d] * wr[c];
        out[t] = (T)acc;
    }
}

// Per-block partial sums of edge lengths over faces.  surface_stencil.py:349-357
__global__ void __launch_bounds__(kBlock)
k_edge_sum(const float *__restrict__ v, const int *__restrict__ faces, int nf,
           double *__restrict__ partial)
{
    __shared__ double red[kBlock / 64];
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (f < nf) {
        for (int e = 0; e < 3; ++e) {
            int a = faces[3 * f + e], b = faces[3 * f + (e + 1) % 3];
            double dx = (double)v[3 * b] - v[3 * a];
            double dy = (double)v[3 * b + 1] - v[3 * a + 1];
            double dz = (double)v[3 * b + 2] - v[3 * a + 2];
            s += sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    double r = block_sum(s, red);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = r;
}

// One heat iteration on geometry: v <- f32((v + Σ_nbr v / 3) / (1 + deg / 3)).
// surface_stencil.py:372-388
//
// The divisors are 3 and den = 1 + deg / 3 (a handful of values), so each
// division x / d is taken from y = 1 / d, computed once (div_known).

struct SmoothDen {
    double den, inv;
};

__device__ __forceinline__ SmoothDen smooth_den(int deg)
{
    const double den = 1.0 + (double)deg / 3.0;
    return {den, 1.0 / den};
}

// f32((c + s / 3) / den) with both divisions through div_known.
__device__ __forceinline__ float smooth_coord(float c, double s, SmoothDen d)
{
    constexpr double kThird = 1.0 / 3.0;
    return (float)div_known((double)c + div_known(s, 3.0, kThird), d.den, d.inv);
}

__global__ void __launch_bounds__(kBlock)
k_smooth_geometry(const float *__restrict__ cur, const int *__restrict__ off,
                  const int *__restrict__ nbr, int n, float *__restrict__ nxt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int b = off[i], e = off[i + 1];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = b; k < e; ++k) {
        const float *p = cur + 3 * nbr[k];
        sx += (double)p[0];
        sy += (double)p[1];
        sz += (double)p[2];
    }
    const SmoothDen den = smooth_den(e - b);
    nxt[3 * i] = smooth_coord(cur[3 * i], sx, den);
    nxt[3 * i + 1] = smooth_coord(cur[3 * i + 1], sy, den);
    nxt[3 * i + 2] = smooth_coord(cur[3 * i + 2], sz, den);
}

// One smoothing step for a batch of meshes of one topology, each with its own
// number of steps: mesh b reads buffer ((start_b + it) & 1) and writes the
// other; meshes whose steps are done are left alone.  Same arithmetic as
// k_smooth_geometry.
__global__ void __launch_bounds__(kBlock)
k_smooth_geometry_batch(float *__restrict__ buf0, float *__restrict__ buf1,
                        const int *__restrict__ off, const int *__restrict__ nbr, int n,
                        SmoothBatch sb, int it)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= n || it >= smooth_iters(sb, b))
        return;
    const bool from1 = ((smooth_iters(sb, b) + it) & 1) != 0;
    const float *cur = (from1 ? buf1 : buf0) + (size_t)b * 3 * n;
    float *nxt = (from1 ? buf0 : buf1) + (size_t)b * 3 * n;
    int bo = off[i], e = off[i + 1];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = bo; k < e; ++k) {
        const float *p = cur + 3 * nbr[k];
        sx += (double)p[0];
        sy += (double)p[1];
        sz += (double)p[2];
    }
    const SmoothDen den = smooth_den(e - bo);
    nxt[3 * i] = smooth_coord(cur[3 * i], sx, den);
    nxt[3 * i + 1] = smooth_coord(cur[3 * i + 1], sy, den);
    nxt[3 * i + 2] = smooth_coord(cur[3 * i + 2], sz, den);
}

// The batched step on padded float4 vertices (x, y, z, 0): one 16-byte load
// per neighbour instead of three 4-byte loads; same arithmetic and order as
// k_smooth_geometry.  Mesh b reads buf[it & 1] and writes buf[(it + 1) & 1]
// while it < iters[b].
__global__ void __launch_bounds__(kBlock)
k_smooth_geometry4(float4 *__restrict__ buf0, float4 *__restrict__ buf1,
                   const int *__restrict__ off, const int *__restrict__ nbr, int n,
                   SmoothBatch sb, int it)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= n || it >= smooth_iters(sb, b))
        return;
    const float4 *cur = ((it & 1) ? buf1 : buf0) + (size_t)b * n;
    float4 *nxt = ((it & 1) ? buf0 : buf1) + (size_t)b * n;
    int bo = off[i], e = off[i + 1];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = bo; k < e; ++k) {
        const float4 p = cur[nbr[k]];
        sx += (double)p.x;
        sy += (double)p.y;
        sz += (double)p.z;
    }
    const float4 c = cur[i];
    const SmoothDen den = smooth_den(e - bo);
    nxt[i] = make_float4(smooth_coord(c.x, sx, den), smooth_coord(c.y, sy, den),
                         smooth_coord(c.z, sz, den), 0.0f);
}

// Up to `rings` iterations of k_smooth_geometry4 in one launch, one
// workgroup per (patch, mesh).  A patch lists its core vertices (ring 0) and
// the rings 1..rings of vertices at growing graph distance; every neighbour of
// a vertex in rings < `rings` is in the patch.  The workgroup stages the
// patch's positions and local adjacency in LDS, and iteration t of m updates
// rings 0..m-1-t (whose neighbours then hold iterate t), so after m
// iterations the core holds exactly the per-iteration kernel's values (same
// arithmetic, neighbours in the same order).  Meshes with fewer than m
// iterations left (m = 0: a copy) are advanced by what they have left.
constexpr int kPatchThreads = 1024;
constexpr int kDenTable = 16;
constexpr int kPatchGather = 6;  // neighbours gathered together (icosphere degrees 5, 6)

__global__ void __launch_bounds__(kPatchThreads)
k_smooth_geometry_patch(const float4 *__restrict__ in, float4 *__restrict__ out, int n,
                        SmoothPatches pp, SmoothBatch sb, int it0)
{
    extern __shared__ float g_patch_lds[];
    // patches (consecutive in Morton order) relabelled XCD by XCD: a patch's
    // halo lies in its neighbours' cores
    const Block2 blk = xcd_block();
    const int p = blk.x, b = blk.y, tid = threadIdx.x;
    const int s0 = pp.vstart[p], L = pp.vstart[p + 1] - s0;
    const int a0 = pp.astart[p], E = pp.astart[p + 1] - a0;
    const int *lv = pp.lverts + s0;
    const int *ring_end = pp.ring_end + p * (pp.rings + 1);
    const float4 *src = in + (size_t)b * n;
    float4 *dst = out + (size_t)b * n;
    int m = smooth_iters(sb, b) - it0;
    m = m < 0 ? 0 : (m > pp.rings ? pp.rings : m);
    const int core = ring_end[0];
    if (m == 0) {
        for (int j = tid; j < core; j += kPatchThreads)
            dst[lv[j]] = src[lv[j]];
        return;
    }
    // LDS: two position buffers as x, y, z planes of L floats, then the local
    // CSR offsets (L + 1) and neighbours (E) as 16-bit indices
    float *buf[2] = {g_patch_lds, g_patch_lds + 3 * L};
    unsigned short *off = reinterpret_cast<unsigned short *>(g_patch_lds + 6 * L);
    unsigned short *adj = off + L + 1;
    const int *aoff = pp.aoff + s0 + p;
    for (int j = tid; j < L; j += kPatchThreads) {
        const float4 v = src[lv[j]];
        buf[0][j] = v.x;
        buf[0][L + j] = v.y;
        buf[0][2 * L + j] = v.z;
    }
    for (int j = tid; j <= L; j += kPatchThreads)
        off[j] = (unsigned short)aoff[j];
    for (int k = tid; k < E; k += kPatchThreads)
        adj[k] = pp.ladj[a0 + k];
    // den and 1 / den of the small degrees, once per workgroup
    __shared__ double den_t[kDenTable], inv_t[kDenTable];
    if (tid < kDenTable) {
        const SmoothDen d = smooth_den(tid);
        den_t[tid] = d.den;
        inv_t[tid] = d.inv;
    }
    __syncthreads();
    for (int t = 0; t < m; ++t) {
        const float *cx = buf[t & 1], *cy = cx + L, *cz = cy + L;
        float *nx = buf[(t + 1) & 1], *ny = nx + L, *nz = ny + L;
        const int count = ring_end[m - 1 - t];
        for (int j = tid; j < count; j += kPatchThreads) {
            const int bo = off[j], e = off[j + 1];
            double sx = 0.0, sy = 0.0, sz = 0.0;
            // the first kPatchGather neighbours' loads are issued together; the
            // sums still run in neighbour order
            int q[kPatchGather];
#pragma unroll
            for (int k = 0; k < kPatchGather; ++k)
                q[k] = bo + k < e ? adj[bo + k] : j;
            float qx[kPatchGather], qy[kPatchGather], qz[kPatchGather];
#pragma unroll
            for (int k = 0; k < kPatchGather; ++k) {
                qx[k] = cx[q[k]];
                qy[k] = cy[q[k]];
                qz[k] = cz[q[k]];
            }
#pragma unroll
            for (int k = 0; k < kPatchGather; ++k)
                if (bo + k < e) {
                    sx += (double)qx[k];
                    sy += (double)qy[k];
                    sz += (double)qz[k];
                }
            for (int k = bo + kPatchGather; k < e; ++k) {
                const int r = adj[k];
                sx += (double)cx[r];
                sy += (double)cy[r];
                sz += (double)cz[r];
            }
            const int deg = e - bo;
            const SmoothDen den = deg < kDenTable ? SmoothDen{den_t[deg], inv_t[deg]}
                                                  : smooth_den(deg);
            nx[j] = smooth_coord(cx[j], sx, den);
            ny[j] = smooth_coord(cy[j], sy, den);
            nz[j] = smooth_coord(cz[j], sz, den);
        }
        __syncthreads();
    }
    const float *fx = buf[m & 1], *fy = fx + L, *fz = fy + L;
    for (int j = tid; j < core; j += kPatchThreads)
        dst[lv[j]] = make_float4(fx[j], fy[j], fz[j], 0.0f);
}

// Up to `rings` iterations of k_smooth_values in one launch, one workgroup per
// (patch, batch row), on the patches of k_smooth_geometry_patch (same ring
// schedule).  Rows stay in vertex order (perm maps a patch entry to its
// vertex); each updated vertex's weights are staged in CSR order with their
// sum, so every iterate equals the per-iteration kernel's.
__global__ void __launch_bounds__(kPatchThreads)
k_smooth_values_patch(const double *__restrict__ in, double *__restrict__ out, int n,
                      const int *__restrict__ goff, const double *__restrict__ w,
                      size_t wstride, SmoothPatches pp, HeatBatch hb, int it0)
{
    extern __shared__ double g_value_lds[];
    const Block2 blk = xcd_block();  // as k_smooth_geometry_patch
    const int p = blk.x, b = blk.y, tid = threadIdx.x;
    const int s0 = pp.vstart[p], L = pp.vstart[p + 1] - s0;
    const int a0 = pp.astart[p], E = pp.astart[p + 1] - a0;
    const int *lv = pp.lverts + s0;
    const int *ring_end = pp.ring_end + p * (pp.rings + 1);
    const double *src = in + (size_t)b * n;
    double *dst = out + (size_t)b * n;
    int m = heat_count(hb, b) - it0;
    m = m < 0 ? 0 : (m > pp.rings ? pp.rings : m);
    const int core = ring_end[0];
    if (m == 0) {
        for (int j = tid; j < core; j += kPatchThreads) {
            const int g = pp.perm[lv[j]];
            dst[g] = src[g];
        }
        return;
    }
    // LDS: two value buffers (L), the weights (E) and weight sums (L), then the
    // local CSR offsets (L + 1) and neighbours (E) as 16-bit indices
    double *buf[2] = {g_value_lds, g_value_lds + L};
    double *wl = g_value_lds + 2 * L, *wsl = wl + E;
    unsigned short *off = reinterpret_cast<unsigned short *>(wsl + L);
    unsigned short *adj = off + L + 1;
    const int *aoff = pp.aoff + s0 + p;
    const double *wb = w + (size_t)b * wstride;
    for (int j = tid; j < L; j += kPatchThreads)
        buf[0][j] = src[pp.perm[lv[j]]];
    for (int j = tid; j <= L; j += kPatchThreads)
        off[j] = (unsigned short)aoff[j];
    for (int k = tid; k < E; k += kPatchThreads)
        adj[k] = pp.ladj[a0 + k];
    const int updated = ring_end[pp.rings - 1];
    for (int j = tid; j < updated; j += kPatchThreads) {
        const int bo = aoff[j], e = aoff[j + 1];
        const double *wk = wb + goff[pp.perm[lv[j]]];
        const int d = e - bo;
        // the first kPatchGather weights are loaded together, then summed in order
        double x[kPatchGather];
#pragma unroll
        for (int k = 0; k < kPatchGather; ++k)
            x[k] = k < d ? wk[k] : 0.0;
        double ws = 0.0;
#pragma unroll
        for (int k = 0; k < kPatchGather; ++k)
            if (k < d) {
                wl[bo + k] = x[k];
                ws += x[k];
            }
        for (int k = kPatchGather; k < d; ++k) {
            const double y = wk[k];
            wl[bo + k] = y;
            ws += y;
        }
        wsl[j] = ws;
    }
    __syncthreads();
    for (int t = 0; t < m; ++t) {
        const double *cv = buf[t & 1];
        double *nv = buf[(t + 1) & 1];
        const int count = ring_end[m - 1 - t];
        for (int j = tid; j < count; j += kPatchThreads) {
            const int bo = off[j], e = off[j + 1];
            // the first kPatchGather neighbours' loads are issued together; the
            // sum still runs in neighbour order
            int q[kPatchGather];
#pragma unroll
            for (int k = 0; k < kPatchGather; ++k)
                q[k] = bo + k < e ? adj[bo + k] : j;
            double qv[kPatchGather], qw[kPatchGather];
#pragma unroll
            for (int k = 0; k < kPatchGather; ++k) {
                qv[k] = cv[q[k]];
                qw[k] = bo + k < e ? wl[bo + k] : 0.0;
            }
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < kPatchGather; ++k)
                if (bo + k < e)
                    s += qw[k] * qv[k];
            for (int k = bo + kPatchGather; k < e; ++k)
                s += wl[k] * cv[adj[k]];
            nv[j] = (cv[j] + s) / (1.0 + wsl[j]);
        }
        __syncthreads();
    }
    const double *fin = buf[m & 1];
    for (int j = tid; j < core; j += kPatchThreads)
        dst[pp.perm[lv[j]]] = fin[j];
}

__global__ void __launch_bounds__(kBlock)
k_pack4(const float *__restrict__ v, int n, float4 *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t b = blockIdx.y;
    if (i >= n)
        return;
    const float *p = v + (b * n + i) * 3;
    out[b * n + i] = make_float4(p[0], p[1], p[2], 0.0f);
}

// out = the buffer holding mesh b's last iterate, buf[iters[b] & 1].
__global__ void __launch_bounds__(kBlock)
k_unpack4(const float4 *__restrict__ buf0, const float4 *__restrict__ buf1, int n,
          SmoothBatch sb, float *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t b = blockIdx.y;
    if (i >= n)
        return;
    const float4 q = ((smooth_iters(sb, (int)b) & 1) ? buf1 : buf0)[b * n + i];
    float *o = out + (b * n + i) * 3;
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
}

// k_pack4 / k_unpack4 through a vertex permutation: working position i of
// mesh b holds vertex perm[i].
__global__ void __launch_bounds__(kBlock)
k_pack4_perm(const float *__restrict__ v, int n, const int *__restrict__ perm,
             float4 *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t b = blockIdx.y;
    if (i >= n)
        return;
    const float *p = v + (b * n + perm[i]) * 3;
    out[b * n + i] = make_float4(p[0], p[1], p[2], 0.0f);
}

__global__ void __launch_bounds__(kBlock)
k_unpack4_perm(const float4 *__restrict__ buf, int n, const int *__restrict__ perm,
               float *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t b = blockIdx.y;
    if (i >= n)
        return;
    const float4 q = buf[b * n + i];
    float *o = out + (b * n + perm[i]) * 3;
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
}

// Angle-weighted vertex normals (float32).  inc[] lists face*3+corner per
// vertex in (corner, face) order, the order index_add_ applies them.
// surface_stencil.py:478-505
__global__ void __launch_bounds__(kBlock)
k_vertex_normals(const float *__restrict__ v, const int *__restrict__ faces,
                 const int *__restrict__ inc_off, const int *__restrict__ inc, int n,
                 float *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float eps = 1.1920928955078125e-07f;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (int k = inc_off[i]; k < inc_off[i + 1]; ++k) {
        int fc = inc[k];
        int f = fc / 3, corner = fc - 3 * f;
        const int *fv = faces + 3 * f;
        const float *p0 = v + 3 * fv[0], *p1 = v + 3 * fv[1], *p2 = v + 3 * fv[2];
        float a0 = p1[0] - p0[0], a1 = p1[1] - p0[1], a2 = p1[2] - p0[2];
        float b0 = p2[0] - p0[0], b1 = p2[1] - p0[1], b2 = p2[2] - p0[2];
        float c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
        float nr = fmaxf(norm3f(c0, c1, c2), eps);
        float n0 = c0 / nr, n1 = c1 / nr, n2 = c2 / nr;
        const float *pc = v + 3 * fv[corner];
        const float *pp = v + 3 * fv[(corner + 2) % 3];
        const float *pn = v + 3 * fv[(corner + 1) % 3];
        float r0 = pp[0] - pc[0], r1 = pp[1] - pc[1], r2 = pp[2] - pc[2];
        float f0 = pn[0] - pc[0], f1 = pn[1] - pc[1], f2 = pn[2] - pc[2];
        float den = fmaxf(norm3f(r0, r1, r2) * norm3f(f0, f1, f2), eps);
        float cosv = (r0 * f0 + r1 * f1 + r2 * f2) / den;
        cosv = fminf(fmaxf(cosv, -1.0f), 1.0f);
        float ang = acosf(cosv);
        ax += n0 * ang;
        ay += n1 * ang;
        az += n2 * ang;
    }
    float nr = fmaxf(norm3f(ax, ay, az), eps);
    out[3 * i] = ax / nr;
    out[3 * i + 1] = ay / nr;
    out[3 * i + 2] = az / nr;
}

// base = f64(Σ_d f32((smoothed - geom)·normals)).  surface_stencil.py:514-519
__global__ void __launch_bounds__(kBlock)
k_type5_base(const float *__restrict__ geom, const float *__restrict__ sm,
             const float *__restrict__ nr, int n, double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float s = 0.0f;
    for (int d = 0; d < 3; ++d)
        s = s + (sm[3 * i + d] - geom[3 * i + d]) * nr[3 * i + d];
    out[i] = (double)s;
}

// Type-2 curvature before smoothing (surface_stencil.py:552-613): at vertex i
// with float32 normals n, the neighbours' coordinate and normal differences
// projected on the tangent plane, in the basis (b0, b1) built from the first
// neighbour (ascending CSR order), fit the shape operator [[a, b], [b, c]] by
// least squares; the value is sqrt((k1^2 + k2^2) / 2) of its eigenvalues, 0 for
// vertices with at most two neighbours.  float64, the reference's evaluation
// order.
__device__ __forceinline__ double dot3d(const double *u, const double *v)
{
    return (u[0] * v[0] + u[1] * v[1]) + u[2] * v[2];
}

__global__ void __launch_bounds__(kBlock)
k_curvature_type2(const float *__restrict__ geom, const float *__restrict__ nr, int n,
                  const int *__restrict__ off, const int *__restrict__ nbr,
                  double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double eps = 2.220446049250313e-16;
    const int k0 = off[i], k1e = off[i + 1], deg = k1e - k0;
    double nn[3], pi[3];
    for (int d = 0; d < 3; ++d) {
        nn[d] = (double)nr[3 * i + d];
        pi[d] = (double)geom[3 * i + d];
    }
    // basis from the first neighbour (the padded table's column 0)
    double b0[3], b1[3];
    {
        const int j = deg > 0 ? nbr[k0] : 0;
        double dc[3];
        for (int d = 0; d < 3; ++d)
            dc[d] = (double)geom[3 * j + d] - pi[d];
        const double t = dot3d(dc, nn);
        for (int d = 0; d < 3; ++d)
            b0[d] = dc[d] - t * nn[d];
        const double len = fmax(sqrt(dot3d(b0, b0)), eps);
        for (int d = 0; d < 3; ++d)
            b0[d] = b0[d] / len;
        const double m0[3] = {-b0[0], -b0[1], -b0[2]};
        b1[0] = m0[1] * nn[2] - m0[2] * nn[1];
        b1[1] = m0[2] * nn[0] - m0[0] * nn[2];
        b1[2] = m0[0] * nn[1] - m0[1] * nn[0];
    }
    double sum1 = 0.0, sum2 = 0.0, sum3 = 0.0, wx = 0.0, wy = 0.0, wxy = 0.0;
    for (int k = k0; k < k1e; ++k) {
        const int j = nbr[k];
        double dc[3], dn[3], pc[3], pn[3];
        for (int d = 0; d < 3; ++d) {
            dc[d] = (double)geom[3 * j + d] - pi[d];
            dn[d] = (double)nr[3 * j + d] - nn[d];
        }
        const double tc = dot3d(dc, nn), tn = dot3d(dn, nn);
        for (int d = 0; d < 3; ++d) {
            pc[d] = dc[d] - tc * nn[d];
            pn[d] = dn[d] - tn * nn[d];
        }
        const double dcx = dot3d(pc, b0), dcy = dot3d(pc, b1), dnx = dot3d(pn, b0),
                     dny = dot3d(pn, b1);
        sum1 += dcx * dnx;
        sum2 += dcx * dny + dcy * dnx;
        sum3 += dcy * dny;
        wx += dcx * dcx;
        wy += dcy * dcy;
        wxy += dcx * dcy;
    }
    const double wx2 = wx * wx, wy2 = wy * wy, wxy2 = wxy * wxy;
    const double den = (wx + wy) * (-wxy2 + wx * wy);
    const double sden = den != 0.0 ? den : 1.0;
    const double a = (sum3 * wxy2 - sum2 * wxy * wy + sum1 * (-wxy2 + wx * wy + wy2)) / sden;
    const double b = (-sum3 * wx * wxy + sum2 * wx * wy - sum1 * wxy * wy) / sden;
    const double c = (-sum2 * wx * wxy + sum1 * wxy2 + sum3 * (wx2 - wxy2 + wx * wy)) / sden;
    const double tr = a + c, det = a * c - b * b;
    const double disc = tr * tr - 4.0 * det;
    const double root = sqrt(fmax(disc, 0.0));
    const double e1 = (tr + root) / 2.0, e2 = (tr - root) / 2.0;
    out[i] = deg > 2 ? sqrt((e1 * e1 + e2 * e2) / 2.0) : 0.0;
}

// Gaussian weights per CSR entry for `nb` sigmas, and their row sums.
// squared_norm: d² as a sum of squares (smooth_values_many) else |d|².
// surface_stencil.py:390-412, :440-452
__global__ void __launch_bounds__(kBlock)
k_heat_weights(const float *__restrict__ geom, const int *__restrict__ off,
               const int *__restrict__ nbr, int n, HeatBatch hb, int squared_norm, size_t wstride,
               double *__restrict__ w, double *__restrict__ wsum)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int b = blockIdx.y;
    double sg = heat_sigma(hb, b);
    geom += (size_t)hb.mesh[b] * 3 * n;
    double gx = geom[3 * i], gy = geom[3 * i + 1], gz = geom[3 * i + 2];
    double ws = 0.0;
    for (int k = off[i]; k < off[i + 1]; ++k) {
        const float *p = geom + 3 * nbr[k];
        double dx = (double)p[0] - gx, dy = (double)p[1] - gy, dz = (double)p[2] - gz;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (!squared_norm) {
            double d = sqrt(d2);
            d2 = d * d;
        }
        double wk = exp(-d2 / (2.0 * sg * sg));
        w[b * wstride + k] = wk;
        ws += wk;
    }
    wsum[(size_t)b * n + i] = ws;
}

// One heat iteration on values for every batch row whose count is not yet
// reached: v <- (v + Σ w·v_nbr) / (1 + Σ w).  surface_stencil.py:404-476
__global__ void __launch_bounds__(kBlock)
k_smooth_values(const double *__restrict__ cur, const int *__restrict__ off,
                const int *__restrict__ nbr, const double *__restrict__ w,
                const double *__restrict__ wsum, int n, size_t wstride, HeatBatch hb,
                int iteration, double *__restrict__ nxt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int b = blockIdx.y;
    const double *cv = cur + (size_t)b * n;
    double *nv = nxt + (size_t)b * n;
    if (iteration >= heat_count(hb, b)) {
        nv[i] = cv[i];
        return;
    }
    const double *wb = w + b * wstride;
    double s = 0.0, ws = 0.0;
    for (int k = off[i]; k < off[i + 1]; ++k) {
        s += wb[k] * cv[nbr[k]];
        ws += wb[k];
    }
    nv[i] = (cv[i] + s) / (1.0 + ws);
}

// Per-block (min, max) of `n` doubles for each batch row.
__global__ void __launch_bounds__(kBlock)
k_minmax_partial(const double *__restrict__ x, size_t n, size_t stride,
                 double *__restrict__ partial, int nblocks)
{
    __shared__ double lo_s[kBlock], hi_s[kBlock];
    const double *xb = x + blockIdx.y * stride;
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        double v = xb[i];
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    lo_s[threadIdx.x] = lo;
    hi_s[threadIdx.x] = hi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            lo_s[threadIdx.x] = fmin(lo_s[threadIdx.x], lo_s[threadIdx.x + s]);
            hi_s[threadIdx.x] = fmax(hi_s[threadIdx.x], hi_s[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * ((size_t)blockIdx.y * nblocks + blockIdx.x)] = lo_s[0];
        partial[2 * ((size_t)blockIdx.y * nblocks + blockIdx.x) + 1] = hi_s[0];
    }
}

// out = (x - min) / max(max - min, floor).  The reference divides by
// (max - min) directly for the curvature maps and clamps by DBL_EPSILON for
// the rotation feature (rotation_feature.py:393-398); floor selects which.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_minmax_apply(const double *__restrict__ x, size_t n, size_t stride,
               const double *__restrict__ partial, int nblocks, double floor_,
               T *__restrict__ out, size_t ostride)
{
    const int b = blockIdx.y;
    double lo = INFINITY, hi = -INFINITY;
    const double *pb = partial + 2 * (size_t)b * nblocks;
    for (int k = 0; k < nblocks; ++k) {
        lo = fmin(lo, pb[2 * k]);
        hi = fmax(hi, pb[2 * k + 1]);
    }
    double range = hi - lo;
    if (floor_ > 0.0)
        range = fmax(range, floor_);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[b * ostride + i] = (T)((x[b * stride + i] - lo) / range);
}

// sheet[b][p] = Σ_c values[b][idx[p,c]]·w[p,c].  surface_stencil.py:308-317
__global__ void __launch_bounds__(kBlock)
k_sheet_gather(const double *__restrict__ vals, size_t vstride, const int *__restrict__ idx,
               const double *__restrict__ w, int m, double *__restrict__ out)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m)
        return;
    const double *v = vals + blockIdx.y * vstride;
    const int *ix = idx + 3 * p;
    const double *wp = w + 3 * p;
    out[(size_t)blockIdx.y * m + p] = v[ix[0]] * wp[0] + v[ix[1]] * wp[1] + v[ix[2]] * wp[2];
}

// ------------------------------------------------------------------- warp

// Sphere points through a flow.  surface_warp.py:491-584
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_apply_flow(const float *__restrict__ unit, int n, const T *__restrict__ flow, int nx, int ny,
             int inverse, T *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const T pi = (T)3.14159265358979323846;
    const int P = nx * ny;
    T x = (T)unit[3 * i], y = (T)unit[3 * i + 1], z = (T)unit[3 * i + 2];
    z = z < (T)-1 ? (T)-1 : (z > (T)1 ? (T)1 : z);
    T theta = acos(z);
    T cw = -atan2(x, y);
    if (cw < (T)0)
        cw = cw + (T)2 * pi;
    T u = (cw + pi / (T)2) / ((T)2 * pi);
    if (u > (T)1)
        u = u - (T)1;
    T v = theta / pi;
    Corners<T> c = corners_at<T>(u * (T)nx - (T)0.5, v * (T)ny - (T)0.5, nx, ny);
    // displacement fields are evaluated at the four corners only
    auto disp = [&](int idx, int comp) -> T {
        int r = idx / nx, col = idx - r * nx;
        T w = sin((((T)r + (T)0.5) / (T)ny) * pi);
        T val = comp == 0 ? (flow[idx] - (T)col - (T)1) / (T)nx
                          : (flow[P + idx] - (T)r - (T)1) / (T)ny;
        if (val >= (T)1)
            val = val - floor(val);
        if (val <= (T)-1)
            val = val + floor(-val);
        if (val >= (T)0.5)
            val = val - (T)1;
        if (val <= (T)-0.5)
            val = val + (T)1;
        return val * w;
    };
    T su = (disp(c.i22, 0) * c.dx2 + disp(c.i12, 0) * c.dx1) * c.dy2 +
           (disp(c.i21, 0) * c.dx2 + disp(c.i11, 0) * c.dx1) * c.dy1;
    T sv = (disp(c.i22, 1) * c.dx2 + disp(c.i12, 1) * c.dx1) * c.dy2 +
           (disp(c.i21, 1) * c.dx2 + disp(c.i11, 1) * c.dx1) * c.dy1;
    if (inverse) {
        su = -su;
        sv = -sv;
    }
    u = u + su;
    v = v + sv;
    if (v < (T)0) {
        u = u + (T)0.5;
        v = -v;
    }
    if (v > (T)1) {
        u = u + (T)0.5;
        v = (T)2 - v;
    }
    u = u - floor(u);
    v = v < (T)0 ? (T)0 : (v > (T)1 ? (T)1 : v);
    T ph = u * ((T)2 * pi), th = v * pi;
    T st = sin(th);
    T wx = st * cos(ph), wy = st * sin(ph), wz = cos(th);
    T eps = sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920928955078125e-07;
    T nrm = sqrt(wx * wx + wy * wy + wz * wz);
    nrm = nrm < eps ? eps : nrm;
    out[3 * i] = wx / nrm;
    out[3 * i + 1] = wy / nrm;
    out[3 * i + 2] = wz / nrm;
}

// Un-rotate the pole-frame result by m (row-major 3x3, p·mᵀ) and blend with
// the main result by colatitude weights.  gpu_pipeline.py:248-271, :1294-1310
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_average_xz(const T *__restrict__ avg, const T *__restrict__ main_, int n, double m00,
             double m01, double m02, double m10, double m11, double m12, double m20, double m21,
             double m22, float *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double pi = 3.14159265358979323846;
    double a0 = (double)avg[3 * i], a1 = (double)avg[3 * i + 1], a2 = (double)avg[3 * i + 2];
    double x[3] = {a0 * m00 + a1 * m01 + a2 * m02, a0 * m10 + a1 * m11 + a2 * m12,
                   a0 * m20 + a1 * m21 + a2 * m22};
    double z[3] = {(double)main_[3 * i], (double)main_[3 * i + 1], (double)main_[3 * i + 2]};
    double px = acos(fmin(fmax(x[0], -1.0), 1.0)) / pi;
    double pz = acos(fmin(fmax(z[2], -1.0), 1.0)) / pi;
    double wx = fmax(exp(-((2.0 * px - 1.0) * (2.0 * px - 1.0)) / 0.1), 1e-19);
    double wz = fmax(exp(-((2.0 * pz - 1.0) * (2.0 * pz - 1.0)) / 0.1), 1e-19);
    double t = wx + wz;
    double r[3];
    for (int d = 0; d < 3; ++d)
        r[d] = (wx * x[d] + wz * z[d]) / t;
    double nrm = fmax(sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]), 2.220446049250313e-16);
    for (int d = 0; d < 3; ++d)
        out[3 * i + d] = (float)(r[d] / nrm);
}

// float32 unit vectors, |p| = sqrt(fma(z, z, fma(y, y, x·x))) clamped by
// FLT_EPSILON (torch's CPU float32 norm).  surface_warp.py:518-522
__global__ void __launch_bounds__(kBlock)
k_normalise_f32(const float *__restrict__ in, int n, float *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float x = in[3 * i], y = in[3 * i + 1], z = in[3 * i + 2];
    float nr = fmaxf(norm3f(x, y, z), 1.1920928955078125e-07f);
    out[3 * i] = x / nr;
    out[3 * i + 1] = y / nr;
    out[3 * i + 2] = z / nr;
}

// Sphere preparation of the pipeline (gpu_pipeline.py:274-305, :373-378).
// The centroid is a strictly sequential float64 sum in vertex order (the
// reference's loop), divided by n: centroid[c] = (Σ_i f64(v[i][c])) / n.  One
// workgroup per sphere: waves 1-3 stream the next chunk of kCentroidChunk
// vertices (coalesced) into one LDS buffer while lanes 0-2 of wave 0 add the
// current chunk, one coordinate per lane, from the other.
constexpr int kCentroidThreads = 256;
constexpr int kCentroidChunk = 256;  // vertices per LDS buffer

__global__ void __launch_bounds__(kCentroidThreads)
k_sequential_centroids(SphereList list, double *__restrict__ centroid)
{
    __shared__ float tile[2][3 * kCentroidChunk];
    const int sphere = blockIdx.x;
    const float *__restrict__ v = list.ptr[sphere];
    const int n = list.n[sphere];
    const long long total = 3LL * n;
    const int tid = threadIdx.x;
    const int chunks = (n + kCentroidChunk - 1) / kCentroidChunk;
    for (int j = tid; j < 3 * kCentroidChunk; j += kCentroidThreads)
        tile[0][j] = j < total ? v[j] : 0.0f;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < chunks; ++k) {
        const int p = k & 1;
        if (tid >= 64 && k + 1 < chunks) {
            const long long base = (long long)(k + 1) * 3 * kCentroidChunk;
            for (int j = tid - 64; j < 3 * kCentroidChunk; j += kCentroidThreads - 64)
                tile[p ^ 1][j] = base + j < total ? v[base + j] : 0.0f;
        }
        if (tid < 3) {
            const int m = min(kCentroidChunk, n - k * kCentroidChunk);
            if (m == kCentroidChunk) {
#pragma unroll 16
                for (int q = 0; q < kCentroidChunk; ++q)
                    s += (double)tile[p][3 * q + tid];
            } else {
                for (int q = 0; q < m; ++q)
                    s += (double)tile[p][3 * q + tid];
            }
        }
        __syncthreads();
    }
    if (tid < 3)
        centroid[3 * sphere + tid] = s / (double)n;
}

// Per vertex: with `centroid`, p = f32(f64(v) - centroid), then
// f32(f64(p) / |f64(p)|) (flagging zero or non-finite lengths in *bad);
// with `m` (row-major 3x3), f32(f64(p) · mᵀ).
__global__ void __launch_bounds__(kBlock)
k_normalise_rotate(const float *__restrict__ v, int n, const double *__restrict__ centroid,
                   Mat3 m, const double *__restrict__ m_dev, int rotate, float *__restrict__ out,
                   int *__restrict__ bad)
{
    if (m_dev)
        for (int k = 0; k < 9; ++k)
            m.a[k] = m_dev[k];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
    if (centroid) {
        const float ex = (float)((double)x - centroid[0]);
        const float ey = (float)((double)y - centroid[1]);
        const float ez = (float)((double)z - centroid[2]);
        const double c0 = ex, c1 = ey, c2 = ez;
        const double len = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        if (!(len > 0.0) || !isfinite(len))
            *bad = 1;
        const double inv = 1.0 / len;
        x = (float)(c0 * inv);
        y = (float)(c1 * inv);
        z = (float)(c2 * inv);
    }
    if (rotate) {
        const double px = x, py = y, pz = z;
        const double rx = px * m.a[0] + py * m.a[1] + pz * m.a[2];
        const double ry = px * m.a[3] + py * m.a[4] + pz * m.a[5];
        const double rz = px * m.a[6] + py * m.a[7] + pz * m.a[8];
        x = (float)rx;
        y = (float)ry;
        z = (float)rz;
    }
    out[3 * i] = x;
    out[3 * i + 1] = y;
    out[3 * i + 2] = z;
}

// ------------------------------------------------------------- launchers

int launch_sphere_centroids(const SphereList &list, double *centroid, hipStream_t s)
{
    if (list.count <= 0)
        return kOk;
    hipLaunchKernelGGL(k_sequential_centroids, dim3(list.count), dim3(kCentroidThreads), 0, s,
                       list, centroid);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_sphere_prepare(const float *v, int n, const double *centroid, const double *m_host,
                          const double *m_dev, float *out, int *bad, hipStream_t s)
{
    if (n <= 0)
        return kOk;
    Mat3 m{};
    if (m_host)
        for (int k = 0; k < 9; ++k)
            m.a[k] = m_host[k];
    hipLaunchKernelGGL(k_normalise_rotate, dim3(grid_for(n)), dim3(kBlock), 0, s, v, n, centroid,
                       m, m_dev, (m_host || m_dev) ? 1 : 0, out, bad);
    CSG_LAUNCH_CHECK();
    return kOk;
}

// Rotation matrices of Euler-angle triples, rotation_matrix_np's arithmetic
// (Rz (Ry Rx) assembled column-major, read row-major, each product sum left to
// right, no contraction), one thread per triple.
__global__ void k_rotation_matrices(const double *__restrict__ angles, int count,
                                    double *__restrict__ out)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count)
        return;
    const double a = angles[3 * t], b = angles[3 * t + 1], g = angles[3 * t + 2];
    const double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
    const double rx[9] = {1.0, 0.0, 0.0, 0.0, ca, sa, 0.0, -sa, ca};
    const double ry[9] = {cb, 0.0, sb, 0.0, 1.0, 0.0, -sb, 0.0, cb};
    const double rz[9] = {cg, sg, 0.0, -sg, cg, 0.0, 0.0, 0.0, 1.0};
    double first[9], res[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            first[r + 3 * c] = (ry[r] * rx[3 * c] + ry[r + 3] * rx[1 + 3 * c]) +
                               ry[r + 6] * rx[2 + 3 * c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            res[r + 3 * c] = (rz[r] * first[3 * c] + rz[r + 3] * first[1 + 3 * c]) +
                             rz[r + 6] * first[2 + 3 * c];
    for (int k = 0; k < 9; ++k)
        out[9 * (size_t)t + k] = res[k];
}

int launch_rotation_matrices(const double *angles, int count, double *out, hipStream_t s)
{
    if (count <= 0)
        return kOk;
    hipLaunchKernelGGL(k_rotation_matrices, dim3(grid_for(count)), dim3(kBlock), 0, s, angles,
                       count, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_normalise_f32(const float *in, int n, float *out, hipStream_t s)
{
    if (n <= 0)
        return kOk;
    hipLaunchKernelGGL(k_normalise_f32, dim3(grid_for(n)), dim3(kBlock), 0, s, in, n, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template <typename T>
int launch_resample(const T *x, const int *idx, const double *w, int n, T *out, hipStream_t s)
{
    if (n <= 0)
        return kOk;
    hipLaunchKernelGGL(k_resample<T>, dim3(grid_for((size_t)3 * n)), dim3(kBlock), 0, s, x, idx,
                       w, n, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}
template int launch_resample<float>(const float *, const int *, const double *, int, float *,
                                    hipStream_t);
template int launch_resample<double>(const double *, const int *, const double *, int, double *,
                                     hipStream_t);

int launch_edge_sum(const float *v, const int *faces, int nf, double *partial, int *nblocks,
                    hipStream_t s)
{
    *nblocks = (int)grid_for(nf);
    hipLaunchKernelGGL(k_edge_sum, dim3(*nblocks), dim3(kBlock), 0, s, v, faces, nf, partial);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_smooth_geometry(const float *cur, const int *off, const int *nbr, int n, float *nxt,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_smooth_geometry, dim3(grid_for(n)), dim3(kBlock), 0, s, cur, off, nbr, n,
                       nxt);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_smooth_geometry_batch(float *buf0, float *buf1, const int *off, const int *nbr,
                                 int n, const SmoothBatch &sb, int it, hipStream_t s)
{
    hipLaunchKernelGGL(k_smooth_geometry_batch, dim3(grid_for(n), sb.count), dim3(kBlock), 0, s,
                       buf0, buf1, off, nbr, n, sb, it);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_smooth_geometry4(const float *v, float4 *buf0, float4 *buf1, const int *off,
                            const int *nbr, int n, const SmoothBatch &sb, int most, float *out,
                            hipStream_t s)
{
    dim3 g(grid_for(n), sb.count);
    hipLaunchKernelGGL(k_pack4, g, dim3(kBlock), 0, s, v, n, buf0);
    CSG_LAUNCH_CHECK();
    for (int it = 0; it < most; ++it) {
        hipLaunchKernelGGL(k_smooth_geometry4, g, dim3(kBlock), 0, s, buf0, buf1, off, nbr, n, sb,
                           it);
        CSG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_unpack4, g, dim3(kBlock), 0, s, buf0, buf1, n, sb, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_smooth_geometry_patched(const float *v, float4 *buf0, float4 *buf1, int n,
                                   const SmoothPatches &pp, const SmoothBatch &sb, int most,
                                   size_t lds_bytes, float *out, hipStream_t s)
{
    static std::once_flag once;
    static size_t limit = 65536;
    std::call_once(once, [] {
        // the dynamic allocation may use what the kernel's static tables leave
        int dev = 0, per_block = 0;
        hipFuncAttributes attr{};
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
                hipSuccess &&
            hipFuncGetAttributes(&attr, (const void *)k_smooth_geometry_patch) == hipSuccess &&
            per_block > (int)attr.sharedSizeBytes) {
            const int dynamic = per_block - (int)attr.sharedSizeBytes;
            if (hipFuncSetAttribute((const void *)k_smooth_geometry_patch,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    dynamic) == hipSuccess)
                limit = (size_t)dynamic;
        }
        (void)hipGetLastError();
    });
    if (lds_bytes > limit)
        return fail(kInvalidArgument, "smoothing patches need %zu bytes of LDS, the device has %zu",
                    lds_bytes, limit);
    dim3 g(grid_for(n), sb.count);
    hipLaunchKernelGGL(k_pack4_perm, g, dim3(kBlock), 0, s, v, n, pp.perm, buf0);
    CSG_LAUNCH_CHECK();
    int launches = 0;
    for (int it0 = 0; it0 < most; it0 += pp.rings, ++launches) {
        const float4 *in = (launches & 1) ? buf1 : buf0;
        float4 *o = (launches & 1) ? buf0 : buf1;
        // reference bytes per vertex and iteration: read and write the float32
        // position (12 + 12, counted as 16 + 16 for the padded form) and the
        // neighbour list (4 bytes each, mean degree 6 on a closed triangle mesh)
        double iters = 0.0;
        for (int m = 0; m < sb.count; ++m)
            iters += std::max(0, std::min(pp.rings, (sb.diters ? most : sb.iters[m]) - it0));
        TimedLaunch timer(kTimedSmoothGeometry, s, iters * n * (32.0 + 4.0 * 6.0));
        hipExtLaunchKernelGGL(k_smooth_geometry_patch, dim3(pp.count, sb.count),
                           dim3(kPatchThreads), lds_bytes, s, timer.start(), timer.stop(), 0, in, o, n, pp, sb, it0);
        CSG_LAUNCH_CHECK();
        timer.done();
    }
    // every mesh's result is in the buffer the last launch wrote
    hipLaunchKernelGGL(k_unpack4_perm, g, dim3(kBlock), 0, s, (launches & 1) ? buf1 : buf0, n,
                       pp.perm, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_vertex_normals(const float *v, const int *faces, const int *inc_off, const int *inc,
                          int n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_vertex_normals, dim3(grid_for(n)), dim3(kBlock), 0, s, v, faces, inc_off,
                       inc, n, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_type5_base(const float *geom, const float *sm, const float *nr, int n, double *out,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_type5_base, dim3(grid_for(n)), dim3(kBlock), 0, s, geom, sm, nr, n, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_curvature_type2(const float *geom, const float *nr, int n, const int *off,
                           const int *nbr, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_curvature_type2, dim3(grid_for(n)), dim3(kBlock), 0, s, geom, nr, n, off,
                       nbr, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

struct FwhmList {
    double f[kMaxHeatFwhm];
};

// One thread: the products and quotients are rounded one by one (no
// contraction), as the host's double arithmetic rounds them.
__global__ void k_heat_parameters(const double *__restrict__ partial, int nblocks, int nf,
                                  FwhmList fw, int nfw, const int *__restrict__ root_n,
                                  const double *__restrict__ root_v, int nroots,
                                  int *__restrict__ iters, double *__restrict__ sigma,
                                  double *__restrict__ em_out)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double em;
    if (nblocks < 0) {  // partial holds the edge mean itself
        em = partial[0];
    } else {
        double total = 0.0;
        for (int k = 0; k < nblocks; ++k)
            total += partial[k];
        em = total / (3.0 * (double)nf);
    }
    if (em_out)
        em_out[0] = em;
    for (int j = 0; j < nfw; ++j) {
        const double f = fw.f[j];
        const double raw = f * f * 0.541011 / (em * em);
        const double c = ceil(raw);
        int it = c >= 2147483647.0 ? 2147483647 : (int)c;
        it = it < 1 ? 1 : it;
        iters[j] = it;
        double root = sqrt((double)it);
        int lo = 0, hi = nroots;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (root_n[mid] < it)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < nroots && root_n[lo] == it)
            root = root_v[lo];
        sigma[j] = f * 0.7355345 / root;
    }
}

int launch_heat_parameters(const double *partial, int nblocks, int nf, const double *fwhm,
                           int nfw, const int *root_n, const double *root_v, int nroots,
                           int *iters, double *sigma, double *em, hipStream_t s)
{
    FwhmList fw{};
    for (int j = 0; j < nfw; ++j)
        fw.f[j] = fwhm[j];
    hipLaunchKernelGGL(k_heat_parameters, dim3(1), dim3(64), 0, s, partial, nblocks, nf, fw, nfw,
                       root_n, root_v, nroots, iters, sigma, em);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_heat_weights(const float *geom, const int *off, const int *nbr, int n,
                        const HeatBatch &hb, int squared_norm, size_t wstride, double *w,
                        double *wsum, hipStream_t s)
{
    hipLaunchKernelGGL(k_heat_weights, dim3(grid_for(n), hb.nb), dim3(kBlock), 0, s, geom, off,
                       nbr, n, hb, squared_norm, wstride, w, wsum);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_smooth_values_patched(const double *cur, const int *goff, const double *w, int n,
                                 size_t wstride, const HeatBatch &hb, const SmoothPatches &pp,
                                 int it0, size_t lds_bytes, double *nxt, hipStream_t s)
{
    static std::once_flag once;
    static int limit = 0;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess)
            limit = 0;
        if (limit > 0)
            (void)hipFuncSetAttribute((const void *)k_smooth_values_patch,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, limit);
        (void)hipGetLastError();
    });
    if (lds_bytes > (size_t)limit)
        return fail(kInvalidArgument, "value smoothing patches need %zu bytes of LDS (limit %d)",
                    lds_bytes, limit);
    hipLaunchKernelGGL(k_smooth_values_patch, dim3(pp.count, hb.nb), dim3(kPatchThreads),
                       lds_bytes, s, cur, nxt, n, goff, w, wstride, pp, hb, it0);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_smooth_values(const double *cur, const int *off, const int *nbr, const double *w,
                         const double *wsum, int n, size_t wstride, const HeatBatch &hb,
                         int iteration, double *nxt, hipStream_t s)
{
    hipLaunchKernelGGL(k_smooth_values, dim3(grid_for(n), hb.nb), dim3(kBlock), 0, s, cur, off,
                       nbr, w, wsum, n, wstride, hb, iteration, nxt);
    CSG_LAUNCH_CHECK();
    return kOk;
}

constexpr int kMinmaxBlocks = 64;

template <typename T>
int launch_minmax_normalise(const double *x, size_t n, int nb, size_t stride, double floor_,
                            double *partial, T *out, size_t ostride, hipStream_t s)
{
    hipLaunchKernelGGL(k_minmax_partial, dim3(kMinmaxBlocks, nb), dim3(kBlock), 0, s, x, n,
                       stride, partial, kMinmaxBlocks);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_minmax_apply<T>, dim3(kMinmaxBlocks, nb), dim3(kBlock), 0, s, x, n,
                       stride, partial, kMinmaxBlocks, floor_, out, ostride);
    CSG_LAUNCH_CHECK();
    return kOk;
}
template int launch_minmax_normalise<float>(const double *, size_t, int, size_t, double,
                                            double *, float *, size_t, hipStream_t);
template int launch_minmax_normalise<double>(const double *, size_t, int, size_t, double,
                                             double *, double *, size_t, hipStream_t);

int launch_sheet_gather(const double *vals, size_t vstride, const int *idx, const double *w,
                        int m, int nb, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sheet_gather, dim3(grid_for(m), nb), dim3(kBlock), 0, s, vals, vstride,
                       idx, w, m, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}


template <typename T>
int launch_apply_flow(const float *unit, int n, const T *flow, int nx, int ny, int inverse,
                      T *out, hipStream_t s)
{
    if (n <= 0)
        return kOk;
    hipLaunchKernelGGL(k_apply_flow<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, unit, n, flow, nx,
                       ny, inverse, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}
template int launch_apply_flow<float>(const float *, int, const float *, int, int, int, float *,
                                      hipStream_t);
template int launch_apply_flow<double>(const float *, int, const double *, int, int, int,
                                       double *, hipStream_t);

template <typename T>
int launch_average_xz(const T *avg, const T *main_, int n, const double *m, float *out,
                      hipStream_t s)
{
    if (n <= 0)
        return kOk;
    hipLaunchKernelGGL(k_average_xz<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, avg, main_, n,
                       m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], out);
    CSG_LAUNCH_CHECK();
    return kOk;
}
template int launch_average_xz<float>(const float *, const float *, int, const double *, float *,
                                      hipStream_t);
template int launch_average_xz<double>(const double *, const double *, int, const double *,
                                       float *, hipStream_t);

}  // namespace csg
