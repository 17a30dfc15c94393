// Depth potential of the rotation feature on gfx950, for a batch of meshes
// that share one topology (source and target stencils):
//   per face:   stable-normal terms, mixed Voronoi area terms, the 9
//               cotangent entries (float64);
//   per vertex: normals and areas gathered from incident (face, corner)
//               entries in the reference's index_add_ order;
//   system:     duplicate (row, col) entries merged into one CSR entry (the
//               sum of its contributions in COO order), Laplacian·points,
//               mean curvature, right-hand side, diagonal;
//   solve:      colour-ordered SOR, all max_iter × colours sweeps in ONE
//               persistent launch (grid capped at what the device holds at
//               once) with a grid barrier between colours (the sweep order of
//               the reference, 1000 × ~7 colours, would otherwise be ~7000
//               dependent launches).
//
// Reference semantics: src/cat_surface_gpu/rotation_feature.py:24-317.
#include "depth_kernels.h"
#include "kernel_timing.h"
#include "xcd_sync.h"

#include <atomic>

namespace csg {

__global__ void __launch_bounds__(kBlock)
k_depth_face_terms(const float *__restrict__ pts, const int *__restrict__ faces, int n, int nf,
                   double *__restrict__ fnormal, double *__restrict__ farea,
                   double *__restrict__ coo)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    const int b = blockIdx.y;
    pts += (size_t)b * 3 * n;
    fnormal += (size_t)b * 9 * nf;
    farea += (size_t)b * 3 * nf;
    coo += (size_t)b * 9 * nf;
    const double eps = 2.220446049250313e-16;
    const int *fv = faces + 3 * f;
    double p[3][3];
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d)
            p[c][d] = (double)pts[3 * fv[c] + d];
    double v1[3], v2[3], v3[3];
    for (int d = 0; d < 3; ++d) {
        v1[d] = p[1][d] - p[0][d];
        v2[d] = p[2][d] - p[1][d];
        v3[d] = p[0][d] - p[2][d];
    }
    auto dot = [](const double *a, const double *c) { return a[0] * c[0] + a[1] * c[1] + a[2] * c[2]; };
    auto cross = [](const double *a, const double *c, double *o) {
        o[0] = a[1] * c[2] - a[2] * c[1];
        o[1] = a[2] * c[0] - a[0] * c[2];
        o[2] = a[0] * c[1] - a[1] * c[0];
    };
    auto norm = [&](const double *a) { return sqrt(dot(a, a)); };
    // _stable_normals (rotation_feature.py:38-73)
    double u1[3], u2[3], u3[3];
    double n1 = fmax(norm(v1), eps), n2 = fmax(norm(v2), eps), n3 = fmax(norm(v3), eps);
    for (int d = 0; d < 3; ++d) {
        u1[d] = v1[d] / n1;
        u2[d] = v2[d] / n2;
        u3[d] = v3[d] / n3;
    }
    double c31[3], c23[3], c12[3];
    cross(u3, u1, c31);
    cross(u2, u3, c23);
    cross(u1, u2, c12);
    double m31 = fmax(norm(c31), eps), m23 = fmax(norm(c23), eps), m12 = fmax(norm(c12), eps);
    double w0 = (1.0 - dot(u3, u1)) / m31;
    double w1 = (1.0 - dot(u1, u2)) / m12;
    double w2 = (1.0 - dot(u2, u3)) / m23;
    double tn[3] = {c12[0] / m12, c12[1] / m12, c12[2] / m12};
    for (int d = 0; d < 3; ++d) {
        fnormal[9 * (size_t)f + d] = tn[d] * w0;
        fnormal[9 * (size_t)f + 3 + d] = tn[d] * w1;
        fnormal[9 * (size_t)f + 6 + d] = tn[d] * w2;
    }
    // _mixed_voronoi_areas (rotation_feature.py:76-128)
    double cr[3];
    cross(v1, v2, cr);
    double area = fmax(norm(cr), eps);
    double d31 = dot(v3, v1), d12 = dot(v1, v2), d23 = dot(v2, v3), d13 = dot(v1, v3);
    bool bad0 = d31 > 0.0, bad1 = d12 > 0.0, bad2 = d23 > 0.0;
    double l1 = dot(v1, v1), l2 = dot(v2, v2), l3 = dot(v3, v3);
    double a0, a1, a2;
    if (bad0 || bad1 || bad2) {
        a0 = a1 = a2 = 0.125 * area;
    } else {
        a0 = -0.125 * (l3 * d12 + l1 * d23) / area;
        a1 = -0.125 * (l2 * d13 + l1 * d23) / area;
        a2 = -0.125 * (l2 * d13 + l3 * d12) / area;
    }
    farea[3 * (size_t)f] = a0 + (bad0 ? 0.125 * area : 0.0);
    farea[3 * (size_t)f + 1] = a1 + (bad1 ? 0.125 * area : 0.0);
    farea[3 * (size_t)f + 2] = a2 + (bad2 ? 0.125 * area : 0.0);
    // _system_entries (rotation_feature.py:131-186)
    double e23 = d23 / area, e13 = d13 / area, e12 = d12 / area;
    double vals[9] = {-0.5 * (e23 + e12), -0.5 * (e13 + e23), -0.5 * (e12 + e13),
                      0.5 * e23, 0.5 * e23, 0.5 * e13, 0.5 * e13, 0.5 * e12, 0.5 * e12};
    for (int k = 0; k < 9; ++k)
        coo[(size_t)k * nf + f] = vals[k];
}

__global__ void __launch_bounds__(kBlock)
k_depth_vertex_terms(const double *__restrict__ fnormal, const double *__restrict__ farea,
                     const int *__restrict__ inc_off, const int *__restrict__ inc, int n, int nf,
                     double *__restrict__ normals, double *__restrict__ areas)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int b = blockIdx.y;
    fnormal += (size_t)b * 9 * nf;
    farea += (size_t)b * 3 * nf;
    const double eps = 2.220446049250313e-16;
    double nx = 0.0, ny = 0.0, nz = 0.0, a = 0.0;
    for (int k = inc_off[i]; k < inc_off[i + 1]; ++k) {
        int fc = inc[k];
        const double *fn = fnormal + 3 * (size_t)fc;
        nx += fn[0];
        ny += fn[1];
        nz += fn[2];
        a += farea[fc];
    }
    double nr = fmax(sqrt(nx * nx + ny * ny + nz * nz), eps);
    double *o = normals + (size_t)b * 3 * n;
    o[3 * i] = nx / nr;
    o[3 * i + 1] = ny / nr;
    o[3 * i + 2] = nz / nr;
    areas[(size_t)b * n + i] = (!isfinite(a) || a == 0.0) ? 1.0 : a;
}

// Merged entry values: lap = Σ face contributions, sys = lap + Σ alpha·area.
__global__ void __launch_bounds__(kBlock)
k_merged_values(const double *__restrict__ coo, const int *__restrict__ contrib_off,
                const int *__restrict__ contrib, const int *__restrict__ row_of_entry_unused,
                int nnz, int nf, int n, const double *__restrict__ areas, double alpha,
                double *__restrict__ lap, double *__restrict__ sys)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nnz)
        return;
    const int b = blockIdx.y;
    coo += (size_t)b * 9 * nf;
    areas += (size_t)b * n;
    const int nface9 = 9 * nf;
    double l = 0.0, s = 0.0;
    for (int k = contrib_off[m]; k < contrib_off[m + 1]; ++k) {
        int e = contrib[k];
        if (e < nface9) {
            l += coo[e];
            s += coo[e];
        } else {
            s += alpha * areas[e - nface9];
        }
    }
    lap[(size_t)b * nnz + m] = l;
    sys[(size_t)b * nnz + m] = s;
}

// Laplacian·points → mean curvature; diagonal of the system.
__global__ void __launch_bounds__(kBlock)
k_depth_rhs_parts(const float *__restrict__ pts, const int *__restrict__ row_off,
                  const int *__restrict__ cols, const double *__restrict__ lap,
                  const double *__restrict__ sys, const double *__restrict__ normals,
                  const double *__restrict__ areas, int n, int nnz, double *__restrict__ mc,
                  double *__restrict__ diag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int b = blockIdx.y;
    pts += (size_t)b * 3 * n;
    lap += (size_t)b * nnz;
    sys += (size_t)b * nnz;
    double lx = 0.0, ly = 0.0, lz = 0.0, dg = 0.0;
    for (int k = row_off[i]; k < row_off[i + 1]; ++k) {
        int j = cols[k];
        double v = lap[k];
        if (j == i)
            dg += sys[k];
        lx += v * (double)pts[3 * j];
        ly += v * (double)pts[3 * j + 1];
        lz += v * (double)pts[3 * j + 2];
    }
    const double *nr = normals + (size_t)b * 3 * n;
    double a = areas[(size_t)b * n + i];
    double h = 0.5 * ((2.0 * lx) * nr[3 * i] + (2.0 * ly) * nr[3 * i + 1] +
                      (2.0 * lz) * nr[3 * i + 2]) / a;
    mc[(size_t)b * n + i] = isfinite(h) ? h : 0.0;
    diag[(size_t)b * n + i] = fmax(dg, 2.220446049250313e-16);
}

// Σ mc·area and Σ area per block.
__global__ void __launch_bounds__(kBlock)
k_weighted_sums(const double *__restrict__ mc, const double *__restrict__ areas, int n,
                double *__restrict__ partial, int nblocks)
{
    __shared__ double red[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    double a = 0.0, w = 0.0;
    if (i < n) {
        a = mc[(size_t)b * n + i] * areas[(size_t)b * n + i];
        w = areas[(size_t)b * n + i];
    }
    double ra = block_sum(a, red);
    double rw = block_sum(w, red);
    if (threadIdx.x == 0) {
        partial[2 * ((size_t)b * nblocks + blockIdx.x)] = ra;
        partial[2 * ((size_t)b * nblocks + blockIdx.x) + 1] = rw;
    }
}

// rhs = (mc - factor)·area, factor = Σ mc·a / Σ a; x = 0.
__global__ void __launch_bounds__(kBlock)
k_depth_rhs(const double *__restrict__ mc, const double *__restrict__ areas, int n,
            const double *__restrict__ partial, int nblocks, double *__restrict__ rhs,
            double *__restrict__ x)
{
    __shared__ double factor;
    const int b = blockIdx.y;
    if (threadIdx.x == 0) {
        double a = 0.0, w = 0.0;
        for (int k = 0; k < nblocks; ++k) {
            a += partial[2 * ((size_t)b * nblocks + k)];
            w += partial[2 * ((size_t)b * nblocks + k) + 1];
        }
        factor = a / w;
    }
    __syncthreads();
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        size_t o = (size_t)b * n + i;
        rhs[o] = (mc[o] - factor) * areas[o];
        x[o] = 0.0;
    }
}

// Colour-ordered copies of the system for the sweeps: ELL values, rhs and
// diagonal indexed by solve-order position r (vertex colour_verts[r]).
__global__ void __launch_bounds__(kBlock)
k_colour_order(const int *__restrict__ verts, const int *__restrict__ ell_src, int W, int n,
               const double *__restrict__ sys, int nnz, const double *__restrict__ rhs,
               const double *__restrict__ diag, double *__restrict__ ell_val,
               double *__restrict__ rhs_c, double *__restrict__ diag_c)
{
    const size_t total = (size_t)W * n + n;
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total)
        return;
    const int b = blockIdx.y;
    if (q < (size_t)W * n) {
        int e = ell_src[q];
        ell_val[(size_t)b * W * n + q] = e >= 0 ? sys[(size_t)b * nnz + e] : 0.0;
    } else {
        int r = (int)(q - (size_t)W * n);
        int i = verts[r];
        rhs_c[(size_t)b * n + r] = rhs[(size_t)b * n + i];
        diag_c[(size_t)b * n + r] = diag[(size_t)b * n + i];
    }
}

// Grid-wide barrier for a co-resident grid, as a two-level tree: blocks
// arrive in groups of kBarrierGroup on their own counter, the last arrival of
// each group arrives at the root, and the last root arrival advances the
// generation word that everyone polls.  Arrivals on one counter serialise, so
// the tree keeps the critical path to about 2·kBarrierGroup atomic round
// trips.  Only thread 0 of a block touches the counters, after the block's
// __syncthreads.  Memory ordering uses as few agent-scope cache operations as
// the chain needs: each arrival is a release (publishes the block's stores);
// a group's last arrival and the root's last arrival take an acquire fence
// before their own release, so the chain arrival -> group -> root ->
// generation carries every block's stores; waiters poll with relaxed loads and
// take one acquire fence after the generation changes.  Returns false if the
// block gave up after kSpinLimit polls; the caller then exits, so every wave
// terminates.
constexpr long kSpinLimit = 1L << 22;
constexpr int kBarrierGroup = 8;

__device__ bool grid_barrier(unsigned *bar, int nblocks, int *timeout)
{
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        const int groups = (nblocks + kBarrierGroup - 1) / kBarrierGroup;
        unsigned *root = bar + (size_t)groups * kBarrierSlotWords;
        unsigned *gen_p = root + kBarrierSlotWords;
        const unsigned gen = __hip_atomic_load(gen_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int g = blockIdx.x / kBarrierGroup;
        const int gsize = min(kBarrierGroup, nblocks - g * kBarrierGroup);
        unsigned *cnt = bar + (size_t)g * kBarrierSlotWords;
        bool released = false;
        unsigned a = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (a == (unsigned)gsize - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned r =
                __hip_atomic_fetch_add(root, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (r == (unsigned)groups - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __hip_atomic_store(root, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(gen_p, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                released = true;
            }
        }
        if (!released) {
            long spins = 0;
            while (__hip_atomic_load(gen_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++spins > kSpinLimit) {
                    __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
    return ok != 0;
}

// All SOR sweeps: for it < max_iter, for each colour, every vertex of the
// colour (in every batch mesh) takes x_i += ω (rhs_i - Σ_j A_ij x_j) / diag_i;
// a grid barrier separates colours.  Row data are read in solve order from
// the colour-ordered ELL copy, so a sweep's only dependent loads are the x
// gathers.  rotation_feature.py:296-311
__global__ void __launch_bounds__(kSorBlock)
k_sor_persistent(const int *__restrict__ verts, SorColours co, const int *__restrict__ ell_cols,
                 const double *__restrict__ ell_val, const double *__restrict__ rhs_c,
                 const double *__restrict__ diag_c, int W, double omega, int n, int batch,
                 int max_iter, double *x, unsigned *bar, int *timeout)
{
    const int nthreads = gridDim.x * blockDim.x;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t wn = (size_t)W * n;
    for (int it = 0; it < max_iter; ++it) {
        for (int c = 0; c < co.count; ++c) {
            const int start = co.off[c], cnt = co.off[c + 1] - start;
            for (int t = tid; t < cnt * batch; t += nthreads) {
                const int b = t / cnt;
                const int r = start + (t - b * cnt);
                const int i = verts[r];
                const double *vb = ell_val + (size_t)b * wn;
                double *xb = x + (size_t)b * n;
                double s = 0.0;
                for (int k = 0; k < W; ++k)
                    s += vb[(size_t)k * n + r] * xb[ell_cols[(size_t)k * n + r]];
                const size_t o = (size_t)b * n + r;
                xb[i] = xb[i] + omega * (rhs_c[o] - s) / diag_c[o];
            }
            if (!grid_barrier(bar, gridDim.x, timeout))
                return;
        }
    }
}

// The same sweeps with every row a thread owns held in registers: when the
// grid has at least as many threads as the largest colour has rows (times the
// batch), thread t owns row t of every colour, so its ELL values, columns,
// rhs and diagonal are loaded once and a sweep's only memory traffic is the x
// gathers and the x update.  Identical arithmetic and order as
// k_sor_persistent.  CMAX / WMAX bound the colours and the ELL width; the
// block is two waves per SIMD so that the rows fit in registers.
constexpr int kSorRegBlock = 512;
constexpr int kSorRegColours = 7, kSorRegWidth = 7;
template <int CMAX, int WMAX>
__global__ void __launch_bounds__(kSorRegBlock)
k_sor_registers(const int *__restrict__ verts, SorColours co, const int *__restrict__ ell_cols,
                const double *__restrict__ ell_val, const double *__restrict__ rhs_c,
                const double *__restrict__ diag_c, int W, double omega, int n, int batch,
                int max_iter, double *x, unsigned *bar, int *timeout)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t wn = (size_t)W * n;
    bool act[CMAX];
    int vi[CMAX];
    int col[CMAX][WMAX];
    double val[CMAX][WMAX], rh[CMAX], dg[CMAX];
#pragma unroll
    for (int c = 0; c < CMAX; ++c) {
        act[c] = false;
        vi[c] = 0;
        rh[c] = 0.0;
        dg[c] = 1.0;
#pragma unroll
        for (int k = 0; k < WMAX; ++k) {
            col[c][k] = 0;
            val[c][k] = 0.0;
        }
        if (c < co.count) {
            const int start = co.off[c], cnt = co.off[c + 1] - start;
            if (tid < cnt * batch) {
                const int b = tid / cnt;
                const int r = start + (tid - b * cnt);
                act[c] = true;
                vi[c] = b * n + verts[r];
                rh[c] = rhs_c[(size_t)b * n + r];
                dg[c] = diag_c[(size_t)b * n + r];
#pragma unroll
                for (int k = 0; k < WMAX; ++k) {
                    if (k < W) {
                        val[c][k] = ell_val[(size_t)b * wn + (size_t)k * n + r];
                        col[c][k] = b * n + ell_cols[(size_t)k * n + r];
                    }
                }
            }
        }
    }
    for (int it = 0; it < max_iter; ++it) {
#pragma unroll
        for (int c = 0; c < CMAX; ++c) {
            if (c >= co.count)
                break;
            if (act[c]) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < WMAX; ++k)
                    if (k < W)
                        s += val[c][k] * x[col[c][k]];
                x[vi[c]] = x[vi[c]] + omega * (rh[c] - s) / dg[c];
            }
            if (!grid_barrier(bar, gridDim.x, timeout))
                return;
        }
    }
}


// ---- XCD-local SOR
//
// The grid barrier above needs agent-scope release/acquire fences that write
// back and invalidate every L2 on each colour, ~7000 times per solve.  The
// sweeps of one mesh only touch that mesh's x, so here each mesh's rows go to
// a group of workgroups on one XCD that synchronise through its L2
// (xcd_sync.h): x is gathered with L1-bypassing device-scope loads and written
// with plain stores (WT = false, the default) or with device-scope stores that
// also write through to memory (WT = true).

template <bool WRITE_THROUGH>
__device__ __forceinline__ void store_x(double *p, double v)
{
    if (WRITE_THROUGH)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

// k_sor_registers with each mesh's rows on one XCD.  blocks[x] workgroups of
// XCD x handle meshes x, x + 8, ... (thread t of the group owns row t of every
// colour, meshes side by side as in k_sor_registers).  Barrier words: slot x
// holds XCD x's claim counter, slot 8 + x its arrival counter.
template <int CMAX, int WMAX, bool WT>
__global__ void __launch_bounds__(kSorRegBlock)
k_sor_xcd_local(const int *__restrict__ verts, SorColours co, const int *__restrict__ ell_cols,
                const double *__restrict__ ell_val, const double *__restrict__ rhs_c,
                const double *__restrict__ diag_c, int W, double omega, int n, int batch,
                int max_iter, SorXcdPlan plan, double *x, unsigned *bar, int *timeout)
{
    int xcd, slot;
    xcd_claim(bar, kBarrierSlotWords, plan.force_xcd, xcd, slot);
    const int nblocks = plan.blocks[xcd];
    if (slot >= nblocks)
        return;
    unsigned *cnt = bar + (size_t)(kXcds + xcd) * kBarrierSlotWords;
    const int meshes = batch > xcd ? (batch - xcd + kXcds - 1) / kXcds : 0;
    const int tid = slot * blockDim.x + threadIdx.x;
    const size_t wn = (size_t)W * n;
    bool act[CMAX];
    int vi[CMAX];
    int col[CMAX][WMAX];
    double val[CMAX][WMAX], rh[CMAX], dg[CMAX];
#pragma unroll
    for (int c = 0; c < CMAX; ++c) {
        act[c] = false;
        vi[c] = 0;
        rh[c] = 0.0;
        dg[c] = 1.0;
#pragma unroll
        for (int k = 0; k < WMAX; ++k) {
            col[c][k] = 0;
            val[c][k] = 0.0;
        }
        if (c < co.count) {
            const int start = co.off[c], count = co.off[c + 1] - start;
            if (tid < count * meshes) {
                const int b = xcd + kXcds * (tid / count);
                const int r = start + tid % count;
                act[c] = true;
                vi[c] = b * n + verts[r];
                rh[c] = rhs_c[(size_t)b * n + r];
                dg[c] = diag_c[(size_t)b * n + r];
#pragma unroll
                for (int k = 0; k < WMAX; ++k) {
                    if (k < W) {
                        val[c][k] = ell_val[(size_t)b * wn + (size_t)k * n + r];
                        col[c][k] = b * n + ell_cols[(size_t)k * n + r];
                    }
                }
            }
        }
    }
    unsigned target = 0;
    for (int it = 0; it < max_iter; ++it) {
#pragma unroll
        for (int c = 0; c < CMAX; ++c) {
            if (c >= co.count)
                break;
            if (act[c]) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < WMAX; ++k)
                    if (k < W)
                        s += val[c][k] * load_l2(x + col[c][k]);
                store_x<WT>(x + vi[c], load_l2(x + vi[c]) + omega * (rh[c] - s) / dg[c]);
            }
            target += (unsigned)nblocks;
            if (!group_barrier(cnt, target, timeout))
                return;
        }
    }
}

// The SOR sweeps blocked in time on mesh patches.  A patch lists its core
// vertices (ring 0) and `rings` rings of their graph neighbourhood, rings =
// the colour count: sweep colour t (t = 0 .. rings-1) only needs the current
// values of its rows' columns, so updating the rows of rings <= rings-1-t in
// phase t leaves the core exact after a whole iteration with no data from
// outside the patch (the same rows, columns and order as k_sor_xcd_local; the
// halo rows repeat what their owners compute).  One workgroup per (patch,
// mesh) on the mesh's XCD (b % 8) stages the patch's x in LDS, runs the
// iteration's colour phases behind workgroup barriers, writes its core and
// meets the XCD group's barrier once per iteration.  Iteration s reads
// x_buf[s & 1] and writes the cores to x_buf[(s + 1) & 1]: a workgroup that is
// ahead never overwrites values another one is still loading for the same
// iteration, so the one barrier suffices.  Thread t holds row-table
// entries t, t + T, .. (kSorPatchRows of them, ELL values in registers); the
// table is sorted by (colour, ring) so that phase t's rows are a prefix of the
// colour's entries.  A workgroup's own core is exact in its LDS after each
// iteration (it is what the workgroup stores), so later iterations load only
// the halo: 7.6 -> 6.6 ms for four 40,962-vertex meshes.
constexpr int kSorPatchThreads = 1024;
constexpr int kSorPatchRows = 3;

template <int WMAX>
__global__ void __launch_bounds__(kSorPatchThreads)
k_sor_patched(SorPatches pp, int ncolours, const double *__restrict__ ell_val,
              const double *__restrict__ rhs_c, const double *__restrict__ diag_c, int W,
              double omega, int n, int max_iter, SorXcdPlan plan, double *x0, double *x1,
              unsigned *bar, int *timeout)
{
    extern __shared__ double g_sor_lds[];
    int xcd, slot;
    xcd_claim(bar, kBarrierSlotWords, plan.force_xcd, xcd, slot);
    const int nblocks = plan.blocks[xcd];
    if (slot >= nblocks)
        return;
    unsigned *cnt = bar + (size_t)(kXcds + xcd) * kBarrierSlotWords;
    const int p = slot % pp.count;
    const int b = xcd + kXcds * (slot / pp.count);
    const int tid = threadIdx.x;
    const int s0 = pp.vstart[p], L = pp.vstart[p + 1] - s0, core = pp.ncore[p];
    const int *lv = pp.lverts + s0;
    const int e0 = pp.estart[p], E = pp.estart[p + 1] - e0;
    const int *cst = pp.cstart + (size_t)p * (ncolours + 1);
    const int *act = pp.active + (size_t)p * pp.rings;
    double *xs = g_sor_lds;
    double *xbuf[2] = {x0 + (size_t)b * n, x1 + (size_t)b * n};
    const size_t wn = (size_t)W * n;
    int ent[kSorPatchRows], lj[kSorPatchRows], lc[kSorPatchRows][WMAX];
    double val[kSorPatchRows][WMAX], rh[kSorPatchRows], dg[kSorPatchRows];
#pragma unroll
    for (int q = 0; q < kSorPatchRows; ++q) {
        const int e = tid + q * kSorPatchThreads;
        ent[q] = e < E ? e : -1;
        lj[q] = 0;
        rh[q] = 0.0;
        dg[q] = 1.0;
#pragma unroll
        for (int k = 0; k < WMAX; ++k) {
            lc[q][k] = 0;
            val[q][k] = 0.0;
        }
        if (e < E) {
            const int r = pp.e_row[e0 + e];
            lj[q] = pp.e_local[e0 + e];
            rh[q] = rhs_c[(size_t)b * n + r];
            dg[q] = diag_c[(size_t)b * n + r];
#pragma unroll
            for (int k = 0; k < WMAX; ++k)
                if (k < W) {
                    val[q][k] = ell_val[(size_t)b * wn + (size_t)k * n + r];
                    lc[q][k] = pp.e_lcol[(size_t)(e0 + e) * W + k];
                }
        }
    }
    const long long total = (long long)max_iter * ncolours;
    unsigned target = 0;
    int step = 0;
    for (long long ph0 = 0; ph0 < total; ph0 += pp.rings, ++step) {
        const double *src = xbuf[step & 1];
        double *dst = xbuf[(step + 1) & 1];
        for (int j = step == 0 ? tid : core + tid; j < L; j += kSorPatchThreads)
            xs[j] = load_l2(src + lv[j]);
        __syncthreads();
        const int m = total - ph0 < pp.rings ? (int)(total - ph0) : pp.rings;
        for (int t = 0; t < m; ++t) {
            const int lo = cst[t], hi = lo + act[t];
#pragma unroll
            for (int q = 0; q < kSorPatchRows; ++q) {
                if (ent[q] >= lo && ent[q] < hi) {
                    double s = 0.0;
#pragma unroll
                    for (int k = 0; k < WMAX; ++k)
                        if (k < W)
                            s += val[q][k] * xs[lc[q][k]];
                    xs[lj[q]] = xs[lj[q]] + omega * (rh[q] - s) / dg[q];
                }
            }
            __syncthreads();
        }
        for (int j = tid; j < core; j += kSorPatchThreads)
            dst[lv[j]] = xs[j];
        target += (unsigned)nblocks;
        if (!group_barrier(cnt, target, timeout))
            return;
    }
}

__global__ void __launch_bounds__(kBlock) k_clear_words(unsigned *__restrict__ w, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        w[i] = 0u;
}

// The whole barrier state cleared by a kernel (a memset node would not be
// safe in a captured graph).
int clear_barrier(unsigned *bar, hipStream_t s)
{
    const int n = kBarrierSlots * kBarrierSlotWords;
    hipLaunchKernelGGL(k_clear_words, dim3(grid_for(n)), dim3(kBlock), 0, s, bar, n);
    CSG_LAUNCH_CHECK();
    return kOk;
}

// Between two launches that share the barrier state: clear the eight claim and
// eight arrival counters (slots 0 .. 15); the timeout flag and the placement
// verdict stay.
__global__ void k_clear_group_counters(unsigned *__restrict__ bar)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * kXcds * kBarrierSlotWords)
        bar[i] = 0u;
}

// After an XCD-local launch: if some XCD claimed fewer workgroups than its
// group needs, that group never completed; record it in the verdict word.
__global__ void k_check_claims(unsigned *__restrict__ bar, SorXcdPlan plan)
{
    const int x = threadIdx.x;
    if (x < kXcds && bar[(size_t)x * kBarrierSlotWords] < (unsigned)plan.blocks[x])
        bar[(size_t)(kBarrierSlots - 2) * kBarrierSlotWords] = 1u;
}

// r = rhs - A x; per-block Σ r², Σ rhs².
__global__ void __launch_bounds__(kBlock)
k_residual_norms(const int *__restrict__ row_off, const int *__restrict__ cols,
                 const double *__restrict__ val, const double *__restrict__ rhs,
                 const double *__restrict__ x, int n, int nnz, double *__restrict__ partial,
                 int nblocks)
{
    __shared__ double red[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    double r2 = 0.0, b2 = 0.0;
    if (i < n) {
        const double *vb = val + (size_t)b * nnz;
        const double *xb = x + (size_t)b * n;
        double s = 0.0;
        for (int k = row_off[i]; k < row_off[i + 1]; ++k)
            s += vb[k] * xb[cols[k]];
        double rr = rhs[(size_t)b * n + i];
        double r = rr - s;
        r2 = r * r;
        b2 = rr * rr;
    }
    double a = block_sum(r2, red);
    double c = block_sum(b2, red);
    if (threadIdx.x == 0) {
        partial[2 * ((size_t)b * nblocks + blockIdx.x)] = a;
        partial[2 * ((size_t)b * nblocks + blockIdx.x) + 1] = c;
    }
}

__global__ void k_residual_final(const double *__restrict__ partial, int nblocks, int batch,
                                 double *__restrict__ out)
{
    int b = threadIdx.x;
    if (b >= batch)
        return;
    double a = 0.0, c = 0.0;
    for (int k = 0; k < nblocks; ++k) {
        a += partial[2 * ((size_t)b * nblocks + k)];
        c += partial[2 * ((size_t)b * nblocks + k) + 1];
    }
    out[b] = sqrt(a) / fmax(sqrt(c), 2.220446049250313e-16);
}

int depth_setup(const DepthArgs &a, hipStream_t s)
{
    const int B = a.batch;
    hipLaunchKernelGGL(k_depth_face_terms, dim3(grid_for(a.nf), B), dim3(kBlock), 0, s, a.pts,
                       a.faces, a.n, a.nf, a.fnormal, a.farea, a.coo);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_depth_vertex_terms, dim3(grid_for(a.n), B), dim3(kBlock), 0, s,
                       a.fnormal, a.farea, a.inc_off, a.inc, a.n, a.nf, a.normals, a.areas);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_merged_values, dim3(grid_for(a.nnz), B), dim3(kBlock), 0, s, a.coo,
                       a.contrib_off, a.contrib, nullptr, a.nnz, a.nf, a.n, a.areas, a.alpha,
                       a.lap, a.sys);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_depth_rhs_parts, dim3(grid_for(a.n), B), dim3(kBlock), 0, s, a.pts,
                       a.row_off, a.cols, a.lap, a.sys, a.normals, a.areas, a.n, a.nnz, a.mc,
                       a.diag);
    CSG_LAUNCH_CHECK();
    int nb = (int)grid_for(a.n);
    hipLaunchKernelGGL(k_weighted_sums, dim3(nb, B), dim3(kBlock), 0, s, a.mc, a.areas, a.n,
                       a.partial, nb);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_depth_rhs, dim3(nb, B), dim3(kBlock), 0, s, a.mc, a.areas, a.n,
                       a.partial, nb, a.rhs, a.x);
    CSG_LAUNCH_CHECK();
    return kOk;
}

namespace {
std::atomic<int> g_sor_xcd_local{2};
std::atomic<int> g_sor_force_xcd{-1};
}

int set_sor_force_xcd(int xcd)
{
    return g_sor_force_xcd.exchange(xcd < 0 ? -1 : xcd);
}

int sor_force_xcd()
{
    return g_sor_force_xcd.load();
}

int set_sor_xcd_local(int mode)
{
    return g_sor_xcd_local.exchange(mode < 0 ? 0 : (mode > 2 ? 2 : mode));
}

int depth_sor(const DepthArgs &a, const SorColours &co, int max_iter, double omega,
              const SorPatches *patches, bool allow_xcd_local, SorRun *run, hipStream_t s)
{
    if (run)
        *run = SorRun{};
    const int W = a.ell_width;
    const size_t total = (size_t)W * a.n + a.n;
    hipLaunchKernelGGL(k_colour_order, dim3(grid_for(total), a.batch), dim3(kBlock), 0, s,
                       a.colour_verts, a.ell_src, W, a.n, a.sys, a.nnz, a.rhs, a.diag, a.ell_val,
                       a.rhs_c, a.diag_c);
    CSG_LAUNCH_CHECK();
    CSG_TRY(clear_barrier(a.barrier, s));
    if (max_iter <= 0)
        return kOk;
    int dev = 0;
    CSG_HIP_CHECK(hipGetDevice(&dev));
    int cus = 0, per_cu = 0;
    CSG_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CSG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sor_persistent,
                                                               kSorBlock, 0));
    int most = 0;
    for (int c = 0; c < co.count; ++c)
        most = (co.off[c + 1] - co.off[c]) > most ? (co.off[c + 1] - co.off[c]) : most;
    int want = (int)grid_for((size_t)most * a.batch, kSorBlock);
    int reg_per_cu = 0;
    CSG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &reg_per_cu, k_sor_registers<kSorRegColours, kSorRegWidth>, kSorRegBlock, 0));
    // XCD-local placement: mesh b on XCD b % 8; every XCD's group must fit on
    // its share of the CUs at the register kernel's occupancy
    SorXcdPlan plan{};
    plan.force_xcd = g_sor_force_xcd.load();
    int local_per_cu = 0;
    CSG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &local_per_cu, k_sor_xcd_local<kSorRegColours, kSorRegWidth, true>, kSorRegBlock, 0));
    const int placement = allow_xcd_local ? g_sor_xcd_local.load() : 0;
    bool local = placement != 0 && co.count <= kSorRegColours &&
                 W <= kSorRegWidth && cus % kXcds == 0;
    int local_grid = 0;
    for (int xcd = 0; xcd < kXcds && local; ++xcd) {
        const int meshes = a.batch > xcd ? (a.batch - xcd + kXcds - 1) / kXcds : 0;
        plan.blocks[xcd] = (int)((size_t)most * meshes + kSorRegBlock - 1) / kSorRegBlock;
        local = plan.blocks[xcd] <= (cus / kXcds) * local_per_cu;
        local_grid = plan.blocks[xcd] > local_grid ? plan.blocks[xcd] : local_grid;
    }
    local_grid = (local_grid + kXcdSpare) * kXcds;
    const int reg_grid = (int)grid_for((size_t)most * a.batch, kSorRegBlock);
    const bool registers = co.count <= kSorRegColours && W <= kSorRegWidth &&
                           reg_grid <= cus * reg_per_cu &&
                           reg_grid <= (kBarrierSlots - 3) * kBarrierGroup;
    int cap = cus * (per_cu > 0 ? per_cu : 1);
    // the barrier state holds one slot per group, the root and the generation word
    int slots_cap = (kBarrierSlots - 3) * kBarrierGroup;
    cap = cap < slots_cap ? cap : slots_cap;
    int grid = want < cap ? want : cap;
    if (grid < 1)
        grid = 1;
    int *timeout = reinterpret_cast<int *>(a.barrier + (size_t)(kBarrierSlots - 1) * kBarrierSlotWords);
    const int *verts = a.colour_verts;
    const int *ell_cols = a.ell_cols;
    const double *ell_val = a.ell_val, *rhs_c = a.rhs_c, *diag_c = a.diag_c;
    int n = a.n, batch = a.batch, width = W;
    double *x = a.x;
    unsigned *bar = a.barrier;
    SorColours colours = co;
    // per row and sweep: the row's ELL values and columns (12 B per slot), its
    // x gathers (8 B per slot), rhs, diagonal and the x update (24 B)
    const double algorithmic = (double)max_iter * batch * n * (20.0 * W + 24.0);
    if (allow_xcd_local && patches && patches->rings == co.count && co.count <= kSorRegColours &&
        W <= kSorRegWidth && cus % kXcds == 0 && a.batch >= 1 &&
        patches->max_rows <= kSorPatchRows * kSorPatchThreads) {
        // every (patch, mesh) workgroup of an XCD must be resident at once: an
        // XCD holds per_xcd meshes, a launch 8 per_xcd; larger batches run as
        // consecutive launches of that many meshes (each mesh's sweeps depend
        // only on its own rows, so the values do not depend on the split)
        const size_t lds = sizeof(double) * (size_t)patches->max_local;
        int patch_per_cu = 0;
        CSG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &patch_per_cu, k_sor_patched<kSorRegWidth>, kSorPatchThreads, lds));
        const int per_xcd = patch_per_cu > 0 ? (cus / kXcds) * patch_per_cu / patches->count : 0;
        if (per_xcd >= 1) {
            const int chunk = per_xcd * kXcds;
            // after an odd number of iterations the result is in the second buffer
            const long long steps = ((long long)max_iter * co.count + patches->rings - 1) /
                                    patches->rings;
            SorXcdPlan pplan{};
            for (int b0 = 0; b0 < batch; b0 += chunk) {
                const int nb = batch - b0 < chunk ? batch - b0 : chunk;
                pplan = SorXcdPlan{};
                pplan.force_xcd = plan.force_xcd;
                int pgrid = 0;
                for (int xcd = 0; xcd < kXcds; ++xcd) {
                    const int meshes = nb > xcd ? (nb - xcd + kXcds - 1) / kXcds : 0;
                    pplan.blocks[xcd] = patches->count * meshes;
                    pgrid = pplan.blocks[xcd] > pgrid ? pplan.blocks[xcd] : pgrid;
                }
                pgrid = (pgrid + kXcdSpare) * kXcds;
                if (b0 > 0) {
                    hipLaunchKernelGGL(k_clear_group_counters,
                                       dim3(grid_for((size_t)2 * kXcds * kBarrierSlotWords)),
                                       dim3(kBlock), 0, s, bar);
                    CSG_LAUNCH_CHECK();
                }
                const size_t off = (size_t)b0 * n;
                TimedLaunch ptimer(kTimedSorPatched, s, algorithmic * nb / batch);
                hipExtLaunchKernelGGL(k_sor_patched<kSorRegWidth>, dim3(pgrid),
                                      dim3(kSorPatchThreads), lds, s, ptimer.start(),
                                      ptimer.stop(), 0, *patches, co.count,
                                      ell_val + off * width, rhs_c + off, diag_c + off, width,
                                      omega, n, max_iter, pplan, x + off, a.x_alt + off, bar,
                                      timeout);
                CSG_LAUNCH_CHECK();
                ptimer.done();
                hipLaunchKernelGGL(k_check_claims, dim3(1), dim3(64), 0, s, bar, pplan);
                CSG_LAUNCH_CHECK();
                if (steps & 1)
                    CSG_HIP_CHECK(hipMemcpyAsync(x + off, a.x_alt + off,
                                                 sizeof(double) * (size_t)nb * n,
                                                 hipMemcpyDeviceToDevice, s));
            }
            if (run)
                *run = SorRun{true, pplan};
            return kOk;
        }
    }
    TimedLaunch timer(local ? kTimedSorXcdLocal : registers ? kTimedSorRegisters : kTimedSor, s,
                      algorithmic);
    // A regular launch: the grid is capped at what the device holds at once
    // (occupancy x CUs), and every barrier wait is bounded, so blocks that were
    // not resident turn into a timeout error rather than a hang.
    if (local && placement == 2)
        hipExtLaunchKernelGGL((k_sor_xcd_local<kSorRegColours, kSorRegWidth, false>),
                           dim3(local_grid), dim3(kSorRegBlock), 0, s, timer.start(), timer.stop(), 0, verts, colours, ell_cols,
                           ell_val, rhs_c, diag_c, width, omega, n, batch, max_iter, plan, x, bar,
                           timeout);
    else if (local)
        hipExtLaunchKernelGGL((k_sor_xcd_local<kSorRegColours, kSorRegWidth, true>),
                           dim3(local_grid), dim3(kSorRegBlock), 0, s, timer.start(), timer.stop(), 0, verts, colours, ell_cols,
                           ell_val, rhs_c, diag_c, width, omega, n, batch, max_iter, plan, x, bar,
                           timeout);
    else if (registers)
        hipExtLaunchKernelGGL((k_sor_registers<kSorRegColours, kSorRegWidth>), dim3(reg_grid),
                           dim3(kSorRegBlock), 0, s, timer.start(), timer.stop(), 0, verts, colours, ell_cols, ell_val, rhs_c,
                           diag_c, width, omega, n, batch, max_iter, x, bar, timeout);
    else
        hipExtLaunchKernelGGL(k_sor_persistent, dim3(grid), dim3(kSorBlock), 0, s, timer.start(), timer.stop(), 0, verts, colours,
                           ell_cols, ell_val, rhs_c, diag_c, width, omega, n, batch, max_iter,
                           x, bar, timeout);
    CSG_LAUNCH_CHECK();
    timer.done();
    if (run && local)
        *run = SorRun{true, plan};
    return kOk;
}

bool sor_placement_complete(const DepthArgs &a, const SorRun &run, hipStream_t s, int *rc)
{
    *rc = kOk;
    if (!run.xcd_local)
        return true;
    // slot x of the barrier state holds XCD x's claim counter: the number of
    // workgroups that ran on that XCD (spares included)
    unsigned claims[kXcds * kBarrierSlotWords];
    const hipError_t e1 = hipMemcpyAsync(claims, a.barrier, sizeof(claims),
                                         hipMemcpyDeviceToHost, s);
    const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(s) : e1;
    if (e2 != hipSuccess) {
        *rc = fail(kRuntimeError, "reading the SOR placement: %s", hipGetErrorString(e2));
        return false;
    }
    for (int x = 0; x < kXcds; ++x)
        if ((long long)claims[(size_t)x * kBarrierSlotWords] < (long long)run.plan.blocks[x])
            return false;
    // earlier launches of a split batch left their verdict in slot kBarrierSlots - 2
    unsigned verdict = 0;
    const hipError_t e3 = hipMemcpyAsync(&verdict, a.barrier + (size_t)(kBarrierSlots - 2) *
                                                                  kBarrierSlotWords,
                                         sizeof(verdict), hipMemcpyDeviceToHost, s);
    const hipError_t e4 = e3 == hipSuccess ? hipStreamSynchronize(s) : e3;
    if (e4 != hipSuccess) {
        *rc = fail(kRuntimeError, "reading the SOR placement: %s", hipGetErrorString(e4));
        return false;
    }
    return verdict == 0;
}

// status[0] = the barrier time-out flag, status[1] = 1 when an XCD-local run's
// groups were not all placed (claims short of the plan, or an earlier launch's
// verdict): what sor_placement_complete and the host's time-out read decide,
// left on the device.
__global__ void k_sor_status(const unsigned *__restrict__ bar, SorXcdPlan plan, int local,
                             int *__restrict__ status)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    status[0] = bar[(size_t)(kBarrierSlots - 1) * kBarrierSlotWords] != 0u ? 1 : 0;
    int bad = 0;
    if (local) {
        for (int x = 0; x < kXcds; ++x)
            bad |= bar[(size_t)x * kBarrierSlotWords] < (unsigned)plan.blocks[x];
        bad |= bar[(size_t)(kBarrierSlots - 2) * kBarrierSlotWords] != 0u;
    }
    status[1] = bad;
}

int sor_status(const DepthArgs &a, const SorRun &run, int *status, hipStream_t s)
{
    hipLaunchKernelGGL(k_sor_status, dim3(1), dim3(64), 0, s, a.barrier, run.plan,
                       run.xcd_local ? 1 : 0, status);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int depth_residual(const DepthArgs &a, double *out, hipStream_t s)
{
    int nb = (int)grid_for(a.n);
    hipLaunchKernelGGL(k_residual_norms, dim3(nb, a.batch), dim3(kBlock), 0, s, a.row_off,
                       a.cols, a.sys, a.rhs, a.x, a.n, a.nnz, a.partial, nb);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(64), 0, s, a.partial, nb, a.batch, out);
    CSG_LAUNCH_CHECK();
    return kOk;
}

}  // namespace csg
