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
    double den = 1.0 + (double)(e - b) / 3.0;
    nxt[3 * i] = (float)(((double)cur[3 * i] + sx / 3.0) / den);
    nxt[3 * i + 1] = (float)(((double)cur[3 * i + 1] + sy / 3.0) / den);
    nxt[3 * i + 2] = (float)(((double)cur[3 * i + 2] + sz / 3.0) / den);
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
    double sg = hb.sigma[b];
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
    if (iteration >= hb.counts[b]) {
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

// ------------------------------------------------------------- depth potential

// Per face: the three unit-edge normals terms of _stable_normals, the mixed
// Voronoi area contributions, and the 9 cotangent entries (float64).
// rotation_feature.py:24-186
__global__ void __launch_bounds__(kBlock)
k_depth_face_terms(const float *__restrict__ pts, const int *__restrict__ faces, int nf,
                   double *__restrict__ fnormal /* [F][3][3] weighted normal per corner */,
                   double *__restrict__ farea /* [F][3] */, double *__restrict__ coo /* [9][F] */)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
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
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto cross = [](const double *a, const double *b, double *c) {
        c[0] = a[1] * b[2] - a[2] * b[1];
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto norm = [&](const double *a) { return sqrt(dot(a, a)); };
    // _stable_normals
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
    // _mixed_voronoi_areas
    double cr[3];
    cross(v1, v2, cr);
    double area = fmax(norm(cr), eps);
    double d31 = dot(v3, v1), d12 = dot(v1, v2), d23 = dot(v2, v3);
    bool bad0 = d31 > 0.0, bad1 = d12 > 0.0, bad2 = d23 > 0.0;
    bool bad = bad0 || bad1 || bad2;
    double l1 = dot(v1, v1), l2 = dot(v2, v2), l3 = dot(v3, v3);
    double d13 = dot(v1, v3);
    double wt0 = l3 * d12 + l1 * d23;
    double wt1 = l2 * d13 + l1 * d23;
    double wt2 = l2 * d13 + l3 * d12;
    double a0, a1, a2;
    if (bad) {
        a0 = a1 = a2 = 0.125 * area;
    } else {
        a0 = -0.125 * wt0 / area;
        a1 = -0.125 * wt1 / area;
        a2 = -0.125 * wt2 / area;
    }
    a0 += bad0 ? 0.125 * area : 0.0;
    a1 += bad1 ? 0.125 * area : 0.0;
    a2 += bad2 ? 0.125 * area : 0.0;
    farea[3 * (size_t)f] = a0;
    farea[3 * (size_t)f + 1] = a1;
    farea[3 * (size_t)f + 2] = a2;
    // _system_entries (cotangent weights)
    double e23 = d23 / area, e13 = d13 / area, e12 = d12 / area;
    double vals[9] = {-0.5 * (e23 + e12), -0.5 * (e13 + e23), -0.5 * (e12 + e13),
                      0.5 * e23, 0.5 * e23, 0.5 * e13, 0.5 * e13, 0.5 * e12, 0.5 * e12};
    for (int k = 0; k < 9; ++k)
        coo[(size_t)k * nf + f] = vals[k];
}

// Per vertex: normal (sum of incident weighted normals, then normalised) and
// mixed area (sum of incident contributions, 1 where invalid).  Entries of
// inc[] are face*3+corner in (corner, face) order.  rotation_feature.py:69-73, :125-128
__global__ void __launch_bounds__(kBlock)
k_depth_vertex_terms(const double *__restrict__ fnormal, const double *__restrict__ farea,
                     const int *__restrict__ inc_off, const int *__restrict__ inc, int n,
                     double *__restrict__ normals, double *__restrict__ areas)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double eps = 2.220446049250313e-16;
    double nx = 0.0, ny = 0.0, nz = 0.0, a = 0.0;
    for (int k = inc_off[i]; k < inc_off[i + 1]; ++k) {
        int fc = inc[k];
        const double *fn = fnormal + 3 * (size_t)fc;  // [F][3 corners][3]
        nx += fn[0];
        ny += fn[1];
        nz += fn[2];
        a += farea[fc];
    }
    double nr = fmax(sqrt(nx * nx + ny * ny + nz * nz), eps);
    normals[3 * i] = nx / nr;
    normals[3 * i + 1] = ny / nr;
    normals[3 * i + 2] = nz / nr;
    areas[i] = (!isfinite(a) || a == 0.0) ? 1.0 : a;
}

// CSR matrix values from per-face COO values: val[k] = coo[perm[k]] for the
// face entries, alpha·area[i] for the diagonal extra entries (perm >= 9F).
__global__ void __launch_bounds__(kBlock)
k_csr_values(const double *__restrict__ coo, const int *__restrict__ perm, int nnz, int nface9,
             const double *__restrict__ areas, double alpha, double *__restrict__ val)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz)
        return;
    int e = perm[k];
    val[k] = e < nface9 ? coo[e] : alpha * areas[e - nface9];
}

// Per vertex: Laplacian·points (without the alpha terms), mean curvature,
// and the matrix diagonal.  rotation_feature.py:205-226, :275-281
__global__ void __launch_bounds__(kBlock)
k_depth_rhs_parts(const float *__restrict__ pts, const int *__restrict__ row_off,
                  const int *__restrict__ cols, const double *__restrict__ val,
                  const unsigned char *__restrict__ is_alpha, const double *__restrict__ normals,
                  const double *__restrict__ areas, int n, double *__restrict__ mc,
                  double *__restrict__ diag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double lx = 0.0, ly = 0.0, lz = 0.0, dg = 0.0;
    for (int k = row_off[i]; k < row_off[i + 1]; ++k) {
        int j = cols[k];
        double v = val[k];
        if (j == i)
            dg += v;
        if (is_alpha[k])
            continue;
        lx += v * (double)pts[3 * j];
        ly += v * (double)pts[3 * j + 1];
        lz += v * (double)pts[3 * j + 2];
    }
    lx *= 2.0;
    ly *= 2.0;
    lz *= 2.0;
    double h = 0.5 * (lx * normals[3 * i] + ly * normals[3 * i + 1] + lz * normals[3 * i + 2]) /
               areas[i];
    mc[i] = isfinite(h) ? h : 0.0;
    diag[i] = fmax(dg, 2.220446049250313e-16);
}

// One colour of SOR: x_i += ω (rhs_i - Σ_j A_ij x_j) / diag_i for the
// vertices of the colour.  rotation_feature.py:296-311
__global__ void __launch_bounds__(kBlock)
k_sor_color(const int *__restrict__ verts, int count, const int *__restrict__ row_off,
            const int *__restrict__ cols, const double *__restrict__ val,
            const double *__restrict__ rhs, const double *__restrict__ diag, double omega,
            double *__restrict__ x)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count)
        return;
    int i = verts[t];
    double s = 0.0;
    for (int k = row_off[i]; k < row_off[i + 1]; ++k)
        s += val[k] * x[cols[k]];
    x[i] = x[i] + omega * (rhs[i] - s) / diag[i];
}

// r_i = rhs_i - Σ_j A_ij x_j and per-block Σ r², Σ rhs².
__global__ void __launch_bounds__(kBlock)
k_residual_norms(const int *__restrict__ row_off, const int *__restrict__ cols,
                 const double *__restrict__ val, const double *__restrict__ rhs,
                 const double *__restrict__ x, int n, double *__restrict__ partial)
{
    __shared__ double red[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0, b2 = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int k = row_off[i]; k < row_off[i + 1]; ++k)
            s += val[k] * x[cols[k]];
        double r = rhs[i] - s;
        r2 = r * r;
        b2 = rhs[i] * rhs[i];
    }
    double a = block_sum(r2, red);
    double b = block_sum(b2, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

// Σ mc·area and Σ area per block (the curvature mean factor).
__global__ void __launch_bounds__(kBlock)
k_weighted_sums(const double *__restrict__ mc, const double *__restrict__ areas, int n,
                double *__restrict__ partial)
{
    __shared__ double red[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < n) {
        a = mc[i] * areas[i];
        b = areas[i];
    }
    double ra = block_sum(a, red);
    double rb = block_sum(b, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ra;
        partial[2 * blockIdx.x + 1] = rb;
    }
}

// rhs_i = (mc_i - factor)·area_i with factor = Σ mc·a / Σ a from the partials.
__global__ void __launch_bounds__(kBlock)
k_depth_rhs(const double *__restrict__ mc, const double *__restrict__ areas, int n,
            const double *__restrict__ partial, int nblocks, double *__restrict__ rhs)
{
    __shared__ double factor;
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < nblocks; ++k) {
            a += partial[2 * k];
            b += partial[2 * k + 1];
        }
        factor = a / b;
    }
    __syncthreads();
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        rhs[i] = (mc[i] - factor) * areas[i];
}

__global__ void k_residual_final(const double *__restrict__ partial, int nblocks,
                                 double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nblocks; ++k) {
        a += partial[2 * k];
        b += partial[2 * k + 1];
    }
    out[0] = sqrt(a) / fmax(sqrt(b), 2.220446049250313e-16);
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

// ------------------------------------------------------------- launchers

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

int launch_heat_weights(const float *geom, const int *off, const int *nbr, int n,
                        const HeatBatch &hb, int squared_norm, size_t wstride, double *w,
                        double *wsum, hipStream_t s)
{
    hipLaunchKernelGGL(k_heat_weights, dim3(grid_for(n), hb.nb), dim3(kBlock), 0, s, geom, off,
                       nbr, n, hb, squared_norm, wstride, w, wsum);
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

int launch_depth_face_terms(const float *pts, const int *faces, int nf, double *fnormal,
                            double *farea, double *coo, hipStream_t s)
{
    hipLaunchKernelGGL(k_depth_face_terms, dim3(grid_for(nf)), dim3(kBlock), 0, s, pts, faces,
                       nf, fnormal, farea, coo);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_depth_vertex_terms(const double *fnormal, const double *farea, const int *inc_off,
                              const int *inc, int n, double *normals, double *areas,
                              hipStream_t s)
{
    hipLaunchKernelGGL(k_depth_vertex_terms, dim3(grid_for(n)), dim3(kBlock), 0, s, fnormal,
                       farea, inc_off, inc, n, normals, areas);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_csr_values(const double *coo, const int *perm, int nnz, int nface9,
                      const double *areas, double alpha, double *val, hipStream_t s)
{
    hipLaunchKernelGGL(k_csr_values, dim3(grid_for(nnz)), dim3(kBlock), 0, s, coo, perm, nnz,
                       nface9, areas, alpha, val);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_depth_rhs(const float *pts, const int *row_off, const int *cols, const double *val,
                     const unsigned char *is_alpha, const double *normals, const double *areas,
                     int n, double *mc, double *diag, double *partial, double *rhs,
                     hipStream_t s)
{
    int nb = (int)grid_for(n);
    hipLaunchKernelGGL(k_depth_rhs_parts, dim3(nb), dim3(kBlock), 0, s, pts, row_off, cols, val,
                       is_alpha, normals, areas, n, mc, diag);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_weighted_sums, dim3(nb), dim3(kBlock), 0, s, mc, areas, n, partial);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_depth_rhs, dim3(nb), dim3(kBlock), 0, s, mc, areas, n, partial, nb, rhs);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_sor_color(const int *verts, int count, const int *row_off, const int *cols,
                     const double *val, const double *rhs, const double *diag, double omega,
                     double *x, hipStream_t s)
{
    if (count <= 0)
        return kOk;
    hipLaunchKernelGGL(k_sor_color, dim3(grid_for(count)), dim3(kBlock), 0, s, verts, count,
                       row_off, cols, val, rhs, diag, omega, x);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_relative_residual(const int *row_off, const int *cols, const double *val,
                             const double *rhs, const double *x, int n, double *partial,
                             double *out, hipStream_t s)
{
    int nb = (int)grid_for(n);
    hipLaunchKernelGGL(k_residual_norms, dim3(nb), dim3(kBlock), 0, s, row_off, cols, val, rhs, x,
                       n, partial);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(64), 0, s, partial, nb, out);
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
