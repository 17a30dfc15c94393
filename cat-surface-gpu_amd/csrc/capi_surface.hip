// C ABI of the surface side of the hot path (declared in include/csg.h).
#include "../../include/csg.h"
#include "depth_kernels.h"
#include "xcd_sync.h"
#include "surface_kernels.h"

#include <algorithm>
#include <atomic>

using namespace csg;

namespace {

struct Carve {
    char *base;
    size_t used;
    template <typename U>
    U *take(size_t count)
    {
        size_t bytes = (count * sizeof(U) + 255) & ~(size_t)255;
        U *p = base ? reinterpret_cast<U *>(base + used) : nullptr;
        used += bytes;
        return p;
    }
};

inline hipStream_t S(void *s) { return (hipStream_t)s; }

// Shared pieces of the heat-value smoothing: workspace layout and run.
struct HeatWs {
    double *w, *wsum, *buf[2];
};

HeatWs carve_heat(Carve &c, int n, int nnz, int nb)
{
    HeatWs h;
    h.w = c.take<double>((size_t)nb * nnz);
    h.wsum = c.take<double>((size_t)nb * n);
    h.buf[0] = c.take<double>((size_t)nb * n);
    h.buf[1] = c.take<double>((size_t)nb * n);
    return h;
}

int make_batch(int nb, const double *sigma, const int *iterations, HeatBatch &hb,
               const int *mesh = nullptr, int meshes = 1)
{
    if (nb < 1 || nb > kMaxHeatBatch)
        return fail(kInvalidArgument, "value smoothing batch must be in [1, %d], got %d",
                    kMaxHeatBatch, nb);
    hb.nb = nb;
    for (int i = 0; i < nb; ++i) {
        if (iterations[i] < 0)
            return fail(kInvalidArgument, "heat iteration counts must be non-negative");
        hb.counts[i] = iterations[i];
        hb.sigma[i] = sigma[i];
        hb.mesh[i] = mesh ? mesh[i] : 0;
        if (hb.mesh[i] < 0 || hb.mesh[i] >= meshes)
            return fail(kInvalidArgument, "value smoothing mesh index %d out of range",
                        hb.mesh[i]);
    }
    return kOk;
}

// The same batch with the counts and widths in device memory (iters_dev,
// sigma_dev: nb entries each); host loops run to `bound` and rows past their
// count are copied through.
int make_batch_dev(int nb, const double *sigma_dev, const int *iters_dev, int bound,
                   HeatBatch &hb, const int *mesh = nullptr, int meshes = 1)
{
    if (nb < 1 || nb > kMaxHeatBatch)
        return fail(kInvalidArgument, "value smoothing batch must be in [1, %d], got %d",
                    kMaxHeatBatch, nb);
    if (!sigma_dev || !iters_dev || bound < 0)
        return fail(kInvalidArgument, "device heat parameters and a non-negative bound are "
                                      "required");
    hb.nb = nb;
    for (int i = 0; i < nb; ++i) {
        hb.counts[i] = bound;
        hb.sigma[i] = 0.0;
        hb.mesh[i] = mesh ? mesh[i] : 0;
        if (hb.mesh[i] < 0 || hb.mesh[i] >= meshes)
            return fail(kInvalidArgument, "value smoothing mesh index %d out of range",
                        hb.mesh[i]);
    }
    hb.dcounts = iters_dev;
    hb.dsigma = sigma_dev;
    return kOk;
}

// values (already in h.buf[0], nb rows of n) smoothed in place; returns the buffer
// holding the result.
int run_heat(const float *geom, const int *off, const int *nbr, int n, int nnz,
             const HeatBatch &hb, int squared_norm, HeatWs &h, double **result, hipStream_t s)
{
    CSG_TRY(launch_heat_weights(geom, off, nbr, n, hb, squared_norm, (size_t)nnz, h.w, h.wsum, s));
    int most = 0;
    for (int i = 0; i < hb.nb; ++i)
        most = hb.counts[i] > most ? hb.counts[i] : most;
    int cur = 0;
    for (int it = 0; it < most; ++it) {
        CSG_TRY(launch_smooth_values(h.buf[cur], off, nbr, h.w, h.wsum, n, (size_t)nnz, hb, it,
                                     h.buf[cur ^ 1], s));
        cur ^= 1;
    }
    *result = h.buf[cur];
    return kOk;
}

// flag[0] = 1 when some value exceeds bound (threads that find one store 1).
__global__ void k_flag_above(const int *__restrict__ values, int n, int bound,
                             int *__restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && values[i] > bound)
        flag[0] = 1;
}

// csg_smooth_values_patched after its batch is set up (host or device counts).
int smooth_values_patched_run(const double *values, int n, int nb, const float *geom,
                              const int *nbr_off, const int *nbr, int nnz, const HeatBatch &hb,
                              int squared_norm, int npatches, int rings, const int *perm,
                              const int *vstart, const int *ring_end, const int *lverts,
                              const int *aoff, const int *astart, const unsigned short *ladj,
                              size_t lds_bytes, double *out, void *ws, size_t ws_bytes,
                              hipStream_t s)
{
    if (ws_bytes < csg_smooth_values_workspace_bytes(n, nnz, nb))
        return fail(kInvalidArgument, "smooth_values workspace too small");
    if (npatches < 1 || npatches > 65535 || rings < 1 || rings > 64 || !perm || !vstart ||
        !ring_end || !lverts || !aoff || !astart || !ladj)
        return fail(kInvalidArgument, "invalid smoothing patches");
    if (lds_bytes < 16 || lds_bytes > 160 * 1024)
        return fail(kInvalidArgument, "patch LDS size must be in [16 B, 160 KiB]");
    Carve c{(char *)ws, 0};
    HeatWs h = carve_heat(c, n, nnz, nb);
    const SmoothPatches pp{npatches, rings, vstart, ring_end, lverts, aoff, astart, perm, ladj};
    CSG_TRY(launch_heat_weights(geom, nbr_off, nbr, n, hb, squared_norm, (size_t)nnz, h.w,
                                h.wsum, s));
    int most = 0;
    for (int i = 0; i < hb.nb; ++i)
        most = hb.counts[i] > most ? hb.counts[i] : most;
    // each launch advances every row by up to `rings` iterations and writes all
    // of its vertices (the patch cores partition the mesh)
    const double *cur = values;
    int which = 0;
    for (int it0 = 0; it0 < most; it0 += rings) {
        double *nxt = (it0 + rings >= most) ? out : h.buf[which];
        CSG_TRY(launch_smooth_values_patched(cur, nbr_off, h.w, n, (size_t)nnz, hb, pp, it0,
                                             lds_bytes, nxt, s));
        cur = nxt;
        which ^= 1;
    }
    if (most == 0)
        CSG_HIP_CHECK(hipMemcpyAsync(out, values, sizeof(double) * n * nb,
                                     hipMemcpyDeviceToDevice, s));
    return kOk;
}

// csg_type5_sheets_batch after its batch is set up (host or device counts).
int type5_batch_run(const float *geom, const float *smoothed, int n, int sides, const int *faces,
                    const int *nbr_off, const int *nbr, int nnz, const int *inc_off,
                    const int *inc, int nfw, const HeatBatch &hb, const int *const *sheet_idx,
                    const double *const *sheet_w, int m, int out_dtype, void *out, void *ws,
                    size_t ws_bytes, hipStream_t s)
{
    const int nb = hb.nb;
    if (ws_bytes < csg_type5_sheets_batch_workspace_bytes(n, nnz, sides, nfw, m))
        return fail(kInvalidArgument, "batched type-5 workspace too small");
    for (int k = 0; k < sides; ++k)
        if (!sheet_idx[k] || !sheet_w[k])
            return fail(kInvalidArgument, "batched type-5: side %d has no sheet stencil", k);
    Carve c{(char *)ws, 0};
    float *nr = c.take<float>((size_t)3 * n * sides);
    double *base = c.take<double>((size_t)n * sides);
    HeatWs h = carve_heat(c, n, nnz, nb);
    double *normed = c.take<double>((size_t)nb * n);
    double *sheets = c.take<double>((size_t)nb * m);
    double *mm = c.take<double>((size_t)nb * 64 * 2 + 64);
    for (int k = 0; k < sides; ++k) {
        const size_t g = (size_t)3 * n * k;
        CSG_TRY(launch_vertex_normals(smoothed + g, faces, inc_off, inc, n, nr + g, s));
        CSG_TRY(launch_type5_base(geom + g, smoothed + g, nr + g, n, base + (size_t)n * k, s));
        for (int f = 0; f < nfw; ++f)
            CSG_HIP_CHECK(hipMemcpyAsync(h.buf[0] + (size_t)(k * nfw + f) * n,
                                         base + (size_t)n * k, sizeof(double) * n,
                                         hipMemcpyDeviceToDevice, s));
    }
    double *res = nullptr;
    CSG_TRY(run_heat(geom, nbr_off, nbr, n, nnz, hb, 1, h, &res, s));
    CSG_TRY(launch_minmax_normalise<double>(res, (size_t)n, nb, (size_t)n, 0.0, mm, normed,
                                            (size_t)n, s));
    for (int k = 0; k < sides; ++k)
        CSG_TRY(launch_sheet_gather(normed + (size_t)k * nfw * n, (size_t)n, sheet_idx[k],
                                    sheet_w[k], m, nfw, sheets + (size_t)k * nfw * m, s));
    if (out_dtype == 0)
        return launch_minmax_normalise<float>(sheets, (size_t)m, nb, (size_t)m, 0.0, mm,
                                              (float *)out, (size_t)m, s);
    return launch_minmax_normalise<double>(sheets, (size_t)m, nb, (size_t)m, 0.0, mm,
                                           (double *)out, (size_t)m, s);
}

}  // namespace

extern "C" {

int csg_resample_vertices(const void *x, int dtype, const int *idx, const double *w, int n_out,
                          void *out, void *stream)
{
    if (n_out < 0)
        return fail(kInvalidArgument, "n_out must be non-negative");
    if (dtype == 0)
        return launch_resample<float>((const float *)x, idx, w, n_out, (float *)out, S(stream));
    if (dtype == 1)
        return launch_resample<double>((const double *)x, idx, w, n_out, (double *)out,
                                       S(stream));
    return fail(kInvalidArgument, "dtype must be 0 or 1");
}

size_t csg_edge_mean_workspace_bytes(int nf) { return (size_t)grid_for(nf) * 8 + 1024; }

int csg_edge_sum(const float *v, const int *faces, int nf, double *partial_out, int *nblocks,
                 void *stream)
{
    if (nf < 1)
        return fail(kInvalidArgument, "at least one face is required");
    return launch_edge_sum(v, faces, nf, partial_out, nblocks, S(stream));
}

int csg_smooth_geometry(const float *v, const int *nbr_off, const int *nbr, int n,
                        int iterations, float *out, void *ws, size_t ws_bytes, void *stream)
{
    if (n < 1 || iterations < 0)
        return fail(kInvalidArgument, "invalid smoothing arguments");
    if (ws_bytes < (size_t)3 * n * sizeof(float))
        return fail(kInvalidArgument, "smooth_geometry workspace too small");
    hipStream_t s = S(stream);
    float *bufs[2] = {out, (float *)ws};
    // choose the starting buffer so that the final iterate lands in `out`
    int cur = (iterations % 2 == 0) ? 0 : 1;
    CSG_HIP_CHECK(hipMemcpyAsync(bufs[cur], v, (size_t)3 * n * sizeof(float),
                                 hipMemcpyDeviceToDevice, s));
    for (int it = 0; it < iterations; ++it) {
        CSG_TRY(launch_smooth_geometry(bufs[cur], nbr_off, nbr, n, bufs[cur ^ 1], s));
        cur ^= 1;
    }
    return kOk;
}

// batch meshes [batch][n][3] of one topology, mesh b smoothed iterations[b]
// (host array) times; ws: 3 n batch floats.
size_t csg_smooth_geometry_batch_workspace_bytes(int n, int batch)
{
    if (n < 1 || batch < 1)
        return 0;
    return (size_t)2 * n * batch * sizeof(float4);
}

int csg_smooth_geometry_batch(const float *v, int batch, const int *nbr_off, const int *nbr, int n,
                              const int *iterations, float *out, void *ws, size_t ws_bytes,
                              void *stream)
{
    if (n < 1 || batch < 1 || batch > kMaxSmoothBatch || iterations == nullptr)
        return fail(kInvalidArgument, "invalid batched smoothing arguments (batch <= %d)",
                    kMaxSmoothBatch);
    if (ws_bytes < (size_t)3 * n * batch * sizeof(float))
        return fail(kInvalidArgument, "smooth_geometry_batch workspace too small");
    SmoothBatch sb{};
    sb.count = batch;
    int most = 0;
    for (int b = 0; b < batch; ++b) {
        if (iterations[b] < 0)
            return fail(kInvalidArgument, "negative smoothing iterations");
        sb.iters[b] = iterations[b];
        most = iterations[b] > most ? iterations[b] : most;
    }
    hipStream_t s = S(stream);
    if (ws_bytes >= csg_smooth_geometry_batch_workspace_bytes(n, batch)) {
        float4 *b0 = (float4 *)ws;
        return launch_smooth_geometry4(v, b0, b0 + (size_t)n * batch, nbr_off, nbr, n, sb, most,
                                       out, s);
    }
    float *bufs[2] = {out, (float *)ws};
    const size_t mesh = (size_t)3 * n;
    // start each mesh in the buffer from which its last step writes `out`
    for (int b = 0; b < batch; ++b)
        CSG_HIP_CHECK(hipMemcpyAsync(bufs[iterations[b] & 1] + b * mesh, v + b * mesh,
                                     mesh * sizeof(float), hipMemcpyDeviceToDevice, s));
    for (int it = 0; it < most; ++it)
        CSG_TRY(launch_smooth_geometry_batch(bufs[0], bufs[1], nbr_off, nbr, n, sb, it, s));
    return kOk;
}

int csg_smooth_geometry_patched(const float *v, int batch, int n, const int *iterations,
                                int npatches, int rings, const int *perm, const int *vstart,
                                const int *ring_end, const int *lverts, const int *aoff,
                                const int *astart, const unsigned short *ladj, size_t lds_bytes,
                                float *out, void *ws, size_t ws_bytes, void *stream)
{
    if (n < 1 || batch < 1 || batch > kMaxSmoothBatch || iterations == nullptr)
        return fail(kInvalidArgument, "invalid patched smoothing arguments (batch <= %d)",
                    kMaxSmoothBatch);
    if (npatches < 1 || npatches > 65535 || rings < 1 || rings > 64 || !perm || !vstart ||
        !ring_end || !lverts || !aoff || !astart || !ladj)
        return fail(kInvalidArgument, "invalid smoothing patches");
    if (lds_bytes < sizeof(float4) || lds_bytes > 160 * 1024)
        return fail(kInvalidArgument, "patch LDS size must be in (0, 160 KiB]");
    if (ws_bytes < csg_smooth_geometry_batch_workspace_bytes(n, batch))
        return fail(kInvalidArgument, "patched smoothing workspace too small");
    SmoothBatch sb{};
    sb.count = batch;
    int most = 0;
    for (int b = 0; b < batch; ++b) {
        if (iterations[b] < 0)
            return fail(kInvalidArgument, "negative smoothing iterations");
        sb.iters[b] = iterations[b];
        most = iterations[b] > most ? iterations[b] : most;
    }
    SmoothPatches pp{npatches, rings, vstart, ring_end, lverts, aoff, astart, perm, ladj};
    float4 *b0 = (float4 *)ws;
    return launch_smooth_geometry_patched(v, b0, b0 + (size_t)n * batch, n, pp, sb, most,
                                          lds_bytes, out, S(stream));
}

int csg_vertex_normals(const float *v, const int *faces, const int *inc_off, const int *inc,
                       int n, float *out, void *stream)
{
    if (n < 1)
        return fail(kInvalidArgument, "at least one vertex is required");
    return launch_vertex_normals(v, faces, inc_off, inc, n, out, S(stream));
}

size_t csg_smooth_values_workspace_bytes(int n, int nnz, int nb)
{
    Carve c{nullptr, 0};
    carve_heat(c, n, nnz, nb);
    return c.used;
}

int csg_smooth_values(const double *values, int n, int nb, const float *geom, int meshes,
                      const int *mesh_of, const int *nbr_off, const int *nbr, int nnz,
                      const double *sigma, const int *iterations, int squared_norm, double *out,
                      void *ws, size_t ws_bytes, void *stream)
{
    HeatBatch hb;
    CSG_TRY(make_batch(nb, sigma, iterations, hb, mesh_of, meshes < 1 ? 1 : meshes));
    if (ws_bytes < csg_smooth_values_workspace_bytes(n, nnz, nb))
        return fail(kInvalidArgument, "smooth_values workspace too small");
    hipStream_t s = S(stream);
    Carve c{(char *)ws, 0};
    HeatWs h = carve_heat(c, n, nnz, nb);
    CSG_HIP_CHECK(hipMemcpyAsync(h.buf[0], values, sizeof(double) * n * nb,
                                 hipMemcpyDeviceToDevice, s));
    double *res = nullptr;
    CSG_TRY(run_heat(geom, nbr_off, nbr, n, nnz, hb, squared_norm, h, &res, s));
    CSG_HIP_CHECK(hipMemcpyAsync(out, res, sizeof(double) * n * nb, hipMemcpyDeviceToDevice, s));
    return kOk;
}

int csg_smooth_values_patched(const double *values, int n, int nb, const float *geom, int meshes,
                              const int *mesh_of, const int *nbr_off, const int *nbr, int nnz,
                              const double *sigma, const int *iterations, int squared_norm,
                              int npatches, int rings, const int *perm, const int *vstart,
                              const int *ring_end, const int *lverts, const int *aoff,
                              const int *astart, const unsigned short *ladj, size_t lds_bytes,
                              double *out, void *ws, size_t ws_bytes, void *stream)
{
    HeatBatch hb;
    CSG_TRY(make_batch(nb, sigma, iterations, hb, mesh_of, meshes < 1 ? 1 : meshes));
    return smooth_values_patched_run(values, n, nb, geom, nbr_off, nbr, nnz, hb, squared_norm,
                                     npatches, rings, perm, vstart, ring_end, lverts, aoff,
                                     astart, ladj, lds_bytes, out, ws, ws_bytes, S(stream));
}

size_t csg_minmax_workspace_bytes(int nb) { return (size_t)nb * 64 * 2 * 8 + 1024; }

int csg_minmax_normalise(const double *x, size_t n, int nb, double floor_, int out_dtype,
                         void *out, void *ws, size_t ws_bytes, void *stream)
{
    if (n < 1 || nb < 1)
        return fail(kInvalidArgument, "min/max normalisation needs data");
    if (ws_bytes < csg_minmax_workspace_bytes(nb))
        return fail(kInvalidArgument, "min/max workspace too small");
    if (out_dtype == 0)
        return launch_minmax_normalise<float>(x, n, nb, n, floor_, (double *)ws, (float *)out, n,
                                              S(stream));
    return launch_minmax_normalise<double>(x, n, nb, n, floor_, (double *)ws, (double *)out, n,
                                           S(stream));
}

int csg_sheet_gather(const double *values, int n_values, int nb, const int *idx, const double *w,
                     int m, double *out, void *stream)
{
    if (m < 1 || nb < 1)
        return fail(kInvalidArgument, "sheet gather needs data");
    return launch_sheet_gather(values, (size_t)n_values, idx, w, m, nb, out, S(stream));
}

size_t csg_type5_workspace_bytes(int n, int nnz, int nb, int m)
{
    Carve c{nullptr, 0};
    c.take<float>((size_t)3 * n);  // smoothed
    c.take<float>((size_t)3 * n);  // ping-pong
    c.take<float>((size_t)3 * n);  // normals
    c.take<double>((size_t)n);     // base values
    carve_heat(c, n, nnz, nb);
    c.take<double>((size_t)nb * n);  // normalised values
    c.take<double>((size_t)nb * m);  // raw sheets
    c.take<double>((size_t)nb * 64 * 2 + 64);
    return c.used;
}

// curvature_type5_to_sheet_many: fwhm-50 geometry smoothing, normals, depth
// along normals, per-fwhm value smoothing, min/max, sheet gather, min/max.
// src/cat_surface_gpu/surface_stencil.py:521-550
int csg_type5_sheets(const float *geom, int n, const int *faces, const int *nbr_off,
                     const int *nbr, int nnz, const int *inc_off, const int *inc,
                     int geom_iterations, const float *smoothed, int nb, const double *sigma,
                     const int *iterations, const int *sheet_idx, const double *sheet_w, int m,
                     int out_dtype, void *out, void *ws, size_t ws_bytes, void *stream)
{
    HeatBatch hb;
    CSG_TRY(make_batch(nb, sigma, iterations, hb));
    if (ws_bytes < csg_type5_workspace_bytes(n, nnz, nb, m))
        return fail(kInvalidArgument, "type-5 workspace too small");
    hipStream_t s = S(stream);
    Carve c{(char *)ws, 0};
    float *sm = c.take<float>((size_t)3 * n);
    float *pp = c.take<float>((size_t)3 * n);
    float *nr = c.take<float>((size_t)3 * n);
    double *base = c.take<double>((size_t)n);
    HeatWs h = carve_heat(c, n, nnz, nb);
    double *normed = c.take<double>((size_t)nb * n);
    double *sheets = c.take<double>((size_t)nb * m);
    double *mm = c.take<double>((size_t)nb * 64 * 2 + 64);
    if (smoothed)
        sm = const_cast<float *>(smoothed);
    else
        CSG_TRY(csg_smooth_geometry(geom, nbr_off, nbr, n, geom_iterations, sm, pp,
                                    (size_t)3 * n * sizeof(float), stream));
    CSG_TRY(launch_vertex_normals(sm, faces, inc_off, inc, n, nr, s));
    CSG_TRY(launch_type5_base(geom, sm, nr, n, base, s));
    for (int b = 0; b < nb; ++b)
        CSG_HIP_CHECK(hipMemcpyAsync(h.buf[0] + (size_t)b * n, base, sizeof(double) * n,
                                     hipMemcpyDeviceToDevice, s));
    double *res = nullptr;
    CSG_TRY(run_heat(geom, nbr_off, nbr, n, nnz, hb, 1, h, &res, s));
    CSG_TRY(launch_minmax_normalise<double>(res, (size_t)n, nb, (size_t)n, 0.0, mm, normed,
                                            (size_t)n, s));
    CSG_TRY(launch_sheet_gather(normed, (size_t)n, sheet_idx, sheet_w, m, nb, sheets, s));
    if (out_dtype == 0)
        return launch_minmax_normalise<float>(sheets, (size_t)m, nb, (size_t)m, 0.0, mm,
                                              (float *)out, (size_t)m, s);
    return launch_minmax_normalise<double>(sheets, (size_t)m, nb, (size_t)m, 0.0, mm,
                                           (double *)out, (size_t)m, s);
}

int csg_curvature_type2(const float *geom, const float *normals, int n, const int *nbr_off,
                        const int *nbr, double *out, void *stream)
{
    if (n < 1 || !geom || !normals || !nbr_off || !nbr || !out)
        return fail(kInvalidArgument, "invalid type-2 curvature arguments");
    return launch_curvature_type2(geom, normals, n, nbr_off, nbr, out, S(stream));
}

size_t csg_curvature_sheet_workspace_bytes(int n, int nnz, int m)
{
    if (n < 1 || m < 1)
        return 0;
    Carve c{nullptr, 0};
    c.take<float>((size_t)3 * n);  // smoothed geometry
    c.take<float>((size_t)3 * n);  // its ping-pong buffer
    c.take<float>((size_t)3 * n);  // normals
    carve_heat(c, n, nnz, 1);
    c.take<double>((size_t)n);      // normalised values
    c.take<double>((size_t)m);      // raw sheet
    c.take<double>(64 * 2 + 64);
    return c.used;
}

// curvature_to_sheet for one fwhm (surface_stencil.py:624-644): type 5 (depth along
// the normals of the fwhm-50 smoothed geometry, :507-519) or type 2 (:552-622), the
// values smoothed with |d|^2 heat weights (smooth_values), normalised, gathered to
// the sheet and normalised again.
int csg_curvature_sheet(const float *geom, int n, const int *faces, const int *nbr_off,
                        const int *nbr, int nnz, const int *inc_off, const int *inc,
                        int curvtype, int geom_iterations, double sigma, int iterations,
                        const int *sheet_idx, const double *sheet_w, int m, int out_dtype,
                        void *out, void *ws, size_t ws_bytes, void *stream)
{
    if (n < 1 || m < 1 || !geom || !faces || !nbr_off || !nbr || !inc_off || !inc ||
        !sheet_idx || !sheet_w || !out)
        return fail(kInvalidArgument, "invalid curvature sheet arguments");
    if (curvtype != 2 && curvtype != 5)
        return fail(kUnsupported, "only upstream curvtype 2 and 5 exist, got %d", curvtype);
    if (out_dtype != 0 && out_dtype != 1)
        return fail(kUnsupported, "curvature sheet: dtype code %d", out_dtype);
    HeatBatch hb;
    CSG_TRY(make_batch(1, &sigma, &iterations, hb));
    if (ws_bytes < csg_curvature_sheet_workspace_bytes(n, nnz, m))
        return fail(kInvalidArgument, "curvature sheet workspace too small");
    hipStream_t s = S(stream);
    Carve c{(char *)ws, 0};
    float *sm = c.take<float>((size_t)3 * n);
    float *pp = c.take<float>((size_t)3 * n);
    float *nr = c.take<float>((size_t)3 * n);
    HeatWs h = carve_heat(c, n, nnz, 1);
    double *normed = c.take<double>((size_t)n);
    double *sheet = c.take<double>((size_t)m);
    double *mm = c.take<double>(64 * 2 + 64);
    if (curvtype == 5) {
        CSG_TRY(csg_smooth_geometry(geom, nbr_off, nbr, n, geom_iterations, sm, pp,
                                    (size_t)3 * n * sizeof(float), stream));
        CSG_TRY(launch_vertex_normals(sm, faces, inc_off, inc, n, nr, s));
        CSG_TRY(launch_type5_base(geom, sm, nr, n, h.buf[0], s));
    } else {
        CSG_TRY(launch_vertex_normals(geom, faces, inc_off, inc, n, nr, s));
        CSG_TRY(launch_curvature_type2(geom, nr, n, nbr_off, nbr, h.buf[0], s));
    }
    double *res = nullptr;
    CSG_TRY(run_heat(geom, nbr_off, nbr, n, nnz, hb, 0, h, &res, s));
    CSG_TRY(launch_minmax_normalise<double>(res, (size_t)n, 1, (size_t)n, 0.0, mm, normed,
                                            (size_t)n, s));
    CSG_TRY(launch_sheet_gather(normed, (size_t)n, sheet_idx, sheet_w, m, 1, sheet, s));
    if (out_dtype == 0)
        return launch_minmax_normalise<float>(sheet, (size_t)m, 1, (size_t)m, 0.0, mm,
                                              (float *)out, (size_t)m, s);
    return launch_minmax_normalise<double>(sheet, (size_t)m, 1, (size_t)m, 0.0, mm,
                                           (double *)out, (size_t)m, s);
}

// ---- the same operators with device-resident heat parameters.  The counts
// (and widths) come from csg_heat_parameters on the device; the host launches
// up to `bound` iterations and rows or meshes past their count are copied
// through, so the values equal the host-count entries' whenever every count
// is at most `bound` (the caller checks that on the device).

int csg_flag_above(const int *values, int n, int bound, int *flag, void *stream)
{
    if (n < 0 || (n > 0 && (!values || !flag)))
        return fail(kInvalidArgument, "invalid flag_above arguments");
    if (n == 0)
        return kOk;
    hipLaunchKernelGGL(k_flag_above, dim3(grid_for(n)), dim3(kBlock), 0, S(stream), values, n,
                       bound, flag);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int csg_heat_parameters(const float *v, const int *faces, int nf, const double *fwhm, int nfw,
                        const int *root_n, const double *root_v, int nroots, int *iters_out,
                        double *sigma_out, double *em_out, void *ws, size_t ws_bytes,
                        void *stream)
{
    if (nf < 1 || nfw < 0 || nfw > kMaxHeatFwhm || (nfw > 0 && (!fwhm || !iters_out ||
                                                                 !sigma_out)))
        return fail(kInvalidArgument, "heat parameters: nf >= 1 and at most %d fwhms",
                    kMaxHeatFwhm);
    if (nroots < 0 || (nroots > 0 && (!root_n || !root_v)))
        return fail(kInvalidArgument, "heat parameters: invalid root table");
    if (ws_bytes < csg_edge_mean_workspace_bytes(nf))
        return fail(kInvalidArgument, "heat parameters workspace too small");
    hipStream_t s = S(stream);
    int nb = 0;
    CSG_TRY(launch_edge_sum(v, faces, nf, (double *)ws, &nb, s));
    return launch_heat_parameters((const double *)ws, nb, nf, fwhm, nfw, root_n, root_v, nroots,
                                  iters_out, sigma_out, em_out, s);
}

int csg_heat_parameters_from_edge_mean(const double *em, const double *fwhm, int nfw,
                                       const int *root_n, const double *root_v, int nroots,
                                       int *iters_out, double *sigma_out, void *stream)
{
    if (!em || nfw < 1 || nfw > kMaxHeatFwhm || !fwhm || !iters_out || !sigma_out)
        return fail(kInvalidArgument, "heat parameters: an edge mean and 1..%d fwhms",
                    kMaxHeatFwhm);
    if (nroots < 0 || (nroots > 0 && (!root_n || !root_v)))
        return fail(kInvalidArgument, "heat parameters: invalid root table");
    return launch_heat_parameters(em, -1, 1, fwhm, nfw, root_n, root_v, nroots, iters_out,
                                  sigma_out, nullptr, S(stream));
}

int csg_smooth_geometry_batch_dev(const float *v, int batch, const int *nbr_off, const int *nbr,
                                  int n, const int *iters_dev, int bound, float *out, void *ws,
                                  size_t ws_bytes, void *stream)
{
    if (n < 1 || batch < 1 || batch > kMaxSmoothBatch || !iters_dev || bound < 0)
        return fail(kInvalidArgument, "invalid batched smoothing arguments (batch <= %d)",
                    kMaxSmoothBatch);
    if (ws_bytes < csg_smooth_geometry_batch_workspace_bytes(n, batch))
        return fail(kInvalidArgument, "smooth_geometry_batch workspace too small");
    SmoothBatch sb{};
    sb.count = batch;
    for (int b = 0; b < batch; ++b)
        sb.iters[b] = bound;
    sb.diters = iters_dev;
    float4 *b0 = (float4 *)ws;
    return launch_smooth_geometry4(v, b0, b0 + (size_t)n * batch, nbr_off, nbr, n, sb, bound, out,
                                   S(stream));
}

int csg_smooth_geometry_patched_dev(const float *v, int batch, int n, const int *iters_dev,
                                    int bound, int npatches, int rings, const int *perm,
                                    const int *vstart, const int *ring_end, const int *lverts,
                                    const int *aoff, const int *astart,
                                    const unsigned short *ladj, size_t lds_bytes, float *out,
                                    void *ws, size_t ws_bytes, void *stream)
{
    if (n < 1 || batch < 1 || batch > kMaxSmoothBatch || !iters_dev || bound < 0)
        return fail(kInvalidArgument, "invalid patched smoothing arguments (batch <= %d)",
                    kMaxSmoothBatch);
    if (npatches < 1 || npatches > 65535 || rings < 1 || rings > 64 || !perm || !vstart ||
        !ring_end || !lverts || !aoff || !astart || !ladj)
        return fail(kInvalidArgument, "invalid smoothing patches");
    if (lds_bytes < sizeof(float4) || lds_bytes > 160 * 1024)
        return fail(kInvalidArgument, "patch LDS size must be in (0, 160 KiB]");
    if (ws_bytes < csg_smooth_geometry_batch_workspace_bytes(n, batch))
        return fail(kInvalidArgument, "patched smoothing workspace too small");
    SmoothBatch sb{};
    sb.count = batch;
    for (int b = 0; b < batch; ++b)
        sb.iters[b] = bound;
    sb.diters = iters_dev;
    SmoothPatches pp{npatches, rings, vstart, ring_end, lverts, aoff, astart, perm, ladj};
    float4 *b0 = (float4 *)ws;
    return launch_smooth_geometry_patched(v, b0, b0 + (size_t)n * batch, n, pp, sb, bound,
                                          lds_bytes, out, S(stream));
}

int csg_smooth_values_dev(const double *values, int n, int nb, const float *geom, int meshes,
                          const int *mesh_of, const int *nbr_off, const int *nbr, int nnz,
                          const double *sigma_dev, const int *iters_dev, int bound,
                          int squared_norm, double *out, void *ws, size_t ws_bytes, void *stream)
{
    HeatBatch hb;
    CSG_TRY(make_batch_dev(nb, sigma_dev, iters_dev, bound, hb, mesh_of,
                           meshes < 1 ? 1 : meshes));
    if (ws_bytes < csg_smooth_values_workspace_bytes(n, nnz, nb))
        return fail(kInvalidArgument, "smooth_values workspace too small");
    hipStream_t s = S(stream);
    Carve c{(char *)ws, 0};
    HeatWs h = carve_heat(c, n, nnz, nb);
    CSG_HIP_CHECK(hipMemcpyAsync(h.buf[0], values, sizeof(double) * n * nb,
                                 hipMemcpyDeviceToDevice, s));
    double *res = nullptr;
    CSG_TRY(run_heat(geom, nbr_off, nbr, n, nnz, hb, squared_norm, h, &res, s));
    CSG_HIP_CHECK(hipMemcpyAsync(out, res, sizeof(double) * n * nb, hipMemcpyDeviceToDevice, s));
    return kOk;
}

int csg_smooth_values_patched_dev(const double *values, int n, int nb, const float *geom,
                                  int meshes, const int *mesh_of, const int *nbr_off,
                                  const int *nbr, int nnz, const double *sigma_dev,
                                  const int *iters_dev, int bound, int squared_norm, int npatches,
                                  int rings, const int *perm, const int *vstart,
                                  const int *ring_end, const int *lverts, const int *aoff,
                                  const int *astart, const unsigned short *ladj,
                                  size_t lds_bytes, double *out, void *ws, size_t ws_bytes,
                                  void *stream)
{
    HeatBatch hb;
    CSG_TRY(make_batch_dev(nb, sigma_dev, iters_dev, bound, hb, mesh_of,
                           meshes < 1 ? 1 : meshes));
    return smooth_values_patched_run(values, n, nb, geom, nbr_off, nbr, nnz, hb, squared_norm,
                                     npatches, rings, perm, vstart, ring_end, lverts, aoff,
                                     astart, ladj, lds_bytes, out, ws, ws_bytes, S(stream));
}

size_t csg_type5_sheets_batch_workspace_bytes(int n, int nnz, int sides, int nfw, int m)
{
    if (n < 1 || sides < 1 || nfw < 1 || m < 1)
        return 0;
    const int nb = sides * nfw;
    Carve c{nullptr, 0};
    c.take<float>((size_t)3 * n * sides);  // normals
    c.take<double>((size_t)n * sides);     // base values
    carve_heat(c, n, nnz, nb);
    c.take<double>((size_t)nb * n);  // normalised values
    c.take<double>((size_t)nb * m);  // raw sheets
    c.take<double>((size_t)nb * 64 * 2 + 64);
    return c.used;
}

// csg_type5_sheets for `sides` meshes of one topology with their fwhm-50
// smoothed copies given (geom, smoothed [sides][n][3]); every side has nfw
// rows (sigma / iterations [sides][nfw]) and its own sheet stencil (host arrays
// of device pointers, m points each); out [sides][nfw][m].  All rows share one
// value-smoothing loop; each row's values equal the single-mesh call's.
int csg_type5_sheets_batch(const float *geom, const float *smoothed, int n, int sides,
                           const int *faces, const int *nbr_off, const int *nbr, int nnz,
                           const int *inc_off, const int *inc, int nfw, const double *sigma,
                           const int *iterations, const int *const *sheet_idx,
                           const double *const *sheet_w, int m, int out_dtype, void *out,
                           void *ws, size_t ws_bytes, void *stream)
{
    if (n < 1 || sides < 1 || nfw < 1 || m < 1 || !geom || !smoothed || !sigma || !iterations ||
        !sheet_idx || !sheet_w || !out)
        return fail(kInvalidArgument, "invalid batched type-5 arguments");
    if (out_dtype != 0 && out_dtype != 1)
        return fail(kUnsupported, "type-5 sheets: dtype code %d", out_dtype);
    if ((long long)sides * nfw > kMaxHeatBatch)
        return fail(kInvalidArgument, "batched type-5: sides x fwhms must be <= %d, got %lld",
                    kMaxHeatBatch, (long long)sides * nfw);
    const int nb = sides * nfw;
    int mesh_of[kMaxHeatBatch];
    for (int b = 0; b < nb; ++b)
        mesh_of[b] = b / nfw;
    HeatBatch hb;
    CSG_TRY(make_batch(nb, sigma, iterations, hb, mesh_of, sides));
    return type5_batch_run(geom, smoothed, n, sides, faces, nbr_off, nbr, nnz, inc_off, inc, nfw,
                           hb, sheet_idx, sheet_w, m, out_dtype, out, ws, ws_bytes, S(stream));
}

int csg_type5_sheets_batch_dev(const float *geom, const float *smoothed, int n, int sides,
                               const int *faces, const int *nbr_off, const int *nbr, int nnz,
                               const int *inc_off, const int *inc, int nfw,
                               const double *sigma_dev, const int *iters_dev, int bound,
                               const int *const *sheet_idx, const double *const *sheet_w, int m,
                               int out_dtype, void *out, void *ws, size_t ws_bytes, void *stream)
{
    if (n < 1 || sides < 1 || nfw < 1 || m < 1 || !geom || !smoothed || !sigma_dev ||
        !iters_dev || !sheet_idx || !sheet_w || !out)
        return fail(kInvalidArgument, "invalid batched type-5 arguments");
    if (out_dtype != 0 && out_dtype != 1)
        return fail(kUnsupported, "type-5 sheets: dtype code %d", out_dtype);
    if ((long long)sides * nfw > kMaxHeatBatch)
        return fail(kInvalidArgument, "batched type-5: sides x fwhms must be <= %d, got %lld",
                    kMaxHeatBatch, (long long)sides * nfw);
    const int nb = sides * nfw;
    int mesh_of[kMaxHeatBatch];
    for (int b = 0; b < nb; ++b)
        mesh_of[b] = b / nfw;
    HeatBatch hb;
    CSG_TRY(make_batch_dev(nb, sigma_dev, iters_dev, bound, hb, mesh_of, sides));
    return type5_batch_run(geom, smoothed, n, sides, faces, nbr_off, nbr, nnz, inc_off, inc, nfw,
                           hb, sheet_idx, sheet_w, m, out_dtype, out, ws, ws_bytes, S(stream));
}

namespace {

// Devices on which an XCD-local SOR left a group without workgroups.
std::atomic<unsigned long long> g_xcd_local_unusable{0};

int current_device()
{
    int dev = 0;
    return hipGetDevice(&dev) == hipSuccess ? dev : 0;
}

bool xcd_local_unusable()
{
    const int dev = current_device();
    return dev < 64 && ((g_xcd_local_unusable.load() >> dev) & 1ull);
}

void mark_xcd_local_unusable()
{
    const int dev = current_device();
    if (dev < 64)
        g_xcd_local_unusable.fetch_or(1ull << dev);
}

DepthArgs carve_depth(Carve &c, int batch, int n, int nf, int nnz, int width)
{
    DepthArgs a{};
    a.fnormal = c.take<double>((size_t)batch * 9 * nf);
    a.farea = c.take<double>((size_t)batch * 3 * nf);
    a.coo = c.take<double>((size_t)batch * 9 * nf);
    a.normals = c.take<double>((size_t)batch * 3 * n);
    a.areas = c.take<double>((size_t)batch * n);
    a.lap = c.take<double>((size_t)batch * nnz);
    a.sys = c.take<double>((size_t)batch * nnz);
    a.mc = c.take<double>((size_t)batch * n);
    a.diag = c.take<double>((size_t)batch * n);
    a.rhs = c.take<double>((size_t)batch * n);
    a.partial = c.take<double>((size_t)batch * 2 * grid_for(n) + 64);
    a.ell_val = c.take<double>((size_t)batch * width * n);
    a.rhs_c = c.take<double>((size_t)batch * n);
    a.diag_c = c.take<double>((size_t)batch * n);
    a.barrier = c.take<unsigned>((size_t)kBarrierSlots * kBarrierSlotWords);
    a.x_alt = c.take<double>((size_t)batch * n);
    return a;
}

}  // namespace

size_t csg_depth_workspace_bytes(int n, int nf, int nnz, int batch, int ell_width)
{
    Carve c{nullptr, 0};
    carve_depth(c, batch, n, nf, nnz, ell_width);
    return c.used;
}

int csg_set_sor_placement(int mode)
{
    if (mode < 0 || mode > 2)
        return fail(kInvalidArgument, "SOR placement must be 0, 1 or 2, got %d", mode);
    return set_sor_xcd_local(mode);
}

int csg_set_sor_force_xcd(int xcd)
{
    if (xcd >= kXcds)
        return fail(kInvalidArgument, "XCD must be below %d, got %d", kXcds, xcd);
    return set_sor_force_xcd(xcd);
}

int csg_reset_sor_xcd_check(void)
{
    const int dev = current_device();
    const unsigned long long bit = dev < 64 ? (1ull << dev) : 0ull;
    return (g_xcd_local_unusable.fetch_and(~bit) & bit) ? 1 : 0;
}

// Depth potential of the rotation feature for `batch` meshes of one topology:
// (L + alpha·diag(area)) x = rhs by colour-ordered SOR (rotation_feature.py:
// 253-317).  Synchronises once at the end to check the persistent solve.
int csg_depth_potential(const float *pts, int batch, int n, const int *faces, int nf,
                        const int *inc_off, const int *inc, const int *row_off, const int *cols,
                        const int *contrib_off, const int *contrib, int nnz,
                        const int *colour_off, const int *colour_verts, int ncolours,
                        const int *ell_cols, const int *ell_src, int ell_width, double alpha,
                        int max_iter, double omega, double *x, double *relres, void *ws,
                        size_t ws_bytes, void *stream)
{
    return csg_depth_potential_patched(pts, batch, n, faces, nf, inc_off, inc, row_off, cols,
                                       contrib_off, contrib, nnz, colour_off, colour_verts,
                                       ncolours, ell_cols, ell_src, ell_width, alpha, max_iter,
                                       omega, nullptr, x, relres, ws, ws_bytes, stream);
}

// csg_depth_potential_patched; with status (device int[2]) nothing waits for
// the host: the time-out and placement verdicts stay on the device and a
// misplaced run is not repeated.
static int depth_patched_impl(const float *pts, int batch, int n, const int *faces, int nf,
                              const int *inc_off, const int *inc, const int *row_off,
                              const int *cols, const int *contrib_off, const int *contrib,
                              int nnz, const int *colour_off, const int *colour_verts,
                              int ncolours, const int *ell_cols, const int *ell_src,
                              int ell_width, double alpha, int max_iter, double omega,
                              const csg_sor_patches *patches, double *x, double *relres,
                              int *status, void *ws, size_t ws_bytes, void *stream)
{
    SorPatches pp{};
    if (patches) {
        const csg_sor_patches &q = *patches;
        if (q.count < 1 || q.rings < 1 || q.max_local < 1 || q.max_rows < 0 || !q.vstart ||
            !q.ncore || !q.lverts || !q.estart || !q.e_row || !q.e_local || !q.cstart ||
            !q.active || !q.e_lcol)
            return fail(kInvalidArgument, "invalid SOR patches");
        pp = SorPatches{q.count,  q.rings, q.max_local, q.max_rows, q.vstart,
                        q.ncore,  q.lverts, q.estart,   q.e_row,    q.e_local,
                        q.cstart, q.active, q.e_lcol};
    }
    if (n < 1 || nf < 1 || batch < 1 || max_iter < 0 || colour_off == nullptr ||
        ell_width < 1)
        return fail(kInvalidArgument, "invalid depth-potential arguments");
    if (ncolours < 1 || ncolours > kMaxColours)
        return fail(kInvalidArgument, "colour count must be in [1, %d], got %d", kMaxColours,
                    ncolours);
    if (ws_bytes < csg_depth_workspace_bytes(n, nf, nnz, batch, ell_width))
        return fail(kInvalidArgument, "depth-potential workspace too small");
    hipStream_t s = S(stream);
    Carve c{(char *)ws, 0};
    DepthArgs a = carve_depth(c, batch, n, nf, nnz, ell_width);
    a.pts = pts;
    a.faces = faces;
    a.inc_off = inc_off;
    a.inc = inc;
    a.row_off = row_off;
    a.cols = cols;
    a.contrib_off = contrib_off;
    a.contrib = contrib;
    a.colour_verts = colour_verts;
    a.ell_cols = ell_cols;
    a.ell_src = ell_src;
    a.ell_width = ell_width;
    a.batch = batch;
    a.n = n;
    a.nf = nf;
    a.nnz = nnz;
    a.alpha = alpha;
    a.x = x;
    SorColours co{};
    co.count = ncolours;
    for (int i = 0; i <= ncolours; ++i)
        co.off[i] = colour_off[i];
    CSG_TRY(depth_setup(a, s));
    SorRun run{};
    const bool allow_local = !xcd_local_unusable();
    CSG_TRY(depth_sor(a, co, max_iter, omega, patches ? &pp : nullptr, allow_local, &run, s));
    if (status) {
        CSG_TRY(sor_status(a, run, status, s));
        if (relres)
            CSG_TRY(depth_residual(a, relres, s));
        return kOk;
    }
    int rc = kOk;
    if (!sor_placement_complete(a, run, s, &rc)) {
        if (rc != kOk)
            return rc;
        // some XCD group never ran (the device does not spread workgroups over
        // eight XCDs): solve again on the device-wide barrier, and keep to it
        mark_xcd_local_unusable();
        CSG_TRY(depth_setup(a, s));
        CSG_TRY(depth_sor(a, co, max_iter, omega, nullptr, false, &run, s));
    }
    if (relres)
        CSG_TRY(depth_residual(a, relres, s));
    int timeout = 0;
    const unsigned *flag = a.barrier + (size_t)(kBarrierSlots - 1) * kBarrierSlotWords;
    CSG_HIP_CHECK(hipMemcpyAsync(&timeout, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    CSG_HIP_CHECK(hipStreamSynchronize(s));
    if (timeout)
        return fail(kRuntimeError, "depth-potential SOR grid barrier timed out");
    return kOk;
}

int csg_depth_potential_patched(const float *pts, int batch, int n, const int *faces, int nf,
                                const int *inc_off, const int *inc, const int *row_off,
                                const int *cols, const int *contrib_off, const int *contrib,
                                int nnz, const int *colour_off, const int *colour_verts,
                                int ncolours, const int *ell_cols, const int *ell_src,
                                int ell_width, double alpha, int max_iter, double omega,
                                const csg_sor_patches *patches, double *x, double *relres,
                                void *ws, size_t ws_bytes, void *stream)
{
    return depth_patched_impl(pts, batch, n, faces, nf, inc_off, inc, row_off, cols, contrib_off,
                              contrib, nnz, colour_off, colour_verts, ncolours, ell_cols, ell_src,
                              ell_width, alpha, max_iter, omega, patches, x, relres, nullptr, ws,
                              ws_bytes, stream);
}

int csg_depth_potential_patched_deferred(const float *pts, int batch, int n, const int *faces,
                                         int nf, const int *inc_off, const int *inc,
                                         const int *row_off, const int *cols,
                                         const int *contrib_off, const int *contrib, int nnz,
                                         const int *colour_off, const int *colour_verts,
                                         int ncolours, const int *ell_cols, const int *ell_src,
                                         int ell_width, double alpha, int max_iter, double omega,
                                         const csg_sor_patches *patches, double *x,
                                         double *relres, int *status, void *ws, size_t ws_bytes,
                                         void *stream)
{
    if (!status)
        return fail(kInvalidArgument, "the deferred depth potential needs a status array");
    return depth_patched_impl(pts, batch, n, faces, nf, inc_off, inc, row_off, cols, contrib_off,
                              contrib, nnz, colour_off, colour_verts, ncolours, ell_cols, ell_src,
                              ell_width, alpha, max_iter, omega, patches, x, relres, status, ws,
                              ws_bytes, stream);
}

size_t csg_depth_pcg_workspace_bytes(int n, int nf, int nnz, int batch)
{
    Carve c{nullptr, 0};
    carve_depth(c, batch, n, nf, nnz, 1);
    return c.used + pcg_workspace_bytes(n, batch);
}

// Depth potential by Jacobi-preconditioned CG (rotation_feature.py:193-250).
// Synchronises once at the end to check the persistent solve and to return the
// iteration counts.
int csg_depth_potential_pcg(const float *pts, int batch, int n, const int *faces, int nf,
                            const int *inc_off, const int *inc, const int *row_off,
                            const int *cols, const int *contrib_off, const int *contrib, int nnz,
                            double alpha, int max_iter, double tol, double *x, double *relres,
                            int *iterations_host, void *ws, size_t ws_bytes, void *stream)
{
    if (n < 1 || nf < 1 || batch < 1 || max_iter < 0 || !x || !relres || !(tol >= 0.0))
        return fail(kInvalidArgument, "invalid depth-potential arguments");
    if (ws_bytes < csg_depth_pcg_workspace_bytes(n, nf, nnz, batch))
        return fail(kInvalidArgument, "depth-potential workspace too small");
    hipStream_t s = S(stream);
    Carve c{(char *)ws, 0};
    DepthArgs a = carve_depth(c, batch, n, nf, nnz, 1);
    double *pcg_ws = c.take<double>(pcg_workspace_bytes(n, batch) / sizeof(double));
    a.pts = pts;
    a.faces = faces;
    a.inc_off = inc_off;
    a.inc = inc;
    a.row_off = row_off;
    a.cols = cols;
    a.contrib_off = contrib_off;
    a.contrib = contrib;
    a.batch = batch;
    a.n = n;
    a.nf = nf;
    a.nnz = nnz;
    a.alpha = alpha;
    a.x = x;
    // iteration counts in the partial-sum area of the setup (not used by the solve)
    int *iters_dev = reinterpret_cast<int *>(a.partial);
    CSG_TRY(depth_setup(a, s));
    SorRun run{};
    const bool allow_local = !xcd_local_unusable();
    CSG_TRY(depth_pcg(a, pcg_ws, max_iter, tol, allow_local, sor_force_xcd(), iters_dev, relres,
                      &run, s));
    int rc = kOk;
    if (!sor_placement_complete(a, run, s, &rc)) {
        if (rc != kOk)
            return rc;
        mark_xcd_local_unusable();
        CSG_TRY(depth_setup(a, s));
        CSG_TRY(depth_pcg(a, pcg_ws, max_iter, tol, false, -1, iters_dev, relres, &run, s));
    }
    int timeout = 0;
    const unsigned *flag = a.barrier + (size_t)(kBarrierSlots - 1) * kBarrierSlotWords;
    CSG_HIP_CHECK(hipMemcpyAsync(&timeout, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    if (iterations_host)
        CSG_HIP_CHECK(hipMemcpyAsync(iterations_host, iters_dev, sizeof(int) * (size_t)batch,
                                     hipMemcpyDeviceToHost, s));
    CSG_HIP_CHECK(hipStreamSynchronize(s));
    if (timeout)
        return fail(kRuntimeError, "depth-potential PCG group barrier timed out");
    return kOk;
}

int csg_apply_flow(const float *unit, int n, const void *flow, int nx, int ny, int dtype,
                   int inverse, void *out, void *stream)
{
    if (nx < 1 || ny < 1 || n < 0)
        return fail(kInvalidArgument, "invalid warp arguments");
    if (dtype == 0)
        return launch_apply_flow<float>(unit, n, (const float *)flow, nx, ny, inverse,
                                        (float *)out, S(stream));
    if (dtype == 1)
        return launch_apply_flow<double>(unit, n, (const double *)flow, nx, ny, inverse,
                                         (double *)out, S(stream));
    return fail(kInvalidArgument, "dtype must be 0 or 1");
}

int csg_normalise_points(const float *in, int n, float *out, void *stream)
{
    if (n < 0)
        return fail(kInvalidArgument, "point count must be non-negative");
    return launch_normalise_f32(in, n, out, S(stream));
}

int csg_sphere_centroids(const float *const *spheres, const int *counts, int count,
                         double *centroids, void *stream)
{
    if (count < 0 || (count > 0 && (!spheres || !counts || !centroids)))
        return fail(kInvalidArgument, "invalid sphere centroid arguments");
    for (int first = 0; first < count; first += kMaxSpheres) {
        SphereList list{};
        list.count = std::min(kMaxSpheres, count - first);
        for (int k = 0; k < list.count; ++k) {
            if (!spheres[first + k] || counts[first + k] < 1)
                return fail(kInvalidArgument, "every sphere needs vertices");
            list.ptr[k] = spheres[first + k];
            list.n[k] = counts[first + k];
        }
        const int rc = launch_sphere_centroids(list, centroids + 3 * first, S(stream));
        if (rc != kOk)
            return rc;
    }
    return kOk;
}

int csg_prepare_sphere(const float *in, int n, const double *centroid, const double *rotation,
                       float *out, int *bad, void *stream)
{
    if (n < 1 || in == nullptr || out == nullptr || (centroid && !bad))
        return fail(kInvalidArgument, "invalid sphere preparation arguments");
    return launch_sphere_prepare(in, n, centroid, rotation, nullptr, out, bad, S(stream));
}

int csg_prepare_sphere_dev(const float *in, int n, const double *centroid,
                           const double *rotation_dev, float *out, int *bad, void *stream)
{
    if (n < 1 || in == nullptr || out == nullptr || rotation_dev == nullptr ||
        (centroid && !bad))
        return fail(kInvalidArgument, "invalid sphere preparation arguments");
    return launch_sphere_prepare(in, n, centroid, nullptr, rotation_dev, out, bad, S(stream));
}

int csg_rotation_matrices(const double *angles, int count, double *out, void *stream)
{
    if (count < 0 || (count > 0 && (!angles || !out)))
        return fail(kInvalidArgument, "invalid rotation matrix arguments");
    return launch_rotation_matrices(angles, count, out, S(stream));
}

int csg_average_xz(const void *avg, const void *main_, int n, int dtype, const double *unrotate,
                   float *out, void *stream)
{
    if (n < 0 || unrotate == nullptr)
        return fail(kInvalidArgument, "invalid average arguments");
    if (dtype == 0)
        return launch_average_xz<float>((const float *)avg, (const float *)main_, n, unrotate,
                                        out, S(stream));
    if (dtype == 1)
        return launch_average_xz<double>((const double *)avg, (const double *)main_, n, unrotate,
                                         out, S(stream));
    return fail(kInvalidArgument, "dtype must be 0 or 1");
}

}  // extern "C"
