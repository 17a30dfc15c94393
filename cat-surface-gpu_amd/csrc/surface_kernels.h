// Launch wrappers for the surface-side kernels (internal to libcsg_hip).
#pragma once

#include "csg_common.h"

namespace csg {

// Up to kMaxHeatBatch value-smoothing problems on meshes of one topology,
// each with its own iteration count, Gaussian width and geometry.
constexpr int kMaxHeatBatch = 32;
// With dcounts / dsigma set, the counts and widths are read from device
// memory instead (computed there by csg_heat_parameters); the host then
// launches up to a bound and rows past their count are copied through.
struct HeatBatch {
    int nb;
    int counts[kMaxHeatBatch];
    double sigma[kMaxHeatBatch];
    int mesh[kMaxHeatBatch];  // geometry of each problem: mesh[b] * 3 n floats in
    const int *dcounts = nullptr;
    const double *dsigma = nullptr;
};

// Up to kMaxSmoothBatch meshes of one topology smoothed together, each with
// its own number of steps (from device memory when diters is set).
constexpr int kMaxSmoothBatch = 32;
struct SmoothBatch {
    int count;
    int iters[kMaxSmoothBatch];
    const int *diters = nullptr;
};

__device__ __forceinline__ int heat_count(const HeatBatch &hb, int b)
{
    return hb.dcounts ? hb.dcounts[b] : hb.counts[b];
}

__device__ __forceinline__ double heat_sigma(const HeatBatch &hb, int b)
{
    return hb.dsigma ? hb.dsigma[b] : hb.sigma[b];
}

__device__ __forceinline__ int smooth_iters(const SmoothBatch &sb, int b)
{
    return sb.diters ? sb.diters[b] : sb.iters[b];
}

// Edge mean and heat-kernel parameters of one mesh on the device: em = (the
// nblocks edge-sum partials added in order) / (3 nf); for each fwhm f,
// iterations = max(1, ceil(f^2 0.541011 / em^2)), sigma = f 0.7355345 /
// iterations^0.5 (surface_stencil.py:349-370).  The host's iterations^0.5 is
// its libm pow, which differs from the correctly rounded square root by one
// ulp for a few integers: those (root_n ascending, root_v their pow values)
// are passed in, and every other count uses sqrt.
constexpr int kMaxHeatFwhm = 8;
int launch_heat_parameters(const double *partial, int nblocks, int nf, const double *fwhm,
                           int nfw, const int *root_n, const double *root_v, int nroots,
                           int *iters, double *sigma, double *em, hipStream_t s);

// Patches for k_smooth_geometry_patch (device arrays).  The working buffers
// hold the vertices in the order perm (working index i = vertex perm[i]), the
// patches' cores one after another; patch p's local vertices are working
// indices lverts[vstart[p] .. vstart[p + 1]) (core first, then rings
// 1..rings); ring_end[p * (rings + 1) + r] = local count of rings 0..r; the
// local adjacency of local vertex j is ladj[astart[p] + aoff[vstart[p] + p + j]
// .. astart[p] + aoff[vstart[p] + p + j + 1]) (local indices, neighbour order
// of the mesh's CSR; empty for the outermost ring).
struct SmoothPatches {
    int count, rings;
    const int *vstart, *ring_end, *lverts, *aoff, *astart, *perm;
    const unsigned short *ladj;
};
// Iterations it0 .. it0 + rings - 1 of launch_smooth_values on mesh patches
// (lds_bytes: 26 L + 10 E of the largest patch, rounded up).
int launch_smooth_values_patched(const double *cur, const int *goff, const double *w, int n,
                                 size_t wstride, const HeatBatch &hb, const SmoothPatches &pp,
                                 int it0, size_t lds_bytes, double *nxt, hipStream_t s);
int launch_smooth_geometry_patched(const float *v, float4 *buf0, float4 *buf1, int n,
                                   const SmoothPatches &pp, const SmoothBatch &sb, int most,
                                   size_t lds_bytes, float *out, hipStream_t s);

template <typename T>
int launch_resample(const T *x, const int *idx, const double *w, int n, T *out, hipStream_t s);
int launch_edge_sum(const float *v, const int *faces, int nf, double *partial, int *nblocks,
                    hipStream_t s);
int launch_smooth_geometry(const float *cur, const int *off, const int *nbr, int n, float *nxt,
                           hipStream_t s);
int launch_smooth_geometry_batch(float *buf0, float *buf1, const int *off, const int *nbr,
                                 int n, const SmoothBatch &sb, int it, hipStream_t s);
// The batched smoothing on padded float4 copies (buf0, buf1: batch x n each):
// pack v, `most` steps, unpack each mesh's last iterate into out.
int launch_smooth_geometry4(const float *v, float4 *buf0, float4 *buf1, const int *off,
                            const int *nbr, int n, const SmoothBatch &sb, int most, float *out,
                            hipStream_t s);
int launch_vertex_normals(const float *v, const int *faces, const int *inc_off, const int *inc,
                          int n, float *out, hipStream_t s);
int launch_type5_base(const float *geom, const float *sm, const float *nr, int n, double *out,
                      hipStream_t s);
int launch_curvature_type2(const float *geom, const float *nr, int n, const int *off,
                           const int *nbr, double *out, hipStream_t s);
int launch_heat_weights(const float *geom, const int *off, const int *nbr, int n,
                        const HeatBatch &hb, int squared_norm, size_t wstride, double *w,
                        double *wsum, hipStream_t s);
int launch_smooth_values(const double *cur, const int *off, const int *nbr, const double *w,
                         const double *wsum, int n, size_t wstride, const HeatBatch &hb,
                         int iteration, double *nxt, hipStream_t s);
template <typename T>
int launch_minmax_normalise(const double *x, size_t n, int nb, size_t stride, double floor_,
                            double *partial, T *out, size_t ostride, hipStream_t s);
int launch_sheet_gather(const double *vals, size_t vstride, const int *idx, const double *w,
                        int m, int nb, double *out, hipStream_t s);
template <typename T>
int launch_apply_flow(const float *unit, int n, const T *flow, int nx, int ny, int inverse,
                      T *out, hipStream_t s);
template <typename T>
int launch_average_xz(const T *avg, const T *main_, int n, const double *m, float *out,
                      hipStream_t s);
int launch_normalise_f32(const float *in, int n, float *out, hipStream_t s);
struct Mat3 {
    double a[9];
};
// Up to kMaxSpheres vertex arrays ([n][3] float32, device) per centroid launch.
constexpr int kMaxSpheres = 64;
struct SphereList {
    int count;
    const float *ptr[kMaxSpheres];
    int n[kMaxSpheres];
};
// centroid[3 s + c] = (strictly sequential float64 sum over sphere s's vertices) / n.
int launch_sphere_centroids(const SphereList &list, double *centroid, hipStream_t s);
// centroid (device, may be null): p = f32(f64(v) - centroid), then float32 unit
// vectors; m_host (row-major, may be null): then p · mᵀ in float64, stored
// float32.  *bad is set to 1 for a vertex of zero or non-finite length (the
// caller clears it).
int launch_sphere_prepare(const float *v, int n, const double *centroid, const double *m_host,
                          const double *m_dev, float *out, int *bad, hipStream_t s);
int launch_rotation_matrices(const double *angles, int count, double *out, hipStream_t s);

}  // namespace csg
