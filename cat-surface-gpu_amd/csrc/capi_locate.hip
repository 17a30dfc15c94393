// C ABI of point location: grid index, rotation costs, closest triangles.
#include "../../include/csg.h"
#include "locate_kernels.h"

#include <mutex>

using namespace csg;

namespace {

inline hipStream_t S(void *s) { return (hipStream_t)s; }

// Clears a device counter with a kernel rather than hipMemsetAsync, so that a
// caller may capture the entry point into a graph.
__global__ void k_clear_int(int *p) { *p = 0; }

int clear_int(int *p, hipStream_t s)
{
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(1), 0, s, p);
    CSG_LAUNCH_CHECK();
    return kOk;
}

// Scratch of csg_rotation_costs_noalloc: the per-block partial sums and the
// uncovered-point counter, in a file-scope device array (one copy per GPU,
// held until the process ends).  In host code `g_rotation_scratch` names a
// host-side stand-in; its device address is looked up once per device with
// hipGetSymbolAddress and only that address is ever passed to kernels.
constexpr size_t kRotationScratchBytes = size_t(4) << 20;
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) unsigned char g_rotation_scratch[kRotationScratchBytes];
void *g_rotation_scratch_ptr[kMaxScratchDevices] = {};
std::mutex g_rotation_scratch_mutex;

int rotation_scratch(void **out)
{
    int dev = 0;
    CSG_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return fail(kRuntimeError, "device %d is beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_rotation_scratch_mutex);
    if (!g_rotation_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(g_rotation_scratch));
        if (e != hipSuccess || ptr == nullptr)
            return fail(kRuntimeError, "hipGetSymbolAddress(rotation scratch): %s",
                        hipGetErrorString(e));
        g_rotation_scratch_ptr[dev] = ptr;
    }
    *out = g_rotation_scratch_ptr[dev];
    return kOk;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct PlanLayout {
    size_t box, counts, offsets, scratch, total;
};

PlanLayout plan_layout(int nf)
{
    PlanLayout l;
    l.box = 0;
    l.counts = l.box + align256(sizeof(int) * 6 * (size_t)nf);
    l.offsets = l.counts + align256(sizeof(int) * ((size_t)nf + 1));
    l.scratch = l.offsets + align256(sizeof(int) * ((size_t)nf + 1));
    size_t scan = grid_sort_scratch_bytes(1);
    l.total = l.scratch + align256(scan);
    return l;
}

}  // namespace

extern "C" {

size_t csg_grid_index_plan_workspace_bytes(int nf) { return plan_layout(nf).total; }

int csg_grid_index_plan(const void *pts, int pts_dtype, const int *faces, int nf, int grid,
                        int margin, void *ws, size_t ws_bytes, int *npairs, void *stream)
{
    if (nf < 1 || grid < 1 || margin < 0 || npairs == nullptr)
        return fail(kInvalidArgument, "invalid grid index arguments");
    if ((long long)grid * grid * grid >= (1LL << 30))
        return fail(kInvalidArgument, "grid size too large");
    PlanLayout l = plan_layout(nf);
    if (ws_bytes < l.total)
        return fail(kInvalidArgument, "grid index plan workspace too small");
    char *b = (char *)ws;
    int *box = (int *)(b + l.box), *counts = (int *)(b + l.counts),
        *offsets = (int *)(b + l.offsets);
    if (pts_dtype == 1)
        return grid_index_count<double>((const double *)pts, faces, nf, grid, margin, box, counts,
                                        offsets, npairs, b + l.scratch, ws_bytes - l.scratch,
                                        S(stream));
    return grid_index_count<float>((const float *)pts, faces, nf, grid, margin, box, counts,
                                   offsets, npairs, b + l.scratch, ws_bytes - l.scratch,
                                   S(stream));
}

size_t csg_grid_index_fill_workspace_bytes(int npairs)
{
    return 3 * align256(sizeof(int) * (size_t)npairs) + align256(grid_sort_scratch_bytes(npairs));
}

int csg_grid_index_fill(const void *plan_ws, int nf, int grid, int npairs, int *cell_off,
                        int *cell_faces, void *ws, size_t ws_bytes, void *stream)
{
    if (npairs < 1)
        return fail(kRuntimeError, "At least one triangle does not cover any spatial grid cell");
    if (ws_bytes < csg_grid_index_fill_workspace_bytes(npairs))
        return fail(kInvalidArgument, "grid index fill workspace too small");
    PlanLayout l = plan_layout(nf);
    const char *pb = (const char *)plan_ws;
    char *b = (char *)ws;
    size_t a = align256(sizeof(int) * (size_t)npairs);
    int *keys_tmp = (int *)b, *vals_tmp = (int *)(b + a), *keys = (int *)(b + 2 * a);
    return grid_index_fill((const int *)(pb + l.box), (const int *)(pb + l.offsets), nf, grid,
                           npairs, keys_tmp, vals_tmp, keys, cell_faces, cell_off, b + 3 * a,
                           ws_bytes - 3 * a, S(stream));
}

size_t csg_grid_index_build_workspace_bytes(int nf, int capacity, int grid)
{
    if (nf < 1 || capacity < 1 || grid < 1 || (long long)grid * grid * grid >= (1LL << 30))
        return 0;
    return plan_layout(nf).scratch + align256(grid_build_scratch_bytes(nf, capacity));
}

int csg_grid_index_build(const void *pts, int pts_dtype, const int *faces, int nf, int grid,
                         int margin, int capacity, int *cell_off, int *cell_faces, int *required,
                         void *ws, size_t ws_bytes, void *stream)
{
    if (nf < 1 || grid < 1 || margin < 0 || capacity < 1 || !required || !cell_off ||
        !cell_faces)
        return fail(kInvalidArgument, "invalid grid index arguments");
    if ((long long)grid * grid * grid >= (1LL << 30))
        return fail(kInvalidArgument, "grid size too large");
    if (ws_bytes < csg_grid_index_build_workspace_bytes(nf, capacity, grid))
        return fail(kInvalidArgument, "grid index build workspace too small");
    PlanLayout l = plan_layout(nf);
    char *b = (char *)ws;
    int *box = (int *)(b + l.box), *counts = (int *)(b + l.counts),
        *offsets = (int *)(b + l.offsets);
    char *tmp = b + l.scratch;
    const size_t tmp_bytes = ws_bytes - l.scratch;
    if (pts_dtype == 1)
        return grid_index_build<double>((const double *)pts, faces, nf, grid, margin, capacity,
                                        box, counts, offsets, cell_faces, cell_off, required,
                                        tmp, tmp_bytes, S(stream));
    return grid_index_build<float>((const float *)pts, faces, nf, grid, margin, capacity, box,
                                   counts, offsets, cell_faces, cell_off, required, tmp,
                                   tmp_bytes, S(stream));
}

int csg_top_k(const double *costs, int n, int k, int *order, void *stream)
{
    if (n < 1 || k < 1 || k > kTopKMax || k > n || !costs || !order)
        return fail(kInvalidArgument, "top-k: 1 <= k <= min(n, %d)", kTopKMax);
    return launch_top_k(costs, n, k, order, S(stream));
}

size_t csg_rotation_costs_workspace_bytes(int ns, int na)
{
    return align256(sizeof(double) * (size_t)rotation_cost_blocks(ns) * na) + 256;
}

int csg_face_spheres(const float *tri, int nf, float *sph, void *stream)
{
    if (nf < 1 || !tri || !sph)
        return fail(kInvalidArgument, "invalid face sphere arguments");
    return launch_face_spheres_f32(tri, nf, reinterpret_cast<float4 *>(sph), S(stream));
}

int csg_rotation_costs(const float *src_pts, const double *src_vals, int ns,
                       const double *tgt_vals, const int *faces, const float *tri,
                       const float *sph, const int *cell_off, const int *cell_faces,
                       int grid, const double *angles, int na, double *costs, int *loc_face,
                       float *loc_bary, void *ws, size_t ws_bytes, int *uncovered_host,
                       int *uncovered_dev, void *stream)
{
    if (ns < 1 || na < 1 || grid < 1)
        return fail(kInvalidArgument, "invalid rotation cost arguments");
    if (ws_bytes < csg_rotation_costs_workspace_bytes(ns, na))
        return fail(kInvalidArgument, "rotation cost workspace too small");
    hipStream_t s = S(stream);
    char *b = (char *)ws;
    RotationCostArgs a{};
    a.src_pts = src_pts;
    a.src_vals = src_vals;
    a.ns = ns;
    a.tgt_vals = tgt_vals;
    a.faces = faces;
    a.tri = tri;
    a.sph = reinterpret_cast<const float4 *>(sph);
    a.cell_off = cell_off;
    a.cell_faces = cell_faces;
    a.grid = grid;
    a.angles = angles;
    a.na = na;
    a.costs = costs;
    a.partial = (double *)b;
    a.uncovered = uncovered_dev
                      ? uncovered_dev
                      : (int *)(b + align256(sizeof(double) * (size_t)rotation_cost_blocks(ns) * na));
    a.loc_face = loc_face;
    a.loc_bary = loc_bary;
    CSG_TRY(clear_int(a.uncovered, s));
    int rc = launch_rotation_costs(a, s);
    if (rc)
        return rc;
    if (uncovered_host) {
        CSG_HIP_CHECK(hipMemcpyAsync(uncovered_host, a.uncovered, sizeof(int),
                                     hipMemcpyDeviceToHost, s));
        CSG_HIP_CHECK(hipStreamSynchronize(s));
    }
    return kOk;
}

size_t csg_rotation_costs_noalloc_max_bytes(void) { return kRotationScratchBytes; }

int csg_rotation_costs_noalloc(const float *src_pts, const double *src_vals, int ns,
                               const double *tgt_vals, const int *faces, const float *tri,
                               const float *sph, const int *cell_off, const int *cell_faces,
                               int grid, const double *angles, int na, double *costs,
                               int *loc_face, float *loc_bary, int *uncovered_dev, void *stream)
{
    if (ns < 1 || na < 1 || grid < 1 || uncovered_dev == nullptr)
        return fail(kInvalidArgument, "invalid rotation cost arguments");
    const size_t need = csg_rotation_costs_workspace_bytes(ns, na);
    if (need > kRotationScratchBytes)
        return fail(kInvalidArgument,
                    "rotation costs for %d points x %d angles need %zu bytes of scratch; the "
                    "no-allocation entry point holds %zu (use csg_rotation_costs with a "
                    "workspace)",
                    ns, na, need, kRotationScratchBytes);
    void *scratch = nullptr;
    CSG_TRY(rotation_scratch(&scratch));
    return csg_rotation_costs(src_pts, src_vals, ns, tgt_vals, faces, tri, sph, cell_off,
                              cell_faces, grid, angles, na, costs, loc_face, loc_bary, scratch,
                              kRotationScratchBytes, nullptr, uncovered_dev, stream);
}

size_t csg_rotation_cost_matrices_workspace_bytes(int ns, int na, int with_gradient)
{
    const size_t rows = (size_t)na * (with_gradient ? kRotationGradTerms : 1);
    return align256(sizeof(double) * (size_t)rotation_cost_blocks(ns) * rows) + 256;
}

int csg_rotation_cost_matrices(const float *src_pts, const double *src_vals, int ns,
                               const double *tgt_vals, const int *faces, const float *tri,
                               const float *sph, const int *cell_off, const int *cell_faces,
                               int grid, const float *mats, int na, int with_gradient, double *out,
                               void *ws, size_t ws_bytes, int *uncovered_dev, void *stream)
{
    if (ns < 1 || na < 1 || grid < 1 || mats == nullptr || out == nullptr ||
        uncovered_dev == nullptr)
        return fail(kInvalidArgument, "invalid rotation cost arguments");
    if (na > 65535)
        return fail(kInvalidArgument, "at most 65535 candidate rotations per call, got %d", na);
    if (ws_bytes < csg_rotation_cost_matrices_workspace_bytes(ns, na, with_gradient))
        return fail(kInvalidArgument, "rotation cost workspace too small");
    hipStream_t s = S(stream);
    RotationCostArgs a{};
    a.src_pts = src_pts;
    a.src_vals = src_vals;
    a.ns = ns;
    a.tgt_vals = tgt_vals;
    a.faces = faces;
    a.tri = tri;
    a.sph = reinterpret_cast<const float4 *>(sph);
    a.cell_off = cell_off;
    a.cell_faces = cell_faces;
    a.grid = grid;
    a.na = na;
    a.costs = out;
    a.partial = (double *)ws;
    a.uncovered = uncovered_dev;
    return launch_rotation_cost_mats(a, mats, with_gradient != 0, s);
}

size_t csg_rotation_map_costs_workspace_bytes(int nx, int ny, int stride, int na)
{
    if (nx < 1 || ny < 1 || stride < 1 || na < 1)
        return 0;
    return align256(sizeof(double) * (size_t)map_cost_blocks(nx, ny, stride) * na) + 256;
}

int csg_rotation_map_costs(const double *src_map, const double *tgt_map, int nx, int ny,
                           int stride, const float *mats, int na, double *costs, void *ws,
                           size_t ws_bytes, void *stream)
{
    if (nx < 2 || ny < 2 || stride < 1 || na < 1 || na > 65535 || src_map == nullptr ||
        tgt_map == nullptr || mats == nullptr || costs == nullptr)
        return fail(kInvalidArgument, "invalid map rotation cost arguments");
    if (ws_bytes < csg_rotation_map_costs_workspace_bytes(nx, ny, stride, na))
        return fail(kInvalidArgument, "map rotation cost workspace too small");
    return launch_rotation_map_costs(src_map, tgt_map, nx, ny, stride, mats, na, (double *)ws,
                                     costs, S(stream));
}

int csg_closest_triangles(const double *q, int nq, const double *verts, const int *faces,
                          const int *cell_off, const int *cell_faces, int grid, double reach,
                          int *face_out, int *idx_out, double *w_out, int *far_host,
                          void *ws, size_t ws_bytes, void *stream)
{
    if (nq < 0 || grid < 1 || ws_bytes < sizeof(int))
        return fail(kInvalidArgument, "invalid closest-triangle arguments");
    hipStream_t s = S(stream);
    int *far_count = (int *)ws;
    CSG_TRY(clear_int(far_count, s));
    int rc = launch_closest_triangles(q, nq, verts, faces, cell_off, cell_faces, grid,
                                      reach * reach, face_out, idx_out, w_out, far_count, s);
    if (rc)
        return rc;
    if (far_host) {
        CSG_HIP_CHECK(hipMemcpyAsync(far_host, far_count, sizeof(int), hipMemcpyDeviceToHost, s));
        CSG_HIP_CHECK(hipStreamSynchronize(s));
    }
    return kOk;
}

size_t csg_closest_triangles_workspace_bytes(int nf)
{
    // far count, face spheres, face-major corner copies
    return nf < 1 ? 0 : 256 + align256(sizeof(double4) * (size_t)nf) +
                            align256(9 * sizeof(double) * (size_t)nf);
}

int csg_closest_triangles_pruned(const double *q, int nq, const double *verts, const int *faces,
                                 int nf, const int *cell_off, const int *cell_faces, int grid,
                                 double reach, int *face_out, int *idx_out, double *w_out,
                                 int *far_dev, int *far_host, void *ws, size_t ws_bytes,
                                 void *stream)
{
    if (nq < 0 || nf < 1 || grid < 1 || !far_dev)
        return fail(kInvalidArgument, "invalid closest-triangle arguments");
    if (ws_bytes < csg_closest_triangles_workspace_bytes(nf))
        return fail(kInvalidArgument, "closest-triangle workspace too small");
    hipStream_t s = S(stream);
    CSG_TRY(clear_int(far_dev, s));
    double4 *sph = (double4 *)((char *)ws + 256);
    double *tri = (double *)((char *)sph + align256(sizeof(double4) * (size_t)nf));
    CSG_TRY(launch_closest_triangles_pruned(q, nq, verts, faces, nf, cell_off, cell_faces, grid,
                                            reach * reach, sph, tri, face_out, idx_out, w_out,
                                            far_dev, s));
    if (far_host) {
        CSG_HIP_CHECK(hipMemcpyAsync(far_host, far_dev, sizeof(int), hipMemcpyDeviceToHost, s));
        CSG_HIP_CHECK(hipStreamSynchronize(s));
    }
    return kOk;
}

int csg_closest_triangles_block(const double *q, int nq, const double *verts, const int *faces,
                                int nf, const int *cell_off, const int *cell_faces, int grid,
                                int span, int *face_out, int *idx_out, double *w_out,
                                int *far_dev, int *far_host, void *ws, size_t ws_bytes,
                                void *stream)
{
    if (nq < 0 || nf < 1 || grid < 1 || span < 0 || !far_dev)
        return fail(kInvalidArgument, "invalid closest-triangle arguments");
    if (ws_bytes < csg_closest_triangles_workspace_bytes(nf))
        return fail(kInvalidArgument, "closest-triangle workspace too small");
    hipStream_t s = S(stream);
    CSG_TRY(clear_int(far_dev, s));
    double4 *sph = (double4 *)((char *)ws + 256);
    double *tri = (double *)((char *)sph + align256(sizeof(double4) * (size_t)nf));
    CSG_TRY(launch_closest_triangles_block(q, nq, verts, faces, nf, cell_off, cell_faces, grid,
                                           span, sph, tri, face_out, idx_out, w_out, far_dev,
                                           s));
    if (far_host) {
        CSG_HIP_CHECK(hipMemcpyAsync(far_host, far_dev, sizeof(int), hipMemcpyDeviceToHost, s));
        CSG_HIP_CHECK(hipStreamSynchronize(s));
    }
    return kOk;
}

}  // extern "C"
