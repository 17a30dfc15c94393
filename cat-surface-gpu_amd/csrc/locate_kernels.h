// Launch wrappers for the point-location kernels (internal to libcsg_hip).
#pragma once

#include "csg_common.h"

namespace csg {

template <typename P>
int grid_index_count(const P *pts, const int *faces, int nf, int g, int margin, int *box,
                     int *counts, int *offsets, int *total, void *tmp, size_t tmp_bytes,
                     hipStream_t s);
size_t grid_sort_scratch_bytes(int npairs);
// The whole index without a host round trip: at most `capacity` pairs are
// listed (cell_faces holds capacity ints) and *required (device) receives the
// count the index needs; the index is complete iff *required <= capacity.
// Built with this library's own scan and stable radix sort (no library call:
// nothing that records a memset into a graph).
template <typename P>
int grid_index_build(const P *pts, const int *faces, int nf, int g, int margin, int capacity,
                     int *box, int *counts, int *offsets, int *cell_faces, int *cell_off,
                     int *required, void *tmp, size_t tmp_bytes, hipStream_t s);
size_t grid_build_scratch_bytes(int nf, int capacity);
int grid_index_fill(const int *box, const int *offsets, int nf, int g, int npairs, int *keys_tmp,
                    int *vals_tmp, int *keys, int *vals, int *cell_off, void *tmp,
                    size_t tmp_bytes, hipStream_t s);

// The k (<= kTopKMax) first of n costs in stable ascending order (NaN last).
constexpr int kTopKMax = 64;
int launch_top_k(const double *costs, int n, int k, int *order, hipStream_t s);

// Lanes that share one query point in the location kernels, each scanning
// every kLocateLanes-th candidate of the point's cell.
constexpr int kLocateLanes = 4;

// Blocks of the rotation-cost kernel per angle.
inline int rotation_cost_blocks(int ns) { return (int)grid_for((size_t)ns * kLocateLanes); }

struct RotationCostArgs {
    const float *src_pts;
    const double *src_vals;
    int ns;
    const double *tgt_vals;
    const int *faces;
    const float *tri;
    const float4 *sph;  // face bounding spheres (may be null: every candidate is scanned)
    const int *cell_off, *cell_faces;
    int grid;
    const double *angles;
    int na;
    double *costs;
    double *partial;  // [na][rotation_cost_blocks(ns)]
    int *uncovered;
    int *loc_face;    // [na][ns] or null
    float *loc_bary;  // [na][ns][3] or null
    // [na][ns] or null (matrix costs only): each point's cell and the position of
    // its face in that cell's list for each candidate from the previous launch
    // (cell << kHintPosBits | position; -1: none), read to start the scan, then
    // updated
    int *hint;
    // [nf][3] or null (matrix costs only): tgt_vals at each face's corners, read
    // instead of faces then tgt_vals
    const double *face_vals;
};
constexpr int kHintPosBits = 10;

int launch_rotation_costs(const RotationCostArgs &a, hipStream_t s);
// Bounding spheres of the faces of tri for the locator's candidate skip.
int launch_face_spheres_f32(const float *tri, int nf, float4 *sph, hipStream_t s);

// Float64 sums per candidate of the matrix-input cost with gradient: cost,
// gradient (3) and Gauss-Newton matrix (xx, yy, zz, xy, xz, yz).
constexpr int kRotationGradTerms = 10;
// a.angles unused; a.costs receives [na] or [na][kRotationGradTerms];
// a.partial needs rows x rotation_cost_blocks(ns) doubles.
// Area-weighted map SSD of candidate rotations (see k_rotation_map_cost).
int map_cost_blocks(int nx, int ny, int stride);
int launch_rotation_map_costs(const double *smap, const double *tmap, int nx, int ny, int stride,
                              const float *mats, int na, double *partial, double *costs,
                              hipStream_t s);
int launch_rotation_cost_mats(const RotationCostArgs &a, const float *mats, bool gradient,
                              hipStream_t s);

int launch_closest_triangles(const double *q, int nq, const double *verts, const int *faces,
                             const int *cell_off, const int *cell_faces, int g, double reach2,
                             int *face_out, int *idx_out, double *w_out, int *far_count,
                             hipStream_t s);
// The same, skipping candidates by their faces' bounding spheres (scratch:
// sph, nf double4; tri, 9·nf doubles for the faces' corners); identical results.
int launch_closest_triangles_pruned(const double *q, int nq, const double *verts,
                                    const int *faces, int nf, const int *cell_off,
                                    const int *cell_faces, int g, double reach2, double4 *sph,
                                    double *tri, int *face_out, int *idx_out, double *w_out,
                                    int *far_count, hipStream_t s);
// The block search on an index built without margin (k_closest_triangles_block):
// the query's cell, then the (2·span + 1)^3 block around it where needed; the
// same scratch as launch_closest_triangles_pruned.
int launch_closest_triangles_block(const double *q, int nq, const double *verts,
                                   const int *faces, int nf, const int *cell_off,
                                   const int *cell_faces, int g, int span, double4 *sph,
                                   double *tri, int *face_out, int *idx_out, double *w_out,
                                   int *far_count, hipStream_t s);

}  // namespace csg
