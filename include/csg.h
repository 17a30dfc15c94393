/*
 * libcsg_hip — C ABI of the MI355X (gfx950) CAT_SurfWarp registration hot path.
 *
 * Conventions for every entry point:
 *   - all array pointers are device memory unless a comment says "host";
 *   - grids are channel-major [C][ny][nx], contiguous; maps are one-based;
 *   - dtype selects the element type of the floating arrays: 0 = float32,
 *     1 = float64 (reductions are always accumulated in float64);
 *   - `stream` is a hipStream_t (NULL = the default stream); work is
 *     stream-ordered and nothing allocates, so callers may capture the calls
 *     into a hipGraph (except where a comment says the call synchronises);
 *   - scratch memory comes from the caller as (ws, ws_bytes); the matching
 *     *_workspace_bytes() call gives the size needed;
 *   - the return value is 0 on success, -1 for an invalid argument, -2 for a
 *     runtime (HIP) failure, -3 for an unsupported option; csg_last_error()
 *     returns the message of the last failure on the calling thread.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the root of ayakacxy/cat-surface-gpu).  The reference binds these seams
 * in Python (kernel="triton" dispatch); INTEGRATION.md shows the ctypes
 * binding a maintainer adds for kernel="hip".
 */
#ifndef CSG_H
#define CSG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- library */

const char *csg_last_error(void);
const char *csg_version(void);
/* 0 if a HIP device is visible; fills *count. */
int csg_device_count(int *count);
/* Multigrid schedule of every FMG below (fmg2_membrane, dartel_step, solve): 1 = fused
 * (fine levels relaxed on LDS tiles, coarse levels in one workgroup per job; the
 * default), 0 = one launch per reference operator call.  Both compute the same values
 * in the same order.  Process-wide; returns the previous setting.  No counterpart in
 * the reference (it has only the per-operator schedule of dartel_grid.py:1211-1417). */
int csg_set_fmg_schedule(int fused);
/* Tile interior (tile_x even, tile_y) of the fused schedule's fine-level relaxation for
 * levels of nx x ny in dtype; 0 x 0 restores the default 32 x 16.  Every shape computes
 * the same values; the shape only changes the speed (set by an autotuner).  Process-wide;
 * launches made (or graphs captured) afterwards use it.  No counterpart in the
 * reference. */
int csg_set_relax_tile(int nx, int ny, int dtype, int tile_x, int tile_y);
int csg_get_relax_tile(int nx, int ny, int dtype, int *tile_x, int *tile_y);
/* Number of levels of an nx x ny FMG (nit relaxation iterations) that the fused schedule
 * relaxes on tiles, i.e. the levels whose tile shape matters; 0 if the fused schedule
 * does not apply; negative on invalid arguments. */
int csg_fmg_tiled_levels(int nx, int ny, int nit, int dtype);
/* Optional HIP-event timing of the hot kernels for roofline reports (no counterpart in the
 * reference).  csg_kernel_timing(1) brackets every later launch of a timed kernel made
 * outside stream capture with two events on its stream and accumulates the launch's
 * algorithmic bytes (what the reference's operator sequence reads and writes once for it);
 * returns the previous setting.  csg_kernel_timing_read synchronises on the recorded events
 * and returns (then clears) the totals of kernel `kernel` (0 .. while csg_kernel_timing_name
 * returns non-NULL). */
int csg_kernel_timing(int enable);
int csg_kernel_timing_read(int kernel, double *ms, double *bytes, long long *launches);
const char *csg_kernel_timing_name(int kernel);

/* ------------------------------------------------------------ DARTEL grid */

/* Bilinear sample of a scalar [ny][nx] field at zero-based (x, y).
 * Replaces dartel_grid.sample_field (src/cat_surface_gpu/dartel_grid.py:510-531). */
int csg_sample_field(const void *field, int nx, int ny, const void *x, const void *y, size_t n,
                     int dtype, void *out, void *stream);

/* c = b o a.  Replaces dartel_grid.composition (dartel_grid.py:552-575). */
int csg_composition(const void *a, const void *b, int nx, int ny, int dtype, void *out,
                    void *stream);

/* (b o a, (Jb o a) Ja).  Replaces dartel_grid.composition_jacobian (dartel_grid.py:578-621). */
int csg_composition_jacobian(const void *a, const void *ja, const void *b, const void *jb, int nx,
                             int ny, int dtype, void *map_out, void *jac_out, void *stream);

/* I + scale * grad v (4 channels).  Replaces dartel_grid.jacobian_of_displacement
 * (dartel_grid.py:624-652). */
int csg_jacobian_of_displacement(const void *v, double scale, int nx, int ny, int dtype,
                                 void *jac_out, void *stream);

/* (id + scale v, I + scale grad v).  Replaces dartel_grid.smalldef_jac (dartel_grid.py:655-676). */
int csg_smalldef_jac(const void *v, double scale, int nx, int ny, int dtype, void *map_out,
                     void *jac_out, void *stream);

/* J (I + scale grad v)^-1.  Replaces dartel_grid.jac_div_smalldef (dartel_grid.py:679-720). */
int csg_jac_div_smalldef(const void *jac, const void *v, double scale, int nx, int ny, int dtype,
                         void *out, void *stream);

/* Scaling and squaring exp(v) with k squarings (+ Jacobian when jac_out != NULL).
 * Replaces dartel_grid.expdef (dartel_grid.py:1619-1666). */
size_t csg_expdef_workspace_bytes(int nx, int ny, int dtype);
int csg_expdef(const void *v, int nx, int ny, int dtype, int k, void *map_out, void *jac_out,
               void *ws, size_t ws_bytes, void *stream);

/* Objective, gradient (rhs, 2ch) and Gauss-Newton Hessian (3ch); *objective (device
 * double) receives 0.5 * sum(d^2 w).  dist may be NULL.
 * Replaces dartel_grid.initialise_objfun (dartel_grid.py:723-775). */
size_t csg_objfun_workspace_bytes(int nx, int ny);
int csg_initialise_objfun(const void *src, const void *tgt, const void *map, const void *jac,
                          const void *dist, int nx, int ny, int dtype, void *rhs, void *hes,
                          double *objective, void *ws, size_t ws_bytes, void *stream);

/* k fused squaring updates; scratch_out receives the reference's map_buffers[0]
 * (the FMG initial guess).  Replaces dartel_grid.squaring_update (dartel_grid.py:778-914)
 * and dartel_triton.squaring_update_triton (src/cat_surface_gpu/dartel_triton.py:644-693). */
size_t csg_squaring_workspace_bytes(int nx, int ny, int dtype);
int csg_squaring_update(const void *rhs, const void *hes, const void *map, const void *jac, int k,
                        int save_transformation, int nx, int ny, int dtype, void *b_out,
                        void *a_out, void *map_out, void *jac_out, void *scratch_out, void *ws,
                        size_t ws_bytes, void *stream);

/* L v for rtype 0/1/2; params (host) = (sx, sy, mu, lambda, id).
 * Replaces dartel_grid.regularization_operator (dartel_grid.py:917-1015). */
int csg_regularization_operator(const void *v, const double *params, int rtype, int nx, int ny,
                                int dtype, void *out, void *stream);

/* (L_membrane + A) u.  Replaces dartel_triton.apply_membrane_system_triton
 * (dartel_triton.py:528-569). */
int csg_apply_membrane_system(const void *hes, const void *u, double w00, double w01, double w10,
                              int nx, int ny, int dtype, void *out, void *stream);

/* nit red-black iterations in place on `solution` (nx even or 1).
 * Replaces dartel_triton.relax_membrane_triton (dartel_triton.py:572-641). */
int csg_relax_membrane(const void *hes, const void *rhs, void *solution, double w00, double w01,
                       double w10, int nit, int nx, int ny, int dtype, void *stream);

/* Quadratic B-spline resize of a [channels][sny][snx] field.
 * Replaces dartel_triton.resize_field_triton (dartel_triton.py:480-525). */
int csg_resize_field(const void *in, int channels, int snx, int sny, int tnx, int tny, int dtype,
                     void *out, void *stream);

/* Full multigrid membrane solve; params (host, 5); initial may be NULL;
 * coarse_init != 0 selects the coarse-initialisation branch.
 * Replaces dartel_grid.fmg2_membrane (dartel_grid.py:1211-1417). */
size_t csg_fmg2_workspace_bytes(int nx, int ny, int batch, int dtype);
int csg_fmg2_membrane(const void *hes, const void *rhs, const double *params, int cycles, int nit,
                      const void *initial, int coarse_init, int nx, int ny, int batch, int dtype,
                      void *out, void *ws, size_t ws_bytes, void *stream);

/* One Gauss-Newton DARTEL step for `batch` independent (src, tgt, v) triples;
 * metrics (device double [batch][3]) = (objective, reg/2, ||rhs||).  dist may be NULL.
 * With assume_nonzero == 0 the call synchronises the stream once (the
 * reference's torch.any(initial) check).
 * Replaces dartel_grid.dartel_step (dartel_grid.py:1420-1616). */
size_t csg_dartel_step_workspace_bytes(int nx, int ny, int batch, int dtype);
int csg_dartel_step(const void *src, const void *tgt, const void *v, const void *dist, int k,
                    const double *params, double lmreg, int cycles, int nit, int code,
                    int assume_nonzero, int nx, int ny, int batch, int dtype, void *v_out,
                    double *metrics, void *ws, size_t ws_bytes, void *stream);

/* Whole DARTEL solve: `batch` maps grouped in jobs of steps_per_job consecutive
 * steps; every step starts from v = 0 and runs loop x its dartel steps; flow
 * (jobs x 2ch) = expdef(v of each job's last step, 10).  velocity / metrics may
 * be NULL.  Replaces surface_warp.solve_dartel_maps
 * (src/cat_surface_gpu/surface_warp.py:239-367). */
size_t csg_solve_workspace_bytes(int nx, int ny, int batch, int dtype);
int csg_solve_dartel_maps(const void *src, const void *tgt, int nx, int ny, int batch,
                          int steps_per_job, int loop, double lmreg, int cycles, int nit, int its,
                          int code, int assume_nonzero, int dtype, void *flow, void *velocity,
                          double *metrics, void *ws, size_t ws_bytes, void *stream);

/* Exp-warp-SSD objective and its hand-written adjoint (no counterpart in the reference,
 * whose DARTEL step uses the push-forward gradient of dartel_grid.py:1509-1592):
 *   E(v) = 1/2 sum (f(phi) - t)^2 + 1/2 sum v . (L v),  phi = exp(v),
 * with K squarings of phi0 = id + v 2^-K (expdef, dartel_grid.py:1619-1666), the bilinear
 * sample of initialise_objfun (:723-775) and L the rtype-1 regulariser with parameters
 * (1, 1, membrane, 0, 0) (:917-955).  v is [2][ny][nx], feature and target [ny][nx], all
 * of dtype (0 float32, 1 float64).  forward writes E to energy[0] (device double) and,
 * if parts != NULL, (SSD, membrane) to parts[0..1].  backward writes dE/dv (times
 * *grad_scale when given) to grad_v.  store = 1 keeps the K + 1 maps of the forward pass
 * in the workspace for backward (pass the same workspace, forward first); store = 0
 * rebuilds them by recomputation (two maps of workspace).  The adjoint accumulates without
 * atomics in a fixed order (stable sort of the scattered terms by destination): the same
 * bits on every run and in both schedules.  At most 2^28 pixels. */
size_t csg_exp_warp_ssd_workspace_bytes(int nx, int ny, int K, int dtype, int store);
int csg_exp_warp_ssd_forward(const void *v, const void *feature, const void *target, int nx,
                             int ny, int K, double membrane, int dtype, int store,
                             double *energy, double *parts, void *ws, size_t ws_bytes,
                             void *stream);
int csg_exp_warp_ssd_backward(const void *v, const void *feature, const void *target, int nx,
                              int ny, int K, double membrane, int dtype, int store,
                              const double *grad_scale, void *grad_v, void *ws,
                              size_t ws_bytes, void *stream);

/* ---------------------------------------------------------------- surface */

/* Stencil resampling of [n_src][3] vertices to [n_out][3] (dtype 0: float32 with the
 * reference's per-corner float32 accumulation; 1: float64).
 * Replaces SurfaceStencilDevice.resample_vertices (src/cat_surface_gpu/surface_stencil.py:282-306). */
int csg_resample_vertices(const void *x, int dtype, const int *idx, const double *w, int n_out,
                          void *out, void *stream);

/* Per-block partial sums of face edge lengths (partial_out: ceil(nf/256) doubles);
 * the mean edge length is their sum / (3 nf).
 * Replaces SurfaceStencilDevice._edge_mean_distance (surface_stencil.py:349-357). */
size_t csg_edge_mean_workspace_bytes(int nf);
int csg_edge_sum(const float *v, const int *faces, int nf, double *partial_out, int *nblocks,
                 void *stream);

/* `iterations` heat iterations of vertex geometry over the ascending CSR neighbour
 * table.  ws: 3 n floats.  Replaces SurfaceStencilDevice.smooth_geometry
 * (surface_stencil.py:372-388). */
int csg_smooth_geometry(const float *v, const int *nbr_off, const int *nbr, int n,
                        int iterations, float *out, void *ws, size_t ws_bytes, void *stream);
/* `batch` meshes [batch][n][3] of one topology, mesh b smoothed iterations[b] (host array)
 * times, all in the same launches (batch <= 32).  ws needs at least 3 n batch floats;
 * with csg_smooth_geometry_batch_workspace_bytes it runs on padded float4 copies (one
 * 16-byte load per neighbour), same values.  Replaces repeated
 * SurfaceStencilDevice.smooth_geometry calls (surface_stencil.py:372-388). */
size_t csg_smooth_geometry_batch_workspace_bytes(int n, int batch);
int csg_smooth_geometry_batch(const float *v, int batch, const int *nbr_off, const int *nbr, int n,
                              const int *iterations, float *out, void *ws, size_t ws_bytes,
                              void *stream);

/* csg_smooth_geometry_batch in launches of `rings` iterations each: one workgroup per
 * (patch, mesh) iterates in LDS on a patch of core vertices and the `rings` rings of
 * their graph neighbourhood (a partition of the mesh computed once per topology, see
 * surface_stencil.smoothing_patches), writing the core.  Identical values to
 * csg_smooth_geometry_batch.  Patch arrays (device): perm [n] (the working order: working
 * index i holds vertex perm[i], the patches' cores one after another), vstart
 * [npatches + 1], ring_end [npatches (rings + 1)] (local counts of rings 0..r), lverts
 * (working indices, core first),
 * aoff [vstart[npatches] + npatches] (per patch, local CSR offsets relative to astart[p]),
 * astart [npatches + 1], ladj (uint16 local neighbour indices in the mesh's CSR order);
 * lds_bytes = the largest patch's 24 L + 2 (L + 1) + 2 E bytes (E < 65536).  Workspace as
 * csg_smooth_geometry_batch_workspace_bytes. */
int csg_smooth_geometry_patched(const float *v, int batch, int n, const int *iterations,
                                int npatches, int rings, const int *perm, const int *vstart,
                                const int *ring_end, const int *lverts, const int *aoff,
                                const int *astart, const unsigned short *ladj, size_t lds_bytes,
                                float *out, void *ws, size_t ws_bytes, void *stream);

/* Angle-weighted vertex normals; inc lists face*3+corner per vertex in (corner, face)
 * order.  Replaces SurfaceStencilDevice.vertex_normals (surface_stencil.py:478-505). */
int csg_vertex_normals(const float *v, const int *faces, const int *inc_off, const int *inc,
                       int n, float *out, void *stream);

/* nb (<= 32) value-smoothing problems on meshes of one topology; sigma, iterations and
 * mesh_of (problem b smooths on geometry geom + mesh_of[b] * 3 n, each < meshes; NULL
 * means every problem uses mesh 0) are host arrays.  squared_norm selects
 * smooth_values_many's distance form.  Replaces
 * SurfaceStencilDevice.smooth_values / smooth_values_many (surface_stencil.py:390-476). */
size_t csg_smooth_values_workspace_bytes(int n, int nnz, int nb);
int csg_smooth_values(const double *values, int n, int nb, const float *geom, int meshes,
                      const int *mesh_of, const int *nbr_off, const int *nbr, int nnz,
                      const double *sigma, const int *iterations, int squared_norm, double *out,
                      void *ws, size_t ws_bytes, void *stream);

/* csg_smooth_values with up to `rings` iterations per launch on mesh patches (the plan of
 * csg_smooth_geometry_patched, built by surface_stencil.smoothing_patches; lds_bytes =
 * the plan's 26 L + 10 E bound).  Values equal csg_smooth_values'.  values and out must
 * not overlap.  No counterpart in the reference (it iterates surface_stencil.py:390-476). */
int csg_smooth_values_patched(const double *values, int n, int nb, const float *geom, int meshes,
                              const int *mesh_of, const int *nbr_off, const int *nbr, int nnz,
                              const double *sigma, const int *iterations, int squared_norm,
                              int npatches, int rings, const int *perm, const int *vstart,
                              const int *ring_end, const int *lverts, const int *aoff,
                              const int *astart, const unsigned short *ladj, size_t lds_bytes,
                              double *out, void *ws, size_t ws_bytes, void *stream);

/* (x - min) / max(max - min, floor) per row of nb rows of n doubles. */
size_t csg_minmax_workspace_bytes(int nb);
int csg_minmax_normalise(const double *x, size_t n, int nb, double floor_, int out_dtype,
                         void *out, void *ws, size_t ws_bytes, void *stream);

/* sheet[b][p] = sum_c values[b][idx[p,c]] w[p,c].  Replaces
 * SurfaceStencilDevice.map_values_to_sheet (surface_stencil.py:308-317). */
int csg_sheet_gather(const double *values, int n_values, int nb, const int *idx, const double *w,
                     int m, double *out, void *stream);

/* Whole curvature_type5_to_sheet_many for nb fwhms (sigma / iterations host arrays).
 * `smoothed` may give the fwhm-50 smoothed geometry already computed (e.g. by
 * csg_smooth_geometry_batch); NULL smooths `geom` geom_iterations times here.
 * Replaces SurfaceStencilDevice.curvature_type5_to_sheet_many (surface_stencil.py:521-550). */
size_t csg_type5_workspace_bytes(int n, int nnz, int nb, int m);
int csg_type5_sheets(const float *geom, int n, const int *faces, const int *nbr_off,
                     const int *nbr, int nnz, const int *inc_off, const int *inc,
                     int geom_iterations, const float *smoothed, int nb, const double *sigma,
                     const int *iterations, const int *sheet_idx, const double *sheet_w, int m,
                     int out_dtype, void *out, void *ws, size_t ws_bytes, void *stream);

/* Type-2 curvature values before smoothing (out: [n] float64) of a mesh with float32
 * vertex normals and ascending neighbour CSR.  Replaces the shape-operator part of
 * SurfaceStencilDevice.curvature_type2 (surface_stencil.py:552-613). */
int csg_curvature_type2(const float *geom, const float *normals, int n, const int *nbr_off,
                        const int *nbr, double *out, void *stream);

/* curvature_to_sheet for one fwhm (sigma / iterations of its heat kernel): curvtype 5 smooths
 * the geometry geom_iterations times (the fwhm-50 smoothing) and takes the depth along the
 * smoothed normals, curvtype 2 the type-2 curvature; the values are smoothed with |d|^2 heat
 * weights, normalised, gathered to the sheet (m points) and normalised again.  Replaces
 * SurfaceStencilDevice.curvature_to_sheet (surface_stencil.py:624-644). */
size_t csg_curvature_sheet_workspace_bytes(int n, int nnz, int m);
int csg_curvature_sheet(const float *geom, int n, const int *faces, const int *nbr_off,
                        const int *nbr, int nnz, const int *inc_off, const int *inc,
                        int curvtype, int geom_iterations, double sigma, int iterations,
                        const int *sheet_idx, const double *sheet_w, int m, int out_dtype,
                        void *out, void *ws, size_t ws_bytes, void *stream);

/* csg_type5_sheets for `sides` meshes of one topology in shared launches, their fwhm-50
 * smoothed copies given: geom and smoothed [sides][n][3]; nfw rows per side (sigma /
 * iterations host arrays [sides][nfw], sides * nfw <= 32); sheet_idx / sheet_w host
 * arrays of each side's device stencil (m points); out [sides][nfw][m].  One
 * value-smoothing loop serves every row; each row equals the single-mesh call's.  Replaces
 * the per-side curvature_type5_to_sheet_many calls of build_maps
 * (surface_warp.py:427-448, surface_stencil.py:521-550). */
size_t csg_type5_sheets_batch_workspace_bytes(int n, int nnz, int sides, int nfw, int m);
int csg_type5_sheets_batch(const float *geom, const float *smoothed, int n, int sides,
                           const int *faces, const int *nbr_off, const int *nbr, int nnz,
                           const int *inc_off, const int *inc, int nfw, const double *sigma,
                           const int *iterations, const int *const *sheet_idx,
                           const double *const *sheet_w, int m, int out_dtype, void *out,
                           void *ws, size_t ws_bytes, void *stream);

/* ---- Device-resident heat parameters.  csg_heat_parameters computes a mesh's mean face-edge
 * length (the csg_edge_sum partials added in order, over 3 nf) and, for each of nfw <= 8
 * fwhms, iterations = max(1, ceil(fwhm^2 0.541011 / em^2)) and sigma = fwhm 0.7355345 /
 * iterations^0.5 into device arrays (em_out optional), without a host round trip.  The
 * host's iterations^0.5 is its libm pow: the integers where that differs from the correctly
 * rounded square root are passed as root_n (ascending) with their pow values root_v (the
 * Python wrapper lists them once).  ws: csg_edge_mean_workspace_bytes(nf).  Replaces
 * _edge_mean_distance + _heat_parameters (surface_stencil.py:349-370) where a schedule
 * must not wait for the host.
 *
 * The _dev smoothing entries equal csg_smooth_geometry_batch / _patched,
 * csg_smooth_values / _patched and csg_type5_sheets_batch with the per-row iteration counts
 * (and widths) read from device arrays: the host launches up to `bound` iterations and rows
 * past their count are copied through, so the results equal the host-count entries' as long
 * as no count exceeds `bound` (callers verify that on the device). */
/* flag[0] = 1 when some of the n device ints exceeds bound (left as is otherwise): the device
 * check that a bounded loop covered every count.  No counterpart in the reference. */
int csg_flag_above(const int *values, int n, int bound, int *flag, void *stream);
int csg_heat_parameters(const float *v, const int *faces, int nf, const double *fwhm, int nfw,
                        const int *root_n, const double *root_v, int nroots, int *iters_out,
                        double *sigma_out, double *em_out, void *ws, size_t ws_bytes,
                        void *stream);
/* The same parameters from an edge mean already on the device (em: one double). */
int csg_heat_parameters_from_edge_mean(const double *em, const double *fwhm, int nfw,
                                       const int *root_n, const double *root_v, int nroots,
                                       int *iters_out, double *sigma_out, void *stream);
int csg_smooth_geometry_batch_dev(const float *v, int batch, const int *nbr_off, const int *nbr,
                                  int n, const int *iters_dev, int bound, float *out, void *ws,
                                  size_t ws_bytes, void *stream);
int csg_smooth_geometry_patched_dev(const float *v, int batch, int n, const int *iters_dev,
                                    int bound, int npatches, int rings, const int *perm,
                                    const int *vstart, const int *ring_end, const int *lverts,
                                    const int *aoff, const int *astart,
                                    const unsigned short *ladj, size_t lds_bytes, float *out,
                                    void *ws, size_t ws_bytes, void *stream);
int csg_smooth_values_dev(const double *values, int n, int nb, const float *geom, int meshes,
                          const int *mesh_of, const int *nbr_off, const int *nbr, int nnz,
                          const double *sigma_dev, const int *iters_dev, int bound,
                          int squared_norm, double *out, void *ws, size_t ws_bytes, void *stream);
int csg_smooth_values_patched_dev(const double *values, int n, int nb, const float *geom,
                                  int meshes, const int *mesh_of, const int *nbr_off,
                                  const int *nbr, int nnz, const double *sigma_dev,
                                  const int *iters_dev, int bound, int squared_norm, int npatches,
                                  int rings, const int *perm, const int *vstart,
                                  const int *ring_end, const int *lverts, const int *aoff,
                                  const int *astart, const unsigned short *ladj,
                                  size_t lds_bytes, double *out, void *ws, size_t ws_bytes,
                                  void *stream);
int csg_type5_sheets_batch_dev(const float *geom, const float *smoothed, int n, int sides,
                               const int *faces, const int *nbr_off, const int *nbr, int nnz,
                               const int *inc_off, const int *inc, int nfw,
                               const double *sigma_dev, const int *iters_dev, int bound,
                               const int *const *sheet_idx, const double *const *sheet_w, int m,
                               int out_dtype, void *out, void *ws, size_t ws_bytes, void *stream);

/* Depth potential (L + alpha diag(area)) x = rhs by colour-ordered SOR for `batch`
 * meshes of one topology (pts [batch][n][3], x and relres [batch]).  The system is a
 * CSR with duplicate (row, col) entries merged (contrib_off / contrib list each merged
 * entry's COO contributions); colour_off is a host array of ncolours + 1 offsets into
 * the device array colour_verts.  ell_cols / ell_src ([ell_width][n], solve order: row r
 * is vertex colour_verts[r]) give each row's merged entries (column, merged index or -1
 * for padding) in CSR order.  All max_iter x ncolours sweeps run in one persistent
 * launch; the call synchronises once at the end.
 * Replaces rotation_feature._solve_depth_potential_colored_sor
 * (src/cat_surface_gpu/rotation_feature.py:253-317). */
/* Placement of csg_depth_potential's SOR sweeps.  2 (default) and 1 put each mesh's rows
 * on workgroups of one XCD (mesh b on XCD b % 8, claimed through the XCC_ID register) with
 * one barrier per mesh, so that colours synchronise through that XCD's L2 without
 * agent-scope cache writeback and invalidation, when every XCD's share fits on its CUs;
 * x is read with L1-bypassing device-scope loads and written with plain stores (2) or
 * device-scope stores written through to memory (1).  0 always uses the device-wide grid
 * barrier.  All compute the same values in the same order.  Process-wide; returns the
 * previous setting.  No counterpart in the reference. */
int csg_set_sor_placement(int mode);
/* The XCD-local placements assume the device spreads workgroups over eight XCDs (workgroup
 * i on XCD i mod 8, each reporting its own XCC_ID).  After each XCD-local solve the claim
 * counters are read back: if an XCD's group received fewer workgroups than planned (a
 * CPX/DPX partition, where every workgroup reports one XCC_ID), the solve is repeated on
 * the device-wide grid barrier and the device keeps to that placement for the rest of the
 * process.  csg_reset_sor_xcd_check forgets that for the current device (returns 1 if it
 * had been recorded).  csg_set_sor_force_xcd (test hook; < 0 turns it off) makes every
 * XCD-local workgroup claim on XCD `xcd`, as such a partition does; returns the previous
 * value.  No counterpart in the reference. */
int csg_set_sor_force_xcd(int xcd);
int csg_reset_sor_xcd_check(void);
size_t csg_depth_workspace_bytes(int n, int nf, int nnz, int batch, int ell_width);
int csg_depth_potential(const float *pts, int batch, int n, const int *faces, int nf,
                        const int *inc_off, const int *inc, const int *row_off, const int *cols,
                        const int *contrib_off, const int *contrib, int nnz,
                        const int *colour_off, const int *colour_verts, int ncolours,
                        const int *ell_cols, const int *ell_src, int ell_width, double alpha,
                        int max_iter, double omega, double *x, double *relres, void *ws,
                        size_t ws_bytes, void *stream);

/* A partition of the mesh for csg_depth_potential_patched (device arrays; built by
 * surface_stencil.sor_patches).  Patch p's local vertices are lverts[vstart[p] ..
 * vstart[p + 1]) (global ids): its ncore[p] core vertices, then the `rings` rings of
 * their graph neighbourhood (rings = the colour count).  Its row table (entries
 * estart[p] .. estart[p + 1]) lists the vertices of rings < `rings`, sorted by (colour,
 * ring): e_row = the vertex's row in solve order, e_local = its local index, e_lcol
 * [entry][ell_width] = the local index of each ELL column; cstart [p][ncolours + 1] are the
 * colours' first entries (relative to estart[p]) and active [p][rings] the number of
 * colour t's entries of rings <= rings - 1 - t.  max_local / max_rows: the largest
 * patch's vertex and entry counts. */
typedef struct csg_sor_patches {
    int count, rings, max_local, max_rows;
    const int *vstart, *ncore, *lverts, *estart, *e_row, *e_local, *cstart, *active;
    const unsigned short *e_lcol;
} csg_sor_patches;

/* csg_depth_potential with the SOR sweeps blocked in time on mesh patches: one workgroup
 * per (patch, mesh), all of a mesh's on one XCD, iterates the colours of each sweep on the
 * patch in LDS and exchanges the cores once per sweep (instead of once per colour).
 * Identical values; falls back to csg_depth_potential's schedule when the patches do not
 * fit the device (rows per patch, resident workgroups per XCD).  No counterpart in the
 * reference (a schedule of rotation_feature.py:253-317). */
int csg_depth_potential_patched(const float *pts, int batch, int n, const int *faces, int nf,
                                const int *inc_off, const int *inc, const int *row_off,
                                const int *cols, const int *contrib_off, const int *contrib,
                                int nnz, const int *colour_off, const int *colour_verts,
                                int ncolours, const int *ell_cols, const int *ell_src,
                                int ell_width, double alpha, int max_iter, double omega,
                                const csg_sor_patches *patches, double *x, double *relres,
                                void *ws, size_t ws_bytes, void *stream);

/* csg_depth_potential_patched without waiting for the host: status (device int[2]) receives
 * the barrier time-out flag and, for XCD-local runs, whether every XCD group was placed; a
 * nonzero status means x is not valid (the host-waiting entry would have repeated the solve
 * on the device-wide barrier instead).  For schedules captured in a graph.  No counterpart
 * in the reference. */
int csg_depth_potential_patched_deferred(const float *pts, int batch, int n, const int *faces,
                                         int nf, const int *inc_off, const int *inc,
                                         const int *row_off, const int *cols,
                                         const int *contrib_off, const int *contrib, int nnz,
                                         const int *colour_off, const int *colour_verts,
                                         int ncolours, const int *ell_cols, const int *ell_src,
                                         int ell_width, double alpha, int max_iter, double omega,
                                         const csg_sor_patches *patches, double *x,
                                         double *relres, int *status, void *ws, size_t ws_bytes,
                                         void *stream);

/* Depth potential of the rotation feature by Jacobi-preconditioned conjugate gradients from
 * x = 0, to relative residual tol or max_iter iterations, for `batch` meshes of one topology
 * (the same system as csg_depth_potential).  Each mesh's workgroups run on one XCD and
 * exchange the search direction and partial sums through its L2; every reduction has a fixed
 * order, so results repeat exactly.  Writes x [batch][n] and relres [batch]; iterations_host
 * (optional) receives the iteration counts.  Synchronises once at the end.
 * Replaces _solve_depth_potential_cg (rotation_feature.py:193-250). */
size_t csg_depth_pcg_workspace_bytes(int n, int nf, int nnz, int batch);
int csg_depth_potential_pcg(const float *pts, int batch, int n, const int *faces, int nf,
                            const int *inc_off, const int *inc, const int *row_off,
                            const int *cols, const int *contrib_off, const int *contrib, int nnz,
                            double alpha, int max_iter, double tol, double *x, double *relres,
                            int *iterations_host, void *ws, size_t ws_bytes, void *stream);

/* Sphere points through a flow (out: [n][3] of dtype).  Replaces
 * surface_warp.apply_flow_to_sphere (src/cat_surface_gpu/surface_warp.py:500-584). */
int csg_apply_flow(const float *unit, int n, const void *flow, int nx, int ny, int dtype,
                   int inverse, void *out, void *stream);

/* float32 unit vectors of [n][3] points, as the warp normalises un-cached points
 * (surface_warp.py:518-522). */
int csg_normalise_points(const float *in, int n, float *out, void *stream);

/* Centroids of `count` spheres (host arrays of device pointers to [n][3] float32 vertices
 * and of vertex counts): centroids[3 s + c] (device, float64) = the strictly sequential
 * float64 sum of sphere s's coordinate c in vertex order, divided by n — the centring of
 * _normalise_sphere (src/cat_surface_gpu/gpu_pipeline.py:274-305).  Does not synchronise. */
int csg_sphere_centroids(const float *const *spheres, const int *counts, int count,
                         double *centroids, void *stream);

/* Sphere preparation of the pipeline on the device.  centroid (device, 3 doubles from
 * csg_sphere_centroids, or NULL to skip): p = f32(f64(v) - centroid), then
 * f32(f64(p) / |f64(p)|), setting *bad (device int, cleared by the caller) for a zero or
 * non-finite length (the reference raises ValueError); replaces _normalise_sphere
 * (gpu_pipeline.py:274-305).  rotation (host, row-major 3x3, may be NULL): then
 * f32(f64(p) m^T), replacing _rotate_vertices (:373-378).  Does not synchronise. */
int csg_prepare_sphere(const float *in, int n, const double *centroid, const double *rotation,
                       float *out, int *bad, void *stream);

/* csg_prepare_sphere with the rotation (row-major 3x3) read from device memory, e.g. a rotation
 * search's result that the host has not seen; rotation_dev is required. */
int csg_prepare_sphere_dev(const float *in, int n, const double *centroid,
                           const double *rotation_dev, float *out, int *bad, void *stream);

/* Rotation matrices (row-major 3x3 each, out [count][9]) of device Euler-angle triples
 * (alpha, beta, gamma): Rz(gamma) (Ry(beta) Rx(alpha)) assembled column-major and read
 * row-major, as surface_rotation.py:223-300 builds them, in float64 on the device. */
int csg_rotation_matrices(const double *angles, int count, double *out, void *stream);

/* Un-rotate the pole-frame warp by `unrotate` (host, row-major 3x3, p m^T) and blend it
 * with the main warp.  Replaces gpu_pipeline._average_xz_surfaces and its caller
 * (src/cat_surface_gpu/gpu_pipeline.py:248-271, :1294-1310). */
int csg_average_xz(const void *avg, const void *main_, int n, int dtype, const double *unrotate,
                   float *out, void *stream);

/* ---------------------------------------------------------------- location */

/* Uniform-grid triangle index on [-1, 1]^3, two phases: plan (synchronises, returns
 * the pair count) then fill (cell_off: grid^3 + 1 ints, cell_faces: npairs ints).
 * Replaces RotationGridIndex.from_geometry (src/cat_surface_gpu/surface_rotation.py:57-120). */
size_t csg_grid_index_plan_workspace_bytes(int nf);
int csg_grid_index_plan(const void *pts, int pts_dtype, const int *faces, int nf, int grid,
                        int margin, void *ws, size_t ws_bytes, int *npairs, void *stream);
size_t csg_grid_index_fill_workspace_bytes(int npairs);
int csg_grid_index_fill(const void *plan_ws, int nf, int grid, int npairs, int *cell_off,
                        int *cell_faces, void *ws, size_t ws_bytes, void *stream);

/* The same index in one asynchronous call (no host round trip): at most `capacity` pairs
 * are listed (cell_faces: capacity ints, cell_off: grid^3 + 1 ints) and *required (device
 * int) receives the pair count the index needs.  The index equals the two-phase one iff
 * *required <= capacity; otherwise it lists a subset (every access stays in bounds) and
 * must be rebuilt with capacity >= *required.  Built with this library's own scan and stable
 * LSD radix sort (no library call that records a memset), so it can be captured in a graph;
 * grid is needed for the workspace size. */
size_t csg_grid_index_build_workspace_bytes(int nf, int capacity, int grid);
int csg_grid_index_build(const void *pts, int pts_dtype, const int *faces, int nf, int grid,
                         int margin, int capacity, int *cell_off, int *cell_faces, int *required,
                         void *ws, size_t ws_bytes, void *stream);

/* CAT rotation SSD for na candidate angle triples (device).  The count of points outside
 * the index is kept in a device int: uncovered_dev if given (the call then does not
 * synchronise), else in the workspace; uncovered_host (optional) receives it after a
 * synchronisation.
 * Replaces RotationGridIndexDevice.cost_batch_prepared (surface_rotation.py:512-546). */
/* order[0..k) = the indices of the first k of n device costs in stable ascending order (ties
 * by index, NaN last), 1 <= k <= min(n, 64): the candidate selection
 * torch.sort(costs, stable=True).indices[:k] of the batched rotation search, without a
 * library sort (so that it can be captured in a graph).  No counterpart in the reference. */
int csg_top_k(const double *costs, int n, int k, int *order, void *stream);

size_t csg_rotation_costs_workspace_bytes(int ns, int na);
/* Face bounding spheres for the rotation locator (tri: [nf][3][3] float32 → sph: [nf][4]
 * float32 centre and radius, enlarged to cover the locator's tolerances).  Passed as `sph`
 * to csg_rotation_costs / _cost_matrices / _refine, it lets the locator skip candidates
 * that cannot be the closest; NULL scans every candidate.  The results are the same. */
int csg_face_spheres(const float *tri, int nf, float *sph, void *stream);

int csg_rotation_costs(const float *src_pts, const double *src_vals, int ns,
                       const double *tgt_vals, const int *faces, const float *tri,
                       const float *sph, const int *cell_off, const int *cell_faces,
                       int grid, const double *angles, int na, double *costs, int *loc_face,
                       float *loc_bary, void *ws, size_t ws_bytes, int *uncovered_host,
                       int *uncovered_dev, void *stream);

/* csg_rotation_costs without a workspace argument, for callers that cannot allocate device
 * memory (code under hipGraph capture, tight Nelder-Mead / serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  It allocates nothing: the partial
 * sums live in a fixed
 * device array of csg_rotation_costs_noalloc_max_bytes() bytes (4 MiB) per GPU, held
 * until the process ends; requests needing more (csg_rotation_costs_workspace_bytes(ns,
 * na) above it, e.g. more than 204 angles at 163,842 points) are rejected with -1.  The
 * count of points outside the index goes to uncovered_dev (required; cleared by a kernel,
 * not a memset, at each call).  Every call on one GPU shares the array, which keeps the
 * last call's partial sums, so calls that overlap on one GPU (different streams or
 * threads) race.  Does not synchronise; the results equal csg_rotation_costs'. */
size_t csg_rotation_costs_noalloc_max_bytes(void);
int csg_rotation_costs_noalloc(const float *src_pts, const double *src_vals, int ns,
                               const double *tgt_vals, const int *faces, const float *tri,
                               const float *sph, const int *cell_off, const int *cell_faces,
                               int grid, const double *angles, int na, double *costs,
                               int *loc_face, float *loc_bary, int *uncovered_dev, void *stream);

/* The same SSD for na candidate rotation matrices (float32 [na][9], row-major, q = M p;
 * na <= 65535) instead of angle triples: cost_batch_prepared's formulation
 * (surface_rotation.py:512-546) for rotations that come as matrices (the Hopf grid of
 * candidates).  With with_gradient, out[a] holds 10 float64 sums: the cost, its gradient
 * with respect to a left axis-angle perturbation exp([d]x) M (3), and the Gauss-Newton
 * matrix (xx, yy, zz, xy, xz, yz), taken with each point's located face held fixed; the
 * gradient has no counterpart in the reference (which refines by Nelder-Mead).  Points
 * outside the index are counted in *uncovered_dev, which the caller clears.  Does not
 * synchronise. */
size_t csg_rotation_cost_matrices_workspace_bytes(int ns, int na, int with_gradient);
int csg_rotation_cost_matrices(const float *src_pts, const double *src_vals, int ns,
                               const double *tgt_vals, const int *faces, const float *tri,
                               const float *sph, const int *cell_off, const int *cell_faces,
                               int grid, const float *mats, int na, int with_gradient, double *out,
                               void *ws, size_t ws_bytes, int *uncovered_dev, void *stream);

/* Rotation SSD of na <= 65535 candidate matrices (float32 [na][9], row-major) on feature
 * maps: src_map and tgt_map are float64 [ny][nx] equirectangular maps (pixel centre
 * (x+1/2, y+1/2) at azimuth 2 pi (x+1/2)/nx and polar angle pi (y+1/2)/ny, the layout of
 * the stencil sheets); for every stride-th pixel x in each direction, costs[a] sums
 * sin(polar) (S(M^T x) - T(x))^2 with S sampled bilinearly (x circular, y clamped).  The
 * continuous form of cost_batch_prepared's SSD (surface_rotation.py:512-546) without point
 * location: a coarse score for many candidates (the Hopf grid) before refinement.  No
 * counterpart in the reference.  Does not synchronise. */
size_t csg_rotation_map_costs_workspace_bytes(int nx, int ny, int stride, int na);
int csg_rotation_map_costs(const double *src_map, const double *tgt_map, int nx, int ny,
                           int stride, const float *mats, int na, double *costs, void *ws,
                           size_t ws_bytes, void *stream);

/* Gradient refinement of m <= 256 candidate rotations (no counterpart in the reference,
 * which refines its best seed by Nelder-Mead, surface_rotation.py:594-699).  start holds
 * m Euler-angle triples (start_is_angles = 1, the reference's angle convention) or m
 * row-major float64 matrices.  Each candidate takes `steps` Levenberg-Marquardt steps on
 * a left axis-angle perturbation driven by csg_rotation_cost_matrices' gradient and
 * Gauss-Newton terms (step clamped to max_step radians, initial damping `damping`; a
 * step is kept only if it lowers the cost).  Writes the final matrices [m][9], their
 * Euler angles [m][3] and those angles' costs in the reference's angle formulation
 * (csg_rotation_costs) [m]; clears and fills *uncovered_dev.  Nothing synchronises. */
size_t csg_rotation_refine_workspace_bytes(int ns, int m);
int csg_rotation_refine(const float *src_pts, const double *src_vals, int ns,
                        const double *tgt_vals, const int *faces, const float *tri,
                        const float *sph, const int *cell_off, const int *cell_faces, int grid,
                        const double *start, int start_is_angles, int m, int steps,
                        double max_step, double damping, double *mats_out, double *angles_out,
                        double *costs_out, void *ws, size_t ws_bytes, int *uncovered_dev,
                        void *stream);

/* Exact closest triangle (first minimum among the query cell's candidates) with
 * barycentric weights.  The count of queries farther than `reach` from their best
 * triangle (their true closest triangle may lie outside the searched cells) is left in
 * the first int of ws (device); far_host, if given, receives it after a synchronisation,
 * else the call does not synchronise.
 * Replaces the stencil helper's find_closest_polygon_point + weights
 * (native/cat_surface_stencil_builder.c:107-170). */
int csg_closest_triangles(const double *q, int nq, const double *verts, const int *faces,
                          const int *cell_off, const int *cell_faces, int grid, double reach,
                          int *face_out, int *idx_out, double *w_out, int *far_host,
                          void *ws, size_t ws_bytes, void *stream);

/* Steps of the in-process stencil builder around its closest-triangle searches
 * (native/cat_surface_stencil_builder.c).  csg_stencil_sphere: pts = float64(v), centred
 * on the mean and scaled to unit length when normalise; out_sphere[j] = (lo + hi) + 0.975 ·
 * mean |pts| · ico[j] with lo / hi the coordinate bounds of pts (params: 8 doubles of
 * device scratch).  csg_resample_unit: res[j] = sum_k w[j][k] pts[idx[j][k]], unit[j] =
 * res[j] / |res[j]|.  csg_unit_points64: out = float32(p / |p|). */
int csg_stencil_sphere(const float *v, int n, int normalise, const double *ico, int m,
                       double *pts, double *out_sphere, double *params, void *stream);
int csg_resample_unit(const double *pts, const int *idx, const double *w, int m, double *res,
                      double *unit, void *stream);
int csg_unit_points64(const double *pts, int n, float *out, void *stream);

/* csg_closest_triangles with each candidate first tested against its face's bounding
 * sphere (computed into ws, csg_closest_triangles_workspace_bytes(nf) bytes, with a
 * face-major copy of the faces' corners): a candidate farther than the scanning lane's
 * best is skipped, which changes no result.  The far count goes to far_dev (device int),
 * and to far_host after a synchronisation if given. */
size_t csg_closest_triangles_workspace_bytes(int nf);
/* Exact closest triangles for queries near the mesh on an index built with margin 0 (each
 * face in the cells its bounding box overlaps; csg_grid_index_build with margin 0): a query
 * scans its own cell, and the (2·span + 1)^3 block of cells around it only when its best
 * distance exceeds the distance to its cell's walls; a query whose best distance exceeds the
 * distance to the block's walls counts as far (far_dev, far_host as in
 * csg_closest_triangles_pruned, same workspace).  Ties go to the smaller face index.  The
 * results equal csg_closest_triangles_pruned's on an index with margin `span` whenever
 * neither counts a query as far.  No counterpart in the reference (a search strategy for the
 * stencil builder's query, native/cat_surface_stencil_builder.c:116-131). */
int csg_closest_triangles_block(const double *q, int nq, const double *verts, const int *faces,
                                int nf, const int *cell_off, const int *cell_faces, int grid,
                                int span, int *face_out, int *idx_out, double *w_out,
                                int *far_dev, int *far_host, void *ws, size_t ws_bytes,
                                void *stream);
int csg_closest_triangles_pruned(const double *q, int nq, const double *verts, const int *faces,
                                 int nf, const int *cell_off, const int *cell_faces, int grid,
                                 double reach, int *face_out, int *idx_out, double *w_out,
                                 int *far_dev, int *far_host, void *ws, size_t ws_bytes,
                                 void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CSG_H */
