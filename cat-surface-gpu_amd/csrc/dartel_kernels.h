// Launch wrappers for the DARTEL grid and FMG kernels (internal to libcsg_hip).
#pragma once

#include "csg_common.h"

namespace csg {

template <typename T>
struct ObjectiveArgs {
    const T *src, *tgt, *v, *map, *jac, *dist;
    double s_div, s_small;
    int nx, ny, batch;
    size_t istride, vstride, mstride, jstride, dstride;
    T *rhs, *hes, *smap, *sjac;
    size_t rstride, hstride;
    double *partial;  // [batch][nblocks], may be null
    bool disp = false;  // maps in displacement form (float32 internal path)
};

inline int objective_blocks(int nx, int ny) { return (int)grid_for((size_t)nx * ny); }

template <typename T>
int launch_expdef_init(const T *v, double scale, int nx, int ny, int batch, size_t vstride,
                       T *map, T *jac, size_t mstride, size_t jstride, hipStream_t s,
                       bool disp = false);
template <typename T>
int launch_compose_self(const T *map, const T *jac, int nx, int ny, int batch, size_t mstride,
                        size_t jstride, T *omap, T *ojac, hipStream_t s, bool disp = false);
template <typename T>
int launch_disp_to_coord(const T *d, int nx, int ny, int batch, size_t stride, T *out,
                         hipStream_t s);
template <typename T>
int launch_compose(const T *a, const T *ja, const T *b, const T *jb, int nx, int ny, T *omap,
                   T *ojac, hipStream_t s);
template <typename T>
int launch_objective(const ObjectiveArgs<T> &a, hipStream_t s);
template <typename T>
int launch_squaring(const T *bin, const T *ain, const T *map, const T *jac, int nx, int ny,
                    int batch, size_t bstride, size_t astride, size_t mstride, size_t jstride,
                    T *bout, T *aout, T *inc, T *omap, T *ojac, hipStream_t s,
                    bool disp = false);
template <typename T>
int launch_combine(const T *v, const T *bf, const T *af, const T *br, const T *ar, double scale,
                   double w00, double w01, double w10, int nx, int ny, int batch, size_t vstride,
                   size_t bstride, size_t astride, T *rhs, T *hes, double *partial_reg,
                   double *partial_norm, int *nblocks_out, hipStream_t s);
template <typename T>
int launch_regularization(const T *v, int rtype, const double *params, int nx, int ny, T *out,
                          hipStream_t s);
int launch_reduce_partials(const double *partial, int nblocks, int batch, double scale,
                           double *out, int out_stride, int accumulate, hipStream_t s);
int launch_sqrt_inplace(double *v, int batch, int stride, hipStream_t s);
template <typename T>
int launch_update_velocity(const T *v, const T *u, const T *dist, int nx, int ny, int batch,
                           size_t vstride, size_t dstride, T *out, hipStream_t s);
template <typename T>
int launch_sample_field(const T *field, int nx, int ny, const T *x, const T *y, size_t n, T *out,
                        hipStream_t s);
template <typename T>
int launch_jac_div(const T *jac, const T *v, double scale, int nx, int ny, T *out, hipStream_t s);

// ---------------------------------------------------------------- FMG

struct MembraneWeights {
    double w00, w01, w10;
};

// Level weights from (sx, sy, mu, lambda, id).  dartel_grid.py:1040-1043, :1101-1104
inline MembraneWeights membrane_weights(const double *p)
{
    double sx = p[0], sy = p[1], mu = p[2];
    return {mu * (2.0 * sx * sx + 2.0 * sy * sy) + p[4], -mu * sy * sy, -mu * sx * sx};
}

template <typename T>
int launch_resize(const T *in, int channels, int snx, int sny, int tnx, int tny, int batch,
                  size_t istride, size_t ostride, T *out, int accumulate, hipStream_t s);
// Two resizes of one source and target grid (each with its own planes and
// strides) in one launch; values equal two launch_resize calls.
template <typename T>
int launch_resize_two(const T *in0, int channels0, size_t istride0, size_t ostride0, T *out0,
                      const T *in1, int channels1, size_t istride1, size_t ostride1, T *out1,
                      int snx, int sny, int tnx, int tny, int batch, hipStream_t s);
template <typename T>
int launch_relax_sweep(const T *a, const T *b, T *u, MembraneWeights w, int nx, int ny,
                       int batch, size_t astride, size_t bstride, int color, hipStream_t s);
template <typename T>
int launch_relax(const T *a, const T *b, T *u, MembraneWeights w, int nx, int ny, int batch,
                 size_t astride, size_t bstride, int nit, hipStream_t s);
template <typename T>
int launch_apply_membrane(const T *a, const T *u, MembraneWeights w, int nx, int ny, int batch,
                          size_t astride, size_t ustride, T *out, const T *b_minus,
                          hipStream_t s);
template <typename T>
int launch_coarse_solve(const T *a, const T *b, double id, int P, int batch, size_t astride,
                        size_t bstride, T *u, hipStream_t s);
template <typename T>
int launch_any_nonzero(const T *u, size_t n, int batch, size_t stride, int *flags,
                       hipStream_t s);

// ---------------------------------------------------------------- fused FMG schedule

constexpr int kFmgMaxLevels = 40;
constexpr int kTileThreads = 512;

// Interior of the fine-level relaxation tiles.  Any shape with an even width
// gives the same values (the halo makes every tile's interior exact); the
// shape only changes the speed, so it is a tuning knob: a process-wide table
// maps (level width, level height, element size) to a shape, default 32 x 16.
struct TileShape {
    int x, y;
};
constexpr TileShape kTileDefault{32, 16};
TileShape relax_tile_shape(int nx, int ny, int elem);
// tx = ty = 0 removes the entry.  Fails unless tx is even, ty >= 1 and a tile
// with the default nit = 3 halo fits the LDS budget for that element size.
int set_relax_tile_shape(int nx, int ny, int elem, int tx, int ty);
constexpr int kCoarseThreads = 1024;

// Levels c0 .. L-1 of the pyramid, relaxed by one workgroup per job in LDS.
// What does not change during one multigrid solve is kept in global memory
// between its coarse launches: `meta` (coarse_meta_bytes) holds the level
// table and the resolved resize taps, apyr[l] (l > c0) level l's A, 3·nx·ny
// values per job.  The first launch of a solve builds and stores them, the
// later ones load them.
template <typename T>
struct CoarseLevels {
    int L, c0;
    int nx[kFmgMaxLevels], ny[kFmgMaxLevels];
    T w00[kFmgMaxLevels], w01[kFmgMaxLevels], w10[kFmgMaxLevels];
    unsigned char *meta;
    T *apyr[kFmgMaxLevels];
};
// Bytes of CoarseLevels::meta for a pyramid of L levels (any first coarse level).
template <typename T>
size_t coarse_meta_bytes(const int *nx, const int *ny, int L);

// Largest dynamic LDS allocation a block may use on the current device (the
// kernels' limits are raised to it once per process).
size_t fmg_lds_budget();
template <typename T>
size_t tile_lds_bytes(int sweeps, TileShape tile);
// Whether the tile's pixels of one colour fit the relaxation kernel's threads.
bool tile_fits_threads(int sweeps, TileShape tile);
template <typename T>
size_t coarse_lds_bytes(const int *nx, const int *ny, int c0, int L);

// `sweeps` red-black half-sweeps of one fine level on LDS tiles with a halo of
// `sweeps` pixels (even nx only); u_out = relaxed(u_in [+ prolong(uc)]).
template <typename T>
int launch_relax_tile(const T *a, const T *b, const T *uin, T *uout, const T *uc, int cnx,
                      int cny, MembraneWeights w, int nx, int ny, int batch, size_t astride,
                      size_t bstride, size_t cstride, int sweeps, hipStream_t s);
// bc = restrict(b - (L + A) u); uc = 0.
template <typename T>
int launch_restrict_residual(const T *a, const T *b, const T *u, MembraneWeights w, int nx,
                             int ny, int cnx, int cny, int batch, size_t astride, size_t bstride,
                             size_t cstride, T *bc, T *uc, hipStream_t s);
// Levels c0.. of the FMG in one workgroup per job: init != 0 runs the coarse
// initialisation (coarse solve, then prolong + `cycles` cycles per level down
// to c0) from b = B0[c0]; otherwise `cycles` cycles from level c0 with b = B[c0]
// and u as the initial guess.  u (level c0) is overwritten.  build != 0 for
// the first launch of a solve (required with init), which stores what the
// solve's later launches load (CoarseLevels::meta, apyr).
template <typename T>
int launch_fmg_coarse(const CoarseLevels<T> &lv, const T *a, const T *b, T *u, size_t astride,
                      size_t bstride, double id, int nit, int cycles, int init, int build,
                      int batch, size_t lds_bytes, hipStream_t s);

}  // namespace csg
