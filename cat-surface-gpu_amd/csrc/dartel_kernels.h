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

}  // namespace csg
