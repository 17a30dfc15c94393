// Host-side DARTEL drivers (internal to libcsg_hip).
#pragma once

#include "dartel_kernels.h"

namespace csg {

constexpr int kMaxLevels = 40;

int fmg_levels(int nx, int ny, int *lnx, int *lny);

template <typename T>
size_t fmg_workspace_bytes(int nx, int ny, int batch);

// coarse_init != 0 runs the reference's coarse-initialisation branch
// (no or all-zero initial guess), else `cycles` V-cycles from `initial`.
// Selects the multigrid schedule for later calls: 1 = fused (LDS tiles and a
// single-workgroup coarse hierarchy, where it applies), 0 = one launch per
// operator call.  Both give identical values.  Returns the previous setting.
int set_fused_fmg_schedule(int enable);
// Levels of an nx x ny FMG that the fused schedule relaxes on LDS tiles (the
// levels above its single-workgroup coarse hierarchy); 0 when the fused
// schedule does not apply.
int fused_tiled_levels(int nx, int ny, int nit, int elem);

template <typename T>
int fmg_run(const T *hes, const T *rhs, const double *params, int cycles, int nit,
            const T *initial, int coarse_init, int nx, int ny, int batch, T *out, void *ws,
            size_t ws_bytes, hipStream_t s);

template <typename T>
struct StepArgs {
    const T *src, *tgt;  // [batch][ny][nx]
    const T *v;          // [batch][2][ny][nx]
    T *v_out;            // [batch][2][ny][nx]
    const T *dist;       // [batch][ny][nx] or null
    size_t dist_batch_stride;
    int k;
    double params[5];
    double lmreg;
    int cycles, nit, code, assume_nonzero;
    int nx, ny, batch;
    double *metrics;  // metrics[b * metrics_stride + {0,1,2}]
    int metrics_stride;
};

template <typename T>
size_t step_workspace_bytes(int nx, int ny, int batch);
template <typename T>
int dartel_step_run(const StepArgs<T> &a, void *ws, size_t ws_bytes, hipStream_t s);

template <typename T>
struct SolveArgs {
    const T *src, *tgt;  // [batch][ny][nx]; consecutive `steps_per_job` elements form a job
    int nx, ny, batch, steps_per_job;
    int loop, cycles, nit, its, code, assume_nonzero;
    double lmreg;
    T *flow;          // [batch / steps_per_job][2][ny][nx] or null
    T *velocity;      // [batch][2][ny][nx] final velocities or null
    double *metrics;  // [batch][loop * its][3] or null
};

template <typename T>
size_t solve_workspace_bytes(int nx, int ny, int batch);
template <typename T>
int solve_run(const SolveArgs<T> &a, void *ws, size_t ws_bytes, hipStream_t s);

}  // namespace csg
