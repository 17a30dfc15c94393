// Depth-potential solve for the rotation feature (internal to libcsg_hip).
#pragma once

#include "csg_common.h"

namespace csg {

constexpr int kSorBlock = 512;
// grid-barrier state: kBarrierSlots slots of kBarrierSlotWords words (one cache line each)
constexpr int kBarrierSlots = 128;
constexpr int kBarrierSlotWords = 32;
constexpr int kMaxColours = 32;

// Colour groups: vertices of colour c are colour_verts[off[c] .. off[c+1]).
struct SorColours {
    int count;
    int off[kMaxColours + 1];
};

// Workgroups of each XCD's group in the XCD-local SOR (0: the XCD holds no mesh).
struct SorXcdPlan {
    int blocks[8];
    int force_xcd;  // < 0: each workgroup claims on the XCD it runs on (see xcd_claim)
};

struct DepthArgs {
    const float *pts;  // [batch][n][3]
    const int *faces;  // [nf][3]
    const int *inc_off, *inc;
    const int *row_off, *cols;          // merged CSR, nnz entries
    const int *contrib_off, *contrib;   // per merged entry, its COO entries
    const int *colour_verts;
    const int *ell_cols, *ell_src;      // colour-ordered ELL, ell_width slots per row
    int ell_width;
    int batch, n, nf, nnz;
    double alpha;
    // workspace
    double *fnormal, *farea, *coo, *normals, *areas, *lap, *sys, *mc, *diag, *rhs, *partial;
    double *ell_val, *rhs_c, *diag_c;   // [batch][ell_width * n], [batch][n] in colour order
    unsigned *barrier;  // kBarrierSlots x kBarrierSlotWords; the last slot holds the timeout flag
    double *x;          // [batch][n]
    double *x_alt;      // [batch][n]: the patch-blocked SOR's second buffer
};

// Patches of the mesh for the LDS-blocked SOR (device arrays, as csg_sor_patches).
struct SorPatches {
    int count, rings, max_local, max_rows;
    const int *vstart, *ncore, *lverts, *estart, *e_row, *e_local, *cstart, *active;
    const unsigned short *e_lcol;
};

// How a depth_sor call placed its sweeps: `xcd_local` when each mesh's
// workgroups were meant to run on XCD b % 8 (plan.blocks[x] workgroups of
// XCD x), which holds only where the device spreads workgroups over eight
// XCDs; sor_placement_complete checks that afterwards.
struct SorRun {
    bool xcd_local;
    SorXcdPlan plan;
};

int depth_setup(const DepthArgs &a, hipStream_t s);
// Zero the barrier state (counters, verdicts, time-out flag) with a kernel.
int clear_barrier(unsigned *bar, hipStream_t s);
// The run's time-out and placement verdicts into status[0..1] on the device
// (no host round trip): nonzero means the solve is not valid.
int sor_status(const DepthArgs &a, const SorRun &run, int *status, hipStream_t s);
// patches (may be null): sweep in LDS on those patches when they fit the
// device (one workgroup per patch and mesh, all resident), else as without.
// allow_xcd_local = false keeps the sweeps on the device-wide grid barrier.
int depth_sor(const DepthArgs &a, const SorColours &co, int max_iter, double omega,
              const SorPatches *patches, bool allow_xcd_local, SorRun *run, hipStream_t s);
// After the stream has drained: whether every XCD group of an XCD-local run
// received all its workgroups.  On a device that does not spread workgroups
// over eight XCDs (a CPX/DPX partition reports one XCC_ID for all of them)
// some groups claim nothing, their meshes are never swept, and nothing waits
// for them: the caller must then solve again without the XCD-local placement.
bool sor_placement_complete(const DepthArgs &a, const SorRun &run, hipStream_t s, int *rc);
int depth_residual(const DepthArgs &a, double *out, hipStream_t s);
// Fixed row chunks of one mesh in the PCG (also the most workgroups per mesh).
constexpr int kPcgChunks = 32;
// Scratch of depth_pcg: r, z, p, q of every mesh and the partial sums.
size_t pcg_workspace_bytes(int n, int batch);
// Jacobi-preconditioned CG on the system depth_setup assembled (x starts at 0),
// to relative residual tol or max_iter iterations; iterations and relres are
// device arrays [batch].  XCD-local groups when allowed and the device fits
// them (run says so, for sor_placement_complete), else one workgroup per mesh.
int depth_pcg(const DepthArgs &d, double *ws, int max_iter, double tol, bool allow_xcd_local,
              int force_xcd, int *iterations, double *relres, SorRun *run, hipStream_t s);
// Process-wide placement of the SOR sweeps: 2 (default) each mesh's rows on
// one XCD, x stored through that XCD's L2; 1 the same with x stores written
// through to memory; 0 the device-wide grid barrier.  1 and 2 apply when the
// meshes fit.  Returns the previous setting.  All compute identical values.
int set_sor_xcd_local(int mode);
// Test hook: every XCD-local SOR workgroup claims on XCD `xcd` (< 0: off).
int set_sor_force_xcd(int xcd);
int sor_force_xcd();

}  // namespace csg
