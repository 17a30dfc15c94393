// Optional HIP-event timing of the hot kernels (internal to libcsg_hip).
//
// When enabled, each launch of a timed kernel outside stream capture goes
// through hipExtLaunchKernelGGL with a start and a stop event, which the
// runtime stamps when the kernel itself starts and ends on the stream it is
// launched on (no dispatch gap inside the interval), and the launch's
// algorithmic bytes (what the reference's operator sequence reads and writes
// once for that work) are accumulated with it.  Reading the totals
// synchronises on the recorded events.  Used for roofline reports.  When
// timing is off both events are null and the launch is an ordinary one.
#pragma once

#include "csg_common.h"

#include <hip/hip_ext.h>

namespace csg {

enum TimedKernel : int {
    kTimedRelaxTile = 0,   // fine-level relax_membrane call on LDS tiles
    kTimedFmgCoarse,       // coarse levels of one FMG cycle (or initialisation)
    kTimedRestrict,        // residual + restriction of one fine level
    kTimedRelaxSweep,      // one half-sweep of the per-operator schedule
    kTimedSor,             // all SOR sweeps of one depth-potential solve (rows from memory)
    kTimedSorRegisters,    // the same with the rows held in registers
    kTimedSorXcdLocal,     // the same with each mesh's rows on one XCD
    kTimedSorPatched,      // the SOR iterated in LDS on mesh patches, one exchange per sweep
    kTimedSquaring,        // one squaring step (b, A and, between steps, phi and J composed)
    kTimedObjective,       // jac_div + initialise_objfun + smalldef_jac, fused
    kTimedComposeSelf,     // phi o phi (and J) of the exponentiation
    kTimedCombine,         // rhs = b·2^-k + L v, A·2^-k, with the step's reductions
    kTimedSmoothGeometry,  // `rings` heat-smoothing iterations of the geometry on mesh patches
    kTimedPcg,             // the whole Jacobi-preconditioned CG of a batch of depth potentials
    kTimedCount
};

const char *timed_kernel_name(int id);
int timing_enable(int on);
int timing_read(int id, double *ms, double *bytes, long long *launches);

// Construct before the launch, pass start() / stop() to hipExtLaunchKernelGGL,
// call done() after it.
class TimedLaunch {
public:
    TimedLaunch(int id, hipStream_t s, double bytes);
    hipEvent_t start() const { return a_; }
    hipEvent_t stop() const { return b_; }
    void done();

private:
    int id_;
    double bytes_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

}  // namespace csg
