// Optional HIP-event timing of the hot kernels (internal to libcsg_hip).
//
// When enabled, each launch of a timed kernel outside stream capture is
// bracketed by two HIP events recorded on the stream it is launched on, and
// the launch's algorithmic bytes (what the reference's operator sequence
// reads and writes once for that work) are accumulated with it.  Reading the
// totals synchronises on the recorded events.  Used for roofline reports.
#pragma once

#include "csg_common.h"

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
    kTimedCount
};

const char *timed_kernel_name(int id);
int timing_enable(int on);
int timing_read(int id, double *ms, double *bytes, long long *launches);

// RAII-style bracket: construct before the launch, call done() after it.
class TimedLaunch {
public:
    TimedLaunch(int id, hipStream_t s, double bytes);
    void done();

private:
    int id_;
    hipStream_t s_;
    double bytes_;
    hipEvent_t a_ = nullptr;
};

}  // namespace csg
