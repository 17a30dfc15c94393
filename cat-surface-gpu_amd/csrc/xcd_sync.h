// Workgroup groups pinned to one XCD and synchronised through its L2
// (internal to libcsg_hip).
//
// MI355X has eight XCDs, each with its own L2; the L2s are not coherent with
// one another, so a device-wide barrier between dependent phases needs
// agent-scope fences that write back and invalidate L2.  When a phase only
// exchanges data among the workgroups of one group, and every workgroup of the
// group runs on the same XCD, that XCD's L2 is the point of coherence: data is
// written with plain stores (L1 is write-through, the lines stay in the L2),
// each wave waits for its stores to complete before the workgroup barrier, and
// readers use device-scope relaxed loads (sc1), which bypass the CU's L1 (not
// coherent between CUs) and are served by the L2.  (Plain loads after an L1
// invalidation per phase do not see the other CUs' stores on MI355X.)
//
// Placement is not assumed from the dispatch order: thread 0 of every
// workgroup reads its XCD from the XCC_ID hardware register and claims the
// next slot of that XCD (xcd_claim); the caller maps (XCD, slot) to its group
// and member, and workgroups beyond the work exit at once.  Launches carry
// spare workgroups per XCD and every wait is bounded (kXcdSpinLimit polls), so
// a placement that leaves a group short ends in a timeout flag, not a hang.
#pragma once

#include "csg_common.h"

namespace csg {

constexpr int kXcds = 8;
constexpr int kXcdSpare = 4;
constexpr long kXcdSpinLimit = 1L << 22;

__device__ __forceinline__ int xcd_id()
{
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return (int)(v & (kXcds - 1));
}

// Thread 0 claims the next slot of its XCD from claim[xcd * stride]; every
// thread returns (xcd, slot).  force_xcd >= 0 makes every workgroup claim as
// if it ran on that XCD (what a partition whose XCDs all report one XCC_ID
// does; set only by the test hook csg_set_sor_force_xcd).
__device__ __forceinline__ void xcd_claim(unsigned *claim, int stride, int force_xcd, int &xcd,
                                          int &slot)
{
    __shared__ int s_xcd, s_slot;
    if (threadIdx.x == 0) {
        const int x = force_xcd >= 0 ? (force_xcd & (kXcds - 1)) : xcd_id();
        s_xcd = x;
        s_slot = (int)__hip_atomic_fetch_add(claim + (size_t)x * stride, 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    xcd = s_xcd;
    slot = s_slot;
}

// Read a value another workgroup of the group stored in this phase's data.
template <typename T>
__device__ __forceinline__ T load_l2(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier of one group on one arrival counter, which only grows: the phase
// whose arrivals bring it to `target` (phase number x group size, modulo
// 2^32) is complete when it reaches that, so nothing is reset.
__device__ __forceinline__ bool group_barrier(unsigned *cnt, unsigned target, int *timeout)
{
    __shared__ int ok;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        long spins = 0;
        while ((int)(__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                     target) < 0) {
            if (++spins > kXcdSpinLimit) {
                __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    return ok != 0;
}

}  // namespace csg
