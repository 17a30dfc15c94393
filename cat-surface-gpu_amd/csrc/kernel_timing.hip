// HIP-event timing of the hot kernels; see kernel_timing.h.
#include "kernel_timing.h"

#include <atomic>
#include <mutex>
#include <vector>

namespace csg {

namespace {

struct Pending {
    hipEvent_t a, b;
    double bytes;
};

struct TimerState {
    std::mutex m;
    std::vector<hipEvent_t> pool;
    std::vector<Pending> pending[kTimedCount];
};

std::atomic<int> g_on{0};
TimerState &state()
{
    static TimerState st;
    return st;
}

hipEvent_t take_event()
{
    TimerState &st = state();
    {
        std::lock_guard<std::mutex> lock(st.m);
        if (!st.pool.empty()) {
            hipEvent_t e = st.pool.back();
            st.pool.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return true;  // unknown: do not record
    }
    return st != hipStreamCaptureStatusNone;
}

const char *kNames[kTimedCount] = {"k_relax_tile",    "k_fmg_coarse",     "k_restrict_residual",
                                   "k_relax_color",   "k_sor_persistent", "k_sor_registers",
                                   "k_sor_xcd_local", "k_sor_patched",    "k_squaring",
                                   "k_objective",     "k_compose_self",   "k_combine",
                                   "k_smooth_geometry_patch", "k_pcg"};

}  // namespace

const char *timed_kernel_name(int id)
{
    return (id >= 0 && id < kTimedCount) ? kNames[id] : nullptr;
}

int timing_enable(int on)
{
    return g_on.exchange(on ? 1 : 0);
}

TimedLaunch::TimedLaunch(int id, hipStream_t s, double bytes) : id_(id), bytes_(bytes)
{
    if (!g_on.load(std::memory_order_relaxed) || capturing(s))
        return;
    a_ = take_event();
    b_ = a_ ? take_event() : nullptr;
    if (!b_) {
        TimerState &st = state();
        std::lock_guard<std::mutex> lock(st.m);
        if (a_)
            st.pool.push_back(a_);
        a_ = nullptr;
    }
}

void TimedLaunch::done()
{
    if (!a_)
        return;
    TimerState &st = state();
    std::lock_guard<std::mutex> lock(st.m);
    st.pending[id_].push_back({a_, b_, bytes_});
    a_ = b_ = nullptr;
}

int timing_read(int id, double *ms, double *bytes, long long *launches)
{
    if (id < 0 || id >= kTimedCount)
        return fail(kInvalidArgument, "no timed kernel %d", id);
    std::vector<Pending> list;
    TimerState &st = state();
    {
        std::lock_guard<std::mutex> lock(st.m);
        list.swap(st.pending[id]);
    }
    double t = 0.0, by = 0.0;
    for (const Pending &p : list) {
        CSG_HIP_CHECK(hipEventSynchronize(p.b));
        float e = 0.0f;
        CSG_HIP_CHECK(hipEventElapsedTime(&e, p.a, p.b));
        t += e;
        by += p.bytes;
    }
    {
        std::lock_guard<std::mutex> lock(st.m);
        for (const Pending &p : list) {
            st.pool.push_back(p.a);
            st.pool.push_back(p.b);
        }
    }
    if (ms)
        *ms = t;
    if (bytes)
        *bytes = by;
    if (launches)
        *launches = (long long)list.size();
    return kOk;
}

}  // namespace csg
