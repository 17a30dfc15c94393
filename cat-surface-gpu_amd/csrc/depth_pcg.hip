// Jacobi-preconditioned conjugate gradients for the rotation feature's depth
// potential (the reference's "pcg" solver, rotation_feature.py:193-250), for a
// batch of meshes sharing one topology, after depth_setup has assembled each
// mesh's merged CSR system, right-hand side and diagonal.
//
// Placement: mesh b's workgroups all run on XCD b % 8 (claimed through the
// XCC_ID register, xcd_sync.h) and split its rows into contiguous chunks; the
// search direction p and the per-workgroup partial sums are exchanged through
// that XCD's L2 with one arrival counter per mesh.  An iteration has three
// such barriers: after q = A p and the partial sums of p.q, after the update
// of x and r with the partial sums of r.r and r.z, and after the update of p.
// Every reduction has a fixed order (a thread's rows in order, a shuffle tree
// per wave, the waves in order, the workgroups in order), and every
// workgroup of a mesh sums the partials itself in the same order, so all of
// them take the same step sizes and stop at the same iteration.  A device
// that does not spread workgroups over eight XCDs runs one workgroup per
// mesh instead (the caller checks the claims afterwards).
#include "depth_kernels.h"
#include "kernel_timing.h"
#include "xcd_sync.h"

namespace csg {

namespace {

constexpr int kPcgThreads = 256;
constexpr int kPcgCounterBase = 16;  // barrier slot of mesh b's arrival counter: 16 + b

struct PcgArgs {
    const int *row_off, *cols;
    const double *sys, *rhs, *diag;  // [batch][nnz], [batch][n], [batch][n]
    int n, nnz, batch;
    double *x, *r, *z, *p, *q;       // [batch][n]
    double *partial;                 // [batch][kPcgChunks][4]
    int max_iter;
    double tol;
    int *iterations;                 // [batch]
    double *relres;                  // [batch]
};

struct PcgPlan {
    int blocks[kXcds];    // workgroups of XCD x in use
    int per_mesh[kXcds];  // workgroups per mesh on XCD x
    int force_xcd;        // < 0: each workgroup claims on the XCD it runs on
};

// Two sums of the workgroup in a fixed order (shuffle tree per wave, then the
// waves in order by thread 0); every thread returns them.
__device__ __forceinline__ void wg_sum2(double a, double b, double &sa, double &sb)
{
    __shared__ double ws[kPcgThreads / 64][2];
    __shared__ double res[2];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        ws[wave][0] = a;
        ws[wave][1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int w = 0; w < kPcgThreads / 64; ++w) {
            x += ws[w][0];
            y += ws[w][1];
        }
        res[0] = x;
        res[1] = y;
    }
    __syncthreads();
    sa = res[0];
    sb = res[1];
    __syncthreads();
}

// GROUP: the mesh's W workgroups exchange through the XCD's L2 (see above);
// otherwise one workgroup per mesh (blockIdx.x = mesh) and workgroup barriers.
// The rows are split into kPcgChunks fixed chunks whatever W is; workgroup w
// handles chunks w, w + W, ..., reduces each chunk on its own (thread t takes
// rows t, t + T, .. of the chunk) and writes the chunk's partial sums; every
// workgroup then adds the chunks' partials in chunk order.  The sums, and so
// the iterates, are the same for any W: batched or alone, grouped or not.
template <bool GROUP>
__global__ void __launch_bounds__(kPcgThreads)
k_pcg(PcgArgs a, PcgPlan plan, unsigned *bar, int *timeout)
{
    int b = blockIdx.x, W = 1, w = 0;
    unsigned *cnt = nullptr;
    if constexpr (GROUP) {
        int xcd, slot;
        xcd_claim(bar, kBarrierSlotWords, plan.force_xcd, xcd, slot);
        if (slot >= plan.blocks[xcd])
            return;
        W = plan.per_mesh[xcd];
        b = xcd + kXcds * (slot / W);
        w = slot % W;
        cnt = bar + (size_t)(kPcgCounterBase + b) * kBarrierSlotWords;
    }
    const int n = a.n;
    const int chunk = (n + kPcgChunks - 1) / kPcgChunks;
    const double *val = a.sys + (size_t)b * a.nnz;
    const double *rhs = a.rhs + (size_t)b * n, *diag = a.diag + (size_t)b * n;
    double *x = a.x + (size_t)b * n, *r = a.r + (size_t)b * n, *z = a.z + (size_t)b * n,
           *p = a.p + (size_t)b * n, *q = a.q + (size_t)b * n;
    double *part = a.partial + (size_t)b * kPcgChunks * 4;
    unsigned target = 0;
    auto rows = [&](int c, int &lo, int &hi) {
        lo = min(n, c * chunk);
        hi = min(n, lo + chunk);
    };
    // after this workgroup's chunks wrote partial slots k, k+1: every chunk's
    // partials summed in chunk order
    auto gather2 = [&](int k, double &ta, double &tb) -> bool {
        if constexpr (GROUP) {
            target += (unsigned)W;
            if (!group_barrier(cnt, target, timeout))
                return false;
        } else {
            __syncthreads();
        }
        double x0 = 0.0, x1 = 0.0;
        for (int c = 0; c < kPcgChunks; ++c) {
            if constexpr (GROUP) {
                x0 += load_l2(part + 4 * c + k);
                x1 += load_l2(part + 4 * c + k + 1);
            } else {
                x0 += part[4 * c + k];
                x1 += part[4 * c + k + 1];
            }
        }
        ta = x0;
        tb = x1;
        return true;
    };
    auto publish2 = [&](int c, int k, double sa, double sb) {
        double ta, tb;
        wg_sum2(sa, sb, ta, tb);
        if (threadIdx.x == 0) {
            part[4 * c + k] = ta;
            part[4 * c + k + 1] = tb;
        }
    };
    auto sync_p = [&]() -> bool {
        if constexpr (GROUP) {
            target += (unsigned)W;
            return group_barrier(cnt, target, timeout);
        } else {
            __syncthreads();
            return true;
        }
    };
    auto load_p = [&](int j) -> double {
        if constexpr (GROUP)
            return load_l2(p + j);
        else
            return p[j];
    };
    // x = 0, r = rhs - A 0, z = r / diag, p = z, rho = r.z (rotation_feature.py:217-229)
    for (int c = w; c < kPcgChunks; c += W) {
        int lo, hi;
        rows(c, lo, hi);
        double s_bb = 0.0, s_rz = 0.0;
        for (int i = lo + threadIdx.x; i < hi; i += kPcgThreads) {
            const double ri = rhs[i], zi = ri / diag[i];
            x[i] = 0.0;
            r[i] = ri;
            z[i] = zi;
            p[i] = zi;
            s_bb += ri * ri;
            s_rz += ri * zi;
        }
        publish2(c, 0, s_bb, s_rz);
    }
    double bb, rho;
    if (!gather2(0, bb, rho))
        return;
    const double rhs_norm = fmax(sqrt(bb), 2.220446049250313e-16);
    double rel = sqrt(bb) / rhs_norm;
    int iterations = 0;
    if (rel > a.tol) {
        for (int it = 1; it <= a.max_iter; ++it) {
            // q = A p, p.q
            for (int c = w; c < kPcgChunks; c += W) {
                int lo, hi;
                rows(c, lo, hi);
                double s_pq = 0.0;
                for (int i = lo + threadIdx.x; i < hi; i += kPcgThreads) {
                    double s = 0.0;
                    for (int k = a.row_off[i]; k < a.row_off[i + 1]; ++k)
                        s += val[k] * load_p(a.cols[k]);
                    q[i] = s;
                    s_pq += load_p(i) * s;
                }
                publish2(c, 2, s_pq, 0.0);
            }
            double pq, unused;
            if (!gather2(2, pq, unused))
                return;
            const double step = rho / pq;
            // x += step p, r -= step q, r.r; z = r / diag, r.z
            for (int c = w; c < kPcgChunks; c += W) {
                int lo, hi;
                rows(c, lo, hi);
                double s_rr = 0.0, s_rz = 0.0;
                for (int i = lo + threadIdx.x; i < hi; i += kPcgThreads) {
                    const double pi = load_p(i);
                    x[i] = x[i] + step * pi;
                    const double ri = r[i] - step * q[i];
                    r[i] = ri;
                    const double zi = ri / diag[i];
                    z[i] = zi;
                    s_rr += ri * ri;
                    s_rz += ri * zi;
                }
                publish2(c, 0, s_rr, s_rz);
            }
            double rr, rz;
            if (!gather2(0, rr, rz))
                return;
            rel = sqrt(rr) / rhs_norm;
            iterations = it;
            if (rel <= a.tol)
                break;
            const double beta = rz / rho;
            rho = rz;
            for (int c = w; c < kPcgChunks; c += W) {
                int lo, hi;
                rows(c, lo, hi);
                for (int i = lo + threadIdx.x; i < hi; i += kPcgThreads)
                    p[i] = z[i] + beta * load_p(i);
            }
            if (!sync_p())
                return;
        }
    }
    if (w == 0 && threadIdx.x == 0) {
        a.iterations[b] = iterations;
        a.relres[b] = rel;
    }
}

}  // namespace

size_t pcg_workspace_bytes(int n, int batch)
{
    return sizeof(double) * ((size_t)5 * batch * n + (size_t)batch * kPcgChunks * 4) + 256;
}

int depth_pcg(const DepthArgs &d, double *ws, int max_iter, double tol, bool allow_xcd_local,
              int force_xcd, int *iterations, double *relres, SorRun *run, hipStream_t s)
{
    if (run)
        *run = SorRun{};
    const int B = d.batch, n = d.n;
    PcgArgs a{};
    a.row_off = d.row_off;
    a.cols = d.cols;
    a.sys = d.sys;
    a.rhs = d.rhs;
    a.diag = d.diag;
    a.n = n;
    a.nnz = d.nnz;
    a.batch = B;
    a.x = d.x;
    a.r = ws;
    a.z = a.r + (size_t)B * n;
    a.p = a.z + (size_t)B * n;
    a.q = a.p + (size_t)B * n;
    a.partial = a.q + (size_t)B * n;
    a.max_iter = max_iter;
    a.tol = tol;
    a.iterations = iterations;
    a.relres = relres;
    CSG_TRY(clear_barrier(d.barrier, s));
    int *timeout = reinterpret_cast<int *>(d.barrier + (size_t)(kBarrierSlots - 1) * kBarrierSlotWords);
    int dev = 0, cus = 0, per_cu = 0;
    CSG_HIP_CHECK(hipGetDevice(&dev));
    CSG_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CSG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg<true>, kPcgThreads,
                                                               0));
    // one row chunk per CU of the XCD, shared by the XCD's meshes; the counters
    // stay below the placement verdict (slot kBarrierSlots - 2) and the timeout
    bool local = allow_xcd_local && cus % kXcds == 0 && per_cu > 0 &&
                 B <= kBarrierSlots - 2 - kPcgCounterBase;
    PcgPlan plan{};
    plan.force_xcd = force_xcd;
    int most = 0;
    for (int x = 0; x < kXcds && local; ++x) {
        const int meshes = B > x ? (B - x + kXcds - 1) / kXcds : 0;
        if (meshes == 0)
            continue;
        int per = (cus / kXcds) / meshes;
        per = per < 1 ? 1 : (per > kPcgChunks ? kPcgChunks : per);
        plan.per_mesh[x] = per;
        plan.blocks[x] = per * meshes;
        local = plan.blocks[x] <= (cus / kXcds) * per_cu;
        most = plan.blocks[x] > most ? plan.blocks[x] : most;
    }
    // timed without algorithmic bytes: the iteration count is known only after the solve
    TimedLaunch timer(kTimedPcg, s, 0.0);
    if (local) {
        hipExtLaunchKernelGGL(k_pcg<true>, dim3((most + kXcdSpare) * kXcds), dim3(kPcgThreads),
                              0, s, timer.start(), timer.stop(), 0, a, plan, d.barrier, timeout);
    } else {
        hipExtLaunchKernelGGL(k_pcg<false>, dim3(B), dim3(kPcgThreads), 0, s, timer.start(),
                              timer.stop(), 0, a, plan, d.barrier, timeout);
    }
    CSG_LAUNCH_CHECK();
    timer.done();
    if (run && local) {
        run->xcd_local = true;
        for (int x = 0; x < kXcds; ++x)
            run->plan.blocks[x] = plan.blocks[x];
    }
    return kOk;
}

}  // namespace csg
