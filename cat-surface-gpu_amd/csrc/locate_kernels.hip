// Point location on triangulated spheres for gfx950: the uniform-grid
// triangle index, the CAT rotation cost (rotate → locate → barycentric →
// float64 SSD, batched over candidate rotations), and the exact
// closest-triangle query used by the stencil builder.
//
// Reference semantics: src/cat_surface_gpu/surface_rotation.py:57-120
// (grid index), :223-313 (rotation matrices), :334-546 (locator and cost);
// native/cat_surface_stencil_builder.c:107-185 (stencil queries).
#include <hipcub/hipcub.hpp>

#include "../../include/csg.h"
#include "locate_kernels.h"

namespace csg {

// ------------------------------------------------------------ grid index

// Per-face cell box (bbox ± margin cells, clipped) and its cell count.
// surface_rotation.py:80-87
template <typename P>
__global__ void __launch_bounds__(kBlock)
k_face_boxes(const P *__restrict__ pts, const int *__restrict__ faces, int nf, int g, int margin,
             int *__restrict__ box, int *__restrict__ counts)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) {
        if (f == nf)
            counts[nf] = 0;  // the exclusive scan reads nf + 1 counts
        return;
    }
    double cell = 2.0 / (double)g;
    int lo[3], hi[3], total = 1;
    for (int d = 0; d < 3; ++d) {
        double a = (double)pts[3 * faces[3 * f] + d], b = (double)pts[3 * faces[3 * f + 1] + d],
               c = (double)pts[3 * faces[3 * f + 2] + d];
        double mn = fmin(fmin(a, b), c), mx = fmax(fmax(a, b), c);
        long l = (long)floor((mn + 1.0) / cell) - margin;
        long h = (long)floor((mx + 1.0) / cell) + margin;
        l = l < 0 ? 0 : (l > g - 1 ? g - 1 : l);
        h = h < 0 ? 0 : (h > g - 1 ? g - 1 : h);
        lo[d] = (int)l;
        hi[d] = (int)h;
        total *= (hi[d] >= lo[d]) ? (hi[d] - lo[d] + 1) : 0;
    }
    for (int d = 0; d < 3; ++d) {
        box[6 * f + d] = lo[d];
        box[6 * f + 3 + d] = hi[d];
    }
    counts[f] = total;
}

// (cell, face) pairs, cells of one face in z, y, x order.
__global__ void __launch_bounds__(kBlock)
k_emit_pairs(const int *__restrict__ box, const int *__restrict__ offsets, int nf, int g,
             int *__restrict__ keys, int *__restrict__ vals)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    int o = offsets[f];
    const int *b = box + 6 * f;
    for (int z = b[2]; z <= b[5]; ++z)
        for (int y = b[1]; y <= b[4]; ++y)
            for (int x = b[0]; x <= b[3]; ++x) {
                keys[o] = x + g * (y + g * z);
                vals[o] = f;
                ++o;
            }
}

// Capacity-bounded emission: pairs past `capacity` are dropped (the caller
// sees the required count in *required and rebuilds with more room).
__global__ void __launch_bounds__(kBlock)
k_emit_pairs_capped(const int *__restrict__ box, const int *__restrict__ offsets, int nf, int g,
                    int capacity, int *__restrict__ keys, int *__restrict__ vals)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    int o = offsets[f];
    const int *b = box + 6 * f;
    for (int z = b[2]; z <= b[5]; ++z)
        for (int y = b[1]; y <= b[4]; ++y)
            for (int x = b[0]; x <= b[3]; ++x) {
                if (o >= capacity)
                    return;
                keys[o] = x + g * (y + g * z);
                vals[o] = f;
                ++o;
            }
}

// Slots [required, capacity) get the key `sentinel` (= the cell count), which
// sorts after every cell; *required = the pair count the index needs.
__global__ void __launch_bounds__(kBlock)
k_pad_keys(const int *__restrict__ total, int capacity, int sentinel, int *__restrict__ keys,
           int *__restrict__ required)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = *total;
    if (i == 0)
        *required = n;
    if (i < capacity && i >= n)
        keys[i] = sentinel;
}

// cell_off[c] = first position of key >= c in the sorted keys.
__global__ void __launch_bounds__(kBlock)
k_cell_offsets(const int *__restrict__ keys, int npairs, int ncells, int *__restrict__ off)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells)
        return;
    int lo = 0, hi = npairs;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    off[c] = lo;
}

template <typename P>
int grid_index_count(const P *pts, const int *faces, int nf, int g, int margin, int *box,
                     int *counts, int *offsets, int *total, void *tmp, size_t tmp_bytes,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_face_boxes<P>, dim3(grid_for((size_t)nf + 1)), dim3(kBlock), 0, s, pts,
                       faces, nf, g, margin, box, counts);
    CSG_LAUNCH_CHECK();
    size_t need = 0;
    CSG_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, counts, offsets, nf + 1, s));
    if (need > tmp_bytes)
        return fail(kInvalidArgument, "grid index scan scratch too small");
    CSG_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, need, counts, offsets, nf + 1, s));
    CSG_HIP_CHECK(hipMemcpyAsync(total, offsets + nf, sizeof(int), hipMemcpyDeviceToHost, s));
    CSG_HIP_CHECK(hipStreamSynchronize(s));
    return kOk;
}

size_t grid_sort_scratch_bytes(int npairs)
{
    // size queries only (no device work); a failed query leaves the size at 0
    // and the sort itself reports the error
    size_t need = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, need, (int *)nullptr, (int *)nullptr,
                                             (int *)nullptr, (int *)nullptr, npairs, 0, 32);
    size_t scan = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (int *)nullptr, (int *)nullptr,
                                           1 << 22);
    return (need > scan ? need : scan) + 1024;
}

int grid_index_fill(const int *box, const int *offsets, int nf, int g, int npairs, int *keys_tmp,
                    int *vals_tmp, int *keys, int *vals, int *cell_off, void *tmp,
                    size_t tmp_bytes, hipStream_t s)
{
    hipLaunchKernelGGL(k_emit_pairs, dim3(grid_for(nf)), dim3(kBlock), 0, s, box, offsets, nf, g,
                       keys_tmp, vals_tmp);
    CSG_LAUNCH_CHECK();
    int bits = 1;
    while ((1LL << bits) < (long long)g * g * g)
        ++bits;
    size_t need = 0;
    CSG_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys_tmp, keys, vals_tmp, vals,
                                                     npairs, 0, bits, s));
    if (need > tmp_bytes)
        return fail(kInvalidArgument, "grid index sort scratch too small");
    CSG_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, need, keys_tmp, keys, vals_tmp, vals,
                                                     npairs, 0, bits, s));
    int ncells = g * g * g;
    hipLaunchKernelGGL(k_cell_offsets, dim3(grid_for((size_t)ncells + 1)), dim3(kBlock), 0, s,
                       keys, npairs, ncells, cell_off);
    CSG_LAUNCH_CHECK();
    return kOk;
}

// ---- the index without library calls (so that it can be captured in a graph:
// the library's scan and sort record memsets): the face-order scan and the
// stable LSD radix sort of the (cell, face) pairs by cell are written here.

__global__ void __launch_bounds__(kBlock) k_fill_int(int *__restrict__ x, int v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        x[i] = v;
}

constexpr int kScanItems = 8;  // per thread; a block scans kBlock * kScanItems values

// Exclusive scan, phase 1: each block's values scanned locally, its total in
// sums[block].
__global__ void __launch_bounds__(kBlock)
k_scan_local(const int *__restrict__ in, int n, int *__restrict__ out, int *__restrict__ sums)
{
    __shared__ int part[kBlock];
    const int base = blockIdx.x * kBlock * kScanItems + threadIdx.x * kScanItems;
    int v[kScanItems], run = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        run += v[k];
    }
    part[threadIdx.x] = run;
    __syncthreads();
    // Hillis-Steele over the threads' totals
    for (int d = 1; d < kBlock; d <<= 1) {
        const int t = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    int acc = part[threadIdx.x] - run;  // exclusive prefix of this thread
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n)
            out[base + k] = acc;
        acc += v[k];
    }
    if (threadIdx.x == kBlock - 1)
        sums[blockIdx.x] = part[kBlock - 1];
}

// Phase 2: the block totals scanned in place by one block, in order.
__global__ void __launch_bounds__(kBlock) k_scan_sums(int *__restrict__ sums, int nblocks)
{
    __shared__ int part[kBlock];
    __shared__ int carry;
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nblocks; b0 += kBlock) {
        const int i = b0 + threadIdx.x;
        const int v = i < nblocks ? sums[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < kBlock; d <<= 1) {
            const int t = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
            __syncthreads();
            part[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nblocks)
            sums[i] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kBlock - 1)
            carry += part[kBlock - 1];
        __syncthreads();
    }
}

// Phase 3: each block's offset added.
__global__ void __launch_bounds__(kBlock)
k_scan_add(int *__restrict__ out, int n, const int *__restrict__ sums)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] += sums[i / (kBlock * kScanItems)];
}

int scan_blocks(int n) { return (n + kBlock * kScanItems - 1) / (kBlock * kScanItems); }

// out[0..n) = exclusive prefix sums of in[0..n) (sums: scan_blocks(n) ints).
int exclusive_scan_int(const int *in, int n, int *out, int *sums, hipStream_t s)
{
    const int nb = scan_blocks(n);
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kBlock), 0, s, in, n, out, sums);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, sums, nb);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_add, dim3(grid_for(n)), dim3(kBlock), 0, s, out, n, sums);
    CSG_LAUNCH_CHECK();
    return kOk;
}

// Stable LSD radix sort of (key, value) int pairs, 8 bits per pass.  A block
// takes a tile of kRadixTile consecutive pairs in rounds of kRadixThreads
// (pair base + r * kRadixThreads + t), so (round, wave, lane) order is the
// input order; the digit offsets are scanned digit-major, block-minor.
constexpr int kRadixThreads = 256;
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixWaves = kRadixThreads / 64;

// hist[d * nblocks + b] = pairs of block b whose digit is d.
__global__ void __launch_bounds__(kRadixThreads)
k_radix_hist(const int *__restrict__ keys, int n, int shift, int nblocks, int *__restrict__ hist)
{
    __shared__ int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRadixTile;
    for (int r = 0; r < kRadixItems; ++r) {
        const int i = base + r * kRadixThreads + threadIdx.x;
        if (i < n)
            atomicAdd(&cnt[((unsigned)keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Each pair goes to offs[d * nblocks + b] + (pairs of digit d before it in the
// block): a wave ranks its lanes among the lanes of the same digit from eight
// ballots, the waves' counts per digit are added in wave order.
__global__ void __launch_bounds__(kRadixThreads)
k_radix_scatter(const int *__restrict__ keys_in, const int *__restrict__ vals_in, int n,
                int shift, int nblocks, const int *__restrict__ offs, int *__restrict__ keys_out,
                int *__restrict__ vals_out)
{
    __shared__ int next[256];
    __shared__ int wcnt[kRadixWaves][256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    next[t] = offs[t * nblocks + blockIdx.x];
    const unsigned long long below = (1ull << lane) - 1ull;
    const int base = blockIdx.x * kRadixTile;
    for (int r = 0; r < kRadixItems; ++r) {
#pragma unroll
        for (int q = 0; q < kRadixWaves; ++q)
            wcnt[q][t] = 0;
        __syncthreads();
        const int i = base + r * kRadixThreads + t;
        const bool valid = i < n;
        int key = 0, val = 0, d = 0;
        if (valid) {
            key = keys_in[i];
            val = vals_in[i];
            d = (int)(((unsigned)key >> shift) & 255u);
        }
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool set = ((d >> bit) & 1) != 0;
            const unsigned long long m = __ballot(set);
            peers &= set ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0)
            wcnt[w][d] = __popcll(peers);
        __syncthreads();
        // thread t: the waves' prefix counts of digit t, and the round's total
        int acc = 0;
        int pre[kRadixWaves];
#pragma unroll
        for (int q = 0; q < kRadixWaves; ++q) {
            pre[q] = acc;
            acc += wcnt[q][t];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kRadixWaves; ++q)
            wcnt[q][t] = pre[q];
        __syncthreads();
        if (valid) {
            const int pos = next[d] + wcnt[w][d] + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        next[t] += acc;
    }
}

int radix_blocks(int n) { return (n + kRadixTile - 1) / kRadixTile; }

size_t radix_scratch_bytes(int n)
{
    const size_t hist = (size_t)256 * radix_blocks(n);
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return 2 * al(sizeof(int) * hist) + al(sizeof(int) * ((size_t)scan_blocks((int)hist) + 1));
}

// Sorts n pairs by the low `bits` bits of the keys; (ka, va) hold the input and
// receive the result, (kb, vb) are the second buffers (n ints each).
int radix_sort_pairs(int *ka, int *va, int *kb, int *vb, int n, int bits, void *tmp,
                     size_t tmp_bytes, hipStream_t s)
{
    if (n <= 0)
        return kOk;
    if (tmp_bytes < radix_scratch_bytes(n))
        return fail(kInvalidArgument, "radix sort scratch too small");
    const int nb = radix_blocks(n);
    const size_t hist_n = (size_t)256 * nb;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char *t = (char *)tmp;
    int *hist = (int *)t;
    int *offs = (int *)(t + al(sizeof(int) * hist_n));
    int *sums = (int *)(t + 2 * al(sizeof(int) * hist_n));
    int *src_k = ka, *src_v = va, *dst_k = kb, *dst_v = vb;
    const int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kRadixThreads), 0, s, src_k, n, 8 * p,
                           nb, hist);
        CSG_LAUNCH_CHECK();
        CSG_TRY(exclusive_scan_int(hist, (int)hist_n, offs, sums, s));
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kRadixThreads), 0, s, src_k, src_v,
                           n, 8 * p, nb, offs, dst_k, dst_v);
        CSG_LAUNCH_CHECK();
        int *tk = src_k, *tv = src_v;
        src_k = dst_k;
        src_v = dst_v;
        dst_k = tk;
        dst_v = tv;
    }
    if (src_k != ka) {
        CSG_HIP_CHECK(hipMemcpyAsync(ka, src_k, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice,
                                     s));
        CSG_HIP_CHECK(hipMemcpyAsync(va, src_v, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice,
                                     s));
    }
    return kOk;
}

size_t grid_build_scratch_bytes(int nf, int capacity)
{
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return 3 * al(sizeof(int) * (size_t)capacity) +                  // keys, values, keys 2
           al(sizeof(int) * ((size_t)scan_blocks(nf + 1) + 1)) +    // face scan sums
           al(radix_scratch_bytes(capacity));
}

template <typename P>
int grid_index_build(const P *pts, const int *faces, int nf, int g, int margin, int capacity,
                     int *box, int *counts, int *offsets, int *cell_faces, int *cell_off,
                     int *required, void *tmp, size_t tmp_bytes, hipStream_t s)
{
    if (tmp_bytes < grid_build_scratch_bytes(nf, capacity))
        return fail(kInvalidArgument, "grid index build scratch too small");
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t a = al(sizeof(int) * (size_t)capacity);
    char *t = (char *)tmp;
    int *keys = (int *)t, *keys2 = (int *)(t + a), *vals2 = (int *)(t + 2 * a);
    int *sums = (int *)(t + 3 * a);
    char *rtmp = t + 3 * a + al(sizeof(int) * ((size_t)scan_blocks(nf + 1) + 1));
    const size_t rbytes = tmp_bytes - (size_t)(rtmp - t);
    hipLaunchKernelGGL(k_face_boxes<P>, dim3(grid_for((size_t)nf + 1)), dim3(kBlock), 0, s, pts,
                       faces, nf, g, margin, box, counts);
    CSG_LAUNCH_CHECK();
    CSG_TRY(exclusive_scan_int(counts, nf + 1, offsets, sums, s));
    const int ncells = g * g * g;
    hipLaunchKernelGGL(k_pad_keys, dim3(grid_for(capacity)), dim3(kBlock), 0, s, offsets + nf,
                       capacity, ncells, keys, required);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_emit_pairs_capped, dim3(grid_for(nf)), dim3(kBlock), 0, s, box, offsets,
                       nf, g, capacity, keys, cell_faces);
    CSG_LAUNCH_CHECK();
    int bits = 1;
    while ((1LL << bits) <= (long long)ncells)  // keys run to the sentinel ncells
        ++bits;
    CSG_TRY(radix_sort_pairs(keys, cell_faces, keys2, vals2, capacity, bits, rtmp, rbytes, s));
    hipLaunchKernelGGL(k_cell_offsets, dim3(grid_for((size_t)ncells + 1)), dim3(kBlock), 0, s,
                       keys, capacity, ncells, cell_off);
    CSG_LAUNCH_CHECK();
    return kOk;
}

template int grid_index_build<float>(const float *, const int *, int, int, int, int, int *, int *,
                                     int *, int *, int *, int *, void *, size_t, hipStream_t);
template int grid_index_build<double>(const double *, const int *, int, int, int, int, int *,
                                      int *, int *, int *, int *, int *, void *, size_t,
                                      hipStream_t);
template int grid_index_count<float>(const float *, const int *, int, int, int, int *, int *,
                                     int *, int *, void *, size_t, hipStream_t);
template int grid_index_count<double>(const double *, const int *, int, int, int, int *, int *,
                                      int *, int *, void *, size_t, hipStream_t);

// ------------------------------------------------------------ top k

// (cost, index) order: ascending cost, NaN after every number, ties by index
// (the order of a stable ascending sort).
__device__ __forceinline__ bool cost_before(double a, int ia, double b, int ib)
{
    const bool na = a != a, nb = b != b;
    if (na != nb)
        return nb;
    if (!na && a != b)
        return a < b;
    return ia < ib;
}

// order[0..k) = the indices of the k first costs in that order: k rounds of a
// block-wide minimum over the costs not yet chosen.  One block.
__global__ void __launch_bounds__(1024)
k_top_k(const double *__restrict__ costs, int n, int k, int *__restrict__ order)
{
    __shared__ double bc[1024];
    __shared__ int bi[1024];
    __shared__ int chosen[kTopKMax];
    const int t = threadIdx.x;
    for (int r = 0; r < k; ++r) {
        double best = 0.0;
        int bidx = -1;
        for (int i = t; i < n; i += blockDim.x) {
            bool taken = false;
            for (int q = 0; q < r; ++q)
                taken |= chosen[q] == i;
            if (taken)
                continue;
            const double c = costs[i];
            if (bidx < 0 || cost_before(c, i, best, bidx)) {
                best = c;
                bidx = i;
            }
        }
        bc[t] = best;
        bi[t] = bidx;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if (t < w) {
                const int oi = bi[t + w];
                if (oi >= 0 && (bi[t] < 0 || cost_before(bc[t + w], oi, bc[t], bi[t]))) {
                    bc[t] = bc[t + w];
                    bi[t] = oi;
                }
            }
            __syncthreads();
        }
        if (t == 0) {
            chosen[r] = bi[0];
            order[r] = bi[0];
        }
        __syncthreads();
    }
}

int launch_top_k(const double *costs, int n, int k, int *order, hipStream_t s)
{
    hipLaunchKernelGGL(k_top_k, dim3(1), dim3(1024), 0, s, costs, n, k, order);
    CSG_LAUNCH_CHECK();
    return kOk;
}

// ------------------------------------------------------------ rotation cost

__device__ __forceinline__ float dot3(const float *a, const float *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void cross3(const float *a, const float *b, float *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ void seg_dist(const float *q, const float *a, const float *b,
                                         float *dist, float *pt)
{
    float dir[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float off[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]};
    float den = dot3(dir, dir);
    float al = den != 0.0f ? dot3(off, dir) / den : 0.0f;
    al = fminf(fmaxf(al, 0.0f), 1.0f);
    for (int d = 0; d < 3; ++d)
        pt[d] = a[d] + al * dir[d];
    float e0 = q[0] - pt[0], e1 = q[1] - pt[1], e2 = q[2] - pt[2];
    *dist = e0 * e0 + e1 * e1 + e2 * e2;
}

// Rotation matrix of (alpha, beta, gamma) in float32, Rz (Ry Rx) assembled
// column-major and read row-major.  surface_rotation.py:223-292
__device__ void rotation_matrix_f32(double a0, double a1, double a2, float *m)
{
    float al = (float)a0, be = (float)a1, ga = (float)a2;
    float ca = cosf(al), sa = sinf(al), cb = cosf(be), sb = sinf(be), cg = cosf(ga),
          sg = sinf(ga);
    float rx[9] = {1, 0, 0, 0, ca, sa, 0, -sa, ca};
    float ry[9] = {cb, 0, sb, 0, 1, 0, -sb, 0, cb};
    float rz[9] = {cg, sg, 0, -sg, cg, 0, 0, 0, 1};
    float first[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            first[r + 3 * c] = ry[r] * rx[3 * c] + ry[r + 3] * rx[1 + 3 * c] +
                               ry[r + 6] * rx[2 + 3 * c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            m[r + 3 * c] = rz[r] * first[3 * c] + rz[r + 3] * first[1 + 3 * c] +
                           rz[r + 6] * first[2 + 3 * c];
}

// Best candidate of one lane: candidates c = lane, lane + lanes, ... of q's
// cell, in increasing order; the first minimum of the squared distance wins.
// Returns the candidate position (or a value >= ncand if none); *face and
// bp[] receive the face and its closest point.  surface_rotation.py:354-447
struct LocateBest {
    float dist;
    int pos, face;
    float bp[3];
};

// Squared distance from q to one face and the closest point (the reference's
// projection test with its 1e-3 barycentric slack, else the nearest vertex's
// two edges).
__device__ __forceinline__ float face_distance(const float *q, const float *__restrict__ tri,
                                               int face, float *cpt)
{
    const float *t = tri + 9 * (size_t)face;
    float p0[3] = {t[0], t[1], t[2]}, p1[3] = {t[3], t[4], t[5]}, p2[3] = {t[6], t[7], t[8]};
    float ea[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    float eb[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    float nrm[3];
    cross3(ea, eb, nrm);
    float nsq = dot3(nrm, nrm);
    float off[3] = {p0[0] - q[0], p0[1] - q[1], p0[2] - q[2]};
    float pt = nsq != 0.0f ? dot3(off, nrm) / nsq : 0.0f;
    float proj[3] = {q[0] + pt * nrm[0], q[1] + pt * nrm[1], q[2] + pt * nrm[2]};
    float po[3] = {proj[0] - p0[0], proj[1] - p0[1], proj[2] - p0[2]};
    float xx = dot3(ea, ea), xy = dot3(ea, eb), xv = dot3(ea, po), yy = dot3(eb, eb),
          yv = dot3(eb, po);
    float den = xx * yy - xy * xy;
    float xp = den != 0.0f ? (xv * yy - yv * xy) / den : 0.0f;
    float yp = den != 0.0f ? (yv * xx - xv * xy) / den : 0.0f;
    bool inside = den != 0.0f && xp >= -1.0e-3f && xp <= 1.0f + 1.0e-3f && yp >= -1.0e-3f &&
                  yp <= 1.0f + 1.0e-3f && xp + yp >= -1.0e-3f && xp + yp <= 1.0f + 1.0e-3f;
    if (inside) {
        cpt[0] = proj[0];
        cpt[1] = proj[1];
        cpt[2] = proj[2];
        return pt * pt * nsq;
    }
    // nearest vertex (first minimum), then the two edges through it; selects
    // instead of indexed pointer tables keep this in registers
    float e0 = q[0] - p0[0], e1 = q[1] - p0[1], e2 = q[2] - p0[2];
    const float dv0 = e0 * e0 + e1 * e1 + e2 * e2;
    e0 = q[0] - p1[0], e1 = q[1] - p1[1], e2 = q[2] - p1[2];
    const float dv1 = e0 * e0 + e1 * e1 + e2 * e2;
    e0 = q[0] - p2[0], e1 = q[1] - p2[1], e2 = q[2] - p2[2];
    const float dv2 = e0 * e0 + e1 * e1 + e2 * e2;
    int cv = 0;
    float dbest = dv0;
    if (dv1 < dbest) {
        cv = 1;
        dbest = dv1;
    }
    if (dv2 < dbest)
        cv = 2;
    float prev[3], cur[3], next[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        cur[d] = cv == 0 ? p0[d] : (cv == 1 ? p1[d] : p2[d]);
        prev[d] = cv == 0 ? p2[d] : (cv == 1 ? p0[d] : p1[d]);
        next[d] = cv == 0 ? p1[d] : (cv == 1 ? p2[d] : p0[d]);
    }
    float d1, d2, q1[3], q2[3];
    seg_dist(q, prev, cur, &d1, q1);
    seg_dist(q, cur, next, &d2, q2);
    if (d1 < d2) {
        cpt[0] = q1[0];
        cpt[1] = q1[1];
        cpt[2] = q1[2];
        return d1;
    }
    cpt[0] = q2[0];
    cpt[1] = q2[1];
    cpt[2] = q2[2];
    return d2;
}

// Ties are broken by the smaller face index, which is the sequential scan's
// first minimum (a cell lists its faces in increasing order).
//
// With face spheres (sph: centre, radius enlarged by 2 % and 1e-6, see
// k_face_spheres_f32), a candidate whose sphere lies farther than
// sqrt(best)·(1 + 1e-4) + 1e-6 from q is skipped.  Its distance would exceed
// the lane's best: outside the triangle the distance is to a point of the
// triangle (>= |q - c| - r); inside, the in-plane point lies within 1.004 r
// of c (the 1e-3 barycentric slack), so the plane distance is >= |q - c| -
// 1.004 r.  Skipping it changes no result.
//
// `start` (a face known to be in this cell's list, at position start_pos, or
// -1) is evaluated first: it is one of the candidates, and a close start lets
// the sphere test skip most of the others.
__device__ __forceinline__ LocateBest scan_candidates(const float *q, const float *__restrict__ tri,
                                                     const int *__restrict__ cand, int ncand,
                                                     int lane, int lanes,
                                                     const float4 *__restrict__ sph,
                                                     int start = -1, int start_pos = 0x7fffffff)
{
    LocateBest r{INFINITY, 0x7fffffff, -1, {0.0f, 0.0f, 0.0f}};
    float reach = INFINITY;
    if (start >= 0) {
        r.dist = face_distance(q, tri, start, r.bp);
        r.face = start;
        r.pos = start_pos;
        reach = sqrtf(r.dist) * 1.0001f + 1e-6f;
    }
    // kScanBatch candidates per lane at a time: their indices, then their
    // spheres, are fetched together (one memory round trip each instead of two
    // per candidate); the order of evaluation does not matter (ties go to the
    // smaller face index)
    constexpr int kScanBatch = 4;
    for (int c0 = lane; c0 < ncand; c0 += lanes * kScanBatch) {
        int fc[kScanBatch];
#pragma unroll
        for (int u = 0; u < kScanBatch; ++u) {
            const int c = c0 + u * lanes;
            fc[u] = c < ncand ? cand[c] : -1;
        }
        float4 bs[kScanBatch];
#pragma unroll
        for (int u = 0; u < kScanBatch; ++u)
            bs[u] = (sph && fc[u] >= 0) ? sph[fc[u]] : make_float4(0.0f, 0.0f, 0.0f, INFINITY);
#pragma unroll
        for (int u = 0; u < kScanBatch; ++u) {
            const int face = fc[u];
            if (face < 0)
                continue;
            if (sph) {
                const float dx = q[0] - bs[u].x, dy = q[1] - bs[u].y, dz = q[2] - bs[u].z;
                const float lim = bs[u].w + reach;
                if (dx * dx + dy * dy + dz * dz > lim * lim)
                    continue;
            }
            float cpt[3];
            const float dist = face_distance(q, tri, face, cpt);
            if (dist < r.dist || (dist == r.dist && face < r.face)) {
                r.dist = dist;
                reach = sqrtf(dist) * 1.0001f + 1e-6f;
                r.pos = c0 + u * lanes;
                r.face = face;
                r.bp[0] = cpt[0];
                r.bp[1] = cpt[1];
                r.bp[2] = cpt[2];
            }
        }
    }
    return r;
}

// Combine the lanes' bests of one query (lanes are consecutive lanes of a
// wave, `lanes` a power of two): smaller distance first, then the earlier
// candidate, which is the sequential loop's first minimum.
__device__ __forceinline__ LocateBest reduce_lanes(LocateBest r, int lanes)
{
    for (int m = 1; m < lanes; m <<= 1) {
        LocateBest o;
        o.dist = __shfl_xor(r.dist, m);
        o.pos = __shfl_xor(r.pos, m);
        o.face = __shfl_xor(r.face, m);
        o.bp[0] = __shfl_xor(r.bp[0], m);
        o.bp[1] = __shfl_xor(r.bp[1], m);
        o.bp[2] = __shfl_xor(r.bp[2], m);
        if (o.dist < r.dist || (o.dist == r.dist && (unsigned)o.face < (unsigned)r.face))
            r = o;
    }
    return r;
}

// Barycentric weights of the chosen point on face bf.
__device__ __forceinline__ void locate_weights(const float *__restrict__ tri, int bf,
                                               const float *bp, float *bary)
{
    const float *t = tri + 9 * (size_t)bf;
    // weight k uses the vertex order (k+1, k+2, k): offsets of the three vertices
    const int first[3] = {3, 6, 0}, second[3] = {6, 0, 3}, third[3] = {0, 3, 6};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float *fi = t + first[k], *se = t + second[k], *th = t + third[k];
        float po[3] = {bp[0] - fi[0], bp[1] - fi[1], bp[2] - fi[2]};
        float ho[3] = {se[0] - fi[0], se[1] - fi[1], se[2] - fi[2]};
        float up[3] = {th[0] - fi[0], th[1] - fi[1], th[2] - fi[2]};
        float nn[3], ve[3];
        cross3(ho, up, nn);
        cross3(nn, ho, ve);
        bary[k] = dot3(po, ve) / dot3(up, ve);
    }
}

__device__ __forceinline__ int cell_of(const float *p, int g)
{
    int c[3];
    for (int d = 0; d < 3; ++d) {
        int k = (int)floorf((p[d] + 1.0f) * ((float)g / 2.0f));
        c[d] = k < 0 ? 0 : (k > g - 1 ? g - 1 : k);
    }
    return c[0] + g * (c[1] + g * c[2]);
}

// kLocateLanes consecutive lanes share one point, each scanning every
// kLocateLanes-th candidate; blockIdx.y is the candidate rotation; per-block
// float64 partial SSD.  `uncovered` counts points whose cell has no triangle.

__global__ void __launch_bounds__(kBlock)
k_rotation_cost(const float *__restrict__ src_pts, const double *__restrict__ src_vals, int ns,
                const double *__restrict__ tgt_vals, const int *__restrict__ faces,
                const float *__restrict__ tri, const float4 *__restrict__ sph,
                const int *__restrict__ cell_off, const int *__restrict__ cell_faces, int g,
                const double *__restrict__ angles,
                double *__restrict__ partial, int nblocks, int *__restrict__ uncovered,
                int *__restrict__ loc_face, float *__restrict__ loc_bary)
{
    __shared__ double red[kBlock / 64];
    __shared__ float m[9];
    const int a = blockIdx.y;
    if (threadIdx.x == 0)
        rotation_matrix_f32(angles[3 * a], angles[3 * a + 1], angles[3 * a + 2], m);
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / kLocateLanes, lane = t % kLocateLanes;
    double e = 0.0;
    // every lane of a point takes part in the shuffles, including padding lanes
    float q[3] = {0.0f, 0.0f, 0.0f};
    int b = 0, nc = 0;
    if (i < ns) {
        const float *p = src_pts + 3 * i;
        for (int r = 0; r < 3; ++r)
            q[r] = p[0] * m[3 * r] + p[1] * m[3 * r + 1] + p[2] * m[3 * r + 2];
        int cell = cell_of(q, g);
        b = cell_off[cell];
        nc = cell_off[cell + 1] - b;
    }
    LocateBest best = scan_candidates(q, tri, cell_faces + b, nc, lane, kLocateLanes, sph);
    best = reduce_lanes(best, kLocateLanes);
    if (i < ns && lane == 0) {
        if (best.face < 0) {
            atomicAdd(uncovered, 1);
        } else {
            const int f = best.face;
            float w[3];
            locate_weights(tri, f, best.bp, w);
            double sv = tgt_vals[faces[3 * f]] * (double)w[0] +
                        tgt_vals[faces[3 * f + 1]] * (double)w[1] +
                        tgt_vals[faces[3 * f + 2]] * (double)w[2];
            double d = src_vals[i] - sv;
            e = d * d;
            if (loc_face) {
                loc_face[(size_t)a * ns + i] = f;
                for (int k = 0; k < 3; ++k)
                    loc_bary[((size_t)a * ns + i) * 3 + k] = w[k];
            }
        }
    }
    double sum = block_sum(e, red);
    if (threadIdx.x == 0)
        partial[(size_t)a * nblocks + blockIdx.x] = sum;
}

// The rotation SSD for candidate matrices (q = M p, M row-major [na][9]) and,
// with GRAD, its gradient and Gauss-Newton matrix with respect to a left
// axis-angle perturbation exp([δ]x)·M.  With r = s - t(q) and g the gradient
// of the located face's linear interpolant of the target values (the face held
// fixed), ∂r/∂δ = -(q x g); per point the kernel adds r², 2 r ∂r/∂δ and
// 2 ∂r/∂δ ∂r/∂δᵀ.  Location and sampling are k_rotation_cost's, so the cost
// equals k_rotation_cost's for the same float32 matrix.  Per block and
// candidate: 1 or 10 float64 partials (cost, gradient xyz, H xx yy zz xy xz yz).
template <bool GRAD>
__global__ void __launch_bounds__(kBlock)
k_rotation_cost_mats(const float *__restrict__ src_pts, const double *__restrict__ src_vals,
                     int ns, const double *__restrict__ tgt_vals, const int *__restrict__ faces,
                     const float *__restrict__ tri, const float4 *__restrict__ sph,
                     const int *__restrict__ cell_off, const int *__restrict__ cell_faces, int g,
                     const float *__restrict__ mats,
                     double *__restrict__ partial, int nblocks, int *__restrict__ uncovered,
                     int *__restrict__ hint, const double *__restrict__ face_vals)
{
    constexpr int kTerms = GRAD ? kRotationGradTerms : 1;
    __shared__ float m[9];
    const int a = blockIdx.y;
    if (threadIdx.x < 9)
        m[threadIdx.x] = mats[9 * (size_t)a + threadIdx.x];
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / kLocateLanes, lane = t % kLocateLanes;
    double term[kTerms];
#pragma unroll
    for (int k = 0; k < kTerms; ++k)
        term[k] = 0.0;
    float q[3] = {0.0f, 0.0f, 0.0f};
    int b = 0, nc = 0, cell = -1;
    if (i < ns) {
        const float *p = src_pts + 3 * i;
        for (int r = 0; r < 3; ++r)
            q[r] = p[0] * m[3 * r] + p[1] * m[3 * r + 1] + p[2] * m[3 * r + 2];
        cell = cell_of(q, g);
        b = cell_off[cell];
        nc = cell_off[cell + 1] - b;
    }
    // the face this point found for this candidate last time starts the scan
    // when the point is still in the same cell (the face is then entry `pos` of
    // the cell's list: one load, no search of the list)
    int start = -1, start_pos = 0x7fffffff;
    if (hint) {
        const int h = i < ns ? hint[(size_t)a * ns + i] : -1;
        if (h >= 0 && (h >> kHintPosBits) == cell && (h & ((1 << kHintPosBits) - 1)) < nc) {
            start_pos = h & ((1 << kHintPosBits) - 1);
            start = cell_faces[b + start_pos];
        }
    }
    LocateBest best = scan_candidates(q, tri, cell_faces + b, nc, lane, kLocateLanes, sph,
                                      start, start_pos);
    best = reduce_lanes(best, kLocateLanes);
    if (hint && i < ns && lane == 0)
        hint[(size_t)a * ns + i] =
            best.face >= 0 && best.pos < (1 << kHintPosBits) && cell < (1 << (31 - kHintPosBits))
                ? (cell << kHintPosBits) | best.pos
                : -1;
    if (i < ns && lane == 0) {
        if (best.face < 0) {
            atomicAdd(uncovered, 1);
        } else {
            const int f = best.face;
            float w[3];
            locate_weights(tri, f, best.bp, w);
            double t0, t1, t2;
            if (face_vals) {
                t0 = face_vals[3 * (size_t)f];
                t1 = face_vals[3 * (size_t)f + 1];
                t2 = face_vals[3 * (size_t)f + 2];
            } else {
                t0 = tgt_vals[faces[3 * f]];
                t1 = tgt_vals[faces[3 * f + 1]];
                t2 = tgt_vals[faces[3 * f + 2]];
            }
            const double d = src_vals[i] - (t0 * (double)w[0] + t1 * (double)w[1] +
                                            t2 * (double)w[2]);
            term[0] = d * d;
            if constexpr (GRAD) {
                const float *tv = tri + 9 * (size_t)f;
                float e1[3] = {tv[3] - tv[0], tv[4] - tv[1], tv[5] - tv[2]};
                float e2[3] = {tv[6] - tv[0], tv[7] - tv[1], tv[8] - tv[2]};
                float xx = dot3(e1, e1), xy = dot3(e1, e2), yy = dot3(e2, e2);
                float den = xx * yy - xy * xy;
                float d1 = (float)(t1 - t0), d2 = (float)(t2 - t0);
                float ca = den != 0.0f ? (d1 * yy - d2 * xy) / den : 0.0f;
                float cb = den != 0.0f ? (d2 * xx - d1 * xy) / den : 0.0f;
                float gv[3] = {ca * e1[0] + cb * e2[0], ca * e1[1] + cb * e2[1],
                               ca * e1[2] + cb * e2[2]};
                float jr[3];
                cross3(q, gv, jr);
                const double j0 = -(double)jr[0], j1 = -(double)jr[1], j2 = -(double)jr[2];
                term[1] = 2.0 * d * j0;
                term[2] = 2.0 * d * j1;
                term[3] = 2.0 * d * j2;
                term[4] = 2.0 * j0 * j0;
                term[5] = 2.0 * j1 * j1;
                term[6] = 2.0 * j2 * j2;
                term[7] = 2.0 * j0 * j1;
                term[8] = 2.0 * j0 * j2;
                term[9] = 2.0 * j1 * j2;
            }
        }
    }
    // all terms reduced together, in block_sum's order: a shuffle tree per
    // wave, then the waves' sums in order (one barrier instead of one per term)
    __shared__ double wave_sums[kBlock / 64][kTerms];
    const int wlane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kTerms; ++k) {
        double v = term[k];
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_down(v, off, 64);
        if (wlane == 0)
            wave_sums[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kTerms) {
        double sum = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            sum += wave_sums[w][threadIdx.x];
        partial[((size_t)a * kTerms + threadIdx.x) * nblocks + blockIdx.x] = sum;
    }
}

// Rotation SSD on feature maps: for every `stride`-th pixel x of the target
// map (equirectangular, pixel centre (x+½, y+½) at azimuth 2π(x+½)/nx and
// polar angle π(y+½)/ny), the source map is sampled bilinearly at Mᵀx (x
// circular, y clamped) and w(x) (S(Mᵀx) - T(x))² is summed in float64 with
// w = sin(polar angle), the pixel's area weight.  The continuous form equals
// the point cost's ∫ (s(p) - t(M p))² dp (a rotation keeps the measure), with
// no point location.  blockIdx.y is the candidate.
__global__ void __launch_bounds__(kBlock)
k_rotation_map_cost(const double *__restrict__ smap, const double *__restrict__ tmap, int nx,
                    int ny, int stride, int sx, int sy, const float *__restrict__ mats,
                    double *__restrict__ partial, int nblocks)
{
    __shared__ double red[kBlock / 64];
    __shared__ float m[9];
    const int a = blockIdx.y;
    if (threadIdx.x < 9)
        m[threadIdx.x] = mats[9 * (size_t)a + threadIdx.x];
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (k < sx * sy) {
        const int yi = (k / sx) * stride, xi = (k % sx) * stride;
        const float kPi = 3.14159265358979f;
        float st, ct, sp, cp;
        __sincosf(kPi * ((float)yi + 0.5f) / (float)ny, &st, &ct);
        __sincosf(2.0f * kPi * ((float)xi + 0.5f) / (float)nx, &sp, &cp);
        const float px = st * cp, py = st * sp, pz = ct;
        // Mᵀ x
        const float qx = m[0] * px + m[3] * py + m[6] * pz;
        const float qy = m[1] * px + m[4] * py + m[7] * pz;
        const float qz = m[2] * px + m[5] * py + m[8] * pz;
        float phi = atan2f(qy, qx);
        if (phi < 0.0f)
            phi += 2.0f * kPi;
        const float theta = acosf(fminf(fmaxf(qz, -1.0f), 1.0f));
        const float u = phi * (0.5f / kPi) * (float)nx - 0.5f;
        const float v = theta * (1.0f / kPi) * (float)ny - 0.5f;
        const float fu = floorf(u), fv = floorf(v);
        const float du = u - fu, dv = v - fv;
        int x0 = (int)fu, y0 = (int)fv;
        int x1 = x0 + 1, y1 = y0 + 1;
        x0 = x0 < 0 ? x0 + nx : (x0 >= nx ? x0 - nx : x0);
        x1 = x1 < 0 ? x1 + nx : (x1 >= nx ? x1 - nx : x1);
        y0 = y0 < 0 ? 0 : (y0 >= ny ? ny - 1 : y0);
        y1 = y1 < 0 ? 0 : (y1 >= ny ? ny - 1 : y1);
        const double s00 = smap[(size_t)y0 * nx + x0], s01 = smap[(size_t)y0 * nx + x1];
        const double s10 = smap[(size_t)y1 * nx + x0], s11 = smap[(size_t)y1 * nx + x1];
        const double sv = (s00 * (1.0 - du) + s01 * du) * (1.0 - dv) +
                          (s10 * (1.0 - du) + s11 * du) * dv;
        const double d = sv - tmap[(size_t)yi * nx + xi];
        e = (double)st * d * d;
    }
    double sum = block_sum(e, red);
    if (threadIdx.x == 0)
        partial[(size_t)a * nblocks + blockIdx.x] = sum;
}

// out[r] = sum of row r of the block partials; one block per row, a fixed
// (deterministic) summation order.
__global__ void __launch_bounds__(kBlock)
k_sum_rows(const double *__restrict__ partial, int nblocks, int rows, double *__restrict__ out)
{
    __shared__ double red[kBlock / 64];
    const int r = blockIdx.x;
    double s = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x)
        s += partial[(size_t)r * nblocks + k];
    s = block_sum(s, red);
    if (threadIdx.x == 0)
        out[r] = s;
}

// Bounding sphere of each face of tri ([nf][3][3] float32): the vertex mean
// and the largest vertex distance enlarged by 2 % and 1e-6 (covering the
// locator's barycentric slack and float32 rounding; unit-sphere scale).
__global__ void __launch_bounds__(kBlock)
k_face_spheres_f32(const float *__restrict__ tri, int nf, float4 *__restrict__ sph)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    const float *t = tri + 9 * (size_t)f;
    float c[3], r2 = 0.0f;
    for (int d = 0; d < 3; ++d)
        c[d] = (t[d] + t[3 + d] + t[6 + d]) / 3.0f;
    for (int v = 0; v < 3; ++v) {
        const float e0 = t[3 * v] - c[0], e1 = t[3 * v + 1] - c[1], e2 = t[3 * v + 2] - c[2];
        const float e = e0 * e0 + e1 * e1 + e2 * e2;
        r2 = e > r2 ? e : r2;
    }
    sph[f] = make_float4(c[0], c[1], c[2], sqrtf(r2) * 1.02f + 1e-6f);
}

int launch_face_spheres_f32(const float *tri, int nf, float4 *sph, hipStream_t s)
{
    if (nf <= 0)
        return kOk;
    hipLaunchKernelGGL(k_face_spheres_f32, dim3(grid_for(nf)), dim3(kBlock), 0, s, tri, nf, sph);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_rotation_costs(const RotationCostArgs &a, hipStream_t s)
{
    int nb = rotation_cost_blocks(a.ns);
    dim3 grid(nb, a.na);
    hipLaunchKernelGGL(k_rotation_cost, grid, dim3(kBlock), 0, s, a.src_pts, a.src_vals, a.ns,
                       a.tgt_vals, a.faces, a.tri, a.sph, a.cell_off, a.cell_faces, a.grid,
                       a.angles,
                       a.partial, nb, a.uncovered, a.loc_face, a.loc_bary);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_rows, dim3(a.na), dim3(kBlock), 0, s, a.partial, nb, a.na, a.costs);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_rotation_cost_mats(const RotationCostArgs &a, const float *mats, bool gradient,
                              hipStream_t s)
{
    int nb = rotation_cost_blocks(a.ns);
    dim3 grid(nb, a.na);
    if (gradient)
        hipLaunchKernelGGL(k_rotation_cost_mats<true>, grid, dim3(kBlock), 0, s, a.src_pts,
                           a.src_vals, a.ns, a.tgt_vals, a.faces, a.tri, a.sph, a.cell_off,
                           a.cell_faces, a.grid, mats, a.partial, nb, a.uncovered, a.hint,
                           a.face_vals);
    else
        hipLaunchKernelGGL(k_rotation_cost_mats<false>, grid, dim3(kBlock), 0, s, a.src_pts,
                           a.src_vals, a.ns, a.tgt_vals, a.faces, a.tri, a.sph, a.cell_off,
                           a.cell_faces, a.grid, mats, a.partial, nb, a.uncovered, a.hint,
                           a.face_vals);
    CSG_LAUNCH_CHECK();
    const int rows = a.na * (gradient ? kRotationGradTerms : 1);
    hipLaunchKernelGGL(k_sum_rows, dim3(rows), dim3(kBlock), 0, s, a.partial, nb, rows, a.costs);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int map_cost_blocks(int nx, int ny, int stride)
{
    const int sx = (nx + stride - 1) / stride, sy = (ny + stride - 1) / stride;
    return (int)grid_for((size_t)sx * sy);
}

int launch_rotation_map_costs(const double *smap, const double *tmap, int nx, int ny, int stride,
                              const float *mats, int na, double *partial, double *costs,
                              hipStream_t s)
{
    const int sx = (nx + stride - 1) / stride, sy = (ny + stride - 1) / stride;
    const int nb = map_cost_blocks(nx, ny, stride);
    hipLaunchKernelGGL(k_rotation_map_cost, dim3(nb, na), dim3(kBlock), 0, s, smap, tmap, nx, ny,
                       stride, sx, sy, mats, partial, nb);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_rows, dim3(na), dim3(kBlock), 0, s, partial, nb, na, costs);
    CSG_LAUNCH_CHECK();
    return kOk;
}

// ------------------------------------------------------------ closest triangle

// Closest point on a triangle with its barycentric weights (float64), by the
// Voronoi-region method of Ericson, Real-Time Collision Detection, 5.1.5.
__device__ double closest_on_triangle(const double *p, const double *a, const double *b,
                                      const double *c, double *w)
{
    double ab[3], ac[3], ap[3];
    for (int d = 0; d < 3; ++d) {
        ab[d] = b[d] - a[d];
        ac[d] = c[d] - a[d];
        ap[d] = p[d] - a[d];
    }
    auto dot = [](const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        w[0] = 1; w[1] = 0; w[2] = 0;
    } else {
        double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
        double d3 = dot(ab, bp), d4 = dot(ac, bp);
        if (d3 >= 0.0 && d4 <= d3) {
            w[0] = 0; w[1] = 1; w[2] = 0;
        } else {
            double vc = d1 * d4 - d3 * d2;
            if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
                double v = d1 / (d1 - d3);
                w[0] = 1 - v; w[1] = v; w[2] = 0;
            } else {
                double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
                double d5 = dot(ab, cp), d6 = dot(ac, cp);
                if (d6 >= 0.0 && d5 <= d6) {
                    w[0] = 0; w[1] = 0; w[2] = 1;
                } else {
                    double vb = d5 * d2 - d1 * d6;
                    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
                        double t = d2 / (d2 - d6);
                        w[0] = 1 - t; w[1] = 0; w[2] = t;
                    } else {
                        double va = d3 * d6 - d5 * d4;
                        if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
                            double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
                            w[0] = 0; w[1] = 1 - t; w[2] = t;
                        } else {
                            double den = 1.0 / (va + vb + vc);
                            double v = vb * den, t = vc * den;
                            w[0] = 1 - v - t; w[1] = v; w[2] = t;
                        }
                    }
                }
            }
        }
    }
    double e = 0.0;
    for (int d = 0; d < 3; ++d) {
        double q = w[0] * a[d] + w[1] * b[d] + w[2] * c[d];
        e += (p[d] - q) * (p[d] - q);
    }
    return e;
}

// Bounding sphere of each face: centre (vertex mean) and a radius enlarged by
// a relative 1e-12 and an absolute 1e-14 (coordinates are within [-1, 1]), so
// that |p - c| - r is a lower bound of p's distance to the face despite
// rounding.
__global__ void __launch_bounds__(kBlock)
k_face_spheres(const double *__restrict__ verts, const int *__restrict__ faces, int nf,
               double4 *__restrict__ sph, double *__restrict__ tri)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    const double *a = verts + 3 * faces[3 * f], *b = verts + 3 * faces[3 * f + 1],
                 *c = verts + 3 * faces[3 * f + 2];
    // the face's corners side by side (one gather per candidate in the scan)
    double *t = tri + 9 * (size_t)f;
    for (int d = 0; d < 3; ++d) {
        t[d] = a[d];
        t[3 + d] = b[d];
        t[6 + d] = c[d];
    }
    double m[3], r2 = 0.0;
    for (int d = 0; d < 3; ++d)
        m[d] = (a[d] + b[d] + c[d]) / 3.0;
    for (const double *v : {a, b, c}) {
        double e = 0.0;
        for (int d = 0; d < 3; ++d)
            e += (v[d] - m[d]) * (v[d] - m[d]);
        r2 = e > r2 ? e : r2;
    }
    sph[f] = make_double4(m[0], m[1], m[2], sqrt(r2) * (1.0 + 1e-12) + 1e-14);
}

// For each query: closest triangle among the candidates of its cell (first
// minimum); `far` counts queries whose best squared distance exceeds
// reach2, i.e. whose true closest triangle may lie outside the searched cells.
// kLocateLanes consecutive lanes share one query, each scanning every
// kLocateLanes-th candidate; the lanes' bests are combined as in the
// sequential scan (smaller distance, then earlier candidate).
__global__ void __launch_bounds__(kBlock)
k_closest_triangles(const double *__restrict__ q, int nq, const double *__restrict__ verts,
                    const int *__restrict__ faces, const int *__restrict__ cell_off,
                    const int *__restrict__ cell_faces, int g, double reach2,
                    int *__restrict__ face_out, int *__restrict__ idx_out,
                    double *__restrict__ w_out, int *__restrict__ far_count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / kLocateLanes, lane = t % kLocateLanes;
    double p[3] = {0.0, 0.0, 0.0};
    int begin = 0, end = 0;
    if (i < nq) {
        p[0] = q[3 * i];
        p[1] = q[3 * i + 1];
        p[2] = q[3 * i + 2];
        double cell = 2.0 / (double)g;
        int c[3];
        for (int d = 0; d < 3; ++d) {
            long k = (long)floor((p[d] + 1.0) / cell);
            c[d] = (int)(k < 0 ? 0 : (k > g - 1 ? g - 1 : k));
        }
        int id = c[0] + g * (c[1] + g * c[2]);
        begin = cell_off[id];
        end = cell_off[id + 1];
    }
    double best = INFINITY, bw[3] = {0, 0, 0};
    int bf = -1, bpos = 0x7fffffff;
    for (int k = begin + lane; k < end; k += kLocateLanes) {
        int f = cell_faces[k];
        const int *fv = faces + 3 * f;
        double w[3];
        double dist = closest_on_triangle(p, verts + 3 * fv[0], verts + 3 * fv[1],
                                          verts + 3 * fv[2], w);
        if (dist < best) {
            best = dist;
            bf = f;
            bpos = k;
            bw[0] = w[0];
            bw[1] = w[1];
            bw[2] = w[2];
        }
    }
    for (int m = 1; m < kLocateLanes; m <<= 1) {
        double ob = __shfl_xor(best, m);
        int of = __shfl_xor(bf, m), op = __shfl_xor(bpos, m);
        double o0 = __shfl_xor(bw[0], m), o1 = __shfl_xor(bw[1], m), o2 = __shfl_xor(bw[2], m);
        if (ob < best || (ob == best && op < bpos)) {
            best = ob;
            bf = of;
            bpos = op;
            bw[0] = o0;
            bw[1] = o1;
            bw[2] = o2;
        }
    }
    if (i >= nq || lane != 0)
        return;
    if (bf < 0 || best > reach2)
        atomicAdd(far_count, 1);
    face_out[i] = bf;
    for (int d = 0; d < 3; ++d) {
        w_out[3 * i + d] = bw[d];
        if (idx_out)
            idx_out[3 * i + d] = bf >= 0 ? faces[3 * bf + d] : 0;
    }
}

// k_closest_triangles with each candidate first tested against its face's
// bounding sphere (sph): a lane skips a candidate whose sphere lies farther
// than sqrt(best)·(1 + 1e-9) + r: its distance then exceeds the lane's best,
// which is at least the final best, so it could neither win nor tie (the skip
// changes no result).  The distances are taken on face-major corner copies
// (tri, 9 values per face, written by k_face_spheres): a face's corners are
// one gather instead of an index and three vertex gathers.  Each lane fetches
// kClosestBatch candidate indices, then their spheres, at once, and evaluates
// them in the order of the sequential scan (increasing position), so the
// result is the same.
constexpr int kClosestBatch = 4;
constexpr int kClosestLanes = 16;

template <int LANES>
__global__ void __launch_bounds__(kBlock)
k_closest_triangles_tri(const double *__restrict__ q, int nq, const double *__restrict__ tri,
                        const int *__restrict__ faces, const int *__restrict__ cell_off,
                        const int *__restrict__ cell_faces, int g, double reach2,
                        const double4 *__restrict__ sph, int *__restrict__ face_out,
                        int *__restrict__ idx_out, double *__restrict__ w_out,
                        int *__restrict__ far_count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / LANES, lane = t % LANES;
    double p[3] = {0.0, 0.0, 0.0};
    int begin = 0, end = 0;
    if (i < nq) {
        p[0] = q[3 * i];
        p[1] = q[3 * i + 1];
        p[2] = q[3 * i + 2];
        double cell = 2.0 / (double)g;
        int c[3];
        for (int d = 0; d < 3; ++d) {
            long k = (long)floor((p[d] + 1.0) / cell);
            c[d] = (int)(k < 0 ? 0 : (k > g - 1 ? g - 1 : k));
        }
        int id = c[0] + g * (c[1] + g * c[2]);
        begin = cell_off[id];
        end = cell_off[id + 1];
    }
    double best = INFINITY, bw[3] = {0, 0, 0}, reach = INFINITY;
    int bf = -1, bpos = 0x7fffffff;
    for (int k0 = begin + lane; k0 < end; k0 += LANES * kClosestBatch) {
        int fb[kClosestBatch];
        double4 sb[kClosestBatch];
#pragma unroll
        for (int b = 0; b < kClosestBatch; ++b) {
            const int k = k0 + b * LANES;
            fb[b] = k < end ? cell_faces[k] : -1;
        }
#pragma unroll
        for (int b = 0; b < kClosestBatch; ++b)
            sb[b] = fb[b] >= 0 ? sph[fb[b]] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
        for (int b = 0; b < kClosestBatch; ++b) {
            if (fb[b] < 0)
                continue;
            const double dx = p[0] - sb[b].x, dy = p[1] - sb[b].y, dz = p[2] - sb[b].z;
            const double lim = sb[b].w + reach;
            if (dx * dx + dy * dy + dz * dz > lim * lim)
                continue;
            const double *tf = tri + 9 * (size_t)fb[b];
            double w[3];
            const double dist = closest_on_triangle(p, tf, tf + 3, tf + 6, w);
            if (dist < best) {
                best = dist;
                reach = sqrt(best) * (1.0 + 1e-9);
                bf = fb[b];
                bpos = k0 + b * LANES;
                bw[0] = w[0];
                bw[1] = w[1];
                bw[2] = w[2];
            }
        }
    }
    for (int m = 1; m < LANES; m <<= 1) {
        double ob = __shfl_xor(best, m);
        int of = __shfl_xor(bf, m), op = __shfl_xor(bpos, m);
        double o0 = __shfl_xor(bw[0], m), o1 = __shfl_xor(bw[1], m), o2 = __shfl_xor(bw[2], m);
        if (ob < best || (ob == best && op < bpos)) {
            best = ob;
            bf = of;
            bpos = op;
            bw[0] = o0;
            bw[1] = o1;
            bw[2] = o2;
        }
    }
    if (i >= nq || lane != 0)
        return;
    if (bf < 0 || best > reach2)
        atomicAdd(far_count, 1);
    face_out[i] = bf;
    for (int d = 0; d < 3; ++d) {
        w_out[3 * i + d] = bw[d];
        if (idx_out)
            idx_out[3 * i + d] = bf >= 0 ? faces[3 * bf + d] : 0;
    }
}

// Exact closest triangle on an index built without margin (each face listed
// in the cells its bounding box overlaps), for queries near the mesh.  A query
// scans its own cell's list; if its best distance is below the distance from
// the query to that cell's walls (less 1e-9 of a cell), every face with a
// closer point would have that point inside the cell and so be in the list,
// and the result is exact.  Otherwise it also scans the other cells of the
// (2·span + 1)^3 block around its cell, exact when the best distance is below
// the distance to the block's walls; beyond that it counts as far.  A wall on
// the grid's boundary is at infinity (points outside the grid fall into, and
// clipped boxes are listed in, the boundary cells).  Candidates are tested
// against their face spheres as in k_closest_triangles_tri; ties go to the
// smaller face index, which is the first minimum in face order (a face may be
// met in several lists).
__device__ __forceinline__ void scan_face_list(const double *p, const double *__restrict__ tri,
                                               const double4 *__restrict__ sph,
                                               const int *__restrict__ list, int n, int lane,
                                               double &best, int &bf, double *bw,
                                               double &reach)
{
    for (int k0 = lane; k0 < n; k0 += kLocateLanes * kClosestBatch) {
        int fb[kClosestBatch];
        double4 sb[kClosestBatch];
#pragma unroll
        for (int b = 0; b < kClosestBatch; ++b) {
            const int k = k0 + b * kLocateLanes;
            fb[b] = k < n ? list[k] : -1;
        }
#pragma unroll
        for (int b = 0; b < kClosestBatch; ++b)
            sb[b] = fb[b] >= 0 ? sph[fb[b]] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
        for (int b = 0; b < kClosestBatch; ++b) {
            if (fb[b] < 0)
                continue;
            const double dx = p[0] - sb[b].x, dy = p[1] - sb[b].y, dz = p[2] - sb[b].z;
            const double lim = sb[b].w + reach;
            if (dx * dx + dy * dy + dz * dz > lim * lim)
                continue;
            const double *tf = tri + 9 * (size_t)fb[b];
            double w[3];
            const double dist = closest_on_triangle(p, tf, tf + 3, tf + 6, w);
            if (dist < best || (dist == best && fb[b] < bf)) {
                best = dist;
                reach = sqrt(best) * (1.0 + 1e-9);
                bf = fb[b];
                bw[0] = w[0];
                bw[1] = w[1];
                bw[2] = w[2];
            }
        }
    }
}

__device__ __forceinline__ void reduce_best(double &best, int &bf, double *bw)
{
    for (int m = 1; m < kLocateLanes; m <<= 1) {
        const double ob = __shfl_xor(best, m);
        const int of = __shfl_xor(bf, m);
        const double o0 = __shfl_xor(bw[0], m), o1 = __shfl_xor(bw[1], m),
                     o2 = __shfl_xor(bw[2], m);
        if (ob < best || (ob == best && (unsigned)of < (unsigned)bf)) {
            best = ob;
            bf = of;
            bw[0] = o0;
            bw[1] = o1;
            bw[2] = o2;
        }
    }
}

// Distance from p to the walls of the cell block [lo, hi] (cell indices per
// axis) of a g^3 grid over [-1, 1]^3 of cell size `cell`; walls on the grid's
// boundary do not count.
__device__ __forceinline__ double block_reach(const double *p, const int *lo, const int *hi,
                                              int g, double cell)
{
    double r = INFINITY;
    for (int d = 0; d < 3; ++d) {
        if (lo[d] > 0)
            r = fmin(r, p[d] - ((double)lo[d] * cell - 1.0));
        if (hi[d] < g - 1)
            r = fmin(r, ((double)(hi[d] + 1) * cell - 1.0) - p[d]);
    }
    return r;
}

__global__ void __launch_bounds__(kBlock)
k_closest_triangles_block(const double *__restrict__ q, int nq, const double *__restrict__ tri,
                          const int *__restrict__ faces, const int *__restrict__ cell_off,
                          const int *__restrict__ cell_faces, int g, int span,
                          const double4 *__restrict__ sph, int *__restrict__ face_out,
                          int *__restrict__ idx_out, double *__restrict__ w_out,
                          int *__restrict__ far_count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / kLocateLanes, lane = t % kLocateLanes;
    const double cell = 2.0 / (double)g, slack = 1e-9 * cell;
    double p[3] = {0.0, 0.0, 0.0};
    int c[3] = {0, 0, 0}, begin = 0, end = 0;
    if (i < nq) {
        p[0] = q[3 * i];
        p[1] = q[3 * i + 1];
        p[2] = q[3 * i + 2];
        for (int d = 0; d < 3; ++d) {
            long k = (long)floor((p[d] + 1.0) / cell);
            c[d] = (int)(k < 0 ? 0 : (k > g - 1 ? g - 1 : k));
        }
        const int id = c[0] + g * (c[1] + g * c[2]);
        begin = cell_off[id];
        end = cell_off[id + 1];
    }
    double best = INFINITY, bw[3] = {0, 0, 0}, reach = INFINITY;
    int bf = -1;
    scan_face_list(p, tri, sph, cell_faces + begin, end - begin, lane, best, bf, bw, reach);
    reduce_best(best, bf, bw);
    const double own = block_reach(p, c, c, g, cell) - slack;
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = c[d] - span < 0 ? 0 : c[d] - span;
        hi[d] = c[d] + span > g - 1 ? g - 1 : c[d] + span;
    }
    // the same decision on every lane of the query (best is reduced)
    if (i < nq && span > 0 && !(bf >= 0 && sqrt(best) <= own)) {
        reach = bf >= 0 ? sqrt(best) * (1.0 + 1e-9) : INFINITY;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    if (x == c[0] && y == c[1] && z == c[2])
                        continue;
                    const int id = x + g * (y + g * z);
                    const int b0 = cell_off[id], b1 = cell_off[id + 1];
                    scan_face_list(p, tri, sph, cell_faces + b0, b1 - b0, lane, best, bf, bw,
                                   reach);
                }
        reduce_best(best, bf, bw);
    }
    if (i >= nq || lane != 0)
        return;
    const double limit = (span > 0 ? block_reach(p, lo, hi, g, cell) : own + slack) - slack;
    if (bf < 0 || !(sqrt(best) <= limit))
        atomicAdd(far_count, 1);
    face_out[i] = bf;
    for (int d = 0; d < 3; ++d) {
        w_out[3 * i + d] = bw[d];
        if (idx_out)
            idx_out[3 * i + d] = bf >= 0 ? faces[3 * bf + d] : 0;
    }
}

int launch_closest_triangles_block(const double *q, int nq, const double *verts,
                                   const int *faces, int nf, const int *cell_off,
                                   const int *cell_faces, int g, int span, double4 *sph,
                                   double *tri, int *face_out, int *idx_out, double *w_out,
                                   int *far_count, hipStream_t s)
{
    if (nq <= 0)
        return kOk;
    hipLaunchKernelGGL(k_face_spheres, dim3(grid_for(nf)), dim3(kBlock), 0, s, verts, faces, nf,
                       sph, tri);
    CSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_closest_triangles_block, dim3(grid_for((size_t)nq * kLocateLanes)),
                       dim3(kBlock), 0, s, q, nq, tri, faces, cell_off, cell_faces, g, span, sph,
                       face_out, idx_out, w_out, far_count);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_closest_triangles(const double *q, int nq, const double *verts, const int *faces,
                             const int *cell_off, const int *cell_faces, int g, double reach2,
                             int *face_out, int *idx_out, double *w_out, int *far_count,
                             hipStream_t s)
{
    if (nq <= 0)
        return kOk;
    hipLaunchKernelGGL(k_closest_triangles, dim3(grid_for((size_t)nq * kLocateLanes)),
                       dim3(kBlock), 0, s, q, nq, verts, faces, cell_off, cell_faces, g, reach2,
                       face_out, idx_out, w_out, far_count);
    CSG_LAUNCH_CHECK();
    return kOk;
}

int launch_closest_triangles_pruned(const double *q, int nq, const double *verts,
                                    const int *faces, int nf, const int *cell_off,
                                    const int *cell_faces, int g, double reach2, double4 *sph,
                                    double *tri, int *face_out, int *idx_out, double *w_out,
                                    int *far_count, hipStream_t s)
{
    if (nq <= 0)
        return kOk;
    hipLaunchKernelGGL(k_face_spheres, dim3(grid_for(nf)), dim3(kBlock), 0, s, verts, faces, nf,
                       sph, tri);
    CSG_LAUNCH_CHECK();
    // sixteen lanes per query: a query of the surface stencil has ~80 candidates
    // and the launch only ~41k queries, so four lanes left the device at ~2.5
    // waves per SIMD with each lane scanning its share in sequence
    hipLaunchKernelGGL(k_closest_triangles_tri<kClosestLanes>,
                       dim3(grid_for((size_t)nq * kClosestLanes)), dim3(kBlock), 0, s, q, nq,
                       tri, faces, cell_off, cell_faces, g, reach2, sph, face_out, idx_out,
                       w_out, far_count);
    CSG_LAUNCH_CHECK();
    return kOk;
}

}  // namespace csg
