// pga_lbvh.hip - linear BVH over terminal AABBs on gfx950 (wave64).
//
// The construction of GPUBVHConstructor.cuh (Karras 2012) in six steps:
//   1. k_morton  - AABB centre normalised to the scene box, 10 bits per axis
//                  interleaved into a 30-bit Morton code (GPUBVHConstructor.cuh:
//                  44-86); the payload is the AABB index;
//   2. LSD radix sort of (code, index) pairs, four 8-bit passes, each
//                  k_hist -> scan (k_scan_reduce, k_scan_top, k_scan_down) ->
//                  k_scatter; the scatter is stable (per-wave ranks from eight
//                  ballots, per-workgroup digit cursors in LDS), so sorting the
//                  codes of indices in ascending order gives the order of the
//                  reference's unique 64-bit keys (code << 32 | index), which
//                  thrust::sort orders (:311-312); the last pass writes those
//                  keys directly;
//   3. k_karras  - one thread per internal node: direction, range and split
//                  by binary search over the common-prefix length of the
//                  64-bit keys (buildRadixTree, :88-213); it also records the
//                  far end of each node's leaf range;
//   4. refit     - climbs from the leaves: the first child to reach a node
//                  stops, the second unites both boxes and goes on
//                  (propagateBVH's ticket, :215-261). k_refit_local completes
//                  the nodes whose leaf range lies inside one workgroup's tile
//                  of sorted leaves, through LDS; k_refit_global the nodes
//                  spanning tiles, through global memory, where boxes are
//                  written and read with agent-scope atomic loads and stores,
//                  coherent across the XCDs' L2 caches without a full
//                  write-back, and each store completes before the ticket of
//                  the parent is taken.
// Children are encoded as in the reference: >= 0 a leaf (sorted position),
// < 0 an internal node (negated index); node 0 is the root.
//
// Device-count mode: every kernel that reads the leaf count n takes the host
// bound n (grid size) and an optional device word nd; with nd the count is
// min(*nd, n), so a gather of the derived terminals, the scene box and the
// whole build run from device state without a host round trip, and can be
// recorded in a HIP graph and replayed while the terminal counts change.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/pga.h"

namespace bvhk {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kItems = 16;
constexpr uint32_t kTile = kThreads * kItems;   // keys per sort workgroup
constexpr uint32_t kScanTile = 4096;            // histogram entries per scan workgroup
constexpr uint32_t kGatherBlocks = 8192;         // grid-stride workgroups of the terminal gather
constexpr uint32_t kRefitTile = 1024;            // sorted leaves per refit workgroup (256 threads)

// leaf count: the host bound, or the device word when there is one
__device__ __forceinline__ uint32_t leaf_count(uint32_t n, const uint32_t* nd) {
    return nd ? min(*nd, n) : n;
}

__device__ __forceinline__ uint32_t f2u_rd(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {     // GPUBVHConstructor.cuh:48-54
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton30(float x, float y, float z) {   // :56-64
    uint32_t vx = min(f2u_rd(x * 1023.0f), 1023u);
    uint32_t vy = min(f2u_rd(y * 1023.0f), 1023u);
    uint32_t vz = min(f2u_rd(z * 1023.0f), 1023u);
    return spread10(vx) | (spread10(vy) << 1) | (spread10(vz) << 2);
}

// AABB of the unit cube (or, planar, the unit square z = 0) under an
// instance's model rows: centre = translation, half extent of row r =
// 0.5 (|m_r0| + |m_r1| (+ |m_r2|)).
__global__ void __launch_bounds__(256) k_instance_aabbs(const pga_instance* __restrict__ inst, uint32_t n,
                                                        uint32_t planar, float* __restrict__ aabbs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* m = reinterpret_cast<const float4*>(inst + i);
    const float4 r0 = m[0], r1 = m[1], r2 = m[2];
    float h0 = fabsf(r0.x) + fabsf(r0.y), h1 = fabsf(r1.x) + fabsf(r1.y), h2 = fabsf(r2.x) + fabsf(r2.y);
    if (!planar) { h0 = h0 + fabsf(r0.z); h1 = h1 + fabsf(r1.z); h2 = h2 + fabsf(r2.z); }
    h0 = h0 * 0.5f; h1 = h1 * 0.5f; h2 = h2 * 0.5f;
    float* o = aabbs + (size_t)i * 6u;
    o[0] = r0.w - h0; o[1] = r1.w - h1; o[2] = r2.w - h2;
    o[3] = r0.w + h0; o[4] = r1.w + h1; o[5] = r2.w + h2;
}

// Boxes of the instances of terminal types 0..ntypes-1, concatenated in type
// order (each buffer in its own order): counts come from the derivation's
// device counters, each clamped to its buffer's capacity. Writes the leaf
// count min(total, capacity) and an overflow flag.
struct GatherArgs {
    const pga_instance* buf[PGA_MAX_TERMINALS];
    unsigned long long cap[PGA_MAX_TERMINALS];
    uint32_t ntypes, planar_mask, capacity;
};

__global__ void __launch_bounds__(256) k_gather_boxes(GatherArgs g, const unsigned long long* __restrict__ counts,
                                                      float* __restrict__ aabbs, uint32_t* __restrict__ nd,
                                                      uint32_t* __restrict__ overflow) {
    __shared__ unsigned long long start[PGA_MAX_TERMINALS + 1];
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (uint32_t t = 0; t < g.ntypes; ++t) {
            start[t] = s;
            const unsigned long long c = counts[t];
            s += c < g.cap[t] ? c : g.cap[t];
        }
        start[g.ntypes] = s;
        if (blockIdx.x == 0) {
            *nd = (uint32_t)(s < g.capacity ? s : g.capacity);
            *overflow = s > g.capacity ? 1u : 0u;
        }
    }
    __syncthreads();
    const unsigned long long total = start[g.ntypes] < g.capacity ? start[g.ntypes] : g.capacity;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        uint32_t t = 0;                       // the last type whose range starts at or before i
        for (uint32_t step = 32; step > 0; step >>= 1)
            if (t + step < g.ntypes && start[t + step] <= i) t += step;
        const float4* m = reinterpret_cast<const float4*>(g.buf[t] + (i - start[t]));
        const float4 r0 = m[0], r1 = m[1], r2 = m[2];
        float h0 = fabsf(r0.x) + fabsf(r0.y), h1 = fabsf(r1.x) + fabsf(r1.y), h2 = fabsf(r2.x) + fabsf(r2.y);
        if (!((g.planar_mask >> t) & 1u)) { h0 = h0 + fabsf(r0.z); h1 = h1 + fabsf(r1.z); h2 = h2 + fabsf(r2.z); }
        h0 = h0 * 0.5f; h1 = h1 * 0.5f; h2 = h2 * 0.5f;
        float* o = aabbs + i * 6u;
        o[0] = r0.w - h0; o[1] = r1.w - h1; o[2] = r2.w - h2;
        o[3] = r0.w + h0; o[4] = r1.w + h1; o[5] = r2.w + h2;
    }
}

// Per-workgroup min/max of the AABB corners (scene box), then one workgroup
// folds the partials.
__global__ void __launch_bounds__(256) k_bounds(const float* __restrict__ aabbs, uint32_t n, const uint32_t* nd,
                                                float* __restrict__ part) {
    __shared__ float red[6][256];
    n = leaf_count(n, nd);
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* a = aabbs + (size_t)i * 6u;
        for (int q = 0; q < 3; ++q) { b[q] = fminf(b[q], a[q]); b[q + 3] = fmaxf(b[q + 3], a[q + 3]); }
    }
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = b[q];
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 6; ++q)
                red[q][threadIdx.x] = q < 3 ? fminf(red[q][threadIdx.x], red[q][threadIdx.x + s])
                                            : fmaxf(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6u + threadIdx.x] = red[threadIdx.x][0];
}
__global__ void __launch_bounds__(256) k_bounds_fold(float* part, uint32_t nparts) {
    __shared__ float red[6][256];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < nparts; b += blockDim.x)
        for (int q = 0; q < 6; ++q) v[q] = q < 3 ? fminf(v[q], part[b * 6u + q]) : fmaxf(v[q], part[b * 6u + q]);
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int q = 0; q < 6; ++q)
                red[q][threadIdx.x] = q < 3 ? fminf(red[q][threadIdx.x], red[q][threadIdx.x + h])
                                            : fmaxf(red[q][threadIdx.x], red[q][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[threadIdx.x] = red[threadIdx.x][0];
}

// scene box from the host (mn, ex) or, with `box`, from the device: min xyz,
// max xyz, an empty extent taken as 1
__global__ void __launch_bounds__(256) k_morton(const float* __restrict__ aabbs, uint32_t n, const uint32_t* nd,
                                                float mn0, float mn1, float mn2, float ex0, float ex1, float ex2,
                                                const float* __restrict__ box,
                                                uint32_t* __restrict__ code, uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= leaf_count(n, nd)) return;
    if (box) {
        mn0 = box[0]; mn1 = box[1]; mn2 = box[2];
        ex0 = box[3] - mn0; ex1 = box[4] - mn1; ex2 = box[5] - mn2;
        ex0 = ex0 > 0.0f ? ex0 : 1.0f; ex1 = ex1 > 0.0f ? ex1 : 1.0f; ex2 = ex2 > 0.0f ? ex2 : 1.0f;
    }
    const float* b = aabbs + (size_t)i * 6u;
    const float cx = (b[0] + b[3]) * 0.5f, cy = (b[1] + b[4]) * 0.5f, cz = (b[2] + b[5]) * 0.5f;
    code[i] = morton30((cx - mn0) / ex0, (cy - mn1) / ex1, (cz - mn2) / ex2);
    idx[i] = i;
}

// ---- radix sort: histogram, scan, stable scatter (8-bit digits) ----
__global__ void __launch_bounds__(kThreads) k_hist(const uint32_t* __restrict__ code, uint32_t n, const uint32_t* nd,
                                                   uint32_t shift, uint32_t* __restrict__ hist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    n = leaf_count(n, nd);
    const uint32_t base = blockIdx.x * kTile;
    if (base >= n) {                          // tile past the leaves: an empty histogram
        hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = 0u;
        return;
    }
    h[threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t it = 0; it < kItems; ++it) {
        const uint32_t i = base + it * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(code[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];   // digit-major
}

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    // 256 threads, 4 waves: wave inclusive scan by shuffles, then wave offsets
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if ((int)lane >= d) x += y;
    }
    if (lane == 63u) sh[w] = x;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t k = 0; k < w; ++k) off += sh[k];
    *total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return off + x - v;
}

__global__ void __launch_bounds__(kThreads) k_scan_reduce(const uint32_t* __restrict__ a, uint32_t n,
                                                          uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[4];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (uint32_t k = threadIdx.x; k < kScanTile; k += kThreads) {
        const uint32_t i = base + k;
        if (i < n) s += a[i];
    }
    uint32_t total;
    (void)block_exclusive_scan(s, sh, &total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kThreads) k_scan_top(uint32_t* part, uint32_t nparts) {
    __shared__ uint32_t sh[4];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += kThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nparts ? part[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, sh, &total);
        if (i < nparts) part[i] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(kThreads) k_scan_down(uint32_t* __restrict__ a, uint32_t n,
                                                        const uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[4];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t carry = part[blockIdx.x];
    for (uint32_t k = 0; k < kScanTile; k += kThreads) {
        const uint32_t i = base + k + threadIdx.x;
        const uint32_t v = i < n ? a[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, sh, &total);
        if (i < n) a[i] = carry + ex;
        carry += total;
    }
}

// keys_out (the last pass): the sorted 64-bit keys (code << 32 | index)
// instead of the code and index arrays
__global__ void __launch_bounds__(kThreads) k_scatter(const uint32_t* __restrict__ code_in,
                                                      const uint32_t* __restrict__ idx_in,
                                                      uint32_t* __restrict__ code_out, uint32_t* __restrict__ idx_out,
                                                      unsigned long long* __restrict__ keys_out,
                                                      uint32_t n, const uint32_t* nd, uint32_t shift,
                                                      const uint32_t* __restrict__ offs, uint32_t nblocks) {
    __shared__ uint32_t base_d[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    n = leaf_count(n, nd);
    if (blockIdx.x * kTile >= n) return;      // whole workgroup: no leaves in this tile
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    base_d[t] = offs[(size_t)t * nblocks + blockIdx.x];
    run[t] = 0u;
    for (int k = 0; k < 4; ++k) wcnt[k][t] = 0u;
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (uint32_t it = 0; it < kItems; ++it) {
        const uint32_t i = blockIdx.x * kTile + it * kThreads + t;
        const bool valid = i < n;
        const uint32_t key = valid ? code_in[i] : 0u;
        const uint32_t val = valid ? idx_in[i] : 0u;
        const uint32_t d = (key >> shift) & 0xFFu;
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long on = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? on : ~on;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0u) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = base_d[d] + run[d] + rank;
            for (uint32_t k = 0; k < w; ++k) off += wcnt[k][d];
            if (off < n) {            // always true when the histograms are consistent
                if (keys_out) {
                    keys_out[off] = ((unsigned long long)key << 32) | (unsigned long long)val;
                } else {
                    code_out[off] = key;
                    idx_out[off] = val;
                }
            }
        }
        __syncthreads();
        run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        for (int k = 0; k < 4; ++k) wcnt[k][t] = 0u;
        __syncthreads();
    }
}

// ---- Karras build ----
// The keys of the workgroup's nodes and kKarrasHalo keys on either side are
// staged in LDS: most nodes span few leaves, so their searches stay in the
// window; a search that leaves it reads the key from global memory.
constexpr int kKarrasHalo = 256;
constexpr int kKarrasWindow = 256 + 2 * kKarrasHalo;

// other_end[i]: the end of node i's leaf range other than i (the range is
// [min(i, j), max(i, j)])
__global__ void __launch_bounds__(256) k_karras(const unsigned long long* __restrict__ keys, uint32_t n,
                                                const uint32_t* nd, int32_t* __restrict__ left,
                                                int32_t* __restrict__ right, int32_t* __restrict__ parent,
                                                int32_t* __restrict__ leaf_parent, int32_t* __restrict__ other_end,
                                                uint32_t* __restrict__ ticket) {
    __shared__ unsigned long long sk[kKarrasWindow];
    const int64_t nn = leaf_count(n, nd);
    const int64_t ws = (int64_t)blockIdx.x * blockDim.x - kKarrasHalo;
    for (int t = threadIdx.x; t < kKarrasWindow; t += blockDim.x) {
        const int64_t k = ws + t;
        sk[t] = (k >= 0 && k < nn) ? keys[k] : 0ull;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn - 1) return;
    ticket[i] = 0u;
    if (i == 0) parent[0] = -1;
    // LDS and global pointers of their own address spaces: the two loads
    // cannot be merged into one load through a selected generic pointer,
    // which would be a flat load
    typedef __attribute__((address_space(3))) const unsigned long long lds_u64;
    typedef __attribute__((address_space(1))) const unsigned long long glb_u64;
    const lds_u64* wk = (const lds_u64*)sk;
    const glb_u64* gk = (const glb_u64*)keys;
    const unsigned long long ki = wk[i - ws];
    auto delta = [&](int64_t jj) -> int {
        if (jj < 0 || jj >= nn) return -1;
        const int64_t o = jj - ws;
        unsigned long long kj;
        if (o >= 0 && o < kKarrasWindow) kj = wk[o];
        else kj = gk[jj];
        return __clzll((long long)(ki ^ kj));
    };
    const int dir = (delta(i + 1) - delta(i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(i - dir);
    int64_t lmax = 2;
    while (delta(i + lmax * dir) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(i + (l + t) * dir) > dmin) l += t;
    const int64_t j = i + l * dir;
    const int dnode = delta(j);
    int64_t s = 0;
    for (int64_t t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (delta(i + (s + t) * dir) > dnode) s += t;
        if (t == 1) break;
    }
    const int64_t gamma = i + s * dir + (dir < 0 ? -1 : 0);
    const int32_t lc = (min(i, j) == gamma) ? (int32_t)gamma : -(int32_t)gamma;
    const int32_t rc = (max(i, j) == gamma + 1) ? (int32_t)(gamma + 1) : -(int32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    other_end[i] = (int32_t)j;
    if (lc < 0) parent[-lc] = (int32_t)i; else leaf_parent[lc] = (int32_t)i;
    if (rc < 0) parent[-rc] = (int32_t)i; else leaf_parent[rc] = (int32_t)i;
}

// ---- refit ----
__device__ __forceinline__ void store_box(float* p, const float b[6]) {
#pragma unroll
    for (int q = 0; q < 6; ++q)
        __hip_atomic_store(reinterpret_cast<uint32_t*>(p) + q, __float_as_uint(b[q]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void load_box(const float* p, float b[6]) {
#pragma unroll
    for (int q = 0; q < 6; ++q)
        b[q] = __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p) + q, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT));
}

// The union is taken in child order, left box first: min(L, R) = R < L ? R : L
// and max(L, R) = L < R ? R : L, as std::min / std::max fold them.
__device__ __forceinline__ void unite(const float b[6], const float o[6], bool mine_left, float u[6]) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const float L = mine_left ? b[q] : o[q], R = mine_left ? o[q] : b[q];
        u[q] = q < 3 ? (R < L ? R : L) : (L < R ? R : L);
    }
}

// a node is local to the tile [s, e) of sorted leaves when its leaf range
// [min(i, j), max(i, j)] lies inside it; both its children then do too
__device__ __forceinline__ bool is_local(int32_t i, int32_t j, uint32_t s, uint32_t e) {
    return (uint32_t)min(i, j) >= s && (uint32_t)max(i, j) < e;
}
__device__ __forceinline__ uint32_t tile_of(uint32_t leaf) { return leaf / kRefitTile; }

// Refit, first kernel: one workgroup per tile of kRefitTile sorted leaves
// completes the tile's local nodes. A local node's index lies in its own
// range, so in the tile: the tile's topology (children, parent, range end,
// leaf parents) and leaf boxes are staged in LDS with coalesced loads, and
// both climbing threads of a local node belong to the workgroup, so its ticket
// and box live in LDS too. A climb stops where it
// would enter a node spanning tiles; the second kernel starts there. Every
// completed box goes to node_aabb with plain stores, which the next kernel
// sees.
__global__ void __launch_bounds__(256) k_refit_local(const float* __restrict__ aabbs,
                                                     const unsigned long long* __restrict__ keys, uint32_t n,
                                                     const uint32_t* nd, const int32_t* __restrict__ left,
                                                     const int32_t* __restrict__ right,
                                                     const int32_t* __restrict__ parent,
                                                     const int32_t* __restrict__ leaf_parent,
                                                     const int32_t* __restrict__ other_end,
                                                     float* __restrict__ node_aabb) {
    __shared__ __align__(16) float lbox[kRefitTile * 6];   // the tile's leaf boxes, sorted order
    __shared__ float nbox[kRefitTile * 6];    // boxes of the tile's local nodes (index - s)
    __shared__ uint32_t ltk[kRefitTile];      // tickets of the tile's local nodes
    __shared__ int32_t tl[kRefitTile], tr[kRefitTile], tp[kRefitTile], te[kRefitTile], tlp[kRefitTile];
    n = leaf_count(n, nd);
    const uint32_t s = blockIdx.x * kRefitTile;
    if (n < 2u || s >= n) return;             // one leaf: no internal node, no parent links
    const uint32_t e = min(s + kRefitTile, n);
    // staging: every load of a thread's slots is issued before any is waited
    // for. Indices are clamped into range so that the loads need no branch;
    // validity only selects what is stored. Leaf boxes are gathered through
    // the sorted keys, 8 bytes at a time (boxes are 24-byte records).
    constexpr uint32_t kSlots = kRefitTile / 256u;
    const uint32_t last_leaf = e - 1u, last_node = n - 2u;
    uint32_t box_of[kSlots];
#pragma unroll
    for (uint32_t u = 0; u < kSlots; ++u)
        box_of[u] = (uint32_t)(keys[min(s + threadIdx.x + u * 256u, last_leaf)] & 0xFFFFFFFFull);
    float2 gb[kSlots][3];
    int32_t vlp[kSlots], vl[kSlots], vr[kSlots], vp[kSlots], ve[kSlots];
#pragma unroll
    for (uint32_t u = 0; u < kSlots; ++u) {
        const uint32_t i = s + threadIdx.x + u * 256u;
        const float2* src = reinterpret_cast<const float2*>(aabbs + (size_t)box_of[u] * 6u);
        gb[u][0] = src[0];
        gb[u][1] = src[1];
        gb[u][2] = src[2];
        const uint32_t li = min(i, last_leaf), ni = min(i, last_node);
        vlp[u] = leaf_parent[li];
        vl[u] = left[ni];
        vr[u] = right[ni];
        vp[u] = parent[ni];
        ve[u] = other_end[ni];
    }
#pragma unroll
    for (uint32_t u = 0; u < kSlots; ++u) {
        const uint32_t k = threadIdx.x + u * 256u, i = s + k;
        const bool internal = i < e && i <= last_node;   // internal nodes 0 .. n-2
        ltk[k] = 0u;
        tlp[k] = vlp[u];
        tl[k] = vl[u];
        tr[k] = vr[u];
        tp[k] = vp[u];
        te[k] = internal ? ve[u] : -1;                   // -1: never local
        float2* d = reinterpret_cast<float2*>(&lbox[k * 6u]);
        d[0] = gb[u][0];
        d[1] = gb[u][1];
        d[2] = gb[u][2];
    }
    __syncthreads();
    for (uint32_t k = s + threadIdx.x; k < e; k += blockDim.x) {
        float b[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) b[q] = lbox[(k - s) * 6u + q];
        int32_t cur = (int32_t)k;             // leaf k, or the local node just completed
        bool cur_leaf = true;
        int32_t node = tlp[k - s];
        // a radix tree over unique 64-bit keys is at most 65 levels deep
        for (int guard = 0; node >= 0 && guard < 96; ++guard) {
            if ((uint32_t)node < s || (uint32_t)node >= e) break;   // outside the tile: spanning
            const uint32_t x = (uint32_t)node - s;
            const int32_t j = te[x];
            if (j < 0 || !is_local(node, j, s, e)) break;
            // only LDS needs ordering here: this thread's LDS box writes are
            // complete before its ticket, and the sibling's box is read after
            // it (a release at workgroup scope would also wait for the global
            // stores of node_aabb)
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            const uint32_t t = __hip_atomic_fetch_add(&ltk[x], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            asm volatile("" ::: "memory");
            if (t == 0u) break;               // the sibling's thread finishes this node
            const int32_t l = tl[x], r = tr[x];
            const bool mine_left = cur_leaf ? (l == cur) : (l == -cur);
            const int32_t sib = mine_left ? r : l;
            const float* src = sib >= 0 ? &lbox[(uint32_t)(sib - (int32_t)s) * 6u]
                                        : &nbox[(uint32_t)(-sib - (int32_t)s) * 6u];
            float o[6], u[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) o[q] = src[q];
            unite(b, o, mine_left, u);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                b[q] = u[q];
                nbox[x * 6u + q] = u[q];
                node_aabb[(size_t)node * 6u + q] = u[q];
            }
            cur = node;
            cur_leaf = false;
            if (node == 0) break;
            node = tp[x];
        }
    }
}

// the box of child c of a node spanning tiles: a leaf's from the input, a
// local node's as the first kernel left it
__device__ __forceinline__ void child_box(const float* aabbs, const unsigned long long* keys, const float* node_aabb,
                                          int32_t c, float o[6]) {
    const float* src = c >= 0 ? aabbs + (size_t)(keys[c] & 0xFFFFFFFFull) * 6u : node_aabb + (size_t)(-c) * 6u;
#pragma unroll
    for (int q = 0; q < 6; ++q) o[q] = src[q];
}

// Refit, second kernel: the nodes spanning tiles. One thread per internal
// node; a spanning node with a child that is a leaf or a local node (both
// complete) starts a climb there - with both children complete it unites them
// at once, with one it takes the node's ticket - and climbs as the reference
// does: the first arriver stops, the second unites and goes on. Boxes that
// climbs exchange are written and read with agent-scope atomics, each store
// complete before the next ticket.
__global__ void __launch_bounds__(256) k_refit_global(const float* __restrict__ aabbs,
                                                      const unsigned long long* __restrict__ keys, uint32_t n,
                                                      const uint32_t* nd, const int32_t* __restrict__ left,
                                                      const int32_t* __restrict__ right,
                                                      const int32_t* __restrict__ parent,
                                                      const int32_t* __restrict__ other_end, float* node_aabb,
                                                      uint32_t* ticket) {
    n = leaf_count(n, nd);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < 2u || i >= n - 1u) return;
    const int32_t node0 = (int32_t)i, j = other_end[i];
    const uint32_t lo = (uint32_t)min(node0, j), hi = (uint32_t)max(node0, j);
    if (tile_of(lo) == tile_of(hi)) return;   // local: done by the first kernel
    const int32_t l = left[i], r = right[i];
    // a child is complete when it is a leaf or a node inside one tile
    auto complete = [&](int32_t c) {
        if (c >= 0) return true;
        const int32_t jj = other_end[-c];
        return tile_of((uint32_t)min(-c, jj)) == tile_of((uint32_t)max(-c, jj));
    };
    const bool lc = complete(l), rc = complete(r);
    if (!lc && !rc) return;                   // both children arrive by climbs
    float b[6];
    int32_t cur, node = node0;
    if (lc && rc) {                           // both known: this node is complete now
        float lb[6], rb[6];
        child_box(aabbs, keys, node_aabb, l, lb);
        child_box(aabbs, keys, node_aabb, r, rb);
        unite(lb, rb, true, b);
        store_box(node_aabb + (size_t)node * 6u, b);
        if (node == 0) return;
        cur = -node;
        node = parent[node];
    } else {
        cur = lc ? l : r;                     // arrive at this node from the complete child
        child_box(aabbs, keys, node_aabb, cur, b);
    }
    for (int guard = 0; node >= 0 && guard < 96; ++guard) {
        // every box this thread wrote has completed before the ticket is taken
        __builtin_amdgcn_s_waitcnt(0);
        const uint32_t t = __hip_atomic_fetch_add(ticket + node, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t == 0u) return;                  // the sibling's thread finishes this node
        const int32_t nl = left[node], nr = right[node];
        const bool mine_left = (nl == cur);
        const int32_t sib = mine_left ? nr : nl;
        // a leaf's box from the input; a node's, written by the first kernel
        // or by a climb, with agent-scope loads (current for both)
        float o[6];
        if (sib >= 0) child_box(aabbs, keys, node_aabb, sib, o);
        else load_box(node_aabb + (size_t)(-sib) * 6u, o);
        float u[6];
        unite(b, o, mine_left, u);
#pragma unroll
        for (int q = 0; q < 6; ++q) b[q] = u[q];
        store_box(node_aabb + (size_t)node * 6u, b);
        if (node == 0) return;
        cur = -node;
        node = parent[node];
    }
}

// The tree in depth-first order with skip indices, for walks without a
// stack: node k is two float4, box min xyz, max x | max y, max z, the index
// after k's subtree, the box index of a leaf (0xFFFFFFFF: internal). The
// node whose sorted leaf range starts at a sits at 2a + (the number of its
// ancestors whose left subtree holds it): the subtrees to its left hold a
// leaves and a - (ancestors it is the right child of) internal nodes, and
// each ancestor comes before it. Its subtree spans 2 (leaves) - 1 nodes.
// Thread t < n: leaf t; n <= t < 2n - 1: internal node t - n.
__global__ void __launch_bounds__(256) k_flatten(const unsigned long long* __restrict__ keys, uint32_t n,
                                                 const uint32_t* nd, const int32_t* __restrict__ left,
                                                 const int32_t* __restrict__ parent,
                                                 const int32_t* __restrict__ leaf_parent,
                                                 const int32_t* __restrict__ other_end,
                                                 const float* __restrict__ node_aabb,
                                                 const float* __restrict__ aabbs, float4* __restrict__ flat) {
    const uint32_t nn = leaf_count(n, nd);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (nn == 0u || t >= 2u * nn - 1u) return;
    int32_t cur, par;
    uint32_t first, len, item;
    const float* box;
    if (t < nn) {
        cur = (int32_t)t;
        par = nn > 1u ? leaf_parent[t] : -1;
        first = t;
        len = 1u;
        item = (uint32_t)keys[t];
        box = aabbs + (size_t)item * 6u;
    } else {
        const int32_t j = (int32_t)(t - nn);
        const int32_t e = other_end[j];
        cur = -j;
        par = j == 0 ? -1 : parent[j];
        first = (uint32_t)min(j, e);
        len = 2u * ((uint32_t)max(j, e) - first + 1u) - 1u;
        item = 0xFFFFFFFFu;
        box = node_aabb + (size_t)j * 6u;
    }
    uint32_t lefts = 0;
    while (par >= 0) {                     // at most the tree's depth
        if (left[par] == cur) ++lefts;
        cur = -par;
        par = parent[par];
    }
    const uint32_t k = 2u * first + lefts;
    flat[2u * k] = make_float4(box[0], box[1], box[2], box[3]);
    flat[2u * k + 1u] = make_float4(box[4], box[5], __uint_as_float(k + len), __uint_as_float(item));
}

} // namespace bvhk

// ---- host API ----
struct pga_lbvh {
    uint32_t capacity = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t *code[2] = {nullptr, nullptr}, *idx[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;       // 256 x sort workgroups
    uint32_t* part = nullptr;       // scan partials
    uint32_t* ticket = nullptr;     // internal nodes
    int32_t* other_end = nullptr;   // internal nodes: the far end of each node's leaf range
    float* bounds = nullptr;        // bounds partials; [0, 6) the folded scene box
    uint32_t* dev_words = nullptr;  // [0] leaf count of the last gather, [1] its overflow flag
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string error;
};

namespace {

using namespace bvhk;

pga_status lfail(pga_lbvh* b, pga_status st, const std::string& m) {
    if (b) b->error = m;
    return st;
}

#define LHIP(b, expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return lfail(b, PGA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

uint32_t sort_blocks(uint32_t n) { return (n + kTile - 1u) / kTile; }

// Morton codes to refit for n leaves (nd: the device count, n its bound),
// with a host scene box (mn, ex) or the device box `box`.
hipError_t enqueue_build(pga_lbvh* b, const float* aabbs, uint32_t n, const uint32_t* nd, const float mn[3],
                         const float ex[3], const float* box, uint64_t* keys, int32_t* left, int32_t* right,
                         int32_t* parent, int32_t* leaf_parent, float* node_aabb) {
    hipStream_t s = b->stream;
    const uint32_t g = (n + 255u) / 256u;
    hipLaunchKernelGGL(k_morton, dim3(g), dim3(256), 0, s, aabbs, n, nd, mn[0], mn[1], mn[2], ex[0], ex[1], ex[2],
                       box, b->code[0], b->idx[0]);
    const uint32_t nb = sort_blocks(n);
    const uint32_t nh = 256u * nb;
    const uint32_t np = (nh + kScanTile - 1u) / kScanTile;
    int cur = 0;
    for (uint32_t pass = 0; pass < 4; ++pass) {     // 30-bit codes: digits at bits 0, 8, 16, 24
        const uint32_t shift = 8u * pass;
        hipLaunchKernelGGL(k_hist, dim3(nb), dim3(kThreads), 0, s, b->code[cur], n, nd, shift, b->hist, nb);
        hipLaunchKernelGGL(k_scan_reduce, dim3(np), dim3(kThreads), 0, s, b->hist, nh, b->part);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kThreads), 0, s, b->part, np);
        hipLaunchKernelGGL(k_scan_down, dim3(np), dim3(kThreads), 0, s, b->hist, nh, b->part);
        unsigned long long* keys_out = pass == 3 ? reinterpret_cast<unsigned long long*>(keys) : nullptr;
        hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kThreads), 0, s, b->code[cur], b->idx[cur], b->code[cur ^ 1],
                           b->idx[cur ^ 1], keys_out, n, nd, shift, b->hist, nb);
        cur ^= 1;
    }
    const unsigned long long* k64 = reinterpret_cast<const unsigned long long*>(keys);
    hipLaunchKernelGGL(k_karras, dim3(g), dim3(256), 0, s, k64, n, nd, left, right, parent, leaf_parent,
                       b->other_end, b->ticket);
    hipLaunchKernelGGL(k_refit_local, dim3((n + kRefitTile - 1u) / kRefitTile), dim3(256), 0, s, aabbs, k64, n, nd,
                       left, right, parent, leaf_parent, b->other_end, node_aabb);
    hipLaunchKernelGGL(k_refit_global, dim3(g), dim3(256), 0, s, aabbs, k64, n, nd, left, right, parent,
                       b->other_end, node_aabb, b->ticket);
    return hipGetLastError();
}

// ---- the allocation-free build (pga_lbvh_build_noalloc) ----
// Its sort and refit scratch is this file-scope device array: two code and
// two index arrays, the tickets and range ends of the internal nodes (4 bytes
// per leaf each), the digit histograms and the scan partials, for at most
// kNoAllocLeaves leaves (about 24.3 MiB, held by every GPU the library runs
// on until the process ends). Host code uses the address that
// hipGetSymbolAddress gives on each device, never the array's name.
constexpr uint32_t kNoAllocLeaves = PGA_LBVH_NOALLOC_MAX_LEAVES;
constexpr size_t kNoAllocHist = (size_t)256 * ((kNoAllocLeaves + kTile - 1u) / kTile);
constexpr size_t kNoAllocPart = (kNoAllocHist + kScanTile - 1u) / kScanTile + 1u;
constexpr size_t kNoAllocWords = 6u * (size_t)kNoAllocLeaves + kNoAllocHist + kNoAllocPart;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[kMaxDevices] = {nullptr};   // the array's address on each device
thread_local std::string g_noalloc_error;      // message of the last failed call on this thread

pga_status noalloc_fail(pga_status st, const std::string& m) {
    g_noalloc_error = m;
    return st;
}

// The array's address on the current device, looked up once per device.
pga_status scratch_on_device(uint32_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail(PGA_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return noalloc_fail(PGA_ERR_UNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return noalloc_fail(PGA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[dev];
    return PGA_OK;
}

} // namespace

extern "C" {

PGA_API pga_status pga_lbvh_build_noalloc(const float* aabbs, uint32_t n, const float scene_min[3],
                                          const float scene_max[3], uint64_t* keys, int32_t* left, int32_t* right,
                                          int32_t* parent, int32_t* leaf_parent, float* node_aabb, void* stream) {
    if (!aabbs || !scene_min || !scene_max || !keys || !left || !right || !parent || !leaf_parent || !node_aabb)
        return noalloc_fail(PGA_ERR_INVALID_ARG, "null argument");
    if (n < 2u || n > kNoAllocLeaves)
        return noalloc_fail(PGA_ERR_INVALID_ARG, "pga_lbvh_build_noalloc: need 2 <= n <= " +
                                                     std::to_string(kNoAllocLeaves) + " leaves, got " +
                                                     std::to_string(n));
    const float ex[3] = {scene_max[0] - scene_min[0], scene_max[1] - scene_min[1], scene_max[2] - scene_min[2]};
    if (!(ex[0] > 0.0f) || !(ex[1] > 0.0f) || !(ex[2] > 0.0f)) return noalloc_fail(PGA_ERR_INVALID_ARG, "empty scene box");
    uint32_t* base = nullptr;
    pga_status st = scratch_on_device(&base);
    if (st) return st;
    // a builder view of the array, sized for n leaves (no events, no allocation)
    pga_lbvh view;
    view.capacity = n;
    view.stream = (hipStream_t)stream;
    const size_t N = kNoAllocLeaves;
    view.code[0] = base;
    view.code[1] = base + N;
    view.idx[0] = base + 2u * N;
    view.idx[1] = base + 3u * N;
    view.ticket = base + 4u * N;
    view.other_end = reinterpret_cast<int32_t*>(base + 5u * N);
    view.hist = base + 6u * N;
    view.part = base + 6u * N + kNoAllocHist;
    const hipError_t e = enqueue_build(&view, aabbs, n, nullptr, scene_min, ex, nullptr, keys, left, right, parent,
                                       leaf_parent, node_aabb);
    if (e != hipSuccess) return noalloc_fail(PGA_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    return PGA_OK;
}

PGA_API pga_status pga_lbvh_create(uint32_t capacity, int32_t device, void* stream, pga_lbvh** out) {
    if (!out || capacity < 2u || capacity > 0x7FFFFFFFu) return PGA_ERR_INVALID_ARG;
    pga_lbvh* b = new pga_lbvh();
    *out = b;
    b->capacity = capacity;
    if (device >= 0) LHIP(b, hipSetDevice(device));
    LHIP(b, hipGetDevice(&b->device));
    b->stream = (hipStream_t)stream;
    for (int k = 0; k < 2; ++k) {
        LHIP(b, hipMalloc((void**)&b->code[k], (size_t)capacity * 4u));
        LHIP(b, hipMalloc((void**)&b->idx[k], (size_t)capacity * 4u));
    }
    const size_t nh = (size_t)256 * sort_blocks(capacity);
    LHIP(b, hipMalloc((void**)&b->hist, nh * 4u));
    LHIP(b, hipMalloc((void**)&b->part, ((nh + kScanTile - 1) / kScanTile + 1) * 4u));
    LHIP(b, hipMalloc((void**)&b->ticket, (size_t)capacity * 4u));
    LHIP(b, hipMalloc((void**)&b->other_end, (size_t)capacity * 4u));
    LHIP(b, hipMalloc((void**)&b->bounds, 1024u * 6u * 4u));
    LHIP(b, hipMalloc((void**)&b->dev_words, 64u));
    LHIP(b, hipMemset(b->dev_words, 0, 64u));
    LHIP(b, hipEventCreate(&b->e0));
    LHIP(b, hipEventCreate(&b->e1));
    return PGA_OK;
}

PGA_API void pga_lbvh_destroy(pga_lbvh* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    else (void)hipDeviceSynchronize();
    for (int k = 0; k < 2; ++k) { (void)hipFree(b->code[k]); (void)hipFree(b->idx[k]); }
    (void)hipFree(b->hist); (void)hipFree(b->part); (void)hipFree(b->ticket); (void)hipFree(b->bounds);
    (void)hipFree(b->dev_words); (void)hipFree(b->other_end);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    delete b;
}

PGA_API const char* pga_lbvh_last_error(const pga_lbvh* b) { return b ? b->error.c_str() : g_noalloc_error.c_str(); }

PGA_API pga_status pga_lbvh_instance_aabbs(pga_lbvh* b, const pga_instance* instances, uint32_t n, int32_t planar,
                                           float* aabbs) {
    if (!b || (n && (!instances || !aabbs))) return lfail(b, PGA_ERR_INVALID_ARG, "null buffer");
    if (n == 0) return PGA_OK;
    hipLaunchKernelGGL(k_instance_aabbs, dim3((n + 255u) / 256u), dim3(256), 0, b->stream, instances, n,
                       planar ? 1u : 0u, aabbs);
    LHIP(b, hipGetLastError());
    return PGA_OK;
}

PGA_API pga_status pga_lbvh_bounds(pga_lbvh* b, const float* aabbs, uint32_t n, float bounds[6]) {
    if (!b || !aabbs || !bounds || n == 0) return lfail(b, PGA_ERR_INVALID_ARG, "bad bounds arguments");
    const uint32_t blocks = std::min<uint32_t>(1024u, (n + 255u) / 256u);
    hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(256), 0, b->stream, aabbs, n, nullptr, b->bounds);
    hipLaunchKernelGGL(k_bounds_fold, dim3(1), dim3(256), 0, b->stream, b->bounds, blocks);
    LHIP(b, hipGetLastError());
    LHIP(b, hipMemcpyAsync(bounds, b->bounds, 6u * 4u, hipMemcpyDeviceToHost, b->stream));
    LHIP(b, hipStreamSynchronize(b->stream));
    return PGA_OK;
}

PGA_API pga_status pga_lbvh_build(pga_lbvh* b, const float* aabbs, uint32_t n, const float scene_min[3],
                                  const float scene_max[3], uint64_t* keys, int32_t* left, int32_t* right,
                                  int32_t* parent, int32_t* leaf_parent, float* node_aabb) {
    if (!b || !aabbs || !scene_min || !scene_max || !keys || !left || !right || !parent || !leaf_parent || !node_aabb)
        return lfail(b, PGA_ERR_INVALID_ARG, "null argument");
    if (n < 2u || n > b->capacity)
        return lfail(b, PGA_ERR_INVALID_ARG, "need 2 <= n <= capacity (" + std::to_string(b->capacity) + ")");
    const float ex[3] = {scene_max[0] - scene_min[0], scene_max[1] - scene_min[1], scene_max[2] - scene_min[2]};
    if (!(ex[0] > 0.0f) || !(ex[1] > 0.0f) || !(ex[2] > 0.0f)) return lfail(b, PGA_ERR_INVALID_ARG, "empty scene box");
    LHIP(b, hipEventRecord(b->e0, b->stream));
    LHIP(b, enqueue_build(b, aabbs, n, nullptr, scene_min, ex, nullptr, keys, left, right, parent, leaf_parent,
                          node_aabb));
    LHIP(b, hipEventRecord(b->e1, b->stream));
    LHIP(b, hipStreamSynchronize(b->stream));
    return PGA_OK;
}

PGA_API pga_status pga_lbvh_gather_instances(pga_lbvh* b, const pga_instance* const* buffers,
                                             const uint64_t* capacities, uint32_t n_types,
                                             const uint64_t* device_counts, uint32_t planar_mask, float* aabbs) {
    if (!b || !buffers || !capacities || !device_counts || !aabbs || n_types == 0 || n_types > PGA_MAX_TERMINALS)
        return lfail(b, PGA_ERR_INVALID_ARG, "bad gather arguments");
    GatherArgs g{};
    for (uint32_t t = 0; t < n_types; ++t) {
        if (!buffers[t] && capacities[t]) return lfail(b, PGA_ERR_INVALID_ARG, "null buffer with nonzero capacity");
        g.buf[t] = buffers[t];
        g.cap[t] = capacities[t];
    }
    g.ntypes = n_types;
    g.planar_mask = planar_mask;
    g.capacity = b->capacity;
    const uint32_t blocks = std::min<uint32_t>(kGatherBlocks, (b->capacity + 255u) / 256u);
    hipLaunchKernelGGL(k_gather_boxes, dim3(blocks), dim3(256), 0, b->stream, g,
                       reinterpret_cast<const unsigned long long*>(device_counts), aabbs, b->dev_words,
                       b->dev_words + 1);
    LHIP(b, hipGetLastError());
    return PGA_OK;
}

PGA_API pga_status pga_lbvh_build_device(pga_lbvh* b, const float* aabbs, const float* scene_min,
                                         const float* scene_max, uint64_t* keys, int32_t* left, int32_t* right,
                                         int32_t* parent, int32_t* leaf_parent, float* node_aabb) {
    if (!b || !aabbs || !keys || !left || !right || !parent || !leaf_parent || !node_aabb)
        return lfail(b, PGA_ERR_INVALID_ARG, "null argument");
    if (!scene_min != !scene_max) return lfail(b, PGA_ERR_INVALID_ARG, "give both scene corners or neither");
    const uint32_t n = b->capacity;
    const uint32_t* nd = b->dev_words;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LHIP(b, hipStreamIsCapturing(b->stream, &cs));
    const bool timed = cs == hipStreamCaptureStatusNone;
    if (timed) LHIP(b, hipEventRecord(b->e0, b->stream));
    if (scene_min) {
        const float ex[3] = {scene_max[0] - scene_min[0], scene_max[1] - scene_min[1], scene_max[2] - scene_min[2]};
        if (!(ex[0] > 0.0f) || !(ex[1] > 0.0f) || !(ex[2] > 0.0f))
            return lfail(b, PGA_ERR_INVALID_ARG, "empty scene box");
        LHIP(b, enqueue_build(b, aabbs, n, nd, scene_min, ex, nullptr, keys, left, right, parent, leaf_parent,
                              node_aabb));
    } else {
        const uint32_t blocks = std::min<uint32_t>(1024u, (n + 255u) / 256u);
        hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(256), 0, b->stream, aabbs, n, nd, b->bounds);
        hipLaunchKernelGGL(k_bounds_fold, dim3(1), dim3(256), 0, b->stream, b->bounds, blocks);
        const float zero[3] = {0.0f, 0.0f, 0.0f}, one[3] = {1.0f, 1.0f, 1.0f};
        LHIP(b, enqueue_build(b, aabbs, n, nd, zero, one, b->bounds, keys, left, right, parent, leaf_parent,
                              node_aabb));
    }
    if (timed) LHIP(b, hipEventRecord(b->e1, b->stream));
    return PGA_OK;
}

} // extern "C"

// The device-count build over a caller's count word (the colliders a
// multi-phase derivation stored, built between its phases by pga_api.cpp):
// min(*nd, capacity) leaves, the scene box from the device.
hipError_t pga_lbvh_enqueue_counted(pga_lbvh* b, const float* aabbs, const uint32_t* nd, uint64_t* keys,
                                    int32_t* left, int32_t* right, int32_t* parent, int32_t* leaf_parent,
                                    float* node_aabb, float4* flat) {
    const uint32_t n = b->capacity;
    const uint32_t blocks = std::min<uint32_t>(1024u, (n + 255u) / 256u);
    hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(256), 0, b->stream, aabbs, n, nd, b->bounds);
    hipLaunchKernelGGL(k_bounds_fold, dim3(1), dim3(256), 0, b->stream, b->bounds, blocks);
    const float zero[3] = {0.0f, 0.0f, 0.0f}, one[3] = {1.0f, 1.0f, 1.0f};
    const hipError_t e = enqueue_build(b, aabbs, n, nd, zero, one, b->bounds, keys, left, right, parent,
                                       leaf_parent, node_aabb);
    if (e != hipSuccess) return e;
    const uint32_t g = (uint32_t)(((uint64_t)2u * n + 255u) / 256u);
    hipLaunchKernelGGL(k_flatten, dim3(g), dim3(256), 0, b->stream, reinterpret_cast<const unsigned long long*>(keys),
                       n, nd, left, parent, leaf_parent, b->other_end, node_aabb, aabbs, flat);
    return hipGetLastError();
}

extern "C" {

PGA_API pga_status pga_lbvh_leaf_count(pga_lbvh* b, uint32_t* n, int32_t* overflow) {
    if (!b || !n) return lfail(b, PGA_ERR_INVALID_ARG, "null argument");
    uint32_t w[2] = {0u, 0u};
    LHIP(b, hipMemcpyAsync(w, b->dev_words, sizeof(w), hipMemcpyDeviceToHost, b->stream));
    LHIP(b, hipStreamSynchronize(b->stream));
    *n = w[0];
    if (overflow) *overflow = w[1] ? 1 : 0;
    return PGA_OK;
}

PGA_API pga_status pga_lbvh_seconds(pga_lbvh* b, double* seconds) {
    if (!b || !seconds) return PGA_ERR_INVALID_ARG;
    float ms = 0.0f;
    LHIP(b, hipEventElapsedTime(&ms, b->e0, b->e1));
    *seconds = ms * 1e-3;
    return PGA_OK;
}

} // extern "C"
