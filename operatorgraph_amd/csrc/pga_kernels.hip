// pga_kernels.hip - derivation kernels for gfx950 (wave64).
//
// One derivation level = every shape record of the current frontier run
// through its fused single-child chain, its fan-out children generated, each
// child settled through its own chain; children that reach another fan-out
// are appended to the next frontier, terminals are written to the instance
// buffers. Appends are aggregated per wave: one ballot, one popcount, one
// global atomic per destination, then per-lane slots from mbcnt.
//
// Two techniques share that level body:
//   k_level  - one launch per level, the host reads the next count
//              (the reference's KERNELS technique, techniqueKernels.cuh);
//   k_mega   - one cooperative persistent launch, levels separated by a grid
//              barrier (the MEGAKERNEL technique, techniqueMegakernel.cuh).
#include "pga_kernels.h"

namespace pga {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t mbcnt(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ void store_rec(Rec* dst, const Rec& r) {
    const float4* s = reinterpret_cast<const float4*>(&r);
    float4* d = reinterpret_cast<float4*>(dst);
#pragma unroll
    for (int i = 0; i < 5; ++i) d[i] = s[i];
}
__device__ __forceinline__ void load_rec(const Rec* src, Rec& r) {
    const float4* s = reinterpret_cast<const float4*>(src);
    float4* d = reinterpret_cast<float4*>(&r);
#pragma unroll
    for (int i = 0; i < 5; ++i) d[i] = s[i];
}

// Appends the records of the lanes with `has` to the next frontier.
// Called by all lanes of the wave (wave-uniform control flow).
__device__ __forceinline__ void emit_record(const Ctx& c, bool has, const Rec& r) {
    unsigned long long mask = __ballot(has);
    if (mask == 0ull) return;
    const uint32_t cnt = (uint32_t)__popcll(mask);
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(c.out_count, cnt);
    base = __shfl(base, leader);
    if (has) {
        uint32_t idx = base + mbcnt(mask);
        if (idx < c.out_cap) store_rec(c.out + idx, r);
        else raise(c, PGA_ERR_OVERFLOW_QUEUE);
    }
}

// GenerationFunction<GenFuncFilter, 1>::perShape (GenerationFunctions.cuh:
// 155-189): instance = model4 * scale(size), custom attribute. Slots come
// from one atomic per terminal type per wave; a full buffer is an error
// (the reference wraps modulo capacity). Capacity 0 counts only.
__device__ __forceinline__ void emit_terminal(const Ctx& c, bool has, const Rec& r) {
    unsigned long long pending = __ballot(has);
    if (pending == 0ull) return;
    uint32_t t = 0;
    if (has) {
        const DevEntry e = c.prog.entries[rec_entry(r)];
        t = f2u_rd(P(c, e, 0, r, r.axiom, r.path, rec_depth(r)));
        if (t >= PGA_MAX_TERMINALS) { raise(c, PGA_ERR_BAD_PROGRAM); t = PGA_MAX_TERMINALS - 1; }
    }
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t tl = __shfl(t, leader);
        const bool mine = has && t == tl;
        const unsigned long long m = __ballot(mine);
        unsigned long long base = 0;
        if ((int)lane_id() == leader) base = atomicAdd(&c.term_counts[tl], (unsigned long long)__popcll(m));
        base = __shfl(base, leader);
        if (mine) {
            unsigned long long idx = base + mbcnt(m);
            const unsigned long long cap = c.inst_cap[tl];
            if (idx < cap) {
                const float b[16] = {r.s[0], 0.0f, 0.0f, 0.0f, 0.0f, r.s[1], 0.0f, 0.0f,
                                     0.0f, 0.0f, r.s[2], 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
                float mm[16];
                mul34x44(r.m, b, mm);
                mm[12] = r.custom;
                mm[13] = __uint_as_float(r.axiom);
                mm[14] = __uint_as_float(r.path);
                mm[15] = __uint_as_float(r.ed);
                float4* d = reinterpret_cast<float4*>(c.inst[tl] + idx);
#pragma unroll
                for (int i = 0; i < 4; ++i) d[i] = make_float4(mm[4 * i], mm[4 * i + 1], mm[4 * i + 2], mm[4 * i + 3]);
            } else if (cap != 0ull) {
                raise(c, PGA_ERR_OVERFLOW_INSTANCES);
            }
        }
        pending &= ~m;
    }
}

__device__ __forceinline__ bool entry_is_leaf(const Ctx& c, const Rec& r) {
    return (c.prog.entries[rec_entry(r)].flags & EF_LEAF) != 0;
}

// Generates every child of the fan-out records of the active lanes, settles
// each child, and routes it. INLINE: leaf fan-out children are expanded in
// place (their children are all terminals), otherwise queued.
template <bool INLINE>
__device__ void expand(const Ctx& c, bool active, const Rec& r) {
    Fan f;
    uint32_t n = 0;
    if (active) {
        fan_init(c, r, f);
        n = f.n;
    }
    for (uint32_t j = 0; __ballot(j < n) != 0ull; ++j) {
        bool has = j < n;
        Rec ch;
        if (has) has = fan_child(c, r, f, j, ch);
        int cs = has ? settle(c, ch) : ST_DROP;
        emit_terminal(c, cs == ST_TERMINAL, ch);
        bool leaf = INLINE && cs == ST_FANOUT && entry_is_leaf(c, ch);
        emit_record(c, (cs == ST_FANOUT && !leaf) || cs == ST_OVERLONG, ch);
        if (INLINE) {
            if (__ballot(leaf) != 0ull) expand<false>(c, leaf, ch);
        }
    }
}

__device__ void process(const Ctx& c, bool valid, Rec r) {
    int st = valid ? settle(c, r) : ST_DROP;
    emit_terminal(c, st == ST_TERMINAL, r);
    emit_record(c, st == ST_OVERLONG, r);
    expand<true>(c, st == ST_FANOUT, r);
}

__device__ __forceinline__ Ctx make_ctx(const KArgs& a, uint32_t level) {
    Ctx c;
    c.prog = a.prog;
    c.aparams = a.aparams;
    c.aparams_per = a.aparams_per;
    c.naxioms_total = a.naxioms_total;
    c.key0 = a.key0;
    c.key1 = a.key1;
    c.inst = a.inst;
    c.inst_cap = a.inst_cap;
    c.term_counts = a.term_counts;
    c.out = ((level + 1u) & 1u) ? a.front1 : a.front0;
    c.out_count = a.level_counts + level + 1u;
    c.out_cap = a.cap;
    c.err = a.err;
    c.max_chain = a.max_chain;
    return c;
}

// One level over the frontier; waves take 64 consecutive records at a time.
__device__ void run_level(const KArgs& a, uint32_t level, uint32_t n, uint32_t wave, uint32_t nwaves) {
    const Ctx c = make_ctx(a, level);
    const Rec* in = (level & 1u) ? a.front1 : a.front0;
    for (uint32_t base = wave * 64u; base < n; base += nwaves * 64u) {
        const uint32_t i = base + lane_id();
        bool valid = i < n;
        Rec r;
        if (level == 0u) {
            if (valid) {
                const pga_axiom& ax = a.axioms[a.first_axiom + i];
                for (int k = 0; k < 12; ++k) r.m[k] = ax.model[k];
                r.s[0] = ax.size[0]; r.s[1] = ax.size[1]; r.s[2] = ax.size[2];
                r.seed = ax.seed;
                r.custom = ax.custom;
                r.axiom = a.first_axiom + i;
                r.path = 0u;
                const int32_t ent = ax.entry_index;
                if (ent < 0) {
                    valid = false;                  // PGAFacade.cu:708
                } else if ((uint32_t)ent >= a.prog.nentries ||
                           (int32_t)a.prog.entries[ent].shape != ax.shape_type) {
                    raise(c, PGA_ERR_BAD_PROGRAM);
                    valid = false;
                } else {
                    r.ed = (uint32_t)ent;
                }
            }
        } else if (valid) {
            load_rec(in + i, r);
        }
        process(c, valid, r);
    }
}

__global__ void __launch_bounds__(256) k_level(KArgs a, uint32_t level) {
    const uint32_t n = level == 0u ? a.n_axioms : min(a.level_counts[level], a.cap);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    run_level(a, level, n, wave, nwaves);
}

// Grid-wide barrier of a cooperative launch: arrive counter bar[0] and
// generation bar[1], agent-scope acquire/release.
__device__ __forceinline__ void grid_barrier(uint32_t* bar, uint32_t nblocks) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1u) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen)
                __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __threadfence();
}

__global__ void __launch_bounds__(256) k_mega(KArgs a) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    uint32_t level = 0;
    for (; level < a.max_levels; ++level) {
        uint32_t n;
        if (level == 0u) n = a.n_axioms;
        else n = min(__hip_atomic_load(&a.level_counts[level], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), a.cap);
        if (n == 0u) break;                     // every block reads the same count
        run_level(a, level, n, wave, nwaves);
        grid_barrier(a.bar, gridDim.x);
    }
    if (level == a.max_levels && blockIdx.x == 0 && threadIdx.x == 0 &&
        __hip_atomic_load(&a.level_counts[level], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
        atomicOr(a.err, 1u << PGA_ERR_MAX_DEPTH);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.level_counts[a.max_levels + 1u] = level;
}

// Zeroes the per-run counters (a kernel, so that the reset can sit in a
// captured graph).
__global__ void k_reset(uint32_t* level_counts, uint32_t n_levels, uint32_t* err, uint32_t* bar,
                        unsigned long long* term_counts, uint32_t reset_terms) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_levels) level_counts[i] = 0u;
    if (i == 0) { *err = 0u; bar[0] = 0u; bar[1] = 0u; }
    if (reset_terms && i < PGA_MAX_TERMINALS) term_counts[i] = 0ull;
}

} // namespace pga

// ---- launchers (host side) ----
extern "C++" {
hipError_t pga_launch_reset(const pga::KArgs& a, int reset_terms, hipStream_t s) {
    const uint32_t n = a.max_levels + 2u;
    const uint32_t blocks = (n + 255u) / 256u;
    hipLaunchKernelGGL(pga::k_reset, dim3(blocks), dim3(256), 0, s, a.level_counts, n, a.err, a.bar,
                       a.term_counts, (uint32_t)reset_terms);
    return hipGetLastError();
}

hipError_t pga_launch_level(const pga::KArgs& a, uint32_t level, uint32_t blocks, uint32_t block_size,
                            hipStream_t s) {
    hipLaunchKernelGGL(pga::k_level, dim3(blocks), dim3(block_size), 0, s, a, level);
    return hipGetLastError();
}

hipError_t pga_mega_occupancy(uint32_t block_size, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, pga::k_mega, (int)block_size, 0);
}

hipError_t pga_launch_mega(const pga::KArgs& a, uint32_t blocks, uint32_t block_size, hipStream_t s) {
    pga::KArgs copy = a;
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel((const void*)pga::k_mega, dim3(blocks), dim3(block_size), args, 0, s);
}
}
