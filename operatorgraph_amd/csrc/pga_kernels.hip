// pga_kernels.hip - derivation kernels for gfx950 (wave64).
//
// A derivation level processes the level's children, one per lane: the
// child is rebuilt from its parent record and 16-byte descriptor, runs its
// fused single-child chain, and then either
//   - writes a terminal instance,
//   - expands a leaf fan-out in place (all its children are terminals), or
//   - becomes a parent of the next level: its record is written once and one
//     descriptor per child (with the sequential running offsets computed
//     here) is appended.
// Level 0's children are the axioms.
//
// Output slots come from exact prefix sums, not atomics: a count pass
// accumulates per-wave totals in LDS (terminals per terminal slot, parents,
// descriptors), blocks publish their totals, a grid barrier, each block
// scans the totals of the blocks before it, and a write pass recomputes the
// children and writes at exact offsets. Same-address global atomics sustain
// only ~1e8/s on this device, so per-wave atomics would bound the kernel;
// the scan also makes the output order deterministic.
//
// Two techniques share that level body:
//   k_level  - one cooperative launch per level, the host reads the next
//              count (the reference's KERNELS technique, techniqueKernels.cuh);
//   k_mega   - one cooperative persistent launch for all levels (the
//              MEGAKERNEL technique, techniqueMegakernel.cuh).
#include "pga_kernels.h"

namespace pga {

constexpr int kWaves = 4;                        // waves per workgroup (256 threads)
constexpr int kCounters = PGA_MAX_TERMINALS + 2; // terminal slots, parents, descriptors

struct LevelLDS {
    uint32_t cnt[kWaves][kCounters];   // count pass: per-wave totals
    uint32_t cur[kWaves][kCounters];   // write pass: per-wave cursors
    unsigned long long term_base[PGA_MAX_TERMINALS];  // running instance count per slot
    uint32_t red[kWaves][2];
    uint32_t next_n;                   // children of the next level (all blocks agree)
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t wave_in_block() { return threadIdx.x >> 6; }

__device__ __forceinline__ uint32_t mbcnt(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ void store_rec(Rec* dst, const Rec& r) {
    float4* d = reinterpret_cast<float4*>(dst);
    d[0] = make_float4(r.m[0], r.m[1], r.m[2], r.m[3]);
    d[1] = make_float4(r.m[4], r.m[5], r.m[6], r.m[7]);
    d[2] = make_float4(r.m[8], r.m[9], r.m[10], r.m[11]);
    d[3] = make_float4(r.s[0], r.s[1], r.s[2], r.seed);
    d[4] = make_float4(r.custom, __uint_as_float(r.ed), __uint_as_float(r.axiom), __uint_as_float(r.path));
}
__device__ __forceinline__ void load_rec(const Rec* src, Rec& r) {
    const float4* s = reinterpret_cast<const float4*>(src);
    float4 a = s[0], b = s[1], c = s[2], d = s[3], e = s[4];
    r.m[0] = a.x; r.m[1] = a.y; r.m[2] = a.z; r.m[3] = a.w;
    r.m[4] = b.x; r.m[5] = b.y; r.m[6] = b.z; r.m[7] = b.w;
    r.m[8] = c.x; r.m[9] = c.y; r.m[10] = c.z; r.m[11] = c.w;
    r.s[0] = d.x; r.s[1] = d.y; r.s[2] = d.z; r.seed = d.w;
    r.custom = e.x; r.ed = __float_as_uint(e.y); r.axiom = __float_as_uint(e.z); r.path = __float_as_uint(e.w);
}

// Claims popc(mask) consecutive slots of counter k for the lanes in `mask`:
// count pass adds to the wave's total, write pass returns the wave cursor
// (the lane's slot is cursor + mbcnt(mask)) and advances it.
template <bool WRITE>
__device__ __forceinline__ uint32_t claim(LevelLDS& L, uint32_t k, unsigned long long mask, int leader) {
    const uint32_t w = wave_in_block();
    const uint32_t cnt = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (WRITE) {
        base = L.cur[w][k];
        if ((int)lane_id() == leader) L.cur[w][k] = base + cnt;
    } else {
        if ((int)lane_id() == leader) L.cnt[w][k] += cnt;
    }
    return base;
}

// GenerationFunction<GenFuncFilter, 1>::perShape (GenerationFunctions.cuh:
// 155-189): instance = model4 * scale(size), custom attribute. A full buffer
// is an error (the reference wraps modulo capacity); capacity 0 counts only.
template <bool WRITE>
__device__ __forceinline__ void emit_terminal(const Ctx& c, LevelLDS& L, bool has, const Rec& r) {
    unsigned long long pending = __ballot(has);
    if (pending == 0ull) return;
    uint32_t slot = 0;
    if (has) {
        const DevEntry e = c.prog.entries[rec_entry(r)];
        slot = e.tslot;
        if (slot == 0xFFu) {                   // terminal index computed per shape
            const uint32_t t = f2u_rd(P(c, e, 0, r, r.axiom, r.path, rec_depth(r)));
            slot = t < PGA_MAX_TERMINALS ? c.prog.term_slot[t] : 0xFFu;
        }
        if (slot >= c.prog.nslots) { raise(c, PGA_ERR_BAD_PROGRAM); slot = 0; }
    }
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t sl = __shfl(slot, leader);
        const bool mine = has && slot == sl;
        const unsigned long long m = __ballot(mine);
        const uint32_t base = claim<WRITE>(L, sl, m, leader);
        if (WRITE && mine) {
            const unsigned long long idx = L.term_base[sl] + base + mbcnt(m);
            const uint32_t t = c.prog.slot_term[sl];
            const unsigned long long cap = c.inst_cap[t];
            if (idx < cap) {
                const float b[16] = {r.s[0], 0.0f, 0.0f, 0.0f, 0.0f, r.s[1], 0.0f, 0.0f,
                                     0.0f, 0.0f, r.s[2], 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
                float mm[12];
                mul34x44(r.m, b, mm);
                float4* d = reinterpret_cast<float4*>(c.inst[t] + idx);
                d[0] = make_float4(mm[0], mm[1], mm[2], mm[3]);
                d[1] = make_float4(mm[4], mm[5], mm[6], mm[7]);
                d[2] = make_float4(mm[8], mm[9], mm[10], mm[11]);
                d[3] = make_float4(r.custom, __uint_as_float(r.axiom), __uint_as_float(r.path),
                                   __uint_as_float(r.ed));
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

// Writes the parent (fan-out) records of the active lanes and one
// descriptor per child; `self` continues an over-long chain on the next
// level (one descriptor that stands for the record itself).
template <bool WRITE>
__device__ __forceinline__ void emit_fanout(const Ctx& c, LevelLDS& L, bool active, bool self, const Rec& r) {
    const uint32_t kP = c.prog.nslots, kD = c.prog.nslots + 1u;
    const unsigned long long pm = __ballot(active);
    if (pm == 0ull) return;
    const int pl = __ffsll((long long)pm) - 1;
    const uint32_t pbase = claim<WRITE>(L, kP, pm, pl);
    uint32_t pidx = ~0u;
    if (WRITE && active) {
        pidx = pbase + mbcnt(pm);
        if (pidx < c.par_cap) store_rec(c.par_out + pidx, r);
        else { raise(c, PGA_ERR_OVERFLOW_QUEUE); pidx = ~0u; }
    }
    Fan f;
    uint32_t n = 0;
    if (active) {
        if (self) {
            n = 1;
        } else {
            fan_init(c, r, f);
            n = f.n;
        }
    }
    // exclusive prefix of n over the wave
    uint32_t incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t v = __shfl_up(incl, d);
        if ((int)lane_id() >= d) incl += v;
    }
    const uint32_t total = __shfl(incl, 63);
    const uint32_t w = wave_in_block();
    if (!WRITE) {
        if (lane_id() == 0) L.cnt[w][kD] += total;
        return;
    }
    const uint32_t dbase = L.cur[w][kD];
    if (lane_id() == 0) L.cur[w][kD] = dbase + total;
    uint32_t slot = dbase + incl - n;
    for (uint32_t j = 0; j < n; ++j, ++slot) {
        uint32_t jx;
        float o = 0.0f, ns = 0.0f;
        if (pidx == ~0u) {
            jx = kDescNone;                      // parent dropped on overflow: no gap
        } else if (self) {
            jx = kDescSelf;
        } else {
            const bool keep = fan_geom(c, r, f, j, o, ns);
            jx = keep ? (j | (f.axis << 30)) : kDescNone;
        }
        if (slot < c.desc_cap)
            *reinterpret_cast<float4*>(c.desc_out + slot) =
                make_float4(__uint_as_float(pidx), __uint_as_float(jx), o, ns);
        else
            raise(c, PGA_ERR_OVERFLOW_QUEUE);
    }
}

// Expands the leaf fan-outs of the active lanes in place: every child is a
// terminal or a discard after its fused chain.
template <bool WRITE>
__device__ void expand_leaf(const Ctx& c, LevelLDS& L, bool active, const Rec& r) {
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
        emit_terminal<WRITE>(c, L, cs == ST_TERMINAL, ch);
        // the load-time leaf analysis bounds every chain below a leaf by the
        // chain limit and excludes fan-outs there
        if (WRITE && (cs == ST_FANOUT || cs == ST_OVERLONG)) raise(c, PGA_ERR_BAD_PROGRAM);
    }
}

// One child per lane: settle, then route.
template <bool WRITE>
__device__ void process_child(const Ctx& c, LevelLDS& L, bool valid, Rec r) {
    int st = valid ? settle(c, r) : ST_DROP;
    emit_terminal<WRITE>(c, L, st == ST_TERMINAL, r);
    const bool fan = st == ST_FANOUT;
    const bool leaf = fan && entry_is_leaf(c, r);
    if (__ballot(leaf) != 0ull) expand_leaf<WRITE>(c, L, leaf, r);
    const bool queue = (fan && !leaf) || st == ST_OVERLONG;
    if (__ballot(queue) != 0ull) emit_fanout<WRITE>(c, L, queue, st == ST_OVERLONG, r);
}

// Program image and terminal tables staged in LDS (dynamic shared memory).
extern __shared__ __align__(16) unsigned char pga_smem[];

// Copies the program image and the terminal buffer tables into LDS when
// they fit (lds_bytes != 0); every operator then reads its entry,
// parameters and successors from LDS instead of the cache hierarchy.
__device__ __forceinline__ void stage_program(const KArgs& a) {
    if (a.lds_bytes == 0u) return;
    uint4* dst = reinterpret_cast<uint4*>(pga_smem);
    const uint32_t n16 = a.image_bytes / 16u;
    for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = a.image[i];
    unsigned long long* tab = reinterpret_cast<unsigned long long*>(pga_smem + a.image_bytes);
    for (uint32_t i = threadIdx.x; i < PGA_MAX_TERMINALS; i += blockDim.x) {
        tab[i] = reinterpret_cast<unsigned long long>(a.inst[i]);
        tab[PGA_MAX_TERMINALS + i] = a.inst_cap[i];
    }
    __syncthreads();
}

__device__ __forceinline__ Ctx make_ctx(const KArgs& a, uint32_t level) {
    Ctx c;
    c.prog = a.prog;
    c.inst = a.inst;
    c.inst_cap = a.inst_cap;
    if (a.lds_bytes != 0u) {
        c.prog.entries = reinterpret_cast<const DevEntry*>(pga_smem);
        c.prog.params = reinterpret_cast<const DevParam*>(pga_smem + a.off_params);
        c.prog.succ = reinterpret_cast<const int32_t*>(pga_smem + a.off_succ);
        c.prog.exp = reinterpret_cast<const uint32_t*>(pga_smem + a.off_exp);
        c.prog.term_slot = pga_smem + a.off_slots;
        c.prog.slot_term = pga_smem + a.off_slots + PGA_MAX_TERMINALS;
        c.inst = reinterpret_cast<pga_instance* const*>(pga_smem + a.image_bytes);
        c.inst_cap = reinterpret_cast<const unsigned long long*>(pga_smem + a.image_bytes + 8u * PGA_MAX_TERMINALS);
    }
    c.aparams = a.aparams;
    c.aparams_per = a.aparams_per;
    c.naxioms_total = a.naxioms_total;
    c.key0 = a.key0;
    c.key1 = a.key1;
    c.term_counts = a.term_counts;
    c.par_out = (level & 1u) ? a.par1 : a.par0;
    c.desc_out = (level & 1u) ? a.desc1 : a.desc0;
    c.par_count = a.par_counts + level;
    c.desc_count = a.desc_counts + level;
    c.par_cap = a.cap_par;
    c.desc_cap = a.cap_desc;
    c.err = a.err;
    c.max_chain = a.max_chain;
    return c;
}

// One pass over the children of a level; waves take 64 consecutive
// children at a time, in a fixed order, so both passes see the same lanes.
template <bool WRITE>
__device__ void run_pass(const KArgs& a, LevelLDS& L, uint32_t level, uint32_t n, uint32_t wave, uint32_t nwaves) {
    const Ctx c = make_ctx(a, level);
    const Rec* par_in = (level & 1u) ? a.par0 : a.par1;      // written by level - 1
    const Desc* desc_in = (level & 1u) ? a.desc0 : a.desc1;
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
                r.ed = 0u;
                const int32_t ent = ax.entry_index;
                if (ent < 0) {
                    valid = false;                  // PGAFacade.cu:708
                } else if ((uint32_t)ent >= a.prog.nentries ||
                           (int32_t)a.prog.entries[ent].shape != ax.shape_type) {
                    if (WRITE) raise(c, PGA_ERR_BAD_PROGRAM);
                    valid = false;
                } else {
                    r.ed = (uint32_t)ent;
                }
            }
        } else if (valid) {
            const float4 dv = *reinterpret_cast<const float4*>(desc_in + i);
            Desc d;
            d.parent = __float_as_uint(dv.x); d.jx = __float_as_uint(dv.y); d.o = dv.z; d.ns = dv.w;
            if (d.parent >= a.cap_par) {
                valid = false;                      // parent was dropped on overflow
            } else {
                Rec p;
                load_rec(par_in + d.parent, p);
                valid = desc_child(c, p, d, r);
            }
        }
        process_child<WRITE>(c, L, valid, r);
    }
}

// Grid-wide barrier of a cooperative launch: arrive counter bar[0] and
// generation bar[1]. The words return to (0, g + 1) after every barrier.
// One agent-scope release fence per block before arriving, relaxed
// polling, one acquire fence after.
__device__ __forceinline__ void grid_barrier(uint32_t* bar, uint32_t nblocks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1u) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen)
                __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Sum over the workgroup of one value per thread (all threads call it).
__device__ __forceinline__ uint32_t block_sum(LevelLDS& L, uint32_t v, int which) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    if (lane_id() == 0) L.red[wave_in_block()][which] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s += L.red[w][which];
    __syncthreads();
    return s;
}

// Phase timestamp of workgroup 0 (instrumentation; off when ticks is NULL).
__device__ __forceinline__ void stamp(const KArgs& a, uint32_t level, uint32_t phase) {
    if (a.ticks != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t i = level * 4u + phase;
        if (i < a.ticks_cap) a.ticks[i] = wall_clock64();
    }
}

// One derivation level: count pass, publish, barrier, scan, write pass.
// Leaves L.next_n = children of the next level; block 0 records the level's
// parent and descriptor counts.
__device__ void run_level(const KArgs& a, LevelLDS& L, uint32_t level, uint32_t n) {
    const uint32_t nslots = a.prog.nslots, ncnt = nslots + 2u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t w = wave_in_block();
    stamp(a, level, 0);
    for (uint32_t k = lane_id(); k < ncnt; k += 64u) L.cnt[w][k] = 0u;
    run_pass<false>(a, L, level, n, wave, nwaves);
    __syncthreads();
    stamp(a, level, 1);
    // block totals; in-block wave offsets go to the cursors
    uint32_t* totals = a.block_totals + (size_t)blockIdx.x * ncnt;
    for (uint32_t k = threadIdx.x; k < ncnt; k += blockDim.x) {
        uint32_t s = 0;
        for (int ww = 0; ww < kWaves; ++ww) {
            L.cur[ww][k] = s;
            s += L.cnt[ww][k];
        }
        totals[k] = s;
    }
    grid_barrier(a.bar, gridDim.x);
    // exclusive prefix over the blocks before this one, and the grand total
    for (uint32_t k = 0; k < ncnt; ++k) {
        uint32_t before = 0, all = 0;
        for (uint32_t b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
            const uint32_t v = a.block_totals[(size_t)b * ncnt + k];
            all += v;
            before += b < blockIdx.x ? v : 0u;
        }
        before = block_sum(L, before, 0);
        all = block_sum(L, all, 1);
        if (threadIdx.x < kWaves) L.cur[threadIdx.x][k] += before;
        if (threadIdx.x == 0) {
            if (k == nslots + 1u) L.next_n = all;
            if (blockIdx.x == 0 && k >= nslots) {
                if (k == nslots) a.par_counts[level] = all;
                else a.desc_counts[level] = all;
            }
        }
        __syncthreads();
        if (k < nslots && threadIdx.x == 0) {
            // terminals accumulate across levels: this level writes after
            // the running base, which then grows by the level total
            L.cnt[0][k] = all;   // stash for the base update after the write pass
        }
    }
    __syncthreads();
    stamp(a, level, 2);
    run_pass<true>(a, L, level, n, wave, nwaves);
    __syncthreads();
    stamp(a, level, 3);
    if (threadIdx.x == 0)
        for (uint32_t k = 0; k < nslots; ++k) L.term_base[k] += L.cnt[0][k];
    __syncthreads();
}

__device__ __forceinline__ void init_bases(const KArgs& a, LevelLDS& L) {
    if (threadIdx.x == 0)
        for (uint32_t k = 0; k < a.prog.nslots; ++k) L.term_base[k] = a.term_counts[a.prog.slot_term[k]];
    __syncthreads();
}

__device__ __forceinline__ void store_bases(const KArgs& a, LevelLDS& L) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (uint32_t k = 0; k < a.prog.nslots; ++k) a.term_counts[a.prog.slot_term[k]] = L.term_base[k];
}

__device__ __forceinline__ uint32_t first_level_size(const KArgs& a, uint32_t level) {
    if (level == 0u) return a.n_axioms;
    return min(a.desc_counts[level - 1u], a.cap_desc);
}

// KERNELS technique: one level per (cooperative) launch.
__global__ void __launch_bounds__(256) k_level(KArgs a, uint32_t level) {
    __shared__ LevelLDS L;
    stage_program(a);
    init_bases(a, L);
    const uint32_t n = first_level_size(a, level);
    if (n != 0u) run_level(a, L, level, n);
    // every block has read the old bases before the barrier inside run_level
    store_bases(a, L);
}

// MEGAKERNEL technique: all levels in one persistent launch.
__global__ void __launch_bounds__(256) k_mega(KArgs a) {
    __shared__ LevelLDS L;
    stage_program(a);
    init_bases(a, L);
    uint32_t level = 0;
    uint32_t n = a.n_axioms;
    for (; level < a.max_levels && n != 0u; ++level) {
        run_level(a, L, level, n);
        n = min(L.next_n, a.cap_desc);
        grid_barrier(a.bar, gridDim.x);       // next level reads this level's records
    }
    store_bases(a, L);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (n != 0u) atomicOr(a.err, 1u << PGA_ERR_MAX_DEPTH);
        a.desc_counts[a.max_levels + 1u] = level;
    }
}

// Zeroes the per-run counters (a kernel, so that the reset can sit in a
// captured graph).
__global__ void k_reset(uint32_t* par_counts, uint32_t* desc_counts, uint32_t n_levels, uint32_t* err, uint32_t* bar,
                        unsigned long long* term_counts, uint32_t reset_terms) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_levels) { par_counts[i] = 0u; desc_counts[i] = 0u; }
    if (i == 0) { *err = 0u; bar[0] = 0u; bar[1] = 0u; }
    if (reset_terms && i < PGA_MAX_TERMINALS) term_counts[i] = 0ull;
}

} // namespace pga

// ---- launchers (host side) ----
hipError_t pga_launch_reset(const pga::KArgs& a, int reset_terms, hipStream_t s) {
    const uint32_t n = a.max_levels + 2u;
    const uint32_t blocks = (n + 255u) / 256u;
    hipLaunchKernelGGL(pga::k_reset, dim3(blocks), dim3(256), 0, s, a.par_counts, a.desc_counts, n, a.err, a.bar,
                       a.term_counts, (uint32_t)reset_terms);
    return hipGetLastError();
}

hipError_t pga_launch_level(const pga::KArgs& a, uint32_t level, uint32_t blocks, uint32_t block_size,
                            hipStream_t s) {
    pga::KArgs copy = a;
    void* args[] = {&copy, &level};
    return hipLaunchCooperativeKernel((const void*)pga::k_level, dim3(blocks), dim3(block_size), args,
                                      a.lds_bytes, s);
}

hipError_t pga_mega_occupancy(uint32_t block_size, uint32_t lds_bytes, int* blocks_per_cu) {
    int a = 0, b = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, pga::k_mega, (int)block_size, lds_bytes);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, pga::k_level, (int)block_size, lds_bytes);
    *blocks_per_cu = a < b ? a : b;
    return e;
}

hipError_t pga_launch_mega(const pga::KArgs& a, uint32_t blocks, uint32_t block_size, hipStream_t s) {
    pga::KArgs copy = a;
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel((const void*)pga::k_mega, dim3(blocks), dim3(block_size), args,
                                      a.lds_bytes, s);
}
