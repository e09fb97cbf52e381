// pga_level.h - the level machinery of the derivation kernels, shared by the
// interpreting kernels (pga_kernels.hip) and the kernels specialised to one
// program at initialisation (pga_spec.h): the LDS layout of a level, the
// wave-level claims of output slots, terminal emission through LDS staging,
// fan-out starts, the program staged in LDS, the grid barrier and the scan
// between the count and write passes.
#pragma once

#include "pga_kernels.h"

namespace pga {

constexpr int kWaves = 16;                       // most waves per workgroup (1024 threads)
constexpr int kCounters = PGA_MAX_TERMINALS + 2; // terminal slots, parents, descriptors

struct LevelLDS {
    uint32_t cnt[kWaves][kCounters];   // count pass: per-wave totals
    uint32_t cur[kWaves][kCounters];   // write pass: per-wave cursors
    uint32_t before[kCounters];        // totals of the preceding blocks
    uint32_t level_total[kCounters];   // totals of the level over the grid
    unsigned long long term_base[PGA_MAX_TERMINALS];  // running instance count per slot
    uint32_t next_n;                   // children of the next level (all blocks agree)
    uint32_t grab;                     // workgroup-local derivation: first axiom handed to this block
    uint32_t spill;                    // ... 1: this level's fan-outs go to the grid's arrays
    uint32_t tp_prev, tp_cur;          // tape: parent-record base of the previous / this level
    uint32_t td_prev, td_cur;          // tape: descriptor base of the previous / this level
};

enum Mode : uint32_t { M_DONE = 0, M_SETTLE = 1, M_FAN = 2 };

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f4 gf4;   // a global-memory f4

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
// the count pass adds to the wave's total, the write pass returns the wave
// cursor (the lane's slot is cursor + mbcnt(mask)) and advances it.
__device__ __forceinline__ uint32_t claim(LevelLDS& L, bool wr, uint32_t k, unsigned long long mask, int leader) {
    const uint32_t w = wave_in_block();
    const uint32_t cnt = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (wr) {
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
// The lanes of one claim get consecutive records, so the write pass stages
// them in the wave's part of LDS (ST, 64 records) in rank order and writes
// them out as consecutive 16-byte nontemporal stores: a lane storing its own
// record would leave every store instruction a 64-byte stride, and partly
// written lines cost more than the coalesced stream (measured 0.91 -> 0.80 ms
// on 16,384 facade lots). The copy reads its (at most four) words per lane
// from LDS before any store, and stores through a global-memory pointer: a
// flat store would also count against the LDS wait of the next read.
// The instance record of a terminal shape as four 16-byte rows: model4 *
// scale(size) (rows 0-2), then custom attribute, axiom, path and entry |
// depth << 16.
__device__ __forceinline__ void instance_rows(const Rec& r, bool inv, f4& q0, f4& q1, f4& q2, f4& q3) {
    const float b[16] = {r.s[0], 0.0f, 0.0f, 0.0f, 0.0f, r.s[1], 0.0f, 0.0f,
                         0.0f, 0.0f, r.s[2], 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    float mm[12];
    if (inv) {
        // an inverted polygon: getAdjustedModel (DynamicPolygon.cuh:68-75)
        // negates the first two columns
        float am[12];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            am[i * 4 + 0] = -r.m[i * 4 + 0];
            am[i * 4 + 1] = -r.m[i * 4 + 1];
            am[i * 4 + 2] = r.m[i * 4 + 2];
            am[i * 4 + 3] = r.m[i * 4 + 3];
        }
        mul34x44(am, b, mm);
    } else {
        mul34x44(r.m, b, mm);
    }
    q0 = f4{mm[0], mm[1], mm[2], mm[3]};
    q1 = f4{mm[4], mm[5], mm[6], mm[7]};
    q2 = f4{mm[8], mm[9], mm[10], mm[11]};
    q3 = f4{r.custom, __uint_as_float(r.axiom), __uint_as_float(r.path), __uint_as_float(r.ed & ~kInvert)};
}

__device__ __forceinline__ void emit_terminal(const Ctx& c, LevelLDS& L, f4* ST, bool wr, bool has, uint32_t slot,
                                              const Rec& r, unsigned long long& idx) {
    f4 q0 = {}, q1 = {}, q2 = {}, q3 = {};
    // c.poly is a constant 0 in kernels specialised to a program without
    // polygons: the inverted form is compiled out there
    if (wr && has) instance_rows(r, c.poly && (r.ed & kInvert) != 0u, q0, q1, q2, q3);
    unsigned long long pending = __ballot(has);
    idx = 0;
    while (pending) {                          // one claim per distinct slot in the wave
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t sl = __shfl(slot, leader);
        const bool mine = has && slot == sl;
        const unsigned long long m = __ballot(mine);
        const uint32_t base = claim(L, wr, sl, m, leader);
        const uint32_t rank = mbcnt(m);
        const unsigned long long first = L.term_base[sl] + base;
        if (mine) idx = first + rank;
        if (wr) {
            // the group's records are consecutive: stage them in rank order
            // and write them out as consecutive 16-byte stores
            if (mine) {
                ST[rank * 4u + 0u] = q0;
                ST[rank * 4u + 1u] = q1;
                ST[rank * 4u + 2u] = q2;
                ST[rank * 4u + 3u] = q3;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t t = c.prog.slot_term[sl];
            const unsigned long long cap = c.inst_cap[t];
            const uint32_t n4 = (uint32_t)__popcll(m) * 4u;
            gf4* dst = (gf4*)(c.inst[t] + first);
            f4 v[4];
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                const uint32_t q = lane_id() + 64u * u;
                v[u] = q < n4 ? ST[q] : f4{};
            }
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                const uint32_t q = lane_id() + 64u * u;
                if (q < n4 && first + (q >> 2) < cap) __builtin_nontemporal_store(v[u], dst + q);
            }
            if (mine && idx >= cap && cap != 0ull) raise(c, PGA_ERR_OVERFLOW_INSTANCES);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        pending &= ~m;
    }
}

// ---- direct leaf fan-outs in the write pass ----
// A direct fan-out (EF_DIRECT: every child is a Generate of a constant
// terminal, or a Discard) does not take each child round the state machine:
// the wave's direct fan-outs emit their children in lock step, child j of
// every such lane in round j (direct_child builds it), each round one
// coalesced emit_terminal - the order in which the state machine would have
// emitted them. The count pass takes the totals from the entry's table
// (count_direct).
template <class E>
__device__ __forceinline__ bool direct_child(const Ctx& c, const Rec& par, const E& e, Fan& f, uint32_t skipm,
                                             uint32_t j, Rec& ch, uint32_t& slot) {
    float o = 0.0f, ns = 0.0f;
    const bool keep = fan_geom(c, par, e, f, j, o, ns);
    const bool skip = j < 32u ? ((skipm >> j) & 1u) != 0u : skipm == 0xFFFFFFFFu;
    if (!keep || skip || !build_child(c, par, e, j, f.axis, o, ns, ch)) return false;
    slot = c.prog.entries[rec_entry(ch)].tslot;
    return true;
}

// Counting pass over a direct leaf fan-out: its children's terminals per
// slot come from the entry's table (times the repetitions of a Repeat).
template <class E>
__device__ __forceinline__ void count_direct(const Ctx& c, LevelLDS& L, const E& e, const Fan& f) {
    const uint32_t mult = e.kind() == K_REPEAT ? f.n : (f.n != 0u ? 1u : 0u);
    const uint32_t tab = entry_aux(c, e, 1);
    const uint32_t w = wave_in_block();
    for (uint32_t k = 0; k < e.ndirect(); ++k) {
        const uint32_t slot = (tab >> (16u * k)) & 0xFFu;
        const unsigned long long add = (unsigned long long)((tab >> (16u * k + 8u)) & 0xFFu) * mult;
        if (add >= 0x80000000ull) { raise(c, PGA_ERR_OVERFLOW_INSTANCES); continue; }
        atomicAdd(&L.cnt[w][slot], (uint32_t)add);
    }
}

// Starts the non-leaf fan-outs of the active lanes: claims and writes each
// parent record (pidx = its index, ~0u when the queue is full) and claims
// `n` consecutive descriptor slots from `dslot` on (unless !claim_desc; the
// count pass counts them either way).
__device__ __forceinline__ void start_fanout(const Ctx& c, LevelLDS& L, bool wr, bool active, const Rec& par,
                                             uint32_t n, uint32_t& pidx, uint32_t& dslot, bool claim_desc = true) {
    const uint32_t kP = c.prog.nslots, kD = c.prog.nslots + 1u;
    const unsigned long long pm = __ballot(active);
    const int pl = __ffsll((long long)pm) - 1;
    const uint32_t pbase = claim(L, wr, kP, pm, pl);
    if (wr && active) {
        pidx = pbase + mbcnt(pm);
        if (pidx < c.par_cap) store_rec(c.par_out + pidx, par);
        else { raise(c, PGA_ERR_OVERFLOW_QUEUE); pidx = ~0u; }
    }
    const uint32_t nn = active ? n : 0u;
    uint32_t incl = nn;       // inclusive prefix of n over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d);
        if ((int)lane_id() >= d) incl += v;
    }
    const uint32_t total = __shfl(incl, 63);
    const uint32_t w = wave_in_block();
    if (!wr) {
        if (lane_id() == 0) L.cnt[w][kD] += total;
        return;
    }
    if (!claim_desc) return;        // the caller claims its descriptors child by child
    const uint32_t dbase = L.cur[w][kD];
    if (lane_id() == 0) L.cur[w][kD] = dbase + total;
    if (active) dslot = dbase + incl - nn;
}

// Program image and terminal tables staged in LDS (dynamic shared memory).
extern __shared__ __align__(16) unsigned char pga_smem[];

// Copies the program image and the terminal buffer tables into LDS; every
// operator then reads its entry, parameters and successors from LDS.
__device__ __forceinline__ void stage_program(const KArgs& a) {
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

// CS: the kernel variant that runs context-sensitive programs (Collider,
// IfCollides, phases); the others see those fields as constants and drop
// the code.
template <bool TAPE, bool CS = false>
__device__ __forceinline__ Ctx make_ctx(const KArgs& a, const LevelLDS& L, uint32_t level, bool wr) {
    Ctx c;
    c.prog.entries = reinterpret_cast<const DevEntry*>(pga_smem);
    c.prog.params = reinterpret_cast<const DevParam*>(pga_smem + a.off_params);
    c.prog.succ = reinterpret_cast<const int32_t*>(pga_smem + a.off_succ);
    c.prog.exp = reinterpret_cast<const uint32_t*>(pga_smem + a.off_exp);
    c.prog.term_slot = pga_smem + a.off_slots;
    c.prog.slot_term = pga_smem + a.off_slots + PGA_MAX_TERMINALS;
    c.prog.nentries = a.prog.nentries;
    c.prog.nslots = a.prog.nslots;
    c.inst = reinterpret_cast<pga_instance* const*>(pga_smem + a.image_bytes);
    c.inst_cap = reinterpret_cast<const unsigned long long*>(pga_smem + a.image_bytes + 8u * PGA_MAX_TERMINALS);
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
    c.tape = TAPE ? 1u : 0u;
    c.prog_lds = 1u;
    if (TAPE) {        // this level's outputs follow the earlier levels' in the tape
        c.par_out = a.tape_par + L.tp_cur;
        c.desc_out = a.tape_desc + L.td_cur;
        c.par_cap = a.cap_par > L.tp_cur ? a.cap_par - L.tp_cur : 0u;
        c.desc_cap = a.cap_desc > L.td_cur ? a.cap_desc - L.td_cur : 0u;
    }
    c.err = a.err;
    c.max_chain = a.max_chain;
    c.geom = wr ? 1u : a.count_geom;
    c.seed = a.use_seed;
    c.poly = a.polygons;
    c.ctxs = CS ? 1u : 0u;
    c.dmask = CS ? a.dmask : 0xFFFFu;
    c.phase = CS ? a.phase : 0u;
    c.wr = wr ? 1u : 0u;
    c.coll = CS ? a.coll : nullptr;
    c.coll_count = CS ? a.coll_count : nullptr;
    c.coll_visible = CS ? a.coll_visible : nullptr;
    c.coll_cap = CS ? a.coll_cap : 0u;
    c.bvh_flat = CS ? a.bvh_flat : nullptr;
    c.coll_bits = CS ? a.coll_bits : nullptr;
    c.coll_bits_cap = CS ? a.coll_bits_cap : 0u;
    return c;
}

// The record of each lane's current fan-out parent lives in LDS (five
// 16-byte words per lane, word k of all lanes contiguous), not in registers:
// the state machine then fits more waves per SIMD.
__device__ __forceinline__ void par_store(float4* PW, const Rec& r) {
    float4* p = PW + threadIdx.x;
    const uint32_t s = blockDim.x;
    p[0] = make_float4(r.m[0], r.m[1], r.m[2], r.m[3]);
    p[s] = make_float4(r.m[4], r.m[5], r.m[6], r.m[7]);
    p[2u * s] = make_float4(r.m[8], r.m[9], r.m[10], r.m[11]);
    p[3u * s] = make_float4(r.s[0], r.s[1], r.s[2], r.seed);
    p[4u * s] = make_float4(r.custom, __uint_as_float(r.ed), __uint_as_float(r.axiom), __uint_as_float(r.path));
}
__device__ __forceinline__ Rec par_load(const float4* PW) {
    const float4* p = PW + threadIdx.x;
    const uint32_t s = blockDim.x;
    const float4 a = p[0], b = p[s], c = p[2u * s], d = p[3u * s], e = p[4u * s];
    Rec r;
    r.m[0] = a.x; r.m[1] = a.y; r.m[2] = a.z; r.m[3] = a.w;
    r.m[4] = b.x; r.m[5] = b.y; r.m[6] = b.z; r.m[7] = b.w;
    r.m[8] = c.x; r.m[9] = c.y; r.m[10] = c.z; r.m[11] = c.w;
    r.s[0] = d.x; r.s[1] = d.y; r.s[2] = d.z; r.seed = d.w;
    r.custom = e.x; r.ed = __float_as_uint(e.y); r.axiom = __float_as_uint(e.z); r.path = __float_as_uint(e.w);
    return r;
}

// Calls f(entry) with entry i read from the program tables at run time (the
// interpreting kernels); the specialised kernels pass their own visitor,
// which hands f the entry as a type of compile-time constants.
struct LoadEntry {
    template <class F>
    __device__ __forceinline__ void operator()(const Ctx& c, uint32_t i, F&& f) const { f(load_entry(c, i)); }
};

// Child i of a level: an axiom (level 0: axiom first + i) or what a
// descriptor names - the child of a parent below `in_cap`, or an axiom
// (kDescAxiom). Returns false when there is none.
// The record of axiom k. Returns false for an axiom without an entry.
__device__ __forceinline__ bool load_axiom(const KArgs& a, const Ctx& c, uint32_t k, bool wr, Rec& r) {
    const pga_axiom& ax = a.axioms[k];
#pragma unroll
    for (int q = 0; q < 12; ++q) r.m[q] = ax.model[q];
    r.s[0] = ax.size[0]; r.s[1] = ax.size[1]; r.s[2] = ax.size[2];
    r.seed = ax.seed;
    r.custom = ax.custom;
    r.axiom = k;
    r.path = 0u;
    r.ed = 0u;
    const int32_t ent = ax.entry_index;
    if (ent < 0) return false;              // PGAFacade.cu:708
    if ((uint32_t)ent >= c.prog.nentries || (int32_t)entry_shape(c, (uint32_t)ent) != ax.shape_type) {
        if (wr) raise(c, PGA_ERR_BAD_PROGRAM);
        return false;
    }
    r.ed = (uint32_t)ent;
    return true;
}

template <class Visit>
__device__ __forceinline__ bool load_child_at(const KArgs& a, const Ctx& c, uint32_t first, const Rec* par_in,
                                              uint32_t in_cap, const Desc* desc_in, bool wr, uint32_t level,
                                              uint32_t i, Rec& r, Visit visit) {
    if (level == 0u) return load_axiom(a, c, first + i, wr, r);
    const uint4 dv = *reinterpret_cast<const uint4*>(desc_in + i);
    const uint32_t parent = dv.x, jx = dv.y;
    const uint32_t j = jx & 0x3FFFFFFFu;
    if (j == kDescAxiom) return load_axiom(a, c, parent, wr, r);
    if (parent >= in_cap) return false;          // parent was dropped on overflow
    if (j == kDescNone) return false;
    Rec p;
    load_rec(par_in + parent, p);
    if (j == kDescSelf) { r = p; return true; }
    bool ok = false;
    visit(c, rec_entry(p), [&](const auto& e) {
        ok = build_child(c, p, e, j, jx >> 30, __uint_as_float(dv.z), __uint_as_float(dv.w), r);
    });
    return ok;
}

// The same over the launch's axioms and the grid's parent arrays.
template <class Visit = LoadEntry>
__device__ __forceinline__ bool load_child(const KArgs& a, const Ctx& c, const Rec* par_in, const Desc* desc_in,
                                           bool wr, uint32_t level, uint32_t i, Rec& r, Visit visit = Visit()) {
    return load_child_at(a, c, a.first_axiom, par_in, a.cap_par, desc_in, wr, level, i, r, visit);
}

// Grid-wide barrier of a cooperative launch. Words (64 bytes apart):
// bar[0] generation, bar[16 (g + 1)] arrivals of block group g = block % 8,
// bar[16 * 9] arrivals of the groups. Spreading the arrivals over nine
// counters keeps same-address atomics short; every counter is back at 0 and
// the generation one higher after each barrier. One agent-scope release
// fence per block before arriving, relaxed polling, one acquire fence after.
__device__ __forceinline__ void grid_barrier(uint32_t* bar, uint32_t nblocks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen = __hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t g = blockIdx.x & 7u;
        const uint32_t in_group = (nblocks - g + 7u) >> 3;          // blocks b < nblocks with b % 8 == g
        const uint32_t groups = nblocks < 8u ? nblocks : 8u;
        bool last = false;
        uint32_t* gc = &bar[16u * (g + 1u)];
        if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == in_group - 1u) {
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t* top = &bar[16u * 9u];
            last = __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == groups - 1u;
            if (last) __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (last) {
            __hip_atomic_store(&bar[0], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen)
                __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Phase timestamp of workgroup 0 (instrumentation; off when ticks is NULL).
__device__ __forceinline__ void stamp(const KArgs& a, uint32_t level, uint32_t phase) {
    if (a.ticks != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t i = level * 4u + phase;
        if (i < a.ticks_cap) a.ticks[i] = wall_clock64();
    }
}

// Between the passes: publish this block's totals, grid barrier, exclusive
// prefix over the preceding blocks into the write cursors, level totals
// (timestamps in row `row`).
__device__ __forceinline__ void scan_counts(const KArgs& a, LevelLDS& L, uint32_t level, uint32_t row) {
    const uint32_t nslots = a.prog.nslots, ncnt = nslots + 2u;
    stamp(a, row, 1);
    const uint32_t nw = blockDim.x >> 6;
    uint32_t* totals = a.block_totals + (size_t)blockIdx.x * ncnt;
    for (uint32_t k = threadIdx.x; k < ncnt; k += blockDim.x) {
        uint32_t s = 0;
        for (uint32_t ww = 0; ww < nw; ++ww) {
            L.cur[ww][k] = s;
            s += L.cnt[ww][k];
        }
        totals[k] = s;
    }
    grid_barrier(a.bar, gridDim.x);
    // wave w sums counters k = w, w + nw, ... over the blocks, 64 blocks a step
    for (uint32_t k = wave_in_block(); k < ncnt; k += nw) {
        uint32_t before = 0, all = 0;
        for (uint32_t b = lane_id(); b < gridDim.x; b += 64u) {
            const uint32_t v = a.block_totals[(size_t)b * ncnt + k];
            all += v;
            before += b < blockIdx.x ? v : 0u;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            before += __shfl_xor(before, d);
            all += __shfl_xor(all, d);
        }
        if (lane_id() == 0) {
            L.before[k] = before;
            L.level_total[k] = all;
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < ncnt; k += blockDim.x) {
        const uint32_t before = L.before[k], all = L.level_total[k];
        for (uint32_t ww = 0; ww < nw; ++ww) L.cur[ww][k] += before;
        if (k == nslots + 1u) L.next_n = all;
        if (blockIdx.x == 0 && k == nslots) a.par_counts[level] = all;
        if (blockIdx.x == 0 && k == nslots + 1u) a.desc_counts[level] = all;
    }
    __syncthreads();
    stamp(a, row, 2);
}

__device__ __forceinline__ void init_bases(const KArgs& a, LevelLDS& L) {
    if (threadIdx.x == 0) {
        for (uint32_t k = 0; k < a.prog.nslots; ++k)
            L.term_base[k] = a.term_counts[pga_smem[a.off_slots + PGA_MAX_TERMINALS + k]];
        L.tp_prev = L.tp_cur = L.td_prev = L.td_cur = 0u;
    }
    __syncthreads();
}

__device__ __forceinline__ void store_bases(const KArgs& a, LevelLDS& L) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (uint32_t k = 0; k < a.prog.nslots; ++k)
            a.term_counts[pga_smem[a.off_slots + PGA_MAX_TERMINALS + k]] = L.term_base[k];
}

__device__ __forceinline__ uint32_t first_level_size(const KArgs& a, uint32_t level) {
    if (level == 0u) return a.n_axioms;
    return min(a.desc_counts[level - 1u], a.cap_desc);
}

} // namespace pga
