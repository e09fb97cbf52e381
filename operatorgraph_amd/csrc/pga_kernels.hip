// pga_kernels.hip - derivation kernels for gfx950 (wave64).
//
// A derivation level processes the level's children, one per lane. Each
// lane runs a small state machine over its child, one operator per step:
//   SETTLE - apply the single-child operator of the current record, or
//            finish it: a Generate emits a terminal instance, a Discard
//            drops it, a fan-out starts FAN;
//   FAN    - child j of the fan-out parent: a leaf fan-out (all of whose
//            children end in terminals) builds the child and settles it in
//            place; any other fan-out writes its parent record once and one
//            16-byte descriptor per child (with the sequential running
//            offsets computed here) for the next level.
// Every operator body exists once in the code, and the two passes of a
// level (below) share it, so the kernel stays small enough for the
// instruction cache and the register file (four waves per SIMD).
// Level 0's children are the axioms.
//
// Output slots come from exact prefix sums, not atomics: a count pass
// accumulates per-wave totals in LDS (terminals per terminal slot, parents,
// descriptors), blocks publish their totals, a grid barrier, each block
// scans the totals of the blocks before it, and a write pass recomputes the
// children and writes at exact offsets. Same-address global atomics sustain
// only ~1e8/s on this device, so per-wave atomics would bound the kernel;
// the scan also makes the output order deterministic. The count pass skips
// model-matrix arithmetic when no decision of the program reads the matrix.
//
// Two techniques share that level body:
//   k_level  - one cooperative launch per level, the host reads the next
//              count (the reference's KERNELS technique, techniqueKernels.cuh);
//   k_mega   - one cooperative persistent launch for all levels (the
//              MEGAKERNEL technique, techniqueMegakernel.cuh).
#include <algorithm>

#include "pga_level.h"

namespace pga {

// One pass over the children of a level; waves take 64 consecutive
// children at a time, in a fixed order, so both passes see the same lanes.
template <bool TAPE, bool CS>
__device__ __forceinline__ void run_pass(const KArgs& a, const Ctx& c, LevelLDS& L, float4* PW, bool wr,
                                         uint32_t level, uint32_t n, uint32_t wave, uint32_t nwaves) {
    // inputs: written by level - 1 (ping-pong buffers, or that level's part of the tape)
    const Rec* par_in = TAPE ? a.tape_par + L.tp_prev : ((level & 1u) ? a.par0 : a.par1);
    const Desc* desc_in = TAPE ? a.tape_desc + L.td_prev : ((level & 1u) ? a.desc0 : a.desc1);
    const uint32_t child_base = level == 0u ? 0u : a.n_axioms + L.td_prev;   // tape outcomes
#pragma unroll 1
    for (uint32_t base = wave * 64u; base < n; base += nwaves * 64u) {
        const uint32_t i = base + lane_id();
        Rec cur;
        bool valid = i < n && load_child(a, c, par_in, desc_in, wr, level, i, cur);
        // an axiom whose entry reaches no phase of this launch adds nothing
        if (CS && valid && level == 0u && multi_phase(c) && c.phase != 0u &&
            c.prog.entries[rec_entry(cur)].aux[1] < c.phase)
            valid = false;
        CollMemo memo;                  // context-sensitive kernels: IfCollides outcomes, count -> write pass
        memo.bits = 0u;
        memo.n = 0u;
        memo.mode = (CS && valid && i < c.coll_bits_cap) ? (wr ? 2u : 1u) : 0u;
        if (CS && memo.mode == 2u) memo.bits = c.coll_bits[i];
        uint32_t out_kind = T_DROP, out_a = 0, out_b = 0, out_c = 0;   // tape outcome
        uint32_t mode = valid ? M_SETTLE : M_DONE;
        bool leaf = false, self = false;
        uint32_t steps = 0, j = 0, dslot = 0, pidx = ~0u;
        uint32_t skipm = 0, nd = 0;   // children that go to a Discard; descriptors to write
        Fan f;
        f.n = 0; f.axis = 0; f.total = 0.0f; f.r2a = 0.0f; f.off = 0.0f;
        f.e1 = f.e2 = f.e3 = 0.0f; f.prev_ns = 0.0f; f.prev_o = 0.0f; f.reps = 0;
#pragma unroll 1
        while (__ballot(mode != M_DONE) != 0ull) {
            bool emit = false, start = false;
            uint32_t slot = 0;
            if (mode == M_FAN) {
                if (j < f.n) {
                    float o = 0.0f, ns = 0.0f;
                    const Rec par = par_load(PW);
                    const EntryV pe = load_entry(c, rec_entry(par));
                    const bool keep = self ? true : fan_geom(c, par, pe, f, j, o, ns);
                    const bool skip = j < 32u ? ((skipm >> j) & 1u) != 0u : skipm == 0xFFFFFFFFu;
                    if (leaf) {
                        if (keep && !skip && build_child(c, par, pe, j, f.axis, o, ns, cur)) {
                            mode = M_SETTLE;
                            steps = 0;
                        }
                    } else if (!skip) {   // write pass only: the count pass claimed nd slots at the start
                        const uint32_t jx = pidx == ~0u ? kDescNone
                                          : self        ? kDescSelf
                                          : keep        ? (j | (f.axis << 30))
                                                        : kDescNone;
                        if (dslot < c.desc_cap)
                            *reinterpret_cast<uint4*>(c.desc_out + dslot) =
                                make_uint4(pidx, jx, __float_as_uint(o), __float_as_uint(ns));
                        else
                            raise(c, PGA_ERR_OVERFLOW_QUEUE);
                        ++dslot;
                    }
                    ++j;
                } else {
                    mode = M_DONE;
                }
            }
            if (mode == M_SETTLE) {
                const EntryV e = load_entry(c, rec_entry(cur));
                const uint32_t k = e.kind();
                const uint32_t back = leaf ? M_FAN : M_DONE;
                if (k == K_GENERATE && rec_phase(c, cur) != c.phase) {
                    mode = back;                    // emitted by the launch of its own phase
                } else if (k == K_GENERATE) {
                    slot = e.tslot();
                    if (slot == 0xFFu) {            // terminal index computed per shape
                        const uint32_t t = f2u_rd(P(c, e, 0, cur, cur.axiom, cur.path, rec_depth(c, cur)));
                        slot = t < PGA_MAX_TERMINALS ? c.prog.term_slot[t] : 0xFFu;
                    }
                    if (slot >= c.prog.nslots) { raise(c, PGA_ERR_BAD_PROGRAM); slot = 0; }
                    emit = true;
                    mode = back;
                } else if (k == K_DISCARD) {
                    mode = back;
                } else if (is_fanout(k)) {
                    if (leaf) {
                        // the load-time leaf analysis excludes fan-outs below a leaf
                        raise(c, PGA_ERR_BAD_PROGRAM);
                        mode = M_FAN;
                    } else {
                        par_store(PW, cur);
                        fan_init(c, cur, e, f);
                        j = 0;
                        leaf = !TAPE && (e.flags() & EF_LEAF) != 0;
                        start = !leaf;
                        const uint32_t sk = entry_aux(c, e, 0);
                        if (k == K_REPEAT) {
                            skipm = (sk & 1u) ? 0xFFFFFFFFu : 0u;
                            nd = skipm ? 0u : f.n;
                        } else {
                            skipm = sk;
                            nd = f.n - (uint32_t)__popc(skipm & (f.n < 32u ? (1u << f.n) - 1u : 0xFFFFFFFFu));
                        }
                        if (!wr && leaf && (e.flags() & EF_DIRECT)) {
                            count_direct(c, L, e, f);
                            mode = M_DONE;
                        } else {
                            mode = (leaf || wr) ? M_FAN : M_DONE;
                        }
                    }
                } else if (steps >= c.max_chain) {
                    if (leaf) {
                        // the leaf analysis bounds every chain below a leaf by the limit
                        raise(c, PGA_ERR_BAD_PROGRAM);
                        mode = M_FAN;
                    } else {
                        // over-long chain: continues on the next level
                        par_store(PW, cur);
                        self = true;
                        f.n = 1;
                        nd = 1;
                        skipm = 0;
                        j = 0;
                        start = true;
                        mode = wr ? M_FAN : M_DONE;
                    }
                } else {
                    uint32_t ord;
                    const int32_t next = apply_single(c, e, cur, &ord, CS ? &memo : nullptr);
                    if (!advance(c, cur, next, ord)) mode = back;
                    ++steps;
                }
            }
            if (__ballot(emit) != 0ull) {
                unsigned long long idx;
                emit_terminal(c, L, reinterpret_cast<f4*>(PW + 5u * blockDim.x) + wave_in_block() * 256u, wr, emit, slot,
                              cur, idx);
                if (TAPE && emit) { out_kind = T_TERM | (slot << 8); out_a = (uint32_t)idx; }
            }
            if (__ballot(start) != 0ull) {
                start_fanout(c, L, wr, start, par_load(PW), nd, pidx, dslot);
                if (TAPE && start) { out_kind = self ? T_SELF : T_FAN; out_a = pidx; out_b = dslot; out_c = nd; }
            }
        }
        if (CS && memo.mode == 1u && memo.n != 0u) c.coll_bits[i] = memo.bits;
        if (TAPE && wr && i < n) {
            if (child_base + i < a.tape_out_cap)
                a.tape_out[child_base + i] = make_uint4(out_kind, out_a, out_b, out_c);
            else
                raise(c, PGA_ERR_OVERFLOW_QUEUE);
        }
    }
}

// One derivation level: count pass, scan, write pass. Leaves L.next_n =
// children of the next level; block 0 records the level's parent and
// descriptor counts.
template <bool TAPE, bool CS>
__device__ __forceinline__ void run_level(const KArgs& a, LevelLDS& L, float4* PW, uint32_t level, uint32_t n) {
    const uint32_t nslots = a.prog.nslots, ncnt = nslots + 2u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t w = wave_in_block();
    stamp(a, level, 0);
    for (uint32_t k = lane_id(); k < ncnt; k += 64u) L.cnt[w][k] = 0u;
#pragma unroll 1
    for (uint32_t pass = 0; pass < 2u; ++pass) {
        const bool wr = pass != 0u;
        const Ctx c = make_ctx<TAPE, CS>(a, L, level, wr);
        run_pass<TAPE, CS>(a, c, L, PW, wr, level, n, wave, nwaves);
        __syncthreads();
        if (!wr) scan_counts(a, L, level, level);
    }
    stamp(a, level, 3);
    // terminals accumulate across levels: this level wrote after the running
    // base, which now grows by the level total
    if (threadIdx.x == 0)
        for (uint32_t k = 0; k < nslots; ++k) L.term_base[k] += L.level_total[k];
    __syncthreads();
}

// KERNELS technique: one level per (cooperative) launch. BS = threads per
// workgroup; the variants differ in the registers the compiler may use
// (BS / 256 waves per SIMD at one workgroup per compute unit).
template <int BS, bool CS>
__global__ void __launch_bounds__(BS) k_level(KArgs a, uint32_t level) {
    __shared__ LevelLDS L;
    __shared__ float4 PW[9 * BS];   // parent records (5 words per lane), then instance staging (4)
    stage_program(a);
    init_bases(a, L);
    const uint32_t n = first_level_size(a, level);
    if (n != 0u) run_level<false, CS>(a, L, PW, level, n);
    // every block has read the old bases before the barrier inside run_level
    store_bases(a, L);
}

// MEGAKERNEL technique: all levels in one persistent launch. TAPE: record
// the derivation tape (reverse mode), one build only.
template <int BS, bool TAPE, bool CS>
__global__ void __launch_bounds__(BS) k_mega(KArgs a) {
    __shared__ LevelLDS L;
    __shared__ float4 PW[9 * BS];   // parent records (5 words per lane), then instance staging (4)
    stage_program(a);
    init_bases(a, L);
    uint32_t level = 0;
    uint32_t n = a.n_axioms;
    for (; level < a.max_levels && n != 0u; ++level) {
        run_level<TAPE, CS>(a, L, PW, level, n);
        n = min(L.next_n, TAPE ? (a.cap_desc > L.td_cur ? a.cap_desc - L.td_cur : 0u) : a.cap_desc);
        if (TAPE && threadIdx.x == 0) {       // the next level's inputs are this level's outputs
            const uint32_t ns = a.prog.nslots;
            L.tp_prev = L.tp_cur;
            L.tp_cur += L.level_total[ns];
            L.td_prev = L.td_cur;
            L.td_cur += L.level_total[ns + 1u];
        }
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
// flags bit 0: zero the terminal counts; bit 1: a later phase of a
// multi-phase run (the error word stays, the colliders stored so far become
// visible). A fresh run zeroes the collider counters.
__global__ void k_reset(uint32_t* par_counts, uint32_t* desc_counts, uint32_t n_levels, uint32_t* err, uint32_t* bar,
                        unsigned long long* term_counts, uint32_t flags, uint32_t* coll_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_levels) { par_counts[i] = 0u; desc_counts[i] = 0u; }
    if (i == 0 && !(flags & 2u)) *err = 0u;
    if (i == 0 && coll_count) {           // [0] appended, [1] visible to this launch
        if (flags & 2u) coll_count[1] = coll_count[0];
        else coll_count[0] = coll_count[1] = 0u;
    }
    if (i < kBarWords) bar[i] = 0u;
    if ((flags & 1u) && i < PGA_MAX_TERMINALS) term_counts[i] = 0ull;
}

// Boxes (shape_box) of the colliders in the store, the leaves of the BVH
// built between phases.
__global__ void __launch_bounds__(256) k_collider_boxes(const float4* __restrict__ coll, uint32_t cap,
                                                        const uint32_t* __restrict__ count, float* __restrict__ aabbs) {
    const uint32_t n = min(*count, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4* k = coll + 5u * i;
        const float4 w0 = k[0], w1 = k[1], w2 = k[2], w3 = k[3];
        const float m[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
        const float sz[3] = {w3.x, w3.y, w3.z};
        float lo[3], hi[3];
        shape_box(m, sz, lo, hi);
        float* o = aabbs + (size_t)i * 6u;
        o[0] = lo[0]; o[1] = lo[1]; o[2] = lo[2];
        o[3] = hi[0]; o[4] = hi[1]; o[5] = hi[2];
    }
}

} // namespace pga

// ---- launchers (host side) ----
hipError_t pga_launch_collider_boxes(const float4* coll, uint32_t cap, const uint32_t* count, float* aabbs,
                                     hipStream_t s) {
    const uint32_t blocks = std::max(1u, std::min(1024u, (cap + 255u) / 256u));
    hipLaunchKernelGGL(pga::k_collider_boxes, dim3(blocks), dim3(256), 0, s, coll, cap, count, aabbs);
    return hipGetLastError();
}

hipError_t pga_launch_reset(const pga::KArgs& a, int flags, hipStream_t s) {
    const uint32_t n = a.max_levels + 2u;
    const uint32_t blocks = (n + 255u) / 256u;
    hipLaunchKernelGGL(pga::k_reset, dim3(blocks), dim3(256), 0, s, a.par_counts, a.desc_counts, n, a.err, a.bar,
                       a.term_counts, (uint32_t)flags, a.coll_count);
    return hipGetLastError();
}

namespace {

template <class F>
bool pick_variant(uint32_t block_size, bool cs, F&& f) {
    if (cs) {
        switch (block_size) {
        case 256: f(pga::k_mega<256, false, true>, pga::k_level<256, true>); return true;
        case 512: f(pga::k_mega<512, false, true>, pga::k_level<512, true>); return true;
        case 768: f(pga::k_mega<768, false, true>, pga::k_level<768, true>); return true;
        case 1024: f(pga::k_mega<1024, false, true>, pga::k_level<1024, true>); return true;
        default: return false;
        }
    }
    switch (block_size) {
    case 256: f(pga::k_mega<256, false, false>, pga::k_level<256, false>); return true;
    case 512: f(pga::k_mega<512, false, false>, pga::k_level<512, false>); return true;
    case 768: f(pga::k_mega<768, false, false>, pga::k_level<768, false>); return true;
    case 1024: f(pga::k_mega<1024, false, false>, pga::k_level<1024, false>); return true;
    default: return false;
    }
}

// the variant for a launch: the context-sensitive one when the program
// stores colliders or has phases
bool context_sensitive(const pga::KArgs& a) { return a.coll != nullptr || a.dmask != 0xFFFFu; }

} // namespace

bool pga_block_size_supported(uint32_t block_size) {
    return pick_variant(block_size, false, [](auto, auto) {});
}

bool pga_needs_coresidency(const void* fn) {
    bool hit = fn == (const void*)pga::k_mega<pga::kTapeBlock, true, false>;
    for (uint32_t bs : {256u, 512u, 768u, 1024u})
        for (bool cs : {false, true})
            pick_variant(bs, cs, [&](auto mega, auto lvl) {
                hit = hit || fn == (const void*)mega || fn == (const void*)lvl;
            });
    return hit;
}

hipError_t pga_launch_level(const pga::KArgs& a, uint32_t level, uint32_t blocks, uint32_t block_size,
                            hipStream_t s) {
    pga::KArgs copy = a;
    void* args[] = {&copy, &level};
    const void* fn = nullptr;
    if (!pick_variant(block_size, context_sensitive(a), [&](auto mega, auto lvl) { (void)mega; fn = (const void*)lvl; }))
        return hipErrorInvalidValue;
    return hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(block_size), args, a.lds_bytes, s);
}

hipError_t pga_mega_occupancy(uint32_t block_size, uint32_t lds_bytes, bool context_sensitive,
                              int* blocks_per_cu) {
    hipError_t e = hipErrorInvalidValue;
    pick_variant(block_size, context_sensitive, [&](auto mega, auto lvl) {
        int a = 0, b = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, mega, (int)block_size, lds_bytes);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, lvl, (int)block_size, lds_bytes);
        *blocks_per_cu = a < b ? a : b;
    });
    return e;
}

hipError_t pga_launch_mega(const pga::KArgs& a, uint32_t blocks, uint32_t block_size, hipStream_t s) {
    pga::KArgs copy = a;
    void* args[] = {&copy};
    const void* fn = nullptr;
    if (!pick_variant(block_size, context_sensitive(a), [&](auto mega, auto lvl) { (void)lvl; fn = (const void*)mega; }))
        return hipErrorInvalidValue;
    return hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(block_size), args, a.lds_bytes, s);
}

hipError_t pga_tape_occupancy(uint32_t lds_bytes, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, pga::k_mega<pga::kTapeBlock, true, false>,
                                                         (int)pga::kTapeBlock, lds_bytes);
}

hipError_t pga_launch_mega_tape(const pga::KArgs& a, uint32_t blocks, hipStream_t s) {
    pga::KArgs copy = a;
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel((const void*)pga::k_mega<pga::kTapeBlock, true, false>, dim3(blocks),
                                      dim3(pga::kTapeBlock), args, a.lds_bytes, s);
}
