// pga_spec.h - derivation kernels specialised to one program.
//
// The library writes, for the program loaded by pga_initialize, one type per
// dispatch entry whose fields are compile-time constants (operator kind,
// flags, inline parameters, successors), and a dispatcher from an entry
// index to those types; hipRTC compiles that source together with this
// header into a persistent kernel for the program. The operators are the
// same templates the interpreting kernels run (pga_ops.h); given a constant
// entry they fold to the operator at hand: no operator-kind dispatch, no
// table reads for constant parameters or successors, the constant parameter
// values as immediates.
//
// The state machine differs from the interpreter's in two ways:
//  * a lane runs its shape until it needs the wave - a terminal to emit or a
//    fan-out to start - or is done: a single-operator chain, a leaf fan-out's
//    child and that child's chain, or all descriptors of a started fan-out,
//    take one trip round the wave loop instead of one trip per operator;
//  * every step reaches its entry through `SpecVisit`: the lanes that share
//    an entry (a group of entries, for large programs) run together under a
//    wave-uniform switch (one pass per distinct entry among the active
//    lanes), so a level whose children share an entry runs that entry's code
//    once per wave with scalar branches.
// Output slots, LDS staging, the count/scan/write structure and the grid
// barrier are the ones of pga_level.h, so the instances are those of the
// interpreting kernels (the same records; the order within a claim may
// differ).
#pragma once

#include "pga_level.h"

namespace pga {

// Fields every constant entry type shares, from D's constants idx, w0, w1,
// pbase, sbase, pv0..pv2, sv0..sv2, aux0, aux1.
template <class D>
struct SpecEntry {
    __device__ __forceinline__ static constexpr uint32_t kind() { return D::w0 & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t shape() { return (D::w0 >> 8) & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t flags() { return (D::w0 >> 16) & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t nparams() { return D::w0 >> 24; }
    __device__ __forceinline__ static constexpr uint32_t nsucc() { return D::w1 & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t inl() { return (D::w1 >> 8) & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t tslot() { return (D::w1 >> 16) & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t ndirect() { return D::w1 >> 24; }
    __device__ __forceinline__ static float pvk(const Ctx&, uint32_t k) {
        return k == 0u ? D::pv0 : (k == 1u ? D::pv1 : D::pv2);
    }
    __device__ __forceinline__ static int32_t svk(const Ctx&, uint32_t k) {
        return k == 0u ? D::sv0 : (k == 1u ? D::sv1 : D::sv2);
    }
    __device__ __forceinline__ static uint32_t auxk(const Ctx&, uint32_t k) { return k == 0u ? D::aux0 : D::aux1; }
    __device__ __forceinline__ static DevParam param(const Ctx& c, uint32_t k) { return c.prog.params[D::pbase + k]; }
};

// The entry type for an index outside the program (a dispatcher's default):
// its kind matches no operator.
constexpr uint32_t K_NONE = 0xFFu;
struct SpecNoEntry : SpecEntry<SpecNoEntry> {
    static constexpr uint32_t idx = 0u, w0 = K_NONE, w1 = 0u, pbase = 0u, sbase = 0u, aux0 = 0u, aux1 = 0u;
    static constexpr float pv0 = 0.0f, pv1 = 0.0f, pv2 = 0.0f;
    static constexpr int32_t sv0 = -1, sv1 = -1, sv2 = -1;
};

// An entry of a group: the entries of a program that share both header words
// (operator kind, shape, flags, parameter and successor counts, inline mask,
// terminal slot, direct fan-out width). Those fields are compile-time
// constants; the index, parameter values, successors and aux words are this
// lane's, read from the LDS copy of the program. Large programs are written
// as groups so that the code to compile grows with the number of distinct
// operator forms rather than with the number of entries.
template <uint32_t W0, uint32_t W1>
struct SpecGroupEntry {
    static constexpr uint32_t w0 = W0, w1 = W1;
    uint32_t idx, pbase, sbase;
    float pv0, pv1, pv2;
    int32_t sv0;
    __device__ __forceinline__ static constexpr uint32_t kind() { return W0 & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t shape() { return (W0 >> 8) & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t flags() { return (W0 >> 16) & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t nparams() { return W0 >> 24; }
    __device__ __forceinline__ static constexpr uint32_t nsucc() { return W1 & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t inl() { return (W1 >> 8) & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t tslot() { return (W1 >> 16) & 0xFFu; }
    __device__ __forceinline__ static constexpr uint32_t ndirect() { return W1 >> 24; }
    __device__ __forceinline__ float pvk(const Ctx&, uint32_t k) const { return k == 0u ? pv0 : (k == 1u ? pv1 : pv2); }
    __device__ __forceinline__ int32_t svk(const Ctx& c, uint32_t k) const {
        return k == 0u ? sv0 : c.prog.entries[idx].sv[k];
    }
    __device__ __forceinline__ uint32_t auxk(const Ctx& c, uint32_t k) const { return c.prog.entries[idx].aux[k]; }
    __device__ __forceinline__ DevParam param(const Ctx& c, uint32_t k) const { return c.prog.params[pbase + k]; }
    __device__ __forceinline__ static SpecGroupEntry at(const Ctx& c, uint32_t i) {
        const uint4* q = reinterpret_cast<const uint4*>(c.prog.entries) + i * 3u;
        const uint4 a = q[0], b = q[1];
        SpecGroupEntry e;
        e.idx = i;
        e.pbase = a.z;
        e.sbase = a.w;
        e.pv0 = __uint_as_float(b.x);
        e.pv1 = __uint_as_float(b.y);
        e.pv2 = __uint_as_float(b.z);
        e.sv0 = (int32_t)b.w;
        return e;
    }
};

// The dispatch key of entry i: its index when every entry has a type of its
// own, its two header words when the program is written as groups (an index
// outside the program gets K_NONE, which no group has).
template <class S>
__device__ __forceinline__ uint64_t spec_key(const Ctx& c, uint32_t i) {
    if (!S::kGrouped) return i;
    if (i >= S::kEntries) return K_NONE;
    const uint2 w = *reinterpret_cast<const uint2*>(c.prog.entries + i);
    return ((uint64_t)w.y << 32) | w.x;
}

// Calls f(entry) for each lane with the entry of index `i` as its type: one
// pass per distinct key among the active lanes, each under a wave-uniform
// switch (S::dispatch).
template <class S>
struct SpecVisit {
    template <class F>
    __device__ __forceinline__ void operator()(const Ctx& c, uint32_t i, F&& f) const {
        const uint64_t key = spec_key<S>(c, i);
        bool pending = true;
#pragma unroll 1
        while (pending) {
            const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)key);
            const uint32_t hi = S::kGrouped ? __builtin_amdgcn_readfirstlane((uint32_t)(key >> 32)) : 0u;
            const uint64_t u = ((uint64_t)hi << 32) | lo;
            if (key == u) {
                S::dispatch(c, u, i, f);
                pending = false;
            }
        }
    }
};

// Where a pass reads its children and writes its fan-outs.
struct PassIO {
    uint32_t first;                 // level 0: the first axiom
    const Rec* par_in;              // parents of the children (levels > 0), `in_cap` of them
    const Desc* desc_in;
    uint32_t in_cap;
    Rec* par_out;                   // fan-outs started by the children
    Desc* desc_out;
    uint32_t par_cap, desc_cap;
};

// The arrays of a grid-wide level: the even / odd level buffers.
__device__ __forceinline__ PassIO grid_io(const KArgs& a, uint32_t level) {
    PassIO io;
    io.first = a.first_axiom;
    io.par_in = (level & 1u) ? a.par0 : a.par1;
    io.desc_in = (level & 1u) ? a.desc0 : a.desc1;
    io.in_cap = a.cap_par;
    io.par_out = (level & 1u) ? a.par1 : a.par0;
    io.desc_out = (level & 1u) ? a.desc1 : a.desc0;
    io.par_cap = a.cap_par;
    io.desc_cap = a.cap_desc;
    return io;
}

// One pass (WR: the write pass) over the n children of a level, 64 a wave,
// the waves numbered `wave` of `nwaves`.
template <class S, bool WR>
__device__ __forceinline__ void spec_pass(const KArgs& a, LevelLDS& L, float4* PW, uint32_t level, uint32_t n,
                                          uint32_t wave, uint32_t nwaves, const PassIO& io) {
    Ctx c = make_ctx<false>(a, L, level, WR);   // one-phase programs without Collider / IfCollides
    c.geom = WR ? 1u : S::kCountGeom;     // constants: the count pass drops the unread arithmetic
    c.seed = S::kUseSeed;
    c.poly = S::kPolygons;
    c.par_out = io.par_out;
    c.desc_out = io.desc_out;
    c.par_cap = io.par_cap;
    c.desc_cap = io.desc_cap;
    const SpecVisit<S> visit;
    f4* ST = reinterpret_cast<f4*>(PW + 5u * blockDim.x) + wave_in_block() * 256u;
#pragma unroll 1
    for (uint32_t base = wave * 64u; base < n; base += nwaves * 64u) {
        const uint32_t i = base + lane_id();
        Rec cur;
        const bool valid = i < n && load_child_at(a, c, io.first, io.par_in, io.in_cap, io.desc_in, WR, level, i,
                                                  cur, visit);
        uint32_t mode = valid ? M_SETTLE : M_DONE;
        bool leaf = false, self = false;
        uint32_t steps = 0, j = 0, dslot = 0, pidx = ~0u, fe = 0;
        uint32_t skipm = 0, nd = 0;
        Fan f;
        f.n = 0; f.axis = 0; f.total = 0.0f; f.r2a = 0.0f; f.off = 0.0f;
        f.e1 = f.e2 = f.e3 = 0.0f; f.prev_ns = 0.0f; f.prev_o = 0.0f; f.reps = 0;
#pragma unroll 1
        while (__ballot(mode != M_DONE) != 0ull) {
            bool emit = false, start = false, direct = false;
            uint32_t slot = 0;
            // this lane alone, until it needs the wave or is done
#pragma unroll 1
            while (mode != M_DONE && !emit && !start && !direct) {
                visit(c, mode == M_FAN ? fe : rec_entry(cur), [&](const auto& e) {
                    const uint32_t k = e.kind();
                    if (k == K_NONE) {          // entries are validated on load: cannot happen
                        raise(c, PGA_ERR_BAD_PROGRAM);
                        mode = M_DONE;
                        return;
                    }
                    if (mode == M_FAN) {
                        // child j of this lane's leaf fan-out (entry e, record in PW)
                        if (j >= f.n || !is_fanout(k)) { mode = M_DONE; return; }
                        float o = 0.0f, ns = 0.0f;
                        const Rec par = par_load(PW);
                        const bool keep = fan_geom(c, par, e, f, j, o, ns);
                        const bool skip = j < 32u ? ((skipm >> j) & 1u) != 0u : skipm == 0xFFFFFFFFu;
                        if (keep && !skip && build_child(c, par, e, j, f.axis, o, ns, cur)) {
                            mode = M_SETTLE;
                            steps = 0;
                        }
                        ++j;
                        return;
                    }
                    // SETTLE: the operator of the lane's current record
                    const uint32_t back = leaf ? (j < f.n ? M_FAN : M_DONE) : M_DONE;
                    if (k == K_GENERATE) {
                        slot = e.tslot();
                        if (slot == 0xFFu) {    // terminal index computed per shape
                            const uint32_t t = f2u_rd(P(c, e, 0, cur, cur.axiom, cur.path, rec_depth(c, cur)));
                            slot = t < PGA_MAX_TERMINALS ? c.prog.term_slot[t] : 0xFFu;
                        }
                        if (slot >= c.prog.nslots) { raise(c, PGA_ERR_BAD_PROGRAM); slot = 0; }
                        emit = true;
                        mode = back;
                    } else if (k == K_DISCARD) {
                        mode = back;
                    } else if (is_fanout(k)) {
                        if (leaf) {             // the load-time leaf analysis excludes fan-outs below a leaf
                            raise(c, PGA_ERR_BAD_PROGRAM);
                            mode = back;
                            return;
                        }
                        par_store(PW, cur);
                        fe = e.idx;
                        fan_init(c, cur, e, f);
                        j = 0;
                        leaf = (e.flags() & EF_LEAF) != 0;
                        start = !leaf;
                        const uint32_t sk = e.auxk(c, 0);
                        if (k == K_REPEAT) {
                            skipm = (sk & 1u) ? 0xFFFFFFFFu : 0u;
                            nd = skipm ? 0u : f.n;
                        } else {
                            skipm = sk;
                            nd = f.n - (uint32_t)__popc(skipm & (f.n < 32u ? (1u << f.n) - 1u : 0xFFFFFFFFu));
                        }
                        if (leaf && (e.flags() & EF_DIRECT)) {
                            if (WR) {                   // the children in lock step with the wave's other ones
                                direct = true;
                            } else {
                                count_direct(c, L, e, f);
                            }
                            mode = M_DONE;
                        } else {
                            // a leaf fan-out takes its children round the state
                            // machine; the write pass writes the descriptors of
                            // any other once the fan-out is started
                            mode = leaf ? M_FAN : M_DONE;
                        }
                    } else if (steps >= c.max_chain) {
                        if (leaf) {             // the leaf analysis bounds every chain below a leaf
                            raise(c, PGA_ERR_BAD_PROGRAM);
                            mode = back;
                            return;
                        }
                        // over-long chain: continues on the next level
                        par_store(PW, cur);
                        fe = e.idx;
                        self = true;
                        f.n = 1;
                        nd = 1;
                        skipm = 0;
                        j = 0;
                        start = true;
                        mode = M_DONE;
                    } else {
                        uint32_t ord;
                        const int32_t next = apply_single(c, e, cur, &ord);
                        if (!advance(c, cur, next, ord)) mode = back;
                        ++steps;
                    }
                });
            }
            if (__ballot(emit) != 0ull) {
                unsigned long long idx;
                emit_terminal(c, L, ST, WR, emit, slot, cur, idx);
            }
            if (__ballot(start) != 0ull) {
                start_fanout(c, L, WR, start, par_load(PW), nd, pidx, dslot, false);
                if (WR && start) {
                    // the fan-out's descriptors in one loop over its children
                    // (the running offsets of fan_geom in child order), child
                    // jj of every lane of the wave side by side: each store
                    // writes consecutive descriptors
                    visit(c, fe, [&](const auto& e) {
                        const Rec par = par_load(PW);
                        const bool geom = !self && is_fanout(e.kind());
                        const uint32_t kD = c.prog.nslots + 1u;
#pragma unroll 1
                        for (uint32_t jj = 0; __ballot(jj < f.n) != 0ull; ++jj) {
                            const bool act = jj < f.n;
                            float o = 0.0f, ns = 0.0f;
                            bool keep = true;
                            if (act && geom) keep = fan_geom(c, par, e, f, jj, o, ns);
                            const bool skip = jj < 32u ? ((skipm >> jj) & 1u) != 0u : skipm == 0xFFFFFFFFu;
                            const bool has = act && !skip;
                            const unsigned long long m = __ballot(has);
                            if (m == 0ull) continue;
                            const uint32_t base = claim(L, true, kD, m, __ffsll((long long)m) - 1);
                            if (has) {
                                const uint32_t ds = base + mbcnt(m);
                                const uint32_t jx = pidx == ~0u ? kDescNone
                                                  : self        ? kDescSelf
                                                  : keep        ? (jj | (f.axis << 30))
                                                                : kDescNone;
                                if (ds < c.desc_cap)
                                    *reinterpret_cast<uint4*>(c.desc_out + ds) =
                                        make_uint4(pidx, jx, __float_as_uint(o), __float_as_uint(ns));
                                else
                                    raise(c, PGA_ERR_OVERFLOW_QUEUE);
                            }
                        }
                    });
                }
            }
            if (__ballot(direct) != 0ull) {
                // a direct lane is done with its own record (the parent is in
                // PW): its children are built into `cur`
#pragma unroll 1
                for (uint32_t jj = 0; __ballot(direct && jj < f.n) != 0ull; ++jj) {
                    uint32_t sl = 0;
                    bool has = false;
                    if (direct && jj < f.n)
                        visit(c, fe, [&](const auto& e) {
                            if (is_fanout(e.kind())) {
                                const Rec par = par_load(PW);
                                has = direct_child(c, par, e, f, skipm, jj, cur, sl);
                            }
                        });
                    unsigned long long idx;
                    emit_terminal(c, L, ST, true, has, sl, cur, idx);
                }
            }
        }
    }
}

// The block scan of a workgroup-local level (see spec_mega): wave offsets
// and totals; the level's instances of each slot claimed after those already
// claimed (one atomic per slot); the fan-outs at 0 of the block's next
// arrays or, when they do not fit (more than kp parents or kd descriptors),
// spilled to the grid's level-1 input (one atomic per array). Returns the
// spill.
__device__ __forceinline__ bool local_scan(const KArgs& a, LevelLDS& L, uint32_t level, uint32_t kp, uint32_t kd) {
    const uint32_t nslots = a.prog.nslots, ncnt = nslots + 2u;
    const uint32_t nw = blockDim.x >> 6;
    for (uint32_t k = threadIdx.x; k < ncnt; k += blockDim.x) {
        uint32_t sum = 0;
        for (uint32_t ww = 0; ww < nw; ++ww) {
            L.cur[ww][k] = sum;
            sum += L.cnt[ww][k];
        }
        L.level_total[k] = sum;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < ncnt; k += blockDim.x) {
        const uint32_t tot = L.level_total[k];
        if (k < nslots) {
            if (tot != 0u)
                L.term_base[k] = __hip_atomic_fetch_add(&a.term_counts[pga_smem[a.off_slots + PGA_MAX_TERMINALS + k]],
                                                        (unsigned long long)tot, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
        } else if (k == nslots) {
            const uint32_t td = L.level_total[nslots + 1u];
            const bool spill = tot > kp || td > kd;
            L.spill = spill ? 1u : 0u;
            L.before[nslots] = spill && tot != 0u ? __hip_atomic_fetch_add(&a.bar[kBarSpillPar], tot,
                                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                                  : 0u;
            L.before[nslots + 1u] = spill && td != 0u ? __hip_atomic_fetch_add(&a.bar[kBarSpillDesc], td,
                                                                               __ATOMIC_RELAXED,
                                                                               __HIP_MEMORY_SCOPE_AGENT)
                                                      : 0u;
            if (tot != 0u) atomicAdd(&a.par_counts[a.max_levels], tot);
            atomicMax(&a.par_counts[a.max_levels + 1u], level + 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = nslots + threadIdx.x; k < ncnt; k += blockDim.x)
        for (uint32_t ww = 0; ww < nw; ++ww) L.cur[ww][k] += L.before[k];
    const bool spill = L.spill != 0u;
    __syncthreads();
    return spill;
}

// The persistent derivation of every level (the MEGAKERNEL technique).
//
// With local arrays (a.loc_par), the levels run first inside each
// workgroup: a block takes loc_grab axioms at a time (an atomic cursor),
// whenever its level is narrower than the block, and derives what it holds
// level by level on its own - the same count and write passes over the
// block's waves, a block scan (local_scan) for the grid scan, the fan-outs
// into the block's double-buffered arrays and each level's terminals claimed
// with one atomic per slot (the instances are the same records; their order
// across blocks is unspecified). New axioms enter a level as kDescAxiom
// descriptors beside the children of the ones taken earlier, so derivations
// of different depths share passes. A level whose fan-outs do not fit the
// block's arrays writes them to the grid's level-1 input instead; when a
// block holds nothing and no axiom is left, the grid-wide levels (count
// pass, grid scan, write pass, grid barrier) take that spill up. A batch of
// small or uneven derivations thus pays no grid barrier per level. Without
// local arrays the grid-wide levels start from the axioms.
template <class S, int BS>
__device__ __forceinline__ void spec_mega(const KArgs& a) {
    __shared__ LevelLDS L;
    __shared__ float4 PW[9 * BS];   // parent records (5 words per lane), then instance staging (4)
    stage_program(a);
    const uint32_t nslots = a.prog.nslots, ncnt = nslots + 2u;
    const uint32_t w = wave_in_block();
    const uint32_t kp = a.loc_kp, kd = a.loc_kd;
    Rec* const lp = a.loc_par + (size_t)blockIdx.x * 2u * kp;
    Desc* const ld = a.loc_desc + (size_t)blockIdx.x * 2u * kd;
    bool local = a.loc_par != nullptr;
    bool more = true;               // local: axioms may be left to hand out
    uint32_t level = 0, buf = 0, since = 0;
    uint32_t n = local ? 0u : a.n_axioms;
    uint32_t tick = 0;              // timestamp row: local levels, then grid levels
    if (!local) init_bases(a, L);
#pragma unroll 1
    for (;;) {
        if (local) {
            // a narrow level takes loc_grab more axioms (as kDescAxiom
            // descriptors after the level's own), so the block's passes stay
            // full while the derivations it holds go deeper
            if (more && n < blockDim.x) {
                if (threadIdx.x == 0)
                    L.grab = __hip_atomic_fetch_add(&a.bar[kBarGrab], a.loc_grab, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
                __syncthreads();
                const uint32_t g = L.grab;
                if (g < a.n_axioms) {
                    const uint32_t cnt = min(a.loc_grab, a.n_axioms - g);
                    Desc* const din = ld + buf * kd;      // n + cnt < blockDim + 64 <= kd
                    for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x)
                        *reinterpret_cast<uint4*>(din + n + i) = make_uint4(a.first_axiom + g + i, kDescAxiom, 0u, 0u);
                    n += cnt;
                    since = 0;
                } else {
                    more = false;
                }
                __syncthreads();
            }
            if (n == 0u) {          // nothing held and nothing left: the grid-wide levels
                grid_barrier(a.bar, gridDim.x);     // every spill and instance claim is done
                const uint32_t spilled =
                    __hip_atomic_load(&a.bar[kBarSpillDesc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (blockIdx.x == 0 && threadIdx.x == 0) a.desc_counts[0] = spilled;
                local = false;
                level = 1;
                n = min(spilled, a.cap_desc);
                init_bases(a, L);
                continue;
            }
            if (since >= a.max_levels) {    // as the grid-wide levels: the frontier is dropped
                if (threadIdx.x == 0) atomicOr(a.err, 1u << PGA_ERR_MAX_DEPTH);
                n = 0;
                continue;
            }
        } else if (level >= a.max_levels || n == 0u) {
            break;
        }
        PassIO io;
        uint32_t wave, nwaves;
        if (local) {
            io.first = 0;
            io.par_in = lp + buf * kp;
            io.desc_in = ld + buf * kd;
            io.in_cap = kp;
            io.par_out = lp + (buf ^ 1u) * kp;
            io.desc_out = ld + (buf ^ 1u) * kd;
            io.par_cap = kp;
            io.desc_cap = kd;
            wave = w;
            nwaves = blockDim.x >> 6;
        } else {
            io = grid_io(a, level);
            wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
            nwaves = (gridDim.x * blockDim.x) >> 6;
        }
        stamp(a, tick, 0);
        const uint32_t plevel = local ? 1u : level;     // local levels read descriptors only
        for (uint32_t k = lane_id(); k < ncnt; k += 64u) L.cnt[w][k] = 0u;
        spec_pass<S, false>(a, L, PW, plevel, n, wave, nwaves, io);
        __syncthreads();
        bool spill = false;
        if (local) {
            stamp(a, tick, 1);
            spill = local_scan(a, L, level, a.loc_cap_p, a.loc_cap_d);
            stamp(a, tick, 2);
            if (spill) {
                io.par_out = a.par0;
                io.desc_out = a.desc0;
                io.par_cap = a.cap_par;
                io.desc_cap = a.cap_desc;
            }
        } else {
            scan_counts(a, L, level, tick);
        }
        spec_pass<S, true>(a, L, PW, plevel, n, wave, nwaves, io);
        __syncthreads();
        stamp(a, tick, 3);
        ++tick;
        if (local) {
            n = spill ? 0u : L.level_total[nslots + 1u];
            buf ^= 1u;
            ++since;
        } else {
            if (threadIdx.x == 0)
                for (uint32_t k = 0; k < nslots; ++k) L.term_base[k] += L.level_total[k];
            __syncthreads();
            n = min(L.next_n, a.cap_desc);
            grid_barrier(a.bar, gridDim.x);       // next level reads this level's records
        }
        ++level;
    }
    store_bases(a, L);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (n != 0u) atomicOr(a.err, 1u << PGA_ERR_MAX_DEPTH);
        a.desc_counts[a.max_levels + 1u] = level;
    }
}

} // namespace pga
