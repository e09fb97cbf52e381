// pga_ops.h - operator execution on one shape record per lane.
//
// Single-child operators (IfSizeLess, If, RandomRule, Translate, Rotate,
// Scale, SwapSize, Extrude, Collider, IfCollides) transform the record in
// registers and continue with the successor entry. Fan-out operators (Subdivide, Repeat,
// ComponentSplit, Replicate) are split into `fan_init` (per-parent prelude:
// axis, relative-to-absolute scale, repetition counts), `fan_geom` (the
// geometry of child j, called for j = 0, 1, ... in order, carrying the
// running offset exactly as the reference's sequential loops do) and
// `build_child` (child j's record from the parent and that geometry).
//
// Entries are read as three 16-byte words and unpacked into scalars, so no
// entry is ever an addressable local. `Ctx::geom` = 0 skips the model-matrix
// arithmetic (the counting pass of a program whose decisions never read the
// matrix).
#pragma once

#include "pga_device.h"

namespace pga {

// The fixed fields of a dispatch entry, unpacked into registers. Fields
// chosen by a run-time index (successor k, inline parameter k, the aux words)
// are read from the LDS entry by index: a select between register copies
// would be turned into an indexed load from a stack copy of the entry.
//
// The operators below take the entry as a template parameter E: EntryV here
// (the interpreting kernels, one copy of every operator), or a type whose
// fields are compile-time constants (kernels specialised to one program,
// where the same operator code folds to the entry at hand). E provides the
// fields idx, w0, w1, pbase, sbase, pv0..pv2, sv0, the accessors below, and
// pvk / svk / auxk / param for fields chosen by index.
struct EntryV {
    uint32_t idx;     // entry index
    uint32_t w0;      // kind | shape << 8 | flags << 16 | nparams << 24
    uint32_t w1;      // nsucc | inl << 8 | tslot << 16 | ndirect << 24
    uint32_t pbase, sbase;
    float pv0, pv1, pv2;
    int32_t sv0;
    __device__ __forceinline__ uint32_t kind() const { return w0 & 0xFFu; }
    __device__ __forceinline__ uint32_t shape() const { return (w0 >> 8) & 0xFFu; }
    __device__ __forceinline__ uint32_t flags() const { return (w0 >> 16) & 0xFFu; }
    __device__ __forceinline__ uint32_t nparams() const { return w0 >> 24; }
    __device__ __forceinline__ uint32_t nsucc() const { return w1 & 0xFFu; }
    __device__ __forceinline__ uint32_t inl() const { return (w1 >> 8) & 0xFFu; }
    __device__ __forceinline__ uint32_t tslot() const { return (w1 >> 16) & 0xFFu; }
    __device__ __forceinline__ uint32_t ndirect() const { return w1 >> 24; }
    // inline constant parameter k < 3, successor k < 3, aux word k
    __device__ __forceinline__ float pvk(const Ctx& c, uint32_t k) const { return c.prog.entries[idx].pv[k]; }
    __device__ __forceinline__ int32_t svk(const Ctx& c, uint32_t k) const { return c.prog.entries[idx].sv[k]; }
    __device__ __forceinline__ uint32_t auxk(const Ctx& c, uint32_t k) const { return c.prog.entries[idx].aux[k]; }
    // parameter k (any k < nparams) from the program tables
    __device__ __forceinline__ DevParam param(const Ctx& c, uint32_t k) const { return c.prog.params[pbase + k]; }
};
static_assert(sizeof(DevEntry) == 48, "DevEntry must be three 16-byte words");

// Entry i of this lane, from the LDS copy of the program.
__device__ __forceinline__ EntryV load_entry(const Ctx& c, uint32_t i) {
    const uint4* q = reinterpret_cast<const uint4*>(c.prog.entries) + i * 3u;
    const uint4 a = q[0], b = q[1];
    EntryV e;
    e.idx = i;
    e.w0 = a.x; e.w1 = a.y; e.pbase = a.z; e.sbase = a.w;
    e.pv0 = __uint_as_float(b.x); e.pv1 = __uint_as_float(b.y); e.pv2 = __uint_as_float(b.z);
    e.sv0 = (int32_t)b.w;
    return e;
}
template <class E>
__device__ __forceinline__ uint32_t entry_aux(const Ctx& c, const E& e, uint32_t k) {
    return e.auxk(c, k);
}

// Shape type of entry i.
__device__ __forceinline__ uint32_t entry_shape(const Ctx& c, uint32_t i) { return c.prog.entries[i].shape; }

__device__ __forceinline__ bool advance(const Ctx& c, Rec& r, int32_t next, uint32_t ordinal) {
    if (next < 0) return false;   // unused dynamic edge: nothing dispatched
    // successors of a multi-phase program carry their phase in bits 16-19
    // (DispatchTableEntry.h:8-36, the successor's phaseIndex), bit 20 for
    // the parent's own; a child bound for an earlier phase than its parent's
    // is never derived (that phase has run), one bound for a later phase than
    // this launch's waits for its own launch (Call.cuh:69-70 dispatches into
    // that phase's queues)
    uint32_t cph = (uint32_t)next >> 16;
    if (multi_phase(c)) {
        const uint32_t pph = r.ed >> 28;
        if (cph & 16u) cph = pph;
        if (cph < pph || cph > c.phase) return false;
        // a shape of an earlier phase whose entry reaches no phase of this
        // launch (the entry's aux[1]) adds nothing here
        if (cph < c.phase && c.prog.entries[(uint32_t)next & 0x7FFFu].aux[1] < c.phase) return false;
    }
    uint32_t depth = rec_depth(c, r) + 1u;
    if (depth > c.dmask) { raise(c, PGA_ERR_MAX_DEPTH); return false; }
    r.ed = ((uint32_t)next & 0x7FFFu) | (c.poly ? (r.ed & kInvert) : 0u) | (depth << 16) | (cph << 28);
    r.path = path_mix(r.path, ordinal);
    return true;
}

__device__ __forceinline__ uint32_t axis_of(const Ctx& c, float v) {
    uint32_t a = f2u_rd(v);
    if (a > 2u) { raise(c, PGA_ERR_BAD_PROGRAM); a = 0u; }
    return a;
}

__device__ __forceinline__ float sel3(const float s[3], uint32_t a) {
    return a == 0 ? s[0] : (a == 1 ? s[1] : s[2]);
}

__device__ __forceinline__ void set_model(Rec& r, const float m[12]) {
#pragma unroll
    for (int i = 0; i < 12; ++i) r.m[i] = m[i];
}

// Parameter k of entry e for the shape r.
template <class E>
__device__ __forceinline__ float P(const Ctx& c, const E& e, uint32_t k, const Rec& r, uint32_t axiom,
                                   uint32_t path, uint32_t depth) {
    if (k < 3u && ((e.inl() >> k) & 1u)) return e.pvk(c, k);
    return eval_param(c, e.param(c, k), r, axiom, path, depth);
}

// Successor k of an entry (the first three are held in the entry).
template <class E>
__device__ __forceinline__ int32_t succ_of(const Ctx& c, const E& e, uint32_t k) {
    if (k < 3u) return e.svk(c, k);
    return c.prog.succ[e.sbase + k];
}

// The first min(nparams, 3) parameters; constants come from the entry, the
// others through one evaluation site.
template <class E>
__device__ __forceinline__ void eval3(const Ctx& c, const E& e, const Rec& r, float& v0, float& v1,
                                      float& v2) {
    const uint32_t np = e.nparams() < 3u ? e.nparams() : 3u;
    v0 = e.pv0; v1 = e.pv1; v2 = e.pv2;
    const uint32_t todo = ((1u << np) - 1u) & ~e.inl();
    if (todo == 0u) return;
    const uint32_t ax = r.axiom, pa = r.path, de = rec_depth(c, r);
#pragma unroll 1
    for (uint32_t k = 0; k < np; ++k) {
        if (!((todo >> k) & 1u)) continue;
        float x = eval_param(c, e.param(c, k), r, ax, pa, de);
        v0 = k == 0 ? x : v0;
        v1 = k == 1 ? x : v1;
        v2 = k == 2 ? x : v2;
    }
}

// RandomRule<DynParams> (RandomRule.cuh:121-147) and the static form
// (RandomRule.cuh:20-111: right-fold sum, zero-probability branches skipped,
// last branch unconditional). Returns the chosen branch or -1.
template <class E>
__device__ __forceinline__ int32_t random_choice(const Ctx& c, const E& e, float chance, const Rec& rr) {
    const uint32_t n = e.nparams();
    const bool is_static = (e.flags() & EF_STATIC) != 0;
    const uint32_t axiom = rr.axiom, path = rr.path, depth = rec_depth(c, rr);
    float sum = 0.0f;
    if (!is_static) {
        for (uint32_t i = 1; i < n; ++i) sum += P(c, e, i, rr, axiom, path, depth);
    } else {
        for (uint32_t i = n - 1; i >= 1; --i) {
            float p = P(c, e, i, rr, axiom, path, depth);
            sum = (i == n - 1) ? p : p + sum;
        }
    }
    float acc = 0.0f;
    for (uint32_t i = 1; i < n; ++i) {
        if (is_static && i == n - 1) return (int32_t)i - 1;
        float np = P(c, e, i, rr, axiom, path, depth) / sum;
        acc += np;
        if ((!is_static || np != 0.0f) && chance <= acc) return (int32_t)i - 1;
    }
    return -1;
}

// ---- context sensitivity (Collider.cuh, IfCollides.cuh, Collision.cuh) ----

// Separating-axis test of two oriented boxes A (a stored collider) and B (the
// shape asking) (Gottschalk 1996; Ericson, Real-Time Collision Detection
// 4.4.1), with the axes in the order of Collision.cuh:35-262: A's three face
// axes, B's three, the nine edge cross products. Models are row-major 3x4,
// the box axes are their columns. af / bf: half extents on the shape's own
// face axes; ap / bp: in projections onto the other shape's axes, where a
// Quad's z half extent drops out (Box x Quad :263-490, Quad x Quad :506-733).
__device__ __forceinline__ bool sat_overlap(const float A[12], const float af[3], const float ap[3],
                                            const float B[12], const float bf[3], const float bp[3]) {
    const float v0 = B[3] - A[3], v1 = B[7] - A[7], v2 = B[11] - A[11];
    float T[3], R[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T[i] = v0 * A[i] + v1 * A[4 + i] + v2 * A[8 + i];
#pragma unroll
        for (int k = 0; k < 3; ++k) R[i][k] = A[i] * B[k] + A[4 + i] * B[4 + k] + A[8 + i] * B[8 + k];
    }
    bool sep = false;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        sep |= fabsf(T[i]) > af[i] + (bp[0] * fabsf(R[i][0]) + bp[1] * fabsf(R[i][1]) + bp[2] * fabsf(R[i][2]));
#pragma unroll
    for (int k = 0; k < 3; ++k)
        sep |= fabsf(T[0] * R[0][k] + T[1] * R[1][k] + T[2] * R[2][k]) >
               (ap[0] * fabsf(R[0][k]) + ap[1] * fabsf(R[1][k]) + ap[2] * fabsf(R[2][k])) + bf[k];
    // L = A_i x B_k: the two A axes other than i and the two B axes other than k
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int i1 = i == 0 ? 1 : 0, i2 = i == 2 ? 1 : 2;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int k1 = k == 0 ? 1 : 0, k2 = k == 2 ? 1 : 2;
            const float ra = ap[i1] * fabsf(R[i2][k]) + ap[i2] * fabsf(R[i1][k]);
            const float rb = bp[k1] * fabsf(R[i][k2]) + bp[k2] * fabsf(R[i][k1]);
            sep |= fabsf(T[i2] * R[i1][k] - T[i1] * R[i2][k]) > ra + rb;
        }
    }
    return !sep;
}

// Collision::check(stored, asking) of two Box / Quad shapes, half extents
// size * 0.5f (Shape.cuh:70-73); Quad x Box is Box x Quad with the shapes
// swapped (Collision.cuh:734-737).
__device__ __forceinline__ bool shapes_overlap(const float am[12], const float as[3], uint32_t at,
                                               const float bm[12], const float bs[3], uint32_t bt) {
    const float ah[3] = {as[0] * 0.5f, as[1] * 0.5f, as[2] * 0.5f};
    const float bh[3] = {bs[0] * 0.5f, bs[1] * 0.5f, bs[2] * 0.5f};
    const float aq[3] = {ah[0], ah[1], 0.0f}, bq[3] = {bh[0], bh[1], 0.0f};
    const bool aquad = at == PGA_SHAPE_QUAD, bquad = bt == PGA_SHAPE_QUAD;
    if (aquad && !bquad) return sat_overlap(bm, bh, bh, am, ah, aq);
    return sat_overlap(am, ah, aquad ? aq : ah, bm, bh, bquad ? bq : bh);
}

// Collider i of the store against the asking shape (bm, bs, shape): the tag
// matches, an earlier phase than the shape's stored it, and the shapes
// collide (IntermediateSymbolsBufferAdapter.cuh:89-136).
__device__ __forceinline__ bool collider_test(const float4* __restrict__ coll, uint32_t i, uint32_t tag,
                                              uint32_t phase, uint32_t shape, const float bm[12], const float bs[3]) {
    const float4* k = coll + 5u * i;
    const float4 w3 = k[3], w4 = k[4];
    if (__float_as_uint(w3.w) != tag || __float_as_uint(w4.y) >= phase) return false;
    const float4 w0 = k[0], w1 = k[1], w2 = k[2];
    const float am[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
    const float as[3] = {w3.x, w3.y, w3.z};
    return shapes_overlap(am, as, __float_as_uint(w4.x), bm, bs, shape);
}

// IfCollides' query: does a collider that an earlier phase than the shape's
// stored with this tag overlap the shape? The store holds `n` such
// colliders. With a BVH (built over them between phases, as
// MultiPhaseEvaluator::execute does, GPUEvaluator.cuh:427-430) the walk of
// BVHTraversal.cuh:21-148 runs without a stack over the nodes in
// depth-first order: a node whose box meets the shape's box is entered (the
// next node), any other skipped (its skip index); at a leaf the collider is
// tested in full. The boxes are conservative, so the answer is the
// existence over all n colliders, which a scan gives for few colliders.
__device__ __forceinline__ bool collider_hit(const Ctx& c, uint32_t n, uint32_t tag, uint32_t phase,
                                             uint32_t shape, float m0, float m1, float m2, float m3, float m4,
                                             float m5, float m6, float m7, float m8, float m9, float m10, float m11,
                                             float s0, float s1, float s2) {
    const float bm[12] = {m0, m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11};
    const float bs[3] = {s0, s1, s2};
    if (c.bvh_flat == nullptr || n <= 32u) {
#pragma unroll 1
        for (uint32_t i = 0; i < n; ++i)
            if (collider_test(c.coll, i, tag, phase, shape, bm, bs)) return true;
        return false;
    }
    float lo[3], hi[3];
    shape_box(bm, bs, lo, hi);
    const uint32_t nodes = 2u * n - 1u;
    uint32_t i = 0;
#pragma unroll 1
    while (i < nodes) {
        const float4 w0 = c.bvh_flat[2u * i], w1 = c.bvh_flat[2u * i + 1u];
        if (w0.x > hi[0] || w0.y > hi[1] || w0.z > hi[2] || w0.w < lo[0] || w1.x < lo[1] || w1.y < lo[2]) {
            i = __float_as_uint(w1.z);          // skip the subtree
            continue;
        }
        const uint32_t item = __float_as_uint(w1.w);
        if (item != 0xFFFFFFFFu && collider_test(c.coll, item, tag, phase, shape, bm, bs)) return true;
        ++i;
    }
    return false;
}

// Collider (Collider.cuh:26-42): a positive tag stores the shape with the
// phase it is derived in; once per shape (the write pass of its own phase).
__device__ __forceinline__ void store_collider(const Ctx& c, const Rec& r, uint32_t shape, uint32_t tag) {
    const uint32_t i = atomicAdd(c.coll_count, 1u);
    if (i >= c.coll_cap) { raise(c, PGA_ERR_OVERFLOW_QUEUE); return; }
    float4* k = c.coll + 5u * i;
    k[0] = make_float4(r.m[0], r.m[1], r.m[2], r.m[3]);
    k[1] = make_float4(r.m[4], r.m[5], r.m[6], r.m[7]);
    k[2] = make_float4(r.m[8], r.m[9], r.m[10], r.m[11]);
    k[3] = make_float4(r.s[0], r.s[1], r.s[2], __uint_as_float(tag));
    k[4] = make_float4(__uint_as_float(shape), __uint_as_float(c.phase), 0.0f, 0.0f);
}

// Applies one single-child operator in place. Returns the successor entry
// (or -1 when nothing is dispatched) and the child ordinal.
template <class E>
__device__ __forceinline__ int32_t apply_single(const Ctx& c, const E& e, Rec& r, uint32_t* ordinal,
                                                CollMemo* memo = nullptr) {
    const uint32_t k = e.kind();
    *ordinal = 0;
    // SetAs (SetAs.cuh:112-150): model, size, seed and custom attribute are
    // kept; the vertex list is the entry's (a table fact, see build_program)
    if (k == K_SETAS) { r.ed &= ~kInvert; return e.sv0; }
    if (!c.geom && (k == K_TRANSLATE || k == K_ROTATE)) return e.sv0;   // model only
    float v0, v1, v2;
    eval3(c, e, r, v0, v1, v2);
    if (k == K_IFSIZELESS) {                   // If.cuh:61-98
        uint32_t axis = axis_of(c, v0);
        uint32_t t = (sel3(r.s, axis) < v1) ? 0u : 1u;
        *ordinal = t;
        return succ_of(c, e, t);
    }
    if (k == K_IF) {                           // If.cuh:23-58 (NaN counts as true)
        uint32_t t = (v0 != 0.0f) ? 0u : 1u;
        *ordinal = t;
        return succ_of(c, e, t);
    }
    if (k == K_RANDOM) {
        float chance = v0 - floorf(v0);        // math::frac (math.h:243-246)
        int32_t ch = random_choice(c, e, chance, r);
        if (ch < 0) return -1;                 // dyn form: round-off can choose nothing
        *ordinal = (uint32_t)ch;
        return succ_of(c, e, (uint32_t)ch);
    }
    if (k == K_TRANSLATE) {                    // Translate.cuh:24-50
        float m[12];
        mul_translate(r.m, v0, v1, v2, m);
        set_model(r, m);
        return e.sv0;
    }
    if (k == K_ROTATE) {                       // Rotate.cuh:24-51, matrix.h:876-888
        const float kd = 3.1415926535897932384626434f / 180.0f;
        float sa, sb, sg, ca, cb, cg;
        sincos_f32(v0 * kd, &sa, &ca);
        sincos_f32(v1 * kd, &sb, &cb);
        sincos_f32(v2 * kd, &sg, &cg);
        const float b[16] = {cb * cg, cg * sa * sb - ca * sg, ca * cg * sb + sa * sg, 0.0f,
                             cb * sg, ca * cg + sa * sb * sg, -cg * sa + ca * sb * sg, 0.0f,
                             -sb, cb * sa, ca * cb, 0.0f,
                             0.0f, 0.0f, 0.0f, 1.0f};
        float m[12];
        mul34x44(r.m, b, m);
        set_model(r, m);
        return e.sv0;
    }
    if (k == K_SCALE) {                        // Scale.cuh:23-50
        const float v[3] = {v0, v1, v2};
#pragma unroll
        for (int i = 0; i < 3; ++i)
            r.s[i] = (((float)(v[i] < 0.0f) * r.s[i]) * v[i]) * -1.0f + ((float)(v[i] > 0.0f) * v[i]);
        return e.sv0;
    }
    if (k == K_SWAPSIZE) {                     // SwapSize.cuh:25-58
        uint32_t a0 = axis_of(c, v0), a1 = axis_of(c, v1), a2 = axis_of(c, v2);
        float s0 = sel3(r.s, a0), s1 = sel3(r.s, a1), s2 = sel3(r.s, a2);
        r.s[0] = s0; r.s[1] = s1; r.s[2] = s2;
        return e.sv0;
    }
    if (c.ctxs && k == K_COLLIDER) {           // Collider.cuh:26-42
        const uint32_t tag = f2u_rd(v0);
        if (c.wr && tag > 0u && rec_phase(c, r) == c.phase) store_collider(c, r, e.shape(), tag);
        return e.sv0;
    }
    if (c.ctxs && k == K_IFCOLLIDES) {         // IfCollides.cuh:26-58
        const uint32_t q = memo ? memo->n : 32u;
        uint32_t t;
        if (memo && memo->mode == 2u && q < 32u) {
            t = (memo->bits >> q) & 1u;         // the count pass's outcome
        } else {
            const uint32_t n = min(*c.coll_visible, c.coll_cap);
            t = (n != 0u && collider_hit(c, n, f2u_rd(v0), rec_phase(c, r), e.shape(), r.m[0], r.m[1], r.m[2],
                                         r.m[3], r.m[4], r.m[5], r.m[6], r.m[7], r.m[8], r.m[9], r.m[10], r.m[11],
                                         r.s[0], r.s[1], r.s[2]))
                    ? 0u : 1u;
            if (memo && memo->mode == 1u && q < 32u) memo->bits |= t << q;
        }
        if (memo) memo->n = q + 1u;
        *ordinal = t;
        return succ_of(c, e, t);
    }
    // K_EXTRUDE, Extrude.cuh:25-71 (Quad -> Box); parameter 0 (axis) unused
    if (c.geom) {
        const float b[16] = {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, -1.0f, 0.0f,
                             0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
        float m[12];
        mul34x44(r.m, b, m);
        set_model(r, m);
    }
    float sx = r.s[0], sy = r.s[1];
    r.s[0] = sx; r.s[1] = v1; r.s[2] = sy;
    if (c.seed) r.seed = next_seed(r.seed, 0u);
    return e.sv0;
}

__device__ __forceinline__ bool is_fanout(uint32_t k) {
    return k == K_SUBDIV || k == K_REPEAT || k == K_COMPSPLIT || k == K_REPLICATE || k == K_PRISMSPLIT;
}

// ---- fan-out operators ----
struct Fan {
    uint32_t n;          // children to iterate (j = 0 .. n-1)
    uint32_t axis;
    float total;
    float r2a;           // Subdivide: relative-to-absolute scale
    float off;           // running offset on the axis
    float e1, e2, e3;    // Repeat extents (start/mid/end)
    float prev_ns;       // Subdivide: size and offset of the previous child
    float prev_o;
    uint32_t reps;       // Repeat repetitions
};

template <class E>
__device__ __forceinline__ void fan_init(const Ctx& c, const Rec& r, const E& e, Fan& f) {
    const uint32_t kind = e.kind();
    f.n = 0;
    f.reps = 0;
    f.axis = 0;
    f.total = 0.0f;
    f.r2a = 0.0f;
    f.off = 0.0f;
    f.e1 = f.e2 = f.e3 = 0.0f;
    f.prev_ns = 0.0f; f.prev_o = 0.0f;
    if (kind == K_SUBDIV) {                    // Subdivide.cuh:142-153 / 181-216
        const uint32_t ax = r.axiom, pa = r.path, de = rec_depth(c, r);
        float abs_sum = 0.0f, rel_sum = 0.0f;
        uint32_t axis = 0;
#pragma unroll 1
        for (uint32_t i = 0; i < e.nparams(); ++i) {
            float fac = P(c, e, i, r, ax, pa, de);
            if (i == 0) {
                axis = f2u_rd(fac);
            } else {
                abs_sum += (float)(fac > 0.0f) * fac;
                rel_sum += (float)(fac < 0.0f) * (fac * -1.0f);
            }
        }
        if (axis > 2u) { raise(c, PGA_ERR_BAD_PROGRAM); return; }
        f.axis = axis;
        f.total = sel3(r.s, axis);
        f.off = f.total * -0.5f;
        f.r2a = (rel_sum == 0.0f) ? 0.0f : (f.total - abs_sum) / rel_sum;
        if (!(e.flags() & EF_STATIC) && f.r2a < 0.0f) {
            raise(c, PGA_ERR_SUBDIV_OVERCONSTRAINED);
            return;
        }
        f.n = e.nparams() - 1;
    } else if (kind == K_REPEAT) {             // Repeat.cuh:70-149
        float v0, ex, v2;
        eval3(c, e, r, v0, ex, v2);
        uint32_t axis = axis_of(c, v0);
        uint32_t mode = f2u_rd(v2);
        float sz = sel3(r.s, axis);
        float extent = (((float)(ex < 0.0f) * sz) * fabsf(ex)) + ((float)(ex > 0.0f) * ex);
        if (extent <= 0.0f || sz < 0.0f) { raise(c, PGA_ERR_REPEAT_BAD_EXTENT); return; }
        f.axis = axis;
        f.total = sz;
        float division = sz / extent;
        float e1, e2, e3;
        uint32_t reps;
        if (mode == PGA_ANCHOR_TO_START || mode == PGA_ANCHOR_TO_END) {
            float rest = fmodf(sz, extent);
            float exact = (float)(rest == 0.0f);
            float cd = ceilf(division);
            reps = cd >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)cd;
            e1 = exact * extent + (1.0f - exact) * rest;
            e2 = e3 = extent;
        } else if (mode == PGA_ADJUST_TO_FILL) {
            float fd = floorf(division);
            reps = fd >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)fd;
            if (reps > 0) e1 = e2 = e3 = sz / (float)reps;
            else e1 = e2 = e3 = 0.0f;
        } else {
            raise(c, PGA_ERR_BAD_PROGRAM);
            return;
        }
        if (reps > (1u << 24)) { raise(c, PGA_ERR_OVERFLOW_QUEUE); return; }
        bool start = mode == PGA_ANCHOR_TO_START;
        f.e1 = (float)start * e1 + (float)(!start) * e3;   // start piece
        f.e2 = e2;                                          // middle pieces
        f.e3 = (float)start * e3 + (float)(!start) * e1;   // end piece
        f.reps = reps;
        f.n = reps > 1 ? reps : 1u;
        f.off = 0.0f;
    } else if (kind == K_COMPSPLIT) {          // ComponentSplit.cuh:234-267
        f.n = 6;
    } else if (kind == K_PRISMSPLIT) {         // the sides, then the bottom and top caps
        f.n = (uint32_t)e.pv1 + 2u;
    } else {                                   // K_REPLICATE, Replicate.cuh:102-146
        f.n = e.nsucc();
    }
}

// ---- child descriptors (one per fan-out child, 16 bytes) ----
// jx = child index | axis << 30; o = translation of the child along the axis
// (Subdivide, Repeat), ns = the child's size along the axis. The fan-out lane
// computes o and ns with the reference's sequential running-offset sums; the
// child lane rebuilds the child from the parent record and the descriptor.
struct __align__(16) Desc {
    uint32_t parent;
    uint32_t jx;
    float o;
    float ns;
};
static_assert(sizeof(Desc) == 16, "Desc must be 16 bytes");
const uint32_t kDescSelf = 0x3FFFFFFFu;   // child = the parent record itself (chain continues)
const uint32_t kDescNone = 0x3FFFFFFEu;   // no child (static Subdivide zero factor)
const uint32_t kDescAxiom = 0x3FFFFFFDu;  // the axiom whose index is in `parent` (workgroup-local levels)

__device__ __forceinline__ void axis_vec(uint32_t axis, float o, float& tx, float& ty, float& tz) {
    tx = axis == 0 ? o : 0.0f;
    ty = axis == 1 ? o : 0.0f;
    tz = axis == 2 ? o : 0.0f;
}

// Geometry of child j of the fan-out parent `r`, j = 0, 1, ... in order:
// the translation o and size ns along the axis, from the reference's
// sequential running offsets (Subdivide.cuh:49-69, 218-236; Repeat.cuh:
// 146-197). Returns false when the child is not dispatched (a zero factor of
// the static Subdivide form).
template <class E>
__device__ __forceinline__ bool fan_geom(const Ctx& c, const Rec& r, const E& e, Fan& f, uint32_t j, float& o,
                                         float& ns, float* fac_out = nullptr) {
    const uint32_t kind = e.kind();
    o = 0.0f;
    ns = 0.0f;
    if (kind == K_SUBDIV) {
        const DevParam p = e.param(c, 1u + j);
        const bool is_static = (e.flags() & EF_STATIC) != 0;
        float fac;
        if (p.type == PGA_PT_SCALAR) {
            fac = p.a;
        } else if (!is_static && j > 0 && shape_dependent(p.type)) {
            // the DynParams form evaluates factor j on the symbol that child
            // j-1 left behind (size, model and seed overwritten in place)
            Rec sym = r;
            float tx, ty, tz;
            axis_vec(f.axis, f.prev_o, tx, ty, tz);
            if (c.geom) mul_translate(r.m, tx, ty, tz, sym.m);
            if (f.axis == 0) sym.s[0] = f.prev_ns; else if (f.axis == 1) sym.s[1] = f.prev_ns; else sym.s[2] = f.prev_ns;
            if (c.seed) sym.seed = next_seed(r.seed, j - 1u);
            fac = eval_param(c, p, sym, r.axiom, r.path, rec_depth(c, r));
        } else {
            fac = eval_param(c, p, r, r.axiom, r.path, rec_depth(c, r));
        }
        if (fac_out) *fac_out = fac;
        if (is_static && fac == 0.0f) return false;
        ns = (((float)(fac < 0.0f) * f.r2a) * fac) * -1.0f + ((float)(fac > 0.0f) * fac);
        o = f.off + (ns * 0.5f);
        f.prev_ns = ns;
        f.prev_o = o;
        f.off = f.off + ns;
        return true;
    }
    if (kind == K_REPEAT) {
        if (j == 0) {
            ns = f.e1;
            f.off = (f.total * -0.5f) + (ns * 0.5f);
        } else {
            if (j == 1) f.off = f.off + f.e1 * 0.5f;
            const bool middle = j + 1 < f.reps;   // pieces i = j + 1 in 2 .. reps-1
            ns = middle ? f.e2 : f.e3;
            f.off = f.off + ns * 0.5f;
        }
        o = f.off;
        if (j > 0 && j + 1 < f.reps) f.off = f.off + ns * 0.5f;
        return true;
    }
    return true;
}

// Child j of a fan-out parent from its geometry (o, ns along `axis`).
// Repeat pieces use seed modifier j + 1 (1 for the first piece, i for the
// middle pieces i = 2 .. reps-1, reps for the last, Repeat.cuh:156-184).
template <class E>
__device__ __forceinline__ bool build_child(const Ctx& c, const Rec& r, const E& e, uint32_t j,
                                            uint32_t axis, float o, float ns, Rec& ch) {
    const uint32_t kind = e.kind();
    ch = r;
    if (kind == K_SUBDIV || kind == K_REPEAT) {
        if (c.geom) {
            float tx, ty, tz;
            axis_vec(axis, o, tx, ty, tz);
            mul_translate(r.m, tx, ty, tz, ch.m);
        }
        if (axis == 0) ch.s[0] = ns; else if (axis == 1) ch.s[1] = ns; else ch.s[2] = ns;
        const bool sub = kind == K_SUBDIV;
        if (c.seed) ch.seed = next_seed(r.seed, sub ? j : j + 1u);
        return advance(c, ch, succ_of(c, e, sub ? j : 0u), j);
    }
    if (kind == K_COMPSPLIT && (e.flags() & EF_PARALLEL)) {   // ParallelFaceSplitter<Box> (ComponentSplit.cuh:179-230)
        // face j from its outward normal n and up vector y, x = cross(y, n);
        // the frame [x y n] and the translation size * n / 2 are formed in
        // FP32 as there (signed zeros included)
        const float sign = (float)(j & 1u) * 2.0f - 1.0f;
        const float fa = (float)(j < 2u);
        const float b1 = (float)(j & 2u) * 0.5f;
        const float b2 = (float)(j & 4u) * 0.25f;
        const float nx = b1 * sign, ny = b2 * sign, nz = fa * sign;
        const float yx = 0.0f, yy = 1.0f - b2, yz = b2 * -sign;
        if (c.geom) {
            const float xx = yy * nz - yz * ny, xy = yz * nx - yx * nz, xz = yx * ny - yy * nx;
            float tmp[12];
            mul_translate(r.m, (r.s[0] * nx) * 0.5f, (r.s[1] * ny) * 0.5f, (r.s[2] * nz) * 0.5f, tmp);
            const float b[16] = {xx, yx, nx, 0.0f, xy, yy, ny, 0.0f, xz, yz, nz, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
            mul34x44(tmp, b, ch.m);
        }
        const float sx = r.s[0], sy = r.s[1], sz = r.s[2];
        ch.s[0] = (sx * (fa + b2) + sy * 0.0f) + sz * b1;    // dot(size, abs(s0)), left to right
        ch.s[1] = (sx * 0.0f + sy * (1.0f - b2)) + sz * b2;  // dot(size, abs(s1))
        ch.s[2] = 1.0f;
        if (c.seed) ch.seed = next_seed(r.seed, (j + 2u) % 6u);
        return advance(c, ch, succ_of(c, e, j < 4 ? 2u : (j == 4 ? 1u : 0u)), j);
    }
    if (kind == K_COMPSPLIT) {                 // FaceSplitter<Box> (ComponentSplit.cuh:77-174)
        if (c.geom) {
            const float hx = r.s[0] * 0.5f, hy = r.s[1] * 0.5f, hz = r.s[2] * 0.5f;
            float tx = 0.0f, ty = 0.0f, tz = 0.0f;
            // face frame b (4x4), identity unless set below
            float b00 = 1.0f, b02 = 0.0f, b06 = 0.0f, b09 = 0.0f, b08 = 0.0f, b10 = 1.0f, b05 = 1.0f;
            bool rotate = true;
            if (j == 0) { tz = -hz; b00 = -1.0f; b10 = -1.0f; }                          // back
            else if (j == 1) { tz = hz; rotate = false; }                                // front
            else if (j == 2) { tx = -hx; b00 = 0.0f; b02 = -1.0f; b08 = 1.0f; b10 = 0.0f; }    // left
            else if (j == 3) { tx = hx; b00 = 0.0f; b02 = 1.0f; b08 = -1.0f; b10 = 0.0f; }     // right
            else if (j == 4) { ty = -hy; b05 = 0.0f; b06 = -1.0f; b09 = 1.0f; b10 = 0.0f; }    // bottom
            else { ty = hy; b05 = 0.0f; b06 = 1.0f; b09 = -1.0f; b10 = 0.0f; }                 // top
            float tmp[12];
            mul_translate(r.m, tx, ty, tz, tmp);
            if (rotate) {
                const float b[16] = {b00, 0.0f, b02, 0.0f, 0.0f, b05, b06, 0.0f,
                                     b08, b09, b10, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
                mul34x44(tmp, b, ch.m);
            } else {
                set_model(ch, tmp);
            }
        }
        const float sx = r.s[0], sy = r.s[1], sz = r.s[2];
        ch.s[0] = (j == 2 || j == 3) ? sz : sx;
        ch.s[1] = j < 4 ? sy : sz;
        ch.s[2] = 1.0f;
        const uint32_t mod = j < 4 ? j + 2u : (j == 4 ? 1u : 0u);
        if (c.seed) ch.seed = next_seed(r.seed, mod);
        return advance(c, ch, succ_of(c, e, j < 4 ? 2u : (j == 4 ? 1u : 0u)), j);
    }
    if (kind == K_PRISMSPLIT) {                // FaceSplitter<DynamicRightPrism> (ComponentSplit.cuh:23-75)
        const uint32_t n = (uint32_t)e.pv1;
        const uint32_t vb = __float_as_uint(e.pv0);
        if (j < n) {
            // side j, from vertex j-1 (n-1 for j = 0) to vertex j: getVertex
            // reverses the order and mirrors y when inverted; both scaled by
            // float2x2::scale(size.xz()) with the zero terms of the 2x2
            // product, y negated, then the edge frame (facing)
            const bool inv = (r.ed & kInvert) != 0u;
            const uint32_t i0 = j == 0u ? n - 1u : j - 1u;
            const DevParam q0 = c.prog.params[vb + (inv ? n - i0 - 1u : i0)];
            const DevParam q1 = c.prog.params[vb + (inv ? n - j - 1u : j)];
            const float a0 = q0.a, a1 = inv ? -q0.b : q0.b;
            const float b0 = q1.a, b1 = inv ? -q1.b : q1.b;
            const float sx = r.s[0], sz = r.s[2];
            const float p0x = sx * a0 + 0.0f * a1, p0y = -(0.0f * a0 + sz * a1);
            const float p1x = sx * b0 + 0.0f * b1, p1y = -(0.0f * b0 + sz * b1);
            float dx = p1x - p0x, dy = p1y - p0y;
            const float x = __builtin_sqrtf(dx * dx + dy * dy);
            const float ox = p0x + (p1x - p0x) * 0.5f, oy = p0y + (p1y - p0y) * 0.5f;
            const float rx = 1.0f / x;
            dx = dx * rx;
            dy = dy * rx;
            if (c.geom) {
                float tmp[12];
                mul_translate(r.m, ox, 0.0f, oy, tmp);
                const float b[16] = {dx, 0.0f, -dy, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f,
                                     dy, 0.0f, dx, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
                mul34x44(tmp, b, ch.m);
            }
            ch.s[0] = x;
            ch.s[2] = 1.0f;
            if (c.seed) ch.seed = next_seed(r.seed, j + 2u);
            ch.ed &= ~kInvert;
            return advance(c, ch, succ_of(c, e, 2u), j);
        }
        // the bottom cap (j = n, inverted) and the top cap (j = n + 1):
        // translate(0, size.y * sign * 0.5, 0), then the cap frame
        const bool bottom = j == n;
        const float sign = bottom ? -1.0f : 1.0f;
        if (c.geom) {
            float tmp[12];
            mul_translate(r.m, 0.0f, (r.s[1] * sign) * 0.5f, 0.0f, tmp);
            const float b[16] = {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, sign, 0.0f,
                                 0.0f, -sign, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
            mul34x44(tmp, b, ch.m);
        }
        ch.s[1] = r.s[2];
        ch.s[2] = 1.0f;
        if (c.seed) ch.seed = next_seed(r.seed, bottom ? 1u : 0u);
        ch.ed = bottom ? (ch.ed | kInvert) : (ch.ed & ~kInvert);
        return advance(c, ch, succ_of(c, e, bottom ? 1u : 0u), j);
    }
    // K_REPLICATE
    return advance(c, ch, succ_of(c, e, j), j);
}

} // namespace pga
