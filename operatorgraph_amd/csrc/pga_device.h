// pga_device.h - device-side data layout and shape-grammar operators for
// gfx950 (wave64). Included by the kernel translation unit only.
//
// Operator arithmetic follows the reference's FP32 operation order exactly
// (the engine is compiled with -ffp-contract=off), so instance records are
// bit-identical to a host derivation of the same grammar.
#pragma once

#ifndef __HIPCC_RTC__   // kernels compiled at run time get the HIP built-ins from hipRTC
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "pga.h"

namespace pga {

// ---- compact program (built from procedures + dispatch entries) ----
enum Kind : uint8_t {
    K_DISCARD = 0,
    K_GENERATE,
    K_TRANSLATE,
    K_ROTATE,
    K_SCALE,
    K_SWAPSIZE,
    K_EXTRUDE,
    K_IFSIZELESS,
    K_IF,
    K_RANDOM,
    K_SUBDIV,
    K_REPEAT,
    K_COMPSPLIT,
    K_REPLICATE,
    K_COLLIDER,       // stores the shape for IfCollides of later phases, one successor
    K_IFCOLLIDES,     // successor 0 if the shape collides with a stored collider, else 1
    K_SETAS,          // the shape becomes a dynamic polygon or right prism, one successor
    K_PRISMSPLIT,     // ComponentSplit of a dynamic right prism: sides, bottom, top
};
enum EntryFlags : uint8_t {
    EF_STATIC = 1,    // static-parameter operator form
    EF_LEAF = 2,      // fan-out whose children never reach another fan-out
    EF_SHAPE_DEP = 4, // some parameter reads the shape (dyn Subdivide re-evaluation)
    EF_DIRECT = 8,    // leaf fan-out whose children are all Generate (constant terminal) or
                      // Discard: its terminal counts are a table (aux[1])
    EF_PARALLEL = 16, // ComponentSplit<true>: ParallelFaceSplitter face frames and seeds
};

// One dispatch entry, 48 bytes. The first three parameters are held inline
// when they are constants (bit k of `inl`), and the first three successors
// always, so most operators cost a single table read.
struct __align__(16) DevEntry {
    uint8_t kind;
    uint8_t shape;      // pga_shape_type of the procedure
    uint8_t flags;
    uint8_t nparams;
    uint8_t nsucc;
    uint8_t inl;        // bit k: parameter k (< 3) is the constant pv[k]
    uint8_t tslot;      // Generate with a constant terminal: its dense slot (else 0xFF)
    uint8_t ndirect;    // direct fan-out: (slot, count) pairs in aux
    uint32_t pbase;     // first DevParam
    uint32_t sbase;     // first successor in the successor pool
    float pv[3];        // inline constant parameters; K_PRISMSPLIT: the bits of the
                        // vertex list's first DevParam, and its vertex count
    int32_t sv[3];      // successors 0..2
    // fan-outs: aux[0] = skip mask, bit j: child j goes to a Discard or
    // nowhere (a Repeat's children all follow bit 0); direct fan-outs:
    // aux[1] bytes k = (terminal slot, count) pair k, count per child set
    // (times the repetitions of a Repeat)
    uint32_t aux[2];
};

struct DevParam {       // 16 bytes
    uint32_t type;      // pga_parameter_type
    float a, b, c;      // SCALAR a; RAND a,b; SHAPE_ATTR a; PHILOX a,b,c; AXIOM_PARAM a;
                        // EXP: bits of a = offset of the expression in the pool
};

struct Prog {
    const DevEntry* __restrict__ entries;
    const DevParam* __restrict__ params;
    const int32_t* __restrict__ succ;
    const uint32_t* __restrict__ exp;   // raw bits of packed expressions
    const uint8_t* __restrict__ term_slot;  // terminal index -> dense slot (0xFF: none)
    const uint8_t* __restrict__ slot_term;  // dense slot -> terminal index
    uint32_t nentries;
    uint32_t nslots;                     // terminal slots used by the program
};

// ---- shape record (80 bytes, 5 x 16 B) ----
struct __align__(16) Rec {
    float m[12];        // row-major 3x4 model, row 3 implicit (0 0 0 1)
    float s[3];         // size
    float seed;
    float custom;
    uint32_t ed;        // entry | invert << 15 | depth << 16; in multi-phase
                        // programs also phase << 28 (depth < 4096); invert: a
                        // dynamic polygon's or prism's vertex order is reversed
    uint32_t axiom;
    uint32_t path;
};
static_assert(sizeof(Rec) == 80, "Rec must be 80 bytes");

const uint32_t kInvert = 0x8000u;   // Rec::ed bit 15
__device__ __forceinline__ uint32_t rec_entry(const Rec& r) { return r.ed & 0x7FFFu; }

struct Desc;

// Per-launch context shared by every lane.
struct Ctx {
    Prog prog;                          // program tables (in LDS during a derivation)
    const float* __restrict__ aparams;  // per-axiom parameters
    uint32_t aparams_per;
    uint32_t naxioms_total;
    uint32_t key0, key1;
    pga_instance* const* inst;          // [PGA_MAX_TERMINALS]
    const unsigned long long* inst_cap; // [PGA_MAX_TERMINALS]
    unsigned long long* term_counts;    // [PGA_MAX_TERMINALS]
    Rec* par_out;                       // parent records of this level
    struct Desc* desc_out;              // child descriptors for the next level
    uint32_t* par_count;
    uint32_t* desc_count;
    uint32_t par_cap, desc_cap;
    uint32_t* err;                      // bit (1 << pga_status) per error raised
    uint32_t max_chain;
    uint32_t geom;                      // 0: skip model-matrix arithmetic (counting only)
    uint32_t seed;                      // 0: no parameter reads a seed, the seed chain is skipped
    uint32_t poly;                      // 1: the program has dynamic polygons (Rec::ed's invert bit)
    uint32_t tape;                      // 1: recording a tape (no in-place leaf expansion)
    uint32_t prog_lds;                  // 1: the program tables are the LDS copy (derivation kernels), 0: global
    // multi-phase programs: depth bits of Rec::ed (0xFFFF: one phase, 0x0FFF:
    // the phase in bits 28-31), the phase this launch derives, and the
    // collider store (5 x 16 bytes per collider: model rows, size and tag,
    // shape type and phase), of which the first *coll_visible were stored by
    // earlier phases
    uint32_t ctxs;                      // 1: the kernel runs Collider / IfCollides and phases
    uint32_t dmask;
    uint32_t phase;
    uint32_t wr;                        // 1: write pass (colliders are stored once, here)
    float4* coll;
    uint32_t* coll_count;
    const uint32_t* coll_visible;
    uint32_t coll_cap;
    // linear BVH over the visible colliders in depth-first order (NULL: none
    // built): per node two 16-byte words, box min xyz, max x | max y, max z,
    // the node after its subtree, the collider of a leaf (~0: internal)
    const float4* bvh_flat;
    // per child of the level: the IfCollides outcomes of the count pass, bit
    // k for the child's k-th IfCollides (k < 32), replayed by the write pass
    uint32_t* coll_bits;
    uint32_t coll_bits_cap;
};

// A lane's IfCollides outcomes for its current child: recorded by the count
// pass, replayed by the write pass (both passes meet the same IfCollides in
// the same order).
struct CollMemo {
    uint32_t bits;
    uint32_t n;        // IfCollides met so far
    uint32_t mode;     // 0: neither, 1: record, 2: replay
};

// Axis-aligned box around a Box or Quad shape (a Quad with its z extent,
// which the collision test keeps on the Quad's own face axis) under its
// row-major 3x4 model, widened by 1e-5 of the magnitudes of its centre and
// half extents, so that FP32 rounding never separates the boxes of shapes
// the collision test finds touching.
__device__ __forceinline__ void shape_box(const float m[12], const float s[3], float lo[3], float hi[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const float h = 0.5f * (fabsf(m[4 * r]) * fabsf(s[0]) + fabsf(m[4 * r + 1]) * fabsf(s[1]) +
                                fabsf(m[4 * r + 2]) * fabsf(s[2]));
        const float c = m[4 * r + 3];
        const float w = h + 1e-5f * (fabsf(c) + h) + 1e-20f;
        lo[r] = c - w;
        hi[r] = c + w;
    }
}

__device__ __forceinline__ uint32_t rec_depth(const Ctx& c, const Rec& r) { return (r.ed >> 16) & c.dmask; }
__device__ __forceinline__ bool multi_phase(const Ctx& c) { return c.ctxs && c.dmask != 0xFFFFu; }
__device__ __forceinline__ uint32_t rec_phase(const Ctx& c, const Rec& r) { return multi_phase(c) ? r.ed >> 28 : 0u; }

__device__ __forceinline__ void raise(const Ctx& c, int status) {
    atomicOr(c.err, 1u << status);
}

// ---- float <-> bits, conversions ----
__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

// round-down float->uint with saturation (the reference's __float2uint_rd)
__device__ __forceinline__ uint32_t f2u_rd(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// ---- hybrid Tausworthe seed chain (Random.cuh:15-57) ----
__device__ __forceinline__ uint32_t taus_step(uint32_t z, int s1, int s2, int s3, uint32_t m) {
    uint32_t b = (((z << s1) ^ z) >> s2);
    return (((z & m) << s3) ^ b);
}
// x % 0xFFFFFFFF for a 32-bit x: only x = 0xFFFFFFFF wraps
__device__ __forceinline__ uint32_t mod_um(uint32_t x) { return x == 0xFFFFFFFFu ? 0u : x; }
__device__ __forceinline__ float next_seed(float base, uint32_t modifier) {
    const uint32_t um = 0xFFFFFFFFu;
    uint32_t z = (f2u_rd(base) * um) + mod_um(modifier * 41231278u);
    uint32_t z1 = mod_um(z + 7873234u);
    uint32_t z2 = mod_um(z1 + 1237217827u);
    uint32_t z3 = mod_um(z2 + 87382u);
    uint32_t z4 = mod_um(z3 + 1410022u);
    uint32_t t = taus_step(z1, 13, 19, 12, 4294967294u) ^ taus_step(z2, 2, 25, 4, 4294967288u) ^
                 taus_step(z3, 3, 11, 17, 4294967280u) ^ (1664525u * z4 + 1013904223u);
    return (float)(2.3283064365387e-10 * (double)t);
}
__device__ __forceinline__ float range(float seed, float mn, float mx) {
    if (seed < 1.0f) return seed * (mx - mn) + mn;
    return mx;
}

// ---- Philox4x32-10 (Salmon et al., SC'11) ----
__device__ __forceinline__ uint32_t philox_x0(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    return c0;
}

// Order-independent identifier of child `ordinal` of a shape (murmur3 fmix32).
__device__ __forceinline__ uint32_t path_mix(uint32_t path, uint32_t ordinal) {
    uint32_t h = path * 0x01000193u ^ (ordinal + 1u) * 0x9E3779B9u;
    h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
    return h;
}

// ---- sin/cos: Cody-Waite reduction by pi/2 (three-float split) and Taylor
// polynomials of sin (to r^9) and cos (to r^10) on |r| <= pi/4 ----
__device__ __forceinline__ void sincos_f32(float x, float* s_out, float* c_out) {
    const float two_over_pi = 0.636619772367581343f;
    const float p1 = 1.5703125f;
    const float p2 = 4.83751296997070312e-4f;
    const float p3 = 7.54978995489188216e-8f;
    float k = __builtin_rintf(x * two_over_pi);
    float r = ((x - k * p1) - k * p2) - k * p3;
    float r2 = r * r;
    float sp = r + r * r2 * (-1.66666666666666667e-1f + r2 * (8.33333333333333333e-3f +
                r2 * (-1.98412698412698413e-4f + r2 * 2.75573192239858907e-6f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (4.16666666666666667e-2f + r2 * (-1.38888888888888889e-3f +
                r2 * (2.48015873015873016e-5f + r2 * -2.75573192239858907e-7f))));
    int q = (int)((long long)k & 3);
    float s, c;
    if (q == 0) { s = sp; c = cp; }
    else if (q == 1) { s = cp; c = -sp; }
    else if (q == 2) { s = -sp; c = -cp; }
    else { s = -cp; c = sp; }
    *s_out = s; *c_out = c;
}

// ---- matrices: c = a(3x4, implicit row 0001) * b(4x4), sums left to right
// as PGA's math::float4x4 operator* (matrix.h:776-781) ----
__device__ __forceinline__ void mul34x44(const float a[12], const float b[16], float c[12]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            c[i * 4 + j] = ((a[i * 4 + 0] * b[j] + a[i * 4 + 1] * b[4 + j]) + a[i * 4 + 2] * b[8 + j]) +
                           a[i * 4 + 3] * b[12 + j];
}
// model4 * translate(t) (matrix.h:904-910): the general product with the
// translate matrix's 0/1 entries, kept term by term for exact signed zeros.
__device__ __forceinline__ void mul_translate(const float a[12], float tx, float ty, float tz, float c[12]) {
    const float b[16] = {1.0f, 0.0f, 0.0f, tx, 0.0f, 1.0f, 0.0f, ty, 0.0f, 0.0f, 1.0f, tz,
                         0.0f, 0.0f, 0.0f, 1.0f};
    mul34x44(a, b, c);
}

__device__ __forceinline__ float shape_at(const Rec& r, uint32_t i) {
    // Shape::at (Shape.cuh:120-171); indices are validated on load
    float v = 0.0f;
#pragma unroll
    for (int k = 0; k < 12; ++k) if (i == (uint32_t)k) v = r.m[k];
    if (i == 12) v = r.s[0];
    if (i == 13) v = r.s[1];
    if (i == 14) v = r.s[2];
    if (i == 15) v = r.seed;
    if (i == 16) v = r.custom;
    return v;
}

// unpackFloat2 (PackUtils.h:39-43)
__device__ __forceinline__ void unpack_float2(float a, float* x, float* y) {
    *y = fmodf(a, 4096.0f) / (4096.0f - 1.0f);
    *x = floorf(a / 4096.0f) / (4096.0f - 1.0f);
}

__device__ __forceinline__ float exp_operate(uint32_t op, float p1, float p2) {
    switch (op) {
    case PGA_OPT_ADD: return p1 + p2;
    case PGA_OPT_SUB: return p1 - p2;
    case PGA_OPT_DIV: return p1 / p2;
    case PGA_OPT_MULTI: return p1 * p2;
    case PGA_OPT_EQ: return (float)(p1 == p2);
    case PGA_OPT_NEQ: return (float)(p1 != p2);
    case PGA_OPT_LT: return (float)(p1 < p2);
    case PGA_OPT_GT: return (float)(p1 > p2);
    case PGA_OPT_LEQ: return (float)(p1 <= p2);
    case PGA_OPT_GEQ: return (float)(p1 >= p2);
    case PGA_OPT_AND: return (float)(p1 != 0.0f && p2 != 0.0f);
    default: return (float)(p1 != 0.0f || p2 != 0.0f);
    }
}

// EncExp::eval (Parameters.cuh:563-664). Expressions are validated on load
// (length <= 15, stack depth <= 4, indices in range), so the four stack
// slots live in registers. Out of line: the shape attributes come in as
// scalars so the record never has to be addressable. One instance per
// address space of the code (WordT: LDS words in the derivation kernels,
// global words in the adjoint), so that its reads are not flat loads.
template <typename WordT>
__device__ __noinline__ float exp_eval(const WordT* __restrict__ b, float m0, float m1, float m2, float m3,
                                       float m4, float m5, float m6, float m7, float m8, float m9, float m10,
                                       float m11, float s0x, float s1x, float s2x, float seed, float custom) {
    float st0 = 0.0f, st1 = 0.0f, st2 = 0.0f, st3 = 0.0f;
    int top = 0;
    uint32_t size = (uint32_t)bitsf(b[0]);
    for (uint32_t i = 1; i <= size; ++i) {
        uint32_t raw = b[i];
        uint32_t type = raw >> 30;
        float x = bitsf(raw << 2);
        float v;
        if (type == PGA_ORT_OP) {
            uint32_t op = (uint32_t)__builtin_rintf(x);
            float p2 = top == 1 ? st0 : top == 2 ? st1 : top == 3 ? st2 : st3;
            float p1 = top == 2 ? st0 : top == 3 ? st1 : st2;
            top -= 2;
            v = exp_operate(op, p1, p2);
        } else if (type == PGA_ORT_SCALAR) {
            v = x;
        } else if (type == PGA_ORT_RAND) {
            float mn, mx;
            unpack_float2(x, &mn, &mx);
            v = range(seed, mn, mx);
        } else {
            uint32_t idx = (uint32_t)__builtin_rintf(x);
            const float a[17] = {m0, m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11, s0x, s1x, s2x, seed, custom};
            v = 0.0f;
#pragma unroll
            for (int k = 0; k < 17; ++k) v = idx == (uint32_t)k ? a[k] : v;
        }
        if (top == 0) st0 = v; else if (top == 1) st1 = v; else if (top == 2) st2 = v; else st3 = v;
        ++top;
    }
    return top == 1 ? st0 : top == 2 ? st1 : top == 3 ? st2 : st3;
}

__device__ __noinline__ float philox_uniform(uint32_t axiom, uint32_t path, uint32_t depth, uint32_t stream,
                                             uint32_t k0, uint32_t k1) {
    uint32_t u = philox_x0(axiom, path, depth, stream, k0, k1);
    return (float)(u >> 8) * 5.9604644775390625e-8f;   // top 24 bits, exact
}

// dynEval (Parameters.cuh:669-711) and the project parameter types.
// (axiom, path, depth) identify the shape whose operator runs.
__device__ __forceinline__ float eval_param(const Ctx& c, const DevParam& p, const Rec& r,
                                            uint32_t axiom, uint32_t path, uint32_t depth) {
    const uint32_t t = p.type;
    if (t == PGA_PT_SCALAR) return p.a;
    if (t == PGA_PT_AXIOM_PARAM) return c.aparams[(size_t)axiom * c.aparams_per + (uint32_t)p.a];
    if (t == PGA_PT_RAND) return range(r.seed, p.a, p.b);
    if (t == PGA_PT_SHAPE_ATTR) return shape_at(r, (uint32_t)p.a);
    if (t == PGA_PT_PHILOX) {
        float uf = philox_uniform(axiom, path, depth, (uint32_t)p.c, c.key0, c.key1);
        return (p.b - p.a) * uf + p.a;
    }
    typedef __attribute__((address_space(3))) const uint32_t lds_u32;
    if (c.prog_lds)
        return exp_eval<lds_u32>((const lds_u32*)(c.prog.exp + fbits(p.a)), r.m[0], r.m[1], r.m[2], r.m[3], r.m[4],
                                 r.m[5], r.m[6], r.m[7], r.m[8], r.m[9], r.m[10], r.m[11], r.s[0], r.s[1], r.s[2],
                                 r.seed, r.custom);
    typedef __attribute__((address_space(1))) const uint32_t glb_u32;   // the adjoint reads the global tables
    return exp_eval<glb_u32>((const glb_u32*)(c.prog.exp + fbits(p.a)), r.m[0], r.m[1], r.m[2], r.m[3], r.m[4],
                             r.m[5], r.m[6], r.m[7], r.m[8], r.m[9], r.m[10], r.m[11], r.s[0], r.s[1], r.s[2],
                             r.seed, r.custom);
}

__device__ __forceinline__ bool shape_dependent(uint32_t type) {
    return type == PGA_PT_RAND || type == PGA_PT_SHAPE_ATTR || type == PGA_PT_EXP;
}

} // namespace pga
