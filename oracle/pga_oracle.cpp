// pga_oracle.cpp - TEST INFRASTRUCTURE ONLY (see pga_oracle.h).
//
// A CPU restatement of the reference's derivation semantics, written from the
// reference's source text. Every function cites the reference lines it
// follows. Build: oracle/Makefile (g++ -O2 -ffp-contract=off).
//
// Nothing in the HIP engine links this file; tests compare the two.
#include "pga_oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Reverse-mode scalar for the adjoint twin: a Wengert tape of binary nodes.
// ---------------------------------------------------------------------------
struct Tape {
    std::vector<int32_t> a, b;
    std::vector<double> wa, wb;
    int32_t push(int32_t ia, double w0, int32_t ib, double w1) {
        a.push_back(ia); b.push_back(ib); wa.push_back(w0); wb.push_back(w1);
        return (int32_t)a.size() - 1;
    }
    int32_t leaf() { return push(-1, 0.0, -1, 0.0); }
};
thread_local Tape* g_tape = nullptr;

struct Var {
    double v;
    int32_t id;
    Var() : v(0.0), id(-1) {}
    Var(double x) : v(x), id(-1) {}
    Var(float x) : v((double)x), id(-1) {}
    Var(int x) : v((double)x), id(-1) {}
    Var(unsigned x) : v((double)x), id(-1) {}
};
inline Var mk(double v, int32_t ia, double wa, int32_t ib, double wb) {
    Var r(v);
    if (ia >= 0 || ib >= 0) r.id = g_tape->push(ia, wa, ib, wb);
    return r;
}
inline Var operator+(Var x, Var y) { return mk(x.v + y.v, x.id, 1.0, y.id, 1.0); }
inline Var operator-(Var x, Var y) { return mk(x.v - y.v, x.id, 1.0, y.id, -1.0); }
inline Var operator*(Var x, Var y) { return mk(x.v * y.v, x.id, y.v, y.id, x.v); }
inline Var operator/(Var x, Var y) { return mk(x.v / y.v, x.id, 1.0 / y.v, y.id, -x.v / (y.v * y.v)); }
inline Var operator-(Var x) { return mk(-x.v, x.id, -1.0, -1, 0.0); }
inline Var& operator+=(Var& x, Var y) { x = x + y; return x; }
inline bool operator<(Var x, Var y) { return x.v < y.v; }
inline bool operator>(Var x, Var y) { return x.v > y.v; }
inline bool operator<=(Var x, Var y) { return x.v <= y.v; }
inline bool operator>=(Var x, Var y) { return x.v >= y.v; }
inline bool operator==(Var x, Var y) { return x.v == y.v; }
inline bool operator!=(Var x, Var y) { return x.v != y.v; }

// ---------------------------------------------------------------------------
// Scalar traits: the same generic code runs in float (reference FP32), double
// (FP64 twin) and Var (adjoint twin).
// ---------------------------------------------------------------------------
template <class T> struct S;
template <> struct S<float> {
    static float c(double x) { return (float)x; }
    static double val(float x) { return x; }
    static float floor_(float x) { return std::floor(x); }
    static float fabs_(float x) { return std::fabs(x); }
    static float fmod_(float x, float y) { return std::fmod(x, y); }
    static float deg2rad() { return 3.1415926535897932384626434f / 180.0f; } // Rotate.cuh:33-35
    static float sqrt_(float x) { return std::sqrt(x); }                      // correctly rounded
};
template <> struct S<double> {
    static double c(double x) { return x; }
    static double val(double x) { return x; }
    static double floor_(double x) { return std::floor(x); }
    static double fabs_(double x) { return std::fabs(x); }
    static double fmod_(double x, double y) { return std::fmod(x, y); }
    static double deg2rad() { return 3.1415926535897932384626434 / 180.0; }
    static double sqrt_(double x) { return std::sqrt(x); }
};
template <> struct S<Var> {
    static Var c(double x) { return Var(x); }
    static double val(Var x) { return x.v; }
    static Var floor_(Var x) { return Var(std::floor(x.v)); }
    static Var fabs_(Var x) { return x.v < 0 ? -x : x; }
    static Var fmod_(Var x, Var y) {
        double q = std::trunc(x.v / y.v);
        return mk(std::fmod(x.v, y.v), x.id, 1.0, y.id, -q);
    }
    static Var deg2rad() { return Var(3.1415926535897932384626434 / 180.0); }
    static Var sqrt_(Var x) {
        const double r = std::sqrt(x.v);
        return mk(r, x.id, r > 0.0 ? 0.5 / r : 0.0, -1, 0.0);
    }
};

// ---------------------------------------------------------------------------
// sin/cos. The reference calls std::sin/std::cos (host) and sinf/cosf
// (device) in rotateXYZ (deps/math/include/math/matrix.h:876-888). The oracle
// uses one fixed algorithm so that its FP32 results are reproducible:
// Cody-Waite reduction by pi/2 split into three floats (P1 has 8 significant
// bits, P2 12), then the Taylor series of sin (to r^9) and cos (to r^10) on
// |r| <= pi/4 in Horner form. Tests pin it against libm within 2 ulp.
// ---------------------------------------------------------------------------
inline void sincos_f32(float x, float* s_out, float* c_out) {
    const float two_over_pi = 0.636619772367581343f;
    const float p1 = 1.5703125f;
    const float p2 = 4.83751296997070312e-4f;
    const float p3 = 7.54978995489188216e-8f;
    float k = std::rint(x * two_over_pi);
    float r = ((x - k * p1) - k * p2) - k * p3;
    float r2 = r * r;
    float sp = r + r * r2 * (-1.66666666666666667e-1f + r2 * (8.33333333333333333e-3f +
                r2 * (-1.98412698412698413e-4f + r2 * 2.75573192239858907e-6f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (4.16666666666666667e-2f + r2 * (-1.38888888888888889e-3f +
                r2 * (2.48015873015873016e-5f + r2 * -2.75573192239858907e-7f))));
    int q = (int)((long long)k & 3);
    float s, c;
    switch (q) {
    case 0: s = sp; c = cp; break;
    case 1: s = cp; c = -sp; break;
    case 2: s = -sp; c = -cp; break;
    default: s = -cp; c = sp; break;
    }
    *s_out = s; *c_out = c;
}
inline void sincos_t(float x, float* s, float* c) { sincos_f32(x, s, c); }
inline void sincos_t(double x, double* s, double* c) { *s = std::sin(x); *c = std::cos(x); }
inline void sincos_t(Var x, Var* s, Var* c) {
    double sv = std::sin(x.v), cv = std::cos(x.v);
    *s = mk(sv, x.id, cv, -1, 0.0);
    *c = mk(cv, x.id, -sv, -1, 0.0);
}

// ---------------------------------------------------------------------------
// Random.cuh:15-57 - hybrid Tausworthe seed chain.
// ---------------------------------------------------------------------------
inline uint32_t taus_step(uint32_t z, int s1, int s2, int s3, uint32_t m) {
    uint32_t b = (((z << s1) ^ z) >> s2);
    return (((z & m) << s3) ^ b);
}
inline uint32_t lcg_step(uint32_t z, uint32_t a, uint32_t c) { return a * z + c; }

// The four generators combined by hybridTaus (Random.cuh:36-43): three
// Tausworthe steps (L'Ecuyer's taus88 components, Math. Comp. 65 (1996)) and
// one linear congruential step. `which` = 0, 1, 2: Tausworthe; 3: LCG.
uint32_t seed_component(uint32_t z, uint32_t which) {
    switch (which) {
    case 0: return taus_step(z, 13, 19, 12, 4294967294u);
    case 1: return taus_step(z, 2, 25, 4, 4294967288u);
    case 2: return taus_step(z, 3, 11, 17, 4294967280u);
    default: return lcg_step(z, 1664525u, 1013904223u);
    }
}

// float -> uint with round-down and saturation (__float2uint_rd, Random.cuh:27)
inline uint32_t f2u_rd(float x) {
    if (!(x > 0.0f)) return 0u;           // negatives and NaN saturate to 0
    if (x >= 4294967296.0f) return UINT_MAX;
    return (uint32_t)x;                   // truncation == floor for x >= 0
}

// The component states z1..z4 that hybridTaus derives from (base, modifier)
// (Random.cuh:26-35).
void seed_state(float base, uint32_t modifier, uint32_t zs[4]) {
    uint32_t z = (f2u_rd(base) * UINT_MAX) + (modifier * 41231278u) % UINT_MAX;
    zs[0] = (z + 7873234u) % UINT_MAX;
    zs[1] = (zs[0] + 1237217827u) % UINT_MAX;
    zs[2] = (zs[1] + 87382u) % UINT_MAX;
    zs[3] = (zs[2] + 1410022u) % UINT_MAX;
}

// The combined 32-bit value to a float (Random.cuh:44-48): a double product
// with 2.3283064365387e-10 (2^-32 rounded up), then a float conversion; the
// largest values round to 1.0f.
inline float seed_unit(uint32_t t) { return (float)(2.3283064365387e-10 * (double)t); }

float hybrid_taus(float base, uint32_t modifier) {
    uint32_t zs[4];
    seed_state(base, modifier, zs);
    return seed_unit(seed_component(zs[0], 0) ^ seed_component(zs[1], 1) ^ seed_component(zs[2], 2) ^
                     seed_component(zs[3], 3));
}
inline float next_seed(float base, uint32_t modifier) { return hybrid_taus(base, modifier); }
// Random::range (Random.cuh:51-57) in the twin's precision
template <class T> T range_t(float seed, T mn, T mx) {
    if (seed < 1.0f) return S<T>::c(seed) * (mx - mn) + mn;
    return mx;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): project extension for
// per-shape rule choice; no reference counterpart.
// ---------------------------------------------------------------------------
void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Identifier of a derived shape, independent of execution order: the child
// `ordinal` of a shape with identifier `path` (murmur3 finaliser).
uint32_t path_mix(uint32_t path, uint32_t ordinal) {
    uint32_t h = path * 0x01000193u ^ (ordinal + 1u) * 0x9E3779B9u;
    h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
    return h;
}

// ---------------------------------------------------------------------------
// PackUtils.h:18-43 - expression operand packing.
// ---------------------------------------------------------------------------
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline void unpack_operand(float a, int* type, float* x) {
    uint32_t r = fbits(a);
    *type = (int)(r >> 30);
    *x = bitsf(r << 2);
}
inline void unpack_float2(float a, float* x, float* y) {
    *y = std::fmod(a, 4096.0f) / (4096.0f - 1.0f);
    *x = std::floor(a / 4096.0f) / (4096.0f - 1.0f);
}

// ---------------------------------------------------------------------------
// Shapes and matrices (Shape.cuh:16-192; deps/math/include/math/matrix.h).
// The model is a row-major 3x4 whose implicit fourth row is (0 0 0 1);
// products keep the reference's left-to-right sum order
// (matrix.h:776-781): c_ij = ((a_i0 b_0j + a_i1 b_1j) + a_i2 b_2j) + a_i3 b_3j.
// ---------------------------------------------------------------------------
template <class T> struct Shape {
    T m[12];
    T size[3];
    float seed;
    float custom;
};
// The vertex list of a dynamic polygon or right prism (DynamicPolygon.cuh:
// 16-100, DynamicRightPrism.cuh:18-130): numSides vertices in the unit
// square, and `invert`, set on the bottom cap of a ComponentSplit, which
// reverses the order, mirrors y and negates the first two model columns of
// its instance.
struct Poly {
    uint32_t n = 0;
    uint32_t invert = 0;
    float v[2 * PGA_MAX_NUM_SIDES] = {};
};
inline bool is_polygon(int32_t st) {
    return st == PGA_SHAPE_DYNAMIC_CONVEX_POLYGON || st == PGA_SHAPE_DYNAMIC_POLYGON;
}
inline bool is_prism(int32_t st) {
    return st == PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM || st == PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
}
inline bool has_vertices(int32_t st) { return is_polygon(st) || is_prism(st); }
// getVertex: vertex i, or with `invert` vertex n-1-i with y negated
inline void poly_vertex(const Poly& p, uint32_t i, float& x, float& y) {
    if (p.invert) {
        x = p.v[2 * (p.n - i - 1)];
        y = -p.v[2 * (p.n - i - 1) + 1];
    } else {
        x = p.v[2 * i];
        y = p.v[2 * i + 1];
    }
}

struct Meta {
    int32_t entry;
    uint32_t axiom, path, depth;
    int32_t shape_type;
    uint32_t phase;    // multi-phase programs: the phase the shape is derived in
    Poly poly;         // dynamic polygons and prisms
};

template <class T> void mul34x44(const T a[12], const T b[16], T c[12]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j)
            c[i * 4 + j] = ((a[i * 4 + 0] * b[0 * 4 + j] + a[i * 4 + 1] * b[1 * 4 + j]) +
                            a[i * 4 + 2] * b[2 * 4 + j]) + a[i * 4 + 3] * b[3 * 4 + j];
}
template <class T> void mat_translate(T x, T y, T z, T o[16]) {        // matrix.h:904-910
    const T one = S<T>::c(1.0), zero = S<T>::c(0.0);
    T v[16] = {one, zero, zero, x, zero, one, zero, y, zero, zero, one, z, zero, zero, zero, one};
    for (int i = 0; i < 16; ++i) o[i] = v[i];
}
template <class T> void mat_scale(T x, T y, T z, T o[16]) {            // matrix.h:912-919
    const T one = S<T>::c(1.0), zero = S<T>::c(0.0);
    T v[16] = {x, zero, zero, zero, zero, y, zero, zero, zero, zero, z, zero, zero, zero, zero, one};
    for (int i = 0; i < 16; ++i) o[i] = v[i];
}
template <class T> void mat_const(const float f[16], T o[16]) {
    for (int i = 0; i < 16; ++i) o[i] = S<T>::c(f[i]);
}
template <class T> void apply_translate(Shape<T>& s, const T base[12], T x, T y, T z) {
    T t[16];
    mat_translate(x, y, z, t);
    mul34x44(base, t, s.m);
}

// Shape::at (Shape.cuh:120-171)
template <class T> T shape_at(const Shape<T>& s, uint32_t i) {
    if (i < 12) return s.m[i];
    if (i < 15) return s.size[i - 12];
    if (i == 15) return S<T>::c(s.seed);
    if (i == 16) return S<T>::c(s.custom);
    throw std::runtime_error("shape attribute index out of range: " + std::to_string(i));
}

// ---------------------------------------------------------------------------
// Derivation context
// ---------------------------------------------------------------------------
struct Dec {
    std::vector<int64_t> v;
    size_t cur = 0;
    bool rec = true;
    int64_t pick(int64_t x) {
        if (rec) { v.push_back(x); return x; }
        if (cur >= v.size()) throw std::runtime_error("decision replay underflow");
        return v[cur++];
    }
};

struct Ctx {
    const pga_procedure* procs;
    uint32_t nprocs;
    const pga_dispatch_entry* entries;
    uint32_t nentries;
    const float* params_f;
    std::vector<double> params_d;
    std::vector<Var> params_v;
    uint32_t params_per_axiom;
    uint32_t naxioms;
    uint32_t key[2];
    std::vector<struct Collider>* colliders;   // stored by Collider, read by IfCollides
    uint32_t phase;                            // the phase being derived
};

template <class T> T axiom_param(const Ctx& c, uint32_t axiom, uint32_t slot);
template <> float axiom_param<float>(const Ctx& c, uint32_t a, uint32_t s) {
    return c.params_f[(size_t)a * c.params_per_axiom + s];
}
template <> double axiom_param<double>(const Ctx& c, uint32_t a, uint32_t s) {
    return c.params_d[(size_t)a * c.params_per_axiom + s];
}
template <> Var axiom_param<Var>(const Ctx& c, uint32_t a, uint32_t s) {
    return c.params_v[(size_t)a * c.params_per_axiom + s];
}

float philox_uniform(const Ctx& c, const Meta& m, uint32_t stream) {
    uint32_t ctr[4] = {m.axiom, m.path, m.depth, stream};
    uint32_t out[4];
    philox(ctr, c.key, out);
    return (float)(out[0] >> 8) * 5.9604644775390625e-8f; // 2^-24
}

// EncExp::operate (Parameters.cuh:521-560)
template <class T> T exp_operate(int op, T p1, T p2) {
    const T one = S<T>::c(1.0), zero = S<T>::c(0.0);
    switch (op) {
    case PGA_OPT_ADD: return p1 + p2;
    case PGA_OPT_SUB: return p1 - p2;
    case PGA_OPT_DIV: return p1 / p2;
    case PGA_OPT_MULTI: return p1 * p2;
    case PGA_OPT_EQ: return (p1 == p2) ? one : zero;
    case PGA_OPT_NEQ: return (p1 != p2) ? one : zero;
    case PGA_OPT_LT: return (p1 < p2) ? one : zero;
    case PGA_OPT_GT: return (p1 > p2) ? one : zero;
    case PGA_OPT_LEQ: return (p1 <= p2) ? one : zero;
    case PGA_OPT_GEQ: return (p1 >= p2) ? one : zero;
    case PGA_OPT_AND: return (p1 != zero && p2 != zero) ? one : zero;
    case PGA_OPT_OR: return (p1 != zero || p2 != zero) ? one : zero;
    default: throw std::runtime_error("unknown expression operation " + std::to_string(op));
    }
}

// EncExp::eval (Parameters.cuh:563-664): postfix stack machine, 4-deep stack.
template <class T> T exp_eval(const Shape<T>& s, const float* b) {
    T stack[PGA_MAX_EXP_STACK];
    int top = 0;
    uint32_t size = (uint32_t)b[0];
    if (size > PGA_MAX_NUM_PARAMETER_VALUES - 1) throw std::runtime_error("expression too long");
    for (uint32_t i = 1; i <= size; ++i) {
        int type; float x;
        unpack_operand(b[i], &type, &x);
        switch (type) {
        case PGA_ORT_SCALAR:
            if (top >= PGA_MAX_EXP_STACK) throw std::runtime_error("expression stack overflow");
            stack[top++] = S<T>::c(x);
            break;
        case PGA_ORT_RAND: {
            float mn, mx;
            unpack_float2(x, &mn, &mx);
            if (top >= PGA_MAX_EXP_STACK) throw std::runtime_error("expression stack overflow");
            stack[top++] = range_t<T>(s.seed, S<T>::c(mn), S<T>::c(mx));
            break;
        }
        case PGA_ORT_SHAPE_ATTR: {
            uint32_t idx = (uint32_t)std::lround(x);   // round(x), Parameters.cuh:611
            if (top >= PGA_MAX_EXP_STACK) throw std::runtime_error("expression stack overflow");
            stack[top++] = shape_at(s, idx);
            break;
        }
        default: { // ORT_OP
            int op = (int)std::lround(x);
            if (top < 2) throw std::runtime_error("expression stack underflow");
            T p2 = stack[--top];
            T p1 = stack[--top];
            stack[top++] = exp_operate(op, p1, p2);
            break;
        }
        }
    }
    if (top < 1) throw std::runtime_error("empty expression");
    return stack[--top];
}

// dynEval (Parameters.cuh:669-711) + project parameter types
template <class T> T dyn_eval(const Ctx& c, const Meta& m, const Shape<T>& s,
                              const pga_dispatch_parameter& p) {
    switch (p.type) {
    case PGA_PT_SCALAR: return S<T>::c(p.value[0]);
    case PGA_PT_RAND: return range_t<T>(s.seed, S<T>::c(p.value[0]), S<T>::c(p.value[1]));
    case PGA_PT_SHAPE_ATTR: return shape_at(s, (uint32_t)p.value[0]);
    case PGA_PT_EXP: return exp_eval(s, p.value);
    case PGA_PT_PHILOX: {
        float u = philox_uniform(c, m, (uint32_t)p.value[2]);
        const T mn = S<T>::c(p.value[0]), mx = S<T>::c(p.value[1]);
        return (mx - mn) * S<T>::c(u) + mn;
    }
    case PGA_PT_AXIOM_PARAM: {
        uint32_t slot = (uint32_t)p.value[0];
        if (slot >= c.params_per_axiom || m.axiom >= c.naxioms)
            throw std::runtime_error("axiom parameter out of range");
        return axiom_param<T>(c, m.axiom, slot);
    }
    default: throw std::runtime_error("unsupported parameter type " + std::to_string(p.type));
    }
}

// Output of one operator execution: children (dispatched to a successor
// entry) and terminal instances.
template <class T> struct Out {
    int kind;          // 0 child, 1 terminal
    uint32_t phase;    // child: the phase of the successor it goes to
    int32_t entry;     // successor entry (child) / generate entry (terminal)
    uint32_t ordinal;  // child ordinal within the operator
    uint32_t terminal; // terminal index
    int32_t shape_type;
    Shape<T> s;
    bool has_poly;     // child: `poly` replaces the parent's vertex list
    Poly poly;
};

struct OpErr : std::runtime_error {
    int code;
    OpErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline uint32_t axis_of(float v) {          // __float2uint_rd / static_cast<unsigned>
    uint32_t a = f2u_rd(v);
    if (a > 2) throw OpErr(PGA_ERR_BAD_PROGRAM, "axis parameter out of range");
    return a;
}

inline int32_t succ(const pga_dispatch_entry& e, uint32_t k) {
    if (k >= e.num_successors) throw OpErr(PGA_ERR_BAD_PROGRAM, "successor index out of range");
    return e.successors[k].entry_index;
}

// The child of successor k of entry e: its entry and the phase the successor
// names (DispatchTableEntry.h:8-36; Call.cuh:69-70 dispatches to that phase).
// PGA_PHASE_INHERIT keeps the parent's phase (resolved by the caller); any
// other negative phase index (an unassigned one) is phase 0.
const uint32_t kPhaseInherit = 0xFFFFFFFFu;
template <class T>
void push_child(std::vector<Out<T>>& out, const pga_dispatch_entry& e, uint32_t k, uint32_t ordinal, int32_t st,
                const Shape<T>& s) {
    Out<T> o;
    o.kind = 0; o.entry = succ(e, k); o.ordinal = ordinal; o.terminal = 0; o.shape_type = st; o.s = s;
    o.has_poly = false;
    const int32_t ph = e.successors[k].phase_index;
    o.phase = ph == PGA_PHASE_INHERIT ? kPhaseInherit : (ph > 0 ? (uint32_t)ph : 0u);
    out.push_back(o);
}
inline float fparam(const pga_dispatch_entry& e, uint32_t k) {
    if (k >= e.num_parameters) throw OpErr(PGA_ERR_BAD_PROGRAM, "parameter index out of range");
    return 0.0f;
}

// Evaluates parameter k in every twin; discrete uses read the float value.
template <class T> T P(const Ctx& c, const Meta& m, const Shape<T>& s, const pga_dispatch_entry& e, uint32_t k) {
    fparam(e, k);
    return dyn_eval<T>(c, m, s, e.parameters[k]);
}

// ------------------------- operators --------------------------------------

// Subdivide<DynParams> (Subdivide.cuh:170-248) and the static form
// Subdivide<Axis, Pair<Factor, Next>...> (Subdivide.cuh:22-168).
template <class T>
void op_subdivide(const Ctx& c, const Meta& m, const Shape<T>& parent, const pga_dispatch_entry& e,
                  bool is_static, Dec& d, std::vector<Out<T>>& out) {
    const T zero = S<T>::c(0.0), one = S<T>::c(1.0), m1 = S<T>::c(-1.0), half = S<T>::c(0.5);
    uint32_t axis = (uint32_t)d.pick(axis_of((float)S<T>::val(P(c, m, parent, e, 0))));
    uint32_t n = e.num_parameters;
    T total = parent.size[axis];
    T offset = total * S<T>::c(-0.5);
    T abs_sum = zero, rel_sum = zero, r2a = zero;
    for (uint32_t i = 1; i < n; ++i) {
        T f = P(c, m, parent, e, i);
        abs_sum += (f > zero ? one : zero) * f;
        rel_sum += (f < zero ? one : zero) * (f * m1);
    }
    if (rel_sum == zero) r2a = zero;
    else r2a = (total - abs_sum) / rel_sum;
    if (!is_static) {
        int neg = (int)d.pick(S<T>::val(r2a) < 0.0 ? 1 : 0);
        if (neg) throw OpErr(PGA_ERR_SUBDIV_OVERCONSTRAINED,
                             "Subdivide: sum of absolute factors exceeds the shape size");
    }
    Shape<T> sym = parent;                 // the DynParams form mutates the symbol in place
    T base_m[12];
    for (int i = 0; i < 12; ++i) base_m[i] = parent.m[i];
    for (uint32_t i = 1, j = 0; i < n; j = i, i++) {
        T f = is_static ? P(c, m, parent, e, i) : P(c, m, sym, e, i);
        if (is_static) {
            int keep = (int)d.pick(S<T>::val(f) != 0.0 ? 1 : 0);
            if (!keep) continue;
        }
        T new_size = (((f < zero ? one : zero) * r2a) * f) * m1 + ((f > zero ? one : zero) * f);
        T off3[3] = {zero, zero, zero};
        off3[axis] = offset + (new_size * half);
        Shape<T> ch = sym;
        ch.size[0] = parent.size[0]; ch.size[1] = parent.size[1]; ch.size[2] = parent.size[2];
        ch.size[axis] = new_size;
        apply_translate(ch, base_m, off3[0], off3[1], off3[2]);
        ch.seed = next_seed(parent.seed, j);
        push_child(out, e, j, j, m.shape_type, ch);
        if (!is_static) sym = ch;
        offset = offset + new_size;
    }
}

// Repeat<false, Axis, Extent, Mode, Next> (Repeat.cuh:63-203)
template <class T>
void op_repeat(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
               Dec& d, std::vector<Out<T>>& out) {
    const T zero = S<T>::c(0.0), one = S<T>::c(1.0), half = S<T>::c(0.5);
    uint32_t axis = (uint32_t)d.pick(axis_of((float)S<T>::val(P(c, m, sh, e, 0))));
    int mode = (int)d.pick((int64_t)f2u_rd((float)S<T>::val(P(c, m, sh, e, 2))));
    T ex = P(c, m, sh, e, 1);
    // getNewExtent (Repeat.cuh:26-31)
    T extent = ((ex < zero ? one : zero) * sh.size[axis] * S<T>::fabs_(ex)) + ((ex > zero ? one : zero) * ex);
    if (d.pick(S<T>::val(extent) <= 0.0 ? 1 : 0))
        throw OpErr(PGA_ERR_REPEAT_BAD_EXTENT, "Repeat: extent parameter cannot be less than or 0");
    T total = sh.size[axis];
    if (d.pick(S<T>::val(total) < 0.0 ? 1 : 0))
        throw OpErr(PGA_ERR_REPEAT_BAD_EXTENT, "Repeat: shape size being subdivided is negative");
    T division = total / extent;
    uint32_t reps;
    T e1, e2, e3;
    if (mode == PGA_ANCHOR_TO_START || mode == PGA_ANCHOR_TO_END) {
        T rest = S<T>::fmod_(total, extent);
        int exact = (int)d.pick(S<T>::val(rest) == 0.0 ? 1 : 0);
        double cdiv = std::ceil(S<T>::val(division));
        reps = (uint32_t)d.pick(cdiv >= 4294967295.0 ? (int64_t)UINT_MAX : (int64_t)cdiv);
        e1 = S<T>::c(exact ? 1.0 : 0.0) * extent + S<T>::c(exact ? 0.0 : 1.0) * rest;
        e2 = e3 = extent;
    } else if (mode == PGA_ADJUST_TO_FILL) {
        double fdiv = std::floor(S<T>::val(division));
        reps = (uint32_t)d.pick(fdiv >= 4294967295.0 ? (int64_t)UINT_MAX : (int64_t)fdiv);
        if (reps > 0) e1 = e2 = e3 = total / S<T>::c((double)(float)reps);
        else e1 = e2 = e3 = zero;
    } else {
        throw OpErr(PGA_ERR_BAD_PROGRAM, "Repeat: unknown repeat mode");
    }
    if (reps > (1u << 24)) throw OpErr(PGA_ERR_OVERFLOW_QUEUE, "Repeat: too many repetitions");
    bool start = mode == PGA_ANCHOR_TO_START;
    T start_ext = S<T>::c(start ? 1.0 : 0.0) * e1 + S<T>::c(start ? 0.0 : 1.0) * e3;
    (void)succ(e, 0);
    Shape<T> ns = sh;
    T base_m[12];
    for (int i = 0; i < 12; ++i) base_m[i] = sh.m[i];
    T off[3] = {zero, zero, zero};
    T half_start = start_ext * half;
    off[axis] = (total * S<T>::c(-0.5)) + half_start;
    ns.size[axis] = start_ext;
    apply_translate(ns, base_m, off[0], off[1], off[2]);
    ns.seed = next_seed(sh.seed, 1u);
    push_child(out, e, 0, 0, m.shape_type, ns);
    off[axis] = off[axis] + half_start;
    T half_mid = e2 * half;
    ns.size[axis] = e2;
    for (uint32_t i = 2; i < reps; ++i) {
        off[axis] = off[axis] + half_mid;
        apply_translate(ns, base_m, off[0], off[1], off[2]);
        ns.seed = next_seed(sh.seed, i);
        push_child(out, e, 0, i - 1, m.shape_type, ns);
        off[axis] = off[axis] + half_mid;
    }
    if (reps > 1) {
        T end_ext = S<T>::c(start ? 1.0 : 0.0) * e3 + S<T>::c(start ? 0.0 : 1.0) * e1;
        off[axis] = off[axis] + end_ext * half;
        ns.size[axis] = end_ext;
        apply_translate(ns, base_m, off[0], off[1], off[2]);
        ns.seed = next_seed(sh.seed, reps);
        push_child(out, e, 0, reps - 1, m.shape_type, ns);
    }
}

// ComponentSplit<false, Top, Bottom, Sides> over a Box with FaceSplitter<Box>
// (ComponentSplit.cuh:77-174, 234-267): sides back/front/left/right, then
// bottom, then top; faces are Quads.
const float kBack[16] = {-1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1};
const float kLeft[16] = {0, 0, -1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1};
const float kRight[16] = {0, 0, 1, 0, 0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 0, 1};
const float kBottom[16] = {1, 0, 0, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 1};
const float kTop[16] = {1, 0, 0, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 1};
const float kExtrude[16] = {1, 0, 0, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 1};

template <class T>
void op_compsplit(const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e, std::vector<Out<T>>& out) {
    if (m.shape_type != PGA_SHAPE_BOX) throw OpErr(PGA_ERR_UNSUPPORTED, "ComponentSplit supports Box shapes only");
    const T zero = S<T>::c(0.0), one = S<T>::c(1.0), half = S<T>::c(0.5);
    T he[3] = {sh.size[0] * half, sh.size[1] * half, sh.size[2] * half};
    (void)succ(e, 0); (void)succ(e, 1); (void)succ(e, 2);
    for (uint32_t f = 0; f < 4; ++f) {
        Shape<T> q;
        T t[16], tmp[12], r[16];
        if (f == 0) { mat_translate(zero, zero, -he[2], t); mat_const(kBack, r); }
        else if (f == 1) { mat_translate(zero, zero, he[2], t); }
        else if (f == 2) { mat_translate(-he[0], zero, zero, t); mat_const(kLeft, r); }
        else { mat_translate(he[0], zero, zero, t); mat_const(kRight, r); }
        if (f == 1) {
            mul34x44(sh.m, t, q.m);
        } else {
            mul34x44(sh.m, t, tmp);
            mul34x44(tmp, r, q.m);
        }
        if (f < 2) { q.size[0] = sh.size[0]; q.size[1] = sh.size[1]; }
        else { q.size[0] = sh.size[2]; q.size[1] = sh.size[1]; }
        q.size[2] = one;
        q.seed = next_seed(sh.seed, f + 2u);
        q.custom = sh.custom;
        push_child(out, e, 2, f, PGA_SHAPE_QUAD, q);
    }
    for (int cap = 0; cap < 2; ++cap) {     // bottom, then top
        bool bottom_face = cap == 0;
        Shape<T> q;
        T t[16], tmp[12], r[16];
        if (bottom_face) { mat_translate(zero, -he[1], zero, t); mat_const(kBottom, r); }
        else { mat_translate(zero, he[1], zero, t); mat_const(kTop, r); }
        mul34x44(sh.m, t, tmp);
        mul34x44(tmp, r, q.m);
        q.size[0] = sh.size[0]; q.size[1] = sh.size[2]; q.size[2] = one;
        q.seed = next_seed(sh.seed, bottom_face ? 1u : 0u);
        q.custom = sh.custom;
        push_child(out, e, bottom_face ? 1u : 0u, 4u + (uint32_t)cap, PGA_SHAPE_QUAD, q);
    }
}

// ComponentSplit<true, Top, Bottom, Sides> over a Box with
// ParallelFaceSplitter<Box>::splitFace (ComponentSplit.cuh:179-230, 269-328):
// face f = 0..5 (back, front, left, right, bottom, top) framed by its normal
// n, up vector y and x = cross(y, n) (vector.h:510-515), translated by
// size * n / 2; sizes are dot(size, |s0|), dot(size, |s1|) (vector.h:505-508);
// seed modifier (f + 2) % 6; faces 4 and 5 go to Bottom and Top.
template <class T>
void op_compsplit_parallel(const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e, std::vector<Out<T>>& out) {
    if (m.shape_type != PGA_SHAPE_BOX) throw OpErr(PGA_ERR_UNSUPPORTED, "ComponentSplit supports Box shapes only");
    (void)succ(e, 0); (void)succ(e, 1); (void)succ(e, 2);
    auto k = [](float v) { return S<T>::c(v); };
    for (uint32_t f = 0; f < 6; ++f) {
        const float sign = ((float)(f & 0x1u) * 2.0f) - 1.0f;
        const float a = (float)(f < 2u);
        const float b1 = (float)(f & 0x2u) * 0.5f;
        const float b2 = (float)(f & 0x4u) * 0.25f;
        const float n[3] = {b1 * sign, b2 * sign, a * sign};
        const float y[3] = {0.0f, 1.0f - b2, b2 * -sign};
        const float x[3] = {y[1] * n[2] - y[2] * n[1], y[2] * n[0] - y[0] * n[2], y[0] * n[1] - y[1] * n[0]};
        const float s0[3] = {std::fabs(a + b2), 0.0f, std::fabs(b1)};
        const float s1[3] = {0.0f, std::fabs(1.0f - b2), std::fabs(b2)};
        Shape<T> q;
        T t[16], tmp[12], r[16];
        mat_translate((sh.size[0] * k(n[0])) * k(0.5f), (sh.size[1] * k(n[1])) * k(0.5f),
                      (sh.size[2] * k(n[2])) * k(0.5f), t);
        const float rot[16] = {x[0], y[0], n[0], 0.0f, x[1], y[1], n[1], 0.0f,
                               x[2], y[2], n[2], 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
        mat_const(rot, r);
        mul34x44(sh.m, t, tmp);
        mul34x44(tmp, r, q.m);
        q.size[0] = sh.size[0] * k(s0[0]) + sh.size[1] * k(s0[1]) + sh.size[2] * k(s0[2]);
        q.size[1] = sh.size[0] * k(s1[0]) + sh.size[1] * k(s1[1]) + sh.size[2] * k(s1[2]);
        q.size[2] = S<T>::c(1.0);
        q.seed = next_seed(sh.seed, (f + 2u) % 6u);
        q.custom = sh.custom;
        push_child(out, e, f == 5u ? 0u : (f == 4u ? 1u : 2u), f, PGA_SHAPE_QUAD, q);
    }
}

// IfSizeLess (If.cuh:61-98) / If (If.cuh:23-58)
template <class T>
void op_ifsizeless(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
                   Dec& d, std::vector<Out<T>>& out) {
    uint32_t axis = (uint32_t)d.pick(axis_of((float)S<T>::val(P(c, m, sh, e, 0))));
    T thr = P(c, m, sh, e, 1);
    int taken = (int)d.pick(sh.size[axis] < thr ? 0 : 1);
    push_child(out, e, (uint32_t)taken, (uint32_t)taken, m.shape_type, sh);
}
template <class T>
void op_if(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
           Dec& d, std::vector<Out<T>>& out) {
    float r = (float)S<T>::val(P(c, m, sh, e, 0));
    int taken = (int)d.pick((r != 0.0f) ? 0 : 1);   // `if (result)`: NaN counts as true
    push_child(out, e, (uint32_t)taken, (uint32_t)taken, m.shape_type, sh);
}

// RandomRule<DynParams> (RandomRule.cuh:121-147) and the static form
// RandomRule<Chance, Pair<P, Next>...> (RandomRule.cuh:20-111). The choice is
// made in FP32 and replayed by the other twins.
template <class T>
void op_random(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
               bool is_static, Dec& d, std::vector<Out<T>>& out) {
    Shape<float> dummy;
    (void)dummy;
    int64_t choice = -1;
    if (d.rec) {
        // decisions are taken in the precision of the recording (f32) pass
        float v0 = (float)S<T>::val(P(c, m, sh, e, 0));
        float chance = v0 - std::floor(v0);      // math::frac (math.h:243-246)
        uint32_t n = e.num_parameters;
        if (!is_static) {
            float sum = 0.0f;
            for (uint32_t i = 1; i < n; ++i) sum += (float)S<T>::val(P(c, m, sh, e, i));
            float acc = 0.0f;
            for (uint32_t i = 1; i < n; ++i) {
                acc += (float)S<T>::val(P(c, m, sh, e, i)) / sum;
                if (chance <= acc) { choice = (int64_t)i - 1; break; }
            }
        } else {
            // getProbabilitiesSum is a right fold: p1 + (p2 + (... + pn))
            float sum = 0.0f;
            for (uint32_t i = n - 1; i >= 1; --i) {
                float p = (float)S<T>::val(P(c, m, sh, e, i));
                sum = (i == n - 1) ? p : p + sum;
            }
            float acc = 0.0f;
            for (uint32_t i = 1; i < n; ++i) {
                if (i == n - 1) { choice = (int64_t)i - 1; break; }
                float np = (float)S<T>::val(P(c, m, sh, e, i)) / sum;
                acc += np;
                if (np != 0.0f && chance <= acc) { choice = (int64_t)i - 1; break; }
            }
        }
    }
    choice = d.pick(choice);
    if (choice >= 0) push_child(out, e, (uint32_t)choice, (uint32_t)choice, m.shape_type, sh);
}

template <class T>
void op_translate(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
                  std::vector<Out<T>>& out) {           // Translate.cuh:24-50
    T x = P(c, m, sh, e, 0), y = P(c, m, sh, e, 1), z = P(c, m, sh, e, 2);
    Shape<T> ns = sh;
    apply_translate(ns, sh.m, x, y, z);
    push_child(out, e, 0, 0, m.shape_type, ns);
}

template <class T>
void op_rotate(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
               std::vector<Out<T>>& out) {              // Rotate.cuh:24-51, matrix.h:876-888
    T x = P(c, m, sh, e, 0), y = P(c, m, sh, e, 1), z = P(c, m, sh, e, 2);
    const T k = S<T>::deg2rad();
    x = x * k; y = y * k; z = z * k;
    T sa, sb, sg, ca, cb, cg;
    sincos_t(x, &sa, &ca);
    sincos_t(y, &sb, &cb);
    sincos_t(z, &sg, &cg);
    const T zero = S<T>::c(0.0), one = S<T>::c(1.0);
    T r[16] = {cb * cg, cg * sa * sb - ca * sg, ca * cg * sb + sa * sg, zero,
               cb * sg, ca * cg + sa * sb * sg, (-cg) * sa + ca * sb * sg, zero,
               -sb, cb * sa, ca * cb, zero,
               zero, zero, zero, one};
    Shape<T> ns = sh;
    mul34x44(sh.m, r, ns.m);
    push_child(out, e, 0, 0, m.shape_type, ns);
}

template <class T>
void op_scale(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
              std::vector<Out<T>>& out) {               // Scale.cuh:23-50
    const T zero = S<T>::c(0.0), one = S<T>::c(1.0), m1 = S<T>::c(-1.0);
    T v[3] = {P(c, m, sh, e, 0), P(c, m, sh, e, 1), P(c, m, sh, e, 2)};
    Shape<T> ns = sh;
    for (int i = 0; i < 3; ++i)
        ns.size[i] = (((v[i] < zero ? one : zero) * sh.size[i]) * v[i]) * m1 + ((v[i] > zero ? one : zero) * v[i]);
    push_child(out, e, 0, 0, m.shape_type, ns);
}

template <class T>
void op_swapsize(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
                 Dec& d, std::vector<Out<T>>& out) {    // SwapSize.cuh:25-58
    uint32_t ax[3];
    for (uint32_t i = 0; i < 3; ++i) ax[i] = (uint32_t)d.pick(axis_of((float)S<T>::val(P(c, m, sh, e, i))));
    Shape<T> ns = sh;
    ns.size[0] = sh.size[ax[0]]; ns.size[1] = sh.size[ax[1]]; ns.size[2] = sh.size[ax[2]];
    push_child(out, e, 0, 0, m.shape_type, ns);
}

// Extrude.cuh:25-71 over a planar shape: Quad -> Box; a dynamic polygon ->
// the right prism of the same vertex list and `invert` (fromPlanarType,
// DynamicRightPrism.cuh:107-115), which the caller carries over.
template <class T>
void op_extrude(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
                std::vector<Out<T>>& out) {
    int32_t to;
    if (m.shape_type == PGA_SHAPE_QUAD) to = PGA_SHAPE_BOX;
    else if (m.shape_type == PGA_SHAPE_DYNAMIC_CONVEX_POLYGON) to = PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM;
    else if (m.shape_type == PGA_SHAPE_DYNAMIC_POLYGON) to = PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
    else throw OpErr(PGA_ERR_UNSUPPORTED, "Extrude supports planar shapes only");
    T extent = P(c, m, sh, e, 1);
    Shape<T> ns;
    T r[16];
    mat_const(kExtrude, r);
    mul34x44(sh.m, r, ns.m);
    ns.size[0] = sh.size[0]; ns.size[1] = extent; ns.size[2] = sh.size[1];
    ns.seed = next_seed(sh.seed, 0u);
    ns.custom = sh.custom;
    push_child(out, e, 0, 0, to, ns);
}

// SetAs<DynParams, NewShape> (SetAs.cuh:112-150): the shape becomes a dynamic
// polygon or right prism with numSides = the entry's parameter count and
// vertex i = dynToFloat2(parameter i) (Parameters.cuh:714-742: a scalar is
// (value, 0); Vec2, Vec4 and Rand give their first two values, unevaluated);
// model, size, seed and custom attribute are kept, `invert` is cleared.
// Fewer than 3 or more than PGA_MAX_NUM_SIDES vertices are refused (the
// reference writes past its vertex array above MaxNumSides).
template <class T>
void op_setas(const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e, int32_t op,
              std::vector<Out<T>>& out) {
    (void)m;
    const uint32_t n = e.num_parameters;
    if (n < 3 || n > PGA_MAX_NUM_SIDES) throw OpErr(PGA_ERR_BAD_PROGRAM, "SetAs needs 3 to 16 vertices");
    Poly p;
    p.n = n;
    for (uint32_t i = 0; i < n; ++i) {
        const pga_dispatch_parameter& q = e.parameters[i];
        if (q.type == PGA_PT_SCALAR) { p.v[2 * i] = q.value[0]; p.v[2 * i + 1] = 0.0f; }
        else if (q.type == PGA_PT_VEC2 || q.type == PGA_PT_VEC4 || q.type == PGA_PT_RAND) {
            p.v[2 * i] = q.value[0]; p.v[2 * i + 1] = q.value[1];
        } else {
            throw OpErr(PGA_ERR_BAD_PROGRAM, "SetAs vertex parameters are scalars, vectors or Rand pairs");
        }
    }
    const int32_t to = op == PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON ? PGA_SHAPE_DYNAMIC_CONVEX_POLYGON
                     : op == PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM ? PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM
                     : op == PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON ? PGA_SHAPE_DYNAMIC_POLYGON
                     : PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
    push_child(out, e, 0, 0, to, sh);
    out.back().has_poly = true;
    out.back().poly = p;
}

// ComponentSplit<false, Top, Bottom, Sides> over a dynamic right prism with
// the generic FaceSplitter (ComponentSplit.cuh:23-75, 234-267): side f =
// 0..n-1 runs from vertex f-1 (n-1 for f = 0) to vertex f, both scaled by
// (size.x, size.z) with a 2x2 product and y negated; the side's frame turns
// x onto that edge (facing), its origin is the edge's midpoint, its size
// (|edge|, size.y, 1) and its seed modifier f + 2. Then the bottom cap
// (inverted, modifier 1, Bottom) and the top cap (modifier 0, Top): the
// model translated by +-size.y / 2 and turned by the cap frame, size
// (size.x, size.z, 1), the prism's raw vertex list.
template <class T>
void op_compsplit_prism(const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e, std::vector<Out<T>>& out) {
    (void)succ(e, 0); (void)succ(e, 1); (void)succ(e, 2);
    const T zero = S<T>::c(0.0), one = S<T>::c(1.0), half = S<T>::c(0.5);
    const uint32_t n = m.poly.n;
    const T sx = sh.size[0], sz = sh.size[2];
    for (uint32_t f = 0; f < n; ++f) {
        float ax, ay, bx, by;
        poly_vertex(m.poly, f == 0 ? n - 1 : f - 1, ax, ay);
        poly_vertex(m.poly, f, bx, by);
        const T a0 = S<T>::c(ax), a1 = S<T>::c(ay), b0 = S<T>::c(bx), b1 = S<T>::c(by);
        // float2x2::scale(size.xz()) * vertex (matrix.h:389-393)
        T p0x = sx * a0 + zero * a1, p0y = zero * a0 + sz * a1;
        T p1x = sx * b0 + zero * b1, p1y = zero * b0 + sz * b1;
        p0y = -p0y;
        p1y = -p1y;
        T dx = p1x - p0x, dy = p1y - p0y;
        const T x = S<T>::sqrt_(dx * dx + dy * dy);
        const T ox = p0x + (p1x - p0x) * half, oy = p0y + (p1y - p0y) * half;
        const T r = one / x;
        dx = dx * r;
        dy = dy * r;
        const T facing[16] = {dx, zero, -dy, zero, zero, one, zero, zero, dy, zero, dx, zero, zero, zero, zero, one};
        T t[16], tmp[12];
        mat_translate(ox, zero, oy, t);
        Shape<T> q;
        mul34x44(sh.m, t, tmp);
        mul34x44(tmp, facing, q.m);
        q.size[0] = x; q.size[1] = sh.size[1]; q.size[2] = one;
        q.seed = next_seed(sh.seed, f + 2u);
        q.custom = sh.custom;
        push_child(out, e, 2, f, PGA_SHAPE_QUAD, q);
    }
    const int32_t cap_shape = m.shape_type == PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM ? PGA_SHAPE_DYNAMIC_CONVEX_POLYGON
                                                                                  : PGA_SHAPE_DYNAMIC_POLYGON;
    for (uint32_t cap = 0; cap < 2; ++cap) {   // bottom, then top
        const bool bottom = cap == 0;
        const T sign = S<T>::c(bottom ? -1.0 : 1.0);   // 1 - 2 * isBottom
        const T frame[16] = {one, zero, zero, zero, zero, zero, sign, zero, zero, -sign, zero, zero, zero, zero, zero, one};
        T t[16], tmp[12];
        mat_translate(zero, (sh.size[1] * sign) * half, zero, t);
        Shape<T> q;
        mul34x44(sh.m, t, tmp);
        mul34x44(tmp, frame, q.m);
        q.size[0] = sh.size[0]; q.size[1] = sh.size[2]; q.size[2] = one;
        q.seed = next_seed(sh.seed, bottom ? 1u : 0u);
        q.custom = sh.custom;
        push_child(out, e, bottom ? 1u : 0u, n + cap, cap_shape, q);
        out.back().has_poly = true;
        out.back().poly = m.poly;
        out.back().poly.invert = bottom ? 1u : 0u;
    }
}

template <class T>
void op_replicate(const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
                  std::vector<Out<T>>& out) {           // Replicate.cuh:102-146
    for (uint32_t i = 0; i < e.num_successors; ++i) push_child(out, e, i, i, m.shape_type, sh);
}

// Generate<false, 1, TerminalIdx> -> GenerationFunction<..., 1>::perShape
// (Generate.cuh:24-40, GenerationFunctions.cuh:155-189)
template <class T>
void op_generate(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e,
                 Dec& d, std::vector<Out<T>>& out) {
    uint32_t term = (uint32_t)d.pick((int64_t)f2u_rd((float)S<T>::val(P(c, m, sh, e, 0))));
    if (term >= PGA_MAX_TERMINALS) throw OpErr(PGA_ERR_BAD_PROGRAM, "terminal index out of range");
    Out<T> o;
    o.kind = 1; o.entry = m.entry; o.ordinal = 0; o.terminal = term; o.shape_type = m.shape_type;
    o.has_poly = false;
    T sc[16];
    mat_scale(sh.size[0], sh.size[1], sh.size[2], sc);
    o.s = sh;
    if (has_vertices(m.shape_type)) {
        // getAdjustedModel (DynamicPolygon.cuh:68-75): the first two columns
        // times sign = -1 when inverted, 1 otherwise
        T am[12];
        const T sign = S<T>::c(m.poly.invert ? -1.0 : 1.0);
        for (int i = 0; i < 3; ++i) {
            am[i * 4 + 0] = sign * sh.m[i * 4 + 0];
            am[i * 4 + 1] = sign * sh.m[i * 4 + 1];
            am[i * 4 + 2] = sh.m[i * 4 + 2];
            am[i * 4 + 3] = sh.m[i * 4 + 3];
        }
        mul34x44(am, sc, o.s.m);
        out.push_back(o);
        return;
    }
    mul34x44(sh.m, sc, o.s.m);
    out.push_back(o);
}

// ---------------------------------------------------------------------------
// Context sensitivity (IfCollides.cuh, Collider.cuh, Collision.cuh,
// IntermediateSymbolsBufferAdapter.cuh)
// ---------------------------------------------------------------------------
struct Collider {
    float m[12];
    float size[3];
    int32_t shape_type;
    uint32_t tag;
    uint32_t phase;            // the phase that stored it
};

inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Collision::check of two shapes A (the stored collider) and B (the shape
// asking), the separating-axis test of two oriented boxes (Gottschalk 1996;
// Ericson, Real-Time Collision Detection 4.4.1) as Box x Box
// (Collision.cuh:35-262) evaluates it: the 3 face axes of A, the 3 of B, the
// 9 edge cross products. A Quad is a box whose z half extent drops out of
// every projection onto the other shape's axes and keeps its own face axis
// (Box x Quad :263-490, Quad x Quad :506-733); Quad x Box is Box x Quad with
// the shapes swapped (:734-737). af / bf: half extents for the shape's own
// face axes; ap / bp: for its projections.
bool sat_check(const float A[12], const float af[3], const float ap[3], const float B[12], const float bf[3],
               const float bp[3]) {
    const float ac[3][3] = {{A[0], A[4], A[8]}, {A[1], A[5], A[9]}, {A[2], A[6], A[10]}};   // column1..3
    const float bc[3][3] = {{B[0], B[4], B[8]}, {B[1], B[5], B[9]}, {B[2], B[6], B[10]}};
    const float v[3] = {B[3] - A[3], B[7] - A[7], B[11] - A[11]};                          // position (column4)
    const float T[3] = {dot3(v, ac[0]), dot3(v, ac[1]), dot3(v, ac[2])};
    float R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) R[i][k] = dot3(ac[i], bc[k]);
    float ra, rb, t;
    for (int i = 0; i < 3; ++i) {
        ra = af[i];
        rb = bp[0] * std::fabs(R[i][0]) + bp[1] * std::fabs(R[i][1]) + bp[2] * std::fabs(R[i][2]);
        t = std::fabs(T[i]);
        if (t > ra + rb) return false;
    }
    for (int k = 0; k < 3; ++k) {
        ra = ap[0] * std::fabs(R[0][k]) + ap[1] * std::fabs(R[1][k]) + ap[2] * std::fabs(R[2][k]);
        rb = bf[k];
        t = std::fabs(T[0] * R[0][k] + T[1] * R[1][k] + T[2] * R[2][k]);
        if (t > ra + rb) return false;
    }
    // L = A0 x B0 .. A2 x B2
    ra = ap[1] * std::fabs(R[2][0]) + ap[2] * std::fabs(R[1][0]);
    rb = bp[1] * std::fabs(R[0][2]) + bp[2] * std::fabs(R[0][1]);
    t = std::fabs(T[2] * R[1][0] - T[1] * R[2][0]);
    if (t > ra + rb) return false;
    ra = ap[1] * std::fabs(R[2][1]) + ap[2] * std::fabs(R[1][1]);
    rb = bp[0] * std::fabs(R[0][2]) + bp[2] * std::fabs(R[0][0]);
    t = std::fabs(T[2] * R[1][1] - T[1] * R[2][1]);
    if (t > ra + rb) return false;
    ra = ap[1] * std::fabs(R[2][2]) + ap[2] * std::fabs(R[1][2]);
    rb = bp[0] * std::fabs(R[0][1]) + bp[1] * std::fabs(R[0][0]);
    t = std::fabs(T[2] * R[1][2] - T[1] * R[2][2]);
    if (t > ra + rb) return false;
    ra = ap[0] * std::fabs(R[2][0]) + ap[2] * std::fabs(R[0][0]);
    rb = bp[1] * std::fabs(R[1][2]) + bp[2] * std::fabs(R[1][1]);
    t = std::fabs(T[0] * R[2][0] - T[2] * R[0][0]);
    if (t > ra + rb) return false;
    ra = ap[0] * std::fabs(R[2][1]) + ap[2] * std::fabs(R[0][1]);
    rb = bp[0] * std::fabs(R[1][2]) + bp[2] * std::fabs(R[1][0]);
    t = std::fabs(T[0] * R[2][1] - T[2] * R[0][1]);
    if (t > ra + rb) return false;
    ra = ap[0] * std::fabs(R[2][2]) + ap[2] * std::fabs(R[0][2]);
    rb = bp[0] * std::fabs(R[1][1]) + bp[1] * std::fabs(R[1][0]);
    t = std::fabs(T[0] * R[2][2] - T[2] * R[0][2]);
    if (t > ra + rb) return false;
    ra = ap[0] * std::fabs(R[1][0]) + ap[1] * std::fabs(R[0][0]);
    rb = bp[1] * std::fabs(R[2][2]) + bp[2] * std::fabs(R[2][1]);
    t = std::fabs(T[1] * R[0][0] - T[0] * R[1][0]);
    if (t > ra + rb) return false;
    ra = ap[0] * std::fabs(R[1][1]) + ap[1] * std::fabs(R[0][1]);
    rb = bp[0] * std::fabs(R[2][2]) + bp[2] * std::fabs(R[2][0]);
    t = std::fabs(T[1] * R[0][1] - T[0] * R[1][1]);
    if (t > ra + rb) return false;
    ra = ap[0] * std::fabs(R[1][2]) + ap[1] * std::fabs(R[0][2]);
    rb = bp[0] * std::fabs(R[2][1]) + bp[1] * std::fabs(R[2][0]);
    t = std::fabs(T[1] * R[0][2] - T[0] * R[1][2]);
    if (t > ra + rb) return false;
    return true;
}

// Collision::check(stored, asking) for Box and Quad shapes, half extents
// size * 0.5f (Shape.cuh:70-73).
bool shapes_collide(const float am[12], const float asz[3], int32_t at, const float bm[12], const float bsz[3],
                    int32_t bt) {
    const float ah[3] = {asz[0] * 0.5f, asz[1] * 0.5f, asz[2] * 0.5f};
    const float bh[3] = {bsz[0] * 0.5f, bsz[1] * 0.5f, bsz[2] * 0.5f};
    const float aq[3] = {ah[0], ah[1], 0.0f}, bq[3] = {bh[0], bh[1], 0.0f};
    const bool aquad = at == PGA_SHAPE_QUAD, bquad = bt == PGA_SHAPE_QUAD;
    if (aquad && !bquad) return sat_check(bm, bh, bh, am, ah, aq);      // Quad x Box = Box x Quad swapped
    return sat_check(am, ah, aquad ? aq : ah, bm, bh, bquad ? bq : bh);
}

// IntermediateSymbolsBufferAdapter::checkCollision (:89-136) over every
// collider a phase before this one stored: the tag matches and the stored
// shape collides with the asking one. The reference prunes the candidates
// with a BVH built between phases (BVHTraversal.cuh:21-148); the answer is
// the same existence.
template <class T> bool collides(const Ctx& c, uint32_t tag, const Meta& m, const Shape<T>& sh) {
    float qm[12], qs[3];
    for (int i = 0; i < 12; ++i) qm[i] = (float)S<T>::val(sh.m[i]);
    for (int i = 0; i < 3; ++i) qs[i] = (float)S<T>::val(sh.size[i]);
    for (const Collider& k : *c.colliders)
        if (k.phase < c.phase && k.tag == tag && shapes_collide(k.m, k.size, k.shape_type, qm, qs, m.shape_type))
            return true;
    return false;
}

// Collider (Collider.cuh:26-42): a positive tag stores the shape; the shape
// goes on to successor 0.
template <class T>
void op_collider(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e, Dec& d,
                 std::vector<Out<T>>& out) {
    const uint32_t tag = (uint32_t)d.pick(f2u_rd((float)S<T>::val(P(c, m, sh, e, 0))));
    if (tag > 0u && d.rec) {
        Collider k;
        for (int i = 0; i < 12; ++i) k.m[i] = (float)S<T>::val(sh.m[i]);
        for (int i = 0; i < 3; ++i) k.size[i] = (float)S<T>::val(sh.size[i]);
        k.shape_type = m.shape_type;
        k.tag = tag;
        k.phase = c.phase;
        c.colliders->push_back(k);
    }
    push_child(out, e, 0, 0, m.shape_type, sh);
}

// IfCollides (IfCollides.cuh:26-58): successor 0 when the shape collides
// with a stored collider of the tag, else successor 1. Decided in FP32 and
// replayed by the twins.
template <class T>
void op_ifcollides(const Ctx& c, const Meta& m, const Shape<T>& sh, const pga_dispatch_entry& e, Dec& d,
                   std::vector<Out<T>>& out) {
    const uint32_t tag = f2u_rd((float)S<T>::val(P(c, m, sh, e, 0)));
    const int taken = (int)d.pick(d.rec ? (collides(c, tag, m, sh) ? 0 : 1) : 0);
    push_child(out, e, (uint32_t)taken, (uint32_t)taken, m.shape_type, sh);
}

template <class T>
void run_op(const Ctx& c, const Meta& m, const Shape<T>& sh, Dec& d, std::vector<Out<T>>& out) {
    const pga_dispatch_entry& e = c.entries[m.entry];
    const pga_procedure& p = c.procs[e.operator_code];
    bool st = (p.flags & PGA_PROC_STATIC) != 0;
    switch (p.op_type) {
    case PGA_OP_SUBDIV: op_subdivide(c, m, sh, e, st, d, out); break;
    case PGA_OP_REPEAT:
        if (p.flags & PGA_PROC_PARALLEL) throw OpErr(PGA_ERR_UNSUPPORTED, "Repeat<true> traps (Repeat.cuh:35-61)");
        op_repeat(c, m, sh, e, d, out);
        break;
    case PGA_OP_COMPSPLIT:
        if (is_prism(m.shape_type)) {
            if (p.flags & PGA_PROC_PARALLEL)
                throw OpErr(PGA_ERR_UNSUPPORTED, "the parallel ComponentSplit splits Boxes only");
            op_compsplit_prism(m, sh, e, out);
        } else if (p.flags & PGA_PROC_PARALLEL) {
            op_compsplit_parallel(m, sh, e, out);
        } else {
            op_compsplit(m, sh, e, out);
        }
        break;
    case PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON:
    case PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM:
    case PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON:
    case PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM:
        op_setas(m, sh, e, p.op_type, out);
        break;
    case PGA_OP_IFSIZELESS: op_ifsizeless(c, m, sh, e, d, out); break;
    case PGA_OP_IF: op_if(c, m, sh, e, d, out); break;
    case PGA_OP_STOCHASTIC: op_random(c, m, sh, e, st, d, out); break;
    case PGA_OP_TRANSLATE: op_translate(c, m, sh, e, out); break;
    case PGA_OP_ROTATE: op_rotate(c, m, sh, e, out); break;
    case PGA_OP_SCALE: op_scale(c, m, sh, e, out); break;
    case PGA_OP_SWAPSIZE: op_swapsize(c, m, sh, e, d, out); break;
    case PGA_OP_EXTRUDE: op_extrude(c, m, sh, e, out); break;
    case PGA_OP_REPLICATE: op_replicate(m, sh, e, out); break;
    case PGA_OP_DISCARD: break;                               // Discard.cuh:15-31
    case PGA_OP_COLLIDER:
    case PGA_OP_IFCOLLIDES:
        if (has_vertices(m.shape_type)) throw OpErr(PGA_ERR_UNSUPPORTED, "collision tests of Box and Quad shapes only");
        if (p.op_type == PGA_OP_COLLIDER) op_collider(c, m, sh, e, d, out);
        else op_ifcollides(c, m, sh, e, d, out);
        break;
    case PGA_OP_GENERATE:
        if (p.gen_func_idx != 0 && p.gen_func_idx != 1)
            throw OpErr(PGA_ERR_UNSUPPORTED, "generation functions 0 and 1 only");
        if (p.gen_func_idx == 0 && has_vertices(m.shape_type))
            throw OpErr(PGA_ERR_UNSUPPORTED, "generation function 0 of Box and Quad shapes only");
        op_generate(c, m, sh, e, d, out);
        break;
    default: throw OpErr(PGA_ERR_UNSUPPORTED, "unsupported operator type " + std::to_string(p.op_type));
    }
}

// GenerationFunction 0 (triangle soup) for Box and Quad, counter-clockwise
// (the default, PGA_CW off): ShapeGenerator<Box, false>::run
// (include/pga/core/ShapeGenerator.cuh:376-511) and ShapeGenerator<Quad,
// false>::run (:670-708). Vertex k of a shape: position = model4 * corner
// (matrix.h:793-799: ((m0 x + m1 y) + m2 z) + m3 w), normal =
// normalize(cross(v0 - v1, v2 - v1)) of its face (vector.h:510-515, 566-574:
// sqrt of ((x x + y y) + z z), times 1 / length), uv per corner. Nine floats
// per vertex: position xyzw, normal xyz, uv.
void xform4(const float m[12], float x, float y, float z, float o[4]) {
    for (int r = 0; r < 3; ++r) o[r] = ((m[r * 4] * x + m[r * 4 + 1] * y) + m[r * 4 + 2] * z) + m[r * 4 + 3] * 1.0f;
    o[3] = ((0.0f * x + 0.0f * y) + 0.0f * z) + 1.0f * 1.0f;
}

void face_vertices(const float m[12], const float c[4][3], std::vector<float>& out) {
    float v[4][4];
    for (int k = 0; k < 4; ++k) xform4(m, c[k][0], c[k][1], c[k][2], v[k]);
    const float a[3] = {v[0][0] - v[1][0], v[0][1] - v[1][1], v[0][2] - v[1][2]};
    const float b[3] = {v[2][0] - v[1][0], v[2][1] - v[1][1], v[2][2] - v[1][2]};
    const float n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const float len = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    const float inv = 1.0f / len;
    static const float uv[4][2] = {{1.0f, 1.0f}, {0.0f, 1.0f}, {0.0f, 0.0f}, {1.0f, 0.0f}};
    for (int k = 0; k < 4; ++k) {
        out.insert(out.end(), v[k], v[k] + 4);
        out.push_back(n[0] * inv);
        out.push_back(n[1] * inv);
        out.push_back(n[2] * inv);
        out.push_back(uv[k][0]);
        out.push_back(uv[k][1]);
    }
}

void shape_mesh(const Shape<float>& s, int32_t shape_type, std::vector<float>& out) {
    const float hx = s.size[0] * 0.5f, hy = s.size[1] * 0.5f, hz = s.size[2] * 0.5f;   // Shape.cuh:70-73
    if (shape_type == PGA_SHAPE_QUAD) {
        const float q[4][3] = {{-hx, hy, 0.0f}, {hx, hy, 0.0f}, {hx, -hy, 0.0f}, {-hx, -hy, 0.0f}};
        face_vertices(s.m, q, out);
        return;
    }
    const float f[6][4][3] = {
        {{-hx, hy, -hz}, {hx, hy, -hz}, {hx, hy, hz}, {-hx, hy, hz}},        // top
        {{-hx, -hy, hz}, {hx, -hy, hz}, {hx, -hy, -hz}, {-hx, -hy, -hz}},    // bottom
        {{-hx, hy, -hz}, {-hx, hy, hz}, {-hx, -hy, hz}, {-hx, -hy, -hz}},    // left
        {{hx, hy, hz}, {hx, hy, -hz}, {hx, -hy, -hz}, {hx, -hy, hz}},        // right
        {{-hx, hy, hz}, {hx, hy, hz}, {hx, -hy, hz}, {-hx, -hy, hz}},        // front
        {{hx, hy, -hz}, {-hx, hy, -hz}, {-hx, -hy, -hz}, {hx, -hy, -hz}}};   // back
    for (int k = 0; k < 6; ++k) face_vertices(s.m, f[k], out);
}

struct Item {
    Meta meta;
    Shape<float> f;
    Shape<double> dd;
    Shape<Var> v;
};

} // namespace

struct pgao_result {
    std::string error;
    uint32_t twins = 0;
    uint64_t steps = 0, executions = 0;
    uint32_t nparams = 0;
    std::vector<std::vector<pga_instance>> inst;
    std::vector<std::vector<float>> mesh;         // generation function 0: 9 floats per vertex
    std::vector<std::vector<double>> inst64;
    std::vector<std::vector<int32_t>> inst_ids;   // 12 tape ids per instance
    Tape tape;
    std::vector<int32_t> param_ids;
    uint32_t ncolliders = 0;   // shapes the Collider operator stored, all phases
};

extern "C" {

int pgao_derive(const pga_procedure* procs, uint32_t num_procs,
                const pga_dispatch_entry* entries, uint32_t num_entries,
                const pga_axiom* axioms, uint32_t num_axioms,
                const float* axiom_params, uint32_t params_per_axiom,
                uint32_t key0, uint32_t key1, uint32_t twins, uint32_t max_steps,
                pgao_result** out) {
    pgao_result* res = new pgao_result();
    *out = res;
    res->twins = twins;
    res->inst.resize(PGA_MAX_TERMINALS);
    res->mesh.resize(PGA_MAX_TERMINALS);
    res->inst64.resize(PGA_MAX_TERMINALS);
    res->inst_ids.resize(PGA_MAX_TERMINALS);
    Ctx c;
    c.procs = procs; c.nprocs = num_procs; c.entries = entries; c.nentries = num_entries;
    c.params_f = axiom_params; c.params_per_axiom = axiom_params ? params_per_axiom : 0;
    c.naxioms = num_axioms; c.key[0] = key0; c.key[1] = key1;
    res->nparams = c.params_per_axiom;
    g_tape = &res->tape;
    if (axiom_params) {
        size_t np = (size_t)num_axioms * params_per_axiom;
        c.params_d.resize(np);
        for (size_t i = 0; i < np; ++i) c.params_d[i] = axiom_params[i];
        if (twins & PGAO_TWIN_ADJ) {
            c.params_v.resize(np);
            res->param_ids.resize(np);
            for (size_t i = 0; i < np; ++i) {
                Var v(c.params_d[i]);
                v.id = res->tape.leaf();
                c.params_v[i] = v;
                res->param_ids[i] = v.id;
            }
        }
    }
    if (max_steps == 0) max_steps = 100000;   // DefaultConfiguration (CPUEvaluator.h:291-295)
    try {
        for (uint32_t i = 0; i < num_procs; ++i)
            if (procs[i].op_type < 1 || procs[i].op_type > 20)
                throw OpErr(PGA_ERR_BAD_PROGRAM, "bad operator type");
        auto check_entry = [&](int32_t ent) {
            if (ent < 0 || (uint32_t)ent >= num_entries)
                throw OpErr(PGA_ERR_BAD_PROGRAM, "entry index out of range: " + std::to_string(ent));
            int32_t op = entries[ent].operator_code;
            if (op < 0 || (uint32_t)op >= num_procs)
                throw OpErr(PGA_ERR_BAD_PROGRAM, "operator code out of range");
        };
        // phases: one more than the largest phase a successor names
        // (MultiPhaseEvaluator, GPUEvaluator.cuh:405-456: every phase runs to
        // completion, then the colliders it stored take part in the next)
        uint32_t nphases = 1;
        for (uint32_t i = 0; i < num_entries; ++i)
            for (uint32_t k = 0; k < entries[i].num_successors && k < PGA_MAX_NUM_SUCCESSORS; ++k) {
                const int32_t ph = entries[i].successors[k].phase_index;
                if (ph >= (int32_t)PGAO_MAX_PHASES) throw OpErr(PGA_ERR_BAD_PROGRAM, "bad phase index");
                nphases = std::max(nphases, ph > 0 ? (uint32_t)ph + 1u : 1u);
            }
        std::vector<Collider> colliders;
        c.colliders = &colliders;
        std::vector<std::vector<std::deque<Item>>> phase_queues(nphases, std::vector<std::deque<Item>>(num_procs));
        // axioms: AxiomGenerator::generateAxiom -> dispatchAxiom (CPUEvaluator.h:240-256)
        for (uint32_t a = 0; a < num_axioms; ++a) {
            const pga_axiom& ax = axioms[a];
            if (ax.entry_index == -1) continue;   // src/interpreter/PGAFacade.cu:708
            check_entry(ax.entry_index);
            Item it;
            it.meta.entry = ax.entry_index; it.meta.axiom = a; it.meta.path = 0; it.meta.depth = 0;
            it.meta.shape_type = ax.shape_type;
            if (has_vertices(ax.shape_type))
                throw OpErr(PGA_ERR_UNSUPPORTED, "axioms are Box or Quad shapes (no vertex list)");
            it.meta.phase = 0;
            for (int i = 0; i < 12; ++i) { it.f.m[i] = ax.model[i]; it.dd.m[i] = ax.model[i]; it.v.m[i] = Var(ax.model[i]); }
            for (int i = 0; i < 3; ++i) { it.f.size[i] = ax.size[i]; it.dd.size[i] = ax.size[i]; it.v.size[i] = Var(ax.size[i]); }
            it.f.seed = it.dd.seed = it.v.seed = ax.seed;
            it.f.custom = it.dd.custom = it.v.custom = ax.custom;
            phase_queues[0][entries[ax.entry_index].operator_code].push_back(it);
        }
        // while (queuesManager.dequeueAll() && ++derivationStep < MaxDerivationSteps)
        // (CPUEvaluator.h:429; RoundRobinQueuesManager::dequeueAll :185-192;
        // Queue::dequeue drains FIFO including self-appended items :45-61)
        uint64_t step = 0;
        for (uint32_t phase = 0; phase < nphases; ++phase) {
        c.phase = phase;
        std::vector<std::deque<Item>>& queues = phase_queues[phase];
        bool proceed = true;
        while (proceed) {
            proceed = false;
            for (uint32_t q = 0; q < num_procs; ++q) {
                if (queues[q].empty()) continue;
                proceed = true;
                while (!queues[q].empty()) {
                    Item it = queues[q].front();
                    queues[q].pop_front();
                    res->executions++;
                    Dec d;
                    std::vector<Out<float>> of;
                    run_op<float>(c, it.meta, it.f, d, of);
                    std::vector<Out<double>> od;
                    std::vector<Out<Var>> ov;
                    if (twins & PGAO_TWIN_F64) { d.rec = false; d.cur = 0; run_op<double>(c, it.meta, it.dd, d, od); }
                    if (twins & PGAO_TWIN_ADJ) { d.rec = false; d.cur = 0; run_op<Var>(c, it.meta, it.v, d, ov); }
                    for (size_t k = 0; k < of.size(); ++k) {
                        const Out<float>& o = of[k];
                        if (o.kind == 1) {
                            pga_instance inst;
                            std::memcpy(inst.model, o.s.m, sizeof(inst.model));
                            inst.custom = o.s.custom;
                            inst.axiom = it.meta.axiom;
                            inst.path = it.meta.path;
                            inst.meta = (it.meta.phase << 28) | (it.meta.depth << 16) | ((uint32_t)it.meta.entry & 0xFFFFu);
                            res->inst[o.terminal].push_back(inst);
                            if (procs[entries[it.meta.entry].operator_code].gen_func_idx == 0)
                                shape_mesh(it.f, it.meta.shape_type, res->mesh[o.terminal]);
                            if (twins & PGAO_TWIN_F64) {
                                for (int i = 0; i < 12; ++i) res->inst64[o.terminal].push_back(od[k].s.m[i]);
                                res->inst64[o.terminal].push_back(od[k].s.custom);
                            }
                            if (twins & PGAO_TWIN_ADJ)
                                for (int i = 0; i < 12; ++i) res->inst_ids[o.terminal].push_back(ov[k].s.m[i].id);
                            continue;
                        }
                        if (o.entry < 0) continue;      // unused dynamic edge
                        check_entry(o.entry);
                        Item ch;
                        ch.meta.entry = o.entry;
                        ch.meta.axiom = it.meta.axiom;
                        ch.meta.path = path_mix(it.meta.path, o.ordinal);
                        ch.meta.depth = it.meta.depth + 1;
                        ch.meta.shape_type = o.shape_type;
                        if (o.has_poly) ch.meta.poly = o.poly;
                        else if (has_vertices(o.shape_type)) ch.meta.poly = it.meta.poly;
                        ch.meta.phase = o.phase == kPhaseInherit ? it.meta.phase : o.phase;
                        // a later phase's shape waits for it; an earlier phase's is
                        // never derived (that phase has run); with phases, the depth
                        // shares its 16 bits of the instance meta with the phase
                        if (ch.meta.phase < phase) continue;
                        if (nphases > 1 && ch.meta.depth >= 4096u)
                            throw OpErr(PGA_ERR_MAX_DEPTH, "depth 4096 reached in a multi-phase program");
                        ch.f = o.s;
                        if (twins & PGAO_TWIN_F64) ch.dd = od[k].s;
                        if (twins & PGAO_TWIN_ADJ) ch.v = ov[k].s;
                        phase_queues[ch.meta.phase][entries[o.entry].operator_code].push_back(ch);
                    }
                }
            }
            if (proceed && ++step >= max_steps) {
                bool left = false;
                for (auto& q : queues) left |= !q.empty();
                if (left) throw OpErr(PGA_ERR_MAX_DEPTH, "maximum number of derivation steps reached");
                break;
            }
        }
        }
        res->steps = step;
        res->ncolliders = (uint32_t)colliders.size();
    } catch (const OpErr& e) {
        res->error = e.what();
        g_tape = nullptr;
        return e.code;
    } catch (const std::exception& e) {
        res->error = e.what();
        g_tape = nullptr;
        return PGA_ERR_BAD_PROGRAM;
    }
    g_tape = nullptr;
    return PGA_OK;
}

void pgao_free(pgao_result* r) { delete r; }
const char* pgao_error(const pgao_result* r) { return r ? r->error.c_str() : "null result"; }
uint32_t pgao_num_colliders(const pgao_result* r) { return r ? r->ncolliders : 0u; }
int pgao_shapes_collide(const float* am, const float* asz, int32_t at, const float* bm, const float* bsz, int32_t bt) {
    return shapes_collide(am, asz, at, bm, bsz, bt) ? 1 : 0;
}
uint32_t pgao_num_terminals(const pgao_result* r) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < PGA_MAX_TERMINALS; ++i) if (!r->inst[i].empty()) n = i + 1;
    return n;
}
uint64_t pgao_count(const pgao_result* r, uint32_t t) { return t < PGA_MAX_TERMINALS ? r->inst[t].size() : 0; }
uint64_t pgao_steps(const pgao_result* r) { return r->steps; }
uint64_t pgao_executions(const pgao_result* r) { return r->executions; }
int pgao_copy_instances(const pgao_result* r, uint32_t t, pga_instance* dst) {
    if (t >= PGA_MAX_TERMINALS) return PGA_ERR_INVALID_ARG;
    if (!r->inst[t].empty()) std::memcpy(dst, r->inst[t].data(), r->inst[t].size() * sizeof(pga_instance));
    return PGA_OK;
}
uint64_t pgao_mesh_floats(const pgao_result* r, uint32_t t) {
    return t < PGA_MAX_TERMINALS ? (uint64_t)r->mesh[t].size() : 0u;
}
int pgao_copy_mesh(const pgao_result* r, uint32_t t, float* dst) {
    if (t >= PGA_MAX_TERMINALS) return PGA_ERR_INVALID_ARG;
    if (!r->mesh[t].empty()) std::memcpy(dst, r->mesh[t].data(), r->mesh[t].size() * sizeof(float));
    return PGA_OK;
}
int pgao_copy_instances_f64(const pgao_result* r, uint32_t t, double* dst) {
    if (t >= PGA_MAX_TERMINALS || !(r->twins & PGAO_TWIN_F64)) return PGA_ERR_INVALID_ARG;
    if (!r->inst64[t].empty()) std::memcpy(dst, r->inst64[t].data(), r->inst64[t].size() * sizeof(double));
    return PGA_OK;
}
int pgao_backward(pgao_result* r, const double* dl_dmodel, double* dl_dparams) {
    if (!(r->twins & PGAO_TWIN_ADJ)) return PGA_ERR_INVALID_ARG;
    const Tape& tp = r->tape;
    std::vector<double> adj(tp.a.size(), 0.0);
    size_t k = 0;
    for (uint32_t t = 0; t < PGA_MAX_TERMINALS; ++t)
        for (int32_t id : r->inst_ids[t]) {
            if (id >= 0) adj[id] += dl_dmodel[k];
            ++k;
        }
    for (size_t i = tp.a.size(); i-- > 0;) {
        double g = adj[i];
        if (g == 0.0) continue;
        if (tp.a[i] >= 0) adj[tp.a[i]] += g * tp.wa[i];
        if (tp.b[i] >= 0) adj[tp.b[i]] += g * tp.wb[i];
    }
    for (size_t i = 0; i < r->param_ids.size(); ++i) dl_dparams[i] = adj[r->param_ids[i]];
    return PGA_OK;
}

float pgao_next_seed(float base, uint32_t modifier) { return next_seed(base, modifier); }
uint32_t pgao_seed_component(uint32_t z, uint32_t which) { return seed_component(z, which); }
void pgao_seed_state(float base, uint32_t modifier, uint32_t* z4) { seed_state(base, modifier, z4); }
float pgao_seed_unit(uint32_t t) { return seed_unit(t); }
float pgao_range(float seed, float mn, float mx) { return range_t<float>(seed, mn, mx); }
void pgao_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }
float pgao_sinf(float x) { float s, c; sincos_f32(x, &s, &c); return s; }
float pgao_cosf(float x) { float s, c; sincos_f32(x, &s, &c); return c; }
uint32_t pgao_path_mix(uint32_t path, uint32_t ordinal) { return path_mix(path, ordinal); }
float pgao_pack_operand(int type, float x) {            // PackUtils::packOperand (PackUtils.h:18-23)
    uint32_t r = fbits(x);
    uint32_t q = (r >> 2) | (((uint32_t)type) << 30);
    return bitsf(q);
}
float pgao_pack_float2(float x, float y) {              // PackUtils::packFloat2 (PackUtils.h:25-28)
    return std::floor(x * (4096.0f - 1.0f)) * 4096.0f + std::floor(y * (4096.0f - 1.0f));
}

// ---------------------------------------------------------------------------
// LBVH (GPUBVHConstructor.cuh:44-261), sequential restatement.
// ---------------------------------------------------------------------------
static inline uint32_t spread(uint32_t v) {                 // :48-54
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
uint32_t pgao_morton30(float x, float y, float z) {         // morton30 (:56-64), COORD_LIM = 1023
    uint32_t vx = f2u_rd(x * 1023.0f), vy = f2u_rd(y * 1023.0f), vz = f2u_rd(z * 1023.0f);
    if (vx > 1023u) vx = 1023u;
    if (vy > 1023u) vy = 1023u;
    if (vz > 1023u) vz = 1023u;
    return spread(vx) | (spread(vy) << 1) | (spread(vz) << 2);
}
static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

int pgao_lbvh(const float* aabbs, uint32_t n, const float smin[3], const float smax[3],
              uint64_t* keys, int32_t* left, int32_t* right, int32_t* parent,
              int32_t* leaf_parent, float* node_aabb) {
    if (n < 2) return PGA_ERR_INVALID_ARG;
    float ext[3] = {smax[0] - smin[0], smax[1] - smin[1], smax[2] - smin[2]};
    for (uint32_t i = 0; i < n; ++i) {                      // generateMortonCodes (:66-86)
        const float* b = aabbs + (size_t)i * 6;
        float cx = (b[0] + b[3]) * 0.5f, cy = (b[1] + b[4]) * 0.5f, cz = (b[2] + b[5]) * 0.5f;
        float nx = (cx - smin[0]) / ext[0], ny = (cy - smin[1]) / ext[1], nz = (cz - smin[2]) / ext[2];
        keys[i] = ((uint64_t)pgao_morton30(nx, ny, nz) << 32) | i;
    }
    std::sort(keys, keys + n);                              // thrust::sort (:311-312); keys unique
    auto delta = [&](int64_t i, int64_t j) -> int {         // delta (:33-36), -1 outside the range
        if (j < 0 || j >= (int64_t)n) return -1;
        return clz64(keys[i] ^ keys[j]);
    };
    for (int64_t i = 0; i < (int64_t)n - 1; ++i) parent[i] = -1;
    for (int64_t i = 0; i < (int64_t)n - 1; ++i) {          // buildRadixTree (:88-213), Karras 2012
        int d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
        int dmin = delta(i, i - d);
        int64_t lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta(i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int dnode = delta(i, j);
        int64_t s = 0;
        for (int64_t t = (l + 1) / 2;; t = (t + 1) / 2) {
            if (delta(i, i + (s + t) * d) > dnode) s += t;
            if (t == 1) break;
        }
        int64_t gamma = i + s * d + std::min(d, 0);
        int32_t lc = (std::min(i, j) == gamma) ? (int32_t)gamma : -(int32_t)gamma;
        int32_t rc = (std::max(i, j) == gamma + 1) ? (int32_t)(gamma + 1) : -(int32_t)(gamma + 1);
        left[i] = lc; right[i] = rc;
        if (lc < 0) parent[-lc] = (int32_t)i; else leaf_parent[lc] = (int32_t)i;
        if (rc < 0) parent[-rc] = (int32_t)i; else leaf_parent[rc] = (int32_t)i;
    }
    (void)clz32;
    // propagateBVH (:215-261): bottom-up union; done here in post-order.
    std::vector<int32_t> order;
    std::vector<int32_t> stack{0};
    while (!stack.empty()) {
        int32_t v = stack.back(); stack.pop_back();
        order.push_back(v);
        if (left[v] < 0) stack.push_back(-left[v]);
        if (right[v] < 0) stack.push_back(-right[v]);
    }
    for (size_t k = order.size(); k-- > 0;) {
        int32_t v = order[k];
        float bb[6];
        for (int side = 0; side < 2; ++side) {
            int32_t ch = side == 0 ? left[v] : right[v];
            const float* src = ch < 0 ? node_aabb + (size_t)(-ch) * 6 : aabbs + (size_t)(keys[ch] & 0xFFFFFFFFu) * 6;
            if (side == 0) { for (int q = 0; q < 6; ++q) bb[q] = src[q]; }
            else {
                for (int q = 0; q < 3; ++q) bb[q] = std::min(bb[q], src[q]);
                for (int q = 3; q < 6; ++q) bb[q] = std::max(bb[q], src[q]);
            }
        }
        for (int q = 0; q < 6; ++q) node_aabb[(size_t)v * 6 + q] = bb[q];
    }
    return PGA_OK;
}

} // extern "C"
