// pga_adjoint.hip - reverse mode over a derivation tape (gfx950).
//
// A tape-mode derivation keeps every level's parent records and child
// descriptors, materialises every fan-out (no in-place leaf expansion), and
// records one outcome per child: the terminal instance it emitted, or the
// parent record and descriptor range of the fan-out it became, or nothing.
// The tape freezes the topology (branches, repetition counts, rule choices).
//
// The backward pass runs one launch per level, last level first, one thread
// per child. Each thread recomputes its child from the tape (the parent
// record and the descriptor: the checkpoint of that depth), replays the
// child's chain of single-child operators, and propagates adjoints back:
//   instance = model4 * scale(size)  <- dL/dinstance of its terminal, or
//   the fan-out it became            <- the adjoints its children left in
//                                       their slots (parent state, offset and
//                                       size along the split axis), through
//                                       the running offsets of the split;
//   chain of single-child operators  (recomputed forward from its start);
//   child construction               -> this child's slot, read by its
//                                       parent's thread in the next launch.
// Parameter adjoints of PGA_PT_AXIOM_PARAM slots are summed with FP64
// atomic adds (many terms of both signs per parameter) and rounded to FP32
// once; every other parameter is a constant. Each slot is written
// by one thread and read by one, so the state adjoints are deterministic.
#include "pga_kernels.h"

namespace pga {

struct Adj {
    float m[12];
    float s[3];
};

__device__ __forceinline__ void adj_zero(Adj& d) {
#pragma unroll
    for (int i = 0; i < 12; ++i) d.m[i] = 0.0f;
    d.s[0] = d.s[1] = d.s[2] = 0.0f;
}

// dL/dparams of one thread. A thread replays one child, so every
// contribution is to the same axiom: up to kGradSlots per-axiom parameters
// accumulate in registers (FP64) and leave once, at the end (flush_grads).
constexpr uint32_t kGradSlots = 8;
struct GradAcc {
    double v[kGradSlots];
    uint32_t axiom;   // ~0u: no contribution yet
};

__device__ __forceinline__ void grad_zero(GradAcc& ga) {
#pragma unroll
    for (uint32_t q = 0; q < kGradSlots; ++q) ga.v[q] = 0.0;
    ga.axiom = ~0u;
}

// dL/dparam k of entry e for the shape of `axiom`: only per-axiom parameters
// are differentiable.
__device__ __forceinline__ void grad_param(const Ctx& c, const AdjArgs& g, GradAcc& ga, const EntryV& e, uint32_t k,
                                           uint32_t axiom, float v) {
    if (v == 0.0f || k >= e.nparams()) return;
    if (k < 3u && ((e.inl() >> k) & 1u)) return;
    const DevParam p = c.prog.params[e.pbase + k];
    if (p.type != PGA_PT_AXIOM_PARAM || c.aparams_per == 0u) return;
    const uint32_t slot = (uint32_t)p.a;
    if (c.aparams_per <= kGradSlots && slot < kGradSlots && (ga.axiom == ~0u || ga.axiom == axiom)) {
        ga.axiom = axiom;
#pragma unroll
        for (uint32_t q = 0; q < kGradSlots; ++q) ga.v[q] += q == slot ? (double)v : 0.0;
    } else {
        atomicAdd(g.grad + (size_t)axiom * c.aparams_per + slot, (double)v);
    }
}

// The accumulated contributions to the FP64 gradient: a full wave whose
// contributions all go to one axiom sums each parameter across its lanes
// first (one atomic per parameter instead of one per lane and operator);
// other waves add lane by lane.
__device__ __forceinline__ void flush_grads(const Ctx& c, const AdjArgs& g, const GradAcc& ga) {
    const uint32_t per = c.aparams_per;
    if (per == 0u || per > kGradSlots) return;
    const bool has = ga.axiom != ~0u;
    const unsigned long long act = __ballot(1), hm = __ballot(has);
    if (hm == 0ull) return;
    const int lead = __ffsll((long long)hm) - 1;
    const uint32_t ax0 = __shfl(ga.axiom, lead);
    const bool uniform = act == ~0ull && __ballot(has && ga.axiom != ax0) == 0ull;
#pragma unroll
    for (uint32_t q = 0; q < kGradSlots; ++q) {   // unrolled: ga.v stays in registers
        if (q >= per || __ballot(ga.v[q] != 0.0) == 0ull) continue;
        if (uniform) {
            double x = ga.v[q];
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d);
            if ((int)(threadIdx.x & 63u) == lead) atomicAdd(g.grad + (size_t)ax0 * per + q, x);
        } else if (ga.v[q] != 0.0) {
            atomicAdd(g.grad + (size_t)ga.axiom * per + q, ga.v[q]);
        }
    }
}

// c = a * translate(t): da, dt from dc.
__device__ __forceinline__ void translate_adj(const float a[12], const float t[3], const float dc[12], float da[12],
                                              float dt[3]) {
    dt[0] = dt[1] = dt[2] = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float d3 = dc[i * 4 + 3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            da[i * 4 + k] = dc[i * 4 + k] + d3 * t[k];
            dt[k] += a[i * 4 + k] * d3;
        }
        da[i * 4 + 3] = d3;
    }
}

// c = a * b (a 3x4 with the implicit row 0 0 0 1, b 4x4): da = dc * b^T.
__device__ __forceinline__ void mul_adj_a(const float b[16], const float dc[12], float da[12]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) s += dc[i * 4 + j] * b[k * 4 + j];
            da[i * 4 + k] = s;
        }
}

// Instance = model4 * scale(size) (GenerationFunctions.cuh:155-189).
__device__ __forceinline__ void emit_adj(const Rec& r, const float* dl, Adj& d) {
    float di[12];   // rows 0..2 of dL/dinstance, three 16-byte global loads
    // (the pointer comes from a table, so without the address space the
    // loads would be flat loads, which also hold up the next LDS wait)
    typedef float vf4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) const vf4 gcf4;
    const gcf4* d4 = (const gcf4*)dl;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const vf4 v = d4[i];
        di[i * 4] = v.x; di[i * 4 + 1] = v.y; di[i * 4 + 2] = v.z; di[i * 4 + 3] = v.w;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            d.m[i * 4 + j] += di[i * 4 + j] * r.s[j];
            d.s[j] += di[i * 4 + j] * r.m[i * 4 + j];
        }
        d.m[i * 4 + 3] += di[i * 4 + 3];
    }
}

// Reverse of one single-child operator applied to the state `r` (before
// the operator): d holds the adjoint after it and is turned into the
// adjoint before it.
__device__ __forceinline__ void single_adj(const Ctx& c, const AdjArgs& g, GradAcc& ga, const EntryV& e, const Rec& r,
                                           Adj& d) {
    const uint32_t k = e.kind();
    if (k == K_IFSIZELESS || k == K_IF || k == K_RANDOM) return;
    float v0, v1, v2;
    eval3(c, e, r, v0, v1, v2);
    if (k == K_TRANSLATE) {
        const float t[3] = {v0, v1, v2};
        float dm[12], dt[3];
        translate_adj(r.m, t, d.m, dm, dt);
#pragma unroll
        for (int i = 0; i < 12; ++i) d.m[i] = dm[i];
        grad_param(c, g, ga, e, 0, r.axiom, dt[0]);
        grad_param(c, g, ga, e, 1, r.axiom, dt[1]);
        grad_param(c, g, ga, e, 2, r.axiom, dt[2]);
        return;
    }
    if (k == K_ROTATE) {                       // m' = m * R(alpha, beta, gamma)
        const float kd = 3.1415926535897932384626434f / 180.0f;
        float sa, sb, sg, ca, cb, cg;
        sincos_f32(v0 * kd, &sa, &ca);
        sincos_f32(v1 * kd, &sb, &cb);
        sincos_f32(v2 * kd, &sg, &cg);
        const float R[16] = {cb * cg, cg * sa * sb - ca * sg, ca * cg * sb + sa * sg, 0.0f,
                             cb * sg, ca * cg + sa * sb * sg, -cg * sa + ca * sb * sg, 0.0f,
                             -sb, cb * sa, ca * cb, 0.0f,
                             0.0f, 0.0f, 0.0f, 1.0f};
        // dR[k][j] = sum_i m[i][k] dm'[i][j] (3x3 part)
        float dR[9];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float s = 0.0f;
#pragma unroll
                for (int i = 0; i < 3; ++i) s += r.m[i * 4 + kk] * d.m[i * 4 + j];
                dR[kk * 3 + j] = s;
            }
        const float Ra[9] = {0.0f, cg * ca * sb + sa * sg, -sa * cg * sb + ca * sg,
                             0.0f, -sa * cg + ca * sb * sg, -cg * ca - sa * sb * sg,
                             0.0f, cb * ca, -sa * cb};
        const float Rb[9] = {-sb * cg, cg * sa * cb, ca * cg * cb,
                             -sb * sg, sa * cb * sg, ca * cb * sg,
                             -cb, -sb * sa, -ca * sb};
        const float Rg[9] = {-cb * sg, -sg * sa * sb - ca * cg, -ca * sg * sb + sa * cg,
                             cb * cg, -ca * sg + sa * sb * cg, sg * sa + ca * sb * cg,
                             0.0f, 0.0f, 0.0f};
        float dal = 0.0f, dbe = 0.0f, dgm = 0.0f;   // dL/d(alpha, beta, gamma) in radians
#pragma unroll
        for (int q = 0; q < 9; ++q) { dal += dR[q] * Ra[q]; dbe += dR[q] * Rb[q]; dgm += dR[q] * Rg[q]; }
        float dm[12];
        mul_adj_a(R, d.m, dm);
#pragma unroll
        for (int i = 0; i < 12; ++i) d.m[i] = dm[i];
        grad_param(c, g, ga, e, 0, r.axiom, dal * kd);
        grad_param(c, g, ga, e, 1, r.axiom, dbe * kd);
        grad_param(c, g, ga, e, 2, r.axiom, dgm * kd);
        return;
    }
    if (k == K_SCALE) {                        // s'_i = v_i > 0 ? v_i : (v_i < 0 ? -s_i v_i : 0)
        const float v[3] = {v0, v1, v2};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float ds = d.s[i];
            if (v[i] < 0.0f) {
                d.s[i] = ds * -v[i];
                grad_param(c, g, ga, e, (uint32_t)i, r.axiom, ds * -r.s[i]);
            } else {
                d.s[i] = 0.0f;
                if (v[i] > 0.0f) grad_param(c, g, ga, e, (uint32_t)i, r.axiom, ds);
            }
        }
        return;
    }
    if (k == K_SWAPSIZE) {                     // s'_k = s[a_k]
        const uint32_t a0 = min(f2u_rd(v0), 2u), a1 = min(f2u_rd(v1), 2u), a2 = min(f2u_rd(v2), 2u);
        const float d0 = d.s[0], d1 = d.s[1], d2 = d.s[2];
        // d.s[a_k] += d_k in k order, written with selects (no indexed local)
#pragma unroll
        for (uint32_t q = 0; q < 3; ++q) {
            float x = 0.0f;
            x = a0 == q ? x + d0 : x;
            x = a1 == q ? x + d1 : x;
            x = a2 == q ? x + d2 : x;
            d.s[q] = x;
        }
        return;
    }
    // K_EXTRUDE: m' = m * B, s' = (s0, v1, s1)
    const float B[16] = {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, -1.0f, 0.0f,
                         0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    float dm[12];
    mul_adj_a(B, d.m, dm);
#pragma unroll
    for (int i = 0; i < 12; ++i) d.m[i] = dm[i];
    const float ds0 = d.s[0], ds1 = d.s[1], ds2 = d.s[2];
    d.s[0] = ds0;
    d.s[1] = ds2;
    d.s[2] = 0.0f;
    grad_param(c, g, ga, e, 1, r.axiom, ds1);
}

// Reverse of build_child: the child state adjoint dc becomes the parent
// state adjoint dp and the adjoints of the child's offset and size.
__device__ __forceinline__ void build_adj(const Rec& p, const EntryV& pe, uint32_t j, uint32_t axis, float o,
                                          const Adj& dc, Adj& dp, float& d_o, float& d_ns) {
    adj_zero(dp);
    d_o = 0.0f;
    d_ns = 0.0f;
    const uint32_t kind = pe.kind();
    if (kind == K_SUBDIV || kind == K_REPEAT) {
        float t[3];
        axis_vec(axis, o, t[0], t[1], t[2]);
        float dt[3];
        translate_adj(p.m, t, dc.m, dp.m, dt);
        // selects, not indexing by `axis`: an indexed local lives in scratch
        d_o = axis == 0u ? dt[0] : (axis == 1u ? dt[1] : dt[2]);
#pragma unroll
        for (uint32_t q = 0; q < 3; ++q) dp.s[q] = q == axis ? 0.0f : dc.s[q];
        d_ns = axis == 0u ? dc.s[0] : (axis == 1u ? dc.s[1] : dc.s[2]);
        return;
    }
    if (kind == K_COMPSPLIT) {                 // m_c = (m * translate(t_j)) * B_j
        const float hx = p.s[0] * 0.5f, hy = p.s[1] * 0.5f, hz = p.s[2] * 0.5f;
        float t[3] = {0.0f, 0.0f, 0.0f};
        float b00 = 1.0f, b02 = 0.0f, b06 = 0.0f, b09 = 0.0f, b08 = 0.0f, b10 = 1.0f, b05 = 1.0f;
        bool rotate = true;
        if (j == 0) { t[2] = -hz; b00 = -1.0f; b10 = -1.0f; }
        else if (j == 1) { t[2] = hz; rotate = false; }
        else if (j == 2) { t[0] = -hx; b00 = 0.0f; b02 = -1.0f; b08 = 1.0f; b10 = 0.0f; }
        else if (j == 3) { t[0] = hx; b00 = 0.0f; b02 = 1.0f; b08 = -1.0f; b10 = 0.0f; }
        else if (j == 4) { t[1] = -hy; b05 = 0.0f; b06 = -1.0f; b09 = 1.0f; b10 = 0.0f; }
        else { t[1] = hy; b05 = 0.0f; b06 = 1.0f; b09 = -1.0f; b10 = 0.0f; }
        float dtmp[12];
        if (rotate) {
            const float B[16] = {b00, 0.0f, b02, 0.0f, 0.0f, b05, b06, 0.0f,
                                 b08, b09, b10, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
            mul_adj_a(B, dc.m, dtmp);
        } else {
#pragma unroll
            for (int i = 0; i < 12; ++i) dtmp[i] = dc.m[i];
        }
        float dt[3];
        translate_adj(p.m, t, dtmp, dp.m, dt);
        const float sgn = (j == 0 || j == 2 || j == 4) ? -0.5f : 0.5f;
        const uint32_t ta = j < 2 ? 2u : (j < 4 ? 0u : 1u);
        dp.s[ta] += sgn * dt[ta];
        // sizes: s_c = ((j == 2 || j == 3) ? sz : sx, j < 4 ? sy : sz, 1)
        if (j == 2 || j == 3) dp.s[2] += dc.s[0]; else dp.s[0] += dc.s[0];
        if (j < 4) dp.s[1] += dc.s[1]; else dp.s[2] += dc.s[1];
        return;
    }
    // K_REPLICATE: a copy
    dp = dc;
}

__device__ __forceinline__ const float* slot_at(const AdjArgs& g, uint32_t child) {
    return g.slots + (size_t)child * kSlotFloats;
}

// Reverse of a fan-out at state `r` (entry e), whose children's slots start
// at child index `first` of the next level.
__device__ __forceinline__ void fan_adj(const Ctx& c, const AdjArgs& g, GradAcc& ga, const EntryV& e, const Rec& r, uint32_t first,
                                        uint32_t nd, bool self, Adj& d) {
    // parent-state adjoints the children left (four 16-byte loads per slot),
    // and the sum of their offset adjoints in child order
    float total_o = 0.0f;
    for (uint32_t q = 0; q < nd; ++q) {
        const float4* sl = reinterpret_cast<const float4*>(slot_at(g, first + q));
        const float4 a = sl[0], b = sl[1], cc = sl[2], dd = sl[3];
        d.m[0] += a.x; d.m[1] += a.y; d.m[2] += a.z; d.m[3] += a.w;
        d.m[4] += b.x; d.m[5] += b.y; d.m[6] += b.z; d.m[7] += b.w;
        d.m[8] += cc.x; d.m[9] += cc.y; d.m[10] += cc.z; d.m[11] += cc.w;
        d.s[0] += dd.x; d.s[1] += dd.y; d.s[2] += dd.z;
        total_o += dd.w;
    }
    if (self) return;
    const uint32_t kind = e.kind();
    if (kind != K_SUBDIV && kind != K_REPEAT) return;
    Fan f;
    fan_init(c, r, e, f);
    if (f.n == 0u) return;
    uint32_t skipm = c.prog.entries[e.idx].aux[0];
    if (kind == K_REPEAT) skipm = (skipm & 1u) ? 0xFFFFFFFFu : 0u;
    // o_j = off_j + ns_j / 2, off_{j+1} = off_j + ns_j, off_0 = -total / 2:
    //   dL/doff_0 = sum_j do_j, dL/dns_j = dns_j + do_j / 2 + sum_{k > j} do_k
    float prefix = 0.0f, d_r2a = 0.0f, d_e1 = 0.0f, d_e2 = 0.0f, d_e3 = 0.0f;
    uint32_t q = 0;
    Fan ff = f;
    for (uint32_t j = 0; j < f.n; ++j) {
        float o, ns, fac = 0.0f;
        const bool keep = fan_geom(c, r, e, ff, j, o, ns, &fac);
        const bool skip = j < 32u ? ((skipm >> j) & 1u) != 0u : skipm == 0xFFFFFFFFu;
        float d_oj = 0.0f, d_nsj = 0.0f;
        if (!skip) {
            if (q < nd) { d_oj = slot_at(g, first + q)[15]; d_nsj = slot_at(g, first + q)[16]; }
            ++q;
        }
        if (!keep) continue;
        prefix += d_oj;
        const float dn = d_nsj + 0.5f * d_oj + (total_o - prefix);
        if (kind == K_SUBDIV) {
            if (fac > 0.0f) {
                grad_param(c, g, ga, e, j + 1u, r.axiom, dn);
            } else if (fac < 0.0f) {
                d_r2a += dn * -fac;
                grad_param(c, g, ga, e, j + 1u, r.axiom, dn * -f.r2a);
            }
        } else {
            if (j == 0u) d_e1 += dn;
            else if (j + 1u < f.reps) d_e2 += dn;
            else d_e3 += dn;
        }
    }
    float d_total = -0.5f * total_o;
    if (kind == K_SUBDIV) {
        // r2a = (total - sum of absolute factors) / (sum of |relative factors|)
        float abs_sum = 0.0f, rel_sum = 0.0f;
        const uint32_t ax = r.axiom, pa = r.path, de = rec_depth(c, r);
        for (uint32_t i = 1; i < e.nparams(); ++i) {
            const float fac = P(c, e, i, r, ax, pa, de);
            abs_sum += (float)(fac > 0.0f) * fac;
            rel_sum += (float)(fac < 0.0f) * (fac * -1.0f);
        }
        if (rel_sum != 0.0f && d_r2a != 0.0f) {
            d_total += d_r2a / rel_sum;
            const float d_abs = -d_r2a / rel_sum, d_rel = -d_r2a * f.r2a / rel_sum;
            for (uint32_t i = 1; i < e.nparams(); ++i) {
                const float fac = P(c, e, i, r, ax, pa, de);
                if (fac > 0.0f) grad_param(c, g, ga, e, i, r.axiom, d_abs);
                else if (fac < 0.0f) grad_param(c, g, ga, e, i, r.axiom, -d_rel);
            }
        }
    } else {
        // Repeat pieces from (size, extent), the repetition count frozen
        float v0, ex, v2;
        eval3(c, e, r, v0, ex, v2);
        const uint32_t mode = f2u_rd(v2);
        const float sz = f.total;
        const float extent = (((float)(ex < 0.0f) * sz) * fabsf(ex)) + ((float)(ex > 0.0f) * ex);
        float d_sz = 0.0f, d_ext = 0.0f;
        if (mode == PGA_ADJUST_TO_FILL) {
            if (f.reps > 0u) d_sz += (d_e1 + d_e2 + d_e3) / (float)f.reps;
        } else {
            // the remainder piece rest = fmod(sz, extent) (extent when exact)
            // is the first piece (ANCHOR_TO_START) or the last (ANCHOR_TO_END)
            const bool start = mode == PGA_ANCHOR_TO_START;
            const float d_rest = start ? d_e1 : d_e3;
            d_ext += start ? (d_e2 + d_e3) : (d_e1 + d_e2);
            const float rest = fmodf(sz, extent);
            if (rest == 0.0f) {
                d_ext += d_rest;
            } else {
                d_sz += d_rest;
                d_ext += d_rest * -truncf(sz / extent);
            }
        }
        if (ex > 0.0f) {
            grad_param(c, g, ga, e, 1, r.axiom, d_ext);
        } else if (ex < 0.0f) {
            d_sz += d_ext * -ex;
            grad_param(c, g, ga, e, 1, r.axiom, d_ext * -sz);
        }
        d_total += d_sz;
    }
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) d.s[q] = q == f.axis ? d.s[q] + d_total : d.s[q];
}

__device__ __forceinline__ Ctx adj_ctx(const KArgs& a) {
    Ctx c;
    c.prog = a.prog;
    c.inst = a.inst;
    c.inst_cap = a.inst_cap;
    c.aparams = a.aparams;
    c.aparams_per = a.aparams_per;
    c.naxioms_total = a.naxioms_total;
    c.key0 = a.key0;
    c.key1 = a.key1;
    c.term_counts = a.term_counts;
    c.par_out = nullptr;
    c.desc_out = nullptr;
    c.par_count = nullptr;
    c.desc_count = nullptr;
    c.par_cap = a.cap_par;
    c.desc_cap = a.cap_desc;
    c.err = a.err;
    c.max_chain = a.max_chain;
    c.geom = 1u;
    c.seed = a.use_seed;
    c.poly = 0u;             // the tape refuses dynamic polygons
    c.tape = 1u;
    c.prog_lds = 0u;
    c.ctxs = 0u;             // one-phase programs without Collider / IfCollides only
    c.dmask = 0xFFFFu;
    c.phase = 0u;
    c.wr = 0u;
    c.coll = nullptr;
    c.coll_count = nullptr;
    c.coll_visible = nullptr;
    c.coll_cap = 0u;
    c.bvh_flat = nullptr;
    c.coll_bits = nullptr;
    c.coll_bits_cap = 0u;
    return c;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_adjoint(KArgs a, AdjArgs g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n_children) return;
    const Ctx c = adj_ctx(a);
    const uint32_t gi = g.child_base + i;
    if (gi >= a.tape_out_cap || gi >= g.slots_cap) return;
    float* my = g.slots + (size_t)gi * kSlotFloats;
    const uint4 out = a.tape_out[gi];
    const uint32_t kind = out.x & 0xFFu;
    // the child's starting state
    Rec s0, par;
    uint32_t j = 0, axis = 0;
    float o = 0.0f;
    bool from_desc = false, self_copy = false;
    if (g.level == 0u) {
        const pga_axiom& ax = a.axioms[a.first_axiom + i];
#pragma unroll
        for (int k = 0; k < 12; ++k) s0.m[k] = ax.model[k];
        s0.s[0] = ax.size[0]; s0.s[1] = ax.size[1]; s0.s[2] = ax.size[2];
        s0.seed = ax.seed;
        s0.custom = ax.custom;
        s0.axiom = a.first_axiom + i;
        s0.path = 0u;
        s0.ed = (uint32_t)max(ax.entry_index, 0);
    } else {
        const uint4 dv = *reinterpret_cast<const uint4*>(a.tape_desc + g.desc_in_base + i);
        const uint32_t jj = dv.y & 0x3FFFFFFFu;
        if (kind == T_DROP || dv.x >= a.cap_par || jj == kDescNone) {
#pragma unroll
            for (uint32_t q = 0; q < kSlotFloats; ++q) my[q] = 0.0f;
            return;
        }
        const Rec* pp = a.tape_par + g.par_in_base + dv.x;
        par = *pp;
        if (jj == kDescSelf) {
            s0 = par;
            self_copy = true;
        } else {
            j = jj;
            axis = dv.y >> 30;
            o = __uint_as_float(dv.z);
            from_desc = true;
            build_child(c, par, load_entry(c, rec_entry(par)), j, axis, o, __uint_as_float(dv.w), s0);
        }
    }
    Adj d;
    adj_zero(d);
    GradAcc ga;
    grad_zero(ga);
    if (kind != T_DROP) {
        // replay the chain of single-child operators to its end
        Rec r = s0;
        uint32_t steps = 0;
        for (; steps < c.max_chain; ++steps) {
            const EntryV e = load_entry(c, rec_entry(r));
            if (!(e.kind() != K_GENERATE && e.kind() != K_DISCARD && !is_fanout(e.kind()))) break;
            uint32_t ord;
            const int32_t next = apply_single(c, e, r, &ord);
            if (!advance(c, r, next, ord)) break;
        }
        // the adjoint of the chain's end
        if (kind == T_TERM) {
            const uint32_t slot = (out.x >> 8) & 0xFFu;
            const uint32_t t = slot < c.prog.nslots ? c.prog.slot_term[slot] : PGA_MAX_TERMINALS;
            if (t < PGA_MAX_TERMINALS && g.dl[t] != nullptr && (unsigned long long)out.y < a.inst_cap[t])
                emit_adj(r, g.dl[t] + (size_t)out.y * 16u, d);
        } else {   // T_FAN / T_SELF: the children of the fan-out are in the next level
            const uint32_t first = g.next_child_base + out.z;
            const uint32_t nd = first <= g.slots_cap && out.w <= g.slots_cap - first ? out.w : 0u;
            fan_adj(c, g, ga, load_entry(c, rec_entry(r)), r, first, nd, kind == T_SELF, d);
        }
        // back through the chain, recomputing each operator's input state
        for (uint32_t t = steps; t-- > 0;) {
            Rec q = s0;
            for (uint32_t u = 0; u < t; ++u) {
                const EntryV e = load_entry(c, rec_entry(q));
                uint32_t ord;
                const int32_t next = apply_single(c, e, q, &ord);
                advance(c, q, next, ord);
            }
            single_adj(c, g, ga, load_entry(c, rec_entry(q)), q, d);
        }
    }
    flush_grads(c, g, ga);
    if (g.level == 0u) return;
    Adj dp;
    float d_o = 0.0f, d_ns = 0.0f;
    if (self_copy) dp = d;
    else if (from_desc) build_adj(par, load_entry(c, rec_entry(par)), j, axis, o, d, dp, d_o, d_ns);
    else adj_zero(dp);
#pragma unroll
    for (int k = 0; k < 12; ++k) my[k] = dp.m[k];
    my[12] = dp.s[0]; my[13] = dp.s[1]; my[14] = dp.s[2];
    my[15] = d_o;
    my[16] = d_ns;
    my[17] = my[18] = my[19] = 0.0f;
}

} // namespace pga

namespace pga {
__global__ void __launch_bounds__(256) k_grad_out(const double* acc, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)acc[i];
}
} // namespace pga

hipError_t pga_launch_grad_out(const double* acc, float* out, uint32_t n, hipStream_t s) {
    if (n == 0u) return hipSuccess;
    hipLaunchKernelGGL(pga::k_grad_out, dim3((n + 255u) / 256u), dim3(256), 0, s, acc, out, n);
    return hipGetLastError();
}

hipError_t pga_launch_adjoint(const pga::KArgs& a, const pga::AdjArgs& g, hipStream_t s) {
    if (g.n_children == 0u) return hipSuccess;
    hipLaunchKernelGGL(pga::k_adjoint, dim3((g.n_children + 255u) / 256u), dim3(256), 0, s, a, g);
    return hipGetLastError();
}
