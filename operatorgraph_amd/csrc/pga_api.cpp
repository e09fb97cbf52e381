// pga_api.cpp - the C ABI (include/pga.h) over the gfx950 derivation kernels.
//
// Owns the device copy of the program, the frontier buffers, the counters and
// the per-context stream. Validation that the reference performs with device
// traps or host exceptions happens here on load, or returns a status after
// execution from the device error word.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "pga_kernels.h"
#include "pga_spec_host.h"

using pga::DevEntry;
using pga::DevParam;

// One derivation recorded into a HIP graph: the argument block and launch
// shape it was captured with (a replay is valid while the context would
// build the same ones).
struct CapturedLaunch {
    pga::KArgs args;
    uint32_t blocks, block_size, first, count;
    const void* kernel;   // the specialised kernel it ran, or NULL
};

struct pga_context {
    pga_config cfg{};
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evk0 = nullptr, evk1 = nullptr;   // around the derivation kernel only
    std::string error;
    uint32_t pending_count = 0;   // axioms of the last asynchronous launch

    // program
    bool initialized = false;
    std::vector<DevEntry> h_entries;
    std::vector<DevParam> h_params;
    std::vector<int32_t> h_succ;
    std::vector<uint32_t> h_exp;
    uint8_t* d_image = nullptr;       // entries | params | succ | exp | slots (16-byte sections)
    uint32_t image_bytes = 0, off_params = 0, off_succ = 0, off_exp = 0, off_slots = 0;
    uint32_t lds_bytes = 0;           // dynamic LDS of the derivation kernels
    uint32_t count_geom = 0;          // 1: some decision of the program reads the model matrix
    uint32_t use_seed = 0;            // 1: some parameter reads a shape seed
    uint32_t nphases = 1;             // 1 + the largest phase a successor names
    bool collision = false;           // the program holds Collider or IfCollides entries
    bool polygons = false;            // the program holds dynamic polygon or prism entries
    DevEntry* d_entries = nullptr;    // views into d_image
    DevParam* d_params = nullptr;
    int32_t* d_succ = nullptr;
    uint32_t* d_exp = nullptr;
    uint8_t* d_slots = nullptr;       // [0,64): terminal -> slot, [64,128): slot -> terminal
    uint32_t nslots = 0;
    int32_t max_param_slot = -1;

    // axioms and parameters
    const pga_axiom* axioms = nullptr;
    pga_axiom* d_axioms_owned = nullptr;
    size_t axioms_owned_cap = 0;
    uint32_t n_axioms = 0;
    float* d_aparams = nullptr;
    size_t aparams_cap = 0;
    uint32_t aparams_per = 0, aparams_axioms = 0;
    uint32_t key[2] = {0, 0};
    float seeds[PGA_MAX_NUM_AXIOM_SEEDS] = {0};

    // terminal buffers
    pga_instance* inst_ptr[PGA_MAX_TERMINALS] = {nullptr};
    unsigned long long inst_cap[PGA_MAX_TERMINALS] = {0};
    bool bind_dirty = true;
    pga_instance** d_inst_ptr = nullptr;
    unsigned long long* d_inst_cap = nullptr;

    // work buffers
    pga::Rec* d_par[2] = {nullptr, nullptr};
    pga::Desc* d_desc[2] = {nullptr, nullptr};
    uint32_t cap = 0;
    unsigned long long* d_term_counts = nullptr;
    uint32_t* d_par_counts = nullptr;
    uint32_t* d_desc_counts = nullptr;
    uint32_t* d_err = nullptr;
    uint32_t* d_bar = nullptr;
    uint32_t* d_block_totals = nullptr;
    uint64_t* ticks = nullptr;      // caller-owned timestamp buffer
    uint32_t ticks_cap = 0;
    uint32_t* h_pinned = nullptr;   // count readback
    uint32_t mega_blocks = 0;
    uint64_t max_level_parents = 0, max_level_children = 0;   // of the last execute

    // collider store of Collider / IfCollides (80 bytes per collider) and its
    // counters: [0] stored, [1] visible to the running phase
    float4* d_coll = nullptr;
    uint32_t coll_cap = 0;
    uint32_t coll_cap_req = 0;        // pga_set_collider_capacity (0: kDefaultColliders)
    uint32_t* d_coll_count = nullptr;
    uint32_t colliders = 0;           // stored by the last execute
    // LBVH over the colliders, built between phases (multi-phase programs):
    // leaf boxes, keys, children, parents and node boxes for coll_cap leaves
    pga_lbvh* coll_bvh = nullptr;
    uint32_t coll_bvh_cap = 0;
    float* d_coll_boxes = nullptr;
    uint64_t* d_coll_keys = nullptr;
    int32_t* d_coll_tree = nullptr;   // left | right | parent | leaf parent, coll_bvh_cap each
    float* d_coll_nodes = nullptr;
    float4* d_coll_flat = nullptr;    // the tree in depth-first order, 2 coll_bvh_cap - 1 nodes
    uint32_t* d_coll_bits = nullptr;  // per child of a level: IfCollides outcomes of the count pass
    uint32_t coll_bits_n = 0;

    // derivation tape (reverse mode)
    pga::Rec* d_tape_par = nullptr;
    pga::Desc* d_tape_desc = nullptr;
    uint4* d_tape_out = nullptr;
    float* d_adj_slots = nullptr;     // per child: the adjoint it leaves for its parent
    const float** d_dl = nullptr;     // [PGA_MAX_TERMINALS] device table of dL/dinstance buffers
    double* d_grad64 = nullptr;       // FP64 accumulators of dL/dparams
    size_t grad64_cap = 0;
    uint32_t tape_par_cap = 0, tape_child_cap = 0, tape_blocks = 0;
    bool tape_valid = false;
    uint64_t tape_gen = 0;            // grows with every recorded tape
    uint32_t tape_axioms = 0;
    std::vector<uint32_t> tape_pc, tape_dc;   // parents / descriptors written by each level

    // graph capture
    bool capturing = false;
    bool capture_failed = false;
    std::string capture_error;
    std::vector<CapturedLaunch> captured;
    std::vector<pga_graph*> graphs;   // live graphs captured on this context

    // kernel specialised to the program (pga_specialize)
    hipModule_t spec_module = nullptr;
    hipFunction_t spec_fn = nullptr;
    // workgroup-local frontiers of the specialised kernel: per workgroup of
    // the full grid two arrays of loc_kp parents and two of loc_kd descriptors
    void* d_loc = nullptr;
    size_t loc_bytes = 0;
    uint32_t loc_kp = 0, loc_kd = 0, loc_blocks = 0;
    bool sched_local = true;       // pga_set_schedule
    uint32_t sched_grab = 0, sched_spill = 0;
    uint32_t spec_bs = 0, spec_blocks = 0;
    bool spec_on = false;

    pga_stats stats{};
};

struct pga_graph {
    pga_context* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<CapturedLaunch> launches;
    uint32_t cap = 0;
    uint32_t last_count = 0;
    uint32_t kernel_nodes = 0, cooperative_nodes = 0;
    ~pga_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

const uint32_t kDefaultCapacity = 1u << 22;
const uint32_t kDefaultColliders = 1u << 20;   // 80 MB of collider store
const uint32_t kDefaultMaxLevels = 4096;
const uint32_t kMaxChain = 256;
const uint32_t kMaxProgramLds = 60u * 1024u;   // program image + terminal tables
const uint32_t kMaxBlocksPerCu = 8;
// workgroup-local frontiers of the specialised kernel, per thread of the
// workgroup: parent records and descriptors a level may write before it
// spills to the grid's arrays; most axioms handed to a workgroup at once
const uint32_t kLocalParentsPerThread = 2, kLocalDescsPerThread = 8, kLocalGrabMax = 64;
const uint32_t kLocalParentsMin = 512, kLocalDescsMin = 2048;
const uint32_t kDefaultBlockSize = 512;        // one workgroup of 8 waves per compute unit

pga_status fail(pga_context* ctx, pga_status st, const std::string& msg) {
    if (ctx) ctx->error = msg;
    return st;
}

pga_status hip_fail(pga_context* ctx, hipError_t e, const char* what) {
    return fail(ctx, PGA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

// Calls that wait for the device or copy synchronously cannot be recorded.
pga_status refuse_in_capture(pga_context* ctx, const char* what) {
    return fail(ctx, PGA_ERR_INVALID_ARG, std::string(what) + " cannot run while a graph capture is open");
}
#define NO_CAPTURE(ctx, what) \
    do { if ((ctx) && (ctx)->capturing) return refuse_in_capture(ctx, what); } while (0)

template <class T> void hfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

bool is_single(uint8_t k) {
    return k == pga::K_IFSIZELESS || k == pga::K_IF || k == pga::K_RANDOM || k == pga::K_TRANSLATE ||
           k == pga::K_ROTATE || k == pga::K_SCALE || k == pga::K_SWAPSIZE || k == pga::K_EXTRUDE ||
           k == pga::K_COLLIDER || k == pga::K_IFCOLLIDES || k == pga::K_SETAS;
}
// shapes with a vertex list: dynamic polygons and right prisms
bool has_vertices(int32_t st) {
    return st == PGA_SHAPE_DYNAMIC_CONVEX_POLYGON || st == PGA_SHAPE_DYNAMIC_POLYGON ||
           st == PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM || st == PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
}
bool is_prism(int32_t st) {
    return st == PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM || st == PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
}
// single-child operators that choose among their successors
bool is_branch(uint8_t k) {
    return k == pga::K_IFSIZELESS || k == pga::K_IF || k == pga::K_RANDOM || k == pga::K_IFCOLLIDES;
}
bool is_fan(uint8_t k) {
    return k == pga::K_SUBDIV || k == pga::K_REPEAT || k == pga::K_COMPSPLIT || k == pga::K_REPLICATE ||
           k == pga::K_PRISMSPLIT;
}

// Checks an expression parameter (Parameters.cuh:563-664): length, operand
// tags, operation codes and attribute indices in range, a stack that never
// exceeds four entries nor underflows. Writes the words to upload to `out`
// (length word first): operation codes and attribute indices are stored
// rounded, as the reference reads them (round(x), Parameters.cuh:615 and
// 632), because generated tables print the packed words with six digits and
// carry values such as 12.99997 for attribute 13.
bool check_exp(const pga_dispatch_parameter& p, std::string& why, std::vector<uint32_t>& out) {
    float lenf = p.value[0];
    if (!(lenf >= 1.0f) || lenf != std::floor(lenf) || lenf > (float)(PGA_MAX_NUM_PARAMETER_VALUES - 1)) {
        why = "bad expression length";
        return false;
    }
    uint32_t len = (uint32_t)lenf;
    uint32_t word;
    std::memcpy(&word, &p.value[0], 4);
    out.push_back(word);
    int depth = 0;
    for (uint32_t i = 1; i <= len; ++i) {
        uint32_t raw;
        std::memcpy(&raw, &p.value[i], 4);
        uint32_t tag = raw >> 30;
        uint32_t xb = raw << 2;
        float x;
        std::memcpy(&x, &xb, 4);
        if (tag == PGA_ORT_OP || tag == PGA_ORT_SHAPE_ATTR) {
            const float r = std::round(x);
            const float hi = tag == PGA_ORT_OP ? 11.0f : 16.0f;
            if (!(r >= 0.0f && r <= hi)) {
                why = tag == PGA_ORT_OP ? "bad expression operation" : "bad expression shape attribute";
                return false;
            }
            uint32_t rb;
            std::memcpy(&rb, &r, 4);
            raw = (tag << 30) | (rb >> 2);   // integral values keep their two lowest mantissa bits at zero
        }
        if (tag == PGA_ORT_OP) {
            if (depth < 2) { why = "expression stack underflow"; return false; }
            depth -= 1;
        } else if (++depth > PGA_MAX_EXP_STACK) {
            why = "expression stack overflow";
            return false;
        }
        out.push_back(raw);
    }
    if (depth < 1) { why = "empty expression"; return false; }
    return true;
}

pga_status ensure_colliders(pga_context* ctx);

// Validates and lowers a program into the context's tables; `upload` = false
// stops before any device call (the tables stay on the host).
pga_status build_program(pga_context* ctx, const pga_procedure* procs, uint32_t nprocs,
                         const pga_dispatch_entry* entries, uint32_t n, bool upload = true) {
    if (!procs || !entries || nprocs == 0 || n == 0) return fail(ctx, PGA_ERR_INVALID_ARG, "empty program");
    // the entry index shares a word of the shape record with a flag (bit 15)
    if (n > 0x7FFFu) return fail(ctx, PGA_ERR_BAD_PROGRAM, "more than 32767 dispatch entries");
    std::vector<DevEntry> E(n);
    std::vector<DevParam> Pm;
    std::vector<int32_t> Sc;
    std::vector<uint32_t> Ex;
    int32_t max_slot = -1;
    bool count_geom = false;
    bool use_seed = false;
    bool collision = false;
    bool polygons = false;
    uint32_t nphases = 1;
    std::vector<uint8_t> Sp;   // phase of each successor in Sc
    for (uint32_t i = 0; i < n; ++i) {
        const pga_dispatch_entry& e = entries[i];
        std::string at = "entry " + std::to_string(i) + ": ";
        if (e.operator_code < 0 || (uint32_t)e.operator_code >= nprocs)
            return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "operator code out of range");
        if (e.num_parameters > PGA_MAX_NUM_PARAMETERS || e.num_successors > PGA_MAX_NUM_SUCCESSORS)
            return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "too many parameters or successors");
        const pga_procedure& p = procs[e.operator_code];
        DevEntry d{};
        uint32_t need_p = 0, need_s = 0;
        switch (p.op_type) {
        case PGA_OP_TRANSLATE: d.kind = pga::K_TRANSLATE; need_p = 3; need_s = 1; break;
        case PGA_OP_ROTATE: d.kind = pga::K_ROTATE; need_p = 3; need_s = 1; break;
        case PGA_OP_SCALE: d.kind = pga::K_SCALE; need_p = 3; need_s = 1; break;
        case PGA_OP_SWAPSIZE: d.kind = pga::K_SWAPSIZE; need_p = 3; need_s = 1; break;
        case PGA_OP_EXTRUDE:
            if (p.shape_type != PGA_SHAPE_QUAD && p.shape_type != PGA_SHAPE_DYNAMIC_CONVEX_POLYGON &&
                p.shape_type != PGA_SHAPE_DYNAMIC_POLYGON)
                return fail(ctx, PGA_ERR_UNSUPPORTED, at + "Extrude needs a planar shape (Quad or dynamic polygon)");
            d.kind = pga::K_EXTRUDE; need_p = 2; need_s = 1; break;
        case PGA_OP_COMPSPLIT:
            if (is_prism(p.shape_type)) {
                if (p.flags & PGA_PROC_PARALLEL)
                    return fail(ctx, PGA_ERR_UNSUPPORTED, at + "the parallel ComponentSplit splits Boxes only");
                d.kind = pga::K_PRISMSPLIT; need_p = 0; need_s = 3; break;
            }
            if (p.shape_type != PGA_SHAPE_BOX)
                return fail(ctx, PGA_ERR_UNSUPPORTED, at + "ComponentSplit needs a Box or a dynamic right prism");
            d.kind = pga::K_COMPSPLIT; need_p = 0; need_s = 3; break;
        case PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON:
        case PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM:
        case PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON:
        case PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM:
            // SetAs<DynParams> (SetAs.cuh:112-150): one vertex per parameter;
            // above PGA_MAX_NUM_SIDES the reference writes past its vertex array
            if (e.num_parameters > PGA_MAX_NUM_SIDES)
                return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "SetAs takes at most 16 vertices");
            d.kind = pga::K_SETAS; need_p = 3; need_s = 1; break;
        case PGA_OP_SUBDIV:
            d.kind = pga::K_SUBDIV; need_p = 1;
            need_s = e.num_parameters > 0 ? e.num_parameters - 1 : 0;
            break;
        case PGA_OP_REPEAT:
            if (p.flags & PGA_PROC_PARALLEL)   // Repeat<true> traps in the reference (Repeat.cuh:35-61)
                return fail(ctx, PGA_ERR_UNSUPPORTED, at + "the parallel form of Repeat is not implemented");
            d.kind = pga::K_REPEAT; need_p = 3; need_s = 1; break;
        case PGA_OP_DISCARD: d.kind = pga::K_DISCARD; break;
        case PGA_OP_IF: d.kind = pga::K_IF; need_p = 1; need_s = 2; break;
        case PGA_OP_IFSIZELESS: d.kind = pga::K_IFSIZELESS; need_p = 2; need_s = 2; break;
        case PGA_OP_STOCHASTIC:
            d.kind = pga::K_RANDOM; need_p = 2;
            need_s = e.num_parameters > 0 ? e.num_parameters - 1 : 0;
            break;
        case PGA_OP_REPLICATE: d.kind = pga::K_REPLICATE; break;
        case PGA_OP_COLLIDER:
        case PGA_OP_IFCOLLIDES:
            if (has_vertices(p.shape_type))
                return fail(ctx, PGA_ERR_UNSUPPORTED, at + "collision tests of Box and Quad shapes only");
            if (p.op_type == PGA_OP_COLLIDER) {
                d.kind = pga::K_COLLIDER; need_p = 1; need_s = 1; collision = true;
            } else {   // decides on the model matrix in the count pass too
                d.kind = pga::K_IFCOLLIDES; need_p = 1; need_s = 2; collision = true; count_geom = true;
            }
            break;
        case PGA_OP_GENERATE:
            // generation function 0 (triangle soup) is derived into instance
            // records as well; pga_tessellate turns them into its geometry
            if (p.gen_func_idx != 0 && p.gen_func_idx != 1)
                return fail(ctx, PGA_ERR_UNSUPPORTED, at + "only generation functions 0 and 1 are supported");
            if (p.gen_func_idx == 0 && has_vertices(p.shape_type))
                return fail(ctx, PGA_ERR_UNSUPPORTED, at + "generation function 0 of Box and Quad shapes only");
            d.kind = pga::K_GENERATE; need_p = 1; break;
        default:
            return fail(ctx, PGA_ERR_UNSUPPORTED,
                        at + "unsupported operator type " + std::to_string(p.op_type));
        }
        if (p.shape_type != PGA_SHAPE_BOX && p.shape_type != PGA_SHAPE_QUAD && !has_vertices(p.shape_type))
            return fail(ctx, PGA_ERR_UNSUPPORTED, at + "only Box, Quad and dynamic polygon and prism shapes are supported");
        if (e.num_parameters < need_p) return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "too few parameters");
        if (e.num_successors < need_s) return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "too few successors");
        d.shape = (uint8_t)p.shape_type;
        polygons |= has_vertices(p.shape_type);
        d.flags = (p.flags & PGA_PROC_STATIC) ? pga::EF_STATIC : 0;
        if (d.kind == pga::K_COMPSPLIT && (p.flags & PGA_PROC_PARALLEL)) d.flags |= pga::EF_PARALLEL;
        d.nparams = (uint8_t)e.num_parameters;
        d.nsucc = (uint8_t)e.num_successors;
        d.pbase = (uint32_t)Pm.size();
        d.sbase = (uint32_t)Sc.size();
        for (uint32_t k = 0; k < e.num_parameters; ++k) {
            const pga_dispatch_parameter& q = e.parameters[k];
            DevParam dp{};
            if (d.kind == pga::K_SETAS) {
                // a vertex, dynToFloat2 (Parameters.cuh:714-742): a scalar is
                // (value, 0); Vec2, Vec4 and Rand give their first two values
                dp.type = PGA_PT_VEC2;
                if (q.type == PGA_PT_SCALAR) { dp.a = q.value[0]; dp.b = 0.0f; }
                else if (q.type == PGA_PT_VEC2 || q.type == PGA_PT_VEC4 || q.type == PGA_PT_RAND) {
                    dp.a = q.value[0]; dp.b = q.value[1];
                } else {
                    return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "SetAs vertex parameters are scalars, vectors or Rand pairs");
                }
                Pm.push_back(dp);
                continue;
            }
            dp.type = (uint32_t)q.type;
            switch (q.type) {
            case PGA_PT_SCALAR: dp.a = q.value[0]; break;
            case PGA_PT_RAND: dp.a = q.value[0]; dp.b = q.value[1]; use_seed = true; break;
            case PGA_PT_SHAPE_ATTR:
                if (!(q.value[0] >= 0.0f) || (uint32_t)q.value[0] > 16u)
                    return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "shape attribute index out of range");
                dp.a = q.value[0];
                if ((uint32_t)q.value[0] < 12u) count_geom = true;   // a model-matrix element
                if ((uint32_t)q.value[0] == 15u) use_seed = true;     // the seed
                break;
            case PGA_PT_EXP: {
                std::string why;
                uint32_t off = (uint32_t)Ex.size();
                if (!check_exp(q, why, Ex)) return fail(ctx, PGA_ERR_BAD_PROGRAM, at + why);
                count_geom = true;   // conservatively: an expression may read the model matrix
                use_seed = true;     // ... and the seed
                std::memcpy(&dp.a, &off, 4);
                break;
            }
            case PGA_PT_PHILOX:
                if (!(q.value[2] >= 0.0f)) return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "bad Philox stream");
                dp.a = q.value[0]; dp.b = q.value[1]; dp.c = q.value[2];
                break;
            case PGA_PT_AXIOM_PARAM:
                if (!(q.value[0] >= 0.0f) || q.value[0] > 65535.0f)
                    return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "bad axiom parameter slot");
                dp.a = q.value[0];
                max_slot = std::max(max_slot, (int32_t)(uint32_t)q.value[0]);
                break;
            default:
                return fail(ctx, PGA_ERR_UNSUPPORTED, at + "unsupported parameter type " + std::to_string(q.type));
            }
            Pm.push_back(dp);
        }
        for (uint32_t k = 0; k < e.num_successors; ++k) {
            int32_t s = e.successors[k].entry_index;
            if (s < -1 || s >= (int32_t)n) return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "successor out of range");
            // PGA_PHASE_INHERIT: the parent's phase (bit 4); any other
            // negative (unassigned) phase index is phase 0
            const int32_t ph = e.successors[k].phase_index;
            if (ph >= (int32_t)PGA_MAX_PHASES) return fail(ctx, PGA_ERR_BAD_PROGRAM, at + "phase index out of range");
            Sc.push_back(s);
            Sp.push_back(ph == PGA_PHASE_INHERIT ? (uint8_t)16 : ph > 0 ? (uint8_t)ph : (uint8_t)0);
            if (ph > 0) nphases = std::max(nphases, (uint32_t)ph + 1u);
        }
        E[i] = d;
    }
    // successor shape consistency: the queue a child lands in holds its shape
    for (uint32_t i = 0; i < n; ++i) {
        const DevEntry& d = E[i];
        uint8_t child_shape = d.shape;
        if (d.kind == pga::K_COMPSPLIT) child_shape = PGA_SHAPE_QUAD;
        if (d.kind == pga::K_EXTRUDE)   // Operator::nextShapeType: the extruded type
            child_shape = d.shape == PGA_SHAPE_QUAD ? PGA_SHAPE_BOX
                        : d.shape == PGA_SHAPE_DYNAMIC_CONVEX_POLYGON ? PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM
                                                                      : PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
        if (d.kind == pga::K_SETAS) {
            const int32_t op = procs[entries[i].operator_code].op_type;
            child_shape = op == PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON ? PGA_SHAPE_DYNAMIC_CONVEX_POLYGON
                        : op == PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM ? PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM
                        : op == PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON ? PGA_SHAPE_DYNAMIC_POLYGON
                                                                       : PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
        }
        for (uint32_t k = 0; k < d.nsucc; ++k) {
            int32_t s = Sc[d.sbase + k];
            // a prism's caps (successors 0 and 1) are its polygons, its sides Quads
            if (d.kind == pga::K_PRISMSPLIT)
                child_shape = k == 2 ? (uint8_t)PGA_SHAPE_QUAD
                            : d.shape == PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM ? (uint8_t)PGA_SHAPE_DYNAMIC_CONVEX_POLYGON
                                                                               : (uint8_t)PGA_SHAPE_DYNAMIC_POLYGON;
            if (s >= 0 && E[s].shape != child_shape)
                return fail(ctx, PGA_ERR_BAD_PROGRAM,
                            "entry " + std::to_string(i) + ": successor " + std::to_string(s) + " expects another shape type");
        }
    }
    // leaf fan-outs: every successor reaches Generate/Discard through
    // single-child operators only (no cycle, bounded chain)
    std::vector<int8_t> memo(n, -1);   // -1 unknown, 0 no, 1 yes, 2 visiting
    std::function<bool(int32_t, uint32_t)> ends = [&](int32_t s, uint32_t len) -> bool {
        if (s < 0) return true;
        if (len > kMaxChain) return false;
        if (memo[s] == 0 || memo[s] == 1) return memo[s] == 1;
        if (memo[s] == 2) return false;
        const DevEntry& d = E[s];
        if (d.kind == pga::K_GENERATE || d.kind == pga::K_DISCARD) { memo[s] = 1; return true; }
        // an IfCollides below a fan-out keeps it off the leaf path: its BVH
        // walk is a chain of dependent loads, which the lanes of the next
        // level run side by side instead of one child after another
        if (!is_single(d.kind) || d.kind == pga::K_IFCOLLIDES) { memo[s] = 0; return false; }
        memo[s] = 2;
        bool ok = true;
        uint32_t ns = is_branch(d.kind) ? d.nsucc : 1u;
        for (uint32_t k = 0; k < ns && ok; ++k) ok = ends(Sc[d.sbase + k], len + 1);
        memo[s] = ok ? 1 : 0;
        return ok;
    };
    for (uint32_t i = 0; i < n; ++i) {
        if (!is_fan(E[i].kind) || E[i].kind == pga::K_PRISMSPLIT) continue;
        bool leaf = true;
        for (uint32_t k = 0; k < E[i].nsucc && leaf; ++k) leaf = ends(Sc[E[i].sbase + k], 1);
        if (leaf) E[i].flags |= pga::EF_LEAF;
    }
    // dense terminal slots: the distinct constant terminal indices of the
    // Generate entries; a computed terminal index needs all 64 slots
    uint8_t slots[2 * PGA_MAX_TERMINALS];
    std::memset(slots, 0xFF, sizeof(slots));
    bool dynamic_terminal = false;
    std::vector<uint32_t> used;
    for (uint32_t i = 0; i < n; ++i) {
        if (E[i].kind != pga::K_GENERATE) continue;
        const pga_dispatch_parameter& q = entries[i].parameters[0];
        if (q.type != PGA_PT_SCALAR) { dynamic_terminal = true; continue; }
        float v = q.value[0];
        uint32_t t = v > 0.0f ? (uint32_t)v : 0u;
        if (!(v < (float)PGA_MAX_TERMINALS))
            return fail(ctx, PGA_ERR_BAD_PROGRAM, "entry " + std::to_string(i) + ": terminal index out of range");
        if (std::find(used.begin(), used.end(), t) == used.end()) used.push_back(t);
    }
    if (dynamic_terminal) {
        used.clear();
        for (uint32_t t = 0; t < PGA_MAX_TERMINALS; ++t) used.push_back(t);
    }
    std::sort(used.begin(), used.end());
    for (uint32_t s = 0; s < used.size(); ++s) {
        slots[used[s]] = (uint8_t)s;
        slots[PGA_MAX_TERMINALS + s] = (uint8_t)used[s];
    }
    // inline constants, successors and terminal slots of each entry
    for (uint32_t i = 0; i < n; ++i) {
        DevEntry& d = E[i];
        d.inl = 0;
        d.tslot = 0xFF;
        for (uint32_t k = 0; k < 3; ++k) {
            d.pv[k] = 0.0f;
            d.sv[k] = k < d.nsucc ? Sc[d.sbase + k] : -1;
            if (k < d.nparams && Pm[d.pbase + k].type == PGA_PT_SCALAR) {
                d.inl |= (uint8_t)(1u << k);
                d.pv[k] = Pm[d.pbase + k].a;
            }
        }
        if (d.kind == pga::K_GENERATE && (d.inl & 1u)) {
            float v = d.pv[0];
            uint32_t t = v > 0.0f ? (uint32_t)v : 0u;
            d.tslot = slots[t];
        }
    }
    // dynamic polygons and prisms: the vertex list an entry's shapes carry is
    // a fact of the table. SetAs gives its successor the list of its
    // parameters; every other operator passes its own on to successors whose
    // shapes have one (Extrude of a polygon, a prism's caps), and a list
    // never changes on the way (operators move and size the shape only). An
    // entry reached with two different lists is refused. A ComponentSplit of
    // a prism keeps its list's first DevParam (bits) and vertex count in
    // pv[0] and pv[1].
    {
        std::vector<int64_t> poly(n, -1);   // first DevParam | count << 32
        auto same = [&](int64_t a, int64_t b) {
            const uint32_t na = (uint32_t)(a >> 32), nb = (uint32_t)(b >> 32);
            if (na != nb) return false;
            for (uint32_t v = 0; v < na; ++v) {
                const DevParam& x = Pm[(uint32_t)a + v];
                const DevParam& y = Pm[(uint32_t)b + v];
                if (std::memcmp(&x.a, &y.a, 8) != 0) return false;   // bitwise: -0 is not 0
            }
            return true;
        };
        std::vector<uint32_t> work;
        std::string clash;
        auto assign = [&](int32_t s, int64_t list) {
            if (s < 0 || !has_vertices(E[s].shape)) return;
            if (poly[s] < 0) { poly[s] = list; work.push_back((uint32_t)s); }
            else if (!same(poly[s], list) && clash.empty())
                clash = "entry " + std::to_string(s) + " is reached by two different vertex lists";
        };
        for (uint32_t i = 0; i < n; ++i)
            if (E[i].kind == pga::K_SETAS) assign(Sc[E[i].sbase], (int64_t)E[i].pbase | ((int64_t)E[i].nparams << 32));
        while (!work.empty()) {
            const uint32_t i = work.back();
            work.pop_back();
            if (E[i].kind == pga::K_SETAS) continue;   // its successor takes the SetAs list
            for (uint32_t k = 0; k < E[i].nsucc; ++k) assign(Sc[E[i].sbase + k], poly[i]);
        }
        if (!clash.empty()) return fail(ctx, PGA_ERR_UNSUPPORTED, clash);
        for (uint32_t i = 0; i < n; ++i) {
            if (E[i].kind != pga::K_PRISMSPLIT) continue;
            const uint32_t first = poly[i] < 0 ? 0u : (uint32_t)poly[i];
            std::memcpy(&E[i].pv[0], &first, 4);
            E[i].pv[1] = poly[i] < 0 ? 0.0f : (float)(uint32_t)(poly[i] >> 32);
        }
    }
    // direct leaf fan-outs: every child is a constant-terminal Generate or a
    // Discard, so a counting pass takes its terminal counts from a table
    for (uint32_t i = 0; i < n; ++i) {
        DevEntry& d = E[i];
        d.ndirect = 0;
        d.aux[0] = d.aux[1] = 0u;
        if (!is_fan(d.kind)) continue;
        // successor of each child, in child order
        std::vector<int32_t> succ_j;
        if (d.kind == pga::K_SUBDIV) {
            for (uint32_t j = 0; j + 1 < d.nparams; ++j) succ_j.push_back(Sc[d.sbase + j]);
        } else if (d.kind == pga::K_REPEAT) {
            succ_j.push_back(Sc[d.sbase]);
        } else if (d.kind == pga::K_COMPSPLIT) {
            for (uint32_t j = 0; j < 6; ++j) succ_j.push_back(Sc[d.sbase + (j < 4 ? 2u : (j == 4 ? 1u : 0u))]);
        } else if (d.kind == pga::K_PRISMSPLIT) {
            const uint32_t ns = (uint32_t)d.pv[1];
            for (uint32_t j = 0; j < ns + 2; ++j) succ_j.push_back(Sc[d.sbase + (j < ns ? 2u : (j == ns ? 1u : 0u))]);
        } else {
            for (uint32_t k = 0; k < d.nsucc; ++k) succ_j.push_back(Sc[d.sbase + k]);
        }
        for (uint32_t j = 0; j < succ_j.size() && j < 32; ++j)
            if (succ_j[j] < 0 || E[succ_j[j]].kind == pga::K_DISCARD) d.aux[0] |= 1u << j;
        if (!(d.flags & pga::EF_LEAF)) continue;
        std::vector<std::pair<int32_t, uint32_t>> kids;   // (successor, children that take it)
        bool ok = true;
        if (d.kind == pga::K_SUBDIV) {
            for (uint32_t j = 0; j + 1 < d.nparams; ++j) {
                const DevParam& q = Pm[d.pbase + 1 + j];
                // the static form skips zero factors: only constant non-zero ones have fixed counts
                if ((d.flags & pga::EF_STATIC) && !(q.type == PGA_PT_SCALAR && q.a != 0.0f)) ok = false;
            }
        }
        for (int32_t s : succ_j) kids.push_back({s, 1u});
        uint32_t cnt[PGA_MAX_TERMINALS] = {0};
        for (auto& kc : kids) {
            if (kc.first < 0 || E[kc.first].kind == pga::K_DISCARD) continue;
            if (E[kc.first].kind != pga::K_GENERATE || E[kc.first].tslot == 0xFF) { ok = false; break; }
            cnt[E[kc.first].tslot] += kc.second;
        }
        uint32_t nd = 0, table = 0;
        for (uint32_t t = 0; t < PGA_MAX_TERMINALS && ok; ++t) {
            if (cnt[t] == 0) continue;
            if (nd == 2 || cnt[t] > 255) { ok = false; break; }
            table |= (t | (cnt[t] << 8)) << (16u * nd);
            ++nd;
        }
        if (!ok) continue;
        d.aux[1] = table;
        d.ndirect = (uint8_t)nd;
        d.flags |= pga::EF_DIRECT;
    }
    // multi-phase programs: each successor carries its phase in bits 16-19
    // (bit 20: the parent's), and no fan-out counts its terminals from a
    // table (a child's phase decides whether a launch emits it). aux[1]
    // holds instead the latest phase a successor reachable from the entry
    // names: a shape of an earlier phase whose entry reaches no later one
    // adds nothing to a launch, which skips it.
    if (nphases > 1) {
        std::vector<uint32_t> reach(n, 0u);
        for (bool changed = true; changed;) {   // a fixed point: at most 16 rises per entry
            changed = false;
            for (uint32_t i = 0; i < n; ++i) {
                uint32_t r = reach[i];
                for (uint32_t k = 0; k < E[i].nsucc; ++k) {
                    const int32_t t = Sc[E[i].sbase + k];
                    if (t < 0) continue;
                    r = std::max(r, std::max(Sp[E[i].sbase + k] & 15u, reach[t]));
                }
                if (r != reach[i]) { reach[i] = r; changed = true; }
            }
        }
        for (size_t k = 0; k < Sc.size(); ++k)
            if (Sc[k] >= 0) Sc[k] |= (int32_t)Sp[k] << 16;
        for (uint32_t i = 0; i < n; ++i) {
            DevEntry& d = E[i];
            d.flags &= (uint8_t)~pga::EF_DIRECT;
            d.ndirect = 0;
            d.aux[1] = reach[i];
            for (uint32_t k = 0; k < 3; ++k) d.sv[k] = k < d.nsucc ? Sc[d.sbase + k] : -1;
        }
    }
    if (Sc.empty()) Sc.push_back(-1);
    if (Pm.empty()) Pm.push_back(DevParam{});
    if (Ex.empty()) Ex.push_back(0u);
    // one program image: entries | params | succ | exp | slots, 16-byte sections
    auto pad16 = [](size_t b) { return (b + 15u) & ~(size_t)15u; };
    const size_t off_params = pad16(E.size() * sizeof(DevEntry));
    const size_t off_succ = off_params + pad16(Pm.size() * sizeof(DevParam));
    const size_t off_exp = off_succ + pad16(Sc.size() * sizeof(int32_t));
    const size_t off_slots = off_exp + pad16(Ex.size() * sizeof(uint32_t));
    const size_t image_bytes = off_slots + sizeof(slots);
    std::vector<uint8_t> image(image_bytes, 0);
    std::memcpy(image.data(), E.data(), E.size() * sizeof(DevEntry));
    std::memcpy(image.data() + off_params, Pm.data(), Pm.size() * sizeof(DevParam));
    std::memcpy(image.data() + off_succ, Sc.data(), Sc.size() * sizeof(int32_t));
    std::memcpy(image.data() + off_exp, Ex.data(), Ex.size() * sizeof(uint32_t));
    std::memcpy(image.data() + off_slots, slots, sizeof(slots));
    if (!upload) {
        ctx->h_entries = E;
        ctx->h_params = Pm;
        ctx->h_succ = Sc;
        ctx->h_exp = Ex;
        ctx->count_geom = count_geom ? 1u : 0u;
        ctx->use_seed = use_seed ? 1u : 0u;
        ctx->max_param_slot = max_slot;
        ctx->nphases = nphases;
        ctx->collision = collision;
        ctx->polygons = polygons;
        return PGA_OK;
    }
    ctx->collision = collision;
    ctx->nphases = nphases;
    if (pga_status st = ensure_colliders(ctx)) return st;
    hfree(ctx->d_image);
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_image, image_bytes));
    HIP_TRY(ctx, hipMemcpy(ctx->d_image, image.data(), image_bytes, hipMemcpyHostToDevice));
    ctx->image_bytes = (uint32_t)image_bytes;
    ctx->off_params = (uint32_t)off_params;
    ctx->off_succ = (uint32_t)off_succ;
    ctx->off_exp = (uint32_t)off_exp;
    ctx->off_slots = (uint32_t)off_slots;
    ctx->d_entries = reinterpret_cast<DevEntry*>(ctx->d_image);
    ctx->d_params = reinterpret_cast<DevParam*>(ctx->d_image + off_params);
    ctx->d_succ = reinterpret_cast<int32_t*>(ctx->d_image + off_succ);
    ctx->d_exp = reinterpret_cast<uint32_t*>(ctx->d_image + off_exp);
    ctx->d_slots = ctx->d_image + off_slots;
    ctx->nslots = (uint32_t)used.size();
    // the image and the terminal tables live in LDS during a derivation
    const uint32_t lds = (uint32_t)image_bytes + 16u * PGA_MAX_TERMINALS;
    if (lds > kMaxProgramLds)
        return fail(ctx, PGA_ERR_UNSUPPORTED, "program image of " + std::to_string(image_bytes) +
                    " bytes exceeds the " + std::to_string(kMaxProgramLds - 16u * PGA_MAX_TERMINALS) +
                    " bytes the derivation kernels stage in LDS");
    ctx->lds_bytes = lds;
    ctx->count_geom = count_geom ? 1u : 0u;
    ctx->use_seed = use_seed ? 1u : 0u;
    int per_cu = 0;
    HIP_TRY(ctx, pga_mega_occupancy(ctx->cfg.block_size, ctx->lds_bytes, collision || nphases > 1, &per_cu));
    if (per_cu < 1) return fail(ctx, PGA_ERR_UNSUPPORTED, "derivation kernel does not fit on a compute unit");
    const uint32_t max_blocks = std::min((uint32_t)per_cu, kMaxBlocksPerCu) * (uint32_t)ctx->num_cus;
    ctx->mega_blocks = ctx->cfg.grid_blocks ? std::min(ctx->cfg.grid_blocks, max_blocks) : max_blocks;
    ctx->h_entries = E;
    ctx->h_params = Pm;
    ctx->h_succ = Sc;
    ctx->h_exp = Ex;
    ctx->max_param_slot = max_slot;
    ctx->nphases = nphases;
    ctx->collision = collision;
    ctx->polygons = polygons;
    ctx->initialized = true;
    return PGA_OK;
}

uint32_t queue_capacity(const pga_context* ctx) {
    return ctx->cfg.queue_capacity ? ctx->cfg.queue_capacity : kDefaultCapacity;
}

pga_status ensure_buffers(pga_context* ctx) {
    uint32_t cap = queue_capacity(ctx);
    if (ctx->cap != cap) {
        for (int b = 0; b < 2; ++b) {
            hfree(ctx->d_par[b]);
            hfree(ctx->d_desc[b]);
            HIP_TRY(ctx, hipMalloc((void**)&ctx->d_par[b], (size_t)cap * sizeof(pga::Rec)));
            HIP_TRY(ctx, hipMalloc((void**)&ctx->d_desc[b], (size_t)cap * sizeof(pga::Desc)));
        }
        ctx->cap = cap;
    }
    if (ctx->bind_dirty) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_inst_ptr, ctx->inst_ptr, sizeof(ctx->inst_ptr), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_inst_cap, ctx->inst_cap, sizeof(ctx->inst_cap), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        ctx->bind_dirty = false;
    }
    return PGA_OK;
}

// Axioms per hand-out of the workgroup-local levels: the scheduled number,
// or an equal share of the axioms per workgroup (one round: every round
// pays the derivation's depth in level latencies).
uint32_t local_grab(const pga_context* ctx, uint32_t count) {
    if (ctx->sched_grab) return std::min(ctx->sched_grab, kLocalGrabMax);
    const uint32_t blocks = std::max(ctx->spec_blocks, 1u);
    return std::max<uint32_t>(1u, std::min<uint32_t>(kLocalGrabMax, (count + blocks - 1u) / blocks));
}

void drop_collider_bvh(pga_context* ctx) {
    if (ctx->coll_bvh) pga_lbvh_destroy(ctx->coll_bvh);
    ctx->coll_bvh = nullptr;
    ctx->coll_bvh_cap = 0;
    hfree(ctx->d_coll_boxes);
    hfree(ctx->d_coll_keys);
    hfree(ctx->d_coll_tree);
    hfree(ctx->d_coll_nodes);
    hfree(ctx->d_coll_flat);
}

// The collider store of a program with Collider / IfCollides (coll_cap_req
// colliders, default kDefaultColliders) and, for more than one phase, the
// LBVH arrays over it; kept while large enough.
pga_status ensure_colliders(pga_context* ctx) {
    if (!ctx->collision) return PGA_OK;
    const uint32_t want = ctx->coll_cap_req ? ctx->coll_cap_req : kDefaultColliders;
    if (!ctx->d_coll || ctx->coll_cap != want) {
        if (ctx->stream) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        hfree(ctx->d_coll);
        ctx->coll_cap = 0;
        HIP_TRY(ctx, hipMalloc((void**)&ctx->d_coll, (size_t)want * 5u * sizeof(float4)));
        ctx->coll_cap = want;
    }
    if (!ctx->d_coll_count) HIP_TRY(ctx, hipMalloc((void**)&ctx->d_coll_count, 2u * sizeof(uint32_t)));
    const uint32_t bits = queue_capacity(ctx);
    if (!ctx->d_coll_bits || ctx->coll_bits_n != bits) {
        if (ctx->stream) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        hfree(ctx->d_coll_bits);
        ctx->coll_bits_n = 0;
        HIP_TRY(ctx, hipMalloc((void**)&ctx->d_coll_bits, (size_t)bits * sizeof(uint32_t)));
        ctx->coll_bits_n = bits;
    }
    if (ctx->nphases < 2 || want < 2u) return PGA_OK;
    if (ctx->coll_bvh && ctx->coll_bvh_cap == want) return PGA_OK;
    if (ctx->stream) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_collider_bvh(ctx);
    pga_lbvh* b = nullptr;
    const pga_status st = pga_lbvh_create(want, ctx->device, ctx->stream, &b);
    if (st) {
        const std::string why = b ? pga_lbvh_last_error(b) : "bad capacity";
        if (b) pga_lbvh_destroy(b);
        return fail(ctx, st, "collider BVH: " + why);
    }
    // the builder is kept only with all its arrays: a failed allocation
    // leaves no BVH, and IfCollides then scans
    hipError_t e = hipMalloc((void**)&ctx->d_coll_boxes, (size_t)want * 6u * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->d_coll_keys, (size_t)want * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->d_coll_tree, (size_t)want * 4u * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->d_coll_nodes, (size_t)want * 6u * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&ctx->d_coll_flat, (size_t)want * 2u * 2u * sizeof(float4));
    ctx->coll_bvh = b;
    if (e != hipSuccess) {
        drop_collider_bvh(ctx);
        return hip_fail(ctx, e, "collider BVH arrays");
    }
    ctx->coll_bvh_cap = want;
    return PGA_OK;
}

pga::KArgs make_args(pga_context* ctx, uint32_t first, uint32_t count) {
    pga::KArgs a{};
    a.prog.entries = ctx->d_entries;
    a.prog.params = ctx->d_params;
    a.prog.succ = ctx->d_succ;
    a.prog.exp = ctx->d_exp;
    a.prog.nentries = (uint32_t)ctx->h_entries.size();
    a.prog.term_slot = ctx->d_slots;
    a.prog.slot_term = ctx->d_slots + PGA_MAX_TERMINALS;
    a.prog.nslots = ctx->nslots;
    a.image = reinterpret_cast<const uint4*>(ctx->d_image);
    a.image_bytes = ctx->image_bytes;
    a.lds_bytes = ctx->lds_bytes;
    a.off_params = ctx->off_params;
    a.off_succ = ctx->off_succ;
    a.off_exp = ctx->off_exp;
    a.off_slots = ctx->off_slots;
    a.block_totals = ctx->d_block_totals;
    a.ticks = reinterpret_cast<unsigned long long*>(ctx->ticks);
    a.ticks_cap = ctx->ticks_cap;
    a.axioms = ctx->axioms;
    a.first_axiom = first;
    a.n_axioms = count;
    a.aparams = ctx->d_aparams;
    a.aparams_per = ctx->aparams_per;
    a.naxioms_total = ctx->n_axioms;
    a.key0 = ctx->key[0];
    a.key1 = ctx->key[1];
    a.inst = ctx->d_inst_ptr;
    a.inst_cap = ctx->d_inst_cap;
    a.term_counts = ctx->d_term_counts;
    a.par0 = ctx->d_par[0];
    a.par1 = ctx->d_par[1];
    a.desc0 = ctx->d_desc[0];
    a.desc1 = ctx->d_desc[1];
    a.cap_par = ctx->cap;
    a.cap_desc = ctx->cap;
    a.par_counts = ctx->d_par_counts;
    a.desc_counts = ctx->d_desc_counts;
    a.err = ctx->d_err;
    a.bar = ctx->d_bar;
    a.max_levels = ctx->cfg.max_levels;
    a.max_chain = kMaxChain;
    a.count_geom = ctx->count_geom;
    a.use_seed = ctx->use_seed;
    a.polygons = ctx->polygons ? 1u : 0u;
    a.dmask = ctx->nphases > 1 ? 0x0FFFu : 0xFFFFu;
    a.phase = 0;
    if (ctx->collision) {
        a.coll = ctx->d_coll;
        a.coll_count = ctx->d_coll_count;
        a.coll_visible = ctx->d_coll_count + 1;
        a.coll_cap = ctx->coll_cap;
        a.coll_bits = ctx->d_coll_bits;
        a.coll_bits_cap = ctx->coll_bits_n;
    }
    if (ctx->spec_on && ctx->sched_local && ctx->d_loc && ctx->spec_blocks <= ctx->loc_blocks) {
        a.loc_par = reinterpret_cast<pga::Rec*>(ctx->d_loc);
        a.loc_desc = reinterpret_cast<pga::Desc*>(reinterpret_cast<char*>(ctx->d_loc) +
                                                  (size_t)ctx->loc_blocks * 2u * ctx->loc_kp * sizeof(pga::Rec));
        a.loc_kp = ctx->loc_kp;
        a.loc_kd = ctx->loc_kd;
        a.loc_cap_p = std::min(ctx->loc_kp, ctx->cap);   // no level holds more than the queues would
        a.loc_cap_d = std::min(ctx->loc_kd, ctx->cap);
        if (ctx->sched_spill) a.loc_cap_d = std::min(a.loc_cap_d, ctx->sched_spill);
        a.loc_grab = local_grab(ctx, count);
    }
    return a;
}

// Workgroup sizes of the specialised kernel: whole waves, 64 to 1024.
bool spec_block_size_supported(uint32_t bs) { return bs >= 64u && bs <= 1024u && bs % 64u == 0u; }

// The specialised kernel's workgroup-local arrays for `blocks` workgroups of
// `bs` threads (kept when large enough).
pga_status ensure_local(pga_context* ctx, uint32_t blocks, uint32_t bs) {
    const uint32_t kp = std::max(kLocalParentsPerThread * bs, kLocalParentsMin);
    const uint32_t kd = std::max(kLocalDescsPerThread * bs, kLocalDescsMin);
    const size_t bytes = (size_t)blocks * 2u * ((size_t)kp * sizeof(pga::Rec) + (size_t)kd * sizeof(pga::Desc));
    if (ctx->d_loc && ctx->loc_kp == kp && ctx->loc_kd == kd && ctx->loc_blocks >= blocks) return PGA_OK;
    if (ctx->stream) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    hfree(ctx->d_loc);
    ctx->d_loc = nullptr;
    ctx->loc_bytes = 0;
    ctx->loc_blocks = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->d_loc, bytes));
    ctx->loc_bytes = bytes;
    ctx->loc_kp = kp;
    ctx->loc_kd = kd;
    ctx->loc_blocks = blocks;
    return PGA_OK;
}

void drop_spec(pga_context* ctx) {
    if (ctx->spec_module) {
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        (void)hipModuleUnload(ctx->spec_module);
    }
    ctx->spec_module = nullptr;
    ctx->spec_fn = nullptr;
    ctx->spec_bs = ctx->spec_blocks = 0;
    ctx->spec_on = false;
}

// The persistent derivation: the specialised kernel when it is on, else the
// interpreting kernel of the configured workgroup size.
hipError_t launch_derivation(pga_context* ctx, const pga::KArgs& a, const void** kernel) {
    if (ctx->spec_on) {
        pga::KArgs copy = a;
        void* params[] = {&copy};
        *kernel = (const void*)ctx->spec_fn;
        return hipModuleLaunchCooperativeKernel(ctx->spec_fn, ctx->spec_blocks, 1, 1, ctx->spec_bs, 1, 1, a.lds_bytes,
                                                ctx->stream, params);
    }
    *kernel = nullptr;
    return pga_launch_mega(a, ctx->mega_blocks, ctx->cfg.block_size, ctx->stream);
}

pga_status status_from_err(pga_context* ctx, uint32_t err) {
    if (!err) return PGA_OK;
    static const pga_status order[] = {PGA_ERR_BAD_PROGRAM, PGA_ERR_SUBDIV_OVERCONSTRAINED,
                                       PGA_ERR_REPEAT_BAD_EXTENT, PGA_ERR_MAX_DEPTH,
                                       PGA_ERR_OVERFLOW_QUEUE, PGA_ERR_OVERFLOW_INSTANCES};
    std::string msg = "device reported:";
    pga_status first = PGA_OK;
    for (pga_status s : order)
        if (err & (1u << s)) {
            msg += std::string(" ") + pga_status_string(s);
            if (first == PGA_OK) first = s;
        }
    if (first == PGA_OK) first = PGA_ERR_BAD_PROGRAM;
    ctx->error = msg;
    return first;
}

} // namespace

extern "C" {

PGA_API uint32_t pga_abi_version(void) { return 1u; }

PGA_API const char* pga_status_string(pga_status s) {
    switch (s) {
    case PGA_OK: return "OK";
    case PGA_ERR_INVALID_ARG: return "INVALID_ARG";
    case PGA_ERR_OVERFLOW_QUEUE: return "OVERFLOW_QUEUE";
    case PGA_ERR_OVERFLOW_INSTANCES: return "OVERFLOW_INSTANCES";
    case PGA_ERR_BAD_PROGRAM: return "BAD_PROGRAM";
    case PGA_ERR_SUBDIV_OVERCONSTRAINED: return "SUBDIV_OVERCONSTRAINED";
    case PGA_ERR_REPEAT_BAD_EXTENT: return "REPEAT_BAD_EXTENT";
    case PGA_ERR_HIP: return "HIP_ERROR";
    case PGA_ERR_MAX_DEPTH: return "MAX_DEPTH";
    case PGA_ERR_UNSUPPORTED: return "UNSUPPORTED";
    case PGA_ERR_NOT_INITIALIZED: return "NOT_INITIALIZED";
    }
    return "UNKNOWN";
}

PGA_API pga_status pga_create(const pga_config* config, pga_context** out) {
    if (!out) return PGA_ERR_INVALID_ARG;
    *out = nullptr;
    pga_context* ctx = new pga_context();
    if (config) ctx->cfg = *config;
    else ctx->cfg.technique = PGA_TECHNIQUE_MEGAKERNEL;
    if (ctx->cfg.technique == PGA_TECHNIQUE_DYN_PAR) {
        delete ctx;
        return PGA_ERR_UNSUPPORTED;
    }
    if (ctx->cfg.technique == PGA_TECHNIQUE_MEGAKERNEL_LOCAL_QUEUES) ctx->cfg.technique = PGA_TECHNIQUE_MEGAKERNEL;
    if (ctx->cfg.technique != PGA_TECHNIQUE_KERNELS && ctx->cfg.technique != PGA_TECHNIQUE_MEGAKERNEL) {
        delete ctx;
        return PGA_ERR_INVALID_ARG;
    }
    if (ctx->cfg.block_size == 0) ctx->cfg.block_size = kDefaultBlockSize;
    if (!pga_block_size_supported(ctx->cfg.block_size)) {
        delete ctx;
        return PGA_ERR_INVALID_ARG;
    }
    if (ctx->cfg.max_levels == 0) ctx->cfg.max_levels = kDefaultMaxLevels;
    *out = ctx;
    if (ctx->cfg.device >= 0) HIP_TRY(ctx, hipSetDevice(ctx->cfg.device));
    HIP_TRY(ctx, hipGetDevice(&ctx->device));
    HIP_TRY(ctx, hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    if (ctx->cfg.stream) {
        ctx->stream = (hipStream_t)ctx->cfg.stream;
    } else {
        HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->own_stream = true;
    }
    HIP_TRY(ctx, hipEventCreate(&ctx->ev0));
    HIP_TRY(ctx, hipEventCreate(&ctx->ev1));
    HIP_TRY(ctx, hipEventCreate(&ctx->evk0));
    HIP_TRY(ctx, hipEventCreate(&ctx->evk1));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_term_counts, PGA_MAX_TERMINALS * sizeof(unsigned long long)));
    HIP_TRY(ctx, hipMemset(ctx->d_term_counts, 0, PGA_MAX_TERMINALS * sizeof(unsigned long long)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_par_counts, (ctx->cfg.max_levels + 2) * sizeof(uint32_t)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_desc_counts, (ctx->cfg.max_levels + 2) * sizeof(uint32_t)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_err, sizeof(uint32_t)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_bar, pga::kBarWords * sizeof(uint32_t)));
    HIP_TRY(ctx, hipMemset(ctx->d_bar, 0, pga::kBarWords * sizeof(uint32_t)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_inst_ptr, PGA_MAX_TERMINALS * sizeof(pga_instance*)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_inst_cap, PGA_MAX_TERMINALS * sizeof(unsigned long long)));
    HIP_TRY(ctx, hipHostMalloc((void**)&ctx->h_pinned, 64 * sizeof(uint32_t), hipHostMallocDefault));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_block_totals,
                           (size_t)ctx->num_cus * kMaxBlocksPerCu * (PGA_MAX_TERMINALS + 2) * sizeof(uint32_t)));
    return PGA_OK;
}

PGA_API void pga_destroy(pga_context* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (pga_graph* g : ctx->graphs) g->ctx = nullptr;   // the graphs stay valid to destroy
    hfree(ctx->d_image);
    hfree(ctx->d_tape_par); hfree(ctx->d_tape_desc); hfree(ctx->d_tape_out); hfree(ctx->d_adj_slots);
    hfree(ctx->d_dl); hfree(ctx->d_grad64);
    hfree(ctx->d_axioms_owned); hfree(ctx->d_aparams);
    hfree(ctx->d_inst_ptr); hfree(ctx->d_inst_cap);
    for (int b = 0; b < 2; ++b) { hfree(ctx->d_par[b]); hfree(ctx->d_desc[b]); }
    hfree(ctx->d_term_counts); hfree(ctx->d_par_counts); hfree(ctx->d_desc_counts);
    hfree(ctx->d_err); hfree(ctx->d_bar); hfree(ctx->d_block_totals); hfree(ctx->d_loc);
    hfree(ctx->d_coll); hfree(ctx->d_coll_count); hfree(ctx->d_coll_bits);
    drop_collider_bvh(ctx);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->evk0) (void)hipEventDestroy(ctx->evk0);
    if (ctx->evk1) (void)hipEventDestroy(ctx->evk1);
    if (ctx->spec_module) (void)hipModuleUnload(ctx->spec_module);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

PGA_API const char* pga_last_error(const pga_context* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

PGA_API pga_status pga_initialize(pga_context* ctx, const pga_procedure* procedures, uint32_t num_procedures,
                                  const pga_dispatch_entry* entries, uint32_t num_entries) {
    NO_CAPTURE(ctx, "pga_initialize");
    if (!ctx) return PGA_ERR_INVALID_ARG;
    ctx->initialized = false;
    ctx->tape_valid = false;
    drop_spec(ctx);
    return build_program(ctx, procedures, num_procedures, entries, num_entries);
}

PGA_API pga_status pga_set_axioms(pga_context* ctx, const pga_axiom* axioms, uint32_t n) {
    NO_CAPTURE(ctx, "pga_set_axioms");
    if (!ctx || (!axioms && n)) return fail(ctx, PGA_ERR_INVALID_ARG, "null axioms");
    size_t bytes = (size_t)std::max<uint32_t>(n, 1) * sizeof(pga_axiom);
    if (bytes > ctx->axioms_owned_cap) {
        hfree(ctx->d_axioms_owned);
        HIP_TRY(ctx, hipMalloc((void**)&ctx->d_axioms_owned, bytes));
        ctx->axioms_owned_cap = bytes;
    }
    ctx->tape_valid = false;
    if (n) HIP_TRY(ctx, hipMemcpy(ctx->d_axioms_owned, axioms, (size_t)n * sizeof(pga_axiom), hipMemcpyHostToDevice));
    ctx->axioms = ctx->d_axioms_owned;
    ctx->n_axioms = n;
    return PGA_OK;
}

PGA_API pga_status pga_set_axioms_device(pga_context* ctx, const pga_axiom* device_axioms, uint32_t n) {
    NO_CAPTURE(ctx, "pga_set_axioms_device");
    if (!ctx || (!device_axioms && n)) return fail(ctx, PGA_ERR_INVALID_ARG, "null axioms");
    ctx->tape_valid = false;
    ctx->axioms = device_axioms;
    ctx->n_axioms = n;
    return PGA_OK;
}

PGA_API pga_status pga_set_seeds(pga_context* ctx, const float* seeds, uint32_t n) {
    NO_CAPTURE(ctx, "pga_set_seeds");
    if (!ctx || (!seeds && n)) return fail(ctx, PGA_ERR_INVALID_ARG, "null seeds");
    uint32_t k = n % (PGA_MAX_NUM_AXIOM_SEEDS + 1);      // GlobalVariables.cuh:28-32
    if (k) std::memcpy(ctx->seeds, seeds, k * sizeof(float));
    return PGA_OK;
}

PGA_API pga_status pga_get_seeds(const pga_context* ctx, float* seeds32) {
    if (!ctx || !seeds32) return PGA_ERR_INVALID_ARG;
    std::memcpy(seeds32, ctx->seeds, sizeof(ctx->seeds));
    return PGA_OK;
}

PGA_API pga_status pga_set_axiom_params(pga_context* ctx, const float* params, uint32_t num_axioms,
                                        uint32_t per_axiom) {
    NO_CAPTURE(ctx, "pga_set_axiom_params");
    if (!ctx || (!params && num_axioms && per_axiom)) return fail(ctx, PGA_ERR_INVALID_ARG, "null parameters");
    size_t bytes = (size_t)num_axioms * per_axiom * sizeof(float);
    ctx->tape_valid = false;
    if (bytes > ctx->aparams_cap) {
        hfree(ctx->d_aparams);
        HIP_TRY(ctx, hipMalloc((void**)&ctx->d_aparams, bytes));
        ctx->aparams_cap = bytes;
    }
    if (bytes) HIP_TRY(ctx, hipMemcpy(ctx->d_aparams, params, bytes, hipMemcpyHostToDevice));
    ctx->aparams_per = per_axiom;
    ctx->aparams_axioms = num_axioms;
    return PGA_OK;
}

PGA_API pga_status pga_set_philox_key(pga_context* ctx, uint32_t k0, uint32_t k1) {
    if (!ctx) return PGA_ERR_INVALID_ARG;
    if (k0 != ctx->key[0] || k1 != ctx->key[1]) ctx->tape_valid = false;   // the adjoint re-evaluates draws
    ctx->key[0] = k0;
    ctx->key[1] = k1;
    return PGA_OK;
}

PGA_API pga_status pga_bind_instances(pga_context* ctx, uint32_t terminal, pga_instance* buf, uint64_t capacity) {
    if (!ctx || terminal >= PGA_MAX_TERMINALS) return fail(ctx, PGA_ERR_INVALID_ARG, "terminal index out of range");
    if (!buf && capacity) return fail(ctx, PGA_ERR_INVALID_ARG, "null buffer with nonzero capacity");
    ctx->inst_ptr[terminal] = buf;
    ctx->inst_cap[terminal] = capacity;
    ctx->bind_dirty = true;
    return PGA_OK;
}

PGA_API pga_status pga_get_counts(pga_context* ctx, uint64_t* counts, uint32_t n) {
    if (!ctx || !counts || n > PGA_MAX_TERMINALS) return fail(ctx, PGA_ERR_INVALID_ARG, "bad counts request");
    NO_CAPTURE(ctx, "pga_get_counts");
    unsigned long long tmp[PGA_MAX_TERMINALS];
    HIP_TRY(ctx, hipMemcpyAsync(tmp, ctx->d_term_counts, sizeof(tmp), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < n; ++i) counts[i] = tmp[i];
    return PGA_OK;
}

static pga_status check_launch(pga_context* ctx, uint32_t first, uint32_t count) {
    if (!ctx) return PGA_ERR_INVALID_ARG;
    if (!ctx->initialized) return fail(ctx, PGA_ERR_NOT_INITIALIZED, "no program loaded");
    if ((uint64_t)first + count > ctx->n_axioms || (count && !ctx->axioms))
        return fail(ctx, PGA_ERR_INVALID_ARG, "axiom range exceeds the bound axioms");
    if (ctx->max_param_slot >= 0 &&
        ((uint32_t)ctx->max_param_slot >= ctx->aparams_per || ctx->aparams_axioms < ctx->n_axioms))
        return fail(ctx, PGA_ERR_INVALID_ARG, "program reads axiom parameters that are not set");
    if (ctx->capturing) {   // nothing may be allocated or copied while the stream records
        if (ctx->bind_dirty || ctx->cap != queue_capacity(ctx))
            return fail(ctx, PGA_ERR_INVALID_ARG, "bindings or queue capacity changed during the capture");
        return PGA_OK;
    }
    return ensure_buffers(ctx);
}

// Reads the error word, level and terminal counters of the finished run.
static pga_status finish(pga_context* ctx, uint32_t count, uint32_t host_levels) {
    uint32_t err = 0;
    HIP_TRY(ctx, hipMemcpy(&err, ctx->d_err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> lc(ctx->cfg.max_levels + 2), pc(ctx->cfg.max_levels + 2);
    HIP_TRY(ctx, hipMemcpy(lc.data(), ctx->d_desc_counts, lc.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(pc.data(), ctx->d_par_counts, pc.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    unsigned long long tc[PGA_MAX_TERMINALS];
    HIP_TRY(ctx, hipMemcpy(tc, ctx->d_term_counts, sizeof(tc), hipMemcpyDeviceToHost));
    uint32_t levels = host_levels;
    if (ctx->cfg.technique == PGA_TECHNIQUE_MEGAKERNEL && count > 0) levels = lc[ctx->cfg.max_levels + 1];
    ctx->colliders = 0;
    if (ctx->collision && ctx->d_coll_count) {
        HIP_TRY(ctx, hipMemcpy(&ctx->colliders, ctx->d_coll_count, sizeof(uint32_t), hipMemcpyDeviceToHost));
        ctx->colliders = std::min(ctx->colliders, ctx->coll_cap);
    }
    // the specialised kernel's workgroup-local levels: the deepest and the
    // parent records they wrote
    ctx->stats.levels = std::max(levels, pc[ctx->cfg.max_levels + 1]);
    ctx->stats.records = pc[ctx->cfg.max_levels];
    ctx->max_level_parents = 0;
    ctx->max_level_children = 0;
    for (uint32_t l = 0; l < levels && l < ctx->cfg.max_levels; ++l) {
        ctx->stats.records += pc[l];
        ctx->max_level_parents = std::max<uint64_t>(ctx->max_level_parents, pc[l]);
        ctx->max_level_children = std::max<uint64_t>(ctx->max_level_children, lc[l]);
        if (pc[l] > ctx->cap || lc[l] > ctx->cap) err |= 1u << PGA_ERR_OVERFLOW_QUEUE;
    }
    ctx->stats.terminals = 0;
    for (int t = 0; t < PGA_MAX_TERMINALS; ++t) ctx->stats.terminals += tc[t];
    return status_from_err(ctx, err);
}

// Programs with context sensitivity run on the interpreting kernels only,
// eagerly: returns an error for the paths that cannot run them.
static pga_status one_phase_only(pga_context* ctx, const char* what) {
    if (ctx->nphases > 1 || ctx->collision)
        return fail(ctx, PGA_ERR_UNSUPPORTED, std::string(what) + " does not run programs with Collider, "
                                              "IfCollides or more than one phase");
    return PGA_OK;
}
// Dynamic polygons and prisms run on the interpreting kernels (eagerly or
// captured) and the specialised kernel; the tape does not take them.
static pga_status no_polygons(pga_context* ctx, const char* what) {
    if (ctx->polygons)
        return fail(ctx, PGA_ERR_UNSUPPORTED, std::string(what) + " does not run programs with dynamic polygons");
    return PGA_OK;
}

// The derivation of a multi-phase program is one persistent launch per phase
// (MultiPhaseEvaluator::execute, GPUEvaluator.cuh:416-440): every launch
// derives from the axioms again, emits the terminals and stores the
// colliders of its own phase only, and carries the shapes of the phases up
// to its own (their IfCollides decisions read the colliders of earlier
// phases, so the re-derivation repeats them). Between launches the reset
// keeps the error word and the terminal counts, and publishes the colliders
// stored so far.
// Before phase p > 0: the LBVH over the colliders stored so far
// (GPUEvaluator.cuh:427-430 builds it after every phase but the last),
// then the reset that keeps the error word and publishes the count.
static pga_status begin_phase(pga_context* ctx, pga::KArgs& a, uint32_t p) {
    a.phase = p;
    if (p == 0) return PGA_OK;
    if (ctx->collision && ctx->coll_bvh) {
        HIP_TRY(ctx, pga_launch_collider_boxes(ctx->d_coll, ctx->coll_cap, ctx->d_coll_count, ctx->d_coll_boxes,
                                               ctx->stream));
        const uint32_t c = ctx->coll_bvh_cap;
        int32_t* t = ctx->d_coll_tree;
        HIP_TRY(ctx, pga_lbvh_enqueue_counted(ctx->coll_bvh, ctx->d_coll_boxes, ctx->d_coll_count, ctx->d_coll_keys,
                                              t, t + c, t + 2u * c, t + 3u * c, ctx->d_coll_nodes,
                                              ctx->d_coll_flat));
        a.bvh_flat = ctx->d_coll_flat;
    }
    HIP_TRY(ctx, pga_launch_reset(a, 2, ctx->stream));
    return PGA_OK;
}

static pga_status launch_phases(pga_context* ctx, pga::KArgs& a, const void** kernel) {
    for (uint32_t p = 0; p < ctx->nphases; ++p) {
        if (pga_status st = begin_phase(ctx, a, p)) return st;
        HIP_TRY(ctx, launch_derivation(ctx, a, kernel));
    }
    a.phase = 0;
    return PGA_OK;
}

static pga_status execute_async(pga_context* ctx, uint32_t first, uint32_t count, int32_t append) {
    pga_status st = check_launch(ctx, first, count);
    if (st) return st;
    if (ctx->cfg.technique != PGA_TECHNIQUE_MEGAKERNEL)
        return fail(ctx, PGA_ERR_UNSUPPORTED, "asynchronous execution needs the MEGAKERNEL technique");
    if (ctx->capturing && (st = one_phase_only(ctx, "graph capture"))) return st;
    pga::KArgs a = make_args(ctx, first, count);
    const bool timed = !ctx->capturing;     // the events time eager launches only
    HIP_TRY(ctx, pga_launch_reset(a, append ? 0 : 1, ctx->stream));
    if (timed) HIP_TRY(ctx, hipEventRecord(ctx->evk0, ctx->stream));
    const void* kernel = nullptr;
    if (count > 0 && (st = launch_phases(ctx, a, &kernel))) return st;
    if (count == 0) kernel = ctx->spec_on ? (const void*)ctx->spec_fn : nullptr;
    if (timed) HIP_TRY(ctx, hipEventRecord(ctx->evk1, ctx->stream));
    if (ctx->capturing) ctx->captured.push_back({a, ctx->mega_blocks, ctx->cfg.block_size, first, count, kernel});
    ctx->pending_count = count;
    return PGA_OK;
}

PGA_API pga_status pga_execute_async(pga_context* ctx, uint32_t first, uint32_t count, int32_t append) {
    pga_status st = execute_async(ctx, first, count, append);
    if (st && ctx && ctx->capturing && !ctx->capture_failed) {
        ctx->capture_failed = true;
        ctx->capture_error = ctx->error;
    }
    return st;
}

PGA_API pga_status pga_synchronize(pga_context* ctx) {
    if (!ctx) return PGA_ERR_INVALID_ARG;
    NO_CAPTURE(ctx, "pga_synchronize");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->evk0, ctx->evk1));
    ctx->stats.seconds = ms * 1e-3;
    return finish(ctx, ctx->pending_count, 0);
}

PGA_API pga_status pga_kernel_seconds(pga_context* ctx, double* seconds) {
    if (!ctx || !seconds) return PGA_ERR_INVALID_ARG;
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->evk0, ctx->evk1));
    *seconds = ms * 1e-3;
    return PGA_OK;
}

PGA_API pga_status pga_execute_range(pga_context* ctx, uint32_t first, uint32_t count, int32_t append,
                                     double* seconds) {
    NO_CAPTURE(ctx, "pga_execute");
    pga_status st = check_launch(ctx, first, count);
    if (st) return st;
    pga::KArgs a = make_args(ctx, first, count);
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, pga_launch_reset(a, append ? 0 : 1, ctx->stream));
    uint32_t levels = 0;
    const uint32_t bs = ctx->cfg.block_size;
    HIP_TRY(ctx, hipEventRecord(ctx->evk0, ctx->stream));
    if (ctx->cfg.technique == PGA_TECHNIQUE_KERNELS) {
        for (uint32_t p = 0; p < ctx->nphases; ++p) {   // one level loop per phase
            if ((st = begin_phase(ctx, a, p))) return st;
            uint32_t n = count, level = 0;
            while (n > 0) {
                if (level >= ctx->cfg.max_levels) {
                    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
                    return fail(ctx, PGA_ERR_MAX_DEPTH, "maximum number of derivation levels reached");
                }
                HIP_TRY(ctx, pga_launch_level(a, level, ctx->mega_blocks, bs, ctx->stream));
                HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->d_desc_counts + level, sizeof(uint32_t),
                                            hipMemcpyDeviceToHost, ctx->stream));
                HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
                n = std::min(ctx->h_pinned[0], ctx->cap);
                ++level;
            }
            levels += level;
        }
    } else if (count > 0) {
        const void* kernel = nullptr;
        if ((st = launch_phases(ctx, a, &kernel))) return st;
    }
    HIP_TRY(ctx, hipEventRecord(ctx->evk1, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.seconds = ms * 1e-3;
    if (seconds) *seconds = ms * 1e-3;
    return finish(ctx, count, levels);
}

PGA_API pga_status pga_execute(pga_context* ctx, uint32_t num_axioms, double* seconds) {
    return pga_execute_range(ctx, 0, num_axioms, 0, seconds);
}

PGA_API pga_status pga_set_timestamps(pga_context* ctx, uint64_t* device_buffer, uint32_t capacity) {
    if (!ctx || (!device_buffer && capacity)) return fail(ctx, PGA_ERR_INVALID_ARG, "null timestamp buffer");
    ctx->ticks = device_buffer;
    ctx->ticks_cap = device_buffer ? capacity : 0u;
    return PGA_OK;
}

PGA_API double pga_timestamp_hz(pga_context* ctx) {
    if (!ctx) return 0.0;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess) return 0.0;
    return (double)khz * 1e3;
}

PGA_API pga_status pga_get_stats(const pga_context* ctx, pga_stats* stats) {
    if (!ctx || !stats) return PGA_ERR_INVALID_ARG;
    *stats = ctx->stats;
    return PGA_OK;
}

PGA_API pga_status pga_set_collider_capacity(pga_context* ctx, uint32_t capacity) {
    if (!ctx) return PGA_ERR_INVALID_ARG;
    NO_CAPTURE(ctx, "pga_set_collider_capacity");
    if (capacity > (1u << 26)) return fail(ctx, PGA_ERR_INVALID_ARG, "more than 2^26 colliders");
    ctx->coll_cap_req = capacity;
    return ensure_colliders(ctx);            // allocated when a program stores colliders
}

PGA_API pga_status pga_num_colliders(const pga_context* ctx, uint32_t* colliders, uint32_t* phases) {
    if (!ctx) return PGA_ERR_INVALID_ARG;
    if (colliders) *colliders = ctx->colliders;
    if (phases) *phases = ctx->initialized ? ctx->nphases : 0u;
    return PGA_OK;
}

// Workgroups of one launch at workgroup size bs: all that fit on the device
// at once (cooperative launch), at most kMaxBlocksPerCu per compute unit.
static pga_status resident_blocks(pga_context* ctx, uint32_t bs, uint32_t* blocks) {
    int per_cu = 0;
    HIP_TRY(ctx, pga_mega_occupancy(bs, ctx->lds_bytes, ctx->collision || ctx->nphases > 1, &per_cu));
    *blocks = per_cu < 1 ? 0u : std::min((uint32_t)per_cu, kMaxBlocksPerCu) * (uint32_t)ctx->num_cus;
    return PGA_OK;
}

PGA_API pga_status pga_specialize(pga_context* ctx, uint32_t block_size, int32_t enable, double* seconds) {
    NO_CAPTURE(ctx, "pga_specialize");
    if (!ctx) return PGA_ERR_INVALID_ARG;
    if (seconds) *seconds = 0.0;
    if (!enable) {
        ctx->spec_on = false;
        return PGA_OK;
    }
    if (!ctx->initialized) return fail(ctx, PGA_ERR_NOT_INITIALIZED, "no program loaded");
    if (pga_status st = one_phase_only(ctx, "the specialised kernel")) return st;
    if (block_size == 0) block_size = kDefaultBlockSize;
    if (!spec_block_size_supported(block_size))
        return fail(ctx, PGA_ERR_INVALID_ARG, "unsupported workgroup size");
    if (ctx->spec_fn && ctx->spec_bs == block_size) {
        ctx->spec_on = true;
        return PGA_OK;
    }
    drop_spec(ctx);
    const auto t0 = std::chrono::steady_clock::now();
    const std::string src = pga_spec_source(ctx->h_entries, ctx->count_geom, ctx->use_seed, block_size);
    std::vector<char> code;
    std::string log;
    bool cached = false;
    if (!pga_spec_compile(src, code, log, &cached))
        return fail(ctx, PGA_ERR_UNSUPPORTED, "run-time compilation of the specialised kernel failed: " + log);
    HIP_TRY(ctx, hipModuleLoadData(&ctx->spec_module, code.data()));
    HIP_TRY(ctx, hipModuleGetFunction(&ctx->spec_fn, ctx->spec_module, "pga_spec_mega"));
    int per_cu = 0;
    HIP_TRY(ctx, hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ctx->spec_fn, (int)block_size,
                                                                    ctx->lds_bytes));
    if (per_cu < 1) {
        drop_spec(ctx);
        return fail(ctx, PGA_ERR_UNSUPPORTED, "the specialised kernel does not fit on a compute unit");
    }
    ctx->spec_bs = block_size;
    ctx->spec_blocks = std::min((uint32_t)per_cu, kMaxBlocksPerCu) * (uint32_t)ctx->num_cus;
    if (pga_status st = ensure_local(ctx, ctx->spec_blocks, block_size)) {
        drop_spec(ctx);
        return st;
    }
    ctx->spec_on = true;
    if (seconds && !cached) seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PGA_OK;
}

PGA_API pga_status pga_set_schedule(pga_context* ctx, int32_t local, uint32_t grab, uint32_t spill) {
    NO_CAPTURE(ctx, "pga_set_schedule");
    if (!ctx) return PGA_ERR_INVALID_ARG;
    ctx->sched_local = local != 0;
    ctx->sched_grab = grab;
    ctx->sched_spill = spill;
    return PGA_OK;
}

PGA_API pga_status pga_specialize_check(const pga_procedure* procedures, uint32_t num_procedures,
                                        const pga_dispatch_entry* entries, uint32_t num_entries, uint32_t block_size,
                                        char* log, uint32_t log_size, double* seconds) {
    if (seconds) *seconds = 0.0;
    if (log && log_size) log[0] = '\0';
    if (block_size == 0) block_size = kDefaultBlockSize;
    if (!spec_block_size_supported(block_size)) return PGA_ERR_INVALID_ARG;
    std::unique_ptr<pga_context> ctx(new pga_context());   // host tables only: no device call
    ctx->cfg.block_size = block_size;
    std::string msg;
    pga_status st = build_program(ctx.get(), procedures, num_procedures, entries, num_entries, false);
    if (st == PGA_OK) st = one_phase_only(ctx.get(), "the specialised kernel");
    if (st != PGA_OK) {
        msg = ctx->error;
    } else {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<char> code;
        bool cached = false;
        if (!pga_spec_compile(pga_spec_source(ctx->h_entries, ctx->count_geom, ctx->use_seed, block_size), code, msg,
                              &cached))
            st = PGA_ERR_UNSUPPORTED;
        else if (seconds && !cached)
            *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (log && log_size) {
        const size_t m = std::min<size_t>(msg.size(), log_size - 1u);
        std::memcpy(log, msg.data(), m);
        log[m] = '\0';
    }
    return st;
}

PGA_API pga_status pga_autotune(pga_context* ctx, uint32_t num_axioms, uint32_t repeats, pga_tune_result* out) {
    NO_CAPTURE(ctx, "pga_autotune");
    if (!ctx || !out) return fail(ctx, PGA_ERR_INVALID_ARG, "null argument");
    if (repeats == 0) repeats = 3;
    const pga_config saved = ctx->cfg;
    const uint32_t saved_blocks = ctx->mega_blocks;
    auto time_config = [&](double* best) -> pga_status {
        *best = 1e30;
        for (uint32_t r = 0; r < repeats; ++r) {
            double s = 0.0;
            pga_status st = pga_execute_range(ctx, 0, num_axioms, 0, &s);
            if (st) return st;
            *best = std::min(*best, s);
        }
        return PGA_OK;
    };
    const bool saved_spec = ctx->spec_on;
    const uint32_t saved_spec_blocks = ctx->spec_blocks;
    const bool saved_sched_local = ctx->sched_local;
    const uint32_t saved_sched_grab = ctx->sched_grab;
    const uint32_t saved_sched_spill = ctx->sched_spill;
    pga_tune_result res{};
    pga_status st = time_config(&res.initial_seconds);
    if (st) return st;
    res.technique = saved.technique;
    res.block_size = saved_spec ? ctx->spec_bs : saved.block_size;
    res.grid_blocks = saved_spec ? saved_spec_blocks : saved_blocks;
    res.specialized = saved_spec ? 1u : 0u;
    res.local_grab = saved_spec && saved_sched_local ? local_grab(ctx, num_axioms) : 0u;
    res.local_spill = saved_spec && saved_sched_local ? saved_sched_spill : 0u;
    res.seconds = res.initial_seconds;
    res.max_level_parents = ctx->max_level_parents;
    res.max_level_children = ctx->max_level_children;
    res.candidates = 1;
    const uint32_t techniques[] = {PGA_TECHNIQUE_MEGAKERNEL, PGA_TECHNIQUE_KERNELS};
    const uint32_t sizes[] = {256u, 512u, 768u, 1024u};
    ctx->spec_on = false;
    if (ctx->spec_fn) {   // the specialised kernel (compiled by pga_specialize) at a full and a half grid
        int per_cu = 0;
        HIP_TRY(ctx, hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ctx->spec_fn, (int)ctx->spec_bs,
                                                                        ctx->lds_bytes));
        const uint32_t full = std::min((uint32_t)std::max(per_cu, 1), kMaxBlocksPerCu) * (uint32_t)ctx->num_cus;
        const uint32_t grids[] = {full, std::max(full / 2u, (uint32_t)ctx->num_cus)};
        ctx->cfg.technique = PGA_TECHNIQUE_MEGAKERNEL;
        for (int g = 0; g < 2 && !st; ++g) {
            if (g == 1 && grids[1] >= full) continue;
            ctx->spec_on = true;
            ctx->spec_blocks = grids[g];
            // level schedules: grid-wide; workgroup-local with the automatic
            // hand-out, half of it and four times it; and the automatic
            // hand-out with levels wider than two rounds of the workgroup
            // left to the grid
            ctx->sched_local = true;
            ctx->sched_grab = 0;
            ctx->sched_spill = 0;
            const uint32_t autog = local_grab(ctx, num_axioms);
            struct Sched { uint32_t grab, spill; };
            const Sched scheds[] = {{0u, 0u}, {autog, 0u}, {std::max(autog / 2u, 1u), 0u},
                                    {std::min(4u * autog, kLocalGrabMax), 0u}, {autog, 2u * ctx->spec_bs}};
            for (int k = 0; k < 5 && !st; ++k) {
                if ((k == 2 || k == 3) && (scheds[k].grab == autog || scheds[k].grab == scheds[k - 1].grab)) continue;
                ctx->sched_local = k != 0;
                ctx->sched_grab = scheds[k].grab;
                ctx->sched_spill = scheds[k].spill;
                double s = 0.0;
                st = time_config(&s);
                ++res.candidates;
                if (!st && s < res.seconds) {
                    res.seconds = s;
                    res.technique = PGA_TECHNIQUE_MEGAKERNEL;
                    res.block_size = ctx->spec_bs;
                    res.grid_blocks = grids[g];
                    res.specialized = 1u;
                    res.local_grab = scheds[k].grab;
                    res.local_spill = scheds[k].spill;
                }
            }
        }
        ctx->sched_local = saved_sched_local;
        ctx->sched_grab = saved_sched_grab;
        ctx->sched_spill = saved_sched_spill;
        ctx->spec_on = false;
        ctx->cfg.technique = saved.technique;
    }
    for (uint32_t tech : techniques) {
        if (st) break;
        for (uint32_t bs : sizes) {
            uint32_t full = 0;
            st = resident_blocks(ctx, bs, &full);
            if (st) break;
            if (full == 0) continue;
            const uint32_t grids[] = {full, std::max(full / 2u, (uint32_t)ctx->num_cus)};
            for (int g = 0; g < 2; ++g) {
                if (g == 1 && grids[1] >= full) continue;
                ctx->cfg.technique = tech;
                ctx->cfg.block_size = bs;
                ctx->mega_blocks = grids[g];
                double s = 0.0;
                st = time_config(&s);
                if (st) break;
                ++res.candidates;
                if (s < res.seconds) {
                    res.seconds = s;
                    res.technique = tech;
                    res.block_size = bs;
                    res.grid_blocks = grids[g];
                }
            }
            if (st) break;
        }
        if (st) break;
    }
    if (st) {   // leave the context as it was
        ctx->cfg = saved;
        ctx->mega_blocks = saved_blocks;
        ctx->spec_on = saved_spec;
        ctx->spec_blocks = saved_spec_blocks;
        return st;
    }
    ctx->cfg.technique = res.technique;
    if (res.specialized) {         // the interpreting kernel keeps its configuration for when it is back
        ctx->cfg.block_size = saved.block_size;
        ctx->mega_blocks = saved_blocks;
        ctx->spec_on = true;
        ctx->spec_blocks = res.grid_blocks;
        ctx->sched_local = res.local_grab != 0u;
        ctx->sched_grab = res.local_grab;
        ctx->sched_spill = res.local_spill;
    } else {
        ctx->cfg.block_size = res.block_size;
        ctx->mega_blocks = res.grid_blocks;
    }
    st = pga_execute_range(ctx, 0, num_axioms, 0, nullptr);   // the buffers hold the chosen run
    *out = res;
    return st;
}

// ---- reverse mode ----

PGA_API pga_status pga_tape_reserve(pga_context* ctx, uint32_t max_records, uint32_t max_children) {
    NO_CAPTURE(ctx, "pga_tape_reserve");
    if (!ctx) return PGA_ERR_INVALID_ARG;
    if (max_records == 0 || max_children == 0) return fail(ctx, PGA_ERR_INVALID_ARG, "empty tape");
    if (ctx->stream) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    hfree(ctx->d_tape_par); hfree(ctx->d_tape_desc); hfree(ctx->d_tape_out); hfree(ctx->d_adj_slots);
    ctx->tape_valid = false;
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_tape_par, (size_t)max_records * sizeof(pga::Rec)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_tape_desc, (size_t)max_children * sizeof(pga::Desc)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_tape_out, (size_t)max_children * sizeof(uint4)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->d_adj_slots, (size_t)max_children * pga::kSlotFloats * sizeof(float)));
    if (!ctx->d_dl) HIP_TRY(ctx, hipMalloc((void**)&ctx->d_dl, PGA_MAX_TERMINALS * sizeof(float*)));
    ctx->tape_par_cap = max_records;
    ctx->tape_child_cap = max_children;
    return PGA_OK;
}

PGA_API pga_status pga_forward_tape(pga_context* ctx, uint32_t num_axioms, double* seconds) {
    NO_CAPTURE(ctx, "pga_forward_tape");
    pga_status st = check_launch(ctx, 0, num_axioms);
    if (st) return st;
    if ((st = one_phase_only(ctx, "the derivation tape"))) return st;
    if ((st = no_polygons(ctx, "the derivation tape"))) return st;
    if (!ctx->d_tape_par) return fail(ctx, PGA_ERR_NOT_INITIALIZED, "no tape: call pga_tape_reserve first");
    if (num_axioms > ctx->tape_child_cap) return fail(ctx, PGA_ERR_OVERFLOW_QUEUE, "tape smaller than the axioms");
    if (ctx->tape_blocks == 0) {
        int per_cu = 0;
        HIP_TRY(ctx, pga_tape_occupancy(ctx->lds_bytes, &per_cu));
        if (per_cu < 1) return fail(ctx, PGA_ERR_UNSUPPORTED, "tape kernel does not fit on a compute unit");
        ctx->tape_blocks = std::min((uint32_t)per_cu, kMaxBlocksPerCu) * (uint32_t)ctx->num_cus;
    }
    ctx->tape_valid = false;
    pga::KArgs a = make_args(ctx, 0, num_axioms);
    a.tape = 1;
    a.tape_par = ctx->d_tape_par;
    a.tape_desc = ctx->d_tape_desc;
    a.tape_out = ctx->d_tape_out;
    a.tape_out_cap = ctx->tape_child_cap;
    a.cap_par = ctx->tape_par_cap;
    a.cap_desc = ctx->tape_child_cap - num_axioms;   // children after the axioms
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, pga_launch_reset(a, 1, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->evk0, ctx->stream));
    if (num_axioms > 0) HIP_TRY(ctx, pga_launch_mega_tape(a, ctx->tape_blocks, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->evk1, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.seconds = ms * 1e-3;
    if (seconds) *seconds = ms * 1e-3;
    // per-level counts of the tape (the stats' queue check uses the ping-pong
    // capacity, the tape its own)
    uint32_t err = 0;
    HIP_TRY(ctx, hipMemcpy(&err, ctx->d_err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> lc(ctx->cfg.max_levels + 2), pc(ctx->cfg.max_levels + 2);
    HIP_TRY(ctx, hipMemcpy(lc.data(), ctx->d_desc_counts, lc.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(pc.data(), ctx->d_par_counts, pc.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    unsigned long long tc[PGA_MAX_TERMINALS];
    HIP_TRY(ctx, hipMemcpy(tc, ctx->d_term_counts, sizeof(tc), hipMemcpyDeviceToHost));
    const uint32_t levels = num_axioms ? lc[ctx->cfg.max_levels + 1] : 0u;
    ctx->stats.levels = levels;
    ctx->stats.records = 0;
    ctx->stats.terminals = 0;
    for (int t = 0; t < PGA_MAX_TERMINALS; ++t) ctx->stats.terminals += tc[t];
    ctx->tape_pc.assign(pc.begin(), pc.begin() + levels);
    ctx->tape_dc.assign(lc.begin(), lc.begin() + levels);
    uint64_t sp = 0, sd = 0;
    for (uint32_t l = 0; l < levels; ++l) { sp += pc[l]; sd += lc[l]; }
    ctx->stats.records = sp;
    if (sp > ctx->tape_par_cap || (uint64_t)num_axioms + sd > ctx->tape_child_cap) err |= 1u << PGA_ERR_OVERFLOW_QUEUE;
    st = status_from_err(ctx, err);
    if (st == PGA_OK) {
        ctx->tape_valid = true;
        ctx->tape_axioms = num_axioms;
        ++ctx->tape_gen;
    }
    return st;
}

PGA_API pga_status pga_tape_generation(const pga_context* ctx, uint64_t* generation) {
    if (!ctx || !generation) return PGA_ERR_INVALID_ARG;
    *generation = ctx->tape_valid ? ctx->tape_gen : 0u;
    return PGA_OK;
}

PGA_API pga_status pga_tape_counts(const pga_context* ctx, uint64_t* parent_records, uint64_t* children) {
    if (!ctx || !parent_records || !children) return PGA_ERR_INVALID_ARG;
    uint64_t sp = 0, sd = ctx->tape_valid ? ctx->tape_axioms : 0u;
    if (ctx->tape_valid)
        for (size_t l = 0; l < ctx->tape_dc.size(); ++l) { sp += ctx->tape_pc[l]; sd += ctx->tape_dc[l]; }
    *parent_records = sp;
    *children = sd;
    return PGA_OK;
}

PGA_API pga_status pga_backward(pga_context* ctx, const float* const* dl_dinstances, float* dl_dparams,
                                double* seconds) {
    NO_CAPTURE(ctx, "pga_backward");
    if (!ctx || !dl_dinstances || !dl_dparams) return fail(ctx, PGA_ERR_INVALID_ARG, "null argument");
    if (!ctx->tape_valid) return fail(ctx, PGA_ERR_NOT_INITIALIZED, "no recorded tape: run pga_forward_tape");
    if (ctx->aparams_per == 0) return fail(ctx, PGA_ERR_INVALID_ARG, "no per-axiom parameters bound");
    for (uint32_t t = 0; t < PGA_MAX_TERMINALS; ++t)   // read as 16-byte rows
        if (((uintptr_t)dl_dinstances[t] & 15u) != 0u)
            return fail(ctx, PGA_ERR_INVALID_ARG, "dL/dinstance buffers must be 16-byte aligned");
    HIP_TRY(ctx, hipMemcpyAsync((void*)ctx->d_dl, (const void*)dl_dinstances, PGA_MAX_TERMINALS * sizeof(float*),
                                hipMemcpyHostToDevice, ctx->stream));
    const uint32_t n = ctx->tape_axioms;
    pga::KArgs a = make_args(ctx, 0, n);
    a.tape = 1;
    a.tape_par = ctx->d_tape_par;
    a.tape_desc = ctx->d_tape_desc;
    a.tape_out = ctx->d_tape_out;
    a.tape_out_cap = ctx->tape_child_cap;
    a.cap_par = ctx->tape_par_cap;
    a.cap_desc = ctx->tape_child_cap - n;
    const uint32_t levels = (uint32_t)ctx->tape_dc.size();
    std::vector<uint32_t> pbase(levels + 1, 0), dbase(levels + 1, 0);
    for (uint32_t l = 0; l < levels; ++l) {
        pbase[l + 1] = pbase[l] + ctx->tape_pc[l];
        dbase[l + 1] = dbase[l] + ctx->tape_dc[l];
    }
    const size_t ng = (size_t)n * ctx->aparams_per;
    if (ng > ctx->grad64_cap) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        hfree(ctx->d_grad64);
        HIP_TRY(ctx, hipMalloc((void**)&ctx->d_grad64, ng * sizeof(double)));
        ctx->grad64_cap = ng;
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_grad64, 0, ng * sizeof(double), ctx->stream));
    for (uint32_t L = levels; L-- > 0;) {
        pga::AdjArgs g{};
        g.dl = ctx->d_dl;
        g.slots = ctx->d_adj_slots;
        g.grad = ctx->d_grad64;
        g.level = L;
        g.n_children = L == 0 ? n : ctx->tape_dc[L - 1];
        g.child_base = L == 0 ? 0u : n + dbase[L - 1];
        g.next_child_base = n + dbase[L];
        g.par_in_base = L == 0 ? 0u : pbase[L - 1];
        g.desc_in_base = L == 0 ? 0u : dbase[L - 1];
        g.slots_cap = ctx->tape_child_cap;
        HIP_TRY(ctx, pga_launch_adjoint(a, g, ctx->stream));
    }
    HIP_TRY(ctx, pga_launch_grad_out(ctx->d_grad64, dl_dparams, (uint32_t)ng, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (seconds) *seconds = ms * 1e-3;
    uint32_t err = 0;
    HIP_TRY(ctx, hipMemcpy(&err, ctx->d_err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return status_from_err(ctx, err);
}

PGA_API pga_status pga_capture_begin(pga_context* ctx) {
    if (!ctx) return PGA_ERR_INVALID_ARG;
    NO_CAPTURE(ctx, "pga_capture_begin");
    if (!ctx->initialized) return fail(ctx, PGA_ERR_NOT_INITIALIZED, "no program loaded");
    if (ctx->cfg.technique != PGA_TECHNIQUE_MEGAKERNEL)
        return fail(ctx, PGA_ERR_UNSUPPORTED, "graphs need the MEGAKERNEL technique");
    pga_status st = ensure_buffers(ctx);        // bindings copied before the stream records
    if (st) return st;
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    ctx->capture_failed = false;
    ctx->capture_error.clear();
    ctx->captured.clear();
    return PGA_OK;
}

PGA_API pga_status pga_capture_end(pga_context* ctx, pga_graph** out) {
    if (!ctx || !out) return fail(ctx, PGA_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!ctx->capturing) return fail(ctx, PGA_ERR_INVALID_ARG, "no graph capture is open");
    ctx->capturing = false;
    std::unique_ptr<pga_graph> g(new pga_graph());
    g->ctx = ctx;
    g->launches.swap(ctx->captured);
    g->cap = ctx->cap;
    g->last_count = g->launches.empty() ? 0u : g->launches.back().count;
    HIP_TRY(ctx, hipStreamEndCapture(ctx->stream, &g->graph));
    if (ctx->capture_failed)
        return fail(ctx, PGA_ERR_INVALID_ARG, "a call failed during the capture: " + ctx->capture_error);
    size_t n = 0;
    HIP_TRY(ctx, hipGraphGetNodes(g->graph, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    if (n) HIP_TRY(ctx, hipGraphGetNodes(g->graph, nodes.data(), &n));
    for (hipGraphNode_t node : nodes) {
        hipGraphNodeType type;
        HIP_TRY(ctx, hipGraphNodeGetType(node, &type));
        if (type != hipGraphNodeTypeKernel) continue;
        ++g->kernel_nodes;
        hipKernelNodeParams kp{};
        HIP_TRY(ctx, hipGraphKernelNodeGetParams(node, &kp));
        hipKernelNodeAttrValue v{};
        HIP_TRY(ctx, hipGraphKernelNodeGetAttribute(node, hipKernelNodeAttributeCooperative, &v));
        const bool spec = ctx->spec_fn != nullptr && kp.func == (void*)ctx->spec_fn;
        if (!v.cooperative && (spec || pga_needs_coresidency(kp.func))) {   // the grid barrier needs every workgroup resident
            v.cooperative = 1;
            HIP_TRY(ctx, hipGraphKernelNodeSetAttribute(node, hipKernelNodeAttributeCooperative, &v));
            HIP_TRY(ctx, hipGraphKernelNodeGetAttribute(node, hipKernelNodeAttributeCooperative, &v));
            if (!v.cooperative) return fail(ctx, PGA_ERR_UNSUPPORTED, "graph node cannot launch cooperatively");
        }
        if (v.cooperative) ++g->cooperative_nodes;
    }
    HIP_TRY(ctx, hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
    ctx->graphs.push_back(*out);
    return PGA_OK;
}

PGA_API pga_status pga_graph_launch(pga_context* ctx, pga_graph* g) {
    if (!ctx || !g) return fail(ctx, PGA_ERR_INVALID_ARG, "null argument");
    NO_CAPTURE(ctx, "pga_graph_launch");
    if (g->ctx != ctx) return fail(ctx, PGA_ERR_INVALID_ARG, "the graph was captured on another context");
    if (ctx->bind_dirty || ctx->cap != g->cap || queue_capacity(ctx) != g->cap)
        return fail(ctx, PGA_ERR_INVALID_ARG, "bindings or queue capacity changed since the capture; capture again");
    for (const CapturedLaunch& l : g->launches) {
        pga_status st = check_launch(ctx, l.first, l.count);
        if (st) return st;
        const pga::KArgs a = make_args(ctx, l.first, l.count);
        const void* kernel = ctx->spec_on ? (const void*)ctx->spec_fn : nullptr;
        if (std::memcmp(&a, &l.args, sizeof(a)) != 0 || l.blocks != ctx->mega_blocks ||
            l.block_size != ctx->cfg.block_size || l.kernel != kernel)
            return fail(ctx, PGA_ERR_INVALID_ARG,
                        "the context changed since the capture (bindings, program layout or configuration); "
                        "capture again");
    }
    HIP_TRY(ctx, hipEventRecord(ctx->evk0, ctx->stream));
    HIP_TRY(ctx, hipGraphLaunch(g->exec, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->evk1, ctx->stream));
    ctx->pending_count = g->last_count;
    return PGA_OK;
}

PGA_API void pga_graph_destroy(pga_graph* g) {
    if (!g) return;
    if (pga_context* ctx = g->ctx) {          // NULL once the context is destroyed
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        ctx->graphs.erase(std::remove(ctx->graphs.begin(), ctx->graphs.end(), g), ctx->graphs.end());
    }
    delete g;
}

PGA_API pga_status pga_graph_info(const pga_graph* g, uint32_t* kernel_nodes, uint32_t* cooperative_nodes) {
    if (!g) return PGA_ERR_INVALID_ARG;
    if (kernel_nodes) *kernel_nodes = g->kernel_nodes;
    if (cooperative_nodes) *cooperative_nodes = g->cooperative_nodes;
    return PGA_OK;
}

PGA_API pga_status pga_device_counts(pga_context* ctx, const uint64_t** device_counts) {
    if (!ctx || !device_counts) return fail(ctx, PGA_ERR_INVALID_ARG, "null argument");
    *device_counts = reinterpret_cast<const uint64_t*>(ctx->d_term_counts);
    return PGA_OK;
}

PGA_API pga_status pga_device_alloc(uint64_t bytes, void** ptr) {
    if (!ptr) return PGA_ERR_INVALID_ARG;
    *ptr = nullptr;
    if (bytes == 0) return PGA_OK;
    return hipMalloc(ptr, bytes) == hipSuccess ? PGA_OK : PGA_ERR_HIP;
}

PGA_API pga_status pga_device_free(void* ptr) {
    if (!ptr) return PGA_OK;
    return hipFree(ptr) == hipSuccess ? PGA_OK : PGA_ERR_HIP;
}

PGA_API pga_status pga_copy_to_host(void* dst, const void* src, uint64_t bytes) {
    if ((!dst || !src) && bytes) return PGA_ERR_INVALID_ARG;
    if (!bytes) return PGA_OK;
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? PGA_OK : PGA_ERR_HIP;
}

PGA_API pga_status pga_get_stream(pga_context* ctx, void** stream) {
    if (!ctx || !stream) return fail(ctx, PGA_ERR_INVALID_ARG, "null argument");
    *stream = (void*)ctx->stream;
    return PGA_OK;
}

} // extern "C"
