// pga_kernels.h - kernel arguments and launchers shared by the kernel and
// host translation units.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "pga_ops.h"

namespace pga {

constexpr uint32_t kBarWords = 192;     // grid barrier: generation + 9 counters, 64 bytes apart (words 0-159)
// words after the grid barrier's, on a line of their own and zeroed with
// them before every launch (k_reset, at most 256 words): the next axiom to
// hand out and the parents / descriptors spilled from workgroup-local
// frontiers to the grid's arrays
constexpr uint32_t kBarGrab = 160, kBarSpillPar = 161, kBarSpillDesc = 162;

struct KArgs {
    Prog prog;                          // tables in the global program image
    const uint4* image;                 // the program image (entries|params|succ|exp|slots)
    uint32_t image_bytes;               // multiple of 16
    uint32_t lds_bytes;                 // dynamic LDS per workgroup: image + terminal tables, 0 = none
    uint32_t off_params, off_succ, off_exp, off_slots;
    const pga_axiom* axioms;
    uint32_t first_axiom;
    uint32_t n_axioms;
    const float* aparams;
    uint32_t aparams_per;
    uint32_t naxioms_total;
    uint32_t key0, key1;
    pga_instance* const* inst;          // device array [PGA_MAX_TERMINALS]
    const unsigned long long* inst_cap; // device array [PGA_MAX_TERMINALS]
    unsigned long long* term_counts;    // device array [PGA_MAX_TERMINALS]
    Rec* par0;                          // parent records, written by even / odd levels
    Rec* par1;
    Desc* desc0;                        // child descriptors, written by even / odd levels
    Desc* desc1;
    uint32_t cap_par, cap_desc;
    uint32_t* par_counts;               // [max_levels + 2]: parents written by level L
    uint32_t* desc_counts;              // [max_levels + 2]: children of level L + 1;
                                        // [max_levels + 1] = levels run by k_mega
    uint32_t* block_totals;             // [grid blocks][nslots + 2] counts of one pass
    unsigned long long* ticks;          // phase timestamps (NULL: off)
    uint32_t ticks_cap;
    uint32_t* err;
    uint32_t* bar;                      // grid barrier words
    uint32_t max_levels;
    uint32_t max_chain;
    uint32_t count_geom;                // 1: some decision reads the model matrix
    uint32_t use_seed;                  // 1: some parameter reads a shape seed
    uint32_t polygons;                  // 1: the program has dynamic polygons or prisms
    // multi-phase programs (one launch per phase): Rec::ed depth bits
    // (0xFFFF: one phase), the phase derived, and the collider store with its
    // append counter and the count stored by earlier phases
    uint32_t dmask;
    uint32_t phase;
    float4* coll;
    uint32_t* coll_count;
    const uint32_t* coll_visible;
    uint32_t coll_cap;
    const float4* bvh_flat;             // BVH over the visible colliders, depth-first (NULL: none)
    uint32_t* coll_bits;                // per child of a level: IfCollides outcomes of the count pass
    uint32_t coll_bits_cap;
    // derivation tape (reverse mode): every level's parent records and
    // descriptors kept (cap_par / cap_desc = their capacities), every
    // fan-out materialised, and one outcome per child
    // workgroup-local frontiers of the specialised kernel: per workgroup two
    // arrays of loc_kp parent records and two of loc_kd descriptors (NULL:
    // every level runs over the grid), of which a level may fill loc_cap_p /
    // loc_cap_d (at most the queue capacity); axioms handed out loc_grab at
    // a time
    Rec* loc_par;
    Desc* loc_desc;
    uint32_t loc_kp, loc_kd, loc_cap_p, loc_cap_d, loc_grab;
    uint32_t tape;
    Rec* tape_par;
    Desc* tape_desc;
    uint4* tape_out;                    // per child: {kind | slot << 8, index, dslot, nd}
    uint32_t tape_out_cap;
};

// Outcome of a child in the tape.
enum TapeKind : uint32_t { T_DROP = 0, T_TERM = 1, T_FAN = 2, T_SELF = 3 };

// One level of the reverse pass.
struct AdjArgs {
    const float* const* dl;       // [PGA_MAX_TERMINALS] dL/dinstance (16 floats per instance) or null
    float* slots;                 // per child: {dparent.m[12], dparent.s[3], d offset, d size, 0, 0, 0}
    double* grad;                 // [axioms][params per axiom], FP64 accumulators
    uint32_t level;
    uint32_t n_children;          // children of this level
    uint32_t child_base;          // first child of this level in the child numbering
    uint32_t next_child_base;     // first child of the next level
    uint32_t par_in_base;         // parent records read by this level (level - 1 outputs)
    uint32_t desc_in_base;        // descriptors read by this level
    uint32_t slots_cap;           // children in the slot array
};

constexpr uint32_t kTapeBlock = 768;    // workgroup size of the tape-recording build (768 threads: 4.5 ms against 5.2 ms at 256 for Menger-5 with 16,384 parameterised lots)
constexpr uint32_t kSlotFloats = 20;    // adjoint slot of a child (80 bytes)

} // namespace pga

#ifndef __HIPCC_RTC__   // host launchers
bool pga_block_size_supported(uint32_t block_size);   // 256, 512, 768 or 1024
// flags: bit 0 zeroes the terminal counts (a fresh run), bit 1 keeps the
// error word and snapshots the collider count (a later phase of the run)
hipError_t pga_launch_reset(const pga::KArgs& a, int flags, hipStream_t s);
hipError_t pga_launch_level(const pga::KArgs& a, uint32_t level, uint32_t blocks, uint32_t block_size,
                            hipStream_t s);
// blocks per compute unit of both interpreting kernels of a workgroup size;
// context_sensitive: the variants that run Collider / IfCollides and phases
hipError_t pga_mega_occupancy(uint32_t block_size, uint32_t lds_bytes, bool context_sensitive, int* blocks_per_cu);
hipError_t pga_launch_mega(const pga::KArgs& a, uint32_t blocks, uint32_t block_size, hipStream_t s);
// true for the kernels whose workgroups wait on one another (grid barrier):
// a graph must launch them cooperatively
bool pga_needs_coresidency(const void* fn);
hipError_t pga_tape_occupancy(uint32_t lds_bytes, int* blocks_per_cu);
hipError_t pga_launch_mega_tape(const pga::KArgs& a, uint32_t blocks, hipStream_t s);
// boxes of the colliders in the store (count in *count, at most cap), for the
// BVH built between phases
hipError_t pga_launch_collider_boxes(const float4* coll, uint32_t cap, const uint32_t* count, float* aabbs,
                                     hipStream_t s);
// the LBVH build over a device count word: leaves min(*nd, capacity), scene
// box from the device, then the tree in depth-first order with skip indices
// (2 * capacity - 1 nodes of two float4 in `flat`; pga_lbvh.hip)
hipError_t pga_lbvh_enqueue_counted(pga_lbvh* b, const float* aabbs, const uint32_t* nd, uint64_t* keys,
                                    int32_t* left, int32_t* right, int32_t* parent, int32_t* leaf_parent,
                                    float* node_aabb, float4* flat);
hipError_t pga_launch_adjoint(const pga::KArgs& a, const pga::AdjArgs& g, hipStream_t s);
hipError_t pga_launch_grad_out(const double* acc, float* out, uint32_t n, hipStream_t s);
#endif
