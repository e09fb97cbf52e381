/*
 * pga_oracle.h - TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's shape-grammar derivation
 * (PGA::CPU::SinglePhaseEvaluator, include/pga/core/CPUEvaluator.h:393-459)
 * used as the parity oracle of the HIP engine. Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it; the product never does.
 *
 * Three twins run in lockstep on every shape:
 *   f32  - the reference's FP32 arithmetic in the reference's operation order
 *          (the bit-exact parity target);
 *   f64  - an FP64 re-derivation that replays the f32 run's discrete
 *          decisions (branches, repetition counts, rule choices), used for
 *          the 1e-5 relative transform check;
 *   adj  - an FP64 reverse-mode (tape) derivation with the same replayed
 *          decisions, whose adjoint gives dL/dparams for PGA_PT_AXIOM_PARAM.
 */
#ifndef PGA_ORACLE_H
#define PGA_ORACLE_H

#include "../include/pga.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PGAO_MAX_PHASES 16u   /* successor phase indices 0 .. 15 */
#define PGAO_TWIN_F64 1u
#define PGAO_TWIN_ADJ 2u

typedef struct pgao_result pgao_result;

/* Runs the derivation of `num_axioms` axioms. Returns a pga_status; on
 * success *out holds the result (free with pgao_free). */
int pgao_derive(const pga_procedure* procs, uint32_t num_procs,
                const pga_dispatch_entry* entries, uint32_t num_entries,
                const pga_axiom* axioms, uint32_t num_axioms,
                const float* axiom_params, uint32_t params_per_axiom,
                uint32_t philox_key0, uint32_t philox_key1,
                uint32_t twins, uint32_t max_steps,
                pgao_result** out);
void pgao_free(pgao_result* r);
const char* pgao_error(const pgao_result* r);
uint32_t pgao_num_terminals(const pgao_result* r);
/* Shapes the Collider operator stored over all phases. */
uint32_t pgao_num_colliders(const pgao_result* r);
/* Collision::check(stored, asking) of two Box/Quad shapes (model 3x4
 * row-major, size, PGA_SHAPE_*): 1 when they collide. */
int pgao_shapes_collide(const float* am, const float* asz, int32_t at, const float* bm, const float* bsz, int32_t bt);
uint64_t pgao_count(const pgao_result* r, uint32_t terminal);
uint64_t pgao_steps(const pgao_result* r);
uint64_t pgao_executions(const pgao_result* r);
/* FP32 instances in derivation order (the reference CPU order) */
int pgao_copy_instances(const pgao_result* r, uint32_t terminal, pga_instance* dst);
/* Generation function 0 (triangle soup, ShapeGenerator.cuh:376-511 Box,
 * 670-708 Quad): the vertices of each such terminal's shapes, in derivation
 * order, 9 floats per vertex (position xyzw, normal xyz, uv), 24 vertices
 * per box and 4 per quad; the index pattern is implied (two triangles
 * (0,1,2) (0,2,3) per face of four vertices). */
uint64_t pgao_mesh_floats(const pgao_result* r, uint32_t terminal);
int pgao_copy_mesh(const pgao_result* r, uint32_t terminal, float* dst);
/* FP64 twin: 13 doubles per instance (model[12], custom) */
int pgao_copy_instances_f64(const pgao_result* r, uint32_t terminal, double* dst);
/* Reverse sweep of the adj twin: dl_dmodel holds 12 doubles per instance of
 * each terminal type in derivation order (terminals concatenated in
 * terminal-index order). Writes num_axioms * params_per_axiom doubles. */
int pgao_backward(pgao_result* r, const double* dl_dmodel, double* dl_dparams);

/* Restated primitives, exported so tests can pin them one by one. */
float pgao_next_seed(float base, uint32_t modifier);         /* Random.cuh:26-49 */
float pgao_range(float seed, float mn, float mx);             /* Random.cuh:51-57 */
uint32_t pgao_seed_component(uint32_t z, uint32_t which);    /* Random.cuh:15-24, 36-43: 0..2 Tausworthe, 3 LCG */
void pgao_seed_state(float base, uint32_t modifier, uint32_t* z4); /* Random.cuh:26-35: z1..z4 */
float pgao_seed_unit(uint32_t t);                            /* Random.cuh:44-48 */
void pgao_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float pgao_sinf(float x);
float pgao_cosf(float x);
uint32_t pgao_path_mix(uint32_t path, uint32_t ordinal);
float pgao_pack_operand(int type, float x);                  /* PackUtils.h:18-23 */
float pgao_pack_float2(float x, float y);                    /* PackUtils.h:25-28 */

/* LBVH over n AABBs (6 floats each: min xyz, max xyz), keys from the AABB
 * centers normalised into [scene_min, scene_max] (GPUBVHConstructor.cuh:
 * 44-86, 88-261). Outputs: keys[n] sorted 64-bit keys, left/right/parent for
 * the n-1 internal nodes (child >= 0: leaf at that sorted position,
 * child < 0: internal node -child; root parent = -1), leaf_parent[n],
 * node_aabb[(n-1)*6]. */
int pgao_lbvh(const float* aabbs, uint32_t n, const float scene_min[3], const float scene_max[3],
              uint64_t* keys, int32_t* left, int32_t* right, int32_t* parent,
              int32_t* leaf_parent, float* node_aabb);
uint32_t pgao_morton30(float x, float y, float z);

#ifdef __cplusplus
}
#endif

#endif
