/*
 * pga.h - C ABI of the MI355X-native shape-grammar derivation engine.
 *
 * The ABI replaces the derivation path of the reference's `pga` C++ API:
 *   - PGA::SinglePhaseEvaluator<ProcList, ...>::initialize/execute/release
 *     (include/pga/core/GPUEvaluator.cuh:475-536, CPUEvaluator.h:393-459)
 *   - the interpreter facade initializePGA/executePGA/releasePGA/destroyPGA
 *     (src/interpreter/PGAFacade.h:17-21, src/interpreter/PGAFacade.cu:897-923)
 *   - the dispatch table (include/pga/core/DispatchTable.h:16-104,
 *     DispatchTableEntry.h:8-36) and procedure list
 *     (include/pga/compiler/ProcedureList.h:18-69).
 *
 * Plain pointers and sizes only: no HIP or torch types cross this boundary.
 * Device pointers are passed as `void*`/typed pointers that the caller
 * obtained from hipMalloc (or torch); a stream is passed as `void*`
 * holding a hipStream_t.
 */
#ifndef PGA_H
#define PGA_H

#ifndef __HIPCC_RTC__   /* runtime-compiled device code has no system headers */
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define PGA_API __attribute__((visibility("default")))
#else
#define PGA_API
#endif

/* ---- limits: PGA::Constants (include/pga/core/GlobalConstants.h:9-14) ---- */
#define PGA_MAX_NUM_PARAMETERS 17
#define PGA_MAX_NUM_PARAMETER_VALUES 16
#define PGA_MAX_NUM_SUCCESSORS 20
/* vertices of a dynamic polygon or right prism (Constants::MaxNumSides) */
#define PGA_MAX_NUM_SIDES 16
#define PGA_MAX_NUM_AXIOM_SEEDS 32
#define PGA_MAX_EXP_STACK 4
/* instanced meshes per scene (include/pga/rendering/RenderingConstants.h) */
#define PGA_MAX_TERMINALS 64
/* phases of a multi-phase program (MultiPhaseEvaluator<..., NumPhasesT, ...>,
 * GPUEvaluator.cuh:395-456): successor phase indices 0..15; the instance
 * meta then holds entry | depth << 16 | phase << 28, with depth < 4096 */
#define PGA_MAX_PHASES 16
/* a successor phase index that continues in the phase of the shape reaching
 * it: an operator chained inside one procedure, which the reference runs
 * without a dispatch (the C++ front end's lowering writes it; no counterpart
 * in the reference's table format). Other negative indices mean phase 0. */
#define PGA_PHASE_INHERIT (-2)

/* ---- status codes (the reference traps or throws; see each code) ---- */
typedef enum pga_status {
    PGA_OK = 0,
    PGA_ERR_INVALID_ARG = 1,
    PGA_ERR_OVERFLOW_QUEUE = 2,      /* queueDistLocks.cuh:65-70 trap */
    PGA_ERR_OVERFLOW_INSTANCES = 3,  /* GenerationFunctions.cuh:175 wrap */
    PGA_ERR_BAD_PROGRAM = 4,         /* invariant checks, Proc.cuh:40-49 */
    PGA_ERR_SUBDIV_OVERCONSTRAINED = 5, /* Subdivide.cuh:208-216 trap */
    PGA_ERR_REPEAT_BAD_EXTENT = 6,   /* Repeat.cuh:73-96 trap */
    PGA_ERR_HIP = 7,                 /* CUDAException.h:8-56 */
    PGA_ERR_MAX_DEPTH = 8,           /* CPUEvaluator MaxDerivationSteps */
    PGA_ERR_UNSUPPORTED = 9,
    PGA_ERR_NOT_INITIALIZED = 10
} pga_status;

/* ---- PGA::ParameterType (include/pga/core/ParameterType.h) ---- */
typedef enum pga_parameter_type {
    PGA_PT_SCALAR = 0,
    PGA_PT_RAND = 1,
    PGA_PT_SHAPE_ATTR = 2,
    PGA_PT_EXP = 3,
    PGA_PT_VEC2 = 4,
    PGA_PT_VEC4 = 5,
    /* project extensions (no reference counterpart):
     *  PHILOX      value[0]=min, value[1]=max, value[2]=stream id;
     *              min + (max-min)*u, u from Philox4x32-10 keyed by the
     *              context key, counter (axiom, path, depth, stream).
     *  AXIOM_PARAM value[0]=slot; the per-axiom parameter
     *              params[axiom * n_params + slot] (differentiable). */
    PGA_PT_PHILOX = 64,
    PGA_PT_AXIOM_PARAM = 65
} pga_parameter_type;

/* ---- expression operand/operation codes (OperandType.h, OperationType.h) */
typedef enum pga_operand_type {
    PGA_ORT_SCALAR = 0,
    PGA_ORT_RAND = 1,
    PGA_ORT_SHAPE_ATTR = 2,
    PGA_ORT_OP = 3
} pga_operand_type;

typedef enum pga_operation_type {
    PGA_OPT_ADD = 0, PGA_OPT_SUB, PGA_OPT_DIV, PGA_OPT_MULTI,
    PGA_OPT_EQ, PGA_OPT_NEQ, PGA_OPT_LT, PGA_OPT_GT,
    PGA_OPT_LEQ, PGA_OPT_GEQ, PGA_OPT_AND, PGA_OPT_OR
} pga_operation_type;

/* ---- PGA::Compiler::OperatorType (include/pga/compiler/OperatorType.h) */
typedef enum pga_operator_type {
    PGA_OP_TRANSLATE = 1,
    PGA_OP_ROTATE = 2,
    PGA_OP_SCALE = 3,
    PGA_OP_EXTRUDE = 4,
    PGA_OP_COMPSPLIT = 5,
    PGA_OP_SUBDIV = 6,
    PGA_OP_REPEAT = 7,
    PGA_OP_DISCARD = 8,
    PGA_OP_IF = 9,
    PGA_OP_IFSIZELESS = 10,
    PGA_OP_IFCOLLIDES = 11,
    PGA_OP_GENERATE = 12,
    PGA_OP_STOCHASTIC = 13,
    PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON = 14,
    PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM = 15,
    PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON = 16,
    PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM = 17,
    PGA_OP_COLLIDER = 18,
    PGA_OP_SWAPSIZE = 19,
    PGA_OP_REPLICATE = 20
} pga_operator_type;

/* ---- PGA::Compiler::ShapeType (include/pga/compiler/ShapeType.h) ---- */
typedef enum pga_shape_type {
    PGA_SHAPE_TRIANGLE = 0,
    PGA_SHAPE_QUAD = 1,
    PGA_SHAPE_DYNAMIC_CONVEX_POLYGON = 6,
    PGA_SHAPE_DYNAMIC_POLYGON = 7,
    PGA_SHAPE_BOX = 9,
    PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM = 14,
    PGA_SHAPE_DYNAMIC_RIGHT_PRISM = 15
} pga_shape_type;

/* ---- PGA::RepeatMode (include/pga/core/RepeatMode.h) ---- */
typedef enum pga_repeat_mode {
    PGA_ANCHOR_TO_START = 0,
    PGA_ANCHOR_TO_END = 1,
    PGA_ADJUST_TO_FILL = 2
} pga_repeat_mode;

/* ---- PGA::GPU::Technique (include/pga/core/GPUTechnique.h) ---- */
typedef enum pga_technique {
    PGA_TECHNIQUE_KERNELS = 0,    /* one launch per derivation level, host loop */
    PGA_TECHNIQUE_DYN_PAR = 1,    /* not available on HIP: PGA_ERR_UNSUPPORTED */
    PGA_TECHNIQUE_MEGAKERNEL = 2, /* one persistent launch for the whole derivation */
    PGA_TECHNIQUE_MEGAKERNEL_LOCAL_QUEUES = 3 /* alias of MEGAKERNEL */
} pga_technique;

/* Byte-identical to PGA::DispatchTableEntry (DispatchTableEntry.h:8-36):
 * 1420 bytes, so a reference caller can pass
 * DispatchTable::toDispatchTableEntriesPtr().get() unchanged. */
typedef struct pga_dispatch_parameter {
    float value[PGA_MAX_NUM_PARAMETER_VALUES];
    int32_t type; /* pga_parameter_type */
} pga_dispatch_parameter;

typedef struct pga_dispatch_successor {
    int32_t entry_index;
    int32_t phase_index;
} pga_dispatch_successor;

typedef struct pga_dispatch_entry {
    int32_t operator_code; /* index into the procedure list */
    uint32_t num_parameters;
    pga_dispatch_parameter parameters[PGA_MAX_NUM_PARAMETERS];
    uint32_t num_successors;
    uint8_t rule_tag_id;
    pga_dispatch_successor successors[PGA_MAX_NUM_SUCCESSORS];
    uint32_t num_edge_indices;
    int32_t edge_indices[PGA_MAX_NUM_SUCCESSORS];
    int32_t sub_graph_index;
} pga_dispatch_entry;

/* One procedure of the procedure list: PGA::Compiler::SingleOperatorProcedure
 * (ProcedureList.h:18-40) plus what the Proc<> template carries.
 * flags: PGA_PROC_STATIC marks the static-parameter form of an operator
 * (Subdivide<Axis, Pair<F,Next>...>, RandomRule<Chance, Pair<P,Next>...>)
 * whose semantics differ from the <DynParams> form (Subdivide.cuh:22-168 vs
 * 170-248; RandomRule.cuh:20-111 vs 121-147). PGA_PROC_PARALLEL marks the
 * <true> form of ComponentSplit, one face per thread through
 * ParallelFaceSplitter<Box> (ComponentSplit.cuh:179-230, 269-328: face
 * frames from normal and up vectors, seed modifier (face + 2) % 6); the
 * <true> form of Repeat traps in the reference (Repeat.cuh:35-61) and is
 * refused with PGA_ERR_UNSUPPORTED. */
#define PGA_PROC_STATIC 1u
#define PGA_PROC_PARALLEL 2u
typedef struct pga_procedure {
    int32_t op_type;      /* pga_operator_type */
    int32_t shape_type;   /* pga_shape_type of the shapes the queue holds */
    int32_t gen_func_idx; /* Generate only: 1 = instanced triangle mesh */
    uint32_t flags;
} pga_procedure;

/* An axiom: the shape record an AxiomGenerator would dispatch
 * (Grid.cuh:188-206, src/interpreter/PGAFacade.cu:697-729).
 * model is the row-major 3x4 model matrix (translation in column 3). */
typedef struct pga_axiom {
    float model[12];
    float size[3];
    float seed;
    float custom;
    int32_t entry_index; /* -1: skipped (PGAFacade.cu:708) */
    int32_t shape_type;  /* pga_shape_type */
    uint32_t reserved;
} pga_axiom; /* 80 bytes */

/* A terminal instance (64 bytes). model holds rows 0..2 of
 * model4 * scale(size) (GenerationFunctions.cuh:180-184, row 3 is 0 0 0 1);
 * the reference stores its transpose as a float4x4 plus `custom` (80 B,
 * InstancedTriangleMeshData.h:9-26). axiom/path/meta identify the derivation
 * path that produced the instance: meta = (depth << 16) | generate entry. */
typedef struct pga_instance {
    float model[12];
    float custom;
    uint32_t axiom;
    uint32_t path;
    uint32_t meta;
} pga_instance;

typedef struct pga_config {
    int32_t technique;       /* pga_technique */
    uint32_t queue_capacity; /* shape records per frontier buffer (0: default) */
    uint32_t block_size;     /* threads per workgroup: 256, 512, 768 or 1024 (0: 512) */
    uint32_t max_levels;     /* derivation-level guard (0: 4096) */
    uint32_t grid_blocks;    /* persistent grid size (0: occupancy-derived) */
    int32_t device;          /* HIP device ordinal (-1: current) */
    void* stream;            /* hipStream_t, NULL: the context creates one */
} pga_config;

typedef struct pga_stats {
    uint64_t levels;          /* derivation levels executed (with workgroup-local levels: the
                                 passes of the busiest workgroup plus the grid-wide levels) */
    uint64_t records;         /* shape records written to frontiers */
    uint64_t terminals;       /* instances written, all terminal types */
    double seconds;           /* device time of the last execute */
} pga_stats;

typedef struct pga_context pga_context;

/* -- context (replaces SinglePhaseEvaluator construction/release) -- */
PGA_API pga_status pga_create(const pga_config* config, pga_context** out);
PGA_API void pga_destroy(pga_context* ctx);
PGA_API const char* pga_last_error(const pga_context* ctx);
PGA_API const char* pga_status_string(pga_status status);
PGA_API uint32_t pga_abi_version(void);

/* -- program: procedure list + dispatch table
 *    (SinglePhaseEvaluator::initialize, GPUEvaluator.cuh:523-528;
 *     initializePGA, src/interpreter/PGAFacade.cu:904-911) -- */
PGA_API pga_status pga_initialize(pga_context* ctx,
                                  const pga_procedure* procedures, uint32_t num_procedures,
                                  const pga_dispatch_entry* entries, uint32_t num_entries);

/* -- axioms (AxiomGenerator::generateAxiom, Grid.cuh:236-284;
 *    Device::setAxioms, src/interpreter/GlobalVariables.cuh:47-59) -- */
PGA_API pga_status pga_set_axioms(pga_context* ctx, const pga_axiom* axioms, uint32_t num_axioms);
/* the caller keeps `device_axioms` alive until the next execute returns */
PGA_API pga_status pga_set_axioms_device(pga_context* ctx, const pga_axiom* device_axioms,
                                         uint32_t num_axioms);
/* PGA::Device::setSeeds (GlobalVariables.cuh:28-32): copies
 * (n % 33) seeds into a 32-slot table, as the reference does. */
PGA_API pga_status pga_set_seeds(pga_context* ctx, const float* seeds, uint32_t num_seeds);
PGA_API pga_status pga_get_seeds(const pga_context* ctx, float* seeds32);
/* per-axiom continuous parameters for PGA_PT_AXIOM_PARAM (host memory) */
PGA_API pga_status pga_set_axiom_params(pga_context* ctx, const float* params,
                                        uint32_t num_axioms, uint32_t params_per_axiom);
PGA_API pga_status pga_set_philox_key(pga_context* ctx, uint32_t key0, uint32_t key1);

/* -- terminal output (Rendering::Device::setInstancedTriangleMeshes,
 *    RenderingGlobalVariables.cuh:26-59): caller-owned device buffers -- */
PGA_API pga_status pga_bind_instances(pga_context* ctx, uint32_t terminal,
                                      pga_instance* device_buffer, uint64_t capacity);
PGA_API pga_status pga_get_counts(pga_context* ctx, uint64_t* counts, uint32_t num_terminals);

/* -- derivation (SinglePhaseEvaluator::execute, GPUEvaluator.cuh:508-521;
 *    executePGA, src/interpreter/PGAFacade.cu:921-923).
 *    Derives axioms [first, first + count) of the bound axiom array and
 *    appends instances after the counts left by earlier calls when
 *    `append` is nonzero (counts start at zero otherwise). -- */
PGA_API pga_status pga_execute(pga_context* ctx, uint32_t num_axioms, double* seconds);
PGA_API pga_status pga_execute_range(pga_context* ctx, uint32_t first_axiom, uint32_t num_axioms,
                                     int32_t append, double* seconds);
PGA_API pga_status pga_get_stats(const pga_context* ctx, pga_stats* stats);

/* -- context sensitivity (MultiPhaseEvaluator::execute,
 *    GPUEvaluator.cuh:416-440; Collider.cuh:26-42, IfCollides.cuh:26-58).
 *    A program whose successors name phases 1..15 runs one derivation per
 *    phase inside each execute: phase p emits the terminals of the shapes
 *    dispatched into phase p, and Collider stores shapes that IfCollides in
 *    later phases tests against (an oriented-box separating-axis test,
 *    Collision.cuh). Such programs, and any program with Collider or
 *    IfCollides, run on the interpreting kernels only: pga_specialize,
 *    pga_forward_tape and graph capture return PGA_ERR_UNSUPPORTED.
 *    pga_set_collider_capacity: colliders one execute may store (0: 2^20,
 *    80 bytes each; more raise PGA_ERR_OVERFLOW_QUEUE). pga_num_colliders:
 *    the colliders the last execute stored and the program's phases. -- */
PGA_API pga_status pga_set_collider_capacity(pga_context* ctx, uint32_t capacity);
PGA_API pga_status pga_num_colliders(const pga_context* ctx, uint32_t* colliders, uint32_t* phases);

/* -- asynchronous form: enqueues the derivation on the context stream and
 *    returns without waiting (MEGAKERNEL technique only). pga_synchronize
 *    waits, then reports the device error word and refreshes the stats;
 *    pga_kernel_seconds gives the device time of the last derivation
 *    kernel alone (events around the persistent launch). -- */
PGA_API pga_status pga_execute_async(pga_context* ctx, uint32_t first_axiom, uint32_t num_axioms,
                                     int32_t append);
PGA_API pga_status pga_synchronize(pga_context* ctx);
PGA_API pga_status pga_kernel_seconds(pga_context* ctx, double* seconds);

/* -- on-device autotuner (stands in for the reference's offline
 *    auto-tuner, which times operator-graph partitions and writes their
 *    configuration into generated code, auto_tuner/: here the candidates are
 *    the launch configurations of this engine, timed on the device with the
 *    caller's own axioms and bindings). Times every technique x workgroup
 *    size x grid size, keeps the fastest for the context and reports the
 *    largest frontiers seen, from which a caller sizes queue_capacity. The
 *    timed derivations write the bound instance buffers. -- */
typedef struct pga_tune_result {
    uint32_t technique;            /* chosen technique */
    uint32_t block_size;           /* chosen workgroup size */
    uint32_t grid_blocks;          /* chosen workgroups per launch */
    uint32_t candidates;           /* configurations timed */
    double seconds;                /* best device time of the chosen configuration */
    double initial_seconds;        /* best device time of the configuration before tuning */
    uint64_t max_level_parents;    /* largest number of parent records written by one level */
    uint64_t max_level_children;   /* largest number of children of one level */
    uint32_t specialized;          /* 1: the chosen kernel is the one specialised to the program */
    uint32_t local_grab;           /* specialised kernel: axioms per workgroup hand-out of its
                                      workgroup-local levels (pga_set_schedule), 0: grid-wide levels */
    uint32_t local_spill;          /* ... and their spill limit (0: the arrays' capacity) */
    uint32_t reserved;
} pga_tune_result;
PGA_API pga_status pga_autotune(pga_context* ctx, uint32_t num_axioms, uint32_t repeats, pga_tune_result* out);

/* -- kernels specialised to the loaded program. The reference compiles each
 *    grammar into its own kernels (the C++ idiom's templates, the code the
 *    analyzer generates per partition, CodeGenerator.cpp:1691-1722, rebuilt
 *    with nvcc). Here the library writes the loaded program as C++ - one
 *    type of compile-time constants per dispatch entry - and compiles it at
 *    run time (hipRTC) with the persistent derivation kernel, workgroups of
 *    block_size threads (a multiple of 64 up to 1024; 0: 512). enable != 0 makes
 *    the MEGAKERNEL technique run it (eager, asynchronous and captured
 *    derivations) and pga_autotune time it beside the interpreting kernels;
 *    enable = 0 goes back to the interpreting kernel. The instances are
 *    those of the interpreting kernel. pga_initialize drops the
 *    specialisation. seconds (optional): the compile time, 0 when the code
 *    object was compiled before in this process. -- */
PGA_API pga_status pga_specialize(pga_context* ctx, uint32_t block_size, int32_t enable, double* seconds);
/* Level scheduling of the specialised kernel. local != 0 (the default): a
 * workgroup takes `grab` axioms at a time (0: an equal share per
 * workgroup, at most 64) whenever its level is narrower than the workgroup,
 * and derives what it holds level by level on its own - the same passes,
 * but a workgroup scan in place of the grid scan and no grid barrier -
 * until a level's fan-outs outgrow its arrays (two parents and eight
 * children per thread, at least 512 / 2048, at most the queue capacity) or
 * number more than `spill` children (0: no such limit); those, and every
 * later level of them, run over the whole grid.
 * local = 0: every level runs over the whole grid. The instances are the
 * same multiset either way; with local levels their order within a
 * terminal buffer differs from run to run. No counterpart in the reference,
 * whose Whippletree queues hand shapes out one procedure queue at a time
 * (queuingPerProc.cuh:524-580); this is its knob for small and uneven
 * batches, and pga_autotune tries both. */
PGA_API pga_status pga_set_schedule(pga_context* ctx, int32_t local, uint32_t grab, uint32_t spill);
/* The same compilation without a device or context: validates the program
 * as pga_initialize does, writes its specialised kernel and compiles it.
 * log (optional) receives the validation error or the compiler's messages,
 * at most log_size bytes. For tests and tools that check a program before a
 * device is at hand. */
PGA_API pga_status pga_specialize_check(const pga_procedure* procedures, uint32_t num_procedures,
                                        const pga_dispatch_entry* entries, uint32_t num_entries, uint32_t block_size,
                                        char* log, uint32_t log_size, double* seconds);

/* -- instrumentation (the reference's clock64 edge/procedure timers,
 *    Instrumentation.cuh:47-76, become phase timestamps): when a device
 *    buffer is set, workgroup 0 of the derivation kernel writes the device
 *    wall clock (ticks of pga_timestamp_hz) at every phase boundary of every
 *    level: [level * 4 + 0] level start, + 1 count pass done, + 2 prefix
 *    sums done, + 3 write pass done; up to capacity / 4 levels. NULL turns
 *    it off. -- */
PGA_API pga_status pga_set_timestamps(pga_context* ctx, uint64_t* device_buffer, uint32_t capacity);
PGA_API double pga_timestamp_hz(pga_context* ctx);

/* -- reverse mode (no counterpart in the reference: the derivation tape and
 *    adjoint of a differentiable build). pga_tape_reserve sizes the tape:
 *    parent records of all levels, and children of all levels counting the
 *    axioms. pga_forward_tape derives the bound axioms [0, n) recording the
 *    tape (every fan-out materialised, the topology frozen) into the bound
 *    instance buffers; pga_backward takes a host array of PGA_MAX_TERMINALS
 *    device pointers (NULL: no gradient) to dL/dinstance in the instance
 *    layout (16 floats per instance, 16-byte aligned, the 12 model entries
 *    are used) and
 *    writes dL/dparams (device, n x params_per_axiom floats) for the
 *    PGA_PT_AXIOM_PARAM parameters. -- */
PGA_API pga_status pga_tape_reserve(pga_context* ctx, uint32_t max_records, uint32_t max_children);
PGA_API pga_status pga_forward_tape(pga_context* ctx, uint32_t num_axioms, double* seconds);
PGA_API pga_status pga_backward(pga_context* ctx, const float* const* dl_dinstances, float* dl_dparams,
                                double* seconds);
/* Generation of the recorded tape: a number that grows with every successful
 * pga_forward_tape, 0 while no tape is valid. pga_set_axioms*,
 * pga_set_axiom_params, pga_set_philox_key, pga_initialize and
 * pga_tape_reserve invalidate the tape (pga_backward then returns
 * PGA_ERR_NOT_INITIALIZED); a caller that keeps a forward's generation can
 * tell whether a later backward would replay that forward's tape. */
PGA_API pga_status pga_tape_generation(const pga_context* ctx, uint64_t* generation);
/* Size of the recorded tape: fan-out parent records of all levels, and
 * children of all levels counting the axioms (one outcome row each). */
PGA_API pga_status pga_tape_counts(const pga_context* ctx, uint64_t* parent_records, uint64_t* children);
/* The loss of the gradient fits: wc * sum |c_i - tc_i|^2 + ws * sum |s_i - ts_i|^2
 * over `count` instances (device, 16 floats each, 16-byte aligned), where
 * c_i is the translation column and s_i the column norms of the 3x3 part;
 * targets are device arrays of count x 3 floats. Writes the gradient rows
 * dL/dinstance (count x 16 floats, the layout pga_backward reads; columns
 * 12..15 zero) and adds the loss to *loss (one device double). One kernel
 * on `stream` (a hipStream_t, NULL: the default stream); no synchronisation.
 * No counterpart in the reference (it has no reverse mode). */
PGA_API pga_status pga_center_size_mse(const float* instances, const float* target_centers,
                                       const float* target_sizes, uint64_t count, float weight_centers,
                                       float weight_sizes, float* dl_dinstances, double* loss, void* stream);

/* -- linear BVH over terminal AABBs (GPUBVHConstructor.cuh:29-333): 30-bit
 *    Morton codes of the box centres normalised to a scene box, a sort of the
 *    unique keys (code << 32 | index), a Karras radix tree and a bottom-up
 *    refit. Children: >= 0 a leaf (sorted position; keys[leaf] & 0xFFFFFFFF
 *    is the box index), < 0 an internal node (negated index); node 0 is the
 *    root. All arrays are device memory: aabbs[n][6] (min xyz, max xyz),
 *    keys[n], left/right/parent[n-1], leaf_parent[n], node_aabb[n-1][6]. -- */
typedef struct pga_lbvh pga_lbvh;
PGA_API pga_status pga_lbvh_create(uint32_t capacity, int32_t device, void* stream, pga_lbvh** out);
PGA_API void pga_lbvh_destroy(pga_lbvh* bvh);
/* message of the builder's last failed call; bvh = NULL: of the last failed
 * pga_lbvh_build_noalloc on this thread */
PGA_API const char* pga_lbvh_last_error(const pga_lbvh* bvh);
/* box of each instance: the unit cube (planar = 0) or the unit square in the
 * z = 0 plane (planar = 1) under the instance's model rows */
PGA_API pga_status pga_lbvh_instance_aabbs(pga_lbvh* bvh, const pga_instance* instances, uint32_t n,
                                           int32_t planar, float* aabbs);
/* union of n boxes, to the host: min xyz, max xyz */
PGA_API pga_status pga_lbvh_bounds(pga_lbvh* bvh, const float* aabbs, uint32_t n, float bounds[6]);
PGA_API pga_status pga_lbvh_build(pga_lbvh* bvh, const float* aabbs, uint32_t n, const float scene_min[3],
                                  const float scene_max[3], uint64_t* keys, int32_t* left, int32_t* right,
                                  int32_t* parent, int32_t* leaf_parent, float* node_aabb);
/* device time of the last build (Morton codes to refit) */
PGA_API pga_status pga_lbvh_seconds(pga_lbvh* bvh, double* seconds);

/* The build of pga_lbvh_build with no builder and no allocation. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code recording a HIP graph, serving loops, harnesses that
 * forbid hipMalloc). The sort and refit scratch is a device array inside the
 * library, one per GPU, looked up with hipGetSymbolAddress on the first call
 * on that GPU and held until the process ends (about 24.3 MiB): at most
 * PGA_LBVH_NOALLOC_MAX_LEAVES leaves, larger n is refused with
 * PGA_ERR_INVALID_ARG (message: pga_lbvh_last_error(NULL)). The build is
 * enqueued on `stream` (a hipStream_t of the current device; NULL: the
 * default stream) and the call returns without waiting. Calls that overlap
 * on one GPU share the array and race: order them on one stream or wait in
 * between. Outputs as pga_lbvh_build. */
#define PGA_LBVH_NOALLOC_MAX_LEAVES (1u << 20)
PGA_API pga_status pga_lbvh_build_noalloc(const float* aabbs, uint32_t n, const float scene_min[3],
                                          const float scene_max[3], uint64_t* keys, int32_t* left, int32_t* right,
                                          int32_t* parent, int32_t* leaf_parent, float* node_aabb, void* stream);

/* -- device-count form (no counterpart in the reference, whose build takes a
 *    host count, GPUBVHConstructor.cuh:275-333): the leaf count stays in
 *    device memory, so a derivation, the gather of its terminals and the
 *    build run without a host round trip and can be recorded in one HIP
 *    graph (pga_capture_begin) on the stream the builder was created with.
 *    pga_lbvh_gather_instances writes the boxes of the instances of terminal
 *    types 0..n_types-1 (buffers/capacities as bound to the context; bit t of
 *    planar_mask: type t is planar) into aabbs (room for `capacity` boxes),
 *    concatenated in type order, reading the counts from device_counts
 *    (pga_device_counts). The leaf count is min(total, capacity); more is an
 *    overflow, reported by pga_lbvh_leaf_count. pga_lbvh_build_device builds
 *    over that count, sized for `capacity`, with the given scene box or, when
 *    both corners are NULL, the union of the boxes computed on the device. -- */
PGA_API pga_status pga_lbvh_gather_instances(pga_lbvh* bvh, const pga_instance* const* buffers,
                                             const uint64_t* capacities, uint32_t n_types,
                                             const uint64_t* device_counts, uint32_t planar_mask, float* aabbs);
PGA_API pga_status pga_lbvh_build_device(pga_lbvh* bvh, const float* aabbs, const float* scene_min,
                                         const float* scene_max, uint64_t* keys, int32_t* left, int32_t* right,
                                         int32_t* parent, int32_t* leaf_parent, float* node_aabb);
/* waits for the stream; the leaf count of the last gather and its overflow flag */
PGA_API pga_status pga_lbvh_leaf_count(pga_lbvh* bvh, uint32_t* n, int32_t* overflow);

/* -- HIP graphs of repeated derivations (no counterpart in the reference,
 *    which launches every derivation from the host, GPUEvaluator.cuh:508-521).
 *    Between pga_capture_begin and pga_capture_end, pga_execute_async calls
 *    (and any other work enqueued on the context stream, such as the
 *    device-count LBVH build) are recorded instead of run. pga_graph_launch
 *    replays the recording on the context stream without waiting;
 *    pga_synchronize then waits and reports errors and stats as for
 *    pga_execute_async. The graph holds the context's buffers, launch
 *    configuration and argument values as they were at capture: values in
 *    those buffers (axioms, axiom parameters, instance memory) may change
 *    between replays, but a replay after a change of bindings, program
 *    layout, queue capacity or configuration is refused (capture again).
 *    pga_device_counts gives the device address of the 64-bit terminal
 *    counters (PGA_MAX_TERMINALS), valid for the context's lifetime. -- */
typedef struct pga_graph pga_graph;
PGA_API pga_status pga_capture_begin(pga_context* ctx);
PGA_API pga_status pga_capture_end(pga_context* ctx, pga_graph** graph);
PGA_API pga_status pga_graph_launch(pga_context* ctx, pga_graph* graph);
/* A graph may outlive its context: pga_destroy detaches the context's graphs,
 * which can then only be destroyed. */
PGA_API void pga_graph_destroy(pga_graph* graph);
/* kernel nodes of the graph and how many of them launch cooperatively */
PGA_API pga_status pga_graph_info(const pga_graph* graph, uint32_t* kernel_nodes, uint32_t* cooperative_nodes);
PGA_API pga_status pga_device_counts(pga_context* ctx, const uint64_t** device_counts);
PGA_API pga_status pga_get_stream(pga_context* ctx, void** stream);

/* -- device memory for callers that link no HIP runtime themselves (the
 *    C++ front end in pga/core/Core.h allocates terminal buffers with
 *    these; the reference's evaluator called cudaMalloc in its headers,
 *    techniqueMegakernel.cuh:607). -- */
PGA_API pga_status pga_device_alloc(uint64_t bytes, void** device_ptr);
PGA_API pga_status pga_device_free(void* device_ptr);
PGA_API pga_status pga_copy_to_host(void* host_dst, const void* device_src, uint64_t bytes);

/* -- generation function 0 (triangle soup): terminals whose procedure uses
 *    generation function 0 are derived into instance buffers like
 *    instanced ones; pga_tessellate then writes the reference's geometry
 *    for `count` instances of one shape type (ShapeGenerator<Box, false>::run,
 *    include/pga/core/ShapeGenerator.cuh:376-511: 24 vertices, 36 indices
 *    per box; ShapeGenerator<Quad, false>::run, :670-708: 4 and 6 per quad;
 *    the default winding, PGA_CW undefined). Vertex records are
 *    TriangleMeshData::VertexAttributes (include/pga/rendering/TriangleMeshData.h:12-17).
 *    Instance i's V vertices are written at vertices[i * V] and its
 *    1.5 V indices at indices[i * 1.5 V]; an index holds
 *    base_vertex + i * V + corner, so calls that fill one mesh share its
 *    index space. Instances keep their buffer order. All pointers are
 *    device memory; `instances` 16-byte aligned, `vertices` and `indices`
 *    4-byte aligned (vertices on a 16-byte boundary are written with
 *    16-byte stores). One kernel on `stream`, no synchronisation. -- */
typedef struct pga_vertex {
    float position[4];
    float normal[3];
    float uv[2];
} pga_vertex; /* 36 bytes */
PGA_API pga_status pga_tessellate(const pga_instance* instances, uint64_t count, int32_t shape_type,
                                  pga_vertex* vertices, uint32_t* indices, uint32_t base_vertex, void* stream);
/* The same with the count read on the device: min(*device_count, capacity)
 *    shapes, e.g. a terminal's counter from pga_device_counts, so that a
 *    derivation and its tessellation run back to back on one stream with no
 *    host synchronisation and can be recorded in one graph (the reference's
 *    generation function 0 writes its mesh during the derivation). Buffers
 *    are sized for `capacity`; `device_count` 8-byte aligned. */
PGA_API pga_status pga_tessellate_device_count(const pga_instance* instances, const uint64_t* device_count,
                                               uint64_t capacity, int32_t shape_type, pga_vertex* vertices,
                                               uint32_t* indices, uint32_t base_vertex, void* stream);

/* -- grammar front end: the PGA grammar language (*.pga) compiled on the
 *    host into a dispatch table over the interpreter's procedure list.
 *    Replaces PGA::Compiler::Parser::parse + DispatchTableGenerator::fromRuleSet
 *    (src/compiler/Parser.cpp:494-540, src/compiler/DispatchTableGenerator.cpp:555-609)
 *    as driven by the interpreter (src/interpreter/PGAFacade.cu:600-660).
 *    pga_interpreter_procedures copies the interpreter's procedure list
 *    (src/interpreter/PGAFacade.cu:24-130: Quad and Box, then the dynamic
 *    polygons and prisms of the list built with INTERPRETER_SUPPORT_ALL_SHAPES;
 *    108 procedures) to out (capacity entries; out may be NULL to query the
 *    count). pga_compile_grammar always sets *out to a
 *    grammar object that the caller destroys; on error it holds no entries
 *    and pga_grammar_error gives the message (with the line number). The
 *    entries and axioms stay valid until pga_grammar_destroy. An axiom whose
 *    symbol has no rule gets entry_index -1 (skipped by the axiom loader). -- */
typedef struct pga_grammar pga_grammar;
typedef struct pga_grammar_axiom {
    int32_t entry_index;
    int32_t shape_type; /* pga_shape_type */
} pga_grammar_axiom;
PGA_API pga_status pga_interpreter_procedures(pga_procedure* out, uint32_t capacity, uint32_t* count);
PGA_API pga_status pga_compile_grammar(const char* source, pga_grammar** out);
PGA_API const char* pga_grammar_error(const pga_grammar* grammar);
PGA_API pga_status pga_grammar_entries(const pga_grammar* grammar, const pga_dispatch_entry** entries,
                                       uint32_t* count);
PGA_API pga_status pga_grammar_axioms(const pga_grammar* grammar, const pga_grammar_axiom** axioms,
                                      uint32_t* count);
PGA_API void pga_grammar_destroy(pga_grammar* grammar);

/* -- partition code (the operator-graph codegen): the grammar's
 *    operator graph, as Context builds it from the first axiom
 *    (src/compiler/Context.cpp:11-253; the library numbers the edges in
 *    depth-first pre-order), cut by `partition` (one '0' keep / '1' cut per
 *    edge; NULL cuts every edge, as the interpreter does,
 *    src/interpreter/InterpreterApp.cpp:214-223), written as C++ in the
 *    form CodeGenerator::fromPartition emits (src/compiler/CodeGenerator.cpp:
 *    1691-1722): Proc<Shape, operator tree, 1> structs P<k>, the procedure
 *    list, a DispatchTable and NumPhases, in `namespace name`. Uncut edges
 *    are inlined; cut edges become FSCall<procedure, phase, -1, entry>.
 *    Phases follow Graph::computePhases (src/compiler/Graph.cpp:149-272):
 *    an IfCollides(tag) runs one phase after the latest Collider(tag), every
 *    other operator in its predecessor's phase. Edges that close a cycle or
 *    enter such an IfCollides must be cut: their flags from
 *    pga_grammar_edges are 1 (cycle) | 2 (phase). Host code only. The result
 *    object holds the code (pga_grammar_code), the edge flags and an error
 *    message; destroy it with pga_grammar_destroy. -- */
PGA_API pga_status pga_generate_partition(const char* source, const char* partition, const char* name,
                                          pga_grammar** out);
PGA_API pga_status pga_grammar_code(const pga_grammar* grammar, const char** code, uint64_t* length);
PGA_API pga_status pga_grammar_edges(const pga_grammar* grammar, uint32_t* num_edges, const uint8_t** cycle_flags);

#ifdef __cplusplus
}
#endif

#endif /* PGA_H */
