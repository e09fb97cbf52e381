"""ctypes mirrors of the C ABI structs in include/pga.h.

The layouts are checked against the C header's sizes by the CPU test suite
(tests/test_abi.py). `pga_dispatch_entry` is byte-identical to the
reference's PGA::DispatchTableEntry (include/pga/core/DispatchTableEntry.h:8-36).
"""
import ctypes as C

import numpy as np

MAX_NUM_PARAMETERS = 17
MAX_NUM_PARAMETER_VALUES = 16
MAX_NUM_SUCCESSORS = 20
MAX_NUM_AXIOM_SEEDS = 32
MAX_TERMINALS = 64
MAX_PHASES = 16
PHASE_INHERIT = -2    # successor phase: the phase of the shape reaching it
MAX_EXP_STACK = 4

# pga_status
OK = 0
ERR_INVALID_ARG = 1
ERR_OVERFLOW_QUEUE = 2
ERR_OVERFLOW_INSTANCES = 3
ERR_BAD_PROGRAM = 4
ERR_SUBDIV_OVERCONSTRAINED = 5
ERR_REPEAT_BAD_EXTENT = 6
ERR_HIP = 7
ERR_MAX_DEPTH = 8
ERR_UNSUPPORTED = 9
ERR_NOT_INITIALIZED = 10

STATUS_NAMES = {
    0: "OK", 1: "INVALID_ARG", 2: "OVERFLOW_QUEUE", 3: "OVERFLOW_INSTANCES",
    4: "BAD_PROGRAM", 5: "SUBDIV_OVERCONSTRAINED", 6: "REPEAT_BAD_EXTENT",
    7: "HIP_ERROR", 8: "MAX_DEPTH", 9: "UNSUPPORTED", 10: "NOT_INITIALIZED",
}

# PGA::ParameterType (include/pga/core/ParameterType.h) + project extensions
PT_SCALAR, PT_RAND, PT_SHAPE_ATTR, PT_EXP, PT_VEC2, PT_VEC4 = 0, 1, 2, 3, 4, 5
PT_PHILOX = 64
PT_AXIOM_PARAM = 65

# PGA::OperandType / OperationType
ORT_SCALAR, ORT_RAND, ORT_SHAPE_ATTR, ORT_OP = 0, 1, 2, 3
(OPT_ADD, OPT_SUB, OPT_DIV, OPT_MULTI, OPT_EQ, OPT_NEQ, OPT_LT, OPT_GT,
 OPT_LEQ, OPT_GEQ, OPT_AND, OPT_OR) = range(12)

# PGA::Compiler::OperatorType (include/pga/compiler/OperatorType.h)
TRANSLATE, ROTATE, SCALE, EXTRUDE, COMPSPLIT, SUBDIV, REPEAT, DISCARD, IF, \
    IFSIZELESS, IFCOLLIDES, GENERATE, STOCHASTIC, SET_AS_DYNAMIC_CONVEX_POLYGON, \
    SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, SET_AS_DYNAMIC_CONCAVE_POLYGON, \
    SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, COLLIDER, SWAPSIZE, REPLICATE = range(1, 21)

# PGA::Compiler::ShapeType (include/pga/compiler/ShapeType.h)
TRIANGLE, QUAD, BOX = 0, 1, 9
DYNAMIC_CONVEX_POLYGON, DYNAMIC_POLYGON = 6, 7
DYNAMIC_CONVEX_RIGHT_PRISM, DYNAMIC_RIGHT_PRISM = 14, 15
MAX_NUM_SIDES = 16

# PGA::RepeatMode
ANCHOR_TO_START, ANCHOR_TO_END, ADJUST_TO_FILL = 0, 1, 2

# PGA::Axis
X, Y, Z = 0, 1, 2

# PGA::GPU::Technique
KERNELS, DYN_PAR, MEGAKERNEL, MEGAKERNEL_LOCAL_QUEUES = 0, 1, 2, 3

PROC_STATIC = 1
PROC_PARALLEL = 2


class DispatchParameter(C.Structure):
    _fields_ = [("value", C.c_float * MAX_NUM_PARAMETER_VALUES), ("type", C.c_int32)]


class DispatchSuccessor(C.Structure):
    _fields_ = [("entry_index", C.c_int32), ("phase_index", C.c_int32)]


class DispatchEntry(C.Structure):
    _fields_ = [
        ("operator_code", C.c_int32),
        ("num_parameters", C.c_uint32),
        ("parameters", DispatchParameter * MAX_NUM_PARAMETERS),
        ("num_successors", C.c_uint32),
        ("rule_tag_id", C.c_uint8),
        ("successors", DispatchSuccessor * MAX_NUM_SUCCESSORS),
        ("num_edge_indices", C.c_uint32),
        ("edge_indices", C.c_int32 * MAX_NUM_SUCCESSORS),
        ("sub_graph_index", C.c_int32),
    ]


class Procedure(C.Structure):
    _fields_ = [("op_type", C.c_int32), ("shape_type", C.c_int32),
                ("gen_func_idx", C.c_int32), ("flags", C.c_uint32)]


class Axiom(C.Structure):
    _fields_ = [("model", C.c_float * 12), ("size", C.c_float * 3), ("seed", C.c_float),
                ("custom", C.c_float), ("entry_index", C.c_int32), ("shape_type", C.c_int32),
                ("reserved", C.c_uint32)]


class GrammarAxiom(C.Structure):
    _fields_ = [("entry_index", C.c_int32), ("shape_type", C.c_int32)]


class Instance(C.Structure):
    _fields_ = [("model", C.c_float * 12), ("custom", C.c_float), ("axiom", C.c_uint32),
                ("path", C.c_uint32), ("meta", C.c_uint32)]


class Vertex(C.Structure):
    _fields_ = [("position", C.c_float * 4), ("normal", C.c_float * 3), ("uv", C.c_float * 2)]


class Config(C.Structure):
    _fields_ = [("technique", C.c_int32), ("queue_capacity", C.c_uint32),
                ("block_size", C.c_uint32), ("max_levels", C.c_uint32),
                ("grid_blocks", C.c_uint32), ("device", C.c_int32), ("stream", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("levels", C.c_uint64), ("records", C.c_uint64),
                ("terminals", C.c_uint64), ("seconds", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("technique", C.c_uint32), ("block_size", C.c_uint32), ("grid_blocks", C.c_uint32),
                ("candidates", C.c_uint32), ("seconds", C.c_double), ("initial_seconds", C.c_double),
                ("max_level_parents", C.c_uint64), ("max_level_children", C.c_uint64),
                ("specialized", C.c_uint32), ("local_grab", C.c_uint32),
                ("local_spill", C.c_uint32), ("reserved", C.c_uint32)]


# numpy views of the 80-byte axiom and 64-byte instance records
AXIOM_DTYPE = np.dtype([("model", "<f4", (12,)), ("size", "<f4", (3,)), ("seed", "<f4"),
                        ("custom", "<f4"), ("entry_index", "<i4"), ("shape_type", "<i4"),
                        ("reserved", "<u4")])
INSTANCE_DTYPE = np.dtype([("model", "<f4", (12,)), ("custom", "<f4"), ("axiom", "<u4"),
                           ("path", "<u4"), ("meta", "<u4")])
assert AXIOM_DTYPE.itemsize == C.sizeof(Axiom) == 80
assert INSTANCE_DTYPE.itemsize == C.sizeof(Instance) == 64
assert C.sizeof(DispatchEntry) == 1420
