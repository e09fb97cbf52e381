"""Host-side mirror of the reference's rule-data API.

`DispatchTable`, `Entry`, `Parameter`, `Successor` follow
PGA::DispatchTable (include/pga/core/DispatchTable.h:16-104): same names,
same argument meaning, same limits and the same errors on overflow.
`ProcedureList` follows PGA::Compiler::ProcedureList
(include/pga/compiler/ProcedureList.h:42-69): a repeated procedure raises,
`index_of` returns -1 when absent.
"""
import ctypes as C
import struct

import numpy as np

from . import abi


class Parameter:
    """PGA::DispatchTable::Parameter (DispatchTable.h:18-37)."""

    def __init__(self, type=abi.PT_SCALAR, values=()):
        self.type = int(type)
        self.values = [float(v) for v in values]

    def fill(self, p):
        if len(self.values) > abi.MAX_NUM_PARAMETER_VALUES:
            raise RuntimeError("values.size() > PGA::Constants::MaxNumParameterValues")
        p.type = self.type
        for i in range(abi.MAX_NUM_PARAMETER_VALUES):
            p.value[i] = 0.0
        for i, v in enumerate(self.values):
            p.value[i] = np.float32(v)  # static_cast<float>(double), DispatchTable.h:35

    def __repr__(self):
        return f"Parameter({self.type}, {self.values})"


class Successor:
    """PGA::DispatchTable::Successor (DispatchTable.h:39-54)."""

    def __init__(self, entry_index=-1, phase_index=-1):
        self.entry_index = int(entry_index)
        self.phase_index = int(phase_index)


class Entry:
    """PGA::DispatchTable::Entry (DispatchTable.h:56-91)."""

    def __init__(self, operator_code=0, parameters=(), successors=(), edge_indices=(),
                 sub_graph_index=-1):
        self.operator_code = int(operator_code)
        self.parameters = [p if isinstance(p, Parameter) else Parameter(*p) for p in parameters]
        self.successors = [s if isinstance(s, Successor) else Successor(*s) for s in successors]
        self.edge_indices = list(edge_indices)
        self.sub_graph_index = int(sub_graph_index)

    def fill(self, e):
        e.operator_code = self.operator_code
        if len(self.parameters) > abi.MAX_NUM_PARAMETERS:
            raise RuntimeError("parameters.size() > PGA::Constants::MaxNumParameters")
        e.num_parameters = len(self.parameters)
        C.memset(C.addressof(e.parameters), 0, C.sizeof(e.parameters))
        for i, p in enumerate(self.parameters):
            p.fill(e.parameters[i])
        if len(self.successors) > abi.MAX_NUM_SUCCESSORS:
            raise RuntimeError("successors.size() > PGA::Constants::MaxNumSuccessors")
        e.num_successors = len(self.successors)
        C.memset(C.addressof(e.successors), 0xFF, C.sizeof(e.successors))
        for i, s in enumerate(self.successors):
            e.successors[i].entry_index = s.entry_index
            e.successors[i].phase_index = s.phase_index
        if len(self.edge_indices) > abi.MAX_NUM_SUCCESSORS:
            raise RuntimeError("edgeIndices.size() > PGA::Constants::MaxNumSuccessors")
        e.num_edge_indices = len(self.edge_indices)
        C.memset(C.addressof(e.edge_indices), 0xFF, C.sizeof(e.edge_indices))
        for i, v in enumerate(self.edge_indices):
            e.edge_indices[i] = v
        e.sub_graph_index = self.sub_graph_index
        e.rule_tag_id = 0


class DispatchTable:
    """PGA::DispatchTable (DispatchTable.h:93-104)."""

    def __init__(self, entries=()):
        self.entries = list(entries)

    def to_dispatch_table_entries(self):
        """toDispatchTableEntriesPtr(): a ctypes array of pga_dispatch_entry."""
        arr = (abi.DispatchEntry * max(1, len(self.entries)))()
        for i, e in enumerate(self.entries):
            e.fill(arr[i])
        return arr


class Procedure:
    """PGA::Compiler::SingleOperatorProcedure plus the Proc<> template's
    static-parameter flag (ProcedureList.h:18-40, Proc.cuh:20-27)."""

    def __init__(self, op_type, shape_type, gen_func_idx=0, flags=0):
        self.op_type = int(op_type)
        self.shape_type = int(shape_type)
        self.gen_func_idx = int(gen_func_idx)
        self.flags = int(flags)

    def key(self):
        return (self.op_type, self.shape_type, self.gen_func_idx, self.flags)

    def __eq__(self, other):
        return self.key() == other.key()

    def __hash__(self):
        return hash(self.key())


class ProcedureList:
    """PGA::Compiler::ProcedureList (ProcedureList.h:42-69)."""

    def __init__(self, procedures):
        self.procedures = []
        for p in procedures:
            p = p if isinstance(p, Procedure) else Procedure(*p)
            if p in self.procedures:
                raise RuntimeError("PGA::Compiler::ProcedureList::ctor(): repeated procedure")
            self.procedures.append(p)

    def index_of(self, procedure):
        procedure = procedure if isinstance(procedure, Procedure) else Procedure(*procedure)
        try:
            return self.procedures.index(procedure)
        except ValueError:
            return -1

    def __len__(self):
        return len(self.procedures)

    def to_array(self):
        arr = (abi.Procedure * max(1, len(self.procedures)))()
        for i, p in enumerate(self.procedures):
            arr[i].op_type = p.op_type
            arr[i].shape_type = p.shape_type
            arr[i].gen_func_idx = p.gen_func_idx
            arr[i].flags = p.flags
        return arr


class Program:
    """A procedure list and the dispatch table whose operator codes index it."""

    def __init__(self, procedures, table, name=""):
        self.procedures = procedures if isinstance(procedures, ProcedureList) else ProcedureList(procedures)
        self.table = table if isinstance(table, DispatchTable) else DispatchTable(table)
        self.name = name

    def arrays(self):
        return self.procedures.to_array(), self.table.to_dispatch_table_entries()


# ---- parameter helpers (the DSL's parameter forms) ----

def Scalar(v):
    return Parameter(abi.PT_SCALAR, [v])


def Vec2(x, y):
    """A PT_VEC2 parameter: a vertex of SetAs* (Parameters.cuh:714-742)."""
    return Parameter(abi.PT_VEC2, [x, y])


def Rand(mn, mx):
    return Parameter(abi.PT_RAND, [mn, mx])


def ShapeAttr(idx):
    return Parameter(abi.PT_SHAPE_ATTR, [idx])


def Philox(mn, mx, stream):
    """Project extension: Philox4x32-10 uniform in [mn, mx)."""
    return Parameter(abi.PT_PHILOX, [mn, mx, stream])


def AxiomParam(slot):
    """Project extension: per-axiom continuous parameter `slot`."""
    return Parameter(abi.PT_AXIOM_PARAM, [slot])


def _f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", float(np.float32(x))))[0]


def _bits_f32(u):
    return struct.unpack("<f", struct.pack("<I", u & 0xFFFFFFFF))[0]


def pack_operand(type, x):
    """PackUtils::packOperand (include/pga/core/PackUtils.h:18-23): the
    operand tag goes to bits 31:30 and the float's bits are shifted right by
    two, dropping its two lowest mantissa bits."""
    r = _f32_bits(x)
    return _bits_f32((r >> 2) | ((int(type) & 3) << 30))


def pack_float2(x, y):
    """PackUtils::packFloat2 (PackUtils.h:25-28): two [0,1] values quantised
    to 12 bits each."""
    f = np.float32
    return float(f(np.floor(f(x) * f(4095.0))) * f(4096.0) + f(np.floor(f(y) * f(4095.0))))


def Exp(*postfix):
    """An expression parameter from a postfix list of operands:
    ('s', value) scalar, ('r', min, max) rand, ('a', index) shape attribute,
    ('op', OPT_*) operation. Layout: value[0] = length, value[1..] packed
    operands (Parameters.cuh:563-568)."""
    vals = [float(len(postfix))]
    for item in postfix:
        kind = item[0]
        if kind == "s":
            vals.append(pack_operand(abi.ORT_SCALAR, item[1]))
        elif kind == "r":
            vals.append(pack_operand(abi.ORT_RAND, pack_float2(item[1], item[2])))
        elif kind == "a":
            vals.append(pack_operand(abi.ORT_SHAPE_ATTR, item[1]))
        elif kind == "op":
            vals.append(pack_operand(abi.ORT_OP, item[1]))
        else:
            raise ValueError(f"unknown expression operand {item!r}")
    if len(vals) > abi.MAX_NUM_PARAMETER_VALUES:
        raise RuntimeError("expression longer than PGA::Constants::MaxNumParameterValues")
    return Parameter(abi.PT_EXP, vals)
