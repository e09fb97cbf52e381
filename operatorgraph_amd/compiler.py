"""The grammar front end: PGA grammar text to a `Program` over the
interpreter's procedure list.

The parser and the table generator are the library's host code
(operatorgraph_amd/csrc/pga_compiler.cpp, declared in include/pga.h); this
module calls them through the C ABI, which needs no GPU. The flow is the
interpreter's (src/interpreter/InterpreterApp.cpp:175-231): parse the
source, build the dispatch table over the interpreter's procedure list,
create one axiom shape per declared axiom (a unit shape at the origin whose
seed and custom attribute are the axiom's seed, PGAFacade.cu:697-729).

    g = compile_grammar(text)       # or compile_file(path)
    g.program                       # Program(INTERPRETER procedures, table)
    g.axioms                        # [(entry_index, shape_type), ...]
    g.axiom_records(seeds)          # pga_axiom records for Engine.set_axioms

Errors in the source raise `GrammarError` with the line number.

`generate_partition(text, partition, name)` writes the grammar's operator
graph, cut by a partition, as C++ partition code (Proc<Shape, operator
tree> structs, the procedure list and a DispatchTable) for the C++ front
end (include/pga/core/Core.h); `operator_graph_edges(text)` gives the
graph's edges with their must-cut flags (edges that close a cycle or enter
a later phase's IfCollides must be cut).
"""
import ctypes as C

import numpy as np

from . import abi
from .program import DispatchTable, Entry, Parameter, Procedure, ProcedureList, Program


class GrammarError(ValueError):
    """A grammar the front end cannot turn into a dispatch table."""


def _lib():
    from .engine import load_library
    return load_library()


def interpreter_procedures():
    """The interpreter's procedure list as the library holds it
    (src/interpreter/PGAFacade.cu:24-62)."""
    L = _lib()
    n = C.c_uint32()
    st = L.pga_interpreter_procedures(None, 0, C.byref(n))
    if st != abi.OK:
        raise GrammarError(f"pga_interpreter_procedures: status {st}")
    arr = (abi.Procedure * n.value)()
    st = L.pga_interpreter_procedures(arr, n.value, C.byref(n))
    if st != abi.OK:
        raise GrammarError(f"pga_interpreter_procedures: status {st}")
    return ProcedureList([Procedure(p.op_type, p.shape_type, p.gen_func_idx, p.flags) for p in arr])


def _entry(e):
    params = []
    for k in range(e.num_parameters):
        p = e.parameters[k]
        values = [float(v) for v in p.value]
        while values and values[-1] == 0.0 and np.signbit(np.float32(values[-1])) == 0:
            values.pop()
        params.append(Parameter(p.type, values))
    succ = [(e.successors[k].entry_index, e.successors[k].phase_index) for k in range(e.num_successors)]
    edges = [e.edge_indices[k] for k in range(e.num_edge_indices)]
    return Entry(e.operator_code, params, succ, edges, e.sub_graph_index)


class CompiledGrammar:
    def __init__(self, program, axioms, raw):
        self.program = program
        self.axioms = axioms
        self.raw = raw          # bytes of the pga_dispatch_entry array, as the library wrote it

    def axiom_records(self, seeds, sizes=(1.0, 1.0, 1.0), positions=None):
        """One pga_axiom per (declared axiom, seed): the interpreter's axiom
        shapes (PGAFacade.cu:697-729), with an optional size and position per
        record. Returns a structured array of abi.AXIOM_DTYPE."""
        from . import axioms as A
        seeds = np.asarray(seeds, dtype=np.float32).reshape(-1)
        n = len(seeds)
        ax = A.empty(n)
        entry, shape = self.axioms[0]
        ax["entry_index"] = entry
        ax["shape_type"] = shape
        ax["seed"] = seeds
        ax["custom"] = seeds
        sizes = np.broadcast_to(np.asarray(sizes, dtype=np.float32), (n, 3))
        ax["size"] = sizes
        m = np.zeros((n, 12), dtype=np.float32)
        m[:, 0] = m[:, 5] = m[:, 10] = 1.0
        if positions is not None:
            pos = np.broadcast_to(np.asarray(positions, dtype=np.float32), (n, 3))
            m[:, 3], m[:, 7], m[:, 11] = pos[:, 0], pos[:, 1], pos[:, 2]
        ax["model"] = m
        return ax


def compile_grammar(source, name="grammar"):
    """Parses PGA grammar text and generates its dispatch table."""
    L = _lib()
    if isinstance(source, str):
        source = source.encode("utf-8")
    h = C.c_void_p()
    st = L.pga_compile_grammar(source, C.byref(h))
    try:
        if st != abi.OK:
            msg = L.pga_grammar_error(h) if h.value else b""
            raise GrammarError(f"{name}: {msg.decode() if msg else abi.STATUS_NAMES.get(st, st)}")
        ep = C.POINTER(abi.DispatchEntry)()
        n = C.c_uint32()
        if L.pga_grammar_entries(h, C.byref(ep), C.byref(n)) != abi.OK:
            raise GrammarError(f"{name}: pga_grammar_entries failed")
        entries = [_entry(ep[i]) for i in range(n.value)]
        raw = C.string_at(ep, C.sizeof(abi.DispatchEntry) * n.value)
        ap = C.POINTER(abi.GrammarAxiom)()
        m = C.c_uint32()
        if L.pga_grammar_axioms(h, C.byref(ap), C.byref(m)) != abi.OK:
            raise GrammarError(f"{name}: pga_grammar_axioms failed")
        axioms = [(ap[i].entry_index, ap[i].shape_type) for i in range(m.value)]
    finally:
        if h.value:
            L.pga_grammar_destroy(h)
    return CompiledGrammar(Program(interpreter_procedures(), DispatchTable(entries), name), axioms, raw)


def compile_file(path, name=None):
    import os
    with open(path, "r", encoding="utf-8") as f:
        text = f.read()
    return compile_grammar(text, name or os.path.basename(os.path.dirname(os.path.abspath(path))))


def _partition(source, partition, name):
    L = _lib()
    if isinstance(source, str):
        source = source.encode("utf-8")
    h = C.c_void_p()
    st = L.pga_generate_partition(source, None if partition is None else partition.encode(), name.encode(),
                                  C.byref(h))
    try:
        if st != abi.OK:
            msg = L.pga_grammar_error(h) if h.value else b""
            raise GrammarError(f"{name}: {msg.decode() if msg else abi.STATUS_NAMES.get(st, st)}")
        code, n = C.c_char_p(), C.c_uint64()
        L.pga_grammar_code(h, C.byref(code), C.byref(n))
        ne, flags = C.c_uint32(), C.POINTER(C.c_uint8)()
        L.pga_grammar_edges(h, C.byref(ne), C.byref(flags))
        cycles = [bool(flags[i]) for i in range(ne.value)]
        return C.string_at(code, n.value).decode(), cycles
    finally:
        if h.value:
            L.pga_grammar_destroy(h)


def operator_graph_edges(source):
    """Must-cut flags of the operator graph's edges, in edge order: the edge
    closes a cycle, or enters an IfCollides of a later phase."""
    return _partition(source, None, "G")[1]


def generate_partition(source, partition=None, name="G"):
    """C++ partition code of the grammar in `namespace name`. `partition` is
    one '0' (keep) or '1' (cut) per operator-graph edge; None cuts every
    edge."""
    return _partition(source, partition, name)[0]
