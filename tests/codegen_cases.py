"""Partition-code cases: grammars in the grammar language, each written as
C++ partition code (compiler.generate_partition) under several partitions
and compiled together with g++ against include/pga/core/Core.h into
tests/cpp/_build/libgen_partitions.so. The library exports
gen_lower(namespace, ...), which returns a case's lowered program: its
procedure list and expanded dispatch table.

Partitions: "all" cuts every edge, the interpreter's choice; "cycles" cuts
only the edges that close a cycle; "random:<seed>" cuts each other edge with
probability 1/2.

Run `python tests/codegen_cases.py` to rebuild the library (build() in
__graft_entry__ does).
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")
SO = os.path.join(BUILD, "libgen_partitions.so")

MENGER = """
axiom Box A;
terminal B (1,0);
A = IfSizeLess(X, 0.111) { B | SubX };
ZDiscard = SubDiv(Z) { -1: A | -1: Discard() | -1: A };
YDiscard = SubDiv(Y) { -1: ZDiscard | -1: Discard() | -1: ZDiscard };
SubZ = SubDiv(Z) { -1: A | -1: A | -1: A };
SubY = SubDiv(Y) { -1: SubZ | -1: ZDiscard | -1: SubZ };
SubX = SubDiv(X) { -1: SubY | -1: YDiscard | -1: SubY };
"""

FACADE = """
axiom Box Lot;
terminal Wall (1, 0);
terminal Window (1, 1);
Lot    = CompSplit() { Discard() | Discard() | Facade };
Facade = Repeat(Y, 3, ADJUST_TO_FILL) Floor;
Floor  = Repeat(X, 2.5, ADJUST_TO_FILL) Tile;
Tile   = IfSizeLess(X, 1.6) { Wall | SubDiv(X) { -1: Wall | 1.2: Window | -1: Wall } };
"""

BUILDING = """
var pitch = 2.5;
axiom Box Lot;
terminal Wall (1, 0);
terminal Window (1, 1);
terminal Roof Cap (1, 2);
Lot = SubDiv(Y) { -1: Body | 1.5: Translate(0, Exp(ShapeSize(Y) * 0.5), 0) Scale(-1.1, -1, -1.1) Roof };
Body = CompSplit() { Cap | Discard() | Facade };
Facade = Repeat(Y, 3, ANCHOR_TO_END) Floor;
Floor = Repeat(X, pitch, ADJUST_TO_FILL) Tile;
Tile = IfSizeLess(Y, 2) { Wall | RandomRule(ShapeSeed()) { 0.6: Glazed | 0.4: Extrude(0.2) SwapSize(X, Z, Y) Roof } };
Glazed = If(Exp(Exp(ShapeNormal(Z) > 0.5) || Exp(ShapeSize(X) >= 2.4))) {
    SubDiv(X) { -1: Wall | 1.2: Rotate(0, 0, Exp(ShapeSeed() * 10)) Window | -1: Wall } |
    Replicate() { Wall | Scale(Rand(0.5, 0.9), -1, -1) Window }
};
"""

# context-sensitive grammars (Collider / IfCollides; their phases come from
# the operator graph, Graph::computePhases): a yard whose props ask for the
# building stored in phase 0; a city block whose buildings also store their
# walls (Quads) under tag 2; three phases, lamps asking for free props (the
# phase walk takes a Replicate's last branch first, so the lamps, which wait
# for the yard's colliders, are listed before the yard)
YARD = """
axiom Box Lot;
terminal Building (1, 0);
terminal Free (1, 1);
terminal Blocked (1, 2);
Lot = Replicate() { Scale(-0.5, -1, -0.5) Collider(1) Building | Yard };
Yard = Repeat(X, 1, ADJUST_TO_FILL) Repeat(Z, 1, ADJUST_TO_FILL) Scale(-0.5, -0.3, -0.5) IfCollides(1) { Blocked | Free };
"""
CITY = """
axiom Box Lot;
terminal Free (1, 1);
terminal Blocked (1, 2);
terminal WallQuad (1, 3);
Lot = Replicate() { Building | Yard };
Building = Scale(-0.45, -1, -0.45) Translate(Rand(-1.5, 1.5), 0, Rand(-1.5, 1.5)) Rotate(0, Rand(0, 90), 0)
           Collider(1) CompSplit() { Discard() | Discard() | Collider(2) WallQuad };
Yard = Repeat(X, 1, ADJUST_TO_FILL) Repeat(Z, 1, ADJUST_TO_FILL) Scale(-0.5, -0.3, -0.5) Prop;
Prop = IfCollides(2) { Blocked | IfCollides(1) { Blocked | Free } };
"""
LAMPS = """
axiom Box Lot;
terminal Building (1, 0);
terminal Free (1, 1);
terminal Blocked (1, 2);
terminal DarkLamp (1, 3);
terminal Lamp (1, 4);
Lot = Replicate() { Scale(-0.5, -1, -0.5) Collider(1) Building | Lamps | Yard };
Yard = Repeat(X, 1, ADJUST_TO_FILL) Repeat(Z, 1, ADJUST_TO_FILL) Scale(-0.5, -0.3, -0.5)
       IfCollides(1) { Blocked | Collider(2) Free };
Lamps = Repeat(X, 2, ADJUST_TO_FILL) Repeat(Z, 2, ADJUST_TO_FILL) Scale(-0.6, -0.3, -0.6) IfCollides(2) { DarkLamp | Lamp };
"""


def context_cases():
    """(case name, grammar text, partition spec) of the context-sensitive
    grammars (compiled like cases(); compared with the oracle, not with an
    interpreter table, which cannot hold Collider / IfCollides)."""
    out = []
    for name, text in (("yard", YARD), ("city", CITY), ("lamps", LAMPS)):
        for spec in ("all", "cycles", "random:5"):
            out.append((f"ctx_{name}_{spec.replace(':', '')}", text, spec))
    return out


# (case name, grammar text, axiom shape, partition spec, lots: "menger" | "facade" | "fuzz:<seed>")
def cases():
    from tests.grammar_fuzz import random_grammar
    out = []
    for spec in ("all", "cycles", "random:1"):
        out.append((f"menger_{spec.replace(':', '')}", MENGER, "Box", spec, "menger"))
    for spec in ("all", "cycles", "random:2"):
        out.append((f"facade_{spec.replace(':', '')}", FACADE, "Box", spec, "facade"))
        out.append((f"building_{spec.replace(':', '')}", BUILDING, "Box", spec, "building"))
    for seed in range(12):
        text, shape = random_grammar(seed)
        spec = ("all", "cycles", f"random:{seed + 10}")[seed % 3]
        out.append((f"fuzz{seed}_{spec.replace(':', '')}", text, shape, spec, f"fuzz:{seed}"))
    return out


def partition_uid(cycles, spec):
    if spec == "all":
        return None
    if spec == "cycles":
        return "".join("1" if c else "0" for c in cycles)
    seed = int(spec.split(":")[1])
    rng = np.random.default_rng(seed)
    return "".join("1" if (c or rng.random() < 0.5) else "0" for c in cycles)


def write_tu(path):
    from operatorgraph_amd.compiler import generate_partition, operator_graph_edges
    parts = ["// generated by tests/codegen_cases.py: partition code of the codegen test cases",
             "#include <pga/core/Core.h>", "", "#include <cstdio>", "#include <cstring>", "",
             "using namespace PGA;", "using namespace PGA::Operators;", "using namespace PGA::Shapes;",
             "using namespace PGA::Parameters;", ""]
    names = []
    for name, text, _, spec, _ in cases():
        uid = partition_uid(operator_graph_edges(text), spec)
        parts.append(generate_partition(text, uid, name))
        names.append(name)
    for name, text, spec in context_cases():
        uid = partition_uid(operator_graph_edges(text), spec)
        parts.append(generate_partition(text, uid, name))
        names.append(name)
    parts.append("""
template <class ProcListT>
static int lower_into(DispatchTable& table, pga_procedure* procs, uint32_t cap_p, uint32_t* np,
                      pga_dispatch_entry* ents, uint32_t cap_e, uint32_t* ne) {
    const auto templates = Lowering::lower<ProcListT>();
    auto entries = table.toDispatchTableEntriesPtr();
    const auto prog = Lowering::expand(templates, reinterpret_cast<const pga_dispatch_entry*>(entries.get()),
                                       table.entries.size());
    *np = (uint32_t)prog.procedures.size();
    *ne = (uint32_t)prog.entries.size();
    if (*np > cap_p || *ne > cap_e) return 2;
    std::memcpy(procs, prog.procedures.data(), prog.procedures.size() * sizeof(pga_procedure));
    std::memcpy(ents, prog.entries.data(), prog.entries.size() * sizeof(pga_dispatch_entry));
    return 0;
}

extern "C" __attribute__((visibility("default"))) int gen_lower(const char* ns, pga_procedure* procs,
                                                                uint32_t cap_p, uint32_t* np,
                                                                pga_dispatch_entry* ents, uint32_t cap_e,
                                                                uint32_t* ne, char* err, uint32_t err_cap) {
    try {""")
    for n in names:
        parts.append(f'        if (!std::strcmp(ns, "{n}")) return lower_into<{n}::ProcedureList>({n}::dispatchTable, '
                     'procs, cap_p, np, ents, cap_e, ne);')
    parts.append("""        return 3;
    } catch (const std::exception& e) {
        std::snprintf(err, err_cap, "%s", e.what());
        return 1;
    }
}
""")
    with open(path, "w", encoding="utf-8") as f:
        f.write("\n".join(parts))
    return names


def build():
    """Generates the cases' partition code and compiles the library."""
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(BUILD, "gen_partitions.cpp")
    write_tu(src)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-fPIC", "-shared", "-I",
                    os.path.join(ROOT, "include"), src, "-L", os.path.join(ROOT, "operatorgraph_amd"), "-lpga_amd",
                    "-Wl,-rpath," + os.path.join(ROOT, "operatorgraph_amd"), "-o", SO], check=True)
    return SO


_lib = None


def library():
    global _lib
    if _lib is None:
        from operatorgraph_amd import engine
        engine.load_library()
        _lib = C.CDLL(SO)
        _lib.gen_lower.restype = C.c_int
    return _lib


def lowered(name):
    """(procedures, entries) ctypes arrays of a case's lowered program."""
    from operatorgraph_amd import abi
    L = library()
    cap = 4096
    procs = (abi.Procedure * cap)()
    ents = (abi.DispatchEntry * cap)()
    np_, ne = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(512)
    rc = L.gen_lower(name.encode(), procs, cap, C.byref(np_), ents, cap, C.byref(ne), err, 512)
    if rc != 0:
        raise RuntimeError(f"gen_lower({name}) = {rc}: {err.value.decode()}")
    return procs, ents, np_.value, ne.value


class LoweredProgram:
    """A lowered program in the shape oracle.derive and Engine.load take."""

    def __init__(self, name):
        from operatorgraph_amd.program import DispatchTable, Procedure, ProcedureList
        from operatorgraph_amd.compiler import _entry
        procs, ents, n_p, n_e = lowered(name)
        self.procedures = ProcedureList([Procedure(p.op_type, p.shape_type, p.gen_func_idx, p.flags)
                                         for p in procs[:n_p]])
        self.table = DispatchTable([_entry(ents[i]) for i in range(n_e)])
        self.name = name

    def arrays(self):
        return self.procedures.to_array(), self.table.to_dispatch_table_entries()


if __name__ == "__main__":
    print(build())
