"""The C++ front end (include/pga/core/Core.h, include/pga/facade/Interpreter.h):
grammars written as the reference's partition code (Proc<Shape, operator
tree> lists, DispatchTable initialisers, Grid axiom generators) compiled
with g++ in tests/cpp/shim_partitions.cpp (loaded in-process as a shared
library).

CPU: the lowered programs are checked against what the reference's forms
mean - the Menger partition lowers to the compiled test-runner table
(src/test_runner/MengerSponge.cuh:28-45, as grammars.menger_test_runner
holds it), the chained static facade derives the terminals of the
interpreter-form facade grammar, and the call forms of the mixed partition
reach the entries SymbolDecorator.cuh:16-123 names (per-terminal counts
worked out by hand below). GPU: the C++ evaluator's derivations equal the
oracle's over the same lowered programs, and the interpreter facade
(initializePGA/executePGA) runs an interpreter-format table."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, grammars
from operatorgraph_amd.program import Exp
from oracle import oracle as O
from tests.helpers import assert_same_multiset

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "shim_partitions.cpp")
SO = os.path.join(ROOT, "tests", "cpp", "_build", "libshim_partitions.so")
HEADERS = [os.path.join(ROOT, "include", p) for p in
           ("pga.h", "pga/core/Core.h", "pga/facade/Interpreter.h")]
_shim = None


def shim():
    """The test program as a shared library (shim_main), built with g++
    against the in-tree library when missing or older than its sources, and
    loaded into this process: cooperative launches of a second process on
    the GPU would be refused while this one holds the device."""
    global _shim
    if _shim is not None:
        return _shim
    lib = os.path.join(ROOT, "operatorgraph_amd", "libpga_amd.so")
    if not os.path.exists(lib):
        pytest.skip("libpga_amd.so is not built")
    srcs = [SRC, lib] + HEADERS
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(s) for s in srcs):
        os.makedirs(os.path.dirname(SO), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-fPIC", "-shared",
                        "-DSHIM_LIBRARY", "-I", os.path.join(ROOT, "include"), SRC, "-L",
                        os.path.join(ROOT, "operatorgraph_amd"), "-lpga_amd",
                        "-Wl,-rpath," + os.path.join(ROOT, "operatorgraph_amd"), "-o", SO], check=True)
    from operatorgraph_amd import engine
    engine.load_library()                    # the same libpga_amd.so instance as the engine's
    _shim = C.CDLL(SO)
    _shim.shim_main.restype = C.c_int
    _shim.shim_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
    return _shim


def run_shim(*args):
    argv = [b"shim_partitions"] + [str(a).encode() for a in args]
    arr = (C.c_char_p * len(argv))(*argv)
    rc = shim().shim_main(len(argv), arr)
    assert rc == 0, f"shim_main{args} returned {rc}"


class Lowered:
    """A lowered program read back from the test program's files, in the
    shape oracle.derive takes (arrays(), procedures, table.entries)."""

    def __init__(self, base):
        pb = open(base + ".procs", "rb").read()
        eb = open(base + ".entries", "rb").read()
        n_p, n_e = len(pb) // C.sizeof(abi.Procedure), len(eb) // C.sizeof(abi.DispatchEntry)
        self.procs = (abi.Procedure * n_p).from_buffer_copy(pb)
        self.entries = (abi.DispatchEntry * n_e).from_buffer_copy(eb)
        self.procedures = list(range(n_p))
        self.table = type("T", (), {"entries": list(range(n_e))})()
        self.name = os.path.basename(base)

    def arrays(self):
        return self.procs, self.entries


def lower(scene, tmp_path):
    run_shim("lower", scene, tmp_path)
    return Lowered(os.path.join(str(tmp_path), scene))


def gen_axioms(tmp_path, scene):
    b = open(os.path.join(str(tmp_path), scene + ".gen_axioms"), "rb").read()
    return np.frombuffer(b, dtype=abi.AXIOM_DTYPE).copy()


def _param(entry, k):
    p = entry.parameters[k]
    return p.type, [p.value[i] for i in range(abi.MAX_NUM_PARAMETER_VALUES)]


def test_menger_partition_lowers_to_the_compiled_table(tmp_path):
    low = lower("menger", tmp_path)
    ref = grammars.menger_test_runner()
    rp, re_ = ref.arrays()
    assert bytes(low.procs) == bytes(rp)[: C.sizeof(low.procs)] and len(low.procs) == len(ref.procedures)
    assert len(low.entries) == len(ref.table.entries)
    for i in range(len(low.entries)):
        a, b = low.entries[i], re_[i]
        assert a.operator_code == b.operator_code and a.num_parameters == b.num_parameters, i
        assert [_param(a, k) for k in range(a.num_parameters)] == [_param(b, k) for k in range(b.num_parameters)], i
        assert [(a.successors[k].entry_index) for k in range(a.num_successors)] == \
               [(b.successors[k].entry_index) for k in range(b.num_successors)], i
    ax = gen_axioms(tmp_path, "menger")          # Grid::Static, one cell: size 9 resting on y = 0
    assert len(ax) == 1 and ax["entry_index"][0] == 2 and ax["size"][0].tolist() == [9.0, 9.0, 9.0]
    assert ax["model"][0][7] == 4.5
    r = O.derive(low, ax)
    assert r.count(0) == 8000
    assert_same_multiset(r.instances(0), O.derive(ref, ax).instances(0), "menger")


def test_static_facade_chain_derives_the_facade(tmp_path):
    low = lower("facade", tmp_path)
    # one inline node per chained operator: Repeat, Repeat, IfSizeLess,
    # Subdivide after the four table entries; the Subdivide is the static form
    kinds = [(p.op_type, p.shape_type, p.flags) for p in low.procs]
    assert (abi.SUBDIV, abi.QUAD, abi.PROC_STATIC) in kinds
    lots = axioms.facade_lots(40)
    walls, windows = axioms.facade_lot_terminals(lots)
    r = O.derive(low, lots)
    f = O.derive(grammars.facade(), lots)
    assert r.count(0) == walls.sum() and r.count(1) == windows.sum()
    for t in range(2):
        a, b = r.instances(t).copy(), f.instances(t).copy()
        a["meta"] &= 0xFFFF0000                  # the Generate entry differs between the two tables
        b["meta"] &= 0xFFFF0000
        assert_same_multiset(a, b, f"terminal {t}")


def test_mixed_partition_call_forms(tmp_path):
    """One Box axiom of size (3, 2, 4). ComponentSplit's faces: the top goes
    to FSCall<1, 0, -1, 2> (procedure 1 on the fixed entry 2: Extrude, then a
    Box terminal 0); the bottom to DCall<0> (entry 0's successor 0 = entry 1,
    procedure 2: size.x = 3 > 2.5, terminal = entry 1's parameter = 1); each
    of the 4 sides (size.x 3 or 4) to Replicate: PSCall<2, 1> (entry 0's
    successor 1 = entry 3, terminal 3) and Scale -> FSCall<3, 0> (no entry:
    stays on entry 0, terminal = entry 0's parameter = 2)."""
    low = lower("mix", tmp_path)
    ax = axioms.empty(1)
    ax["size"] = [3.0, 2.0, 4.0]
    ax["entry_index"] = 0
    ax["shape_type"] = abi.BOX
    r = O.derive(low, ax)
    assert [r.count(t) for t in range(4)] == [1, 1, 4, 4]
    b = r.instances(0)["model"][0]             # the extruded top: a Box of depth 2
    assert np.isclose(np.linalg.norm(b[[0, 4, 8]]) * np.linalg.norm(b[[1, 5, 9]]) *
                      np.linalg.norm(b[[2, 6, 10]]), 3.0 * 4.0 * 2.0, rtol=1e-5)
    # constants: the translate's expression, the rotation angle and the If condition
    e0 = low.entries[0]
    assert low.procs[e0.operator_code].op_type == abi.TRANSLATE
    want = Exp(("a", 13), ("s", 0.5), ("op", abi.OPT_MULTI))
    t, v = _param(e0, 1)
    assert t == abi.PT_EXP and np.array_equal(np.float32(v[:4]), np.float32(want.values))
    rot = low.entries[e0.successors[0].entry_index]
    assert low.procs[rot.operator_code].op_type == abi.ROTATE and _param(rot, 1)[1][0] == 30.0
    e1 = low.entries[1]
    t, v = _param(e1, 0)
    want = Exp(("a", 12), ("s", 2.5), ("op", abi.OPT_GT))
    assert low.procs[e1.operator_code].op_type == abi.IF and t == abi.PT_EXP
    assert np.array_equal(np.float32(v[:4]), np.float32(want.values))


# ---- on the GPU: the C++ evaluator against the oracle ----

def _read_instances(path):
    return np.frombuffer(open(path, "rb").read(), dtype=abi.INSTANCE_DTYPE).copy()


def _note(msg):
    print(msg, flush=True)


@pytest.mark.gpu
def test_cpp_evaluator_menger(tmp_path):
    _note("C++ SinglePhaseEvaluator: Menger partition")
    low = lower("menger", tmp_path)
    ax = gen_axioms(tmp_path, "menger")
    run_shim("run", "menger", tmp_path)
    got = _read_instances(os.path.join(str(tmp_path), "menger.t0"))
    assert len(got) == 8000
    assert_same_multiset(got, O.derive(low, ax).instances(0), "menger")


@pytest.mark.gpu
@pytest.mark.parametrize("scene,n", [("facade", 96), ("mix", 6)])
def test_cpp_evaluator_partitions(tmp_path, scene, n):
    _note(f"C++ SinglePhaseEvaluator: {scene} partition, {n} axioms")
    low = lower(scene, tmp_path)
    if scene == "facade":
        ax = axioms.facade_lots(n)
    else:
        rng = np.random.default_rng(3)
        ax = axioms.empty(n)
        ax["size"] = rng.uniform(1.0, 5.0, size=(n, 3)).astype(np.float32)
        ax["model"][:, [3, 7, 11]] = rng.uniform(-9, 9, size=(n, 3)).astype(np.float32)
        ax["entry_index"] = 0
        ax["shape_type"] = abi.BOX
    r = O.derive(low, ax)
    nt = 2 if scene == "facade" else 4
    caps = np.array([max(1, r.count(t)) for t in range(nt)], dtype=np.uint64)
    base = os.path.join(str(tmp_path), scene)
    open(base + ".axioms", "wb").write(np.ascontiguousarray(ax).tobytes())
    open(base + ".caps", "wb").write(caps.tobytes())
    run_shim("run", scene, tmp_path)
    for t in range(nt):
        got = _read_instances(f"{base}.t{t}")
        assert_same_multiset(got, r.instances(t), f"{scene} terminal {t}")


@pytest.mark.gpu
def test_interpreter_facade_runs_an_interpreter_table(tmp_path):
    """initializePGA/executePGA over the interpreter-format Menger table
    (grammars.menger_interpreter) with two interpreter axioms (a Box at
    entry 0 and one skipped with entry -1)."""
    _note("C++ interpreter facade: Menger interpreter table")
    prog = grammars.menger_interpreter()
    _, entries = prog.arrays()
    d = str(tmp_path)
    open(os.path.join(d, "interp.table"), "wb").write(bytes(entries))
    open(os.path.join(d, "interp.axioms"), "wb").write(np.array([abi.BOX, 0, abi.BOX, -1], dtype=np.int32).tobytes())
    open(os.path.join(d, "interp.caps"), "wb").write(np.array([8000], dtype=np.uint64).tobytes())
    run_shim("interp", d)
    got = _read_instances(os.path.join(d, "interp.t0"))
    ax = axioms.menger_axiom(1.0, 0, seed=0.0)   # the interpreter's axiom: origin, size 1, seed table zero
    ref = O.derive(prog, ax).instances(0)
    assert len(got) == 8000
    assert_same_multiset(got, ref, "interpreter Menger")


FACADE_TEXT = """
// the batched facade in the grammar language
axiom Box Lot;
terminal Wall (1, 0);
terminal Window (1, 1);
Lot    = CompSplit() { Discard() | Discard() | Facade };
Facade = Repeat(Y, 3, ADJUST_TO_FILL) Floor;
Floor  = Repeat(X, 2.5, ADJUST_TO_FILL) Tile;
Tile   = IfSizeLess(X, 1.6) { Wall | SubDiv(X) { -1: Wall | 1.2: Window | -1: Wall } };
"""


def test_compile_grammar_from_cpp(tmp_path):
    """compileGrammar (the interpreter's parse step in C++) gives, through
    DispatchTable::toDispatchTableEntriesPtr, the same entry bytes as the C
    ABI's compile, and the axiom at entry 0."""
    from operatorgraph_amd.compiler import compile_grammar
    d = str(tmp_path)
    open(os.path.join(d, "grammar.pga"), "w").write(FACADE_TEXT)
    run_shim("grammar", d)
    got = open(os.path.join(d, "grammar.table"), "rb").read()
    assert got == compile_grammar(FACADE_TEXT).raw


def test_compile_grammar_from_cpp_reports_the_line(tmp_path, capfd):
    d = str(tmp_path)
    open(os.path.join(d, "grammar.pga"), "w").write("axiom Box A;\nA = Translate(0, 0, 0)\n")
    argv = [b"shim_partitions", b"grammar", d.encode()]
    rc = shim().shim_main(len(argv), (C.c_char_p * len(argv))(*argv))
    assert rc == 1
    assert "parsing error: line 3" in capfd.readouterr().err


@pytest.mark.gpu
def test_interpreter_flow_from_grammar_text(tmp_path):
    """Grammar text -> compileGrammar -> setAxioms/initializePGA/executePGA,
    all in C++: the interpreter's Box axiom (origin, size 1, seed table
    zero) becomes one facade lot, checked against the oracle."""
    from operatorgraph_amd.compiler import compile_grammar
    _note("C++ interpreter flow: facade grammar text")
    d = str(tmp_path)
    open(os.path.join(d, "grammar.pga"), "w").write(FACADE_TEXT.replace("1.6", "0.2").replace("2.5", "0.25")
                                                    .replace("(Y, 3,", "(Y, 0.2,").replace("1.2:", "0.1:"))
    open(os.path.join(d, "grammar.caps"), "wb").write(np.array([4096, 4096], dtype=np.uint64).tobytes())
    run_shim("grammar-run", d)
    g = compile_grammar(open(os.path.join(d, "grammar.pga")).read())
    ax = g.axiom_records([0.0])
    r = O.derive(g.program, ax)
    for t in range(2):
        got = _read_instances(os.path.join(d, f"grammar.t{t}"))
        assert len(got) > 0
        assert_same_multiset(got, r.instances(t), f"interpreter flow terminal {t}")


@pytest.mark.gpu
def test_interpreter_flow_triangle_soup(tmp_path):
    """The same flow with generation-function-0 terminals: the evaluator's
    getTriangleMesh gives, per terminal, the oracle's mesh of its shapes
    (canonical instance order) and the reference's index pattern."""
    from oracle.oracle import mesh_indices
    from operatorgraph_amd.compiler import compile_grammar
    from tests.helpers import canon_order
    _note("C++ interpreter flow: facade grammar text, triangle soup")
    d = str(tmp_path)
    text = (FACADE_TEXT.replace("1.6", "0.2").replace("2.5", "0.25").replace("(Y, 3,", "(Y, 0.2,")
            .replace("1.2:", "0.1:").replace("(1, 0)", "(0, 0)").replace("(1, 1)", "(0, 1)"))
    open(os.path.join(d, "grammar.pga"), "w").write(text)
    open(os.path.join(d, "grammar.caps"), "wb").write(np.array([4096, 4096], dtype=np.uint64).tobytes())
    open(os.path.join(d, "grammar.mesh"), "wb").write(np.array([abi.QUAD, abi.QUAD], dtype=np.int32).tobytes())
    run_shim("grammar-run", d)
    g = compile_grammar(text)
    r = O.derive(g.program, g.axiom_records([0.0]))
    for t in range(2):
        got = _read_instances(os.path.join(d, f"grammar.t{t}"))
        n = len(got)
        assert n > 0
        assert_same_multiset(got, r.instances(t), f"triangle soup terminal {t}")
        v = np.fromfile(os.path.join(d, f"grammar.t{t}.vtx"), dtype=np.float32).reshape(n, 4, 9)
        idx = np.fromfile(os.path.join(d, f"grammar.t{t}.idx"), dtype=np.uint32)
        ref = r.mesh(t).reshape(n, 4, 9)
        assert np.array_equal(v[canon_order(got)], ref[canon_order(r.instances(t))])
        assert np.array_equal(idx, mesh_indices(n, abi.QUAD))


def test_yard_partition_lowers_its_phases(tmp_path):
    """The two-phase Yard partition: the lot's DCall keeps the table's phase
    0, its FSCall<2, 1, -1, 2> goes to entry 2 in phase 1, the operators
    chained inside each procedure run in their shape's phase
    (PGA_PHASE_INHERIT), and the yard's PSCalls keep the table's phase 1.
    On a 4 x 3 x 4 lot the building (2 x 3 x 2 at the centre) blocks the
    four inner props of sixteen."""
    low = lower("yard", tmp_path)
    phases = {}
    for i in range(len(low.entries)):
        e = low.entries[i]
        op = low.procs[e.operator_code].op_type
        phases[i] = (op, [(e.successors[k].entry_index, e.successors[k].phase_index) for k in range(e.num_successors)])
    assert phases[0][0] == abi.REPLICATE and [p for _, p in phases[0][1]] == [0, 1]
    assert phases[0][1][1][0] == 2
    assert phases[2][0] == abi.REPEAT and phases[2][1][0][1] == abi.PHASE_INHERIT
    ifc = [i for i, (op, _) in phases.items() if op == abi.IFCOLLIDES]
    assert len(ifc) == 1 and [p for _, p in phases[ifc[0]][1]] == [1, 1]
    col = [i for i, (op, _) in phases.items() if op == abi.COLLIDER]
    assert len(col) == 1 and [p for _, p in phases[col[0]][1]] == [abi.PHASE_INHERIT]
    ax = axioms.empty(1)
    ax["size"] = (4.0, 3.0, 4.0)
    ax["entry_index"] = 0
    ax["shape_type"] = abi.BOX
    r = O.derive(low, ax)
    assert [r.count(t) for t in range(3)] == [1, 12, 4] and r.colliders == 1
    assert np.all(r.instances(1)["meta"] >> 28 == 1) and np.all(r.instances(0)["meta"] >> 28 == 0)


@pytest.mark.gpu
def test_cpp_multi_phase_evaluator_yard(tmp_path):
    _note("C++ MultiPhaseEvaluator: Yard partition, 64 lots")
    low = lower("yard", tmp_path)
    rng = np.random.default_rng(5)
    n = 64
    ax = axioms.empty(n)
    ax["size"][:, 0] = rng.uniform(3.0, 9.0, n).astype(np.float32)
    ax["size"][:, 1] = 3.0
    ax["size"][:, 2] = rng.uniform(3.0, 9.0, n).astype(np.float32)
    ax["model"][:, 3] = (np.arange(n) % 8 * 12.0).astype(np.float32)
    ax["model"][:, 11] = (np.arange(n) // 8 * 12.0).astype(np.float32)
    ax["entry_index"] = 0
    ax["shape_type"] = abi.BOX
    r = O.derive(low, ax)
    caps = np.array([max(1, r.count(t)) for t in range(3)], dtype=np.uint64)
    base = os.path.join(str(tmp_path), "yard")
    open(base + ".axioms", "wb").write(np.ascontiguousarray(ax).tobytes())
    open(base + ".caps", "wb").write(caps.tobytes())
    run_shim("run", "yard", tmp_path)
    for t in range(3):
        assert_same_multiset(_read_instances(f"{base}.t{t}"), r.instances(t), f"yard terminal {t}")
    assert r.count(2) > 0 and r.count(1) > 0
