"""The grammar front end (pga_compile_grammar, host code behind the C ABI):
syntax, entry layout and numbering, errors, and the reference's interpreter
grammars through the oracle. No GPU is used.

What pins it:
* the menger_sponge_3 table against the hand-compiled
  `grammars.menger_interpreter()` (written from the reference's entry
  builders), differing only where fromRuleSet gives each inline operator an
  entry of its own;
* terminal counts of the reference grammars against closed forms of their
  recursions and against the counts their configuration files state;
* operand encodings through what a derivation does with them (positions and
  sizes of terminals), not through a restatement of the encoder;
* the stored tables (tests/golden/grammar_tables.json) against a fresh
  compile of the reference's grammar files, where those are present."""
import hashlib
import os

import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, compiler, grammars
from operatorgraph_amd.compiler import GrammarError, compile_grammar
from operatorgraph_amd.program import Exp, Parameter
from tests.helpers import (TREE_COUNTS, assert_same_multiset, fixture_grammar, grammar_fixture,
                           grammar_lots)

REF_GRAMMARS = "/root/reference/src/interpreter/grammars"
HAVE_REF = os.path.isdir(REF_GRAMMARS)
has_reference = pytest.mark.skipif(not HAVE_REF, reason="reference grammars not present")

B, Q = abi.BOX, abi.QUAD


def derive(g, ax, twins=0):
    from oracle import oracle as O
    return O.derive(g.program, ax, twins=twins)


def op_of(g, i):
    return grammars.INTERPRETER_PROCEDURES.procedures[g.program.table.entries[i].operator_code]


def unit_axiom(g, size=(1.0, 1.0, 1.0), pos=(0.0, 0.0, 0.0), seed=0.25):
    return g.axiom_records([seed], sizes=size, positions=pos)


def test_procedure_list_is_the_interpreters():
    got = [p.key() for p in compiler.interpreter_procedures().procedures]
    assert got == [p.key() for p in grammars.INTERPRETER_PROCEDURES.procedures]
    assert len(got) == 36 + 4 * 18


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


def _menger_expected_bytes():
    """menger_interpreter() with fromRuleSet's numbering: ZDiscard's inline
    Discard() is entry 8 of its own (YDiscard's is entry 7)."""
    prog = grammars.menger_interpreter()
    entries = list(prog.table.entries)
    z = entries[6]
    entries[6] = type(z)(z.operator_code, z.parameters, [(0, 0), (8, 0), (0, 0)])
    entries.append(entries[7])
    from operatorgraph_amd.program import DispatchTable
    return bytes(DispatchTable(entries).to_dispatch_table_entries())


def test_menger_table_bytes():
    g = compile_grammar(MENGER, "menger")
    assert g.axioms == [(0, B)]
    assert g.raw == _menger_expected_bytes()
    r = derive(g, axioms.menger_axiom(1.0, 0))
    assert r.count(0) == 20 ** 3


@has_reference
def test_menger_reference_file():
    g = compiler.compile_file(os.path.join(REF_GRAMMARS, "menger_sponge_3", "grammar.pga"))
    assert g.raw == _menger_expected_bytes()


def test_menger_same_instances_as_hand_table():
    g = compile_grammar(MENGER)
    ax = axioms.menger_axiom(1.0, 0)
    a = derive(g, ax).instances(0)
    b = derive(type(g)(grammars.menger_interpreter(), g.axioms, b""), ax).instances(0)
    a["meta"] &= 0xFFFF0000     # the Generate entry is 1 in both; depth kept
    b["meta"] &= 0xFFFF0000
    assert_same_multiset(a, b, "menger: compiled vs hand table")


def test_entry_numbering_follows_from_rule_set():
    """DispatchTableGenerator::fromRuleSet order: the first axiom's symbol is
    entry 0; successors get entries when their parent is visited (symbols
    once, inline operators each time); the queue is breadth-first; a second
    axiom reached before its turn keeps the entry it was given."""
    src = """
    axiom Box Root;
    axiom Box Second;
    terminal T (1, 0);
    Root = Replicate() { Mid | Translate(1, 0, 0) Second | Tail };
    Mid = Scale(-1, -1, -1) T;
    Second = Replicate() { Tail | T };
    Tail = Discard();
    """
    g = compile_grammar(src)
    kinds = [op_of(g, i).op_type for i in range(len(g.program.table.entries))]
    # 0 Root (Replicate): Mid=1, inline Translate=2, Tail=3; the Translate:
    # Second=4 (queued as an axiom, no entry yet); Second: Tail (3), T=5;
    # Mid: T (5); Tail: Discard
    assert kinds == [abi.REPLICATE, abi.SCALE, abi.TRANSLATE, abi.DISCARD, abi.REPLICATE, abi.GENERATE]
    succ = [[s.entry_index for s in e.successors] for e in g.program.table.entries]
    assert succ == [[1, 2, 3], [5], [4], [], [3, 5], []]
    assert g.axioms == [(0, B), (4, B)]


def test_operand_forms_by_derivation():
    """Each operand form, checked by where the terminal lands: the instance
    model is rows 0..2 of model * scale(size), so column 3 is the position
    and the column norms are the sizes."""
    src = """
    var half = 0.5;
    var two = 2;
    axiom Box A;
    terminal T (1, 0);
    terminal U V (1, 1);
    A = Translate(Exp(ShapeSize(Y) * half), ShapePos(X), Exp(Exp(1 + two) * 2)) B;
    B = Scale(ShapeSize(Z), Exp(ShapeSize(X) / 4), two) C;
    C = IfSizeLess(Y, 0.4) { U | T };
    """
    g = compile_grammar(src)
    ax = unit_axiom(g, size=(8.0, 4.0, 3.0), pos=(2.0, 0.0, 0.0))
    r = derive(g, ax)
    assert r.count(0) == 1 and r.count(1) == 0
    m = r.instances(0)["model"][0].reshape(3, 4)
    # Translate(4*0.5, posX=2, (1+2)*2) from (2, 0, 0)
    np.testing.assert_array_equal(m[:, 3], [4.0, 2.0, 6.0])
    # Scale(sizeZ=3, sizeX/4=2, 2): absolute sizes
    np.testing.assert_array_equal(np.linalg.norm(m[:, :3], axis=0), [3.0, 2.0, 2.0])


def test_if_expressions_and_normals():
    """CompSplit's faces by their normals: If(Exp(ShapeNormal(Y) == 1))
    keeps the top face (PGA: face normals in the shape's frame)."""
    src = """
    axiom Box A;
    terminal Top (1, 0);
    terminal Other (1, 1);
    A = CompSplit() { F | F | F };
    F = If(Exp(ShapeNormal(Y) == 1)) { Top | Other };
    """
    g = compile_grammar(src)
    r = derive(g, unit_axiom(g, size=(2.0, 3.0, 4.0)))
    assert r.count(0) == 1 and r.count(1) == 5


def test_rand_within_bounds():
    src = """
    axiom Box A;
    terminal T (1, 0);
    A = Repeat(X, 1, ADJUST_TO_FILL) Scale(Rand(0.5, 1.5), -1, -1) T;
    """
    g = compile_grammar(src)
    r = derive(g, unit_axiom(g, size=(64.0, 1.0, 1.0)))
    sx = np.linalg.norm(r.instances(0)["model"].reshape(-1, 3, 4)[:, :, 0], axis=1)
    assert len(sx) == 64
    assert sx.min() >= 0.5 - 1e-6 and sx.max() <= 1.5 + 1e-6
    assert len(np.unique(sx)) > 16


def test_repeat_modes_and_extrude():
    src = """
    axiom Quad A;
    terminal T (1, 0);
    terminal S (1, 1);
    A = Repeat(X, 3, ANCHOR_TO_END) Extrude(0.5) SwapSize(Z, X, Y) T;
    """
    g = compile_grammar(src)
    assert g.axioms == [(0, Q)]
    e = g.program.table.entries[1]
    assert op_of(g, 1).op_type == abi.EXTRUDE
    assert [p.values for p in e.parameters] == [[2.0], [0.5]]     # implicit Z axis first
    r = derive(g, unit_axiom(g, size=(10.0, 2.0, 0.0)))
    # 10 = 3 + 3 + 3 + 1: ceil(10/3) = 4 tiles
    assert r.count(0) == 4
    sizes = np.sort(np.linalg.norm(r.instances(0)["model"].reshape(-1, 3, 4)[:, :, :3], axis=1), axis=0)
    # Extrude: (x, y) quad -> (x, extent, y) box (Extrude.cuh:42-53);
    # SwapSize(Z, X, Y) -> (2, tile, 0.5)
    np.testing.assert_allclose(sizes[:, 0], [2.0] * 4)
    np.testing.assert_allclose(sorted(sizes[:, 1]), [1.0, 3.0, 3.0, 3.0])
    np.testing.assert_allclose(sizes[:, 2], [0.5] * 4)


def test_parameter_layouts():
    src = """
    var w = 2;
    axiom Box A;
    terminal T (1, 3, 0.25, 9);
    A = SubDiv(Y) { 1: B | -1: B | w: B | -3: B };
    B = RandomRule(ShapeSeed()) { 0.25: T | 0.75: Discard() };
    """
    g = compile_grammar(src)
    e = g.program.table.entries
    # [axis, absolute sum, relative sum], then [factor, absolute and
    # relative sums before it] (trailing zeros are not listed)
    assert [p.values for p in e[0].parameters] == [
        [1.0, 3.0, 4.0], [1.0], [-1.0, 1.0], [2.0, 1.0, 1.0], [-3.0, 3.0, 1.0]]
    assert [p.type for p in e[1].parameters] == [abi.PT_SHAPE_ATTR, abi.PT_SCALAR, abi.PT_SCALAR]
    assert [p.values for p in e[1].parameters] == [[15.0], [0.25], [0.75]]
    gen = e[e[1].successors[0].entry_index]
    assert op_of(g, e[1].successors[0].entry_index).gen_func_idx == 1
    assert [p.values for p in gen.parameters] == [[3.0], [0.25], [9.0]]


def test_expression_encoding_matches_parameter_helper():
    g = compile_grammar("""
    axiom Box A;
    terminal T (1, 0);
    A = Translate(0, Exp(Exp(ShapeSize(Y) * -0.5) + Rand(0.25, 0.75)), Exp(ShapeRotation(X, Y) >= (3, 4))) T;
    """)
    p = g.program.table.entries[0].parameters
    want = Exp(("a", 13), ("s", -0.5), ("op", abi.OPT_MULTI), ("r", 0.25, 0.75), ("op", abi.OPT_ADD))
    assert p[1].type == abi.PT_EXP and np.array_equal(np.float32(p[1].values), np.float32(want.values))
    want = Exp(("a", 4), ("s", 5.0), ("op", abi.OPT_GEQ))
    assert np.array_equal(np.float32(p[2].values), np.float32(want.values))


def test_syntax_forms():
    src = """
    /* block
       comment */
    var a = 1.5;   // trailing
    var b = a;
    axiom Box S;
    terminal P Q2 (1, 0);
    S --> Replicate() { P1 | P2 | P3 | P4 | P5 };
    P1 -> Translate(b, 0, 0) P;
    P2 => Translate(0, b, 0) P;
    P3 ==> Translate(0, 0, b) P;
    P4 >> Q2;
    P5 = P4;
    """
    g = compile_grammar(src)
    r = derive(g, unit_axiom(g))
    # P, Q2 (one declaration): P1..P3 -> P; P4 is Q2's alias and P5 P4's,
    # so both share Q2's entry
    assert r.count(0) == 5
    pos = np.sort(r.instances(0)["model"].reshape(-1, 3, 4)[:, :, 3], axis=0)
    np.testing.assert_array_equal(pos.max(axis=0), [1.5, 1.5, 1.5])
    succ = [s.entry_index for s in g.program.table.entries[0].successors]
    assert succ[3] == succ[4] and len(g.program.table.entries) == 1 + 3 + 2


def test_generate_without_declaration_is_generation_function_0():
    """A symbol that is neither a rule nor a declared terminal, and a rule
    Generate() whose name no terminal declares, compile to Generate with
    generation function 0 and terminal index 0 (visitTerminalSymbol)."""
    g = compile_grammar("""
    axiom Box S;
    S = Replicate() { Undeclared | P6 };
    P6 = Generate();
    """)
    gens = [i for i in range(len(g.program.table.entries)) if op_of(g, i).op_type == abi.GENERATE]
    assert len(gens) == 2
    for i in gens:
        assert op_of(g, i).gen_func_idx == 0
        assert g.program.table.entries[i].parameters[0].values == []


def test_generate_in_rule_uses_its_terminal_declaration():
    g = compile_grammar("""
    axiom Box A;
    terminal A2 (1, 5);
    A = Scale(2, 2, 2) A2;
    A2 = Generate();
    """)
    e = g.program.table.entries[1]
    assert op_of(g, 1).op_type == abi.GENERATE and op_of(g, 1).gen_func_idx == 1
    assert e.parameters[0].values == [5.0]


@pytest.mark.parametrize("src,msg", [
    ("terminal T (1,0); A = T;", "expecting an axiom"),
    ("axiom Box A; terminal T (1,0);", "expecting a production rule"),
    ("axiom Cube A; A = Discard();", "unknown shape type 'Cube'"),
    ("axiom Box A; A = Translate(w, 0, 0) Discard();", "unknown variable 'w'"),
    ("axiom Box A;\nA = Discard()", "line 2"),
    ("axiom Box A; A = Discard(); A = Discard();", "defined twice"),
    ("axiom Box A; A = 50%: B else: C;", "stochastic"),
    ("axiom Box A; A = Translate(0, 0) Discard();", "at least 3 parameters"),
    ("axiom Box A; A = IfSizeLess(X, 1) { Discard() };", "requires 2 successors"),
    ("axiom Box A; A = SubDiv(X) { Discard() | Discard() };", "factor before each successor"),
    ("axiom Box A; A = Replicate() { 1: Discard() };", "cannot have successor parameters"),
    ("axiom Box A; terminal T (1,0); A = Replicate() { T | CompSplit() { T | T | T } };", "but also gets shapes of type Quad"),
    ("axiom Box A; A = CompSplit() { Discard() | Discard() | CompSplit() { Discard() | Discard() | Discard() } };",
     "unsupported procedure (CompSplit on Quad"),
    ("axiom Box A; A = Extrude(1) Discard();", "unsupported procedure (Extrude on Box"),
    ("axiom Box A; A = SetAsDynamicConvexRightPrism((0,1),(1,0)) Discard();", "at least 3 parameters"),
    ("axiom Box A; A = IfCollides(1) { Discard() | Discard() };", "not supported"),
    ("axiom Box A; A = SetAsDynamicConvexPolygon((0,1),(1,0),(0,0)) CompSplit() { Discard() | Discard() | Discard() };",
     "unsupported procedure (CompSplit on DynamicConvexPolygon"),
    ("axiom Box A; A = B; B = A;", "cycle"),
    ("axiom Box A; A = Translate(ShapeNormal(X), 0, 0) Discard();", "only valid inside Exp"),
    ("axiom Box A; /* open", "unterminated comment"),
])
def test_errors(src, msg):
    with pytest.raises(GrammarError, match=msg.replace("(", r"\(")):
        compile_grammar(src)


def test_error_line_number():
    with pytest.raises(GrammarError, match="line 4"):
        compile_grammar("axiom Box A;\nterminal T (1, 0);\nA = Replicate() {\n T | Bogus(1) };\n")


def test_facade_grammar_text_matches_hand_table():
    """The facade grammar written out in the language gives the same
    instances as grammars.facade() (entry numbers aside)."""
    src = """
    axiom Box Lot;
    terminal Wall (1, 0);
    terminal Window (1, 1);
    Lot    = CompSplit() { Discard() | Discard() | Facade };
    Facade = Repeat(Y, 3, ADJUST_TO_FILL) Floor;
    Floor  = Repeat(X, 2.5, ADJUST_TO_FILL) Tile;
    Tile   = IfSizeLess(X, 1.6) { Wall | SubDiv(X) { -1: Wall | 1.2: Window | -1: Wall } };
    """
    g = compile_grammar(src)
    ax = axioms.facade_lots(64)
    a = derive(g, ax)
    b = derive(type(g)(grammars.facade(), [(0, B)], b""), ax)
    for t in range(2):
        x, y = a.instances(t), b.instances(t)
        x["meta"] &= 0xFFFF0000
        y["meta"] &= 0xFFFF0000
        assert len(x) > 0
        assert_same_multiset(x, y, f"facade terminal {t}")


# ---- the reference's interpreter grammars (compiled tables in tests/golden) ----

FIXTURE = grammar_fixture()
SUPPORTED = sorted(k for k, v in FIXTURE.items() if "refused" not in v)


def test_fixture_covers_the_interpreter_grammars():
    # all nine, balcony and suburban_house with the dynamic-shape procedures
    # of the list built with INTERPRETER_SUPPORT_ALL_SHAPES
    assert SUPPORTED == ["3d_tree_12_2", "3d_tree_4_3", "3d_tree_8_3", "balcony", "commercial", "mc_skyscraper",
                         "mc_spaceship", "menger_sponge_3", "suburban_house"]


@pytest.mark.parametrize("name", ["balcony", "suburban_house"])
def test_polygon_grammars_match_the_compiled_scenes(name):
    """balcony and suburban_house derive what the reference's own code
    generator made of them (src/test_runner/Balcony.cuh, SuburbanHouse.cuh)
    on the scenes' lots, terminal by terminal. The generator wrote the
    roof triangle as (0, 0.5), (-0.499, -0.499), (0.499, -0.499); with the
    grammar's (-0.5, -0.5), (0.5, -0.5) the suburban roofs' exact normal
    tests (ShapeNormal(Y) == -1) come out differently for two sides a lot,
    so that grammar is compared with the generator's vertices."""
    from tests.helpers import SceneProgram, scene_axioms
    scene = {"balcony": "Balcony", "suburban_house": "SuburbanHouse"}[name]
    text = open(os.path.join(REF_GRAMMARS, name, "grammar.pga"), encoding="utf-8").read() if HAVE_REF else None
    g = fixture_grammar(name)
    ax = scene_axioms(scene, 6)
    ref = derive(type(g)(SceneProgram(scene), [(0, B)], b""), ax)
    ax["entry_index"] = g.axioms[0][0]
    got = derive(g, ax)
    c_got = [got.count(t) for t in range(abi.MAX_TERMINALS)]
    c_ref = [ref.count(t) for t in range(abi.MAX_TERMINALS)]
    if name == "balcony":
        assert c_got == c_ref
    else:
        roof = 5
        assert [c for t, c in enumerate(c_got) if t != roof] == [c for t, c in enumerate(c_ref) if t != roof]
        assert c_got[roof] < c_ref[roof]
        if text is not None:
            g2 = compile_grammar(text.replace("(0,0.5),(-0.5,-0.5),(0.5,-0.5)", "(0,0.5),(-0.499,-0.499),(0.499,-0.499)"))
            ax["entry_index"] = g2.axioms[0][0]
            assert [derive(g2, ax).count(t) for t in range(abi.MAX_TERMINALS)] == c_ref


def test_setas_orders_vertices_counter_clockwise():
    """buildSetAs*Entry orders the vertices with orderVertices_CCW: the
    bottom-left one first, the rest by the angle from it."""
    g = compile_grammar("axiom Box A; terminal T (1, 0);"
                        "A = SetAsDynamicConvexRightPrism((0.5, 0.5), (-0.5, 0.5), (0, -0.5), (0.5, -0.25)) T;")
    setas = [e for e in g.program.table.entries if len(e.parameters) == 4][0]
    assert [tuple(p.values[:2]) for p in setas.parameters] == [(0.0, -0.5), (0.5, -0.25), (0.5, 0.5), (-0.5, 0.5)]
    assert all(p.type == abi.PT_VEC2 for p in setas.parameters)


@has_reference
@pytest.mark.parametrize("name", sorted(FIXTURE))
def test_fixture_is_a_fresh_compile(name):
    path = os.path.join(REF_GRAMMARS, name, "grammar.pga")
    text = open(path, encoding="utf-8").read()
    assert hashlib.sha256(text.encode()).hexdigest() == FIXTURE[name]["sha256"]
    if "refused" in FIXTURE[name]:
        with pytest.raises(GrammarError):
            compile_grammar(text, name)
        return
    assert compile_grammar(text, name).raw == fixture_grammar(name).raw


@pytest.mark.parametrize("name", sorted(TREE_COUNTS))
def test_tree_counts(name):
    g = fixture_grammar(name)
    r = derive(g, grammar_lots(g, name, 2))
    trunks, leaves = TREE_COUNTS[name]
    assert (r.count(0), r.count(1)) == (2 * trunks, 2 * leaves)
    if name == "3d_tree_4_3":     # configuration.xml maxNumElements
        cfg = FIXTURE[name]["config_max_elements"]
        assert (trunks, leaves) == (cfg["0"], cfg["1"])


def test_spaceship_and_menger_counts():
    """mc_spaceship per ship: M1 = 1 + 2*2 cubes and 2*2 cylinders; M2 =
    1 + 2*2 cylinders and a Top tower of 5 cubes (heights 2.5 down to 0.5 in
    steps of 0.5); M3 = 4 cylinders and a taper."""
    g = fixture_grammar("mc_spaceship")
    r = derive(g, grammar_lots(g, "mc_spaceship", 3))
    assert [r.count(t) for t in range(3)] == [3 * 10, 3 * 13, 3 * 1]
    g = fixture_grammar("menger_sponge_3")
    r = derive(g, grammar_lots(g, "menger_sponge_3", 2))
    assert r.count(0) == 2 * 20 ** 3


@pytest.mark.parametrize("name", ["commercial", "mc_skyscraper"])
def test_building_grammars_derive(name):
    g = fixture_grammar(name)
    r = derive(g, grammar_lots(g, name, 8), twins=1)
    total = sum(r.count(t) for t in range(abi.MAX_TERMINALS))
    assert total > 1000
    from tests.helpers import rel_err
    for t in range(abi.MAX_TERMINALS):
        if r.count(t):
            assert rel_err(r.instances(t)["model"], r.instances_f64(t)[:, :12]) < 1e-5
