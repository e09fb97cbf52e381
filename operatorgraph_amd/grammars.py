"""Grammars and procedure lists used as workloads and fixtures.

* `INTERPRETER_PROCEDURES`: the interpreter's fixed Box/Quad procedure list
  (src/interpreter/PGAFacade.cu:24-62); operator codes of interpreter
  dispatch tables index this list.
* `menger_interpreter(threshold)`: src/interpreter/grammars/menger_sponge_3/
  grammar.pga compiled by hand into an interpreter dispatch table.
* `menger_test_runner()`: the procedure list and dispatch table of
  src/test_runner/MengerSponge.cuh:28-45 (the "Menger-5" scene when the
  axiom has size 81).
* `facade()`: the batched-facade grammar (two terminal types).
* `stochastic()`: per-shape Philox rule choice and split ratios.
* `facade_params()`: the facade with 8 per-lot differentiable parameters.
"""
from . import abi
from .program import (AxiomParam, DispatchTable, Entry, Parameter, Philox, Procedure, Rand,
                      ProcedureList, Program, Scalar, ShapeAttr)

B, Q = abi.BOX, abi.QUAD

# src/interpreter/PGAFacade.cu:24-130: Quad, Box, then the dynamic polygons and prisms
INTERPRETER_PROCEDURES = ProcedureList([
    # QUAD
    Procedure(abi.DISCARD, Q), Procedure(abi.EXTRUDE, Q), Procedure(abi.GENERATE, Q, 0),
    Procedure(abi.GENERATE, Q, 1), Procedure(abi.IF, Q), Procedure(abi.IFSIZELESS, Q),
    Procedure(abi.STOCHASTIC, Q), Procedure(abi.REPEAT, Q), Procedure(abi.REPLICATE, Q),
    Procedure(abi.ROTATE, Q), Procedure(abi.SCALE, Q),
    Procedure(abi.SET_AS_DYNAMIC_CONVEX_POLYGON, Q),
    Procedure(abi.SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, Q),
    Procedure(abi.SET_AS_DYNAMIC_CONCAVE_POLYGON, Q),
    Procedure(abi.SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, Q),
    Procedure(abi.SUBDIV, Q), Procedure(abi.SWAPSIZE, Q), Procedure(abi.TRANSLATE, Q),
    # BOX
    Procedure(abi.COMPSPLIT, B), Procedure(abi.DISCARD, B), Procedure(abi.GENERATE, B, 0),
    Procedure(abi.GENERATE, B, 1), Procedure(abi.IF, B), Procedure(abi.IFSIZELESS, B),
    Procedure(abi.STOCHASTIC, B), Procedure(abi.REPEAT, B), Procedure(abi.REPLICATE, B),
    Procedure(abi.ROTATE, B), Procedure(abi.SCALE, B),
    Procedure(abi.SET_AS_DYNAMIC_CONVEX_POLYGON, B),
    Procedure(abi.SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, B),
    Procedure(abi.SET_AS_DYNAMIC_CONCAVE_POLYGON, B),
    Procedure(abi.SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, B),
    Procedure(abi.SUBDIV, B), Procedure(abi.SWAPSIZE, B), Procedure(abi.TRANSLATE, B),
] + [
    # the dynamic shapes of the list built with INTERPRETER_SUPPORT_ALL_SHAPES
    # (PGAFacade.cu:65-130, 395-470), after Box as there
    Procedure(o, s, g)
    for s, first in ((abi.DYNAMIC_CONVEX_POLYGON, abi.EXTRUDE), (abi.DYNAMIC_POLYGON, abi.EXTRUDE),
                     (abi.DYNAMIC_CONVEX_RIGHT_PRISM, abi.COMPSPLIT), (abi.DYNAMIC_RIGHT_PRISM, abi.COMPSPLIT))
    for o, g in ([(abi.COMPSPLIT, 0), (abi.DISCARD, 0)] if first == abi.COMPSPLIT else [(abi.DISCARD, 0), (abi.EXTRUDE, 0)])
    + [(abi.GENERATE, 0), (abi.GENERATE, 1), (abi.IF, 0), (abi.IFSIZELESS, 0), (abi.STOCHASTIC, 0),
       (abi.REPEAT, 0), (abi.REPLICATE, 0), (abi.ROTATE, 0), (abi.SCALE, 0),
       (abi.SET_AS_DYNAMIC_CONVEX_POLYGON, 0), (abi.SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, 0),
       (abi.SET_AS_DYNAMIC_CONCAVE_POLYGON, 0), (abi.SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, 0),
       (abi.SUBDIV, 0), (abi.SWAPSIZE, 0), (abi.TRANSLATE, 0)]
])


def op(op_type, shape, gen_func=0):
    """Operator code of an interpreter procedure (ProcedureList::indexOf)."""
    idx = INTERPRETER_PROCEDURES.index_of(Procedure(op_type, shape, gen_func))
    if idx < 0:
        raise KeyError((op_type, shape, gen_func))
    return idx


def subdiv_entry(shape, axis, factors, successors):
    """DispatchTableGenerator::buildSubdivideEntry
    (src/compiler/DispatchTableGenerator.cpp:355-383): parameter 0 holds the
    axis and the absolute/relative sums, parameter j the factor and the
    running sums before it."""
    params = []
    abs_sum = rel_sum = 0.0
    for f in factors:
        params.append(Parameter(abi.PT_SCALAR, [f, abs_sum, rel_sum]))
        if f > 0:
            abs_sum += f
        else:
            rel_sum -= f
    p0 = Parameter(abi.PT_SCALAR, [axis, abs_sum, rel_sum])
    return Entry(op(abi.SUBDIV, shape), [p0] + params, [(s, 0) for s in successors])


def menger_interpreter(threshold=0.111):
    """menger_sponge_3/grammar.pga:
        A = IfSizeLess(X, 0.111) { B | SubX };
        ZDiscard = SubDiv(Z) { -1: A | -1: Discard() | -1: A };
        YDiscard = SubDiv(Y) { -1: ZDiscard | -1: Discard() | -1: ZDiscard };
        SubZ = SubDiv(Z) { -1: A | -1: A | -1: A };
        SubY = SubDiv(Y) { -1: SubZ | -1: ZDiscard | -1: SubZ };
        SubX = SubDiv(X) { -1: SubY | -1: YDiscard | -1: SubY };
    terminal B (1,0). Axiom entry 0."""
    A, Bt, SUBX, SUBY, YDISC, SUBZ, ZDISC, DISC = range(8)
    r = [-1.0, -1.0, -1.0]
    table = DispatchTable([
        Entry(op(abi.IFSIZELESS, B), [Scalar(abi.X), Scalar(threshold)], [(Bt, 0), (SUBX, 0)]),
        Entry(op(abi.GENERATE, B, 1), [Scalar(0)], []),
        subdiv_entry(B, abi.X, r, [SUBY, YDISC, SUBY]),
        subdiv_entry(B, abi.Y, r, [SUBZ, ZDISC, SUBZ]),
        subdiv_entry(B, abi.Y, r, [ZDISC, DISC, ZDISC]),
        subdiv_entry(B, abi.Z, r, [A, A, A]),
        subdiv_entry(B, abi.Z, r, [A, DISC, A]),
        Entry(op(abi.DISCARD, B), [], []),
    ])
    return Program(INTERPRETER_PROCEDURES, table, "menger_interpreter")


def menger_test_runner():
    """src/test_runner/MengerSponge.cuh:28-45, verbatim: a four-procedure
    list (static-parameter Subdivide, Discard, IfSizeLess, instanced Generate)
    and a nine-entry table; axioms start at entry 2."""
    procs = ProcedureList([
        Procedure(abi.SUBDIV, B, 0, abi.PROC_STATIC),
        Procedure(abi.DISCARD, B),
        Procedure(abi.IFSIZELESS, B),
        Procedure(abi.GENERATE, B, 1),
    ])
    S = abi.PT_SCALAR
    table = DispatchTable([
        Entry(2, [(S, [0]), (S, [1])], [(1, 0), (2, 0)]),
        Entry(3, [(S, [0])], []),
        Entry(0, [(S, [0]), (S, [-1]), (S, [-1]), (S, [-1])], [(3, 0), (7, 0), (3, 0)]),
        Entry(0, [(S, [1]), (S, [-1]), (S, [-1]), (S, [-1])], [(4, 0), (5, 0), (4, 0)]),
        Entry(0, [(S, [2]), (S, [-1]), (S, [-1]), (S, [-1])], [(0, 0), (0, 0), (0, 0)]),
        Entry(0, [(S, [2]), (S, [-1]), (S, [-1]), (S, [-1])], [(0, 0), (6, 0), (0, 0)]),
        Entry(1, [], []),
        Entry(0, [(S, [1]), (S, [-1]), (S, [-1]), (S, [-1])], [(5, 0), (8, 0), (5, 0)]),
        Entry(1, [], []),
    ])
    return Program(procs, table, "menger_test_runner")


MENGER_TEST_RUNNER_AXIOM_ENTRY = 2


def facade():
    """Batched-facade grammar (two terminal types, Wall=0 and Window=1):
        Lot    = CompSplit() { Discard() | Discard() | Facade };
        Facade = Repeat(Y, 3, ADJUST_TO_FILL) Floor;
        Floor  = Repeat(X, 2.5, ADJUST_TO_FILL) Tile;
        Tile   = IfSizeLess(X, 1.6) { Wall | SubDiv(X) { -1: Wall | 1.2: Window | -1: Wall } };
    Axiom entry 0 (a Box lot)."""
    LOT, QDISC, FACADE, FLOOR, TILE, WALL, SUBWIN, WINDOW = range(8)
    table = DispatchTable([
        Entry(op(abi.COMPSPLIT, B), [], [(QDISC, 0), (QDISC, 0), (FACADE, 0)]),
        Entry(op(abi.DISCARD, Q), [], []),
        Entry(op(abi.REPEAT, Q), [Scalar(abi.Y), Scalar(3.0), Scalar(abi.ADJUST_TO_FILL)], [(FLOOR, 0)]),
        Entry(op(abi.REPEAT, Q), [Scalar(abi.X), Scalar(2.5), Scalar(abi.ADJUST_TO_FILL)], [(TILE, 0)]),
        Entry(op(abi.IFSIZELESS, Q), [Scalar(abi.X), Scalar(1.6)], [(WALL, 0), (SUBWIN, 0)]),
        Entry(op(abi.GENERATE, Q, 1), [Scalar(0)], []),
        subdiv_entry(Q, abi.X, [-1.0, 1.2, -1.0], [WALL, WINDOW, WALL]),
        Entry(op(abi.GENERATE, Q, 1), [Scalar(1)], []),
    ])
    return Program(INTERPRETER_PROCEDURES, table, "facade")


def stochastic():
    """Stochastic grammar (one terminal type). The axiom's Y size is its
    depth budget: every split halves Y, derivation stops once Y < 1.5.
        S      = IfSizeLess(Y, 1.5) { Leaf | Choose };
        Choose = RandomRule(Philox(0,1,#0)) { 0.2: Leaf | 0.4: SplitX | 0.4: SplitZ };
        SplitX = SubDiv(X) { Philox(-0.8,-0.2,#1): Half | -0.5: Half };
        SplitZ = SubDiv(Z) { Philox(-0.8,-0.2,#1): Half | -0.5: Half };
        Half   = Scale(-1, -0.5, -1) S;
        Leaf   = Generate(0);
    Axiom entry 0."""
    S_, CHOOSE, SPLITX, SPLITZ, HALF, LEAF = range(6)
    split = lambda axis: Entry(op(abi.SUBDIV, B), [Scalar(axis), Philox(-0.8, -0.2, 1), Scalar(-0.5)],
                               [(HALF, 0), (HALF, 0)])
    table = DispatchTable([
        Entry(op(abi.IFSIZELESS, B), [Scalar(abi.Y), Scalar(1.5)], [(LEAF, 0), (CHOOSE, 0)]),
        Entry(op(abi.STOCHASTIC, B), [Philox(0.0, 1.0, 0), Scalar(0.2), Scalar(0.4), Scalar(0.4)],
              [(LEAF, 0), (SPLITX, 0), (SPLITZ, 0)]),
        split(abi.X),
        split(abi.Z),
        Entry(op(abi.SCALE, B), [Scalar(-1.0), Scalar(-0.5), Scalar(-1.0)], [(S_, 0)]),
        Entry(op(abi.GENERATE, B, 1), [Scalar(0)], []),
    ])
    return Program(INTERPRETER_PROCEDURES, table, "stochastic")


def combine(*parts, name="combined"):
    """One dispatch table holding several grammars over the interpreter's
    procedure list, so that axioms of all of them derive in one launch (an
    axiom's entry_index selects its grammar). parts: (program,
    terminal_offset) pairs. Successor indices move with their grammar's
    first entry; Generate's terminal index (parameter 0) moves by
    terminal_offset. Returns (program, first entry index of each part)."""
    entries, starts = [], []
    for prog, toff in parts:
        if [p.key() for p in prog.procedures.procedures] != \
                [p.key() for p in INTERPRETER_PROCEDURES.procedures]:
            raise ValueError(f"{prog.name}: not over the interpreter procedure list")
        base = len(entries)
        starts.append(base)
        for e in prog.table.entries:
            params = [Parameter(p.type, list(p.values)) for p in e.parameters]
            if INTERPRETER_PROCEDURES.procedures[e.operator_code].op_type == abi.GENERATE and toff:
                params[0] = Parameter(abi.PT_SCALAR, [params[0].values[0] + toff])
            succ = [(s.entry_index + base if s.entry_index >= 0 else s.entry_index, s.phase_index)
                    for s in e.successors]
            entries.append(Entry(e.operator_code, params, succ, e.edge_indices, e.sub_graph_index))
    return Program(INTERPRETER_PROCEDURES, DispatchTable(entries), name), starts


# per-lot parameter slots of facade_params()
FACADE_PARAM_NAMES = ("left_weight", "right_weight", "floor_height", "tile_width",
                      "window_width", "window_recess", "sill_height", "lintel_height")
FACADE_PARAM_DEFAULTS = (-1.0, -1.0, 2.9, 2.45, 1.2, 0.15, 0.9, 0.4)


def facade_params():
    """Facade with 8 per-lot parameters (PT_AXIOM_PARAM slots):
        Lot    = CompSplit() { Discard() | Discard() | Facade };
        Facade = Repeat(Y, p2, ADJUST_TO_FILL) Floor;
        Floor  = Repeat(X, p3, ADJUST_TO_FILL) Tile;
        Tile   = SubDiv(X) { p0: Wall | p4: Column | p1: Wall };
        Column = SubDiv(Y) { p6: Wall | -1: Recess | p7: Wall };
        Recess = Translate(0, 0, p5) Window;
    Slots (FACADE_PARAM_NAMES): p0/p1 relative wall weights (negative
    values), p2 floor height, p3 tile width, p4 window width, p5 window
    recess offset, p6 sill height, p7 lintel height. ADJUST_TO_FILL keeps
    every floor >= p2 and every tile >= p3, so the absolute factors always
    fit; the repetition counts are discrete, so dL/dp2 and dL/dp3 vanish
    almost everywhere. Terminals: Wall=0, Window=1."""
    LOT, QDISC, FACADE, FLOOR, TILE, WALL, COLUMN, RECESS, WINDOW = range(9)
    table = DispatchTable([
        Entry(op(abi.COMPSPLIT, B), [], [(QDISC, 0), (QDISC, 0), (FACADE, 0)]),
        Entry(op(abi.DISCARD, Q), [], []),
        Entry(op(abi.REPEAT, Q), [Scalar(abi.Y), AxiomParam(2), Scalar(abi.ADJUST_TO_FILL)], [(FLOOR, 0)]),
        Entry(op(abi.REPEAT, Q), [Scalar(abi.X), AxiomParam(3), Scalar(abi.ADJUST_TO_FILL)], [(TILE, 0)]),
        Entry(op(abi.SUBDIV, Q), [Scalar(abi.X), AxiomParam(0), AxiomParam(4), AxiomParam(1)],
              [(WALL, 0), (COLUMN, 0), (WALL, 0)]),
        Entry(op(abi.GENERATE, Q, 1), [Scalar(0)], []),
        Entry(op(abi.SUBDIV, Q), [Scalar(abi.Y), AxiomParam(6), Scalar(-1.0), AxiomParam(7)],
              [(WALL, 0), (RECESS, 0), (WALL, 0)]),
        Entry(op(abi.TRANSLATE, Q), [Scalar(0.0), Scalar(0.0), AxiomParam(5)], [(WINDOW, 0)]),
        Entry(op(abi.GENERATE, Q, 1), [Scalar(1)], []),
    ])
    return Program(INTERPRETER_PROCEDURES, table, "facade_params")


# ---- context-sensitive grammars (Collider / IfCollides, two phases) ----

CONTEXT_PROCEDURES = ProcedureList([
    Procedure(abi.REPLICATE, B), Procedure(abi.SCALE, B), Procedure(abi.TRANSLATE, B),
    Procedure(abi.ROTATE, B), Procedure(abi.COLLIDER, B), Procedure(abi.GENERATE, B, 1),
    Procedure(abi.REPEAT, B), Procedure(abi.IFCOLLIDES, B), Procedure(abi.COMPSPLIT, B),
    Procedure(abi.DISCARD, Q), Procedure(abi.COLLIDER, Q), Procedure(abi.GENERATE, Q, 1),
    Procedure(abi.REPLICATE, Q), Procedure(abi.SCALE, Q), Procedure(abi.TRANSLATE, Q),
    Procedure(abi.ROTATE, Q), Procedure(abi.IFCOLLIDES, Q),
])


def _cop(op_type, shape, gen_func=0):
    idx = CONTEXT_PROCEDURES.index_of(Procedure(op_type, shape, gen_func))
    if idx < 0:
        raise KeyError((op_type, shape, gen_func))
    return idx


def gabled_houses():
    """A house per lot with a gabled roof, over the interpreter procedures:

        Lot    = SubDiv(Y) { -0.7: Body | -0.3: Roof };
        Body   = CompSplit() { Discard() | Discard() | Repeat(X, 2, ADJUST_TO_FILL) Wall };
        Roof   = SetAsDynamicConvexRightPrism((-0.5, -0.5), (0.5, -0.5), (0, 0.5))
                 Rotate(90, 0, 0) Scale(ShapeSize(X), ShapeSize(Z), ShapeSize(Y))
                 CompSplit() { Gable | Gable | RoofSide };
        Wall (1, 0); Gable (1, 1); RoofSide (1, 2)

    (the roof of the balcony grammar, RoofStructure1, on the top 30 % of the
    lot): every lot gives two gables and three roof sides, and
    2 max(1, floor(w / 2)) + 2 max(1, floor(d / 2)) wall tiles."""
    B_, Q_ = abi.BOX, abi.QUAD
    CR, CP = abi.DYNAMIC_CONVEX_RIGHT_PRISM, abi.DYNAMIC_CONVEX_POLYGON
    tri = [(-0.5, -0.5), (0.5, -0.5), (0.0, 0.5)]
    e = [
        subdiv_entry(B_, 1, [-0.7, -0.3], [1, 2]),                                              # 0 Lot
        Entry(op(abi.COMPSPLIT, B_), [], [(3, 0), (3, 0), (4, 0)]),                             # 1 Body
        Entry(op(abi.SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, B_), [Parameter(abi.PT_VEC2, [x, y]) for x, y in tri],
              [(6, 0)]),                                                                        # 2 Roof
        Entry(op(abi.DISCARD, Q_), [], []),                                                     # 3
        Entry(op(abi.REPEAT, Q_), [Scalar(0.0), Scalar(2.0), Scalar(float(abi.ADJUST_TO_FILL))], [(5, 0)]),  # 4
        Entry(op(abi.GENERATE, Q_, 1), [Scalar(0.0)], []),                                      # 5 Wall
        Entry(op(abi.ROTATE, CR), [Scalar(90.0), Scalar(0.0), Scalar(0.0)], [(7, 0)]),          # 6
        Entry(op(abi.SCALE, CR), [ShapeAttr(12), ShapeAttr(14), ShapeAttr(13)], [(8, 0)]),      # 7
        Entry(op(abi.COMPSPLIT, CR), [], [(9, 0), (9, 0), (10, 0)]),                            # 8
        Entry(op(abi.GENERATE, CP, 1), [Scalar(1.0)], []),                                      # 9 Gable
        Entry(op(abi.GENERATE, Q_, 1), [Scalar(2.0)], []),                                      # 10 RoofSide
    ]
    return Program(INTERPRETER_PROCEDURES, DispatchTable(e), "gabled_houses")


def context_city(walls=False):
    """A two-phase city block (the reference's context-sensitive operators,
    IfCollides.cuh / Collider.cuh, under MultiPhaseEvaluator):
        Lot      = Replicate { Building (phase 0) | Yard (phase 1) };
        Building = Scale(-0.45, -1, -0.45) Translate(r, 0, r)
                   Rotate(0, 90 r', 0) Collider(1) Generate(0);
        Yard     = Repeat(X, 1, ADJUST_TO_FILL) Repeat(Z, 1, ADJUST_TO_FILL)
                   Scale(-0.5, -0.3, -0.5) IfCollides(1) { Generate(2) | Generate(1) };
    r, r' are Rand(-1.5, 1.5) and Rand(0, 1) of the shape's seed. Phase 0
    derives the buildings and stores them as colliders; phase 1 places one
    prop per yard cell, blocked (terminal 2) where it overlaps a building,
    free (terminal 1) elsewhere. walls=True also splits each building into
    faces stored under tag 2 (Quad colliders) and tests props against them
    first: IfCollides(2) { Generate(2) | IfCollides(1) {...} }."""
    (LOT, BUILDING, MOVE, TURN, STORE, GEN_B, YARD_X, YARD_Z, PROP, TEST, GEN_BLOCK, GEN_FREE,
     SPLIT, QDISC, WALL, GEN_WALL, TEST_WALL) = range(17)
    P0, P1 = 0, 1
    after_store = SPLIT if walls else GEN_B
    first_test = TEST_WALL if walls else TEST
    table = DispatchTable([
        Entry(_cop(abi.REPLICATE, B), [], [(BUILDING, P0), (YARD_X, P1)]),
        Entry(_cop(abi.SCALE, B), [Scalar(-0.45), Scalar(-1.0), Scalar(-0.45)], [(MOVE, P0)]),
        Entry(_cop(abi.TRANSLATE, B), [Rand(-1.5, 1.5), Scalar(0.0), Rand(-1.5, 1.5)], [(TURN, P0)]),
        Entry(_cop(abi.ROTATE, B), [Scalar(0.0), Rand(0.0, 90.0), Scalar(0.0)], [(STORE, P0)]),
        Entry(_cop(abi.COLLIDER, B), [Scalar(1)], [(after_store, P0)]),
        Entry(_cop(abi.GENERATE, B, 1), [Scalar(0)], []),
        Entry(_cop(abi.REPEAT, B), [Scalar(abi.X), Scalar(1.0), Scalar(abi.ADJUST_TO_FILL)], [(YARD_Z, P1)]),
        Entry(_cop(abi.REPEAT, B), [Scalar(abi.Z), Scalar(1.0), Scalar(abi.ADJUST_TO_FILL)], [(PROP, P1)]),
        Entry(_cop(abi.SCALE, B), [Scalar(-0.5), Scalar(-0.3), Scalar(-0.5)], [(first_test, P1)]),
        Entry(_cop(abi.IFCOLLIDES, B), [Scalar(1)], [(GEN_BLOCK, P1), (GEN_FREE, P1)]),
        Entry(_cop(abi.GENERATE, B, 1), [Scalar(2)], []),
        Entry(_cop(abi.GENERATE, B, 1), [Scalar(1)], []),
        Entry(_cop(abi.COMPSPLIT, B), [], [(QDISC, P0), (QDISC, P0), (WALL, P0)]),
        Entry(_cop(abi.DISCARD, Q), [], []),
        Entry(_cop(abi.COLLIDER, Q), [Scalar(2)], [(GEN_WALL, P0)]),
        Entry(_cop(abi.GENERATE, Q, 1), [Scalar(3)], []),
        Entry(_cop(abi.IFCOLLIDES, B), [Scalar(2)], [(GEN_BLOCK, P1), (TEST, P1)]),
    ])
    return Program(CONTEXT_PROCEDURES, table, "context_city_walls" if walls else "context_city")


# per-axiom slots of collision_pairs(): size, rotation (degrees), translation, tag
COLLISION_PAIR_SLOTS = 10


def collision_pairs():
    """Shape pairs, the collision test itself as a program. Each pair is two
    axioms with per-axiom parameters (PT_AXIOM_PARAM) p0..p2 size, p3..p5
    rotation in degrees, p6..p8 translation, p9 tag:
        A = Scale(p0..p2) Rotate(p3..p5) Translate(p6..p8) Collider(p9) Generate(0);
        B = Replicate { Moved (phase 1) };
        Moved = Scale(p0..p2) Rotate(p3..p5) Translate(p6..p8)
                IfCollides(p9) { Generate(1) | Generate(2) };
    A stores its shape in phase 0, B's shape asks in phase 1; a tag per pair
    keeps each pair to itself. Entry points (COLLISION_PAIR_ENTRIES): A and B
    on a Box or on a Quad; a Quad's third size component is its z extent (0:
    planar)."""
    from .program import AxiomParam as AP
    entries = []
    for s in (B, Q):
        base = len(entries)
        SC, RO, TR, CO, GA = (base + k for k in range(5))
        entries += [
            Entry(_cop(abi.SCALE, s), [AP(0), AP(1), AP(2)], [(RO, 0)]),
            Entry(_cop(abi.ROTATE, s), [AP(3), AP(4), AP(5)], [(TR, 0)]),
            Entry(_cop(abi.TRANSLATE, s), [AP(6), AP(7), AP(8)], [(CO, 0)]),
            Entry(_cop(abi.COLLIDER, s), [AP(9)], [(GA, 0)]),
            Entry(_cop(abi.GENERATE, s, 1), [Scalar(0)], []),
        ]
    for s in (B, Q):
        base = len(entries)
        RE, SC, RO, TR, IC, G1, G2 = (base + k for k in range(7))
        entries += [
            Entry(_cop(abi.REPLICATE, s), [], [(SC, 1)]),
            Entry(_cop(abi.SCALE, s), [AP(0), AP(1), AP(2)], [(RO, 1)]),
            Entry(_cop(abi.ROTATE, s), [AP(3), AP(4), AP(5)], [(TR, 1)]),
            Entry(_cop(abi.TRANSLATE, s), [AP(6), AP(7), AP(8)], [(IC, 1)]),
            Entry(_cop(abi.IFCOLLIDES, s), [AP(9)], [(G1, 1), (G2, 1)]),
            Entry(_cop(abi.GENERATE, s, 1), [Scalar(1)], []),
            Entry(_cop(abi.GENERATE, s, 1), [Scalar(2)], []),
        ]
    return Program(CONTEXT_PROCEDURES, DispatchTable(entries), "collision_pairs")


# collision_pairs() entry points: (role, shape type) -> entry index
COLLISION_PAIR_ENTRIES = {("A", B): 0, ("A", Q): 5, ("B", B): 10, ("B", Q): 17}
