"""Hand-built programs over dynamic polygons and right prisms (SetAs*,
ComponentSplit of a prism, Extrude of a polygon, inverted caps), and the
axioms they run on.

* `prism_program(vertices)`: Box -> SetAsDynamicConvexRightPrism(vertices)
  -> ComponentSplit { top: Generate 0 | bottom: Generate 1 | sides:
  Generate 2 }.
* `roof_program()`: the balcony grammar's roof
  (src/interpreter/grammars/balcony/grammar.pga, RoofStructure1 and Roof1):
  the triangle prism translated, rotated and scaled, its caps RoofSide, its
  sides split by ShapeNormal into a discarded bottom and two roof halves.
* `recap_program(vertices)`: a prism whose caps are extruded again and
  split, so the bottom cap's inverted vertex order reaches a second split.
* `polygon_ops_program(vertices)`: the generic operators (Subdivide,
  Repeat, IfSizeLess, RandomRule, SwapSize, Rotate, Translate, Scale) on
  polygon and prism shapes, and SetAs on a Quad.
"""
import numpy as np

from operatorgraph_amd import abi, axioms
from operatorgraph_amd.program import (DispatchTable, Entry, Exp, Procedure, ProcedureList, Program, Scalar, ShapeAttr,
                                       Vec2)

B, Q = abi.BOX, abi.QUAD
CP, CR = abi.DYNAMIC_CONVEX_POLYGON, abi.DYNAMIC_CONVEX_RIGHT_PRISM
P, R = abi.DYNAMIC_POLYGON, abi.DYNAMIC_RIGHT_PRISM

_OPS = [abi.DISCARD, abi.GENERATE, abi.IF, abi.IFSIZELESS, abi.STOCHASTIC, abi.REPEAT, abi.REPLICATE, abi.ROTATE,
        abi.SCALE, abi.SUBDIV, abi.SWAPSIZE, abi.TRANSLATE, abi.SET_AS_DYNAMIC_CONVEX_POLYGON,
        abi.SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, abi.SET_AS_DYNAMIC_CONCAVE_POLYGON,
        abi.SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM]


def _procedures():
    procs = []
    for shape in (Q, B, CP, P, CR, R):
        for op in _OPS:
            procs.append(Procedure(op, shape, 1 if op == abi.GENERATE else 0))
        if shape in (Q, CP, P):
            procs.append(Procedure(abi.EXTRUDE, shape))
        if shape in (B, CR, R):
            procs.append(Procedure(abi.COMPSPLIT, shape))
    return ProcedureList(procs)


PROCEDURES = _procedures()


def op(o, shape):
    i = PROCEDURES.index_of(Procedure(o, shape, 1 if o == abi.GENERATE else 0))
    assert i >= 0, (o, shape)
    return i


def gen(shape, t):
    return Entry(op(abi.GENERATE, shape), [Scalar(float(t))], [])


def regular(n, phase=0.0, radius=0.5):
    """A counter-clockwise regular n-gon in the unit square."""
    a = phase + 2.0 * np.pi * np.arange(n) / n
    return [(float(np.float32(radius * np.cos(t))), float(np.float32(radius * np.sin(t)))) for t in a]


def random_convex(n, rng):
    """n vertices of a random convex polygon (sorted angles on an ellipse),
    counter-clockwise, in [-0.5, 0.5]^2."""
    t = np.sort(rng.uniform(0.0, 2.0 * np.pi, n))
    rx, ry = rng.uniform(0.2, 0.5, 2)
    return [(float(np.float32(rx * np.cos(a))), float(np.float32(ry * np.sin(a)))) for a in t]


def setas(shape_from, vertices, succ, prism=True, convex=True):
    o = {(True, True): abi.SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, (False, True): abi.SET_AS_DYNAMIC_CONVEX_POLYGON,
         (True, False): abi.SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, (False, False): abi.SET_AS_DYNAMIC_CONCAVE_POLYGON}
    return Entry(op(o[(prism, convex)], shape_from), [Vec2(x, y) for x, y in vertices], [(succ, 0)])


def prism_program(vertices, convex=True):
    pr, cap = (CR, CP) if convex else (R, P)
    return Program(PROCEDURES, DispatchTable([
        setas(B, vertices, 1, convex=convex),                       # 0
        Entry(op(abi.COMPSPLIT, pr), [], [(2, 0), (3, 0), (4, 0)]),  # 1
        gen(cap, 0),                                                # 2 top
        gen(cap, 1),                                                # 3 bottom
        gen(Q, 2),                                                  # 4 sides
    ]), f"prism{len(vertices)}")


def roof_program():
    """RoofStructure1 = SetAsDynamicConvexRightPrism((0, 0.5), (-0.5, -0.5),
    (0.5, -0.5)) Translate(0, 2.4, 0) Rotate(90, 0, 0) Scale(ShapeSize(X),
    ShapeSize(Z), 5) CompSplit() { RoofSide | RoofSide | Roof1 };
    Roof1 = If(ShapeNormal(Y) == -1) { Discard() | If(ShapeNormal(X) < 0)
    { Translate(0.5, 0, 0) Scale(ShapeSize(X) + 1, -1, -1) Roof |
      Translate(-0.5, 0, 0) Scale(ShapeSize(X) + 1, -1, -1) Roof } }
    with RoofSide terminal 0 and Roof terminal 1."""
    normal_x, normal_y = 2, 6            # Shape::at indices: model elements (0, 2) and (1, 2)
    size_x, size_z = 12, 14
    e = [
        setas(B, [(0.0, 0.5), (-0.5, -0.5), (0.5, -0.5)], 1),                                   # 0
        Entry(op(abi.TRANSLATE, CR), [Scalar(0.0), Scalar(2.4), Scalar(0.0)], [(2, 0)]),        # 1
        Entry(op(abi.ROTATE, CR), [Scalar(90.0), Scalar(0.0), Scalar(0.0)], [(3, 0)]),          # 2
        Entry(op(abi.SCALE, CR), [ShapeAttr(size_x), ShapeAttr(size_z), Scalar(5.0)], [(4, 0)]),      # 3
        Entry(op(abi.COMPSPLIT, CR), [], [(5, 0), (5, 0), (6, 0)]),                            # 4
        gen(CP, 0),                                                                             # 5 RoofSide
        Entry(op(abi.IF, Q), [Exp(("a", normal_y), ("s", -1.0), ("op", abi.OPT_EQ))],
              [(7, 0), (8, 0)]),                                                                # 6 Roof1
        Entry(op(abi.DISCARD, Q), [], []),                                                      # 7
        Entry(op(abi.IF, Q), [Exp(("a", normal_x), ("s", 0.0), ("op", abi.OPT_LT))],
              [(9, 0), (12, 0)]),                                                               # 8
        Entry(op(abi.TRANSLATE, Q), [Scalar(0.5), Scalar(0.0), Scalar(0.0)], [(10, 0)]),        # 9
        Entry(op(abi.SCALE, Q), [Exp(("a", size_x), ("s", 1.0), ("op", abi.OPT_ADD)),
                                 Scalar(-1.0), Scalar(-1.0)], [(11, 0)]),                       # 10
        gen(Q, 1),                                                                              # 11 Roof
        Entry(op(abi.TRANSLATE, Q), [Scalar(-0.5), Scalar(0.0), Scalar(0.0)], [(13, 0)]),       # 12
        Entry(op(abi.SCALE, Q), [Exp(("a", size_x), ("s", 1.0), ("op", abi.OPT_ADD)),
                                 Scalar(-1.0), Scalar(-1.0)], [(11, 0)]),                       # 13
    ]
    return Program(PROCEDURES, DispatchTable(e), "roof")


def recap_program(vertices):
    """Box -> prism -> ComponentSplit { top: Extrude(1) -> T | bottom:
    Extrude(1) -> T | sides: Generate 0 }; T = ComponentSplit { Generate 1 |
    Generate 2 | Generate 3 } (the caps' prisms, the bottom one inverted)."""
    e = [
        setas(B, vertices, 1),                                                     # 0
        Entry(op(abi.COMPSPLIT, CR), [], [(2, 0), (2, 0), (3, 0)]),               # 1
        Entry(op(abi.EXTRUDE, CP), [Scalar(2.0), Scalar(1.0)], [(4, 0)]),         # 2
        gen(Q, 0),                                                                 # 3
        Entry(op(abi.COMPSPLIT, CR), [], [(5, 0), (6, 0), (7, 0)]),               # 4
        gen(CP, 1),                                                                # 5
        gen(CP, 2),                                                                # 6
        gen(Q, 3),                                                                 # 7
    ]
    return Program(PROCEDURES, DispatchTable(e), "recap")


def polygon_ops_program(vertices):
    """Generic operators on polygons and prisms:
    Box -> Subdivide(X) { -1: A | -1: SetAs polygon on a Quad face path }
    A = SetAs prism -> Repeat(Y, 1.5, ADJUST_TO_FILL) -> RandomRule(50%) {
        Rotate(0, 30, 0) S | SwapSize(Z, Y, X) S }
    S = IfSizeLess(X, 2) { ComponentSplit -> terminals | Scale(-0.5, -1, -1)
        Translate(0.2, 0, 0.1) ComponentSplit -> terminals }
    The second half: ComponentSplit (Box) -> sides -> SetAs polygon ->
    Subdivide(X) { -1: Generate 4 | -1: Extrude(0.5) -> prism split }."""
    e = [
        Entry(op(abi.SUBDIV, B), [Scalar(0.0), Scalar(-1.0), Scalar(-1.0)], [(1, 0), (11, 0)]),      # 0
        setas(B, vertices, 2),                                                                      # 1
        Entry(op(abi.REPEAT, CR), [Scalar(1.0), Scalar(1.5), Scalar(float(abi.ADJUST_TO_FILL))],
              [(3, 0)]),                                                                            # 2
        Entry(op(abi.STOCHASTIC, CR), [Scalar(0.0), Scalar(50.0), Scalar(50.0)], [(4, 0), (5, 0)]),  # 3
        Entry(op(abi.ROTATE, CR), [Scalar(0.0), Scalar(30.0), Scalar(0.0)], [(6, 0)]),               # 4
        Entry(op(abi.SWAPSIZE, CR), [Scalar(2.0), Scalar(1.0), Scalar(0.0)], [(6, 0)]),              # 5
        Entry(op(abi.IFSIZELESS, CR), [Scalar(0.0), Scalar(2.0)], [(7, 0), (8, 0)]),                 # 6
        Entry(op(abi.COMPSPLIT, CR), [], [(10, 0), (10, 0), (16, 0)]),                              # 7
        Entry(op(abi.SCALE, CR), [Scalar(-0.5), Scalar(-1.0), Scalar(-1.0)], [(9, 0)]),              # 8
        Entry(op(abi.TRANSLATE, CR), [Scalar(0.2), Scalar(0.0), Scalar(0.1)], [(7, 0)]),             # 9
        gen(CP, 0),                                                                                 # 10
        Entry(op(abi.COMPSPLIT, B), [], [(12, 0), (17, 0), (12, 0)]),                               # 11
        setas(Q, vertices, 13, prism=False),                                                        # 12
        Entry(op(abi.SUBDIV, CP), [Scalar(0.0), Scalar(-1.0), Scalar(-1.0)], [(14, 0), (15, 0)]),    # 13
        gen(CP, 4),                                                                                 # 14
        Entry(op(abi.EXTRUDE, CP), [Scalar(2.0), Scalar(0.5)], [(7, 0)]),                           # 15
        gen(Q, 2),                                                                                  # 16
        Entry(op(abi.DISCARD, Q), [], []),                                                          # 17
    ]
    return Program(PROCEDURES, DispatchTable(e), "polygon_ops")


def boxes(n, seed, rotate=True):
    """n Box axioms at entry 0: sizes in [1, 6), random yaw (or none),
    spread over a line."""
    rng = np.random.default_rng(seed)
    a = axioms.empty(n)
    a["size"] = rng.uniform(1.0, 6.0, (n, 3)).astype(np.float32)
    yaw = rng.uniform(0.0, 2.0 * np.pi, n) if rotate else np.zeros(n)
    c, s = np.cos(yaw).astype(np.float32), np.sin(yaw).astype(np.float32)
    a["model"][:, 0], a["model"][:, 2] = c, s
    a["model"][:, 8], a["model"][:, 10] = -s, c
    a["model"][:, 3] = (np.arange(n) * 8.0).astype(np.float32)
    a["seed"] = rng.random(n).astype(np.float32)
    a["custom"] = a["seed"]
    a["entry_index"] = 0
    a["shape_type"] = B
    return a
