"""Seeded random context-sensitive programs (Collider, IfCollides, phases),
for parity tests of the multi-phase derivation.

A program is a tree of entries (no cycles, at most three children per
fan-out) over Box and Quad shapes: ComponentSplit turns Boxes into Quads,
Extrude Quads into Boxes. The axiom entry replicates its Box into two or
three branches of phases 0, 1 and 2, so later branches meet the colliders
of earlier ones. Below that every successor names a phase: mostly its
parent's (explicitly or with PGA_PHASE_INHERIT), sometimes a later one (up
to 3), now and then an earlier one (whose shapes the derivation drops) or
an unassigned -1 (phase 0). Collider and IfCollides draw their tags from
{1, 2} (mostly 1), so tests meet colliders of their tag from earlier phases, of the
same phase (invisible) and of none. Translations reach a few units and
rotations 90 degrees, so a test's outcome varies over the axioms.
"""
import numpy as np

from operatorgraph_amd import abi
from operatorgraph_amd.program import DispatchTable, Entry, Procedure, ProcedureList, Program, Scalar

B, Q = abi.BOX, abi.QUAD

PROCEDURES = ProcedureList([
    Procedure(abi.REPLICATE, B), Procedure(abi.SCALE, B), Procedure(abi.TRANSLATE, B), Procedure(abi.ROTATE, B),
    Procedure(abi.COLLIDER, B), Procedure(abi.IFCOLLIDES, B), Procedure(abi.GENERATE, B, 1),
    Procedure(abi.DISCARD, B), Procedure(abi.SUBDIV, B), Procedure(abi.REPEAT, B), Procedure(abi.COMPSPLIT, B),
    Procedure(abi.REPLICATE, Q), Procedure(abi.SCALE, Q), Procedure(abi.TRANSLATE, Q), Procedure(abi.ROTATE, Q),
    Procedure(abi.COLLIDER, Q), Procedure(abi.IFCOLLIDES, Q), Procedure(abi.GENERATE, Q, 1),
    Procedure(abi.DISCARD, Q), Procedure(abi.EXTRUDE, Q),
])


def _op(op, shape, gen=0):
    i = PROCEDURES.index_of(Procedure(op, shape, gen))
    assert i >= 0, (op, shape)
    return i


class _Gen:
    def __init__(self, seed):
        self.rng = np.random.default_rng(seed)
        self.entries = []

    def phase(self, cur):
        u = self.rng.random()
        if u < 0.35:
            return cur
        if u < 0.65:
            return abi.PHASE_INHERIT
        if u < 0.85:
            return min(cur + 1 + int(self.rng.integers(2)), 3)
        if u < 0.93:
            return -1
        return max(cur - 1, 0)

    def tag(self):
        return 1.0 if self.rng.random() < 0.8 else 2.0

    def new(self):
        self.entries.append(None)
        return len(self.entries) - 1

    def child(self, shape, depth, cur):
        ph = self.phase(cur)
        nxt = cur if ph < 0 else ph
        return (self.node(shape, depth - 1, nxt), ph)

    def node(self, shape, depth, cur):
        r = self.rng
        idx = self.new()
        if depth <= 0:
            self.entries[idx] = Entry(_op(abi.GENERATE, shape, 1), [Scalar(float(r.integers(4)))], [])
            return idx
        kinds = ["scale", "translate", "rotate", "collider", "ifcollides", "replicate", "generate"]
        kinds += ["subdiv", "repeat", "compsplit"] if shape == B else ["extrude"]
        weights = np.array([1, 2.5, 1, 2, 3, 1, 0.4] + ([1, 0.7, 0.7] if shape == B else [0.7]), dtype=float)
        k = kinds[r.choice(len(kinds), p=weights / weights.sum())]
        if k == "generate":
            self.entries[idx] = Entry(_op(abi.GENERATE, shape, 1), [Scalar(float(r.integers(4)))], [])
        elif k == "scale":
            v = [float(-r.uniform(0.4, 1.3)) if r.random() < 0.7 else float(r.uniform(0.5, 4.0)) for _ in range(3)]
            if shape == Q:
                v[2] = 0.0 if r.random() < 0.5 else float(r.uniform(0.2, 1.0))
            self.entries[idx] = Entry(_op(abi.SCALE, shape), [Scalar(x) for x in v],
                                      [self.child(shape, depth, cur)])
        elif k == "translate":
            v = [float(r.uniform(-4.0, 4.0)) for _ in range(3)]
            self.entries[idx] = Entry(_op(abi.TRANSLATE, shape), [Scalar(x) for x in v],
                                      [self.child(shape, depth, cur)])
        elif k == "rotate":
            v = [float(r.uniform(0, 90)) if r.random() < 0.5 else 0.0 for _ in range(3)]
            self.entries[idx] = Entry(_op(abi.ROTATE, shape), [Scalar(x) for x in v],
                                      [self.child(shape, depth, cur)])
        elif k == "collider":
            self.entries[idx] = Entry(_op(abi.COLLIDER, shape), [Scalar(self.tag())],
                                      [self.child(shape, depth, cur)])
        elif k == "ifcollides":
            self.entries[idx] = Entry(_op(abi.IFCOLLIDES, shape), [Scalar(self.tag())],
                                      [self.child(shape, depth, cur), self.child(shape, depth, cur)])
        elif k == "replicate":
            n = int(r.integers(2, 4))
            self.entries[idx] = Entry(_op(abi.REPLICATE, shape), [],
                                      [self.child(shape, depth, cur) for _ in range(n)])
        elif k == "subdiv":
            from operatorgraph_amd.grammars import subdiv_entry
            n = int(r.integers(2, 4))
            succ = [self.child(shape, depth, cur) for _ in range(n)]
            e = subdiv_entry(shape, int(r.integers(3)), [-1.0] * n, [s for s, _ in succ])
            e.successors[:] = [type(e.successors[0])(s, p) for s, p in succ]
            e.operator_code = _op(abi.SUBDIV, shape)
            self.entries[idx] = e
        elif k == "repeat":
            self.entries[idx] = Entry(_op(abi.REPEAT, shape),
                                      [Scalar(float(r.integers(3))), Scalar(float(r.uniform(1.5, 3.0))),
                                       Scalar(float(abi.ADJUST_TO_FILL))],
                                      [self.child(shape, depth, cur)])
        elif k == "compsplit":
            self.entries[idx] = Entry(_op(abi.COMPSPLIT, B), [],
                                      [self.child(Q, depth, cur), (self.discard(Q), cur), self.child(Q, depth, cur)])
        else:   # extrude
            self.entries[idx] = Entry(_op(abi.EXTRUDE, Q), [Scalar(2.0), Scalar(float(r.uniform(0.5, 2.0)))],
                                      [self.child(B, depth, cur)])
        return idx

    def discard(self, shape):
        idx = self.new()
        self.entries[idx] = Entry(_op(abi.DISCARD, shape), [], [])
        return idx


def random_program(seed, depth=5):
    """A seeded random multi-phase program with its axiom entry at 0."""
    g = _Gen(seed)
    root = g.new()
    n = 2 + int(g.rng.integers(2))
    g.entries[root] = Entry(_op(abi.REPLICATE, B), [], [(g.node(B, depth - 1, ph), ph) for ph in range(n)])
    return Program(PROCEDURES, DispatchTable(g.entries), f"ctx_fuzz{seed}")


def axioms(n, seed):
    """n Box axioms at entry 0: sizes in [2, 6), spread over a square of side
    2 sqrt(n) (neighbours overlap now and then)."""
    from operatorgraph_amd import axioms as A
    rng = np.random.default_rng(1000 + seed)
    a = A.empty(n)
    a["size"] = rng.uniform(2.0, 6.0, (n, 3)).astype(np.float32)
    side = 2.0 * np.sqrt(n)
    a["model"][:, 3] = rng.uniform(0, side, n).astype(np.float32)
    a["model"][:, 11] = rng.uniform(0, side, n).astype(np.float32)
    a["seed"] = rng.random(n).astype(np.float32)
    a["custom"] = a["seed"]
    a["entry_index"] = 0
    a["shape_type"] = B
    return a


def subtree(program, e):
    """The entries reachable from entry e (the programs are trees)."""
    out, todo = set(), [e]
    while todo:
        k = todo.pop()
        if k not in out:
            out.add(k)
            todo += [s.entry_index for s in program.table.entries[k].successors]
    return out


def branch_outcomes(program, result, n_terms=4):
    """For each IfCollides entry: the terminals derived below its first
    (collides) and its second successor, counted by their entries."""
    ifc = {PROCEDURES.index_of(Procedure(abi.IFCOLLIDES, s)) for s in (B, Q)}
    per_entry = {}
    for t in range(n_terms):
        for m in result.instances(t)["meta"]:
            per_entry[int(m) & 0xFFFF] = per_entry.get(int(m) & 0xFFFF, 0) + 1
    out = []
    for i, e in enumerate(program.table.entries):
        if e.operator_code in ifc:
            yes, no = (sum(per_entry.get(k, 0) for k in subtree(program, s.entry_index)) for s in e.successors)
            out.append((i, yes, no))
    return out
