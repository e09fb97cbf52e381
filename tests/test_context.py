"""Context sensitivity in the oracle: the collision test (Collision.cuh) and
multi-phase derivation with Collider / IfCollides (MultiPhaseEvaluator,
GPUEvaluator.cuh:395-456; Collider.cuh; IfCollides.cuh).

The separating-axis test is pinned against a linear program (two convex
boxes or planar rectangles meet iff a point lies in both; scipy's HiGHS
decides it), on pairs that clear the boundary by a margin; the Quad
z-extent rule, which departs from the geometry, against the closed form of
the reference's face-axis test. Phase semantics are pinned by closed-form
counts of small programs."""
import numpy as np
import pytest
from scipy.optimize import linprog

from operatorgraph_amd import abi, axioms, grammars
from operatorgraph_amd.grammars import CONTEXT_PROCEDURES, _cop
from operatorgraph_amd.program import DispatchTable, Entry, Program, Scalar
from oracle import oracle as O

B, Q = abi.BOX, abi.QUAD


def rotation(rng):
    """A uniformly random rotation (unit quaternion), FP32."""
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def model(rot, pos):
    m = np.zeros((3, 4))
    m[:, :3] = rot
    m[:, 3] = pos
    return m.astype(np.float32).reshape(12)


def lp_margin(am, ah, bm, bh):
    """Largest t such that a point lies in both shapes shrunk to (1 - t)
    of their half extents (t < 0: grown); shapes are centre + columns of
    the model times half extents, zero half extents for a planar side."""
    A = am.astype(np.float64).reshape(3, 4)
    Bm = bm.astype(np.float64).reshape(3, 4)
    CA = A[:, :3] * np.asarray(ah, dtype=np.float64)
    CB = Bm[:, :3] * np.asarray(bh, dtype=np.float64)
    # variables u(3), w(3), t: A.c + CA u = B.c + CB w, |u_i|, |w_i| <= 1 - t
    A_eq = np.hstack([CA, -CB, np.zeros((3, 1))])
    b_eq = Bm[:, 3] - A[:, 3]
    rows = []
    for k in range(6):
        for s in (1.0, -1.0):
            r = np.zeros(7)
            r[k] = s
            r[6] = 1.0
            rows.append(r)
    res = linprog(c=[0] * 6 + [-1.0], A_ub=np.array(rows), b_ub=np.ones(12), A_eq=A_eq, b_eq=b_eq,
                  bounds=[(None, None)] * 6 + [(-5.0, 1.0)], method="highs")
    if res.status == 2:       # infeasible even grown 6x: far apart
        return -5.0
    assert res.status == 0, res.message
    return -res.fun


@pytest.mark.parametrize("ta,tb", [(B, B), (B, Q), (Q, B), (Q, Q)])
def test_separating_axis_test_matches_linear_program(ta, tb):
    rng = np.random.default_rng(100 + 10 * ta + tb)
    robust = {True: 0, False: 0}
    for _ in range(400):
        sa = rng.uniform(0.4, 2.0, 3).astype(np.float32)
        sb = rng.uniform(0.4, 2.0, 3).astype(np.float32)
        if ta == Q:
            sa[2] = 0.0
        if tb == Q:
            sb[2] = 0.0
        am = model(rotation(rng), rng.uniform(-0.2, 0.2, 3))
        bm = model(rotation(rng), rng.uniform(-1.6, 1.6, 3))
        t = lp_margin(am, sa * 0.5, bm, sb * 0.5)
        if abs(t) < 1e-3:
            continue
        got = O.shapes_collide(am, sa, ta, bm, sb, tb)
        assert got == (t > 0), (ta, tb, t, am, sa, bm, sb)
        assert O.shapes_collide(bm, sb, tb, am, sa, ta) == got   # the test is symmetric off the boundary
        robust[t > 0] += 1
    assert robust[True] > 25 and robust[False] > 50, robust


def test_quad_keeps_its_depth_on_its_own_face_axis():
    """A Quad's z half extent drops out of its projections onto the other
    shape's axes but stays on its own face axis (Collision.cuh:263-490:
    rb = b[2] there): a quad of depth 1 in front of a box's oblique corner
    counts as touching up to the box's projected radius plus 0.5; a planar
    one only up to the radius. The quad is wide enough that no other axis
    separates."""
    rng = np.random.default_rng(7)
    rot = rotation(rng)
    box = model(rot, [0.0, 0.0, 0.0])
    half = np.array([0.5, 0.7, 0.9])
    radius = float(np.sum(half * np.abs(rot[2, :])))     # projection onto z
    sb = np.array([2 * half[0], 2 * half[1], 2 * half[2]], dtype=np.float32)
    for depth, d, expect in [(0.0, radius + 0.25, False), (1.0, radius + 0.25, True),
                             (1.0, radius + 0.45, True), (1.0, radius + 0.55, False),
                             (0.0, radius - 0.05, True)]:
        quad = model(np.eye(3), [0.0, 0.0, d])
        sq = np.array([40.0, 40.0, depth], dtype=np.float32)
        assert O.shapes_collide(box, sb, B, quad, sq, Q) == expect, (depth, d)
        assert O.shapes_collide(quad, sq, Q, box, sb, B) == expect, (depth, d)


def test_touching_boxes_collide():
    """t > ra + rb separates: boxes whose faces touch exactly collide."""
    one = np.ones(3, dtype=np.float32)
    assert O.shapes_collide(model(np.eye(3), [0, 0, 0]), one, B, model(np.eye(3), [1.0, 0, 0]), one, B)
    assert not O.shapes_collide(model(np.eye(3), [0, 0, 0]), one, B, model(np.eye(3), [1.0001, 0, 0]), one, B)


# ---- phases ----

def yard_program(yard_phase=1, tag=1, late_generate_phase=None, lamps=False, lamp_tag=2):
    """Lot = Replicate { Building (phase 0) | Yard (phase yard_phase) };
    Building = Scale(-0.5, -1, -0.5) Collider(1) Generate(0);
    Yard = Repeat(X, 1) Repeat(Z, 1) Scale(-0.5, -0.3, -0.5)
           IfCollides(tag) { Generate(2) | Free };
    Free = Collider(2) Generate(1)      (lamps: and a phase-2 grid
    Lamps = Repeat(X, 2) Repeat(Z, 2) Scale(-0.6, -0.3, -0.6)
            IfCollides(lamp_tag) { Generate(3) | Generate(4) }).
    late_generate_phase: the free props' Generate is reached through a
    successor of that phase."""
    P1 = yard_phase
    (LOT, BLD, BSTORE, GB, YX, YZ, PROP, TEST, G2, FREE, G1, LX, LZ, LAMP, LTEST, G3, G4) = range(17)
    gphase = P1 if late_generate_phase is None else late_generate_phase
    lot_succ = [(BLD, 0), (YX, P1)] + ([(LX, 2)] if lamps else [])
    entries = [
        Entry(_cop(abi.REPLICATE, B), [], lot_succ),
        Entry(_cop(abi.SCALE, B), [Scalar(-0.5), Scalar(-1.0), Scalar(-0.5)], [(BSTORE, 0)]),
        Entry(_cop(abi.COLLIDER, B), [Scalar(1)], [(GB, 0)]),
        Entry(_cop(abi.GENERATE, B, 1), [Scalar(0)], []),
        Entry(_cop(abi.REPEAT, B), [Scalar(abi.X), Scalar(1.0), Scalar(abi.ADJUST_TO_FILL)], [(YZ, P1)]),
        Entry(_cop(abi.REPEAT, B), [Scalar(abi.Z), Scalar(1.0), Scalar(abi.ADJUST_TO_FILL)], [(PROP, P1)]),
        Entry(_cop(abi.SCALE, B), [Scalar(-0.5), Scalar(-0.3), Scalar(-0.5)], [(TEST, P1)]),
        Entry(_cop(abi.IFCOLLIDES, B), [Scalar(tag)], [(G2, P1), (FREE, P1)]),
        Entry(_cop(abi.GENERATE, B, 1), [Scalar(2)], []),
        Entry(_cop(abi.COLLIDER, B), [Scalar(2)], [(G1, gphase)]),
        Entry(_cop(abi.GENERATE, B, 1), [Scalar(1)], []),
        Entry(_cop(abi.REPEAT, B), [Scalar(abi.X), Scalar(2.0), Scalar(abi.ADJUST_TO_FILL)], [(LZ, 2)]),
        Entry(_cop(abi.REPEAT, B), [Scalar(abi.Z), Scalar(2.0), Scalar(abi.ADJUST_TO_FILL)], [(LAMP, 2)]),
        Entry(_cop(abi.SCALE, B), [Scalar(-0.6), Scalar(-0.3), Scalar(-0.6)], [(LTEST, 2)]),
        Entry(_cop(abi.IFCOLLIDES, B), [Scalar(lamp_tag)], [(G3, 2), (G4, 2)]),
        Entry(_cop(abi.GENERATE, B, 1), [Scalar(3)], []),
        Entry(_cop(abi.GENERATE, B, 1), [Scalar(4)], []),
    ]
    return Program(CONTEXT_PROCEDURES, DispatchTable(entries), "yard")


def lot_4x4():
    a = axioms.empty(1)
    a["size"] = (4.0, 3.0, 4.0)
    a["entry_index"] = 0
    a["shape_type"] = B
    return a


def counts(r, n=5):
    return tuple(len(r.instances(t)) for t in range(n))


def test_props_under_the_building_are_blocked():
    # building 2 x 3 x 2 at the centre of a 4 x 4 lot; 16 props of 0.5 x 0.5
    # at cell centres +-0.5, +-1.5: the four inner ones overlap it
    r = O.derive(yard_program(), lot_4x4())
    assert counts(r) == (1, 12, 4, 0, 0)
    assert r.colliders == 1 + 12
    for t, phase in ((0, 0), (1, 1), (2, 1)):
        assert np.all(r.instances(t)["meta"] >> 28 == phase)


def test_colliders_of_the_same_phase_are_invisible():
    # the yard in phase 0 asks while phase 0 runs: nothing stored before it
    r = O.derive(yard_program(yard_phase=0), lot_4x4())
    assert counts(r) == (1, 16, 0, 0, 0)
    assert np.all(r.instances(1)["meta"] >> 28 == 0)


def test_tags_keep_colliders_apart():
    r = O.derive(yard_program(tag=3), lot_4x4())
    assert counts(r) == (1, 16, 0, 0, 0)


def test_shapes_sent_to_an_earlier_phase_are_dropped():
    r = O.derive(yard_program(late_generate_phase=0), lot_4x4())
    assert counts(r) == (1, 0, 4, 0, 0)
    assert r.colliders == 13          # stored before the dropped successor


def test_three_phases_see_every_earlier_phase():
    # phase-2 lamps (1.2 x 1.2 at +-1) overlap the free phase-1 props at
    # (+-1.5, +-1.5): all four are blocked by tag 2, and by the phase-0
    # building (tag 1) too; no collider carries tag 5
    assert counts(O.derive(yard_program(lamps=True), lot_4x4())) == (1, 12, 4, 4, 0)
    assert counts(O.derive(yard_program(lamps=True, lamp_tag=1), lot_4x4())) == (1, 12, 4, 4, 0)
    assert counts(O.derive(yard_program(lamps=True, lamp_tag=5), lot_4x4())) == (1, 12, 4, 0, 4)


def test_phase_index_limits():
    prog = yard_program()
    prog.table.entries[0].successors[1].phase_index = 16
    with pytest.raises(O.OracleError) as e:
        O.derive(prog, lot_4x4())
    assert e.value.status == abi.ERR_BAD_PROGRAM
    r = O.derive(yard_program(yard_phase=-1), lot_4x4())      # unassigned: phase 0
    assert counts(r) == (1, 16, 0, 0, 0)
    assert np.all(r.instances(1)["meta"] >> 28 == 0)


def test_context_city_counts():
    """Every yard cell holds one prop, blocked or free; one building per lot;
    the walls variant adds four wall quads per building."""
    lots = axioms.city_lots(16)
    cells = sum(int(np.floor(a["size"][0])) * int(np.floor(a["size"][2])) for a in lots)
    r = O.derive(grammars.context_city(), lots)
    c = counts(r, 4)
    assert c[0] == 16 and c[1] + c[2] == cells and c[3] == 0
    assert 0 < c[2] < cells
    rw = O.derive(grammars.context_city(walls=True), lots)
    cw = counts(rw, 4)
    assert cw[0] == 0 and cw[3] == 64 and cw[1] + cw[2] == cells
    assert cw[2] >= c[2]              # walls only add blockers
    assert rw.colliders == 16 + 64


def test_collision_pairs_against_linear_program():
    """The pair program's decisions (terminal 1: collide) agree with the
    linear program on the instance geometry wherever the margin is clear."""
    rng = np.random.default_rng(3)
    kinds = [(ta, tb) for ta in (B, Q) for tb in (B, Q)] * 40
    ax, p = axioms.collision_pair_axioms(kinds, rng)
    r = O.derive(grammars.collision_pairs(), ax, params=p)
    a_inst = r.instances(0)
    assert len(a_inst) == len(kinds) and r.colliders == len(kinds)
    by_pair = {int(i["axiom"]) // 2: i for i in a_inst}
    hits = {int(i["axiom"]) // 2: True for i in r.instances(1)}
    hits.update({int(i["axiom"]) // 2: False for i in r.instances(2)})
    assert len(hits) == len(kinds)
    checked = 0
    for b in list(r.instances(1)) + list(r.instances(2)):
        pair = int(b["axiom"]) // 2
        a = by_pair[pair]
        # instance models carry the size: columns scaled, half extents 0.5
        t = lp_margin(a["model"], [0.5] * 3, b["model"], [0.5] * 3)
        if abs(t) < 1e-3:
            continue
        assert hits[pair] == (t > 0), (pair, kinds[pair], t)
        checked += 1
    assert checked > 120


# ---- host-side validation of the library (no device needed) ----

def test_library_refuses_to_specialise_context_sensitive_programs():
    from operatorgraph_amd import engine
    for prog in (grammars.collision_pairs(), grammars.context_city(), yard_program(yard_phase=0)):
        st, log, _ = engine.specialize_check(prog, 512)
        assert st == abi.ERR_UNSUPPORTED and "Collider" in log, (prog.name, st, log)


def test_library_validates_phase_indices():
    from operatorgraph_amd import engine
    prog = yard_program()
    prog.table.entries[0].successors[1].phase_index = 16
    st, log, _ = engine.specialize_check(prog, 512)
    assert st == abi.ERR_BAD_PROGRAM and "phase" in log, (st, log)
