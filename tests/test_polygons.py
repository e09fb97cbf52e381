"""Dynamic polygons and right prisms in the oracle (SetAs.cuh:112-150,
ComponentSplit.cuh:23-75, DynamicPolygon.cuh:68-100, Extrude.cuh:25-71),
pinned by what the geometry fixes: a prism's sides are its scaled edges,
centred on their midpoints, facing outward, as tall as the prism; its caps
lie at +-size.y / 2 facing +-y, the bottom one inverted (the first two model
columns of its instance negated); an inverted polygon's extrusion splits
into the same edges in reverse order; the balcony roof's counts. Then the
library's validation of polygon programs (no GPU needed)."""
import numpy as np
import pytest

from operatorgraph_amd import abi
from operatorgraph_amd.program import DispatchTable, Entry, Program, Scalar, Vec2
from tests import polygon_cases as PC


def cols(inst):
    m = inst["model"].reshape(-1, 3, 4).astype(np.float64)
    return m[:, :, 0], m[:, :, 1], m[:, :, 2], m[:, :, 3]


def edges(vertices, sx, sz):
    """Scaled edges (p0, p1) of side j = 0..n-1 in the prism's (x, z) plane:
    vertex j-1 to vertex j, y mirrored onto -z."""
    v = np.array(vertices, dtype=np.float64) * [sx, sz]
    v[:, 1] = -v[:, 1]
    return np.roll(v, 1, axis=0), v


@pytest.mark.parametrize("n", [3, 4, 5, 7, 16])
def test_prism_sides_and_caps(n):
    from oracle import oracle as O
    verts = PC.regular(n, phase=0.3)
    ax = PC.boxes(3, n, rotate=False)
    r = O.derive(PC.prism_program(verts), ax)
    top, bottom, sides = r.instances(0), r.instances(1), r.instances(2)
    assert (len(top), len(bottom), len(sides)) == (3, 3, 3 * n)
    for a in range(3):
        sx, sy, sz = (float(v) for v in ax["size"][a])
        origin = np.array([ax["model"][a][3], ax["model"][a][7], ax["model"][a][11]], dtype=np.float64)
        s = sides[sides["axiom"] == a]
        c0, c1, c2, c3 = cols(s)
        p0, p1 = edges(verts, sx, sz)
        width = np.linalg.norm(p1 - p0, axis=1)
        mid = (p0 + p1) / 2
        got = c3 - origin
        # each side: width |edge| along its x axis, the prism's height along
        # +y, centred on the edge's midpoint, unit normal facing away from
        # the centre (the polygon contains the origin)
        order = np.lexsort((np.round(got[:, 2], 4), np.round(got[:, 0], 4)))
        want = np.lexsort((np.round(mid[:, 1], 4), np.round(mid[:, 0], 4)))
        np.testing.assert_allclose(got[order][:, [0, 2]], mid[want], atol=1e-5)
        np.testing.assert_allclose(got[:, 1], 0.0, atol=1e-6)
        np.testing.assert_allclose(np.sort(np.linalg.norm(c0, axis=1)), np.sort(width), rtol=1e-5)
        np.testing.assert_allclose(c1, np.tile([0.0, sy, 0.0], (n, 1)), rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(np.linalg.norm(c2, axis=1), 1.0, rtol=1e-6)
        assert np.all(np.einsum("ij,ij->i", c2, got) > 0)
        np.testing.assert_allclose(np.einsum("ij,ij->i", c0, c2), 0.0, atol=1e-5)
        # caps: top at +sy/2 facing +y with x along +x, y along -z; the
        # bottom (inverted) at -sy/2 facing -y, its first two columns negated
        t, b = top[top["axiom"] == a], bottom[bottom["axiom"] == a]
        t0, t1, t2, t3 = cols(t)
        b0, b1, b2, b3 = cols(b)
        np.testing.assert_allclose(t3[0] - origin, [0.0, sy / 2, 0.0], atol=1e-6)
        np.testing.assert_allclose(b3[0] - origin, [0.0, -sy / 2, 0.0], atol=1e-6)
        np.testing.assert_allclose(t0[0], [sx, 0, 0], atol=1e-6)
        np.testing.assert_allclose(t1[0], [0, 0, -sz], atol=1e-6)
        np.testing.assert_allclose(t2[0], [0, 1, 0], atol=1e-6)
        np.testing.assert_array_equal(b0[0], -t0[0])
        np.testing.assert_array_equal(b1[0], t1[0])
        np.testing.assert_array_equal(b2[0], -t2[0])


def test_rotated_axioms_move_every_face_rigidly():
    """The faces of a yawed axiom are those of the unyawed one turned by the
    same rotation."""
    from oracle import oracle as O
    verts = PC.random_convex(9, np.random.default_rng(4))
    a0 = PC.boxes(1, 3, rotate=False)
    a1 = a0.copy()
    c, s = np.float32(np.cos(0.7)), np.float32(np.sin(0.7))
    a1["model"][:, 0], a1["model"][:, 2], a1["model"][:, 8], a1["model"][:, 10] = c, s, -s, c
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], dtype=np.float64)
    prog = PC.prism_program(verts)
    r0, r1 = O.derive(prog, a0), O.derive(prog, a1)
    o = np.array([a0["model"][0][3], a0["model"][0][7], a0["model"][0][11]], dtype=np.float64)
    for t in range(3):
        m0 = r0.instances(t)["model"].reshape(-1, 3, 4).astype(np.float64)
        m1 = r1.instances(t)["model"].reshape(-1, 3, 4).astype(np.float64)
        m0[:, :, 3] -= o
        m1[:, :, 3] -= o
        np.testing.assert_allclose(np.einsum("ij,njk->nik", R, m0), m1, atol=1e-5)


@pytest.mark.parametrize("seed", range(4))
def test_inverted_cap_splits_into_reversed_edges(seed):
    """The bottom cap's extrusion is an inverted prism: its sides run over
    the same edges (the same widths), in reverse order."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    verts = PC.random_convex(int(rng.integers(3, 17)), rng)
    n = len(verts)
    ax = PC.boxes(1, 10 + seed, rotate=False)
    r = O.derive(PC.recap_program(verts), ax)
    assert [len(r.instances(t)) for t in range(4)] == [n, 2, 2, 2 * n]
    sides = r.instances(3)
    # the two second-level prisms: one from the top cap (its sides above
    # the box), one from the bottom cap (below)
    c0, _, _, c3 = cols(sides)
    y0 = ax["model"][0][7]
    up, down = c3[:, 1] > y0, c3[:, 1] < y0
    assert up.sum() == n and down.sum() == n
    wu = np.linalg.norm(c0[up], axis=1)
    wd = np.linalg.norm(c0[down], axis=1)
    np.testing.assert_allclose(np.sort(wu), np.sort(wd), rtol=1e-5)
    # side j of each second-level split, by its path (path_mix of the
    # ordinals: the SetAs child, cap n (bottom) or n + 1 (top), the
    # extrusion, side j): the top cap's prism runs edge (j-1, j) as the
    # first split did, the inverted one edge (n-1-j, n-j), since getVertex
    # reverses the list, so its widths are the others' at (n - j) mod n
    from oracle.oracle import path_mix
    p1 = path_mix(0, 0)
    width = dict(zip(sides["path"].tolist(), np.linalg.norm(c0, axis=1)))
    top = [width[path_mix(path_mix(path_mix(p1, n + 1), 0), j)] for j in range(n)]
    bottom = [width[path_mix(path_mix(path_mix(p1, n), 0), j)] for j in range(n)]
    first = r.instances(0)
    w_first = dict(zip(first["path"].tolist(), np.linalg.norm(cols(first)[0], axis=1)))
    np.testing.assert_allclose(top, [w_first[path_mix(p1, j)] for j in range(n)], rtol=1e-6)
    np.testing.assert_allclose(bottom, [top[(n - j) % n] for j in range(n)], rtol=1e-6)


def test_roof_counts():
    """The balcony roof: two RoofSide caps and two Roof halves per lot. The
    prism's bottom side is discarded where its normal's y is exactly -1,
    that is where the edge direction size.x * (1 / |edge|) rounds to 1 in
    FP32; elsewhere the normal misses -1 by an ulp and that side becomes a
    Roof too."""
    from oracle import oracle as O
    ax = PC.boxes(8, 1, rotate=True)
    r = O.derive(PC.roof_program(), ax)
    assert len(r.instances(0)) == 16
    roofs = r.instances(1)
    _, _, c2, _ = cols(roofs)
    up = c2[:, 1] > 0.1                    # the two slanted halves of every lot
    assert up.sum() == 16
    assert np.all(c2[~up, 1] < -0.999999)
    sx = ax["size"][:, 0]
    x = np.sqrt(sx * sx)                   # float32 throughout
    inexact = sx * (np.float32(1.0) / x) != np.float32(1.0)
    assert (~up).sum() == int(np.sum(inexact)) and 0 < (~up).sum() < 8


def test_generic_operators_on_polygons():
    from oracle import oracle as O
    r = O.derive(PC.polygon_ops_program(PC.regular(6)), PC.boxes(4, 2))
    counts = [len(r.instances(t)) for t in range(5)]
    assert counts[0] > 0 and counts[2] > 0 and counts[4] == 4 * 5 and counts[1] == counts[3] == 0


def _two_lists():
    """Two SetAs with different vertex lists into one prism entry."""
    return Program(PC.PROCEDURES, DispatchTable([
        Entry(PC.op(abi.REPLICATE, PC.B), [], [(1, 0), (2, 0)]),
        PC.setas(PC.B, PC.regular(3), 3),
        PC.setas(PC.B, PC.regular(4), 3),
        Entry(PC.op(abi.COMPSPLIT, PC.CR), [], [(4, 0), (4, 0), (5, 0)]),
        PC.gen(PC.CP, 0),
        PC.gen(PC.Q, 1),
    ]), "two_lists")


def test_oracle_refusals():
    from oracle import oracle as O
    ax = PC.boxes(1, 0)
    for nv in (2, 17):
        prog = Program(PC.PROCEDURES, DispatchTable([
            Entry(PC.op(abi.SET_AS_DYNAMIC_CONVEX_POLYGON, PC.B), [Vec2(0.1 * i, 0.0) for i in range(nv)], [(1, 0)]),
            PC.gen(PC.CP, 0)]), "bad_setas")
        with pytest.raises((O.OracleError, RuntimeError)):
            O.derive(prog, ax)
    # the oracle derives each list where it lands (the library refuses this)
    r = O.derive(_two_lists(), ax)
    assert len(r.instances(1)) == 3 + 4


def test_library_validation():
    """The library's validation (pga_specialize_check validates as
    pga_initialize does, then compiles the specialised kernel): a
    vertex-list clash and too few or too many vertices are errors; the roof
    and the generic-operator program compile."""
    from operatorgraph_amd import engine as E
    st, log, _ = E.specialize_check(_two_lists())
    assert st == abi.ERR_UNSUPPORTED and "two different vertex lists" in log
    for prog in (PC.roof_program(), PC.polygon_ops_program(PC.regular(5)), PC.recap_program(PC.regular(4))):
        st, log, _ = E.specialize_check(prog)
        assert st == abi.OK, log
    for nv in (2, 17):
        prog = Program(PC.PROCEDURES, DispatchTable([
            Entry(PC.op(abi.SET_AS_DYNAMIC_CONVEX_POLYGON, PC.B), [Vec2(0.1 * i, 0.0) for i in range(nv)], [(1, 0)]),
            PC.gen(PC.CP, 0)]), "bad_setas")
        st, log, _ = E.specialize_check(prog)
        assert st == abi.ERR_BAD_PROGRAM, log


def test_random_polygon_programs_derive():
    """The seeded random polygon programs (tests/polygon_fuzz.py) derive in
    the oracle; more than half of them split prisms."""
    from oracle import oracle as O
    from tests import polygon_fuzz as F
    splits = 0
    for seed in range(24):
        p = F.random_program(seed)
        O.derive(p, PC.boxes(16, seed))
        splits += any(e.operator_code in (PC.op(abi.COMPSPLIT, PC.CR), PC.op(abi.COMPSPLIT, PC.R))
                      for e in p.table.entries)
    assert splits >= 12


def test_gabled_houses():
    """grammars.gabled_houses on city lots: 2 max(1, floor(w / 2)) +
    2 max(1, floor(d / 2)) wall tiles, two gables and three roof sides a
    lot; the gables stand at the lot's front and back faces, up to the top."""
    from oracle import oracle as O
    from operatorgraph_amd import axioms, grammars
    lots = axioms.city_lots(64)
    r = O.derive(grammars.gabled_houses(), lots)
    w, d = lots["size"][:, 0], lots["size"][:, 2]
    walls = int(np.sum(2 * np.maximum(1, np.floor(w / 2)) + 2 * np.maximum(1, np.floor(d / 2))))
    assert [len(r.instances(t)) for t in range(3)] == [walls, 2 * 64, 3 * 64]
    g = r.instances(1)
    _, _, c2, c3 = cols(g)
    a = g["axiom"]
    np.testing.assert_allclose(np.abs(c2[:, 2]), 1.0, atol=1e-6)     # gables face along z
    np.testing.assert_allclose(np.abs(c3[:, 2] - lots["model"][a, 11]), d[a] / 2, rtol=1e-5)


def test_refused_polygon_forms():
    """Triangle soup (generation function 0) and the collision operators of
    polygon shapes are refused on load (PGA_ERR_UNSUPPORTED), in the oracle
    too; so is a parallel ComponentSplit of a prism."""
    from oracle import oracle as O
    from operatorgraph_amd import engine as E
    from operatorgraph_amd.program import Procedure, ProcedureList
    procs = ProcedureList([Procedure(abi.SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, PC.B),
                           Procedure(abi.GENERATE, PC.CR, 0), Procedure(abi.COLLIDER, PC.CR),
                           Procedure(abi.GENERATE, PC.CR, 1), Procedure(abi.COMPSPLIT, PC.CR, 0, abi.PROC_PARALLEL),
                           Procedure(abi.GENERATE, PC.CP, 1), Procedure(abi.GENERATE, PC.Q, 1)])
    tri = [Vec2(0.0, 0.5), Vec2(-0.5, -0.5), Vec2(0.5, -0.5)]
    cases = {
        "generation function 0": [Entry(0, tri, [(1, 0)]), Entry(1, [Scalar(0.0)], [])],
        "collision": [Entry(0, tri, [(1, 0)]), Entry(2, [Scalar(1.0)], [(2, 0)]), Entry(3, [Scalar(0.0)], [])],
        "parallel": [Entry(0, tri, [(1, 0)]), Entry(4, [], [(2, 0), (2, 0), (3, 0)]), Entry(5, [Scalar(0.0)], []),
                     Entry(6, [Scalar(1.0)], [])],
    }
    for what, entries in cases.items():
        prog = Program(procs, DispatchTable(entries), what)
        st, log, _ = E.specialize_check(prog)
        assert st == abi.ERR_UNSUPPORTED, (what, log)
        with pytest.raises(O.OracleError) as e:
            O.derive(prog, PC.boxes(2, 0))
        assert e.value.status == abi.ERR_UNSUPPORTED, what
