"""GPU parity of dynamic polygons and right prisms (SetAs*, ComponentSplit of
a prism, Extrude of a polygon, inverted caps; tests/polygon_cases.py): the
interpreting kernels through the C ABI against the oracle, bit-exact
instance multisets, eagerly, replayed from a captured graph and in the
kernel specialised to the program; the tape refuses them."""
import sys

import numpy as np
import pytest

from operatorgraph_amd import abi
from tests import polygon_cases as PC
from tests.helpers import assert_same_multiset

pytestmark = pytest.mark.gpu

TECHNIQUES = [abi.KERNELS, abi.MEGAKERNEL]
N_TERMS = 5


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def E():
    from operatorgraph_amd import engine
    return engine


def check(E, program, ax, technique=abi.MEGAKERNEL, **cfg):
    from oracle import oracle as O
    ref = O.derive(program, ax)
    eng = E.Engine(technique=technique, **cfg)
    try:
        eng.load(program).set_axioms(ax)
        for t in range(N_TERMS):
            eng.bind(t, max(1, len(ref.instances(t))))
        eng.execute()
        for t in range(N_TERMS):
            assert_same_multiset(eng.instances(t), ref.instances(t), f"{program.name} terminal {t}")
    finally:
        eng.close()
    return ref


@pytest.mark.parametrize("technique", TECHNIQUES)
@pytest.mark.parametrize("n", [3, 4, 5, 8, 16])
@pytest.mark.parametrize("convex", [True, False])
def test_regular_prisms(E, technique, n, convex):
    _note(f"regular prism n={n} convex={convex} technique={technique}")
    ref = check(E, PC.prism_program(PC.regular(n, phase=0.1 * n), convex=convex), PC.boxes(256, n), technique)
    assert len(ref.instances(2)) == 256 * n


@pytest.mark.parametrize("seed", range(8))
def test_random_convex_prisms(E, seed):
    rng = np.random.default_rng(100 + seed)
    verts = PC.random_convex(int(rng.integers(3, 17)), rng)
    _note(f"random convex prism seed={seed} sides={len(verts)}")
    check(E, PC.prism_program(verts), PC.boxes(512, 200 + seed))


@pytest.mark.parametrize("technique", TECHNIQUES)
@pytest.mark.parametrize("name", ["roof", "recap", "polygon_ops"])
def test_programs(E, technique, name):
    _note(f"polygon program {name} technique={technique}")
    prog = {"roof": PC.roof_program, "recap": lambda: PC.recap_program(PC.random_convex(7, np.random.default_rng(3))),
            "polygon_ops": lambda: PC.polygon_ops_program(PC.regular(6))}[name]()
    check(E, prog, PC.boxes(1024, 7))


def test_captured_graph_replays(E):
    _note("polygon program captured and replayed")
    from oracle import oracle as O
    prog = PC.recap_program(PC.regular(5))
    ax = PC.boxes(128, 11)
    ref = O.derive(prog, ax)
    eng = E.Engine()
    try:
        eng.load(prog).set_axioms(ax)
        for t in range(N_TERMS):
            eng.bind(t, max(1, len(ref.instances(t))))
        eng.execute()
        with eng.capture() as g:
            eng.execute_async()
        for _ in range(3):
            g.launch()
            eng.synchronize()
            for t in range(4):
                assert_same_multiset(eng.instances(t), ref.instances(t), f"replay terminal {t}")
    finally:
        eng.close()


@pytest.mark.parametrize("name", ["roof", "recap", "polygon_ops", "prism16"])
def test_specialised_kernel(name):
    from tests.test_spec_gpu import spec_check
    _note(f"polygon program {name}: specialised kernel")
    prog = {"roof": PC.roof_program, "recap": lambda: PC.recap_program(PC.random_convex(7, np.random.default_rng(3))),
            "polygon_ops": lambda: PC.polygon_ops_program(PC.regular(6)),
            "prism16": lambda: PC.prism_program(PC.regular(16, phase=0.2))}[name]()
    spec_check(prog, PC.boxes(512, 9), N_TERMS)


def test_tape_refused(E):
    _note("polygon programs: the tape refused")
    eng = E.Engine()
    try:
        eng.load(PC.roof_program()).set_axioms(PC.boxes(4, 0))
        for t in range(2):
            eng.bind(t, 64)
        eng.tape_reserve(1 << 12, 1 << 12)
        with pytest.raises(E.PGAError) as e:
            eng.forward_tape()
        assert e.value.status == abi.ERR_UNSUPPORTED
        eng.execute()
        assert int(eng.counts(2)[0]) == 8
    finally:
        eng.close()


@pytest.mark.parametrize("technique", TECHNIQUES)
@pytest.mark.parametrize("seed", range(16))
def test_random_polygon_programs(E, seed, technique):
    from tests import polygon_fuzz as F
    _note(f"random polygon program seed={seed} technique={technique}")
    check(E, F.random_program(seed), PC.boxes(512, 300 + seed), technique)


@pytest.mark.parametrize("seed", range(8))
def test_random_polygon_programs_specialised(seed):
    from tests import polygon_fuzz as F
    from tests.test_spec_gpu import spec_check
    _note(f"random polygon program seed={seed}: specialised kernel")
    spec_check(F.random_program(seed), PC.boxes(512, 300 + seed), N_TERMS)


@pytest.mark.parametrize("technique", TECHNIQUES)
def test_gabled_houses(E, technique):
    from operatorgraph_amd import axioms, grammars
    _note(f"gabled houses, 1024 lots, technique={technique}")
    ref = check(E, grammars.gabled_houses(), axioms.city_lots(1024), technique)
    assert len(ref.instances(1)) == 2048


def test_gabled_houses_specialised():
    from operatorgraph_amd import axioms, grammars
    from tests.test_spec_gpu import spec_check
    _note("gabled houses, 1024 lots, specialised kernel")
    spec_check(grammars.gabled_houses(), axioms.city_lots(1024), 3)


def _two_phase_roofs():
    """Phase 0 stores each lot as a collider (tag 1) and emits it; phase 1
    tests the lot against the colliders and, where it collides (its own
    box always does), turns it into a triangular prism and splits it: the
    invert bit of the bottom cap beside the phase bits of a multi-phase
    record."""
    from operatorgraph_amd.program import Procedure, ProcedureList
    procs = ProcedureList([Procedure(abi.REPLICATE, PC.B), Procedure(abi.COLLIDER, PC.B),
                           Procedure(abi.GENERATE, PC.B, 1), Procedure(abi.IFCOLLIDES, PC.B),
                           Procedure(abi.SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, PC.B), Procedure(abi.DISCARD, PC.B),
                           Procedure(abi.COMPSPLIT, PC.CR), Procedure(abi.GENERATE, PC.CP, 1),
                           Procedure(abi.GENERATE, PC.Q, 1)])
    from operatorgraph_amd.program import DispatchTable, Entry, Program, Scalar, Vec2
    tri = [Vec2(0.0, 0.5), Vec2(-0.5, -0.5), Vec2(0.5, -0.5)]
    return Program(procs, DispatchTable([
        Entry(0, [], [(1, 0), (3, 1)]),                 # 0 Replicate: phase 0 | phase 1
        Entry(1, [Scalar(1.0)], [(2, 0)]),              # 1 Collider(1)
        Entry(2, [Scalar(0.0)], []),                    # 2 Generate 0
        Entry(3, [Scalar(1.0)], [(4, 1), (5, 1)]),      # 3 IfCollides(1)
        Entry(4, tri, [(6, abi.PHASE_INHERIT)]),        # 4 SetAs prism
        Entry(5, [], []),                               # 5 Discard
        Entry(6, [], [(7, abi.PHASE_INHERIT), (7, abi.PHASE_INHERIT), (8, abi.PHASE_INHERIT)]),  # 6 split
        Entry(7, [Scalar(1.0)], []),                    # 7 caps
        Entry(8, [Scalar(2.0)], []),                    # 8 sides
    ]), "two_phase_roofs")


@pytest.mark.parametrize("technique", TECHNIQUES)
def test_polygons_in_a_later_phase(E, technique):
    _note(f"two-phase roofs technique={technique}")
    ref = check(E, _two_phase_roofs(), PC.boxes(256, 21), technique)
    assert [len(ref.instances(t)) for t in range(3)] == [256, 512, 768]
    assert np.all(ref.instances(1)["meta"] >> 28 == 1)
