"""GPU parity of context-sensitive, multi-phase derivation (Collider,
IfCollides; MultiPhaseEvaluator, GPUEvaluator.cuh:395-456): the interpreting
kernels through the C ABI against the oracle, bit-exact instance multisets
and equal collider counts, and the refusals of the paths that do not run
such programs."""
import sys

import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, grammars
from tests.helpers import assert_same_multiset
from tests.test_context import lot_4x4, yard_program

pytestmark = pytest.mark.gpu

B, Q = abi.BOX, abi.QUAD
TECHNIQUES = [abi.KERNELS, abi.MEGAKERNEL]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def E():
    from operatorgraph_amd import engine
    return engine


def run(E, program, ax, n_terms, technique=abi.MEGAKERNEL, params=None, cap=1 << 20, **cfg):
    eng = E.Engine(technique=technique, **cfg)
    try:
        eng.load(program).set_axioms(ax)
        if params is not None:
            eng.set_axiom_params(params)
        for t in range(n_terms):
            eng.bind(t, cap)
        eng.execute()
        return [eng.instances(t) for t in range(n_terms)], eng.colliders()
    finally:
        eng.close()


def check(E, program, ax, n_terms, technique=abi.MEGAKERNEL, params=None, **cfg):
    from oracle import oracle as O
    ref = O.derive(program, ax, params=params)
    got, (ncoll, phases) = run(E, program, ax, n_terms, technique, params=params, **cfg)
    for t in range(n_terms):
        assert_same_multiset(got[t], ref.instances(t), f"{program.name} terminal {t}")
    assert ncoll == ref.colliders
    return got, ref, phases


@pytest.mark.parametrize("technique", TECHNIQUES)
@pytest.mark.parametrize("variant", ["two_phases", "same_phase", "other_tag", "dropped", "three_phases",
                                     "three_phases_tag1", "three_phases_tag5", "unassigned"])
def test_yard_programs(E, technique, variant):
    _note(f"yard {variant} technique={technique}")
    kw = {"two_phases": {}, "same_phase": {"yard_phase": 0}, "other_tag": {"tag": 3},
          "dropped": {"late_generate_phase": 0}, "three_phases": {"lamps": True},
          "three_phases_tag1": {"lamps": True, "lamp_tag": 1}, "three_phases_tag5": {"lamps": True, "lamp_tag": 5},
          "unassigned": {"yard_phase": -1}}[variant]
    got, ref, phases = check(E, yard_program(**kw), lot_4x4(), 5, technique)
    assert phases == 3          # the lamp entries name phase 2 even where unreachable
    if variant == "two_phases":
        assert [len(g) for g in got] == [1, 12, 4, 0, 0]


@pytest.mark.parametrize("quad_depth", [0.0, 0.7])
def test_collision_pairs(E, quad_depth):
    """Every pairing of Box and Quad, random sizes, rotations and offsets:
    the device's decision for each pair is the oracle's."""
    _note(f"collision pairs quad_depth={quad_depth}")
    rng = np.random.default_rng(11 + int(quad_depth * 10))
    kinds = [(ta, tb) for ta in (B, Q) for tb in (B, Q)] * 256
    ax, p = axioms.collision_pair_axioms(kinds, rng, quad_depth=quad_depth)
    got, ref, phases = check(E, grammars.collision_pairs(), ax, 3, params=p)
    assert phases == 2
    assert len(got[0]) == len(kinds) and len(got[1]) + len(got[2]) == len(kinds)
    assert 0.2 * len(kinds) < len(got[1]) < 0.8 * len(kinds)


@pytest.mark.parametrize("technique", TECHNIQUES)
@pytest.mark.parametrize("walls", [False, True])
def test_context_city(E, technique, walls):
    _note(f"context city walls={walls} technique={technique}")
    lots = axioms.city_lots(64)
    got, ref, phases = check(E, grammars.context_city(walls=walls), lots, 4, technique)
    cells = sum(int(np.floor(a["size"][0])) * int(np.floor(a["size"][2])) for a in lots)
    assert len(got[1]) + len(got[2]) == cells and len(got[2]) > 0


@pytest.mark.parametrize("block_size", [256, 768, 1024])
def test_context_city_block_sizes(E, block_size):
    _note(f"context city block_size={block_size}")
    check(E, grammars.context_city(walls=True), axioms.city_lots(256), 4, block_size=block_size)


def test_repeated_and_appended_executes(E):
    """Colliders start empty at every execute; an appending execute adds the
    same instances again."""
    _note("context city execute twice, then append")
    from oracle import oracle as O
    lots = axioms.city_lots(32)
    prog = grammars.context_city()
    ref = O.derive(prog, lots)
    eng = E.Engine()
    try:
        eng.load(prog).set_axioms(lots)
        for t in range(4):
            eng.bind(t, 1 << 18)
        for _ in range(2):
            eng.execute()
            for t in range(3):
                assert_same_multiset(eng.instances(t), ref.instances(t), f"terminal {t}")
            assert eng.colliders() == (ref.colliders, 2)
        eng.execute(append=True)
        for t in range(3):
            assert_same_multiset(eng.instances(t), np.concatenate([ref.instances(t)] * 2), f"appended {t}")
        assert eng.colliders() == (ref.colliders, 2)
    finally:
        eng.close()


def test_collider_capacity_overflow(E):
    _note("collider capacity 4 with 13 colliders")
    eng = E.Engine()
    try:
        eng.load(yard_program()).set_axioms(lot_4x4())
        for t in range(5):
            eng.bind(t, 1024)
        eng.set_collider_capacity(4)
        with pytest.raises(E.PGAError) as e:
            eng.execute()
        assert e.value.status == abi.ERR_OVERFLOW_QUEUE
        eng.set_collider_capacity(0)
        eng.execute()
        assert [int(c) for c in eng.counts(5)] == [1, 12, 4, 0, 0]
    finally:
        eng.close()


def test_refusals(E):
    _note("context-sensitive programs: specialise, tape and capture refused; phase 16 rejected")
    eng = E.Engine()
    try:
        eng.load(grammars.context_city()).set_axioms(axioms.city_lots(4))
        for t in range(4):
            eng.bind(t, 4096)
        with pytest.raises(E.PGAError) as e:
            eng.specialize()
        assert e.value.status == abi.ERR_UNSUPPORTED
        eng.tape_reserve(1 << 16, 1 << 16)
        with pytest.raises(E.PGAError) as e:
            eng.forward_tape()
        assert e.value.status == abi.ERR_UNSUPPORTED
        with pytest.raises(E.PGAError) as e:
            with eng.capture():
                eng.execute_async()
        assert e.value.status == abi.ERR_UNSUPPORTED
        eng.execute()                       # still runs eagerly
        assert eng.colliders()[0] == 4
        bad = yard_program()
        bad.table.entries[0].successors[1].phase_index = 16
        with pytest.raises(E.PGAError) as e:
            eng.load(bad)
        assert e.value.status == abi.ERR_BAD_PROGRAM
    finally:
        eng.close()


def test_depth_limit_of_multi_phase_programs(E):
    """Depth shares the instance meta with the phase: a multi-phase program
    stops at depth 4096 with PGA_ERR_MAX_DEPTH, as the oracle does."""
    _note("multi-phase depth limit")
    from oracle import oracle as O
    from operatorgraph_amd.grammars import CONTEXT_PROCEDURES, _cop
    from operatorgraph_amd.program import DispatchTable, Entry, Program, Scalar
    # Lot = Replicate { Loop (phase 1) }; Loop = Translate(0, 0, 0) Loop
    prog = Program(CONTEXT_PROCEDURES, DispatchTable([
        Entry(_cop(abi.REPLICATE, B), [], [(1, 1)]),
        Entry(_cop(abi.TRANSLATE, B), [Scalar(0.0), Scalar(0.0), Scalar(0.0)], [(1, 1)]),
    ]), "loop")
    with pytest.raises(O.OracleError) as e:
        O.derive(prog, lot_4x4())
    assert e.value.status == abi.ERR_MAX_DEPTH
    eng = E.Engine()
    try:
        eng.load(prog).set_axioms(lot_4x4())
        eng.bind(0, 16)
        with pytest.raises(E.PGAError) as e:
            eng.execute()
        assert e.value.status == abi.ERR_MAX_DEPTH
    finally:
        eng.close()
