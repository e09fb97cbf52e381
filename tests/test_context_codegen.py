"""Context-sensitive grammars in the grammar language, through the operator
graph's phases (Graph::computePhases, src/compiler/Graph.cpp:149-272) and
partition code (tests/codegen_cases.py context_cases): the lowered programs
derive, in the oracle, the closed-form counts of the yard and the lamps and
exactly the instances of the hand-built tables of the same grammars
(grammars.context_city, tests/test_context.yard_program; entry numbers in
`meta` aside); on the GPU, exactly the oracle's instances."""
import os
import sys

import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, grammars
from operatorgraph_amd.compiler import GrammarError, compile_grammar, generate_partition, operator_graph_edges
from tests import codegen_cases as CC
from tests.helpers import assert_same_multiset
from tests.test_context import lot_4x4, yard_program


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(CC.SO):
        CC.build()
    return True


def masked(inst):
    inst = inst.copy()
    inst["meta"] &= 0xFFFF0000
    return inst


def counts(r, n):
    return tuple(len(r.instances(t)) for t in range(n))


def names(kind):
    return [n for n, _, _ in CC.context_cases() if n.startswith(f"ctx_{kind}_")]


def test_phase_edges_must_be_cut():
    flags = operator_graph_edges(CC.YARD)
    assert sum(flags) == 1                      # the edge into IfCollides(1); the grammar has no cycle
    keep = "".join("0" for _ in flags)
    with pytest.raises(GrammarError, match="IfCollides"):
        generate_partition(CC.YARD, keep, "Y")
    code = generate_partition(CC.YARD, None, "Y")
    assert "NumPhases = 2" in code and "IfCollides<" in code and "Collider<" in code
    assert "NumPhases = 3" in generate_partition(CC.LAMPS, None, "L")
    # the walk reaches the lamps' IfCollides(2) before the yard's Collider(2)
    # when the yard comes last, and retries it only once
    with pytest.raises(GrammarError, match="before its colliders"):
        generate_partition(CC.LAMPS.replace("Building | Lamps | Yard", "Building | Yard | Lamps"), None, "L")


def test_interpreter_tables_refuse_collision_operators():
    # the reference's interpreter procedure list has no such procedures
    with pytest.raises(GrammarError, match="IfCollides|Collider"):
        compile_grammar(CC.YARD, "yard")


@pytest.mark.parametrize("name", names("yard"))
def test_yard_partition_code(built, name):
    from oracle import oracle as O
    prog = CC.LoweredProgram(name)
    r = O.derive(prog, lot_4x4())
    assert counts(r, 3) == (1, 12, 4) and r.colliders == 1
    # the hand-built table stores the free props as colliders on the way
    # (one more operator in their path): compare their geometry and axioms
    ref = O.derive(yard_program(), lot_4x4())
    for t in range(3):
        a, b = r.instances(t).copy(), ref.instances(t).copy()
        a["meta"] = b["meta"] = 0
        a["path"] = b["path"] = 0
        assert_same_multiset(a, b, f"{name} terminal {t}")


@pytest.mark.parametrize("name", names("lamps"))
def test_lamps_partition_code(built, name):
    from oracle import oracle as O
    r = O.derive(CC.LoweredProgram(name), lot_4x4())
    # the phase-2 lamps (1.2 x 1.2 at +-1) overlap the free phase-1 props
    assert counts(r, 5) == (1, 12, 4, 4, 0) and r.colliders == 13
    assert np.all(r.instances(3)["meta"] >> 28 == 2)


@pytest.mark.parametrize("name", names("city"))
def test_city_partition_code(built, name):
    from oracle import oracle as O
    lots = axioms.city_lots(16)
    r = O.derive(CC.LoweredProgram(name), lots)
    ref = O.derive(grammars.context_city(walls=True), lots)
    for t in range(4):
        assert_same_multiset(masked(r.instances(t)), masked(ref.instances(t)), f"{name} terminal {t}")
    assert r.colliders == ref.colliders == 16 * 5


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n, _, _ in CC.context_cases()])
def test_context_partition_code_on_the_gpu(built, name):
    from oracle import oracle as O
    from operatorgraph_amd import engine
    _note(f"context partition code {name}")
    prog = CC.LoweredProgram(name)
    if "_city_" in name:
        ax = axioms.city_lots(64)
    else:
        rng = np.random.default_rng(9)
        ax = axioms.empty(16)
        ax["size"][:, 0] = rng.uniform(3.0, 9.0, 16).astype(np.float32)
        ax["size"][:, 1] = 3.0
        ax["size"][:, 2] = rng.uniform(3.0, 9.0, 16).astype(np.float32)
        ax["model"][:, 3] = (np.arange(16) % 4 * 12.0).astype(np.float32)
        ax["model"][:, 11] = (np.arange(16) // 4 * 12.0).astype(np.float32)
        ax["entry_index"] = 0
        ax["shape_type"] = abi.BOX
    ref = O.derive(prog, ax)
    eng = engine.Engine()
    try:
        eng.load(prog).set_axioms(ax)
        for t in range(5):
            eng.bind(t, max(1, len(ref.instances(t))))
        eng.execute()
        for t in range(5):
            assert_same_multiset(eng.instances(t), ref.instances(t), f"{name} terminal {t}")
        assert eng.colliders()[0] == ref.colliders
    finally:
        eng.close()
