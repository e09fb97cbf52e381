"""Seeded random multi-phase programs (tests/context_fuzz.py): Collider and
IfCollides at random places, successors that keep, inherit, raise, drop or
leave unassigned their phase. On the CPU, the generator's coverage and two
rewrites whose outcome the semantics of MultiPhaseEvaluator
(GPUEvaluator.cuh:395-456) and IfCollides (Operators.cuh, IfCollides) fix;
on the GPU, both techniques against the oracle: bit-exact instance
multisets and equal collider counts."""
import sys

import numpy as np
import pytest

from operatorgraph_amd import abi
from operatorgraph_amd.program import DispatchTable, Entry, Program, Scalar
from tests import context_fuzz as F
from tests.helpers import assert_same_multiset

SEEDS = range(16)


def _replicate_of(e, k):
    shape = F.PROCEDURES.procedures[e.operator_code].shape_type
    return Entry(F._op(abi.REPLICATE, shape), [], [e.successors[k]])


def _rewrite(program, fn):
    entries = [fn(e) for e in program.table.entries]
    return Program(F.PROCEDURES, DispatchTable(entries), program.name + "_rw")


def _no_path(inst):
    inst = inst.copy()
    inst["path"] = 0
    return inst


def test_generator_is_deterministic():
    a, b = F.random_program(5), F.random_program(5)
    assert len(a.table.entries) == len(b.table.entries)
    for x, y in zip(a.table.entries, b.table.entries):
        assert x.operator_code == y.operator_code
        assert [p.values for p in x.parameters] == [p.values for p in y.parameters]
        assert [(s.entry_index, s.phase_index) for s in x.successors] == \
               [(s.entry_index, s.phase_index) for s in y.successors]
    assert np.array_equal(F.axioms(64, 5), F.axioms(64, 5))


def test_programs_cover_the_phase_machinery():
    from oracle import oracle as O
    late = stored = mixed = taken = 0
    for seed in range(40):
        p = F.random_program(seed)
        r = O.derive(p, F.axioms(256, seed))
        phases = np.concatenate([r.instances(t)["meta"] >> 28 for t in range(4)])
        late += bool(np.any(phases >= 2))
        stored += r.colliders > 32              # enough for the device's BVH walk
        bo = F.branch_outcomes(p, r)
        mixed += sum(yes > 0 and no > 0 for _, yes, no in bo)
        taken += sum(yes > 0 for _, yes, _ in bo)
    assert late >= 15 and stored >= 25 and mixed >= 4 and taken >= 20, (late, stored, mixed, taken)


@pytest.mark.parametrize("seed", SEEDS)
def test_tests_without_colliders_take_the_second_successor(seed):
    """Every IfCollides asked for a tag no Collider stores derives what a
    Replicate of its second successor derives (paths aside: they hash the
    successor's ordinal), and so does the program with no Collider at all."""
    from oracle import oracle as O
    p = F.random_program(seed)
    ax = F.axioms(128, seed)
    ifc = {F._op(abi.IFCOLLIDES, s) for s in (F.B, F.Q)}
    col = {F._op(abi.COLLIDER, s) for s in (F.B, F.Q)}

    def untagged(e):
        if e.operator_code in ifc:
            return Entry(e.operator_code, [Scalar(7.0)], [(s.entry_index, s.phase_index) for s in e.successors])
        return e

    def falses(e):
        return _replicate_of(e, 1) if e.operator_code in ifc else e

    def bare(e):
        if e.operator_code in col:
            return _replicate_of(e, 0)
        return falses(e)

    r1 = O.derive(_rewrite(p, untagged), ax)
    r2 = O.derive(_rewrite(p, falses), ax)
    r3 = O.derive(_rewrite(p, bare), ax)
    for t in range(4):
        assert_same_multiset(_no_path(r1.instances(t)), _no_path(r2.instances(t)), f"seed {seed} terminal {t}")
        assert_same_multiset(_no_path(r3.instances(t)), _no_path(r2.instances(t)), f"seed {seed} terminal {t}")
    assert r1.colliders == r2.colliders and r3.colliders == 0


@pytest.mark.parametrize("seed", SEEDS)
def test_dropped_successors_derive_nothing(seed):
    """A successor of a lower phase than its parent's is dropped: replacing
    it by a Discard changes nothing (paths included)."""
    from oracle import oracle as O
    p = F.random_program(seed)
    ax = F.axioms(128, seed)
    phase = {0: 0}
    for i, e in enumerate(p.table.entries):      # parents precede children in the generator's order
        for s in e.successors:
            ph = s.phase_index
            phase[s.entry_index] = phase[i] if ph == abi.PHASE_INHERIT else max(ph, 0)
    discards = {s: F._op(abi.DISCARD, s) for s in (F.B, F.Q)}
    entries = list(p.table.entries)
    n_dropped = 0
    for i, e in enumerate(p.table.entries):
        succ = []
        for s in e.successors:
            if s.phase_index != abi.PHASE_INHERIT and max(s.phase_index, 0) < phase[i]:
                shape = F.PROCEDURES.procedures[p.table.entries[s.entry_index].operator_code].shape_type
                entries.append(Entry(discards[shape], [], []))
                succ.append((len(entries) - 1, s.phase_index))
                n_dropped += 1
            else:
                succ.append((s.entry_index, s.phase_index))
        entries[i] = Entry(e.operator_code, [(q.type, q.values) for q in e.parameters], succ)
    q = Program(F.PROCEDURES, DispatchTable(entries), p.name + "_dropped")
    r, rq = O.derive(p, ax), O.derive(q, ax)
    for t in range(4):
        assert_same_multiset(r.instances(t), rq.instances(t), f"seed {seed} terminal {t}")
    assert r.colliders == rq.colliders


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
@pytest.mark.parametrize("technique", [abi.KERNELS, abi.MEGAKERNEL])
@pytest.mark.parametrize("seed", SEEDS)
def test_random_programs_on_the_gpu(seed, technique):
    from oracle import oracle as O
    from operatorgraph_amd import engine
    _note(f"context fuzz seed={seed} technique={technique}")
    p = F.random_program(seed)
    ax = F.axioms(2048, seed)
    ref = O.derive(p, ax)
    eng = engine.Engine(technique=technique)
    try:
        eng.load(p).set_axioms(ax)
        for t in range(4):
            eng.bind(t, max(1, len(ref.instances(t))))
        eng.execute()
        for t in range(4):
            assert_same_multiset(eng.instances(t), ref.instances(t), f"seed {seed} terminal {t}")
        assert eng.colliders()[0] == ref.colliders
    finally:
        eng.close()
