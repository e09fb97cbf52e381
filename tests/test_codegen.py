"""Partition code (pga_generate_partition): the operator graph of a grammar,
cut by a partition, written as C++ for the C++ front end, compiled with
g++ (tests/codegen_cases.py) and lowered. Every case must derive, through
the oracle, exactly the instances of the grammar's interpreter table
(entry numbers in `meta` aside), whatever the partition."""
import os

import numpy as np
import pytest

from operatorgraph_amd import axioms, compiler
from operatorgraph_amd.compiler import GrammarError, compile_grammar
from tests import codegen_cases as CC
from tests.grammar_fuzz import lots as fuzz_lots
from tests.helpers import assert_same_multiset


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(CC.SO):
        CC.build()
    return True


def case_axioms(g, kind):
    if kind == "menger":
        return axioms.menger_axiom(1.0, 0)
    if kind == "facade":
        return axioms.facade_lots(16)
    if kind == "building":
        ax = axioms.facade_lots(16)
        ax["size"][:, 1] += 1.5
        return ax
    seed = int(kind.split(":")[1])
    return fuzz_lots(g, 8, seed, quad=g.axioms[0][1] == 1)


def masked(inst):
    inst = inst.copy()
    inst["meta"] &= 0xFFFF0000
    return inst


@pytest.mark.parametrize("case", CC.cases(), ids=lambda c: c[0])
def test_partition_code_derives_the_interpreter_instances(built, case):
    from oracle import oracle as O
    name, text, shape, spec, kind = case
    g = compile_grammar(text, name)
    prog = CC.LoweredProgram(name)
    ax = case_axioms(g, kind)
    try:
        want = O.derive(g.program, ax)
    except O.OracleError as e:
        with pytest.raises(O.OracleError) as info:
            O.derive(prog, ax)
        assert info.value.status == e.status
        return
    got = O.derive(prog, ax)
    total = 0
    for t in range(3):
        a, b = got.instances(t), want.instances(t)
        total += len(b)
        assert_same_multiset(masked(a), masked(b), f"{name} terminal {t}")
    assert total > 0 or kind.startswith("fuzz")


def test_partition_shapes():
    """All cut: one procedure per operator vertex; only the cycle edges cut:
    one procedure for the whole Menger sponge, which calls itself."""
    code_all = compiler.generate_partition(CC.MENGER, None, "M")
    assert code_all.count(": public Proc<") == 9
    cyc = compiler.operator_graph_edges(CC.MENGER)
    assert len(cyc) == 17 and sum(cyc) == 5     # SubZ's three and ZDiscard's two edges back to A
    code_min = compiler.generate_partition(CC.MENGER, "".join("1" if c else "0" for c in cyc), "M")
    assert code_min.count(": public Proc<") == 1
    assert code_min.count("FSCall<0, 0, -1, 0>") == 20   # the 20 of 27 cells a step keeps
    assert compiler.operator_graph_edges(CC.FACADE) == [False] * 10   # 3 + 1 + 1 + 2 + 3, no cycle


def test_partition_errors():
    cyc = compiler.operator_graph_edges(CC.MENGER)
    with pytest.raises(GrammarError, match="closes a cycle and must be cut"):
        compiler.generate_partition(CC.MENGER, "0" * len(cyc), "M")
    with pytest.raises(GrammarError, match="names 3 edges; the graph has 17"):
        compiler.generate_partition(CC.MENGER, "101", "M")
    with pytest.raises(GrammarError, match="must be a C\\+\\+ identifier"):
        compiler.generate_partition(CC.MENGER, None, "9x")
