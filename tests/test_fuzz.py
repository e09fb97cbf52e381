"""Seeded random grammars (tests/grammar_fuzz.py) through the front end and
the oracle, CPU only: each compiles; batching is invisible (the batch of
eight lots derives the same instances as the lots one at a time); a second
derivation is byte-identical; every instance points at a Generate entry of
its own terminal type."""
import numpy as np
import pytest

from operatorgraph_amd import abi, grammars
from operatorgraph_amd.compiler import compile_grammar
from tests.grammar_fuzz import lots, random_grammar
from tests.helpers import assert_same_multiset

SEEDS = range(0, 120)


def _derive(g, ax):
    from oracle import oracle as O
    return O.derive(g.program, ax)


@pytest.mark.parametrize("seed", SEEDS)
def test_fuzz_grammar_oracle_properties(seed):
    from oracle.oracle import OracleError
    text, shape = random_grammar(seed)
    g = compile_grammar(text, f"fuzz{seed}")
    ax = lots(g, 8, seed, quad=shape == "Quad")
    try:
        r = _derive(g, ax)
    except OracleError as e:     # a Subdivide the random sizes over-constrain
        assert "SUBDIV_OVERCONSTRAINED" in str(e)
        return
    again = _derive(g, ax)
    procs = grammars.INTERPRETER_PROCEDURES.procedures
    entries = g.program.table.entries
    for t in range(3):
        a = r.instances(t)
        assert a.tobytes() == again.instances(t).tobytes()
        single = np.concatenate([_derive(g, ax[i:i + 1]).instances(t) for i in range(len(ax))])
        single["axiom"] = np.repeat(np.arange(len(ax), dtype=np.uint32),
                                    [_derive(g, ax[i:i + 1]).count(t) for i in range(len(ax))])
        assert_same_multiset(a, single, f"seed {seed} terminal {t}")
        for m in np.unique(a["meta"] & 0xFFFF):
            e = entries[int(m)]
            assert procs[e.operator_code].op_type == abi.GENERATE
            assert int(e.parameters[0].values[0] if e.parameters[0].values else 0) == t
