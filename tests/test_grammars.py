"""Host-side grammar tables: grammars.combine merges dispatch tables so that
one launch derives axioms of several grammars. Checked on the oracle: the
merged derivation equals the separate derivations, terminal indices moved."""
import numpy as np

from operatorgraph_amd import abi, axioms, grammars
from oracle import oracle as O
from tests.helpers import assert_same_multiset


def test_combine_equals_separate_derivations():
    lots = axioms.facade_lots(12)
    cube = axioms.menger_axiom(1.0, 0)
    prog, starts = grammars.combine((grammars.facade(), 0), (grammars.menger_interpreter(), 2))
    assert starts == [0, len(grammars.facade().table.entries)]
    both = np.concatenate([lots, cube])
    both["entry_index"][len(lots):] = starts[1]
    r = O.derive(prog, both)
    f = O.derive(grammars.facade(), lots)
    m = O.derive(grammars.menger_interpreter(), cube)
    assert r.count(2) == 8000 and m.count(0) == 8000
    for t in range(2):
        assert_same_multiset(r.instances(t), f.instances(t), f"facade terminal {t}")
    got = r.instances(2)
    want = m.instances(0).copy()
    want["axiom"] += len(lots)            # the cube is axiom 12 of the merged batch
    want["meta"] += starts[1]             # meta = (depth << 16) | Generate entry, moved with the table
    assert_same_multiset(got, want, "menger terminal")


def test_combine_rejects_other_procedure_lists():
    import pytest
    with pytest.raises(ValueError):
        grammars.combine((grammars.menger_test_runner(), 0))
