"""GPU parity for grammars that come through the front end: the reference's
interpreter grammars (tables compiled by pga_compile_grammar and stored in
tests/golden/grammar_tables.json) and a grammar compiled on the spot, each
derived by the HIP engine through the C ABI and checked against the oracle:
terminal multisets bit-exact after a canonical sort, transforms within 1e-5
(normwise relative) of the oracle's FP64 re-derivation."""
import sys

import numpy as np
import pytest

from operatorgraph_amd import abi, axioms
from operatorgraph_amd.compiler import compile_grammar
from tests.helpers import (TREE_COUNTS, assert_same_multiset, canon_order, fixture_grammar, grammar_fixture,
                           grammar_lots, rel_err)

pytestmark = pytest.mark.gpu

SUPPORTED = sorted(k for k, v in grammar_fixture().items() if "refused" not in v)


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check(program, ax, technique, n_terms=abi.MAX_TERMINALS):
    from oracle import oracle as O
    from operatorgraph_amd import engine
    r = O.derive(program, ax, twins=O.TWIN_F64)
    counts = [r.count(t) for t in range(n_terms)]
    eng = engine.Engine(technique=technique)
    eng.load(program).set_axioms(ax)
    used = [t for t in range(n_terms) if counts[t]]
    for t in used:
        eng.bind(t, counts[t] + 1024)
    eng.execute()
    got = {t: eng.instances(t) for t in used}
    eng.close()
    for t in used:
        ref = r.instances(t)
        assert_same_multiset(got[t], ref, f"{program.name} terminal {t}")
        f64 = r.instances_f64(t)[:, :12]
        assert rel_err(got[t][canon_order(got[t])]["model"], f64[canon_order(ref)]) < 1e-5
    return counts


@pytest.mark.parametrize("technique", [abi.KERNELS, abi.MEGAKERNEL])
@pytest.mark.parametrize("name", SUPPORTED)
def test_reference_grammar(name, technique):
    n = 32
    _note(f"reference grammar {name}: {n} lots, technique={technique}")
    g = fixture_grammar(name)
    counts = check(g.program, grammar_lots(g, name, n), technique)
    assert sum(counts) > 0
    if name in TREE_COUNTS:
        assert tuple(counts[:2]) == tuple(n * c for c in TREE_COUNTS[name])
    if name == "mc_spaceship":
        assert counts[:3] == [n * 10, n * 13, n]
    if name == "menger_sponge_3":
        assert counts[0] == n * 20 ** 3


def test_compiled_on_the_spot():
    """A grammar compiled in this process (every operand form and operator
    the engine runs), then derived on the GPU."""
    src = """
    var pitch = 2.5;
    axiom Box Lot;
    terminal Wall (1, 0);
    terminal Window (1, 1);
    terminal Roof Cap (1, 2);
    Lot = SubDiv(Y) { -1: Body | 1.5: Translate(0, Exp(ShapeSize(Y) * 0.5), 0) Scale(-1.1, -1, -1.1) Roof };
    Body = CompSplit() { Cap | Discard() | Facade };
    Facade = Repeat(Y, 3, ANCHOR_TO_END) Floor;
    Floor = Repeat(X, pitch, ADJUST_TO_FILL) Tile;
    Tile = IfSizeLess(Y, 2) { Wall | RandomRule(ShapeSeed()) { 0.6: Glazed | 0.4: Extrude(0.2) SwapSize(X, Z, Y) Roof } };
    Glazed = If(Exp(Exp(ShapeNormal(Z) > 0.5) || Exp(ShapeSize(X) >= 2.4))) {
        SubDiv(X) { -1: Wall | 1.2: Rotate(0, 0, Exp(ShapeSeed() * 10)) Window | -1: Wall } |
        Replicate() { Wall | Scale(Rand(0.5, 0.9), -1, -1) Window }
    };
    """
    _note("compiled-on-the-spot grammar, 48 lots")
    g = compile_grammar(src, "on_the_spot")
    ax = axioms.facade_lots(48)
    ax["size"][:, 1] += 1.5
    counts = check(g.program, ax, abi.MEGAKERNEL, n_terms=3)
    assert all(c > 0 for c in counts)
