"""The on-device autotuner (pga_autotune): it times every launch
configuration on the bound axioms, keeps the fastest, and the derivation it
leaves in the buffers is the oracle's, bit for bit."""
import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, grammars
from tests.helpers import assert_same_multiset

pytestmark = pytest.mark.gpu


def test_autotune_facade_keeps_parity():
    from oracle import oracle as O
    from operatorgraph_amd import engine
    print("autotune facade 256 lots", flush=True)
    lots = axioms.facade_lots(256)
    walls, windows = axioms.facade_lot_terminals(lots)
    eng = engine.Engine(queue_capacity=int(windows.sum()) + 4096)
    eng.load(grammars.facade()).set_axioms(lots)
    eng.bind(0, int(walls.sum()))
    eng.bind(1, int(windows.sum()))
    res = eng.autotune(repeats=2)
    assert res["technique"] in (abi.KERNELS, abi.MEGAKERNEL)
    assert res["block_size"] in (256, 512, 768, 1024)
    assert res["grid_blocks"] > 0
    assert res["candidates"] >= 9
    assert 0 < res["seconds"] <= res["initial_seconds"]
    # the tiles are the largest frontier (one child each); a window tile emits
    # wall|window|wall and a narrow one a single wall, so tiles = walls - windows
    assert res["max_level_children"] == int(walls.sum() - windows.sum())
    ref = O.derive(grammars.facade(), lots)
    for t in range(2):
        assert_same_multiset(eng.instances(t), ref.instances(t), f"terminal {t}")
    # the chosen configuration stays in force
    eng.execute()
    assert eng.stats()["terminals"] == int(walls.sum() + windows.sum())
    eng.close()
