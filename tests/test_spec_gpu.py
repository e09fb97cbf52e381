"""GPU parity of the kernels specialised to one program (pga_specialize):
the program written as constant entry types and compiled at run time must
derive the oracle's instances bit for bit (canonical sort; transforms within
1e-5 of its FP64 twin), over every operator, the reference's interpreter
grammars, random grammars, the error statuses, appended and count-only runs,
graph replays and the autotuner's choice."""
import sys

import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, grammars
from operatorgraph_amd.compiler import compile_grammar
from tests.grammar_fuzz import lots, random_grammar
from tests.helpers import (assert_same_multiset, canon_order, fixture_grammar, grammar_fixture, grammar_lots,
                           rel_err)
from tests.test_engine_gpu import _compsplit_program, _kitchen_sink

pytestmark = pytest.mark.gpu

B = abi.BOX


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def spec_check(program, ax, n_terms, params=None, key=None, block_size=512, cap=1 << 20, twins=True):
    from oracle import oracle as O
    from operatorgraph_amd import engine
    r = O.derive(program, ax, params=params, key=key or (0, 0), twins=O.TWIN_F64 if twins else 0)
    eng = engine.Engine(queue_capacity=1 << 22)
    eng.load(program).set_axioms(ax)
    if params is not None:
        eng.set_axiom_params(params)
    if key is not None:
        eng.set_philox_key(key)
    for t in range(n_terms):
        eng.bind(t, cap)
    sec = eng.specialize(block_size)
    _note(f"  compiled in {sec:.2f} s")
    eng.execute()
    got = [eng.instances(t) for t in range(n_terms)]
    st = eng.stats()
    eng.close()
    for t in range(n_terms):
        ref = r.instances(t)
        assert_same_multiset(got[t], ref, f"{program.name} terminal {t}")
        if twins and len(ref):
            f64 = r.instances_f64(t)[:, :12]
            assert rel_err(got[t][canon_order(got[t])]["model"], f64[canon_order(ref)]) < 1e-5
    return got, st


def _boxes(n, seed):
    rng = np.random.default_rng(seed)
    ax = axioms.empty(n)
    ax["size"] = rng.uniform(0.5, 3.0, size=(n, 3)).astype(np.float32)
    ax["model"][:, [3, 7, 11]] = rng.uniform(-5, 5, size=(n, 3)).astype(np.float32)
    ax["seed"] = rng.uniform(0, 1, n).astype(np.float32)
    ax["custom"] = ax["seed"]
    ax["entry_index"] = 0
    ax["shape_type"] = B
    return ax


def test_spec_menger_and_facade():
    _note("specialised: menger-3, menger test-runner table, facade 1024 lots")
    got, _ = spec_check(grammars.menger_interpreter(), axioms.menger_axiom(1.0, 0), 1)
    assert len(got[0]) == 8000
    ax = axioms.menger_axiom(27.0, grammars.MENGER_TEST_RUNNER_AXIOM_ENTRY, centred=False)
    got, _ = spec_check(grammars.menger_test_runner(), ax, 1)
    assert len(got[0]) == 20 ** 4
    lots = axioms.facade_lots(1024)
    walls, windows = axioms.facade_lot_terminals(lots)
    got, st = spec_check(grammars.facade(), lots, 2, cap=int(walls.sum()))
    assert len(got[0]) == walls.sum() and len(got[1]) == windows.sum()


@pytest.mark.parametrize("block_size", [256, 512, 768, 1024])
def test_spec_every_operator(block_size):
    _note(f"specialised: kitchen sink, stochastic, parameterised facade, parallel split; block {block_size}")
    spec_check(_kitchen_sink(), _boxes(64, 11), 4, block_size=block_size)
    sx, _ = axioms.stochastic_axioms(300)
    spec_check(grammars.stochastic(), sx, 1, key=axioms.STOCHASTIC_KEY, block_size=block_size)
    ax, params = axioms.param_lots(20, perturb=0.05)
    spec_check(grammars.facade_params(), ax, 2, params=params, block_size=block_size)
    for parallel in (False, True):
        spec_check(_compsplit_program(parallel), _boxes(50, 17), 3, block_size=block_size)


@pytest.mark.parametrize("name", sorted(k for k, v in grammar_fixture().items() if "refused" not in v))
def test_spec_reference_grammar(name):
    _note(f"specialised: reference grammar {name}")
    g = fixture_grammar(name)
    from oracle import oracle as O
    ax = grammar_lots(g, name, 16)
    r = O.derive(g.program, ax)
    used = [t for t in range(abi.MAX_TERMINALS) if r.count(t)]
    spec_check(g.program, ax, max(used) + 1, cap=max(r.count(t) for t in used) + 1024, twins=False)


@pytest.mark.parametrize("seed", list(range(0, 96, 6)) + [246])
def test_spec_fuzz_grammar(seed):
    from oracle import oracle as O
    from operatorgraph_amd import engine
    text, shape = random_grammar(seed)
    g = compile_grammar(text, f"fuzz{seed}")
    ax = lots(g, 8, seed, quad=shape == "Quad")
    _note(f"specialised: fuzz grammar seed={seed}")
    try:
        r = O.derive(g.program, ax)
        oracle_error = None
    except O.OracleError as e:
        r, oracle_error = None, e.status
    eng = engine.Engine()
    try:
        eng.load(g.program).set_axioms(ax)
        for t in range(3):
            eng.bind(t, 1 << 16)
        eng.specialize()
        if oracle_error is not None:
            with pytest.raises(engine.PGAError) as info:
                eng.execute()
            assert info.value.status == oracle_error
            return
        eng.execute()
        for t in range(3):
            assert_same_multiset(eng.instances(t), r.instances(t), f"seed {seed} terminal {t}")
    finally:
        eng.close()


@pytest.mark.parametrize("local,grab", [(False, 0), (True, 0), (True, 1), (True, 7), (True, 64)])
def test_spec_level_schedules(local, grab):
    """Every level schedule (pga_set_schedule: grid-wide levels, or
    workgroup-local levels with any hand-out size) derives the oracle's
    multisets: lots of different depths, Philox rule choices, and a sponge
    whose frontier outgrows the workgroup arrays (the spill to the grid-wide
    levels)."""
    from oracle import oracle as O
    from operatorgraph_amd import engine
    _note(f"specialised: schedule local={local} grab={grab}")
    cases = [(grammars.facade(), axioms.facade_lots(300), None, 2),
             (grammars.stochastic(), axioms.stochastic_axioms(2000)[0], axioms.STOCHASTIC_KEY, 1),
             (grammars.menger_test_runner(),
              axioms.menger_axiom(27.0, grammars.MENGER_TEST_RUNNER_AXIOM_ENTRY, centred=False), None, 1)]
    for program, ax, key, n_terms in cases:
        r = O.derive(program, ax, key=key or (0, 0))
        eng = engine.Engine(queue_capacity=1 << 22)
        eng.load(program).set_axioms(ax)
        if key is not None:
            eng.set_philox_key(key)
        for t in range(n_terms):
            eng.bind(t, r.count(t) + 64)
        eng.specialize(256)
        eng.set_schedule(local, grab)
        for rep in range(2):
            eng.execute()
            for t in range(n_terms):
                assert_same_multiset(eng.instances(t), r.instances(t), f"{program.name} terminal {t} run {rep}")
        eng.close()


def test_spec_errors_counts_append_and_switch_back():
    from operatorgraph_amd import engine
    from operatorgraph_amd.engine import PGAError
    _note("specialised: overflow statuses, count-only, append, back to the interpreter")
    lots = axioms.facade_lots(64)
    walls, windows = axioms.facade_lot_terminals(lots)
    eng = engine.Engine()
    eng.load(grammars.facade()).set_axioms(lots)
    eng.specialize()
    eng.bind(0, 16)
    eng.bind(1, 16)
    with pytest.raises(PGAError) as e:
        eng.execute()
    assert e.value.status == abi.ERR_OVERFLOW_INSTANCES
    assert list(eng.counts(2)) == [walls.sum(), windows.sum()]
    eng.bind(0, 0)
    eng.bind(1, 0)
    eng.execute()
    assert list(eng.counts(2)) == [walls.sum(), windows.sum()]
    eng.bind(0, 1 << 18)
    eng.bind(1, 1 << 18)
    eng.execute(32, first=0)
    eng.execute(32, first=32, append=True)
    halves = [eng.instances(t) for t in range(2)]
    eng.specialize(enable=False)
    eng.execute()
    whole = [eng.instances(t) for t in range(2)]
    for t in range(2):
        assert_same_multiset(halves[t], whole[t], "append")
    eng.close()
    small = engine.Engine(queue_capacity=64)
    small.load(grammars.menger_interpreter()).set_axioms(axioms.menger_axiom(1.0, 0))
    small.bind(0, 1 << 14)
    small.specialize()
    with pytest.raises(PGAError) as e:
        small.execute()
    assert e.value.status == abi.ERR_OVERFLOW_QUEUE
    small.close()


def test_spec_graph_replay_and_autotune():
    """Replays of a graph recorded with the specialised kernel equal the
    eager run, with eager work and synchronisations between them; the
    autotuner times the specialised kernel beside the interpreting ones and
    its choice derives the same instances."""
    import torch
    from operatorgraph_amd import engine
    _note("specialised: graph replays and autotune")
    lots = axioms.facade_lots(256)
    eng = engine.Engine()
    eng.load(grammars.facade()).set_axioms(lots)
    eng.bind(0, 1 << 20)
    eng.bind(1, 1 << 20)
    eng.specialize()
    eng.execute()
    ref = [eng.instances(t) for t in range(2)]
    with eng.capture() as g:
        eng.execute_async()
    k, coop = g.info()
    assert coop >= 1
    for rep in range(3):
        for t in range(2):
            eng.buffer(t).zero_()
        torch.cuda.synchronize()
        g.launch()
        eng.synchronize()
        for t in range(2):
            assert_same_multiset(eng.instances(t), ref[t], f"replay {rep} terminal {t}")
        eng.execute()            # eager work between replays
    g.close()
    res = eng.autotune(repeats=2)
    _note(f"  autotune chose specialized={res['specialized']} block={res['block_size']} grab={res['local_grab']}")
    assert res["candidates"] >= 5
    eng.execute()
    for t in range(2):
        assert_same_multiset(eng.instances(t), ref[t], f"tuned terminal {t}")
    eng.close()
