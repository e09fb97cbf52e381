"""GPU parity: the HIP derivation engine (through the C ABI) against the CPU
oracle on the same seeded inputs. Terminal multisets are bit-exact after a
canonical sort; transforms are within 1e-5 (normwise relative) of the
oracle's FP64 re-derivation."""
import sys

import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, grammars
from operatorgraph_amd.grammars import op
from operatorgraph_amd.program import DispatchTable, Entry, Exp, Program, Rand, Scalar, ShapeAttr
from tests.helpers import assert_same_multiset, canon_order, rel_err

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


def gpu_derive(E, program, ax, n_terms, technique, params=None, key=None, cap=None, **cfg):
    eng = E.Engine(technique=technique, **cfg)
    eng.load(program).set_axioms(ax)
    if params is not None:
        eng.set_axiom_params(params)
    if key is not None:
        eng.set_philox_key(key)
    for t in range(n_terms):
        eng.bind(t, cap if cap is not None else 1 << 20)
    eng.execute()
    out = [eng.instances(t) for t in range(n_terms)]
    st = eng.stats()
    eng.close()
    return out, st


def check_against_oracle(E, program, ax, n_terms, technique, params=None, key=None, twins=True, **cfg):
    from oracle import oracle as O
    r = O.derive(program, ax, params=params, key=key or (0, 0), twins=O.TWIN_F64 if twins else 0)
    got, st = gpu_derive(E, program, ax, n_terms, technique, params=params, key=key, **cfg)
    for t in range(n_terms):
        ref = r.instances(t)
        assert_same_multiset(got[t], ref, f"terminal {t}")
        if twins and len(ref):
            f64 = r.instances_f64(t)[:, :12]
            gi = got[t][canon_order(got[t])]
            ri = canon_order(ref)
            assert rel_err(gi["model"], f64[ri]) < 1e-5
    return got, st


@pytest.mark.parametrize("technique", TECHNIQUES)
def test_menger3_interpreter(E, technique):
    _note(f"menger3 interpreter technique={technique}")
    got, st = check_against_oracle(E, grammars.menger_interpreter(), axioms.menger_axiom(1.0, 0), 1, technique)
    assert len(got[0]) == 8000
    assert st["terminals"] == 8000


@pytest.mark.parametrize("technique", TECHNIQUES)
def test_menger_test_runner_table(E, technique):
    _note(f"menger test-runner table size 27 technique={technique}")
    ax = axioms.menger_axiom(27.0, grammars.MENGER_TEST_RUNNER_AXIOM_ENTRY, centred=False)
    got, _ = check_against_oracle(E, grammars.menger_test_runner(), ax, 1, technique)
    assert len(got[0]) == 20 ** 4


@pytest.mark.parametrize("technique", TECHNIQUES)
def test_facade_lots(E, technique):
    _note(f"facade 96 lots technique={technique}")
    lots = axioms.facade_lots(96)
    got, _ = check_against_oracle(E, grammars.facade(), lots, 2, technique)
    walls, windows = axioms.facade_lot_terminals(lots)
    assert len(got[0]) == walls.sum() and len(got[1]) == windows.sum()


@pytest.mark.parametrize("technique", TECHNIQUES)
def test_stochastic_philox(E, technique):
    _note(f"stochastic 300 axioms technique={technique}")
    ax, _ = axioms.stochastic_axioms(300)
    check_against_oracle(E, grammars.stochastic(), ax, 1, technique, key=axioms.STOCHASTIC_KEY)


@pytest.mark.parametrize("technique", TECHNIQUES)
def test_param_facade(E, technique):
    _note(f"parametrised facade technique={technique}")
    ax, params = axioms.param_lots(20, perturb=0.05)
    check_against_oracle(E, grammars.facade_params(), ax, 2, technique, params=params)


def _kitchen_sink():
    """Every supported operator and parameter form in one grammar."""
    gen = lambda t, s=B: Entry(op(abi.GENERATE, s, 1), [Scalar(t)], [])
    e_half = Exp(("a", 12), ("s", -0.25), ("op", abi.OPT_MULTI))
    e_big = Exp(("a", 13), ("s", 1.5), ("op", abi.OPT_GT))
    return Program(grammars.INTERPRETER_PROCEDURES, DispatchTable([
        # 0: box -> translate(Exp) -> rotate -> scale -> if(Exp) -> {compsplit | swapsize}
        Entry(op(abi.TRANSLATE, B), [e_half, Scalar(0.5), Scalar(-1)], [(1, 0)]),
        Entry(op(abi.ROTATE, B), [Scalar(30), Scalar(45), Scalar(-60)], [(2, 0)]),
        Entry(op(abi.SCALE, B), [Scalar(-1), Scalar(-1.5), Scalar(2)], [(3, 0)]),
        Entry(op(abi.IF, B), [e_big], [(4, 0), (9, 0)]),
        Entry(op(abi.COMPSPLIT, B), [], [(5, 0), (6, 0), (7, 0)]),
        # 5 top: extrude -> box random rule
        Entry(op(abi.EXTRUDE, Q), [Scalar(0), Scalar(0.3)], [(10, 0)]),
        Entry(op(abi.DISCARD, Q), [], []),
        # 7 sides: repeat anchor-to-end then rand subdivision
        Entry(op(abi.REPEAT, Q), [Scalar(0), Scalar(0.7), Scalar(abi.ANCHOR_TO_END)], [(8, 0)]),
        Entry(op(abi.SUBDIV, Q), [Scalar(1), Scalar(-1), Rand(0.1, 0.3), Scalar(-2)],
              [(12, 0), (13, 0), (12, 0)]),
        # 9: swapsize -> replicate
        Entry(op(abi.SWAPSIZE, B), [Scalar(2), Scalar(0), Scalar(1)], [(11, 0)]),
        # 10: random rule on the seed
        Entry(op(abi.STOCHASTIC, B), [ShapeAttr(15), Scalar(0.3), Scalar(0.7)], [(14, 0), (15, 0)]),
        Entry(op(abi.REPLICATE, B), [], [(14, 0), (15, 0), (14, 0)]),
        gen(0, Q), gen(1, Q), gen(2), gen(3),
    ]))


@pytest.mark.parametrize("technique", TECHNIQUES)
def test_kitchen_sink_operators(E, technique):
    _note(f"kitchen sink technique={technique}")
    rng = np.random.default_rng(11)
    ax = axioms.empty(64)
    ax["size"] = rng.uniform(0.5, 3.0, size=(64, 3)).astype(np.float32)
    ax["model"][:, [3, 7, 11]] = rng.uniform(-5, 5, size=(64, 3)).astype(np.float32)
    ax["seed"] = rng.uniform(0, 1, 64).astype(np.float32)
    ax["custom"] = ax["seed"]
    ax["entry_index"] = 0
    ax["shape_type"] = B
    got, _ = check_against_oracle(E, _kitchen_sink(), ax, 4, technique)
    assert sum(len(g) for g in got) > 64


@pytest.mark.parametrize("block_size", [256, 768, 1024])
def test_workgroup_size_variants(E, block_size):
    # each workgroup size is its own kernel build (register budget 256/168/128)
    _note(f"workgroup size variants block_size={block_size}")
    rng = np.random.default_rng(12)
    ax = axioms.empty(64)
    ax["size"] = rng.uniform(0.5, 3.0, size=(64, 3)).astype(np.float32)
    ax["model"][:, [3, 7, 11]] = rng.uniform(-5, 5, size=(64, 3)).astype(np.float32)
    ax["seed"] = rng.uniform(0, 1, 64).astype(np.float32)
    ax["entry_index"] = 0
    ax["shape_type"] = B
    for technique in TECHNIQUES:
        check_against_oracle(E, _kitchen_sink(), ax, 4, technique, block_size=block_size)
    sx, _ = axioms.stochastic_axioms(200)
    check_against_oracle(E, grammars.stochastic(), sx, 1, abi.MEGAKERNEL, key=axioms.STOCHASTIC_KEY,
                         block_size=block_size)
    lots = axioms.facade_lots(48)
    check_against_oracle(E, grammars.facade(), lots, 2, abi.MEGAKERNEL, block_size=block_size)


def test_error_paths(E):
    _note("error paths")
    from operatorgraph_amd.engine import PGAError
    lots = axioms.facade_lots(8)
    # instance buffer too small -> OVERFLOW_INSTANCES, counts still complete
    eng = E.Engine()
    eng.load(grammars.facade()).set_axioms(lots)
    eng.bind(0, 16)
    eng.bind(1, 16)
    with pytest.raises(PGAError) as e:
        eng.execute()
    assert e.value.status == abi.ERR_OVERFLOW_INSTANCES
    walls, windows = axioms.facade_lot_terminals(lots)
    assert list(eng.counts(2)) == [walls.sum(), windows.sum()]
    eng.close()
    # frontier too small -> OVERFLOW_QUEUE
    for tech in TECHNIQUES:
        eng = E.Engine(technique=tech, queue_capacity=64)
        eng.load(grammars.menger_interpreter()).set_axioms(axioms.menger_axiom(1.0, 0))
        eng.bind(0, 1 << 14)
        with pytest.raises(PGAError) as e:
            eng.execute()
        assert e.value.status == abi.ERR_OVERFLOW_QUEUE
        eng.close()
    # over-constrained subdivision and bad repeat extent raise, kernels end
    gen = Entry(op(abi.GENERATE, B, 1), [Scalar(0)], [])
    over = Program(grammars.INTERPRETER_PROCEDURES, DispatchTable([
        Entry(op(abi.SUBDIV, B), [Scalar(0), Scalar(3), Scalar(-1)], [(1, 0), (1, 0)]), gen]))
    bad = Program(grammars.INTERPRETER_PROCEDURES, DispatchTable([
        Entry(op(abi.REPEAT, B), [Scalar(0), Scalar(0.0), Scalar(0)], [(1, 0)]), gen]))
    a = axioms.empty(4)
    a["size"] = (2, 1, 1)
    a["entry_index"] = 0
    a["shape_type"] = B
    for prog, status in ((over, abi.ERR_SUBDIV_OVERCONSTRAINED), (bad, abi.ERR_REPEAT_BAD_EXTENT)):
        eng = E.Engine()
        eng.load(prog).set_axioms(a)
        eng.bind(0, 64)
        with pytest.raises(PGAError) as e:
            eng.execute()
        assert e.value.status == status
        eng.close()
    # program validation happens on load
    eng = E.Engine()
    with pytest.raises(PGAError) as e:
        eng.load(Program(grammars.INTERPRETER_PROCEDURES, DispatchTable([
            Entry(op(abi.COMPSPLIT, B), [], [(1, 0), (1, 0), (1, 0)]), gen])))
    assert e.value.status == abi.ERR_BAD_PROGRAM    # a Quad child sent to a Box entry
    eng.close()


def test_count_only_and_append(E):
    _note("count-only and append")
    lots = axioms.facade_lots(64)
    eng = E.Engine()
    eng.load(grammars.facade()).set_axioms(lots)
    eng.bind(0, 0)
    eng.bind(1, 0)
    eng.execute()
    walls, windows = axioms.facade_lot_terminals(lots)
    assert list(eng.counts(2)) == [walls.sum(), windows.sum()]
    # two halves appended equal one run
    eng.bind(0, 1 << 18)
    eng.bind(1, 1 << 18)
    eng.execute(32, first=0)
    eng.execute(32, first=32, append=True)
    assert list(eng.counts(2)) == [walls.sum(), windows.sum()]
    halves = [eng.instances(t) for t in range(2)]
    eng.execute()
    whole = [eng.instances(t) for t in range(2)]
    for t in range(2):
        assert_same_multiset(halves[t], whole[t], "append")
    eng.close()


def test_menger5_full_size_properties(E):
    """Menger-5 (test-runner table, axiom of size 81): 3.2M cubes; checked
    by size-independent properties."""
    _note("menger5 full size")
    ax = axioms.menger_axiom(81.0, grammars.MENGER_TEST_RUNNER_AXIOM_ENTRY, centred=False)
    eng = E.Engine(queue_capacity=1 << 22)
    eng.load(grammars.menger_test_runner()).set_axioms(ax)
    buf = eng.bind(0, 3_200_000)
    eng.execute()
    assert int(eng.counts(1)[0]) == 3_200_000
    m = buf.view(-1, 16)[:, :12].double().cpu().numpy()
    s = np.float32(81.0)
    for _ in range(5):
        s = np.float32(s / np.float32(3.0))      # FP32 chain of relative splits
    assert np.all(m[:, [0, 5, 10]] == float(s))
    # symmetric about the axiom centre (0, 40.5, 0)
    c = m[:, [3, 7, 11]]
    assert np.allclose(c.mean(axis=0), (0.0, 40.5, 0.0), atol=1e-3)
    # all distinct cells
    cells = np.rint((c - (0.0, 40.5, 0.0) + 40.5) * 3.0 - 0.5).astype(np.int64)
    keys = (cells[:, 0] * 243 + cells[:, 1]) * 243 + cells[:, 2]
    assert len(np.unique(keys)) == 3_200_000
    eng.close()


def test_facade_lots_1024_full_size(E):
    """The headline workload at its own size: 1,024 facade lots (about 1.2M
    terminals), bit-exact against the oracle, transforms within 1e-5 of its
    FP64 twin, and the closed-form counts per lot."""
    _note("facade 1024 lots (headline workload) megakernel")
    lots = axioms.facade_lots(1024)
    walls, windows = axioms.facade_lot_terminals(lots)
    got, st = check_against_oracle(E, grammars.facade(), lots, 2, abi.MEGAKERNEL,
                                   cap=int(max(walls.sum(), windows.sum())), queue_capacity=1 << 22)
    assert len(got[0]) == walls.sum() and len(got[1]) == windows.sum()
    assert np.array_equal(np.bincount(got[1]["axiom"], minlength=1024), windows)
    assert st["terminals"] == walls.sum() + windows.sum()


def _compsplit_program(parallel):
    from operatorgraph_amd.program import Procedure, ProcedureList
    procs = ProcedureList([Procedure(abi.COMPSPLIT, B, 0, abi.PROC_PARALLEL if parallel else 0),
                           Procedure(abi.TRANSLATE, Q), Procedure(abi.GENERATE, Q, 1)])
    # each face is moved along its normal by its seed, so that the seeds show in the instances
    tr = lambda nxt: Entry(1, [Scalar(0), Scalar(0), ShapeAttr(15)], [(nxt, 0)])
    return Program(procs, DispatchTable([Entry(0, [], [(1, 0), (2, 0), (3, 0)]), tr(4), tr(5), tr(6),
                                         Entry(2, [Scalar(0)], []), Entry(2, [Scalar(1)], []),
                                         Entry(2, [Scalar(2)], [])]))


@pytest.mark.parametrize("technique", TECHNIQUES)
@pytest.mark.parametrize("parallel", [False, True])
def test_parallel_compsplit(E, technique, parallel):
    """ComponentSplit<true> (ParallelFaceSplitter<Box>) next to the serial
    form, bit-exact against the oracle (signed zeros of its face frames
    included) on boxes of random size, pose and seed."""
    _note(f"compsplit parallel={parallel} technique={technique}")
    rng = np.random.default_rng(17)
    n = 200
    a = axioms.empty(n)
    a["size"] = rng.uniform(0.5, 5.0, size=(n, 3)).astype(np.float32)
    a["model"][:, [3, 7, 11]] = rng.uniform(-10, 10, size=(n, 3)).astype(np.float32)
    a["model"][:, [0, 5, 10]] = 1.0
    a["model"][: n // 2, [1, 4]] = (0.6, -0.6)      # some sheared frames: no zero products
    a["seed"] = rng.random(n).astype(np.float32)
    a["entry_index"] = 0
    a["shape_type"] = B
    check_against_oracle(E, _compsplit_program(parallel), a, 3, technique)


def test_parallel_repeat_is_refused(E):
    """Repeat<true> traps in the reference (Repeat.cuh:35-61): refused on load."""
    from operatorgraph_amd.engine import PGAError
    from operatorgraph_amd.program import Procedure, ProcedureList
    procs = ProcedureList([Procedure(abi.REPEAT, B, 0, abi.PROC_PARALLEL), Procedure(abi.GENERATE, B, 1)])
    prog = Program(procs, DispatchTable([Entry(0, [Scalar(0), Scalar(0.3), Scalar(abi.ADJUST_TO_FILL)], [(1, 0)]),
                                         Entry(1, [Scalar(0)], [])]))
    eng = E.Engine()
    with pytest.raises(PGAError) as e:
        eng.load(prog)
    assert e.value.status == abi.ERR_UNSUPPORTED
    eng.close()


def test_set_seeds_copies_n_mod_33(E):
    """Device::setSeeds copies (n % 33) * 4 bytes (GlobalVariables.cuh:28-32):
    33 seeds copy nothing, 1,024 seeds copy one, 5 seeds copy five."""
    eng = E.Engine()
    base = np.arange(1, 33, dtype=np.float32) / 64.0
    eng.set_seeds(base)
    assert np.array_equal(eng.seeds(), base)
    eng.set_seeds(np.full(33, 0.75, dtype=np.float32))
    assert np.array_equal(eng.seeds(), base)
    big = np.full(1024, 0.5, dtype=np.float32)
    eng.set_seeds(big)
    want = base.copy()
    want[0] = 0.5
    assert np.array_equal(eng.seeds(), want)
    eng.set_seeds(np.full(5, 0.25, dtype=np.float32))
    want[:5] = 0.25
    assert np.array_equal(eng.seeds(), want)
    eng.close()
