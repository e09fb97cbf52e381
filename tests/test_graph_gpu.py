"""HIP graphs of repeated derivations (pga_capture_begin, pga_graph_launch)
and the device-count LBVH recorded in the same graph. Every replay must equal
an eager run bit for bit; replays alternate with eager work (torch fills,
other derivations) and host synchronizations, and parameter values updated
in place between replays must be picked up (including a change of the
terminal counts, which the device-count LBVH follows)."""
import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, grammars
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)


def _snapshot(eng, n_types):
    """Instance records as int32 words (axiom, path and meta are integers
    stored in float slots: some are NaN patterns, so compare bits)."""
    import torch
    c = eng.counts(n_types)
    return [eng.buffer(t)[: int(c[t])].clone().view(torch.int32) for t in range(n_types)], c


def test_graph_replays_equal_eager():
    import torch
    from operatorgraph_amd import engine
    _note("graph replay: 512 facade lots, 3 replays with eager work between")
    lots = axioms.facade_lots(512)
    walls, windows = axioms.facade_lot_terminals(lots)
    eng = engine.Engine(queue_capacity=1 << 20)
    eng.load(grammars.facade()).set_axioms(lots)
    eng.bind(0, int(walls.sum()))
    eng.bind(1, int(windows.sum()))
    eng.execute()
    ref, rc = _snapshot(eng, 2)
    assert rc[0] == walls.sum() and rc[1] == windows.sum()
    with eng.capture() as g:
        eng.execute_async()
    kernels, cooperative = g.info()
    assert kernels >= 2 and cooperative >= 1, (kernels, cooperative)
    for rep in range(3):
        for t in range(2):
            eng.buffer(t).zero_()
        torch.cuda.synchronize()
        g.launch()
        eng.synchronize()
        got, gc = _snapshot(eng, 2)
        assert np.array_equal(gc, rc), (rep, gc[:2], rc[:2])
        for t in range(2):
            assert torch.equal(got[t], ref[t]), (rep, t)
        eng.execute(n=100)                       # eager derivation between replays
        assert eng.counts(2)[1] == windows[:100].sum()
    g.close()
    eng.close()


def test_graph_of_repeated_derivations_appends():
    """Three derivations of disjoint axiom ranges in one graph, the later two
    appending: the counters and buffers equal one eager run of all axioms."""
    import torch
    from operatorgraph_amd import engine
    _note("graph of 3 appended derivations")
    lots = axioms.facade_lots(300)
    walls, windows = axioms.facade_lot_terminals(lots)
    eng = engine.Engine(queue_capacity=1 << 20)
    eng.load(grammars.facade()).set_axioms(lots)
    eng.bind(0, int(walls.sum()))
    eng.bind(1, int(windows.sum()))
    eng.execute()
    ref = [x.cpu().numpy().view(abi.INSTANCE_DTYPE).reshape(-1) for x in _snapshot(eng, 2)[0]]
    with eng.capture() as g:
        eng.execute_async(n=100, first=0)
        eng.execute_async(n=120, first=100, append=True)
        eng.execute_async(n=80, first=220, append=True)
    for rep in range(2):
        g.launch()
        eng.synchronize()
        got, gc = _snapshot(eng, 2)
        assert gc[0] == walls.sum() and gc[1] == windows.sum()
        for t in range(2):
            from tests.helpers import assert_same_multiset
            assert_same_multiset(got[t].cpu().numpy().view(abi.INSTANCE_DTYPE).reshape(-1), ref[t], f"type {t}")
        torch.cuda.synchronize()
    g.close()
    eng.close()


def test_graph_refuses_changed_bindings():
    from operatorgraph_amd import engine
    from operatorgraph_amd.engine import PGAError
    _note("graph refuses a replay after a rebind")
    lots = axioms.facade_lots(16)
    walls, windows = axioms.facade_lot_terminals(lots)
    eng = engine.Engine(queue_capacity=1 << 18)
    eng.load(grammars.facade()).set_axioms(lots)
    eng.bind(0, int(walls.sum()))
    eng.bind(1, int(windows.sum()))
    with eng.capture() as g:
        eng.execute_async()
    g.launch()
    eng.synchronize()
    eng.bind(1, int(windows.sum()) + 8)          # a new buffer: the recorded address is stale
    with pytest.raises(PGAError) as e:
        g.launch()
    assert e.value.status == abi.ERR_INVALID_ARG
    with pytest.raises(PGAError):                # calls that wait are refused while recording
        with eng.capture():
            eng.counts(2)
    g.close()
    eng.close()


def _eager_boxes(eng, bvh, counts):
    import torch
    parts = [bvh.instance_aabbs(eng.buffer(t)[: int(counts[t])], planar=True) for t in range(2) if counts[t]]
    return torch.cat(parts, dim=0)


def test_graph_derivation_and_lbvh_follow_parameter_updates():
    """One graph = derivation + terminal gather + LBVH over the device count
    with the scene box computed on the device. Parameter values change in
    place between replays, which changes the terminal counts; each replay's
    boxes equal the eager boxes and its tree equals the oracle's."""
    import torch
    from operatorgraph_amd import engine
    from operatorgraph_amd.lbvh import LBVH
    _note("graph: derivation + device-count LBVH, parameters updated between replays")
    ax, p0 = axioms.param_lots(48, perturb=0.05)
    p1 = p0.copy()
    p1[:, 2] = 2.3                               # lower floors and narrower tiles: more terminals
    p1[:, 3] = 1.9
    prog = grammars.facade_params()
    counts = []
    for p in (p0, p1):
        r = O.derive(prog, ax, params=p)
        counts.append((r.count(0), r.count(1)))
    cap = [max(c[t] for c in counts) for t in range(2)]
    assert counts[1][0] > counts[0][0]
    eng = engine.Engine(queue_capacity=1 << 20)
    eng.load(prog).set_axioms(ax).set_axiom_params(p0)
    eng.bind(0, cap[0])
    eng.bind(1, cap[1])
    bvh = LBVH(cap[0] + cap[1], engine=eng)
    out = bvh.alloc()
    with eng.capture() as g:
        eng.execute_async()
        bvh.gather_engine(eng, 2, 0b11, out)
        bvh.build_device(out)
    eager = LBVH(cap[0] + cap[1])
    for rep, p in enumerate((p0, p1, p0)):
        eng.set_axiom_params(p)                  # same shape: updated in place
        out["keys"].fill_(-1)
        torch.cuda.synchronize()
        g.launch()
        eng.synchronize()
        n, overflow = bvh.leaf_count()
        c = eng.counts(2)
        want = counts[0] if rep != 1 else counts[1]
        assert (int(c[0]), int(c[1])) == want and not overflow and n == sum(want), (rep, c[:2], n)
        boxes = _eager_boxes(eng, eager, c)
        assert torch.equal(out["aabbs"][:n].view(torch.int32), boxes.view(torch.int32)), rep
        b = boxes.cpu().numpy()
        smin, smax = b[:, :3].min(0), b[:, 3:].max(0)
        ref = O.lbvh(b, smin, smax)
        got = {k: v.cpu().numpy() for k, v in out.items()}
        assert np.array_equal(got["keys"][:n].view(np.uint64), ref["keys"]), rep
        assert np.array_equal(got["left"][: n - 1], ref["left"]), rep
        assert np.array_equal(got["right"][: n - 1], ref["right"]), rep
        assert np.array_equal(got["parent"][1 : n - 1], ref["parent"][1:]), rep
        assert np.array_equal(got["leaf_parent"][:n], ref["leaf_parent"]), rep
        assert np.array_equal(got["node_aabb"][: n - 1].view(np.uint32), ref["node_aabb"].view(np.uint32)), rep
        eng.execute(n=5)                         # eager derivation between replays
    g.close()
    bvh.close()
    eager.close()
    eng.close()


def test_device_count_build_equals_host_count_build():
    """pga_lbvh_build_device with a host scene box over a gathered count
    equals pga_lbvh_build with the same box and the host count."""
    import torch
    from operatorgraph_amd import engine
    from operatorgraph_amd.lbvh import LBVH
    _note("device-count LBVH vs host-count LBVH, stochastic terminals")
    ax, _ = axioms.stochastic_axioms(512)
    eng = engine.Engine(queue_capacity=1 << 20)
    eng.load(grammars.stochastic()).set_axioms(ax).set_philox_key(axioms.STOCHASTIC_KEY)
    r = O.derive(grammars.stochastic(), ax, key=axioms.STOCHASTIC_KEY)
    nt = [r.count(t) for t in range(abi.MAX_TERMINALS)]
    ntypes = max(t for t in range(abi.MAX_TERMINALS) if nt[t]) + 1
    for t in range(ntypes):
        eng.bind(t, max(nt[t], 1))
    eng.execute()
    total = sum(nt[:ntypes])
    bvh = LBVH(total + 1000, engine=eng)   # capacity above the count
    out = bvh.alloc()
    bvh.gather_engine(eng, ntypes, 0, out)
    smin, smax = (-500.0, -10.0, -500.0), (2000.0, 5000.0, 2000.0)
    bvh.build_device(out, smin, smax)
    n, overflow = bvh.leaf_count()
    assert n == total and not overflow
    host = LBVH(total)
    res = host.build(out["aabbs"][:n].clone(), smin, smax)
    torch.cuda.synchronize()
    for k in ("keys", "leaf_parent"):
        assert torch.equal(out[k][:n], res[k]), k
    for k in ("left", "right"):
        assert torch.equal(out[k][: n - 1], res[k]), k
    assert torch.equal(out["node_aabb"][: n - 1].view(torch.int32), res["node_aabb"].view(torch.int32))
    assert torch.equal(out["parent"][1 : n - 1], res["parent"][1:])
    host.close()
    bvh.close()
    eng.close()


def test_combined_grammars_in_one_launch():
    """Facade lots and a Menger sponge (grammars.combine) derived by one
    persistent launch, recorded in a graph: equal to the oracle per type."""
    from operatorgraph_amd import engine
    from tests.helpers import assert_same_multiset
    _note("combined facade + menger grammars, one launch in a graph")
    lots = axioms.facade_lots(64)
    cube = axioms.menger_axiom(1.0, 0)
    prog, starts = grammars.combine((grammars.facade(), 0), (grammars.menger_interpreter(), 2))
    both = np.concatenate([cube, lots])
    both["entry_index"][0] = starts[1]
    r = O.derive(prog, both)
    eng = engine.Engine(queue_capacity=1 << 20)
    eng.load(prog).set_axioms(both)
    for t in range(3):
        eng.bind(t, r.count(t))
    with eng.capture() as g:
        eng.execute_async()
    g.launch()
    eng.synchronize()
    for t in range(3):
        assert_same_multiset(eng.instances(t), r.instances(t), f"terminal {t}")
    g.close()
    eng.close()
