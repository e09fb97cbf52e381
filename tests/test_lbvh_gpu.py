"""GPU LBVH (csrc/pga_lbvh.hip) against the oracle's sequential restatement
of GPUBVHConstructor.cuh: keys, topology and node boxes bit-exact; at
full size, the tree's invariants."""
import numpy as np
import pytest

from operatorgraph_amd import axioms, grammars
from oracle import oracle as O

pytestmark = pytest.mark.gpu

f32 = np.float32


def _note(msg):
    print(msg, flush=True)


def _boxes(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:   # many centres share a Morton cell: ties resolved by the index
        lo = rng.integers(0, 8, size=(n, 3)).astype(f32) * f32(0.01) + f32(40.0)
    else:
        lo = rng.uniform(0, 90, size=(n, 3)).astype(f32)
    ext = rng.uniform(0.1, 5, size=(n, 3)).astype(f32)
    return np.concatenate([lo, lo + ext], axis=1).astype(f32)


def _gpu_build(boxes, smin, smax):
    import torch
    from operatorgraph_amd.lbvh import LBVH
    b = LBVH(len(boxes))
    res = b.build(torch.from_numpy(boxes).cuda(), smin, smax)
    out = {k: v.cpu().numpy() for k, v in res.items()}
    b.close()
    return out


def _compare(got, ref, n):
    assert np.array_equal(got["keys"].view(np.uint64), ref["keys"])
    assert np.array_equal(got["left"], ref["left"])
    assert np.array_equal(got["right"], ref["right"])
    assert np.array_equal(got["parent"][1:], ref["parent"][1:]) and got["parent"][0] == -1
    assert np.array_equal(got["leaf_parent"], ref["leaf_parent"])
    assert np.array_equal(got["node_aabb"].view(np.uint32), ref["node_aabb"].view(np.uint32))


@pytest.mark.parametrize("n,clustered", [(2, False), (3, False), (64, False), (1000, True), (70000, False),
                                         (70000, True)])
def test_lbvh_matches_oracle(n, clustered):
    _note(f"lbvh n={n} clustered={clustered}")
    boxes = _boxes(n, n, clustered)
    smin, smax = (0.0, 0.0, 0.0), (100.0, 100.0, 100.0)
    ref = O.lbvh(boxes, smin, smax)
    got = _gpu_build(boxes, smin, smax)
    _compare(got, ref, n)


def _np_instance_boxes(inst, planar):
    m = inst["model"].astype(f32)
    h = [np.abs(m[:, 4 * r]) + np.abs(m[:, 4 * r + 1]) for r in range(3)]
    if not planar:
        h = [h[r] + np.abs(m[:, 4 * r + 2]) for r in range(3)]
    h = [x * f32(0.5) for x in h]
    c = [m[:, 4 * r + 3] for r in range(3)]
    return np.stack([c[0] - h[0], c[1] - h[1], c[2] - h[2], c[0] + h[0], c[1] + h[1], c[2] + h[2]],
                    axis=1).astype(f32)


def test_lbvh_over_derived_terminals():
    import torch
    from operatorgraph_amd import engine
    from operatorgraph_amd.lbvh import LBVH
    _note("lbvh over facade terminals")
    lots = axioms.facade_lots(64)
    walls, windows = axioms.facade_lot_terminals(lots)
    eng = engine.Engine(queue_capacity=int(windows.sum()) + 4096)
    eng.load(grammars.facade()).set_axioms(lots)
    buf = eng.bind(1, int(windows.sum()))
    eng.execute()
    inst = eng.instances(1)
    b = LBVH(len(inst))
    gb = b.instance_aabbs(buf[:len(inst)], planar=True)
    boxes = gb.cpu().numpy()
    assert np.array_equal(boxes.view(np.uint32), _np_instance_boxes(inst, True).view(np.uint32))
    bounds = b.bounds(gb)
    assert np.allclose(bounds[:3], boxes[:, :3].min(0)) and np.allclose(bounds[3:], boxes[:, 3:].max(0))
    smin = [v - 1.0 for v in bounds[:3]]
    smax = [v + 1.0 for v in bounds[3:]]
    ref = O.lbvh(boxes, smin, smax)
    res = b.build(gb, smin, smax)
    got = {k: v.cpu().numpy() for k, v in res.items()}
    _compare(got, ref, len(boxes))
    b.close()
    eng.close()


def test_lbvh_full_size_invariants():
    """4M boxes: sorted unique keys, every internal node the union of its
    children, parent links consistent, root = the union of all boxes."""
    import torch
    from operatorgraph_amd.lbvh import LBVH
    n = 4_000_000
    _note(f"lbvh invariants n={n}")
    g = torch.Generator(device="cuda").manual_seed(5)
    lo = torch.rand((n, 3), device="cuda", generator=g) * 1000.0
    boxes = torch.cat([lo, lo + torch.rand((n, 3), device="cuda", generator=g) * 2.0], dim=1).contiguous()
    b = LBVH(n)
    r = b.build(boxes, (0.0, 0.0, 0.0), (1002.0, 1002.0, 1002.0))
    keys = r["keys"]
    assert bool((keys[1:] > keys[:-1]).all())                      # signed compare: codes < 2^30
    idx = (keys & 0xFFFFFFFF).long()
    assert torch.equal(torch.sort(idx).values, torch.arange(n, device="cuda"))
    left, right, parent = r["left"].long(), r["right"].long(), r["parent"].long()
    node = r["node_aabb"]
    leafbox = boxes[idx]
    def child_box(c):
        return torch.where((c >= 0).unsqueeze(1), leafbox[c.clamp(min=0)], node[(-c).clamp(min=0)])
    lb, rb = child_box(left), child_box(right)
    assert torch.equal(node[:, :3], torch.minimum(lb[:, :3], rb[:, :3]))
    assert torch.equal(node[:, 3:], torch.maximum(lb[:, 3:], rb[:, 3:]))
    internal = torch.arange(n - 1, device="cuda")
    for c in (left, right):
        m = c < 0
        assert torch.equal(parent[(-c[m])], internal[m])
        assert torch.equal(r["leaf_parent"].long()[c[~m]], internal[~m])
    assert torch.equal(node[0, :3], boxes[:, :3].min(0).values)
    assert torch.equal(node[0, 3:], boxes[:, 3:].max(0).values)
    b.close()
