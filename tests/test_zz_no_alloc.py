"""The allocation-free LBVH build (pga_lbvh_build_noalloc, the library's
per-device scratch array) against the builder's build (pga_lbvh_build) of
the same boxes: keys, topology and node boxes bit-equal, on two calls in a
row with the second smaller (the array keeps the first call's contents),
and a request beyond the array's size refused with a message."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

f32 = np.float32


def _boxes(n, seed):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(0, 90, size=(n, 3)).astype(f32)
    lo[: n // 4] = lo[0]                      # shared Morton cells: ties resolved by the index
    ext = rng.uniform(0.1, 5, size=(n, 3)).astype(f32)
    return np.concatenate([lo, lo + ext], axis=1).astype(f32)


def _outputs(torch, n):
    dev = "cuda"
    return dict(keys=torch.full((n,), -7, dtype=torch.int64, device=dev),
                left=torch.full((n - 1,), -7, dtype=torch.int32, device=dev),
                right=torch.full((n - 1,), -7, dtype=torch.int32, device=dev),
                parent=torch.full((n - 1,), -7, dtype=torch.int32, device=dev),
                leaf_parent=torch.full((n,), -7, dtype=torch.int32, device=dev),
                node_aabb=torch.full((n - 1, 6), -7.0, dtype=torch.float32, device=dev))


def test_noalloc_build_equals_builder_build():
    import torch
    from operatorgraph_amd import lbvh
    from operatorgraph_amd.engine import PGAError
    smin, smax = (0.0, 0.0, 0.0), (100.0, 100.0, 100.0)
    for n, seed in ((200_000, 1), (5_000, 2)):    # the second call smaller than the first
        print(f"noalloc lbvh n={n}", flush=True)
        boxes = torch.from_numpy(_boxes(n, seed)).cuda()
        b = lbvh.LBVH(n)
        ref = b.build(boxes, smin, smax)
        b.close()
        out = _outputs(torch, n)
        lbvh.build_noalloc(boxes, smin, smax, out)
        torch.cuda.synchronize()
        for k in ("keys", "left", "right", "leaf_parent"):
            assert torch.equal(out[k], ref[k]), k
        assert torch.equal(out["parent"][1:], ref["parent"][1:])
        assert torch.equal(out["node_aabb"].view(torch.int32), ref["node_aabb"].view(torch.int32))
    big = lbvh.NOALLOC_MAX_LEAVES + 1
    boxes = torch.zeros((big, 6), dtype=torch.float32, device="cuda")
    with pytest.raises(PGAError) as e:
        lbvh.build_noalloc(boxes, smin, smax, _outputs(torch, 2))
    assert "1048576" in str(e.value)
