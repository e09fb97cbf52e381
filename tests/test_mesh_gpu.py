"""GPU parity of generation function 0 (triangle soup): grammars whose
terminals use generation function 0 are derived by the HIP engine, their
instance buffers tessellated by pga_tessellate through the C ABI, and the
vertices compared value for value with the oracle's mesh of the same
shapes (canonical instance order; zeros may differ in sign), the indices
with the reference's (0, 1, 2), (0, 2, 3) per face. At full size
(2^20 random boxes) the check is the numpy restatement on a sample plus
the whole index buffer."""
import ctypes as C
import sys

import numpy as np
import pytest

from operatorgraph_amd import abi
from operatorgraph_amd.compiler import compile_grammar
from tests.helpers import assert_same_multiset, canon_order
from tests.test_mesh import BOX_ONE, MENGER0, QUAD_ONE, random_lots

pytestmark = pytest.mark.gpu

FACADE0 = """
axiom Box Lot;
terminal Wall (0, 0);
terminal Window (0, 1);
Lot    = CompSplit() { Discard() | Discard() | Facade };
Facade = Repeat(Y, 3, ADJUST_TO_FILL) Floor;
Floor  = Repeat(X, 2.5, ADJUST_TO_FILL) Tile;
Tile   = IfSizeLess(X, 1.6) { Wall | SubDiv(X) { -1: Wall | 1.2: Window | -1: Wall } };
"""

# (grammar, lots, terminal shape types)
CASES = {
    "box_one": (BOX_ONE, 256, [abi.BOX]),
    "quad_one": (QUAD_ONE, 256, [abi.QUAD]),
    "menger3": (MENGER0, 4, [abi.BOX]),
    "facade": (FACADE0, 24, [abi.QUAD, abi.QUAD]),
}


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def facade_lots(g, n, seed):
    rng = np.random.default_rng(seed)
    sizes = np.stack([rng.uniform(8, 30, n), rng.uniform(9, 40, n), rng.uniform(8, 30, n)], 1)
    pos = np.stack([np.arange(n) * 40.0, np.zeros(n), np.zeros(n)], 1)
    return g.axiom_records(rng.random(n).astype(np.float32), sizes=sizes, positions=pos)


def derived(case, technique):
    """(oracle result, engine) after deriving the case on both."""
    from oracle import oracle as O
    from operatorgraph_amd import engine
    src, n, shapes = CASES[case]
    g = compile_grammar(src, case)
    ax = facade_lots(g, n, 11) if case == "facade" else random_lots(g, n, 11)
    r = O.derive(g.program, ax)
    eng = engine.Engine(technique=technique)
    eng.load(g.program).set_axioms(ax)
    for t in range(len(shapes)):
        eng.bind(t, r.count(t) + 64)
    eng.execute()
    for t in range(len(shapes)):
        assert r.count(t) > 0
        assert_same_multiset(eng.instances(t), r.instances(t), f"{case} terminal {t}")
    return r, eng, shapes


@pytest.mark.parametrize("technique", [abi.KERNELS, abi.MEGAKERNEL])
@pytest.mark.parametrize("case", sorted(CASES))
def test_tessellate_matches_oracle(case, technique):
    import torch

    from oracle.oracle import mesh_from_instances, mesh_indices
    from operatorgraph_amd.mesh import tessellate
    _note(f"tessellate {case}, technique={technique}")
    r, eng, shapes = derived(case, technique)
    for t, shape in enumerate(shapes):
        n = r.count(t)
        V = 4 if shape == abi.QUAD else 24
        verts, idx = tessellate(eng.buffer(t)[:n], shape)
        torch.cuda.synchronize()
        gv = verts.cpu().numpy().reshape(n, V, 9)
        gi = idx.cpu().numpy().view(np.uint32)
        got, ref = eng.instances(t), r.instances(t)
        assert np.array_equal(gv[canon_order(got)], r.mesh(t).reshape(n, V, 9)[canon_order(ref)])
        assert np.array_equal(gv.reshape(-1, 9), mesh_from_instances(got["model"], shape))
        assert np.array_equal(gi, mesh_indices(n, shape))
    eng.close()


def test_two_calls_share_one_index_space(tmp_path):
    """Walls then windows into one vertex/index space (base_vertex), written
    as one OBJ file and read back."""
    import torch

    from oracle.oracle import mesh_indices
    from operatorgraph_amd.mesh import read_obj, tessellate, write_obj
    _note("tessellate facade walls + windows, one index space")
    r, eng, _ = derived("facade", abi.MEGAKERNEL)
    nw, nq = r.count(0), r.count(1)
    v0, i0 = tessellate(eng.buffer(0)[:nw], abi.QUAD)
    v1, i1 = tessellate(eng.buffer(1)[:nq], abi.QUAD, base_vertex=4 * nw)
    verts, idx = torch.cat([v0, v1]), torch.cat([i0, i1])
    torch.cuda.synchronize()
    want = np.concatenate([mesh_indices(nw, abi.QUAD), mesh_indices(nq, abi.QUAD, 4 * nw)])
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), want)
    path = tmp_path / "facade.obj"
    write_obj(str(path), verts, idx.cpu().numpy().view(np.uint32))
    v, vn, vt, tri = read_obj(str(path))
    host = verts.cpu().numpy()
    assert np.array_equal(v, host[:, :3]) and np.array_equal(vn, host[:, 4:7]) and np.array_equal(vt, host[:, 7:9])
    assert np.array_equal(tri.reshape(-1), want.astype(np.int64))
    eng.close()


def test_full_size_random_boxes():
    """2^20 random instance records (1.3 GB of output): a sample of 4096
    shapes equals the numpy restatement value for value, including records
    with a zero column (degenerate faces: NaN normals on both sides); the
    whole index buffer is checked on the GPU."""
    import torch

    from oracle.oracle import mesh_from_instances
    from operatorgraph_amd.mesh import tessellate
    n = 1 << 20
    _note(f"tessellate {n} random boxes")
    g = torch.Generator(device="cuda").manual_seed(5)
    inst = torch.randn((n, 16), device="cuda", generator=g) * 10.0
    inst[::97, 0:12:4] = 0.0                       # zero first column: degenerate faces
    verts, idx = tessellate(inst, abi.BOX)
    torch.cuda.synchronize()
    sample = np.random.default_rng(6).choice(n, 4096, replace=False)
    sample[:8] = np.arange(0, 8 * 97, 97)
    s = torch.from_numpy(sample).cuda()
    gv = verts.view(n, 24, 9)[s].cpu().numpy()
    ref = mesh_from_instances(inst[s, :12].cpu().numpy(), abi.BOX).reshape(-1, 24, 9)
    assert np.array_equal(gv, ref, equal_nan=True)
    deg = sample % 97 == 0
    assert np.isnan(gv[deg]).any(axis=(1, 2)).all() and not np.isnan(gv[~deg]).any()
    j = torch.arange(n * 36, device="cuda", dtype=torch.int64)
    corner = torch.tensor([0, 1, 2, 0, 2, 3], device="cuda")[j % 6]
    want = (j // 36) * 24 + ((j % 36) // 6) * 4 + corner
    assert torch.equal(idx.to(torch.int64) & 0xFFFFFFFF, want)


def test_tessellate_argument_errors():
    import torch

    from operatorgraph_amd.engine import load_library
    from operatorgraph_amd.mesh import tessellate
    _note("tessellate argument errors")
    L = load_library()
    inst = torch.zeros((8, 16), device="cuda")
    verts = torch.empty((8 * 24, 9), device="cuda")
    idx = torch.empty((8 * 36,), dtype=torch.int32, device="cuda")
    p = [C.c_void_p(x.data_ptr()) for x in (inst, verts, idx)]
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.pga_tessellate(p[0], 8, abi.TRIANGLE, p[1], p[2], 0, s) == abi.ERR_UNSUPPORTED
    assert L.pga_tessellate(None, 8, abi.BOX, p[1], p[2], 0, s) == abi.ERR_INVALID_ARG
    assert L.pga_tessellate(C.c_void_p(inst.data_ptr() + 4), 7, abi.BOX, p[1], p[2], 0, s) == abi.ERR_INVALID_ARG
    # 32-bit vertex indices: more than 2^32 vertices is refused before any launch
    assert L.pga_tessellate(p[0], (1 << 32) // 24 + 1, abi.BOX, p[1], p[2], 0, s) == abi.ERR_INVALID_ARG
    assert L.pga_tessellate(p[0], 8, abi.BOX, p[1], p[2], 0xFFFFFFFF - 100, s) == abi.ERR_INVALID_ARG
    assert L.pga_tessellate(None, 0, abi.BOX, None, None, 0, s) == abi.OK
    v, i = tessellate(inst[:0], abi.BOX)
    assert v.shape == (0, 9) and i.shape == (0,)
    with pytest.raises(ValueError):
        tessellate(inst[:, :12], abi.BOX)


def test_vertex_buffer_off_a_16_byte_boundary():
    """A vertex buffer 4 bytes past a 16-byte boundary takes the dword copy
    of the tile; its vertices equal those of the aligned buffer, and a
    ragged last tile (count not a multiple of the tile) stays in bounds."""
    import torch

    from operatorgraph_amd.engine import load_library
    from operatorgraph_amd.mesh import tessellate
    _note("tessellate into an unaligned vertex buffer")
    L = load_library()
    for shape, n in ((abi.BOX, 1000), (abi.QUAD, 1001)):
        V = 4 if shape == abi.QUAD else 24
        inst = torch.randn((n, 16), device="cuda")
        want, want_i = tessellate(inst, shape)
        flat = torch.full((n * V * 9 + 4 + 64,), 7.0, device="cuda")
        idx = torch.empty((n * V * 3 // 2,), dtype=torch.int32, device="cuda")
        s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert L.pga_tessellate(C.c_void_p(inst.data_ptr()), n, shape, C.c_void_p(flat.data_ptr() + 4),
                                C.c_void_p(idx.data_ptr()), 0, s) == abi.OK
        torch.cuda.synchronize()
        assert torch.equal(flat[1:1 + n * V * 9].view(-1, 9), want)
        assert torch.equal(flat[0], torch.tensor(7.0, device="cuda"))
        assert bool((flat[1 + n * V * 9:] == 7.0).all())
        assert torch.equal(idx, want_i)


def test_graph_of_derivation_and_device_count_tessellation():
    """One graph = derivation + tessellation of both terminal types with the
    counts read on the device. The lots change in place between replays,
    which changes the counts; every replay's mesh equals the oracle's, and
    the buffers past the count stay untouched. Eager derivations run
    between the replays."""
    import torch

    from oracle import oracle as O
    from oracle.oracle import mesh_indices
    from operatorgraph_amd import engine
    from operatorgraph_amd.mesh import tessellate_terminal
    _note("graph: derivation + device-count tessellation, lots changed between replays")
    g = compile_grammar(FACADE0, "facade")
    lots = [facade_lots(g, 24, s) for s in (11, 12)]
    refs = [O.derive(g.program, ax) for ax in lots]
    counts = [(r.count(0), r.count(1)) for r in refs]
    assert counts[0] != counts[1]
    cap = [max(c[t] for c in counts) + 16 for t in range(2)]
    dev = torch.from_numpy(lots[0].view(np.uint8).copy()).cuda()
    eng = engine.Engine(technique=abi.MEGAKERNEL)
    eng.load(g.program).set_axioms(dev)
    for t in range(2):
        eng.bind(t, cap[t])
    outs = [(torch.empty((cap[t] * 4, 9), device="cuda"), torch.empty((cap[t] * 6,), dtype=torch.int32, device="cuda"))
            for t in range(2)]
    with eng.capture() as gr:
        eng.execute_async()
        for t in range(2):
            tessellate_terminal(eng, t, abi.QUAD, out=outs[t])
    for rep, k in enumerate((0, 1, 0)):
        dev.copy_(torch.from_numpy(lots[k].view(np.uint8).copy()).cuda())
        for t in range(2):
            outs[t][0].fill_(7.0)
        torch.cuda.synchronize()
        gr.launch()
        eng.synchronize()
        assert tuple(int(x) for x in eng.counts(2)) == counts[k], rep
        for t in range(2):
            n = counts[k][t]
            got, ref = eng.instances(t), refs[k].instances(t)
            assert_same_multiset(got, ref, f"replay {rep} terminal {t}")
            v = outs[t][0][: 4 * n].cpu().numpy().reshape(n, 4, 9)
            assert np.array_equal(v[canon_order(got)], refs[k].mesh(t).reshape(n, 4, 9)[canon_order(ref)]), rep
            assert np.array_equal(outs[t][1][: 6 * n].cpu().numpy().view(np.uint32), mesh_indices(n, abi.QUAD)), rep
            assert bool((outs[t][0][4 * n:] == 7.0).all()), rep
        eng.execute(n=5)                         # eager derivation between replays
    gr.close()
    eng.close()


def test_device_count_is_clamped_to_the_capacity():
    """A device count above the capacity tessellates `capacity` shapes and
    writes nothing past the buffers; a count of zero writes nothing."""
    import torch

    from operatorgraph_amd.engine import load_library
    _note("device-count tessellation: clamping")
    L = load_library()
    cap = 100
    inst = torch.randn((cap, 16), device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    from operatorgraph_amd.mesh import tessellate
    want, want_i = tessellate(inst, abi.BOX)
    for dc, n in ((cap + 100, cap), (0, 0), (37, 37)):
        cnt = torch.tensor([dc], dtype=torch.int64, device="cuda")
        verts = torch.full((cap * 24 + 64, 9), 7.0, device="cuda")
        idx = torch.full((cap * 36 + 64,), -1, dtype=torch.int32, device="cuda")
        assert L.pga_tessellate_device_count(C.c_void_p(inst.data_ptr()), C.c_void_p(cnt.data_ptr()), cap, abi.BOX,
                                             C.c_void_p(verts.data_ptr()), C.c_void_p(idx.data_ptr()), 0, s) == abi.OK
        torch.cuda.synchronize()
        assert torch.equal(verts[: 24 * n], want[: 24 * n]) and bool((verts[24 * n:] == 7.0).all()), dc
        assert torch.equal(idx[: 36 * n], want_i[: 36 * n]) and bool((idx[36 * n:] == -1).all()), dc
    assert L.pga_tessellate_device_count(C.c_void_p(inst.data_ptr()), None, cap, abi.BOX, C.c_void_p(verts.data_ptr()),
                                         C.c_void_p(idx.data_ptr()), 0, s) == abi.ERR_INVALID_ARG


def _instance_shapes(program, inst):
    """Shape type of each instance record: its entry (meta & 0xFFFF) runs a
    Generate procedure of that shape."""
    procs = program.procedures.to_array()
    entries = program.table.entries
    return np.array([procs[entries[e].operator_code].shape_type for e in (inst["meta"] & 0xFFFF)], dtype=np.int32)


@pytest.mark.parametrize("seed", list(range(32)))
def test_fuzz_grammar_triangle_soup(seed):
    """Seeded random grammars (tests/grammar_fuzz.py) with every terminal on
    generation function 0: the GPU derivation's instances, tessellated per
    shape type (a terminal index may collect boxes and quads), give the
    oracle's mesh shape for shape, value for value."""
    import re

    import torch

    from oracle import oracle as O
    from oracle.oracle import mesh_indices
    from operatorgraph_amd import engine
    from operatorgraph_amd.mesh import tessellate
    from tests.grammar_fuzz import lots, random_grammar
    text, shape = random_grammar(seed)
    text = re.sub(r"terminal (\w+) (\w+) \(1, ", r"terminal \1 \2 (0, ", text)
    g = compile_grammar(text, f"fuzz{seed}_soup")
    ax = lots(g, 4, seed, quad=shape == "Quad")
    technique = abi.KERNELS if seed % 4 == 3 else abi.MEGAKERNEL
    _note(f"triangle soup of fuzz grammar seed={seed} shape={shape} technique={technique}")
    try:
        r = O.derive(g.program, ax)
    except O.OracleError:
        return                                   # the error status is checked by test_fuzz_gpu.py
    eng = engine.Engine(technique=technique)
    try:
        eng.load(g.program).set_axioms(ax)
        for t in range(3):
            eng.bind(t, r.count(t) + 64)
        eng.execute()
        for t in range(3):
            got, ref = eng.instances(t), r.instances(t)
            assert_same_multiset(got, ref, f"seed {seed} terminal {t}")
            if not len(ref):
                continue
            sg, sr = _instance_shapes(g.program, got), _instance_shapes(g.program, ref)
            off = np.concatenate([[0], np.cumsum(np.where(sr == abi.QUAD, 4, 24))])
            mesh = r.mesh(t)
            assert mesh.shape[0] == off[-1]
            blocks = [None] * len(got)
            for st in (abi.BOX, abi.QUAD):
                sel = np.nonzero(sg == st)[0]
                if not len(sel):
                    continue
                V = 4 if st == abi.QUAD else 24
                tens = torch.from_numpy(np.ascontiguousarray(got[sel]).view(np.float32).reshape(-1, 16).copy()).cuda()
                v, i = tessellate(tens, st)
                vv = v.cpu().numpy()
                for k, gi in enumerate(sel):
                    blocks[gi] = vv[k * V:(k + 1) * V]
                assert np.array_equal(i.cpu().numpy().view(np.uint32), mesh_indices(len(sel), st))
            for a, b in zip(canon_order(got), canon_order(ref)):
                assert np.array_equal(blocks[a], mesh[off[b]:off[b + 1]]), (seed, t, a, b)
    finally:
        eng.close()
