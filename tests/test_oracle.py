"""Pins the CPU oracle (oracle/pga_oracle.cpp) against published vectors,
closed forms, textbook geometry and brute force, so that a dropped term, a
wrong sign or a transposed operand in the restatement fails here."""
import math
import struct

import numpy as np
import pytest

from operatorgraph_amd import abi, axioms, grammars
from operatorgraph_amd.grammars import op
from operatorgraph_amd.philox import philox4x32_10, uniform24
from operatorgraph_amd.program import (DispatchTable, Entry, Exp, Parameter, Procedure, ProcedureList,
                                       Program, Rand, Scalar, ShapeAttr, pack_float2, pack_operand)
from oracle import oracle as O
from tests.helpers import canon, rel_err

B, Q = abi.BOX, abi.QUAD
f32 = np.float32


# ---------------------------------------------------------------- Philox
# Known-answer vectors of Philox4x32-10 published with the Random123
# reference implementation (Salmon et al., SC'11): tests/golden.
def _load_kat():
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "philox4x32_10_kat.json")
    with open(path) as f:
        vec = json.load(f)["vectors"]
    h = lambda xs: tuple(int(x, 16) for x in xs)
    return [(h(v["ctr"]), h(v["key"]), h(v["out"])) for v in vec]


PHILOX_KAT = _load_kat()


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    assert tuple(O.philox(ctr, key)) == expect
    assert tuple(philox4x32_10(np.array([ctr], dtype=np.uint32), key)[0]) == expect


def test_philox_numpy_matches_oracle():
    rng = np.random.default_rng(7)
    ctr = rng.integers(0, 2 ** 32, size=(64, 4), dtype=np.uint64).astype(np.uint32)
    key = (0x5EED0003, 0x1234)
    ref = philox4x32_10(ctr, key)
    for i in range(64):
        assert tuple(O.philox(tuple(int(x) for x in ctr[i]), key)) == tuple(int(x) for x in ref[i])


# ---------------------------------------------------------------- seeds
# hybridTaus (Random.cuh:15-49) XORs three Tausworthe generators and one LCG.
# Each is pinned by what the mathematics fixes about it, using the oracle's
# own component functions as black boxes:
#  * the Tausworthe steps are L'Ecuyer's taus88 components (Math. Comp. 65,
#    1996): GF(2)-linear maps whose state is the top k bits (k = 31, 29, 28),
#    each of maximal order 2^k - 1; a wrong shift, mask or operand order
#    breaks linearity, the state projection or the period;
#  * the LCG z -> 1664525 z + 1013904223 has full period 2^32 (Hull-Dobell:
#    odd increment, multiplier = 1 mod 4);
#  * the final conversion multiplies by 2^-32 (rounded up in the 13th digit)
#    in double and converts to float, so values below 2^24 map exactly and
#    the largest round to 1.0f.
# What no reference-held vector fixes (no run of the reference is possible
# here): the additive offsets of the state chain (Random.cuh:26-35) and the
# per-operator modifiers; they are restated from the source text.
_M32 = 0xFFFFFFFF


def _prime_factors(n):
    out, d = [], 2
    while d * d <= n:
        while n % d == 0:
            out.append(d)
            n //= d
        d += 1
    if n > 1:
        out.append(n)
    return sorted(set(out))


def _gf2_mul(a_cols, b_cols):
    """Columns of A.B over GF(2); a column is an int bitmask."""
    out = []
    for col in b_cols:
        acc, j = 0, 0
        while col:
            if col & 1:
                acc ^= a_cols[j]
            col >>= 1
            j += 1
        out.append(acc)
    return out


def _gf2_pow(cols, e):
    k = len(cols)
    res = [1 << j for j in range(k)]
    base = cols
    while e:
        if e & 1:
            res = _gf2_mul(base, res)
        base = _gf2_mul(base, base)
        e >>= 1
    return res


@pytest.mark.parametrize("which,k", [(0, 31), (1, 29), (2, 28)])
def test_tausworthe_component_is_a_maximal_period_gf2_map(which, k):
    comp = lambda z: O.seed_component(z, which)
    rng = np.random.default_rng(11 + which)
    # linear over GF(2)
    assert comp(0) == 0
    for a, b in rng.integers(0, 2 ** 32, size=(200, 2), dtype=np.uint64):
        a, b = int(a), int(b)
        assert comp(a ^ b) == comp(a) ^ comp(b)
    # the state is the top k bits: the bits below never reach them
    shift = 32 - k
    for i in range(shift):
        assert comp(1 << i) >> shift == 0, i
    # the induced k x k map has order exactly 2^k - 1
    cols = [comp(1 << (j + shift)) >> shift for j in range(k)]
    ident = [1 << j for j in range(k)]
    n = 2 ** k - 1
    assert _gf2_pow(cols, n) == ident
    for p in _prime_factors(n):
        assert _gf2_pow(cols, n // p) != ident, p


def test_lcg_component_has_full_period():
    f = lambda z: O.seed_component(z, 3)
    c = f(0)
    a = (f(1) - c) & _M32
    rng = np.random.default_rng(5)
    for z in rng.integers(0, 2 ** 32, size=500, dtype=np.uint64):
        assert f(int(z)) == (a * int(z) + c) & _M32     # affine mod 2^32
    assert c % 2 == 1 and a % 4 == 1                    # Hull-Dobell for m = 2^32


def test_seed_unit_conversion():
    u = O.seed_unit
    assert u(0) == 0.0
    assert u(0xFFFFFFFF) == 1.0                         # can round to 1.0f
    for t in [1, 2, 3, 1000, 2 ** 23 + 5, 2 ** 24 - 1]:
        assert u(t) == t * 2.0 ** -32                   # exact below 2^24
    rng = np.random.default_rng(3)
    ts = sorted(int(t) for t in rng.integers(0, 2 ** 32, size=3000, dtype=np.uint64))
    vals = [u(t) for t in ts]
    assert all(x <= y for x, y in zip(vals, vals[1:]))  # monotone
    for t, v in zip(ts, vals):
        exact = t * 2.0 ** -32
        ulp = float(np.spacing(np.float32(max(exact, 2.0 ** -126))))
        assert abs(v - exact) <= 0.5 * ulp * (1 + 1e-9) + exact * 1e-12


def test_seed_chain_composes_the_pinned_components():
    for m in list(range(0, 40)) + [1000, 65535, 2 ** 31, 2 ** 32 - 1]:
        for base in (0.0, 0.25, 0.999, 1.0, 1.5, 7.0):
            z = O.seed_state(base, m)
            t = O.seed_component(z[0], 0) ^ O.seed_component(z[1], 1) ^ \
                O.seed_component(z[2], 2) ^ O.seed_component(z[3], 3)
            assert O.next_seed(base, m) == O.seed_unit(t), (base, m)
            # the state depends on the base through its integer part only
            assert z == O.seed_state(float(math.floor(base)), m)


def test_seed_chain_properties():
    # for bases in [0, 1) the child seed depends on the modifier only
    for m in range(50):
        vals = {O.next_seed(b, m) for b in (0.0, 0.1, 0.5, 0.99999)}
        assert len(vals) == 1
    u = np.array([O.next_seed(0.3, m) for m in range(20000)])
    assert np.all((u >= 0.0) & (u <= 1.0))
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1.0 / 12.0) < 0.01
    assert O.range_(1.0, 2.0, 5.0) == 5.0 and O.range_(0.5, 2.0, 6.0) == 4.0


# ---------------------------------------------------------------- sin / cos
def _ulp_err(a, b):
    a32 = np.float32(a)
    return abs(float(a32) - b) / float(np.spacing(np.float32(abs(b)) if b != 0 else np.float32(1e-30)))


def test_sincos_against_libm():
    xs = np.concatenate([np.linspace(-20, 20, 20001, dtype=np.float32),
                         np.float32(np.deg2rad(np.arange(-720, 721, 15)))])
    worst = 0.0
    for x in xs:
        s, c = O.sinf(float(x)), O.cosf(float(x))
        assert abs(s - math.sin(float(x))) <= 2.5e-7 and abs(c - math.cos(float(x))) <= 2.5e-7
        if abs(math.sin(float(x))) > 1e-3:
            worst = max(worst, _ulp_err(s, math.sin(float(x))))
    assert worst <= 2.0


# ---------------------------------------------------------------- packing
def test_pack_operand_matches_packutils():
    for t in range(4):
        for x in (0.0, 1.0, 2.5, 16.0, 0.111, 3.0e5):
            a = O.pack_operand(t, x)
            assert struct.pack("<f", a) == struct.pack("<f", pack_operand(t, x))
            bits = struct.unpack("<I", struct.pack("<f", a))[0]
            assert bits >> 30 == t
            back = struct.unpack("<f", struct.pack("<I", (bits << 2) & 0xFFFFFFFF))[0]
            assert back == float(np.float32(x)) or abs(back - x) <= abs(x) * 2.0 ** -21
    for x, y in ((0.0, 0.0), (1.0, 1.0), (0.25, 0.75)):
        assert O.pack_float2(x, y) == pack_float2(x, y)


# ---------------------------------------------------------------- Menger
def _menger_cells(depth):
    """Closed form: a cell (i, j, k) of the 3^d lattice is kept iff no base-3
    digit position has two or more of its three digits equal to 1."""
    n = 3 ** depth
    cells = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                a, b, c, ok = i, j, k, True
                for _ in range(depth):
                    if int(a % 3 == 1) + int(b % 3 == 1) + int(c % 3 == 1) >= 2:
                        ok = False
                        break
                    a, b, c = a // 3, b // 3, c // 3
                if ok:
                    cells.append((i, j, k))
    return np.array(cells)


@pytest.mark.parametrize("depth,threshold", [(1, 0.5), (2, 0.2), (3, 0.111)])
def test_menger_closed_form(depth, threshold):
    r = O.derive(grammars.menger_interpreter(threshold), axioms.menger_axiom(1.0, 0), twins=O.TWIN_F64)
    assert r.count(0) == 20 ** depth
    inst = r.instances(0)
    m = inst["model"]
    # size: FP32 chain s_{k+1} = s_k / 3 (r2a = (s - 0) / 3, times |-1|)
    s = f32(1.0)
    for _ in range(depth):
        s = f32(s / f32(3.0))
    assert np.all(m[:, [0, 5, 10]] == s)
    assert np.all(m[:, [1, 2, 4, 6, 8, 9]] == 0.0)
    # centres on the lattice (i + 1/2) / 3^d - 1/2, exactly the kept cells
    n = 3 ** depth
    cells = np.rint((m[:, [3, 7, 11]].astype(np.float64) + 0.5) * n - 0.5).astype(int)
    assert np.max(np.abs(m[:, [3, 7, 11]] - ((cells + 0.5) / n - 0.5))) < 1e-6
    got = {tuple(c) for c in cells}
    assert got == {tuple(c) for c in _menger_cells(depth)}
    assert len(got) == 20 ** depth
    # FP64 twin: the exact lattice
    d = r.instances_f64(0)
    assert np.max(np.abs(d[:, [3, 7, 11]] - ((cells + 0.5) / n - 0.5))) < 1e-12
    assert rel_err(m, d[:, :12]) < 1e-5


def test_menger_static_and_dyn_forms_agree():
    """The test-runner table (static Subdivide form) and the interpreter
    table (DynParams form) derive the same instances for scalar factors."""
    a = axioms.menger_axiom(9.0, grammars.MENGER_TEST_RUNNER_AXIOM_ENTRY, centred=False)
    r1 = O.derive(grammars.menger_test_runner(), a)
    assert r1.count(0) == 20 ** 3   # 9 -> 3 -> 1 -> 1/3 < 1
    b = a.copy()
    b["entry_index"] = 0
    r2 = O.derive(grammars.menger_interpreter(1.0), b)
    i1, i2 = r1.instances(0), r2.instances(0)
    assert np.array_equal(np.sort(i1["model"].view(np.uint32), axis=0), np.sort(i2["model"].view(np.uint32), axis=0))


# ---------------------------------------------------------------- facade
def test_facade_closed_form_counts_and_geometry():
    lots = axioms.facade_lots(40)
    r = O.derive(grammars.facade(), lots, twins=O.TWIN_F64)
    walls, windows = axioms.facade_lot_terminals(lots)
    assert r.count(0) == walls.sum() and r.count(1) == windows.sum()
    win = r.instances(1)
    per_lot = np.bincount(win["axiom"], minlength=40)
    assert np.array_equal(per_lot, windows)
    # every window is 1.2 wide (absolute factor) and one floor (3 m) high:
    # column norms of model * scale(size)
    m = win["model"].reshape(-1, 3, 4).astype(np.float64)
    assert np.allclose(np.linalg.norm(m[:, :, 0], axis=1), 1.2, atol=1e-5)
    assert np.allclose(np.linalg.norm(m[:, :, 1], axis=1), 3.0, atol=1e-5)
    # windows lie on the lot's side faces
    a = lots[win["axiom"]]
    centre = m[:, :, 3] - a["model"][:, [3, 7, 11]]
    half = a["size"] / 2.0
    on_x = np.isclose(np.abs(centre[:, 0]), half[:, 0], atol=1e-3)
    on_z = np.isclose(np.abs(centre[:, 2]), half[:, 2], atol=1e-3)
    assert np.all(on_x | on_z)
    assert rel_err(win["model"], r.instances_f64(1)[:, :12]) < 1e-5


# ---------------------------------------------------------------- small grammars
def _prog(entries, procs=None):
    return Program(procs or grammars.INTERPRETER_PROCEDURES, DispatchTable(entries))


def _box(size=(1, 1, 1), pos=(0, 0, 0), entry=0, seed=0.5, shape=B):
    a = axioms.empty(1)
    a["size"] = size
    a[0]["model"][[3, 7, 11]] = pos
    a["seed"] = seed
    a["custom"] = seed
    a["entry_index"] = entry
    a["shape_type"] = shape
    return a


def test_compsplit_faces_textbook():
    """ComponentSplit of a box (2,3,4): side, bottom and top quads centred on
    the faces with outward normals (the model's z column)."""
    gen = lambda t: Entry(op(abi.GENERATE, Q, 1), [Scalar(t)], [])
    p = _prog([Entry(op(abi.COMPSPLIT, B), [], [(1, 0), (2, 0), (3, 0)]), gen(0), gen(1), gen(2)])
    r = O.derive(p, _box((2, 3, 4)))
    sides = r.instances(2)["model"].reshape(-1, 3, 4)
    assert len(sides) == 4
    normals = sides[:, :, 2]   # size z = 1, so column 2 is the unit normal
    centres = sides[:, :, 3]
    expect = {(0, 0, -1): (0, 0, -2), (0, 0, 1): (0, 0, 2), (-1, 0, 0): (-1, 0, 0), (1, 0, 0): (1, 0, 0)}
    got = {tuple(np.rint(n).astype(int)): tuple(c) for n, c in zip(normals, centres)}
    assert set(got) == set(expect)
    for n, c in expect.items():
        assert np.allclose(got[n], c)
    widths = sorted(float(np.linalg.norm(s[:, 0])) for s in sides)
    assert widths == [2.0, 2.0, 4.0, 4.0]
    bottom = r.instances(1)["model"].reshape(3, 4)
    top = r.instances(0)["model"].reshape(3, 4)
    assert np.allclose(bottom[:, 3], (0, -1.5, 0)) and np.allclose(bottom[:, 2], (0, -1, 0))
    assert np.allclose(top[:, 3], (0, 1.5, 0)) and np.allclose(top[:, 2], (0, 1, 0))
    # right-handed frames: det(R) = +1 on every face
    for s in list(sides) + [bottom, top]:
        R = s[:, :3] / np.linalg.norm(s[:, :3], axis=0)
        assert np.isclose(np.linalg.det(R), 1.0, atol=1e-6)


def _compsplit_prog(parallel, term_of_face=(0, 1, 2)):
    procs = ProcedureList([Procedure(abi.COMPSPLIT, B, 0, abi.PROC_PARALLEL if parallel else 0),
                           Procedure(abi.GENERATE, Q, 1)])
    gen = lambda t: Entry(1, [Scalar(t)], [])
    return Program(procs, DispatchTable([Entry(0, [], [(1, 0), (2, 0), (3, 0)])] +
                                        [gen(t) for t in term_of_face]))


def test_parallel_compsplit_faces_textbook():
    """ComponentSplit<true> (ParallelFaceSplitter<Box>, ComponentSplit.cuh:
    179-230): the same six faces as the serial splitter - centred on the box
    faces, outward normals, right-handed frames, the serial form's sizes and
    models up to the sign of zero - with the caps' seed modifiers swapped:
    (f + 2) % 6 gives bottom 0 and top 1, where the serial splitter gives
    bottom 1 and top 0 (ComponentSplit.cuh:132, 169)."""
    size = (2.0, 3.0, 4.0)
    a = _box(size, pos=(5.0, -1.0, 2.0), seed=0.5)
    rp, rs = O.derive(_compsplit_prog(True), a), O.derive(_compsplit_prog(False), a)
    for t in range(3):
        ip, is_ = rp.instances(t), rs.instances(t)
        kp, ks = np.argsort(ip["path"]), np.argsort(is_["path"])
        assert np.array_equal(ip["path"][kp], is_["path"][ks])      # same faces (ordinals)
        assert np.array_equal(ip["model"][kp], is_["model"][ks])    # equal values (+0 == -0)
    top, bottom = rp.instances(0)["model"].reshape(3, 4), rp.instances(1)["model"].reshape(3, 4)
    assert np.allclose(bottom[:, 3], (5, -2.5, 2)) and np.allclose(bottom[:, 2], (0, -1, 0))
    assert np.allclose(top[:, 3], (5, 0.5, 2)) and np.allclose(top[:, 2], (0, 1, 0))
    sides = rp.instances(2)["model"].reshape(-1, 3, 4)
    centres = {tuple(np.rint(s[:, 2]).astype(int)): tuple(s[:, 3]) for s in sides}
    assert centres == {(0, 0, -1): (5, -1, 0), (0, 0, 1): (5, -1, 4), (-1, 0, 0): (4, -1, 2),
                       (1, 0, 0): (6, -1, 2)}
    for s in list(sides) + [bottom, top]:
        R = s[:, :3] / np.linalg.norm(s[:, :3], axis=0)
        assert np.isclose(np.linalg.det(R), 1.0, atol=1e-6)
    # the seeds as custom attribute would not show them: route each face's
    # seed into a Translate by ShapeSeed and read it back from the instance
    seeds = {}
    for par in (True, False):
        procs = ProcedureList([Procedure(abi.COMPSPLIT, B, 0, abi.PROC_PARALLEL if par else 0),
                               Procedure(abi.TRANSLATE, Q), Procedure(abi.GENERATE, Q, 1)])
        tr = lambda nxt: Entry(1, [Scalar(0), Scalar(0), ShapeAttr(15)], [(nxt, 0)])
        p = Program(procs, DispatchTable([Entry(0, [], [(1, 0), (2, 0), (3, 0)]), tr(4), tr(5), tr(6),
                                          Entry(2, [Scalar(0)], []), Entry(2, [Scalar(1)], []),
                                          Entry(2, [Scalar(2)], [])]))
        r = O.derive(p, _box((1, 1, 1), seed=0.5))
        # translate z by the seed in the face frame: the z column is the normal
        seeds[par] = [r.instances(t)["model"].reshape(-1, 3, 4) for t in range(3)]
    def lifted(m):   # seed = displacement along the normal minus the half size (0.5)
        return np.round(np.einsum("ij,ij->i", m[:, :, 3], m[:, :, 2]) - 0.5, 6)
    top_p, bot_p = lifted(seeds[True][0])[0], lifted(seeds[True][1])[0]
    top_s, bot_s = lifted(seeds[False][0])[0], lifted(seeds[False][1])[0]
    assert np.isclose(top_p, O.next_seed(0.5, 1), atol=1e-6) and np.isclose(bot_p, O.next_seed(0.5, 0), atol=1e-6)
    assert np.isclose(top_s, O.next_seed(0.5, 0), atol=1e-6) and np.isclose(bot_s, O.next_seed(0.5, 1), atol=1e-6)
    assert sorted(lifted(seeds[True][2])) == sorted(lifted(seeds[False][2]))   # sides: f + 2 in both


def test_parallel_repeat_is_refused():
    """Repeat<true> traps in the reference (Repeat.cuh:35-61)."""
    procs = ProcedureList([Procedure(abi.REPEAT, B, 0, abi.PROC_PARALLEL), Procedure(abi.GENERATE, B, 1)])
    p = Program(procs, DispatchTable([Entry(0, [Scalar(0), Scalar(0.3), Scalar(abi.ADJUST_TO_FILL)], [(1, 0)]),
                                      Entry(1, [Scalar(0)], [])]))
    with pytest.raises(O.OracleError) as e:
        O.derive(p, _box())
    assert e.value.status == abi.ERR_UNSUPPORTED


@pytest.mark.parametrize("mode,total,extent", [
    (abi.ADJUST_TO_FILL, 10.0, 3.0), (abi.ANCHOR_TO_START, 10.0, 3.0), (abi.ANCHOR_TO_END, 10.0, 3.0),
    (abi.ANCHOR_TO_START, 9.0, 3.0), (abi.ADJUST_TO_FILL, 2.0, 3.0)])
def test_repeat_closed_form(mode, total, extent):
    p = _prog([Entry(op(abi.REPEAT, B), [Scalar(abi.X), Scalar(extent), Scalar(mode)], [(1, 0)]),
               Entry(op(abi.GENERATE, B, 1), [Scalar(0)], [])])
    r = O.derive(p, _box((total, 1, 1)))
    m = r.instances(0)["model"].reshape(-1, 3, 4)
    widths = m[:, 0, 0]
    xs = m[:, 0, 3]
    order = np.argsort(xs)
    widths, xs = widths[order], xs[order]
    if mode == abi.ADJUST_TO_FILL:
        n = math.floor(total / extent)
        expect = [total / n] * n if n else [0.0]   # n == 0 still emits the first piece
    else:
        n = math.ceil(total / extent)
        rest = math.fmod(total, extent)
        pieces = [extent] * n
        if rest:
            if mode == abi.ANCHOR_TO_START:
                pieces[0] = rest
            else:
                pieces[-1] = rest
        expect = pieces
    assert np.allclose(widths, expect, atol=1e-6)
    # pieces tile [-total/2, total/2] left to right
    edges = np.concatenate([[-total / 2.0], np.cumsum(expect) - total / 2.0])
    if len(expect) > 1 or expect[0] > 0:
        assert np.allclose(xs, (edges[:-1] + edges[1:]) / 2.0, atol=1e-5)


def test_subdivide_forms_zero_factor_and_overconstrained():
    gen = Entry(op(abi.GENERATE, B, 1), [Scalar(0)], [])
    dyn = _prog([Entry(op(abi.SUBDIV, B), [Scalar(0), Scalar(-1), Scalar(0), Scalar(-1)],
                       [(1, 0), (1, 0), (1, 0)]), gen])
    r = O.derive(dyn, _box((4, 1, 1)))
    assert r.count(0) == 3   # DynParams form dispatches the zero-size child
    procs = ProcedureList([Procedure(abi.SUBDIV, B, 0, abi.PROC_STATIC), Procedure(abi.GENERATE, B, 1)])
    st = Program(procs, DispatchTable([Entry(0, [Scalar(0), Scalar(-1), Scalar(0), Scalar(-1)],
                                              [(1, 0), (1, 0), (1, 0)]), Entry(1, [Scalar(0)], [])]))
    r = O.derive(st, _box((4, 1, 1)))
    assert r.count(0) == 2   # static form skips it
    assert sorted(r.instances(0)["model"][:, 0]) == [2.0, 2.0]
    # absolute factors alone never trap (relative sum 0 => scale 0)
    assert O.derive(_prog([Entry(op(abi.SUBDIV, B), [Scalar(0), Scalar(3), Scalar(2)], [(1, 0), (1, 0)]), gen]),
                    _box((4, 1, 1))).count(0) == 2
    over = _prog([Entry(op(abi.SUBDIV, B), [Scalar(0), Scalar(3), Scalar(-1)], [(1, 0), (1, 0)]), gen])
    with pytest.raises(O.OracleError) as e:
        O.derive(over, _box((2, 1, 1)))
    assert e.value.status == abi.ERR_SUBDIV_OVERCONSTRAINED
    bad = _prog([Entry(op(abi.REPEAT, B), [Scalar(0), Scalar(0.0), Scalar(0)], [(1, 0)]), gen])
    with pytest.raises(O.OracleError) as e:
        O.derive(bad, _box((4, 1, 1)))
    assert e.value.status == abi.ERR_REPEAT_BAD_EXTENT


def test_transform_operators_textbook():
    """Translate then Rotate(0, 90, 0) then Scale(2, -0.5, -1): the instance
    equals T * R_y(90) * diag(scaled size) from textbook matrices."""
    p = _prog([
        Entry(op(abi.TRANSLATE, B), [Scalar(1), Scalar(2), Scalar(3)], [(1, 0)]),
        Entry(op(abi.ROTATE, B), [Scalar(0), Scalar(90), Scalar(0)], [(2, 0)]),
        Entry(op(abi.SCALE, B), [Scalar(2), Scalar(-0.5), Scalar(-1)], [(3, 0)]),
        Entry(op(abi.SWAPSIZE, B), [Scalar(1), Scalar(0), Scalar(2)], [(4, 0)]),
        Entry(op(abi.GENERATE, B, 1), [Scalar(0)], []),
    ])
    r = O.derive(p, _box((4, 6, 8)), twins=O.TWIN_F64)
    m = r.instances(0)["model"].reshape(3, 4).astype(np.float64)
    Ry = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], dtype=np.float64)
    size = np.array([3.0, 2.0, 8.0])          # scale -> (2, 3, 8), swap x/y -> (3, 2, 8)
    assert np.allclose(m[:, :3], Ry @ np.diag(size), atol=1e-6)
    assert np.allclose(m[:, 3], (1, 2, 3))
    assert rel_err(r.instances(0)["model"], r.instances_f64(0)[:, :12]) < 1e-5


def test_extrude_and_replicate():
    p = _prog([
        Entry(op(abi.COMPSPLIT, B), [], [(1, 0), (2, 0), (2, 0)]),
        Entry(op(abi.EXTRUDE, Q), [Scalar(0), Scalar(0.5)], [(3, 0)]),
        Entry(op(abi.DISCARD, Q), [], []),
        Entry(op(abi.REPLICATE, B), [], [(4, 0), (4, 0), (5, 0)]),
        Entry(op(abi.GENERATE, B, 1), [Scalar(0)], []),
        Entry(op(abi.GENERATE, B, 1), [Scalar(1)], []),
    ])
    r = O.derive(p, _box((2, 2, 2)))
    assert r.count(0) == 2 and r.count(1) == 1
    m = r.instances(1)["model"].reshape(3, 4)
    # top face extruded upwards: a 2 x 0.5 x 2 box whose base sits on y = 1
    assert np.allclose(np.abs(np.linalg.norm(m[:, :3], axis=0)), (2, 0.5, 2))
    assert np.allclose(m[:, 3], (0, 1, 0))


def test_expression_and_if():
    """If(Exp(ShapeSize(X) > 2)) and Translate(Exp(ShapeSize(X) * -0.5), ...)."""
    e_gt = Exp(("a", 12), ("s", 2.0), ("op", abi.OPT_GT))
    e_half = Exp(("a", 12), ("s", -0.5), ("op", abi.OPT_MULTI))
    p = _prog([
        Entry(op(abi.IF, B), [e_gt], [(1, 0), (2, 0)]),
        Entry(op(abi.TRANSLATE, B), [e_half, Scalar(0), Scalar(0)], [(2, 0)]),
        Entry(op(abi.GENERATE, B, 1), [Scalar(0)], []),
    ])
    big = O.derive(p, _box((6, 1, 1))).instances(0)["model"]
    small = O.derive(p, _box((1, 1, 1))).instances(0)["model"]
    assert big[0][3] == -3.0 and small[0][3] == 0.0


def test_random_rule_frequencies():
    """RandomRule over the seed chain: branch frequencies follow the weights."""
    gen = lambda t: Entry(op(abi.GENERATE, B, 1), [Scalar(t)], [])
    p = _prog([
        Entry(op(abi.REPEAT, B), [Scalar(0), Scalar(1.0), Scalar(abi.ADJUST_TO_FILL)], [(1, 0)]),
        Entry(op(abi.STOCHASTIC, B), [ShapeAttr(15), Scalar(1), Scalar(3)], [(2, 0), (3, 0)]),
        gen(0), gen(1),
    ])
    r = O.derive(p, _box((4000, 1, 1)))
    n0, n1 = r.count(0), r.count(1)
    assert n0 + n1 == 4000
    assert abs(n0 / 4000 - 0.25) < 0.03


# ---------------------------------------------------------------- LBVH
def _brute_radix_tree(keys):
    """Top-down binary radix tree over sorted unique keys: node id of a range
    [first, last] is `first` when it is a right child's... numbering of
    Karras 2012: internal node gamma covers a range with one end at gamma."""
    n = len(keys)
    left = np.zeros(n - 1, dtype=np.int64)
    right = np.zeros(n - 1, dtype=np.int64)
    stack = [(0, n - 1, 0)]
    while stack:
        first, last, node = stack.pop()
        common = 64 - int(keys[first] ^ keys[last]).bit_length()
        split = first
        step = last - first
        # largest index whose prefix with `first` exceeds `common`
        for i in range(first, last):
            if 64 - int(keys[first] ^ keys[i + 1]).bit_length() > common:
                split = i + 1
        lc, rc = split, split + 1
        left[node] = lc if lc == first else -lc
        right[node] = rc if rc == last else -rc
        if lc != first:
            stack.append((first, lc, lc))
        if rc != last:
            stack.append((rc, last, rc))
    return left, right


@pytest.mark.parametrize("n", [2, 3, 17, 1000])
def test_lbvh_matches_brute_force(n):
    rng = np.random.default_rng(n)
    lo = rng.uniform(0, 90, size=(n, 3)).astype(np.float32)
    ext = rng.uniform(0.1, 5, size=(n, 3)).astype(np.float32)
    aabb = np.concatenate([lo, lo + ext], axis=1)
    res = O.lbvh(aabb, (0, 0, 0), (100, 100, 100))
    keys = res["keys"]
    assert np.all(np.diff(keys.astype(np.float64)) > 0)
    left, right = _brute_radix_tree([int(k) for k in keys])
    assert np.array_equal(res["left"], left) and np.array_equal(res["right"], right)
    # parents consistent, root covers the union of all boxes
    for i in range(n - 1):
        for ch in (res["left"][i], res["right"][i]):
            if ch < 0:
                assert res["parent"][-ch] == i
            else:
                assert res["leaf_parent"][ch] == i
    assert np.allclose(res["node_aabb"][0, :3], aabb[:, :3].min(0))
    assert np.allclose(res["node_aabb"][0, 3:], aabb[:, 3:].max(0))
    # refit: every node's box is the min / max over the boxes of the leaves
    # below it (collected from the brute-force topology)
    order = (keys & 0xFFFFFFFF).astype(np.int64)
    below = {}

    def leaves(node):
        if node not in below:
            out = []
            for ch in (left[node], right[node]):
                out += [int(ch)] if ch >= 0 else leaves(int(-ch))
            below[node] = out
        return below[node]

    for i in range(n - 1):
        b = aabb[order[leaves(i)]]
        assert np.array_equal(res["node_aabb"][i, :3], b[:, :3].min(0)), i
        assert np.array_equal(res["node_aabb"][i, 3:], b[:, 3:].max(0)), i
    # Morton code of the first leaf from its centre
    c = (aabb[keys[0] & 0xFFFFFFFF, :3] + aabb[keys[0] & 0xFFFFFFFF, 3:]) * f32(0.5) / f32(100)
    assert int(keys[0]) >> 32 == O.morton30(*[float(v) for v in c])


def test_morton_bit_interleave():
    # x bits land on positions 0,3,6..; y on 1,4,..; z on 2,5,..
    assert O.morton30(1.0, 0.0, 0.0) == int("".join("001" for _ in range(10)), 2)
    assert O.morton30(0.0, 1.0, 0.0) == int("".join("010" for _ in range(10)), 2)
    assert O.morton30(0.0, 0.0, 1.0) == int("".join("100" for _ in range(10)), 2)


# ---------------------------------------------------------------- stochastic
def test_stochastic_grammar_philox_choices():
    ax, depth = axioms.stochastic_axioms(200)
    r = O.derive(grammars.stochastic(), ax, key=axioms.STOCHASTIC_KEY)
    inst = r.instances(0)
    per = np.bincount(inst["axiom"], minlength=200)
    assert np.all(per >= 1) and np.all(per <= 2 ** depth)
    # every terminal's Y size is below 1.5 or came from an early leaf choice
    assert r.count(0) == per.sum()
    r2 = O.derive(grammars.stochastic(), ax, key=axioms.STOCHASTIC_KEY)
    assert np.array_equal(canon(inst), canon(r2.instances(0)))
    r3 = O.derive(grammars.stochastic(), ax, key=(1, 2))
    assert r3.count(0) != r.count(0) or not np.array_equal(canon(inst), canon(r3.instances(0)))


# ---------------------------------------------------------------- gradients
def test_adjoint_twin_matches_fd():
    """The FP64 reverse-mode twin against central differences of the FP64
    twin (decisions frozen by the FP32 pass) on the parametrised facade."""
    ax, params = axioms.param_lots(3)
    prog = grammars.facade_params()
    r = O.derive(prog, ax, params=params, twins=O.TWIN_F64 | O.TWIN_ADJ)
    rng = np.random.default_rng(3)
    g = [rng.standard_normal((r.count(t), 12)) for t in range(2)]
    grad = r.backward(g)

    def loss(p):
        rr = O.derive(prog, ax, params=p, twins=O.TWIN_F64)
        assert rr.count(0) == r.count(0) and rr.count(1) == r.count(1)
        return sum(float(np.sum(rr.instances_f64(t)[:, :12] * g[t])) for t in range(2))

    h = 1e-3
    for lot in range(3):
        for k in (0, 1, 4, 5, 6, 7):
            pp, pm = params.copy(), params.copy()
            pp[lot, k] += h
            pm[lot, k] -= h
            # parameters are FP32: use the perturbation actually applied
            hh = (float(pp[lot, k]) - float(pm[lot, k])) / 2.0
            fd = (loss(pp) - loss(pm)) / (2 * hh)
            assert abs(fd - grad[lot, k]) <= 2e-3 * max(1.0, abs(fd)), (lot, k, fd, grad[lot, k])
    # repetition counts are piecewise constant in the repeat extents
    assert np.all(grad[:, 2] == 0.0) and np.all(grad[:, 3] == 0.0)
    assert np.all(np.abs(grad[:, [0, 1, 4, 5, 6, 7]]) > 0.0)
