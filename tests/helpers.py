"""Shared test helpers: canonical ordering of instance records, the
compiled reference grammars."""
import os

import numpy as np

from operatorgraph_amd import abi


def canon(inst):
    """Instances sorted by (axiom, path, meta, model bits...): a canonical
    order independent of execution order (the GPU appends in any order)."""
    inst = np.ascontiguousarray(inst, dtype=abi.INSTANCE_DTYPE)
    u = inst.view(np.uint32).reshape(len(inst), 16)
    keys = [u[:, k] for k in range(12, -1, -1)] + [u[:, 15], u[:, 14], u[:, 13]]
    order = np.lexsort(keys)
    return u[order]


def canon_order(inst):
    inst = np.ascontiguousarray(inst, dtype=abi.INSTANCE_DTYPE)
    u = inst.view(np.uint32).reshape(len(inst), 16)
    keys = [u[:, k] for k in range(12, -1, -1)] + [u[:, 15], u[:, 14], u[:, 13]]
    return np.lexsort(keys)


def assert_same_multiset(a, b, what=""):
    ca, cb = canon(a), canon(b)
    assert ca.shape == cb.shape, f"{what}: {ca.shape[0]} vs {cb.shape[0]} instances"
    if not np.array_equal(ca, cb):
        bad = np.nonzero(np.any(ca != cb, axis=1))[0]
        raise AssertionError(f"{what}: {len(bad)} of {len(ca)} instances differ, first {bad[:5]}:\n"
                             f"{ca[bad[0]]}\n{cb[bad[0]]}")


def rel_err(a, b):
    """Largest normwise relative error over rows: max_i |a_i - b_i|_inf / |b_i|_inf."""
    a = np.asarray(a, dtype=np.float64).reshape(len(a), -1)
    b = np.asarray(b, dtype=np.float64).reshape(len(b), -1)
    if a.size == 0:
        return 0.0
    num = np.max(np.abs(a - b), axis=1)
    den = np.maximum(np.max(np.abs(b), axis=1), 1e-30)
    return float(np.max(num / den))


# ---- compiled reference grammars (tests/golden/grammar_tables.json) ----
GRAMMAR_TABLES = os.path.join(os.path.dirname(__file__), "golden", "grammar_tables.json")


def grammar_fixture():
    import json
    with open(GRAMMAR_TABLES, encoding="utf-8") as f:
        return json.load(f)


def fixture_grammar(name):
    """The CompiledGrammar stored for one reference grammar."""
    import base64
    import ctypes as C
    import zlib

    from operatorgraph_amd import compiler
    from operatorgraph_amd.grammars import INTERPRETER_PROCEDURES
    from operatorgraph_amd.program import DispatchTable, Program
    rec = grammar_fixture()[name]
    raw = zlib.decompress(base64.b64decode(rec["entries"]))
    n = len(raw) // C.sizeof(abi.DispatchEntry)
    arr = (abi.DispatchEntry * n).from_buffer_copy(raw)
    table = DispatchTable([compiler._entry(arr[i]) for i in range(n)])
    return compiler.CompiledGrammar(Program(INTERPRETER_PROCEDURES, table, name),
                                    [tuple(a) for a in rec["axioms"]], raw)


def grammar_lots(g, name, n, seed=7):
    """n lots for a reference grammar: unit boxes (the interpreter's axiom)
    for grammars that scale themselves, building lots for the facade
    grammars. The skyscraper's heights are 3k + 2.75 so that every floor
    band its Repeat(Y, 3, ANCHOR_TO_END) leaves is tall enough for the
    absolute window factors (the reference traps otherwise,
    Subdivide.cuh:207-216)."""
    rng = np.random.default_rng(seed)
    seeds = rng.random(n).astype(np.float32)
    pos = np.stack([np.arange(n) * 60.0, np.zeros(n), (np.arange(n) % 4) * 60.0], 1)
    if name == "commercial":
        sizes = np.stack([rng.uniform(20, 40, n), rng.uniform(40, 90, n), rng.uniform(20, 40, n)], 1)
    elif name == "mc_skyscraper":
        sizes = np.stack([rng.uniform(20, 40, n), 3.0 * rng.integers(10, 25, n) + 2.75, rng.uniform(20, 40, n)], 1)
    else:
        sizes = np.ones((n, 3))
    return g.axiom_records(seeds, sizes=sizes, positions=pos)


# 3d_tree_<iterations>_<branches>: (trunks, leaves) from the recursion's
# closed form: a trunk per branch node, (b^(d+1) - 1) / (b - 1) in all, and
# L leaves per tip, L * b^d. 3d_tree_4_3 (d=4, b=3, L=2) is 121 / 162, the
# maxNumElements of its configuration.xml; 3d_tree_12_2 (d=11, b=2, L=3) is
# 4095 / 6144, the "(x4095)" / "(x6144)" its configuration.xml notes.
TREE_COUNTS = {
    "3d_tree_4_3": ((3 ** 5 - 1) // 2, 2 * 3 ** 4),
    "3d_tree_8_3": ((3 ** 9 - 1) // 2, 2 * 3 ** 8),
    "3d_tree_12_2": (2 ** 12 - 1, 3 * 2 ** 11),
}
