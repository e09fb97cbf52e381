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
    elif name in ("balcony", "suburban_house"):   # the compiled scenes' AxiomTraits sizes
        sizes = np.broadcast_to(scene_fixture()["Balcony" if name == "balcony" else "SuburbanHouse"]["axiom"]["size"],
                                (n, 3))
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


# ---- the reference's compiled test-runner scenes (tests/golden/scene_tables.json) ----
SCENE_TABLES = os.path.join(os.path.dirname(__file__), "golden", "scene_tables.json")

# MengerSponge.cuh:55-59 sizes its axiom 3^(attribute % MaxIterations); the
# scene's MaxIterations is 5.
MENGER_SCENE_MAX_ITERATIONS = 5


def scene_fixture():
    import json
    with open(SCENE_TABLES, encoding="utf-8") as f:
        return json.load(f)


class SceneProgram:
    """A scene's lowered procedure and entry arrays, in the shape
    oracle.derive and Engine.load take (arrays(), procedures, table.entries)."""

    def __init__(self, name, rec=None):
        import base64
        import ctypes as C
        import zlib
        rec = rec or scene_fixture()[name]
        pb = zlib.decompress(base64.b64decode(rec["procedures"]))
        eb = zlib.decompress(base64.b64decode(rec["entries"]))
        assert len(pb) == rec["num_procedures"] * C.sizeof(abi.Procedure)
        assert len(eb) == rec["num_entries"] * C.sizeof(abi.DispatchEntry)
        self.procs = (abi.Procedure * rec["num_procedures"]).from_buffer_copy(pb)
        self.entries = (abi.DispatchEntry * rec["num_entries"]).from_buffer_copy(eb)
        self.procedures = list(range(rec["num_procedures"]))
        self.table = type("Table", (), {"entries": list(range(rec["num_entries"]))})()
        self.name = name
        self.axiom = rec["axiom"]

    def arrays(self):
        return self.procs, self.entries


def scene_axioms(name, n, seed=5, menger_attribute=2):
    """n axioms of one scene as its AxiomTraits make them: the constant size
    and entry index the header returns, or its seed-driven size
    (MC_Skyscrapers.cuh:91-102: side (round(2s) + 2) * 5, height
    (round(3s) + 2) * 15) and MengerSponge's 3^(attribute % 5). Lots sit on
    a 28-unit grid, as the skyscraper scene's getPosition places them."""
    from operatorgraph_amd import axioms as A
    rec = scene_fixture()[name]
    rng = np.random.default_rng(seed)
    s = rng.random(n).astype(np.float32)
    ax = A.empty(n)
    if name == "MC_Skyscrapers":
        side = (np.round(2.0 * s) + 2.0) * 5.0
        height = (np.round(3.0 * s) + 2.0) * 15.0
        ax["size"] = np.stack([side, height, side], 1).astype(np.float32)
    elif name == "MengerSponge":
        ax["size"] = float(3 ** (menger_attribute % MENGER_SCENE_MAX_ITERATIONS))
    else:
        ax["size"] = rec["axiom"]["size"]
    i = np.arange(n)
    ax["model"][:, 3] = (i % 16) * 28.0 + 14.0
    ax["model"][:, 7] = ax["size"][:, 1] * 0.5
    ax["model"][:, 11] = (i // 16) * 28.0 + 14.0
    ax["seed"] = s
    ax["custom"] = s
    ax["entry_index"] = rec["axiom"]["entry"]
    ax["shape_type"] = abi.BOX
    return ax
