"""Shared test helpers: canonical ordering of instance records."""
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
