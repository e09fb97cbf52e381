"""Axiom generators: arrays of 80-byte `pga_axiom` records.

* `interpreter_axioms` mirrors the interpreter's AxiomGenerator
  (src/interpreter/PGAFacade.cu:697-729): origin, size 1, seed and custom
  attribute from the 32-slot seed table.
* `grid_axioms` mirrors Grid::Static (include/pga/core/Grid.cuh:188-258),
  including its row/column quirks.
* `facade_lots`, `stochastic_axioms`, `param_lots` draw the project
  workloads from Philox4x32-10 with fixed keys.
"""
import numpy as np

from . import abi
from .philox import philox4x32_10, uniform24

IDENTITY = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], dtype=np.float32)

FACADE_KEY = (0x5EED0002, 0)
STOCHASTIC_KEY = (0x5EED0003, 0)
GRADIENT_KEY = (0x5EED0004, 0)


def empty(n):
    a = np.zeros(n, dtype=abi.AXIOM_DTYPE)
    a["model"] = IDENTITY
    a["size"] = 1.0
    return a


def set_seeds_table(seeds):
    """PGA::Host::setSeeds (include/pga/core/GlobalVariables.cuh:38-44):
    copies (n % 33) seeds into a zero-initialised 32-slot table."""
    table = np.zeros(abi.MAX_NUM_AXIOM_SEEDS, dtype=np.float32)
    seeds = np.asarray(seeds, dtype=np.float32)
    k = len(seeds) % (abi.MAX_NUM_AXIOM_SEEDS + 1)
    table[:k] = seeds[:k]
    return table


def interpreter_axioms(entry_indices, shape_types, seed_table):
    """AxiomGenerator::generateAxiom (src/interpreter/PGAFacade.cu:697-729):
    position (0,0,0), size (1,1,1), seed = custom = Seeds[id % 32]."""
    n = len(entry_indices)
    a = empty(n)
    ids = np.arange(n)
    a["seed"] = seed_table[ids % abi.MAX_NUM_AXIOM_SEEDS]
    a["custom"] = a["seed"]
    a["entry_index"] = entry_indices
    a["shape_type"] = shape_types
    return a


def grid_axioms(rows, columns, entry_index, size, seed_table, shape_type=abi.BOX):
    """Grid::Static<Traits, 1, rows, columns>::generateAxiom (Grid.cuh:236-258)
    with DefaultAxiomTraits::getPosition (Grid.cuh:141-145). The reference
    computes row = groupId / Rows and passes (column, row) into the
    (row, column) slots of ForEachStartingAxiom::generate; both quirks kept."""
    n = rows * columns
    a = empty(n)
    sx, sy, sz = (np.float32(s) for s in size)
    for i in range(n):
        group = i
        column = group % columns
        row = group // rows
        r_slot, c_slot = column, row          # generate(id, axiomId, column, row)
        seed = seed_table[i % abi.MAX_NUM_AXIOM_SEEDS]
        a[i]["seed"] = seed
        a[i]["custom"] = seed
        a[i]["size"] = (sx, sy, sz)
        m = IDENTITY.copy()
        m[3] = np.float32(c_slot) * (sx * np.float32(1.5))
        m[7] = sy * np.float32(0.5)
        m[11] = np.float32(r_slot) * (sz * np.float32(1.5))
        a[i]["model"] = m
    a["entry_index"] = entry_index
    a["shape_type"] = shape_type
    return a


def menger_axiom(size, entry_index, seed=0.5, centred=True):
    """One Box axiom of edge `size`; at the origin (interpreter style) or
    resting on the ground plane (Grid::Static single cell, y = size/2)."""
    a = empty(1)
    a["size"] = np.float32(size)
    a["seed"] = np.float32(seed)
    a["custom"] = np.float32(seed)
    if not centred:
        a[0]["model"][7] = np.float32(size) * np.float32(0.5)
    a["entry_index"] = entry_index
    a["shape_type"] = abi.BOX
    return a


def facade_lots(n, key=FACADE_KEY, columns=None, pitch=50.0, entry_index=0):
    """Building lots: width W = 10 + 30 u0, depth D = 10 + 20 u1,
    floors F = 3 + floor(18 u2) in {3..20}, height H = 3 F (all FP32);
    (u0, u1, u2) = Philox4x32-10(counter = (lot, 0, 0, 0), key). Lots sit on a
    square grid of `pitch` metres, on the ground (y = H/2); seed and custom
    attribute are u3."""
    ids = np.arange(n, dtype=np.uint32)
    ctr = np.zeros((n, 4), dtype=np.uint32)
    ctr[:, 0] = ids
    u = uniform24(philox4x32_10(ctr, key))
    f32 = np.float32
    w = f32(10.0) + f32(30.0) * u[:, 0]
    d = f32(10.0) + f32(20.0) * u[:, 1]
    floors = (f32(3.0) + np.floor(f32(18.0) * u[:, 2])).astype(np.float32)
    h = f32(3.0) * floors
    if columns is None:
        columns = max(1, int(np.ceil(np.sqrt(n))))
    a = empty(n)
    a["size"][:, 0] = w
    a["size"][:, 1] = h
    a["size"][:, 2] = d
    a["model"][:, 3] = (ids % columns).astype(np.float32) * f32(pitch)
    a["model"][:, 7] = h * f32(0.5)
    a["model"][:, 11] = (ids // columns).astype(np.float32) * f32(pitch)
    a["seed"] = u[:, 3]
    a["custom"] = u[:, 3]
    a["entry_index"] = entry_index
    a["shape_type"] = abi.BOX
    return a


def facade_lot_terminals(axioms):
    """Closed-form terminal counts of `grammars.facade()` per lot:
    floors = floor(H/3) per side, tiles = floor(W/2.5) or floor(D/2.5), every
    tile is at least 2.5 wide (>= 1.6), so 2 walls + 1 window per tile."""
    f32 = np.float32
    w, h, d = (axioms["size"][:, i].astype(np.float32) for i in range(3))
    floors = np.floor(h / f32(3.0))
    tw = np.floor(w / f32(2.5))
    td = np.floor(d / f32(2.5))
    tiles = 2 * floors * (tw + td)
    return (2 * tiles).astype(np.int64), tiles.astype(np.int64)   # walls, windows


def stochastic_axioms(n, key=STOCHASTIC_KEY, mean_log_terminals=3.03, sigma=1.5, columns=None,
                      pitch=24.0, entry_index=0):
    """Axioms of `grammars.stochastic()`: log T ~ N(mean, sigma) drawn by
    Box-Muller from Philox, depth budget d = clamp(round(log2 T), 1, 12),
    axiom size (16, 2^d, 16)."""
    ids = np.arange(n, dtype=np.uint32)
    ctr = np.zeros((n, 4), dtype=np.uint32)
    ctr[:, 0] = ids
    ctr[:, 1] = 1
    u = uniform24(philox4x32_10(ctr, key)).astype(np.float64)
    z = np.sqrt(-2.0 * np.log(np.maximum(u[:, 0], 2.0 ** -24))) * np.cos(2.0 * np.pi * u[:, 1])
    logt = mean_log_terminals + sigma * z
    depth = np.clip(np.rint(logt / np.log(2.0)), 1, 12).astype(np.int64)
    if columns is None:
        columns = max(1, int(np.ceil(np.sqrt(n))))
    a = empty(n)
    a["size"][:, 0] = 16.0
    a["size"][:, 1] = np.exp2(depth).astype(np.float32)
    a["size"][:, 2] = 16.0
    a["model"][:, 3] = (ids % columns).astype(np.float32) * np.float32(pitch)
    a["model"][:, 11] = (ids // columns).astype(np.float32) * np.float32(pitch)
    a["seed"] = u[:, 2].astype(np.float32)
    a["custom"] = a["seed"]
    a["entry_index"] = entry_index
    a["shape_type"] = abi.BOX
    return a, depth


def param_lots(n, key=GRADIENT_KEY, perturb=0.0, defaults=None):
    """Lots of `grammars.facade_params()` with 8 parameters each:
    parameters = defaults + perturb * N(0, 1) (Box-Muller over Philox, counter
    (lot, 2, k, 0)). Returns (axioms, params[n, 8] float32)."""
    from .grammars import FACADE_PARAM_DEFAULTS
    defaults = np.asarray(defaults if defaults is not None else FACADE_PARAM_DEFAULTS, dtype=np.float64)
    a = facade_lots(n, key=key)
    p = np.tile(defaults, (n, 1))
    if perturb:
        ids = np.arange(n, dtype=np.uint32)
        ctr = np.zeros((n, 4), dtype=np.uint32)
        ctr[:, 0] = ids
        ctr[:, 1] = 2
        noise = np.zeros((n, 8))
        for k in range(4):
            ctr[:, 2] = k
            u = uniform24(philox4x32_10(ctr, key)).astype(np.float64)
            r = np.sqrt(-2.0 * np.log(np.maximum(u[:, 0], 2.0 ** -24)))
            noise[:, 2 * k] = r * np.cos(2.0 * np.pi * u[:, 1])
            noise[:, 2 * k + 1] = r * np.sin(2.0 * np.pi * u[:, 1])
        p = p + perturb * noise
    return a, p.astype(np.float32)


CITY_KEY = (0x5EED0005, 0)


def city_lots(n, key=CITY_KEY, columns=None, pitch=16.0, entry_index=0):
    """Lots of `grammars.context_city()`: width and depth uniform in [6, 12),
    height 3, seeds uniform in [0, 1) (Philox4x32-10, counter (lot, 3, 0, 0)),
    on a grid of `pitch` (lots never overlap)."""
    ids = np.arange(n, dtype=np.uint32)
    ctr = np.zeros((n, 4), dtype=np.uint32)
    ctr[:, 0] = ids
    ctr[:, 1] = 3
    u = uniform24(philox4x32_10(ctr, key)).astype(np.float64)
    if columns is None:
        columns = max(1, int(np.ceil(np.sqrt(n))))
    a = empty(n)
    a["size"][:, 0] = (6.0 + 6.0 * u[:, 0]).astype(np.float32)
    a["size"][:, 1] = 3.0
    a["size"][:, 2] = (6.0 + 6.0 * u[:, 1]).astype(np.float32)
    a["model"][:, 3] = (ids % columns).astype(np.float32) * np.float32(pitch)
    a["model"][:, 11] = (ids // columns).astype(np.float32) * np.float32(pitch)
    a["seed"] = u[:, 2].astype(np.float32)
    a["custom"] = a["seed"]
    a["entry_index"] = entry_index
    a["shape_type"] = abi.BOX
    return a


def collision_pair_axioms(kinds, rng, spread=1.6, quad_depth=0.0):
    """Axioms and per-axiom parameters of `grammars.collision_pairs()`: pair i
    is axioms 2i (A) and 2i + 1 (B) with shape types kinds[i] = (type_a,
    type_b). Sizes uniform in [0.4, 2), rotations uniform in [0, 360)
    degrees, A's translation uniform in [-0.2, 0.2)^3 and B's in
    [-spread, spread)^3; Quads get z extent `quad_depth`. Tag i + 1.
    Returns (axioms, params[2n, COLLISION_PAIR_SLOTS])."""
    from .grammars import COLLISION_PAIR_ENTRIES, COLLISION_PAIR_SLOTS
    n = len(kinds)
    a = empty(2 * n)
    p = np.zeros((2 * n, COLLISION_PAIR_SLOTS), dtype=np.float32)
    for i, (ta, tb) in enumerate(kinds):
        for r, (role, t) in enumerate((("A", ta), ("B", tb))):
            k = 2 * i + r
            a[k]["entry_index"] = COLLISION_PAIR_ENTRIES[(role, t)]
            a[k]["shape_type"] = t
            size = rng.uniform(0.4, 2.0, 3)
            if t == abi.QUAD:
                size[2] = quad_depth
            p[k, 0:3] = size
            p[k, 3:6] = rng.uniform(0.0, 360.0, 3)
            p[k, 6:9] = rng.uniform(-0.2, 0.2, 3) if role == "A" else rng.uniform(-spread, spread, 3)
            p[k, 9] = i + 1
    return a, p
