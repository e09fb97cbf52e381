"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle library
(oracle/_build/libpga_oracle.so, built by oracle/Makefile).

Only tests/, the smoke check and bench.py's CPU-baseline leg use this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from operatorgraph_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpga_oracle.so")

TWIN_F64 = 1
TWIN_ADJ = 2

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.pgao_derive.restype = C.c_int
        L.pgao_derive.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.c_uint32, P, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(P)]
        L.pgao_free.argtypes = [P]
        L.pgao_error.restype = C.c_char_p
        L.pgao_error.argtypes = [P]
        L.pgao_num_terminals.restype = C.c_uint32
        L.pgao_num_terminals.argtypes = [P]
        L.pgao_count.restype = C.c_uint64
        L.pgao_count.argtypes = [P, C.c_uint32]
        L.pgao_steps.restype = C.c_uint64
        L.pgao_steps.argtypes = [P]
        L.pgao_executions.restype = C.c_uint64
        L.pgao_executions.argtypes = [P]
        L.pgao_copy_instances.argtypes = [P, C.c_uint32, P]
        L.pgao_copy_instances_f64.argtypes = [P, C.c_uint32, P]
        L.pgao_mesh_floats.restype = C.c_uint64
        L.pgao_mesh_floats.argtypes = [P, C.c_uint32]
        L.pgao_copy_mesh.argtypes = [P, C.c_uint32, P]
        L.pgao_backward.argtypes = [P, P, P]
        L.pgao_next_seed.restype = C.c_float
        L.pgao_next_seed.argtypes = [C.c_float, C.c_uint32]
        L.pgao_range.restype = C.c_float
        L.pgao_seed_component.restype = C.c_uint32
        L.pgao_seed_component.argtypes = [C.c_uint32, C.c_uint32]
        L.pgao_seed_state.argtypes = [C.c_float, C.c_uint32, P]
        L.pgao_seed_unit.restype = C.c_float
        L.pgao_seed_unit.argtypes = [C.c_uint32]
        L.pgao_range.argtypes = [C.c_float, C.c_float, C.c_float]
        L.pgao_philox4x32_10.argtypes = [P, P, P]
        L.pgao_sinf.restype = C.c_float
        L.pgao_sinf.argtypes = [C.c_float]
        L.pgao_cosf.restype = C.c_float
        L.pgao_cosf.argtypes = [C.c_float]
        L.pgao_path_mix.restype = C.c_uint32
        L.pgao_path_mix.argtypes = [C.c_uint32, C.c_uint32]
        L.pgao_pack_operand.restype = C.c_float
        L.pgao_pack_operand.argtypes = [C.c_int, C.c_float]
        L.pgao_pack_float2.restype = C.c_float
        L.pgao_pack_float2.argtypes = [C.c_float, C.c_float]
        L.pgao_lbvh.restype = C.c_int
        L.pgao_lbvh.argtypes = [P, C.c_uint32, P, P, P, P, P, P, P, P]
        L.pgao_num_colliders.restype = C.c_uint32
        L.pgao_num_colliders.argtypes = [P]
        L.pgao_shapes_collide.restype = C.c_int
        L.pgao_shapes_collide.argtypes = [P, P, C.c_int32, P, P, C.c_int32]
        L.pgao_morton30.restype = C.c_uint32
        L.pgao_morton30.argtypes = [C.c_float, C.c_float, C.c_float]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Result:
    def __init__(self, handle, n_axioms, n_params):
        self._h = handle
        self.n_axioms = n_axioms
        self.n_params = n_params

    def __del__(self):
        if self._h and _lib is not None:
            _lib.pgao_free(self._h)
            self._h = None

    @property
    def num_terminals(self):
        return lib().pgao_num_terminals(self._h)

    def count(self, t):
        return lib().pgao_count(self._h, t)

    @property
    def steps(self):
        return lib().pgao_steps(self._h)

    @property
    def colliders(self):
        """Shapes the Collider operator stored, over all phases."""
        return lib().pgao_num_colliders(self._h)

    @property
    def executions(self):
        return lib().pgao_executions(self._h)

    def instances(self, t):
        n = self.count(t)
        out = np.zeros(n, dtype=abi.INSTANCE_DTYPE)
        if n:
            lib().pgao_copy_instances(self._h, t, out.ctypes.data)
        return out

    def mesh(self, t):
        """Generation-function-0 vertices of terminal t: (n_vertices, 9)
        float32 (position xyzw, normal xyz, uv), in derivation order."""
        n = lib().pgao_mesh_floats(self._h, t)
        out = np.zeros(n, dtype=np.float32)
        if n:
            lib().pgao_copy_mesh(self._h, t, out.ctypes.data)
        return out.reshape(-1, 9)

    def instances_f64(self, t):
        n = self.count(t)
        out = np.zeros((n, 13), dtype=np.float64)
        if n:
            st = lib().pgao_copy_instances_f64(self._h, t, out.ctypes.data)
            if st:
                raise OracleError(st, "f64 twin not recorded")
        return out

    def backward(self, dl_dmodel_per_terminal):
        """dl_dmodel_per_terminal: list (terminal order) of (count, 12) arrays."""
        parts = []
        for t in range(abi.MAX_TERMINALS):
            n = self.count(t)
            if t < len(dl_dmodel_per_terminal) and dl_dmodel_per_terminal[t] is not None:
                g = np.ascontiguousarray(dl_dmodel_per_terminal[t], dtype=np.float64).reshape(n, 12)
            else:
                g = np.zeros((n, 12))
            parts.append(g.reshape(-1))
        flat = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(0))
        out = np.zeros(self.n_axioms * self.n_params, dtype=np.float64)
        st = lib().pgao_backward(self._h, flat.ctypes.data, out.ctypes.data)
        if st:
            raise OracleError(st, "adjoint twin not recorded")
        return out.reshape(self.n_axioms, self.n_params)


def derive(program, axioms, params=None, key=(0, 0), twins=0, max_steps=0):
    """Runs the oracle derivation of `program` (operatorgraph_amd.program.Program)
    over `axioms` (numpy AXIOM_DTYPE array)."""
    L = lib()
    procs, entries = program.arrays()
    axioms = np.ascontiguousarray(axioms, dtype=abi.AXIOM_DTYPE)
    n_params = 0
    pptr = None
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        n_params = params.shape[1]
        pptr = params.ctypes.data
    h = C.c_void_p()
    st = L.pgao_derive(C.addressof(procs), len(program.procedures), C.addressof(entries),
                       len(program.table.entries), axioms.ctypes.data, len(axioms),
                       pptr, n_params, key[0], key[1], twins, max_steps, C.byref(h))
    res = Result(h.value, len(axioms), n_params)
    if st:
        msg = L.pgao_error(h.value).decode()
        raise OracleError(st, msg)
    return res


def shapes_collide(a_model, a_size, a_type, b_model, b_size, b_type):
    """Collision::check(stored a, asking b) of two Box/Quad shapes: model as
    12 floats (3x4 row-major), size as 3 floats, abi.BOX or abi.QUAD."""
    am = np.ascontiguousarray(a_model, dtype=np.float32)
    asz = np.ascontiguousarray(a_size, dtype=np.float32)
    bm = np.ascontiguousarray(b_model, dtype=np.float32)
    bsz = np.ascontiguousarray(b_size, dtype=np.float32)
    return bool(lib().pgao_shapes_collide(am.ctypes.data, asz.ctypes.data, int(a_type), bm.ctypes.data,
                                          bsz.ctypes.data, int(b_type)))


def next_seed(base, modifier):
    return lib().pgao_next_seed(base, modifier)


def seed_component(z, which):
    """One of hybridTaus's four generators: 0..2 Tausworthe steps, 3 the LCG."""
    return lib().pgao_seed_component(z, which)


def seed_state(base, modifier):
    """The component states z1..z4 hybridTaus derives from (base, modifier)."""
    z = (C.c_uint32 * 4)()
    lib().pgao_seed_state(base, modifier, C.addressof(z))
    return list(z)


def seed_unit(t):
    """hybridTaus's final 32-bit value -> float conversion."""
    return lib().pgao_seed_unit(t)


def range_(seed, mn, mx):
    return lib().pgao_range(seed, mn, mx)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().pgao_philox4x32_10(C.addressof(c), C.addressof(k), C.addressof(o))
    return list(o)


def sinf(x):
    return lib().pgao_sinf(x)


def cosf(x):
    return lib().pgao_cosf(x)


def path_mix(path, ordinal):
    return lib().pgao_path_mix(path, ordinal)


def pack_operand(t, x):
    return lib().pgao_pack_operand(t, x)


def pack_float2(x, y):
    return lib().pgao_pack_float2(x, y)


def morton30(x, y, z):
    return lib().pgao_morton30(x, y, z)


def lbvh(aabbs, scene_min, scene_max):
    aabbs = np.ascontiguousarray(aabbs, dtype=np.float32)
    n = len(aabbs)
    smin = np.ascontiguousarray(scene_min, dtype=np.float32)
    smax = np.ascontiguousarray(scene_max, dtype=np.float32)
    keys = np.zeros(n, dtype=np.uint64)
    left = np.zeros(n - 1, dtype=np.int32)
    right = np.zeros(n - 1, dtype=np.int32)
    parent = np.zeros(n - 1, dtype=np.int32)
    leaf_parent = np.zeros(n, dtype=np.int32)
    node_aabb = np.zeros((n - 1, 6), dtype=np.float32)
    st = lib().pgao_lbvh(aabbs.ctypes.data, n, smin.ctypes.data, smax.ctypes.data, keys.ctypes.data,
                         left.ctypes.data, right.ctypes.data, parent.ctypes.data,
                         leaf_parent.ctypes.data, node_aabb.ctypes.data)
    if st:
        raise OracleError(st, "lbvh failed")
    return dict(keys=keys, left=left, right=right, parent=parent, leaf_parent=leaf_parent,
                node_aabb=node_aabb)


# Corner signs (x, y, z) of the reference's generation-function-0 faces, in
# half-size units, default winding (PGA_CW undefined): top, bottom, left,
# right, front, back of ShapeGenerator<Box, false>::run
# (include/pga/core/ShapeGenerator.cuh:376-511) and the one face of
# ShapeGenerator<Quad, false>::run (:670-708).
BOX_FACES = np.array([
    [[-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]],
    [[-1, -1, 1], [1, -1, 1], [1, -1, -1], [-1, -1, -1]],
    [[-1, 1, -1], [-1, 1, 1], [-1, -1, 1], [-1, -1, -1]],
    [[1, 1, 1], [1, 1, -1], [1, -1, -1], [1, -1, 1]],
    [[-1, 1, 1], [1, 1, 1], [1, -1, 1], [-1, -1, 1]],
    [[1, 1, -1], [-1, 1, -1], [-1, -1, -1], [1, -1, -1]]], dtype=np.float32)
QUAD_FACE = np.array([[[-1, 1, 0], [1, 1, 0], [1, -1, 0], [-1, -1, 0]]], dtype=np.float32)
FACE_UV = np.array([[1, 1], [0, 1], [0, 0], [1, 0]], dtype=np.float32)   # uv0..uv3


def mesh_from_instances(models, shape_type):
    """Generation-function-0 vertices of instance records, vectorised in
    numpy float32 for sizes the C oracle is slow at: (n, 12) row-major
    I = model4 * scale(size) -> (n * V, 9). A corner is I * (+-0.5, 1)
    summed as ((a x + b y) + c z) + d w (math::float4x4 * float4,
    include/pga/math/matrix.h:793-799); the normal is
    normalize(cross(v0 - v1, v2 - v1)) (vector.h:510-515, 566-574). Each
    numpy float32 operation rounds once, with no fused multiply-add, so
    the values are those of the C oracle's shape_mesh up to the sign of a
    zero (the C oracle scales the model by size / 2 instead)."""
    faces = QUAD_FACE if shape_type == abi.QUAD else BOX_FACES
    m = np.ascontiguousarray(models, dtype=np.float32).reshape(-1, 3, 4)
    half = faces * np.float32(0.5)                                     # (F, 4, 3)
    x, y, z = half[..., 0], half[..., 1], half[..., 2]
    one = np.float32(1.0)
    p = np.empty((len(m), len(faces), 4, 4), dtype=np.float32)
    for r in range(3):
        a, b, c, d = (m[:, r, k][:, None, None] for k in range(4))
        p[..., r] = ((a * x + b * y) + c * z) + d * one
    p[..., 3] = ((np.float32(0) * x + np.float32(0) * y) + np.float32(0) * z) + one * one
    e1 = p[:, :, 0, :3] - p[:, :, 1, :3]
    e2 = p[:, :, 2, :3] - p[:, :, 1, :3]
    n = np.stack([e1[..., 1] * e2[..., 2] - e1[..., 2] * e2[..., 1],
                  e1[..., 2] * e2[..., 0] - e1[..., 0] * e2[..., 2],
                  e1[..., 0] * e2[..., 1] - e1[..., 1] * e2[..., 0]], -1)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = one / np.sqrt((n[..., 0] * n[..., 0] + n[..., 1] * n[..., 1]) + n[..., 2] * n[..., 2])
        n = n * inv[..., None]
    out = np.empty((len(m), len(faces), 4, 9), dtype=np.float32)
    out[..., 0:4] = p
    out[..., 4:7] = n[:, :, None, :]
    out[..., 7:9] = FACE_UV
    return out.reshape(-1, 9)


def mesh_indices(n, shape_type, base_vertex=0):
    """Index buffer of n shapes: (0, 1, 2), (0, 2, 3) per face of four
    vertices (ShapeGenerator.cuh:505-509, 703-704), offset by base_vertex."""
    faces = 1 if shape_type == abi.QUAD else 6
    q = (np.arange(n * faces, dtype=np.int64) * 4)[:, None] + np.array([0, 1, 2, 0, 2, 3])
    return (q.reshape(-1) + base_vertex).astype(np.uint32)
