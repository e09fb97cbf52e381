"""Linear BVH over terminal boxes on the GPU (pga_lbvh_* in include/pga.h).

The reference builds it in GPUBVHConstructor.cuh:29-333 (Morton codes,
thrust::sort, Karras radix tree, ticket refit); this wrapper drives the
gfx950 kernels of csrc/pga_lbvh.hip on torch device tensors."""
import ctypes as C
import sys

from . import abi
from .engine import PGAError, load_library

_SIGS = {
    "pga_lbvh_create": (C.c_int, [C.c_uint32, C.c_int32, C.c_void_p, C.POINTER(C.c_void_p)]),
    "pga_lbvh_destroy": (None, [C.c_void_p]),
    "pga_lbvh_last_error": (C.c_char_p, [C.c_void_p]),
    "pga_lbvh_instance_aabbs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p]),
    "pga_lbvh_bounds": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float * 6)]),
    "pga_lbvh_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float * 3),
                                 C.POINTER(C.c_float * 3)] + [C.c_void_p] * 6),
    "pga_lbvh_seconds": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "pga_lbvh_gather_instances": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p * abi.MAX_TERMINALS),
                                            C.POINTER(C.c_uint64 * abi.MAX_TERMINALS), C.c_uint32, C.c_void_p,
                                            C.c_uint32, C.c_void_p]),
    "pga_lbvh_build_device": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float * 3),
                                        C.POINTER(C.c_float * 3)] + [C.c_void_p] * 6),
    "pga_lbvh_leaf_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]),
    "pga_lbvh_build_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float * 3),
                                         C.POINTER(C.c_float * 3)] + [C.c_void_p] * 7),
}

NOALLOC_MAX_LEAVES = 1 << 20   # PGA_LBVH_NOALLOC_MAX_LEAVES


def build_noalloc(aabbs, scene_min, scene_max, out, stream=None):
    """pga_lbvh_build_noalloc: the build into the caller's device arrays
    `out` (keys, left, right, parent, leaf_parent, node_aabb, as from
    LBVH.alloc() or LBVH.build()) with the library's own per-device scratch;
    nothing is allocated. Enqueued on `stream` (torch's current stream when
    None); the caller synchronises before reading the outputs."""
    import torch
    L = _lib()
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    smin = (C.c_float * 3)(*[float(v) for v in scene_min])
    smax = (C.c_float * 3)(*[float(v) for v in scene_max])
    ptr = lambda t: C.c_void_p(t.data_ptr())
    st = L.pga_lbvh_build_noalloc(ptr(aabbs), int(aabbs.shape[0]), C.byref(smin), C.byref(smax), ptr(out["keys"]),
                                  ptr(out["left"]), ptr(out["right"]), ptr(out["parent"]), ptr(out["leaf_parent"]),
                                  ptr(out["node_aabb"]), C.c_void_p(stream))
    if st != abi.OK:
        raise PGAError(st, L.pga_lbvh_last_error(None).decode())


def _lib():
    L = load_library()
    if not getattr(L, "_lbvh_bound", False):
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._lbvh_bound = True
    return L


class LBVH:
    """Builder for up to `capacity` boxes; scratch lives on the device.

    `engine`: build on that Engine's stream (so that a derivation and the
    build record into one graph); the builder then keeps the engine, whose
    stream it uses, alive until it is closed."""

    def __init__(self, capacity, device=-1, stream=None, engine=None):
        import torch
        if not torch.cuda.is_available():
            raise PGAError(abi.ERR_HIP, "no GPU visible: the LBVH has no CPU path")
        self.torch = torch
        self.L = _lib()
        self._engine = engine
        if engine is not None:
            stream = engine.stream()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        h = C.c_void_p()
        st = self.L.pga_lbvh_create(int(capacity), device, C.c_void_p(stream), C.byref(h))
        self._h = h.value
        self._check(st)
        self.capacity = int(capacity)

    def _check(self, st):
        if st != abi.OK:
            msg = self.L.pga_lbvh_last_error(self._h).decode() if self._h else ""
            raise PGAError(st, msg)

    def close(self):
        if getattr(self, "_h", None):
            self.L.pga_lbvh_destroy(self._h)
            self._h = None
        self._engine = None

    def __del__(self, _finalizing=sys.is_finalizing):   # bound now: module globals are gone at shutdown
        # at interpreter shutdown the HIP runtime may already be torn down:
        # the process releases the device memory
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def instance_aabbs(self, instances, planar):
        """Boxes of a (n, 16) float32 instance tensor (64-byte records)."""
        torch = self.torch
        n = instances.shape[0]
        out = torch.empty((n, 6), dtype=torch.float32, device=instances.device)
        self._check(self.L.pga_lbvh_instance_aabbs(self._h, C.c_void_p(instances.data_ptr()), n,
                                                   1 if planar else 0, C.c_void_p(out.data_ptr())))
        return out

    def bounds(self, aabbs):
        out = (C.c_float * 6)()
        self._check(self.L.pga_lbvh_bounds(self._h, C.c_void_p(aabbs.data_ptr()), aabbs.shape[0], C.byref(out)))
        return list(out)

    def build(self, aabbs, scene_min, scene_max):
        """Returns a dict of device tensors: keys (uint64 as int64), left, right,
        parent (n-1), leaf_parent (n), node_aabb (n-1, 6)."""
        torch = self.torch
        n = aabbs.shape[0]
        dev = aabbs.device
        keys = torch.empty(n, dtype=torch.int64, device=dev)
        left = torch.empty(max(n - 1, 1), dtype=torch.int32, device=dev)
        right = torch.empty_like(left)
        parent = torch.empty_like(left)
        leaf_parent = torch.empty(n, dtype=torch.int32, device=dev)
        node_aabb = torch.empty((max(n - 1, 1), 6), dtype=torch.float32, device=dev)
        smin = (C.c_float * 3)(*[float(v) for v in scene_min])
        smax = (C.c_float * 3)(*[float(v) for v in scene_max])
        ptr = lambda t: C.c_void_p(t.data_ptr())
        self._check(self.L.pga_lbvh_build(self._h, ptr(aabbs.contiguous()), n, C.byref(smin), C.byref(smax),
                                          ptr(keys), ptr(left), ptr(right), ptr(parent), ptr(leaf_parent),
                                          ptr(node_aabb)))
        return dict(keys=keys, left=left, right=right, parent=parent, leaf_parent=leaf_parent,
                    node_aabb=node_aabb)

    # -- device-count form: the leaf count stays on the device --
    def alloc(self):
        """Device arrays for a build of up to `capacity` leaves: aabbs and the
        build_device outputs (a graph records their addresses)."""
        torch = self.torch
        n = self.capacity
        dev = "cuda"
        return dict(aabbs=torch.empty((n, 6), dtype=torch.float32, device=dev),
                    keys=torch.empty(n, dtype=torch.int64, device=dev),
                    left=torch.empty(n - 1, dtype=torch.int32, device=dev),
                    right=torch.empty(n - 1, dtype=torch.int32, device=dev),
                    parent=torch.empty(n - 1, dtype=torch.int32, device=dev),
                    leaf_parent=torch.empty(n, dtype=torch.int32, device=dev),
                    node_aabb=torch.empty((n - 1, 6), dtype=torch.float32, device=dev))

    def gather_engine(self, engine, n_types, planar_mask, out):
        """Boxes of the instances of terminal types 0..n_types-1 bound to
        `engine`, counted on the device (pga_lbvh_gather_instances)."""
        bufs = (C.c_void_p * abi.MAX_TERMINALS)()
        caps = (C.c_uint64 * abi.MAX_TERMINALS)()
        for t in range(n_types):
            b = engine.buffer(t)
            if b is not None:
                bufs[t] = b.data_ptr()
                caps[t] = b.shape[0]
        self._check(self.L.pga_lbvh_gather_instances(self._h, C.byref(bufs), C.byref(caps), n_types,
                                                     C.c_void_p(engine.device_counts()), planar_mask,
                                                     C.c_void_p(out["aabbs"].data_ptr())))

    def build_device(self, out, scene_min=None, scene_max=None):
        """Builds over the leaf count of the last gather into `out` (from
        alloc()); the scene box is the union of the boxes when not given."""
        ptr = lambda k: C.c_void_p(out[k].data_ptr())
        smin = smax = None
        if scene_min is not None:
            smin = C.byref((C.c_float * 3)(*[float(v) for v in scene_min]))
            smax = C.byref((C.c_float * 3)(*[float(v) for v in scene_max]))
        self._check(self.L.pga_lbvh_build_device(self._h, ptr("aabbs"), smin, smax, ptr("keys"), ptr("left"),
                                                 ptr("right"), ptr("parent"), ptr("leaf_parent"),
                                                 ptr("node_aabb")))

    def leaf_count(self):
        """Waits for the builder's stream; (leaf count, overflowed)."""
        n, o = C.c_uint32(0), C.c_int32(0)
        self._check(self.L.pga_lbvh_leaf_count(self._h, C.byref(n), C.byref(o)))
        return n.value, bool(o.value)

    def seconds(self):
        s = C.c_double(0.0)
        self._check(self.L.pga_lbvh_seconds(self._h, C.byref(s)))
        return s.value
