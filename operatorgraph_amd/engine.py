"""ctypes binding of libpga_amd.so (include/pga.h).

`Engine` mirrors the reference's evaluator life cycle
(SinglePhaseEvaluator::initialize / execute / release,
include/pga/core/GPUEvaluator.cuh:475-536): load a program, bind axioms and
terminal buffers, execute. Device memory comes from torch (plumbing only);
every derivation runs in the HIP kernels of libpga_amd.so. There is no CPU
fallback: without the built library, or without a GPU, construction raises.
"""
import contextlib
import ctypes as C
import os
import sys

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpga_amd.so")

_lib = None


class PGAError(RuntimeError):
    def __init__(self, status, message=""):
        name = abi.STATUS_NAMES.get(status, str(status))
        super().__init__(f"{name}: {message}" if message else name)
        self.status = status


def load_library():
    """Loads the in-tree libpga_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PGAError(abi.ERR_NOT_INITIALIZED,
                       f"{LIB_PATH} is missing: build it with `python -m operatorgraph_amd.build`")
    # torch first: the library then binds to the HIP runtime torch loaded,
    # so both see the same devices in one process (loaded the other way
    # round, a later pga_create found no device)
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sigs = {
        "pga_create": (C.c_int, [C.POINTER(abi.Config), C.POINTER(P)]),
        "pga_destroy": (None, [P]),
        "pga_last_error": (C.c_char_p, [P]),
        "pga_status_string": (C.c_char_p, [C.c_int]),
        "pga_abi_version": (C.c_uint32, []),
        "pga_initialize": (C.c_int, [P, P, C.c_uint32, P, C.c_uint32]),
        "pga_set_axioms": (C.c_int, [P, P, C.c_uint32]),
        "pga_set_axioms_device": (C.c_int, [P, P, C.c_uint32]),
        "pga_set_seeds": (C.c_int, [P, P, C.c_uint32]),
        "pga_get_seeds": (C.c_int, [P, P]),
        "pga_set_axiom_params": (C.c_int, [P, P, C.c_uint32, C.c_uint32]),
        "pga_set_philox_key": (C.c_int, [P, C.c_uint32, C.c_uint32]),
        "pga_bind_instances": (C.c_int, [P, C.c_uint32, P, C.c_uint64]),
        "pga_get_counts": (C.c_int, [P, P, C.c_uint32]),
        "pga_execute": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_double)]),
        "pga_execute_range": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_double)]),
        "pga_get_stats": (C.c_int, [P, C.POINTER(abi.Stats)]),
        "pga_set_collider_capacity": (C.c_int, [P, C.c_uint32]),
        "pga_num_colliders": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "pga_execute_async": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_int32]),
        "pga_synchronize": (C.c_int, [P]),
        "pga_kernel_seconds": (C.c_int, [P, C.POINTER(C.c_double)]),
        "pga_set_timestamps": (C.c_int, [P, P, C.c_uint32]),
        "pga_timestamp_hz": (C.c_double, [P]),
        "pga_autotune": (C.c_int, [P, C.c_uint32, C.c_uint32, C.POINTER(abi.TuneResult)]),
        "pga_specialize": (C.c_int, [P, C.c_uint32, C.c_int32, C.POINTER(C.c_double)]),
        "pga_set_schedule": (C.c_int, [P, C.c_int32, C.c_uint32, C.c_uint32]),
        "pga_specialize_check": (C.c_int, [P, C.c_uint32, P, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32,
                                           C.POINTER(C.c_double)]),
        "pga_tape_reserve": (C.c_int, [P, C.c_uint32, C.c_uint32]),
        "pga_forward_tape": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_double)]),
        "pga_backward": (C.c_int, [P, C.POINTER(C.c_void_p * abi.MAX_TERMINALS), P, C.POINTER(C.c_double)]),
        "pga_tape_generation": (C.c_int, [P, C.POINTER(C.c_uint64)]),
        "pga_tape_counts": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "pga_capture_begin": (C.c_int, [P]),
        "pga_capture_end": (C.c_int, [P, C.POINTER(P)]),
        "pga_graph_launch": (C.c_int, [P, P]),
        "pga_graph_destroy": (None, [P]),
        "pga_graph_info": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "pga_device_counts": (C.c_int, [P, C.POINTER(P)]),
        "pga_get_stream": (C.c_int, [P, C.POINTER(P)]),
        "pga_interpreter_procedures": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_uint32)]),
        "pga_compile_grammar": (C.c_int, [C.c_char_p, C.POINTER(P)]),
        "pga_grammar_error": (C.c_char_p, [P]),
        "pga_grammar_entries": (C.c_int, [P, C.POINTER(C.POINTER(abi.DispatchEntry)), C.POINTER(C.c_uint32)]),
        "pga_grammar_axioms": (C.c_int, [P, C.POINTER(C.POINTER(abi.GrammarAxiom)), C.POINTER(C.c_uint32)]),
        "pga_grammar_destroy": (None, [P]),
        "pga_center_size_mse": (C.c_int, [P, P, P, C.c_uint64, C.c_float, C.c_float, P, P, P]),
        "pga_generate_partition": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(P)]),
        "pga_tessellate": (C.c_int, [P, C.c_uint64, C.c_int32, P, P, C.c_uint32, P]),
        "pga_tessellate_device_count": (C.c_int, [P, P, C.c_uint64, C.c_int32, P, P, C.c_uint32, P]),
        "pga_grammar_code": (C.c_int, [P, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64)]),
        "pga_grammar_edges": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.POINTER(C.c_uint8))]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class Engine:
    """One derivation context (pga_context)."""

    def __init__(self, technique=abi.MEGAKERNEL, queue_capacity=0, block_size=0, max_levels=0,
                 grid_blocks=0, device=-1, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise PGAError(abi.ERR_HIP, "no GPU visible: the engine has no CPU path")
        self.torch = torch
        self.L = load_library()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        cfg = abi.Config(technique, queue_capacity, block_size, max_levels, grid_blocks, device,
                         C.c_void_p(stream))
        h = C.c_void_p()
        st = self.L.pga_create(C.byref(cfg), C.byref(h))
        self._h = h.value
        self._check(st)
        self._keep = {}
        self._ext = None
        self._capturing = False
        self.n_axioms = 0
        self.params_per_axiom = 0

    def close(self):
        """Destroys the context; graphs captured on it can then only be closed."""
        if getattr(self, "_h", None):
            self.L.pga_destroy(self._h)
            self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):   # bound now: module globals are gone at shutdown
        # at interpreter shutdown the HIP runtime may already be torn down:
        # the process releases the context's device memory and stream
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    # -- ordering against torch --
    # With no stream given, the library runs the context on a stream of its
    # own, which does not wait for torch's streams: device inputs that torch
    # wrote (dL/dinstance, device axioms, buffers) need an explicit order.
    # Results of the asynchronous calls are read after synchronize().
    def _after_torch(self):
        """The context stream waits for the work enqueued on torch's current
        stream, unless that stream is already idle."""
        if self._capturing:
            return
        cur = self.torch.cuda.current_stream()
        s = self.stream()
        if cur.cuda_stream == s or cur.query():
            return
        if self._ext is None:
            self._ext = self.torch.cuda.ExternalStream(s)
        self._ext.wait_stream(cur)

    def _check(self, st):
        if st != abi.OK:
            msg = self.L.pga_last_error(self._h).decode() if getattr(self, "_h", None) else ""
            raise PGAError(st, msg)

    # -- program --
    def load(self, program):
        procs, entries = program.arrays()
        self._keep["program"] = (procs, entries)
        self._check(self.L.pga_initialize(self._h, C.addressof(procs), len(program.procedures),
                                          C.addressof(entries), len(program.table.entries)))
        self.program = program
        return self

    # -- inputs --
    def set_axioms(self, axioms):
        """axioms: numpy AXIOM_DTYPE array (copied) or a torch CUDA tensor of
        80-byte records (uint8 [n, 80] or float32 [n, 20], not copied)."""
        torch = self.torch
        if isinstance(axioms, torch.Tensor):
            assert axioms.is_cuda and axioms.is_contiguous()
            n = axioms.numel() * axioms.element_size() // 80
            self._keep["axioms"] = axioms
            self._check(self.L.pga_set_axioms_device(self._h, C.c_void_p(axioms.data_ptr()), n))
        else:
            a = np.ascontiguousarray(axioms, dtype=abi.AXIOM_DTYPE)
            n = len(a)
            self._check(self.L.pga_set_axioms(self._h, a.ctypes.data, n))
        self.n_axioms = n
        return self

    def set_axiom_params(self, params):
        if hasattr(params, "detach"):
            params = params.detach().cpu().numpy()
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.L.pga_set_axiom_params(self._h, p.ctypes.data, p.shape[0], p.shape[1]))
        self.params_per_axiom = p.shape[1]
        return self

    def set_philox_key(self, key):
        self._check(self.L.pga_set_philox_key(self._h, int(key[0]), int(key[1])))
        return self

    def set_seeds(self, seeds):
        s = np.ascontiguousarray(seeds, dtype=np.float32)
        self._check(self.L.pga_set_seeds(self._h, s.ctypes.data, len(s)))

    def seeds(self):
        out = np.zeros(abi.MAX_NUM_AXIOM_SEEDS, dtype=np.float32)
        self._check(self.L.pga_get_seeds(self._h, out.ctypes.data))
        return out

    # -- outputs --
    def bind(self, terminal, capacity, buffer=None):
        """Binds a [capacity, 16] float32 CUDA tensor as terminal buffer
        (allocated when `buffer` is None). capacity 0 counts only."""
        torch = self.torch
        if capacity == 0:
            self._keep[("inst", terminal)] = None
            self._check(self.L.pga_bind_instances(self._h, terminal, None, 0))
            return None
        if buffer is None:
            buffer = torch.empty((capacity, 16), dtype=torch.float32, device="cuda")
        assert buffer.is_cuda and buffer.is_contiguous() and buffer.numel() * 4 >= capacity * 64
        self._keep[("inst", terminal)] = buffer
        self._check(self.L.pga_bind_instances(self._h, terminal, C.c_void_p(buffer.data_ptr()), capacity))
        return buffer

    def buffer(self, terminal):
        return self._keep.get(("inst", terminal))

    def counts(self, n=abi.MAX_TERMINALS):
        out = np.zeros(n, dtype=np.uint64)
        self._check(self.L.pga_get_counts(self._h, out.ctypes.data, n))
        return out

    # -- execution --
    def execute(self, n=None, first=0, append=False):
        """Derives axioms [first, first + n); returns device seconds."""
        self.torch.cuda.current_stream().synchronize()
        if n is None:
            n = self.n_axioms - first
        sec = C.c_double(0.0)
        self._check(self.L.pga_execute_range(self._h, first, n, 1 if append else 0, C.byref(sec)))
        return sec.value

    def execute_async(self, n=None, first=0, append=False):
        """Enqueues the derivation on the engine's stream without waiting
        (after the work torch has enqueued); read the results after
        synchronize()."""
        if n is None:
            n = self.n_axioms - first
        self._after_torch()
        self._check(self.L.pga_execute_async(self._h, first, n, 1 if append else 0))

    def synchronize(self):
        self._check(self.L.pga_synchronize(self._h))

    def kernel_seconds(self):
        sec = C.c_double(0.0)
        self._check(self.L.pga_kernel_seconds(self._h, C.byref(sec)))
        return sec.value

    # -- HIP graphs --
    def stream(self):
        """The context's HIP stream (pga_get_stream) as an integer handle:
        create an LBVH on it to record its device-count build in a graph."""
        h = C.c_void_p()
        self._check(self.L.pga_get_stream(self._h, C.byref(h)))
        return h.value or 0

    def device_counts(self):
        """Device address of the 64-bit terminal counters (pga_device_counts)."""
        h = C.c_void_p()
        self._check(self.L.pga_device_counts(self._h, C.byref(h)))
        return h.value

    @contextlib.contextmanager
    def capture(self):
        """Records the work enqueued on the context stream inside the block
        (execute_async, device-count LBVH calls on self.stream()) into a
        Graph instead of running it:

            with eng.capture() as g:
                eng.execute_async()
            g.launch(); eng.synchronize()

        Values in the bound buffers may change between replays (update them
        while no replay runs: after synchronize()); a change of bindings,
        program layout or configuration makes launch() refuse the graph."""
        g = Graph(self)
        self._after_torch()
        self._check(self.L.pga_capture_begin(self._h))
        self._capturing = True
        try:
            yield g
        except BaseException:
            self._capturing = False
            h = C.c_void_p()
            self.L.pga_capture_end(self._h, C.byref(h))
            if h.value:
                self.L.pga_graph_destroy(h.value)
            raise
        self._capturing = False
        h = C.c_void_p()
        self._check(self.L.pga_capture_end(self._h, C.byref(h)))
        g._h = h.value

    def specialize(self, block_size=0, enable=True):
        """Compiles the loaded program into its own derivation kernel
        (pga_specialize) and runs it from now on; returns the compile time
        in seconds (0 when this process compiled the program before)."""
        sec = C.c_double(0.0)
        self._check(self.L.pga_specialize(self._h, int(block_size), 1 if enable else 0, C.byref(sec)))
        return sec.value

    def set_schedule(self, local=True, grab=0, spill=0):
        """Level scheduling of the specialised kernel (pga_set_schedule):
        workgroup-local levels with `grab` axioms per hand-out (0: automatic)
        and levels of more than `spill` children (0: no limit) left to the
        grid, or every level over the grid (local=False)."""
        self._check(self.L.pga_set_schedule(self._h, 1 if local else 0, int(grab), int(spill)))
        return self

    def autotune(self, n=None, repeats=3):
        """Times every launch configuration on the bound axioms and keeps the
        fastest (pga_autotune); returns the chosen configuration."""
        r = abi.TuneResult()
        self._after_torch()
        self._check(self.L.pga_autotune(self._h, self.n_axioms if n is None else n, repeats, C.byref(r)))
        return {name: getattr(r, name) for name, _ in abi.TuneResult._fields_}

    # -- reverse mode --
    def tape_reserve(self, max_records, max_children):
        """Sizes the derivation tape: parent records of all levels, children
        of all levels counting the axioms."""
        self._check(self.L.pga_tape_reserve(self._h, int(max_records), int(max_children)))
        return self

    def forward_tape(self, n=None):
        """Derivation of axioms [0, n) recording the tape; returns seconds."""
        sec = C.c_double(0.0)
        self._after_torch()
        self._check(self.L.pga_forward_tape(self._h, self.n_axioms if n is None else n, C.byref(sec)))
        return sec.value

    def tape_generation(self):
        """Generation of the recorded tape (0: none valid); see
        pga_tape_generation."""
        g = C.c_uint64(0)
        self._check(self.L.pga_tape_generation(self._h, C.byref(g)))
        return g.value

    def tape_counts(self):
        """(parent records, children counting the axioms) of the recorded tape."""
        p, c = C.c_uint64(0), C.c_uint64(0)
        self._check(self.L.pga_tape_counts(self._h, C.byref(p), C.byref(c)))
        return p.value, c.value

    def backward(self, dl_dinstances, n=None):
        """dl_dinstances: {terminal: (count, 16) float32 cuda tensor}.
        Returns dL/dparams as an (n, params_per_axiom) cuda tensor."""
        torch = self.torch
        n = self.n_axioms if n is None else n
        ptrs = (C.c_void_p * abi.MAX_TERMINALS)()
        keep = []
        for t, g in dl_dinstances.items():
            if g is None:
                continue
            g = g.detach().contiguous().to(torch.float32)
            keep.append(g)
            ptrs[t] = g.data_ptr()
        grad = torch.empty((n, self.params_per_axiom), dtype=torch.float32, device="cuda")
        sec = C.c_double(0.0)
        self._after_torch()   # dL/dinstance comes from torch's stream
        self._check(self.L.pga_backward(self._h, C.byref(ptrs), C.c_void_p(grad.data_ptr()), C.byref(sec)))
        self.last_backward_seconds = sec.value
        return grad

    def enable_timestamps(self, capacity=4 * 64):
        """Phase timestamps of the derivation kernel (see pga_set_timestamps)."""
        torch = self.torch
        buf = torch.zeros(capacity, dtype=torch.int64, device="cuda")
        self._keep["ticks"] = buf
        self._check(self.L.pga_set_timestamps(self._h, C.c_void_p(buf.data_ptr()), capacity))
        return buf

    def phase_times(self):
        """Per-level (count, scan, write) durations in microseconds of the
        last run with timestamps enabled."""
        buf = self._keep.get("ticks")
        if buf is None:
            return []
        hz = self.L.pga_timestamp_hz(self._h)
        t = buf.cpu().numpy().reshape(-1, 4).astype(np.float64)
        out = []
        for lv in range(len(t)):
            if t[lv, 0] == 0:
                break
            d = np.diff(t[lv]) / hz * 1e6
            gap = (t[lv + 1, 0] - t[lv, 3]) / hz * 1e6 if lv + 1 < len(t) and t[lv + 1, 0] else 0.0
            out.append({"count_us": d[0], "scan_us": d[1], "write_us": d[2], "barrier_us": gap})
        return out

    def set_collider_capacity(self, capacity):
        """Colliders one execute may store (pga_set_collider_capacity, 0:
        the default of 2^20)."""
        self._check(self.L.pga_set_collider_capacity(self._h, int(capacity)))
        return self

    def colliders(self):
        """(colliders the last execute stored, phases of the program)."""
        n, p = C.c_uint32(), C.c_uint32()
        self._check(self.L.pga_num_colliders(self._h, C.byref(n), C.byref(p)))
        return n.value, p.value

    def stats(self):
        s = abi.Stats()
        self._check(self.L.pga_get_stats(self._h, C.byref(s)))
        return {"levels": s.levels, "records": s.records, "terminals": s.terminals, "seconds": s.seconds}

    def instances(self, terminal, count=None):
        """Copies the instances of `terminal` to host as an INSTANCE_DTYPE array."""
        buf = self.buffer(terminal)
        if count is None:
            count = int(self.counts(terminal + 1)[terminal])
        if buf is None or count == 0:
            return np.zeros(0, dtype=abi.INSTANCE_DTYPE)
        count = min(count, buf.shape[0])
        host = buf[:count].cpu().numpy()
        return host.view(abi.INSTANCE_DTYPE).reshape(count)


def specialize_check(program, block_size=0):
    """Compiles a program's specialised kernel without a device
    (pga_specialize_check); returns (status, compiler log, seconds)."""
    L = load_library()
    procs, entries = program.arrays()
    log = C.create_string_buffer(1 << 16)
    sec = C.c_double(0.0)
    st = L.pga_specialize_check(C.addressof(procs), len(program.procedures), C.addressof(entries),
                                len(program.table.entries), int(block_size), log, len(log), C.byref(sec))
    return st, log.value.decode(errors="replace"), sec.value


def derive(program, axioms, capacities, technique=abi.MEGAKERNEL, params=None, key=None, **cfg):
    """One-shot derivation: returns (list of instance arrays, stats)."""
    eng = Engine(technique=technique, **cfg)
    eng.load(program).set_axioms(axioms)
    if params is not None:
        eng.set_axiom_params(params)
    if key is not None:
        eng.set_philox_key(key)
    for t, cap in enumerate(capacities):
        eng.bind(t, cap)
    eng.execute()
    out = [eng.instances(t) for t in range(len(capacities))]
    st = eng.stats()
    eng.close()
    return out, st


class Graph:
    """A recorded sequence of derivations (pga_graph); see Engine.capture."""

    def __init__(self, engine):
        self.engine = engine
        self._h = None

    def launch(self):
        """Replays the recording on the context stream without waiting;
        Engine.synchronize() waits and reports errors and stats."""
        e = self.engine
        e._after_torch()
        e._check(e.L.pga_graph_launch(e._h, self._h))

    def info(self):
        """(kernel nodes, cooperative kernel nodes) of the graph."""
        k, c = C.c_uint32(0), C.c_uint32(0)
        self.engine._check(self.engine.L.pga_graph_info(self._h, C.byref(k), C.byref(c)))
        return k.value, c.value

    def close(self):
        if self._h:
            self.engine.L.pga_graph_destroy(self._h)
            self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):   # bound now: module globals are gone at shutdown
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
