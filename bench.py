#!/usr/bin/env python
"""bench.py - shape-grammar derivation throughput on one MI355X.

Workload (one step): batched derivation of the facade grammar over 1,024
seeded building lots (W in [10,40] m, D in [10,30] m, 3..20 floors; Philox
key 0x5EED0002) in one persistent launch, inputs resident in HBM, about
1.2M instanced terminals of 64 bytes per step.

Prints ONE JSON line. Beside the headline (batched terminals/s) it reports:
  roofline      - HBM roofline of the derivation kernel: algorithmic bytes
                  (64 B per instance written + 80 B per axiom read) over the
                  kernel's average duration (HIP events on its stream);
  plain         - the one-axiom-per-call path (pga_execute per lot), the
                  reference's executePGA usage, on a sample of lots;
  batched_16384 - the same grammar over 16,384 lots in one launch;
  graph         - the headline derivation recorded once in a HIP graph and
                  replayed every step;
  stochastic_65536_lbvh - per-shape Philox rule choice over 65,536 axioms,
                  then the LBVH over the terminal boxes;
  fwd_bwd_gradient_256 - tape forward, the fused centre/size loss kernel
                  and the adjoint over 256 parameterised lots (dL/dparams
                  per second);
  mixed_menger5_16384_lbvh - a Menger sponge of depth 5 and 16,384 lots in
                  one launch plus the LBVH over all terminals from the device
                  counts, one HIP graph per step;
  mixed_fwd_bwd_menger5_16384_lbvh - the same batch with parameterised lots:
                  tape forward, fused loss, adjoint and the LBVH rebuild per
                  step;
  triangle_soup_16384 - the facade grammar with triangle-soup terminals
                  (generation function 0) over 16,384 lots: derivation plus
                  one tessellation launch (4 vertices, 6 indices per quad)
                  per step, with the tessellation kernel's HBM roofline;
  context_city_16384 - the two-phase city block (Collider / IfCollides)
                  over 16,384 lots: phase 0 stores each building and its
                  four walls as colliders, the LBVH over them is built, and
                  phase 1 tests one prop per yard cell against them;
  cpu_baseline  - the CPU restatement (oracle) on the host, one thread.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "terminals/sec batched vs plain at 1 GPU; fwd+bwd grads/sec; % HBM peak"
HBM_PEAK_GBS = 8000.0
AXIOM_BYTES = 80
INSTANCE_BYTES = 64
CPU_THREADS_MAX = 16     # the host share of one GPU


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--lots", type=int, default=1024)
    p.add_argument("--big-lots", type=int, default=16384)
    p.add_argument("--plain-sample", type=int, default=256)
    p.add_argument("--stochastic-axioms", type=int, default=65536)
    p.add_argument("--gradient-lots", type=int, default=256)
    p.add_argument("--mixed-lots", type=int, default=16384)
    p.add_argument("--mixed-grad-lots", type=int, default=16384)
    p.add_argument("--mesh-lots", type=int, default=16384)
    p.add_argument("--context-lots", type=int, default=16384)
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--only-kernel", action="store_true",
                   help="run the timed batched steps only (for profilers)")
    p.add_argument("--spec-block", type=int, default=768,
                   help="workgroup size of the kernel specialised to the program")
    p.add_argument("--workload", choices=("facade", "stochastic", "context"), default="facade",
                   help="with --only-kernel: the facade lots (--lots), the stochastic batch "
                        "(--stochastic-axioms) or the context-sensitive city block (--context-lots)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="device-to-device copies of MIB MiB (scale HBM byte counters); alone, or after "
                        "the steps of --only-kernel")
    return p.parse_args()


def make_engine(torch, engine_mod, grammars, axioms, n_lots, key=None):
    from operatorgraph_amd import abi
    lots = axioms.facade_lots(n_lots)
    walls, windows = axioms.facade_lot_terminals(lots)
    T = int(walls.sum() + windows.sum())
    # largest level: one child descriptor per tile (= one per window)
    queue = max(1 << 16, int(windows.sum() * 1.05))
    eng = engine_mod.Engine(technique=abi.MEGAKERNEL, queue_capacity=queue)
    eng.load(grammars.facade())
    dev_axioms = torch.from_numpy(lots.view(np.uint8).copy()).cuda()
    eng.set_axioms(dev_axioms)
    eng.bind(0, int(walls.sum()))
    eng.bind(1, int(windows.sum()))
    tune = tune_engine(eng, 3)
    return eng, lots, T, tune


SPEC_BLOCK = 768   # workgroup size of the kernel specialised to the program (--spec-block)
# the kernel sources a traffic measurement under profiles/ was taken with
KERNEL_SOURCES = ["pga.h", "pga_device.h", "pga_ops.h", "pga_kernels.h", "pga_level.h", "pga_spec.h",
                  "pga_spec.cpp", "pga_kernels.hip"]


def tune_engine(eng, repeats):
    """Compiles the program's specialised kernel (pga_specialize), then lets
    the autotuner time it beside the interpreting kernels and keep the
    fastest configuration."""
    compile_s = eng.specialize(SPEC_BLOCK)
    tune = eng.autotune(repeats=repeats)
    tune["compile_seconds"] = compile_s
    return tune


def timed_steps(torch, eng, steps, technique):
    from operatorgraph_amd import abi
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        if technique == abi.MEGAKERNEL:
            eng.execute_async()
        else:
            eng.execute()     # per-level launches read each level's count on the host
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if technique == abi.MEGAKERNEL:
        eng.synchronize()     # raises on any device error of the last step
    return (t1 - t0) / steps


def graph_steps(torch, eng, steps):
    """The derivation recorded once in a HIP graph (pga_capture_begin) and
    replayed `steps` times back to back; returns (wall seconds per step,
    device seconds of the last replay)."""
    with eng.capture() as g:
        eng.execute_async()
    g.launch()
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eng.synchronize()
    dev = eng.kernel_seconds()
    g.close()
    return (t1 - t0) / steps, dev


def kernel_seconds(torch, eng, steps):
    times = []
    for _ in range(steps):
        eng.execute()
        times.append(eng.kernel_seconds())
    return float(np.mean(times)), times


def cpu_baseline(lots_fn, seconds, threads=1):
    from operatorgraph_amd import axioms, grammars
    from oracle import oracle as O
    prog = grammars.facade()
    lots = lots_fn()
    parts = np.array_split(np.arange(len(lots)), threads)
    done = []
    t0 = time.perf_counter()
    while True:
        counts = [0] * threads

        def work(i):
            sub = lots[parts[i]]
            r = O.derive(prog, sub)
            counts[i] = r.count(0) + r.count(1)

        ts = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
        s = time.perf_counter()
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        done.append((sum(counts), time.perf_counter() - s))
        if time.perf_counter() - t0 >= seconds:
            break
    rates = [c / dt for c, dt in done]
    return float(np.median(rates)), len(done), sum(c for c, _ in done)


def copy_calibration(torch, mib, launches):
    """Device-to-device copies of `mib` MiB: a kernel of known traffic to
    scale the HBM byte counters of a profiler."""
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(launches):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": launches}), flush=True)


LBVH_BYTES_PER_LEAF = 204   # 24 B box read + 8 B key write + 4 sort passes x 16 B + Karras 24 + 12 + refit 72


def stochastic_workload(torch, engine, grammars, axioms, n, steps):
    """65,536-axiom stochastic batch (per-shape Philox rule choice and split
    ratios, lognormal terminal counts), autotuned, then the LBVH over all
    terminal boxes."""
    from operatorgraph_amd.lbvh import LBVH
    ax, _ = axioms.stochastic_axioms(n)
    eng = engine.Engine(queue_capacity=1 << 23)
    eng.load(grammars.stochastic()).set_axioms(ax)
    eng.set_philox_key(axioms.STOCHASTIC_KEY)
    eng.bind(0, 0)                           # count-only run sizes the buffer
    eng.execute()
    T = int(eng.counts(1)[0])
    buf = eng.bind(0, T)
    tune = tune_engine(eng, 3)
    eng.execute()
    sec = timed_steps(torch, eng, steps, tune["technique"])
    ksec, _ = kernel_seconds(torch, eng, 3)
    lb = LBVH(T)
    boxes = lb.instance_aabbs(buf[:T], planar=False)
    bounds = lb.bounds(boxes)
    smin, smax = [v - 1.0 for v in bounds[:3]], [v + 1.0 for v in bounds[3:]]
    lb.build(boxes, smin, smax)
    times = []
    for _ in range(max(3, steps)):
        lb.build(boxes, smin, smax)
        times.append(lb.seconds())
    lsec = float(np.median(times))
    out = {"axioms": n, "terminals_per_step": T, "ms_per_step": round(sec * 1e3, 4),
           "terminals_per_s": round(T / sec), "kernel_ms": round(ksec * 1e3, 4),
           "hbm_frac": round((INSTANCE_BYTES * T + AXIOM_BYTES * n) / ksec / 1e9 / HBM_PEAK_GBS, 4),
           "kernel": kernel_name(tune), "autotune": tune_summary(tune),
           "lbvh": {"leaves": T, "ms": round(lsec * 1e3, 4), "leaves_per_s": round(T / lsec),
                    "algorithmic_bytes": LBVH_BYTES_PER_LEAF * T,
                    "hbm_frac": round(LBVH_BYTES_PER_LEAF * T / lsec / 1e9 / HBM_PEAK_GBS, 4)}}
    lb.close()
    eng.close()
    return out


def context_engine(engine, grammars, axioms, n_lots):
    from operatorgraph_amd import abi
    eng = engine.Engine(queue_capacity=1 << 22)
    eng.load(grammars.context_city(walls=True)).set_axioms(axioms.city_lots(n_lots))
    eng.set_collider_capacity(1 << 17)
    for t in range(4):
        eng.bind(t, 0)                       # count-only run sizes the buffers
    eng.execute()
    counts = [int(c) for c in eng.counts(4)]
    for t in range(4):
        eng.bind(t, max(counts[t], 1))
    tune = eng.autotune(repeats=3)
    eng.execute()
    return eng, counts, tune


def context_workload(torch, engine, grammars, axioms, n_lots, steps):
    """The two-phase city block with walls (grammars.context_city) over
    n_lots lots, autotuned over the interpreting kernels (programs with
    Collider / IfCollides are not specialised); the oracle's single-thread
    rate on 256 of the lots beside it."""
    from operatorgraph_amd import abi
    prog = grammars.context_city(walls=True)
    eng, counts, tune = context_engine(engine, grammars, axioms, n_lots)
    T = sum(counts)
    sec = timed_steps(torch, eng, steps, tune["technique"])
    colliders, phases = eng.colliders()
    eng.close()
    from oracle import oracle as O
    sample = axioms.city_lots(256)
    t0 = time.perf_counter()
    r = O.derive(prog, sample)
    cpu_sec = time.perf_counter() - t0
    cpu_T = sum(len(r.instances(t)) for t in range(4))
    return {"lots": n_lots, "phases": phases, "colliders": colliders, "props": counts[1] + counts[2],
            "props_blocked": counts[2], "terminals_per_step": T, "ms_per_step": round(sec * 1e3, 4),
            "terminals_per_s": round(T / sec), "queries_per_s": round((counts[1] + counts[2]) / sec),
            "autotune": {"technique": "MEGAKERNEL" if tune["technique"] == abi.MEGAKERNEL else "KERNELS",
                         "workgroup_size": tune["block_size"], "workgroups": tune["grid_blocks"],
                         "candidates": tune["candidates"]},
            "cpu_oracle": {"lots": 256, "terminals": cpu_T, "terminals_per_s": round(cpu_T / cpu_sec), "cores": 1}}


def houses_workload(torch, engine, grammars, axioms, n_lots, steps):
    """Gabled houses (grammars.gabled_houses: dynamic-polygon roofs, a
    triangular right prism split into gables and roof sides) over n_lots
    city lots, autotuned over the interpreting and the specialised kernels;
    the oracle's single-thread rate on 256 of the lots beside it."""
    prog = grammars.gabled_houses()
    lots = axioms.city_lots(n_lots)
    eng = engine.Engine(queue_capacity=1 << 23)
    eng.load(prog).set_axioms(lots)
    for t in range(3):
        eng.bind(t, 0)
    eng.execute()
    counts = [int(c) for c in eng.counts(3)]
    for t in range(3):
        eng.bind(t, counts[t])
    tune = tune_engine(eng, 3)
    sec = timed_steps(torch, eng, steps, tune["technique"])
    eng.close()
    T = sum(counts)
    from oracle import oracle as O
    sample = axioms.city_lots(256)
    t0 = time.perf_counter()
    r = O.derive(prog, sample)
    cpu_sec = time.perf_counter() - t0
    cpu_T = sum(len(r.instances(t)) for t in range(3))
    return {"lots": n_lots, "walls": counts[0], "gables": counts[1], "roof_sides": counts[2],
            "terminals_per_step": T, "ms_per_step": round(sec * 1e3, 4), "terminals_per_s": round(T / sec),
            "kernel": kernel_name(tune), "autotune": tune_summary(tune),
            "cpu_oracle": {"lots": 256, "terminals": cpu_T, "terminals_per_s": round(cpu_T / cpu_sec), "cores": 1}}


def context_steps(torch, engine, grammars, axioms, n_lots, steps, warmup):
    """The context-sensitive city block alone, `steps` timed executes (for
    profilers)."""
    eng, counts, tune = context_engine(engine, grammars, axioms, n_lots)
    for _ in range(warmup):
        eng.execute()
    sec = timed_steps(torch, eng, steps, tune["technique"])
    print(json.dumps({"workload": f"context_city_{n_lots}", "terminals_per_step": sum(counts),
                      "ms_per_step": sec * 1e3, "block_size": tune["block_size"]}), flush=True)
    eng.close()


def stochastic_steps(torch, engine, grammars, axioms, n, steps, warmup):
    """The stochastic batch's tuned derivation alone, `steps` timed launches
    (for profilers)."""
    ax, _ = axioms.stochastic_axioms(n)
    eng = engine.Engine(queue_capacity=1 << 23)
    eng.load(grammars.stochastic()).set_axioms(ax)
    eng.set_philox_key(axioms.STOCHASTIC_KEY)
    eng.bind(0, 0)
    eng.execute()
    T = int(eng.counts(1)[0])
    eng.bind(0, T)
    tune = tune_engine(eng, 3)
    for _ in range(warmup):
        eng.execute()
    sec = timed_steps(torch, eng, steps, tune["technique"])
    print(json.dumps({"workload": f"stochastic_{n}", "terminals_per_step": T, "ms_per_step": sec * 1e3,
                      "kernel": kernel_name(tune), "autotune": tune_summary(tune)}), flush=True)
    eng.close()


MENGER5_CUBES = 20 ** 5


def mixed_workload(torch, engine, grammars, axioms, n_lots, steps):
    """A Menger sponge of depth 5 (one Box axiom of size 81, IfSizeLess(X, 1):
    3.2M cubes) and n_lots facade lots derived by one persistent launch
    (grammars.combine), then the boxes of all terminals gathered from the
    device counts and the LBVH built over them with the scene box computed
    on the device; the whole iteration is one HIP graph, replayed per step."""
    from operatorgraph_amd import abi
    from operatorgraph_amd.lbvh import LBVH
    prog, starts = grammars.combine((grammars.facade(), 0), (grammars.menger_interpreter(1.0), 2))
    lots = axioms.facade_lots(n_lots)
    walls, windows = axioms.facade_lot_terminals(lots)
    both = np.concatenate([axioms.menger_axiom(81.0, starts[1]), lots])
    caps = [int(walls.sum()), int(windows.sum()), MENGER5_CUBES]
    T = sum(caps)
    eng = engine.Engine(technique=abi.MEGAKERNEL,
                        queue_capacity=int(windows.sum() * 1.05) + MENGER5_CUBES + (1 << 16))
    eng.load(prog)
    eng.set_axioms(torch.from_numpy(both.view(np.uint8).copy()).cuda())
    for t in range(3):
        eng.bind(t, caps[t])
    tune = tune_engine(eng, 2)
    eng.execute()
    got = [int(c) for c in eng.counts(3)]
    assert got == caps, (got, caps)
    dsec, _ = kernel_seconds(torch, eng, 3)
    bvh = LBVH(T, engine=eng)
    out = bvh.alloc()
    bvh.gather_engine(eng, 3, 0b011, out)
    bvh.build_device(out)
    times = []
    for _ in range(3):
        bvh.build_device(out)
        bvh.leaf_count()
        times.append(bvh.seconds())
    lsec = float(np.median(times))
    with eng.capture() as g:
        eng.execute_async()
        bvh.gather_engine(eng, 3, 0b011, out)
        bvh.build_device(out)
    g.launch()
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    eng.synchronize()
    gsec = eng.kernel_seconds()
    n, overflow = bvh.leaf_count()
    assert n == T and not overflow, (n, T, overflow)
    assert [int(c) for c in eng.counts(3)] == caps
    alg = (INSTANCE_BYTES + LBVH_BYTES_PER_LEAF) * T + AXIOM_BYTES * len(both)
    res = {"lots": n_lots, "menger_cubes": MENGER5_CUBES, "terminals_per_step": T,
           "ms_per_step": round(sec * 1e3, 4), "terminals_per_s": round(T / sec),
           "graph_device_ms": round(gsec * 1e3, 4), "derive_kernel_ms": round(dsec * 1e3, 4),
           "lbvh_ms": round(lsec * 1e3, 4),
           "hbm_frac": round(alg / gsec / 1e9 / HBM_PEAK_GBS, 4), "kernel": kernel_name(tune),
           "autotune": tune_summary(tune)}
    g.close()
    bvh.close()
    eng.close()
    return res


FACADE_MESH = """
axiom Box Lot;
terminal Wall (0, 0);
terminal Window (0, 1);
Lot    = CompSplit() { Discard() | Discard() | Facade };
Facade = Repeat(Y, 3, ADJUST_TO_FILL) Floor;
Floor  = Repeat(X, 2.5, ADJUST_TO_FILL) Tile;
Tile   = IfSizeLess(X, 1.6) { Wall | SubDiv(X) { -1: Wall | 1.2: Window | -1: Wall } };
"""
QUAD_MESH_BYTES = 48 + 4 * 36 + 6 * 4   # instance rows read; 4 vertices, 6 indices written


def mesh_workload(torch, engine, axioms, n_lots, steps):
    """The facade grammar with triangle-soup terminals (generation function
    0, compiled from grammar text) over n_lots lots. One step derives the
    lots into one instance buffer (walls, then windows) and tessellates it
    with one pga_tessellate launch into one vertex/index buffer, all on the
    engine's stream with no host synchronisation between steps."""
    from operatorgraph_amd import abi
    from operatorgraph_amd.compiler import compile_grammar
    from operatorgraph_amd.mesh import tessellate
    g = compile_grammar(FACADE_MESH, "facade_mesh")
    lots = axioms.facade_lots(n_lots)
    lots["entry_index"] = g.axioms[0][0]
    walls, windows = axioms.facade_lot_terminals(lots)
    nw, nq = int(walls.sum()), int(windows.sum())
    n = nw + nq
    eng = engine.Engine(technique=abi.MEGAKERNEL, queue_capacity=int(nq * 1.05) + (1 << 16))
    eng.load(g.program)
    eng.set_axioms(torch.from_numpy(lots.view(np.uint8).copy()).cuda())
    inst = torch.empty((n, 16), dtype=torch.float32, device="cuda")
    eng.bind(0, nw, buffer=inst[:nw])
    eng.bind(1, nq, buffer=inst[nw:])
    tune = tune_engine(eng, 2)
    eng.execute()
    assert [int(c) for c in eng.counts(2)] == [nw, nq]
    verts = torch.empty((4 * n, 9), dtype=torch.float32, device="cuda")
    idx = torch.empty((6 * n,), dtype=torch.int32, device="cuda")
    ext = torch.cuda.ExternalStream(eng.stream())
    with torch.cuda.stream(ext):
        tessellate(inst, abi.QUAD, out=(verts, idx))
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        reps = max(3, steps)
        e[0].record(ext)
        for _ in range(reps):
            tessellate(inst, abi.QUAD, out=(verts, idx))
        e[1].record(ext)
        torch.cuda.synchronize()
        tsec = e[0].elapsed_time(e[1]) / 1e3 / reps
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.execute_async()
            tessellate(inst, abi.QUAD, out=(verts, idx))
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
    eng.synchronize()
    assert [int(c) for c in eng.counts(2)] == [nw, nq]
    assert int(idx[-1]) == 4 * n - 1
    # the same step with the counts read on the device, recorded once as a
    # HIP graph: derivation, then one tessellation launch per terminal type
    from operatorgraph_amd.mesh import tessellate_terminal
    outs = [(verts[: 4 * nw], idx[: 6 * nw]), (verts[4 * nw:], idx[6 * nw:])]
    with eng.capture() as gr:
        eng.execute_async()
        tessellate_terminal(eng, 0, abi.QUAD, out=outs[0])
        tessellate_terminal(eng, 1, abi.QUAD, base_vertex=4 * nw, out=outs[1])
    gr.launch()
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        gr.launch()
    torch.cuda.synchronize()
    gsec = (time.perf_counter() - t0) / steps
    eng.synchronize()
    gr.close()
    assert [int(c) for c in eng.counts(2)] == [nw, nq]
    assert int(idx[-1]) == 4 * n - 1
    alg = QUAD_MESH_BYTES * n
    res = {"lots": n_lots, "quads_per_step": n, "triangles_per_step": 2 * n, "ms_per_step": round(sec * 1e3, 4),
           "triangles_per_s": round(2 * n / sec),
           "graph": {"ms_per_step": round(gsec * 1e3, 4), "triangles_per_s": round(2 * n / gsec),
                     "note": "derivation + device-count tessellation per terminal type, one HIP graph"},
           "tessellate": {"kernel": "k_tessellate", "kernel_ms": round(tsec * 1e3, 4),
                          "algorithmic_bytes_per_launch": alg, "bound": "hbm",
                          "achieved": round(alg / tsec / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(alg / tsec / 1e9 / HBM_PEAK_GBS, 4)},
           "kernel": kernel_name(tune)}
    eng.close()
    return res


def mixed_gradient_workload(torch, engine, grammars, axioms, n_lots, steps):
    """The mixed batch forward and backward: a Menger sponge of depth 5 and
    n_lots parameterised facade lots (8 parameters each) in one tape forward,
    the MSE of instance centres and sizes against targets from perturbed
    split parameters, the adjoint back to dL/dparams, and the LBVH over all
    terminals rebuilt from the device counts, every iteration."""
    from operatorgraph_amd import abi
    from operatorgraph_amd.autograd import center_size_mse, centers_and_sizes, derive_instances
    from operatorgraph_amd.lbvh import LBVH
    prog, starts = grammars.combine((grammars.facade_params(), 0), (grammars.menger_interpreter(1.0), 2))
    lots, lparams = axioms.param_lots(n_lots)
    both = np.concatenate([axioms.menger_axiom(81.0, starts[1]), lots])
    params = np.zeros((len(both), lparams.shape[1]), dtype=np.float32)
    params[1:] = lparams                     # the sponge reads no parameter
    rng = np.random.default_rng(4)
    tparams = params.copy()
    tparams[1:, [0, 1, 4, 5, 6, 7]] += rng.normal(0, 0.05, size=(n_lots, 6)).astype(np.float32)
    # the widest level: three children per tile (about 19M for 16,384 lots)
    eng = engine.Engine(technique=abi.MEGAKERNEL, queue_capacity=1 << 25)
    eng.load(prog).set_axioms(torch.from_numpy(both.view(np.uint8).copy()).cuda()).set_axiom_params(tparams)
    eng.bind(0, 0)                           # count-only run sizes the buffers
    eng.bind(1, 0)
    eng.bind(2, 0)
    eng.execute()
    caps = [int(c) for c in eng.counts(3)]
    assert caps[2] == MENGER5_CUBES, caps
    for t in range(3):
        eng.bind(t, caps[t])
    eng.tape_reserve(1 << 25, 1 << 26)      # every fan-out's record; every child of every level
    eng.forward_tape()
    targets = [tuple(x.contiguous() for x in centers_and_sizes(eng.buffer(t)[:caps[t]].clone())) for t in range(3)]
    T = sum(caps)
    bvh = LBVH(T, engine=eng)
    out = bvh.alloc()
    p = torch.from_numpy(params).cuda()

    def step():
        q = p.clone().requires_grad_(True)
        outs = derive_instances(q, eng, (0, 1, 2))
        loss = 0.0
        for inst, (ct, st) in zip(outs, targets):
            loss = loss + center_size_mse(inst, ct, st)   # one kernel: loss and dL/dinstance
        loss.backward()
        torch.cuda.synchronize()
        bvh.gather_engine(eng, 3, 0b011, out)
        bvh.build_device(out)
        return q.grad

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd = bwd = lb = 0.0
    for _ in range(steps):
        step()
        fwd += eng.stats()["seconds"]
        bwd += eng.last_backward_seconds
        n, overflow = bvh.leaf_count()
        lb += bvh.seconds()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    assert n == T and not overflow, (n, T)
    _, S = eng.tape_counts()
    res = {"lots": n_lots, "menger_cubes": MENGER5_CUBES, "params": n_lots * lparams.shape[1], "terminals": T,
           "ms_per_step": round(sec * 1e3, 4), "grads_per_s": round(n_lots * lparams.shape[1] / sec),
           "terminals_per_s": round(T / sec), "forward_tape_ms": round(fwd / steps * 1e3, 4),
           "backward_ms": round(bwd / steps * 1e3, 4), "lbvh_ms": round(lb / steps * 1e3, 4),
           "roofline": fwd_bwd_roofline(T, len(both), S, len(both) * lparams.shape[1], (fwd + bwd) / steps)}
    bvh.close()
    eng.close()
    return res


def gradient_workload(torch, engine, grammars, axioms, n, steps):
    """256 parameterised facade lots x 8 parameters: forward with tape, MSE of
    instance centres and sizes against targets from perturbed split
    parameters (same topology), backward to dL/dparams."""
    from operatorgraph_amd.autograd import center_size_mse, centers_and_sizes, derive_instances
    ax, params = axioms.param_lots(n)
    rng = np.random.default_rng(4)
    tparams = params.copy()
    tparams[:, [0, 1, 4, 5, 6, 7]] += rng.normal(0, 0.05, size=(n, 6)).astype(np.float32)
    prog = grammars.facade_params()
    eng = engine.Engine(queue_capacity=1 << 21)
    eng.load(prog).set_axioms(ax).set_axiom_params(tparams)
    eng.bind(0, 1 << 21)
    eng.bind(1, 1 << 21)
    eng.tape_reserve(1 << 21, 1 << 22)
    # targets: the tape forward with the perturbed parameters (the same
    # topology, so the same instance order as the fitted run)
    eng.forward_tape()
    counts = eng.counts(2)
    targets = [tuple(x.contiguous() for x in centers_and_sizes(eng.buffer(t)[: int(counts[t])].clone()))
               for t in range(2)]
    p = torch.from_numpy(params).cuda()

    def step():
        q = p.clone().requires_grad_(True)
        outs = derive_instances(q, eng, (0, 1))
        loss = 0.0
        for inst, (ct, st) in zip(outs, targets):
            loss = loss + center_size_mse(inst, ct, st)   # one kernel: loss and dL/dinstance
        loss.backward()
        return q.grad

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd = bwd = 0.0
    for _ in range(steps):
        step()
        fwd += eng.stats()["seconds"]
        bwd += eng.last_backward_seconds
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    T = int(eng.stats()["terminals"])
    _, S = eng.tape_counts()
    out = {"lots": n, "params": n * params.shape[1], "terminals": T, "ms_per_step": round(sec * 1e3, 4),
           "grads_per_s": round(n * params.shape[1] / sec),
           "forward_tape_ms": round(fwd / steps * 1e3, 4), "backward_ms": round(bwd / steps * 1e3, 4),
           "roofline": fwd_bwd_roofline(T, n, S, n * params.shape[1], (fwd + bwd) / steps)}
    eng.close()
    return out


def fwd_bwd_roofline(T, A, S, P, seconds):
    """HBM roofline of a tape forward plus its backward: algorithmic bytes
    64 T + 80 A (instances written, axioms read), 8 S written and 8 S read
    for the tape (S shapes recorded), 48 T of dL/dinstance read (the 12 model
    entries), 64 T recomputed transforms (one checkpoint per depth), 4 P
    parameter gradients written; over the device time of both calls."""
    b = INSTANCE_BYTES * T + AXIOM_BYTES * A + 16 * S + 48 * T + 64 * T + 4 * P
    achieved = b / seconds / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": b,
            "shapes_on_tape": S, "device_ms": round(seconds * 1e3, 4),
            "formula": "64T + 80A + 16S + 48T + 64T + 4P over forward-tape + backward device time"}


def cpu_info():
    """Host CPU model and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0)), os.cpu_count()


def kernel_name(tune):
    from operatorgraph_amd import abi
    if tune.get("specialized"):
        return f"pga_spec_mega<{tune['block_size']}>"
    k = "k_mega" if tune["technique"] == abi.MEGAKERNEL else "k_level"
    return f"pga::{k}<{tune['block_size']}>"


def tune_summary(tune):
    from operatorgraph_amd import abi
    tech = "MEGAKERNEL (cooperative persistent launch)" if tune["technique"] == abi.MEGAKERNEL \
        else "KERNELS (one cooperative launch per level)"
    if tune.get("specialized"):
        tech = "MEGAKERNEL specialised to the program (hipRTC)"
    return {"technique": tech, "workgroup_size": tune["block_size"], "workgroups": tune["grid_blocks"],
            "candidates": tune["candidates"], "best_ms": round(tune["seconds"] * 1e3, 4),
            "untuned_ms": round(tune["initial_seconds"] * 1e3, 4),
            "specialized": bool(tune.get("specialized")),
            "level_schedule": (f"workgroup-local levels, {tune['local_grab']} axioms per hand-out"
                               + (f", levels over {tune['local_spill']} children to the grid"
                                  if tune.get("local_spill") else "")
                               if tune.get("specialized") and tune.get("local_grab") else "grid-wide levels"),
            "compile_s": round(tune.get("compile_seconds", 0.0), 3)}


def load_traffic(workload, kernel=None):
    """HBM bytes per launch from the committed counter summary (the entry of
    `kernel`, else the summary's main kernel) and whether the summary was
    measured with the kernel sources of this tree (their digest matches)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    value = d.get("per_kernel", {}).get(kernel, d.get("traffic_bytes_per_launch"))
    return value, d.get("kernel_sources_sha256") == kernel_sources_digest()


def kernel_sources_digest():
    """SHA-256 over the derivation kernel's sources, in KERNEL_SOURCES order."""
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        for d in (os.path.join(ROOT, "operatorgraph_amd", "csrc"), os.path.join(ROOT, "include")):
            path = os.path.join(d, name)
            if os.path.exists(path):
                with open(path, "rb") as f:
                    h.update(f.read())
                break
    return h.hexdigest()


def main():
    global SPEC_BLOCK
    args = parse()
    SPEC_BLOCK = args.spec_block
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the derivation shards by axiom with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_calibration and not args.only_kernel:
        copy_calibration(torch, args.copy_calibration, args.steps)
        return

    from operatorgraph_amd import abi, axioms, engine, grammars

    if args.only_kernel and args.workload == "context":
        context_steps(torch, engine, grammars, axioms, args.context_lots, args.steps, args.warmup)
        return
    if args.only_kernel and args.workload == "stochastic":
        stochastic_steps(torch, engine, grammars, axioms, args.stochastic_axioms, args.steps, args.warmup)
        return
    eng, lots, T, tune = make_engine(torch, engine, grammars, axioms, args.lots)
    A = len(lots)
    for _ in range(args.warmup):
        eng.execute()
    assert eng.stats()["terminals"] == T, (eng.stats(), T)
    sec = timed_steps(torch, eng, args.steps, tune["technique"])
    value = T / sec
    if args.only_kernel:
        print(json.dumps({"terminals_per_step": T, "ms_per_step": sec * 1e3, "kernel": kernel_name(tune)}),
              flush=True)
        eng.close()
        if args.copy_calibration:
            copy_calibration(torch, args.copy_calibration, args.steps)
        return
    ksec, _ = kernel_seconds(torch, eng, max(5, args.steps // 2))
    graph = None
    if tune["technique"] == abi.MEGAKERNEL:
        gsec, gdev = graph_steps(torch, eng, args.steps)
        graph = {"ms_per_step": round(gsec * 1e3, 4), "terminals_per_s": round(T / gsec),
                 "device_ms": round(gdev * 1e3, 4)}
    alg_bytes = INSTANCE_BYTES * T + AXIOM_BYTES * A
    traffic, traffic_current = load_traffic(f"facade_lots_{args.lots}", kernel_name(tune))
    achieved = alg_bytes / ksec / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_measured_on_these_sources": traffic_current,
                "kernel": kernel_name(tune), "kernel_ms": round(ksec * 1e3, 4),
                "algorithmic_bytes_per_launch": alg_bytes}

    # plain: one axiom per call (synchronous pga_execute_range per lot)
    n_plain = min(args.plain_sample, A)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_T = 0
    for i in range(n_plain):
        eng.execute(1, first=i)
        plain_T += int(eng.stats()["terminals"])
    plain_sec = time.perf_counter() - t0
    plain_rate = plain_T / plain_sec
    eng.close()

    big = None
    if args.big_lots:
        beng, blots, BT, btune = make_engine(torch, engine, grammars, axioms, args.big_lots)
        for _ in range(2):
            beng.execute()
        bsec = timed_steps(torch, beng, max(3, args.steps // 4), btune["technique"])
        bk, _ = kernel_seconds(torch, beng, 3)
        bbytes = INSTANCE_BYTES * BT + AXIOM_BYTES * len(blots)
        big = {"lots": len(blots), "terminals_per_step": BT, "ms_per_step": round(bsec * 1e3, 4),
               "terminals_per_s": round(BT / bsec), "kernel_ms": round(bk * 1e3, 4),
               "hbm_frac": round(bbytes / bk / 1e9 / HBM_PEAK_GBS, 4),
               "vs_plain": round((BT / bsec) / plain_rate, 1), "kernel": kernel_name(btune),
               "autotune": tune_summary(btune)}
        beng.close()

    stoch = None
    if args.stochastic_axioms:
        stoch = stochastic_workload(torch, engine, grammars, axioms, args.stochastic_axioms,
                                    max(3, args.steps // 4))

    mixed = None
    if args.mixed_lots:
        mixed = mixed_workload(torch, engine, grammars, axioms, args.mixed_lots, max(3, args.steps // 4))

    mixed_grad = None
    if args.mixed_grad_lots:
        mixed_grad = mixed_gradient_workload(torch, engine, grammars, axioms, args.mixed_grad_lots,
                                             max(3, args.steps // 4))

    mesh = None
    if args.mesh_lots:
        mesh = mesh_workload(torch, engine, axioms, args.mesh_lots, max(3, args.steps // 4))

    context = None
    if args.context_lots:
        context = context_workload(torch, engine, grammars, axioms, args.context_lots, max(3, args.steps // 4))

    houses = None
    if args.context_lots:
        houses = houses_workload(torch, engine, grammars, axioms, args.context_lots, max(3, args.steps // 4))

    grad = None
    if args.gradient_lots:
        grad = gradient_workload(torch, engine, grammars, axioms, args.gradient_lots, max(5, args.steps // 2))

    cpu = None
    if not args.no_cpu:
        model, usable, online = cpu_info()
        threads = max(1, min(usable, CPU_THREADS_MAX))
        rate, runs, total = cpu_baseline(lambda: axioms.facade_lots(args.lots), args.cpu_seconds, threads)
        rate1, runs1, total1 = cpu_baseline(lambda: axioms.facade_lots(args.lots), args.cpu_seconds, 1)
        cpu = {"value": round(rate), "unit": "terminals/s", "cores": threads, "kind": "port",
               "sample": f"facade grammar over the same {args.lots} lots, {runs} full derivations "
                         f"({total} terminals), the lots split over {threads} host threads each running "
                         f"the oracle (g++ -O3 -march=x86-64-v3 -ffp-contract=off), median rate",
               "cpu_model": model, "cores_usable": usable, "cores_online": online,
               "single_thread": {"value": round(rate1), "cores": 1, "runs": runs1, "terminals": total1}}

    out = {
        "metric": METRIC, "value": round(value), "unit": "terminals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"facade_lots_{args.lots}", "lots": A, "terminals_per_step": T,
                   "grammar": "Lot=CompSplit{Discard|Discard|Facade}; Facade=Repeat(Y,3); "
                              "Floor=Repeat(X,2.5); Tile=IfSizeLess(X,1.6){Wall|SubDiv(X){-1|1.2|-1}}",
                   "autotune": tune_summary(tune)},
        "roofline": roofline,
        "plain": {"lots": n_plain, "terminals": plain_T, "terminals_per_s": round(plain_rate),
                  "ms_per_lot": round(plain_sec / n_plain * 1e3, 4)},
        "batched_vs_plain": round(value / plain_rate, 1),
        "graph": graph,
        "batched_16384": big,
        "stochastic_65536_lbvh": stoch,
        "fwd_bwd_gradient_256": grad,
        "mixed_menger5_16384_lbvh": mixed,
        "mixed_fwd_bwd_menger5_16384_lbvh": mixed_grad,
        "triangle_soup_16384": mesh,
        "context_city_16384": context,
        "gabled_houses_16384": houses,
        "cpu_baseline": cpu,
    }
    torch.cuda.synchronize()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
