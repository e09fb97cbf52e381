"""Terminal consumers: the triangle-soup geometry of instance buffers
(generation function 0, pga_tessellate) and an OBJ writer.

    vertices, indices = tessellate(instances, abi.BOX)   # (n*24, 9), (n*36,)
    write_obj("out.obj", vertices, indices)

`instances` is an (n, 16) float32 cuda tensor of instance records (an
engine's buffer). Vertices are the reference's VertexAttributes (position
xyzw, normal xyz, uv), 24 per box and 4 per quad; indices are uint32 (held
in int32 tensors), triangles (0, 1, 2) and (0, 2, 3) per face of four
vertices, the reference's default winding. The OBJ writer
plays the part of the reference's OBJExporter (src/rendering/OBJExporter.cpp):
positions, normals and uvs with one face line per triangle.
"""
import ctypes as C

import numpy as np

from . import abi


def tessellate(instances, shape_type, base_vertex=0, out=None):
    """Runs pga_tessellate on the current torch stream; returns
    (vertices (n * V, 9) float32, indices (n * 1.5 V,) int32) cuda tensors,
    written into `out` = (vertices, indices) when given."""
    import torch

    from .engine import PGAError, load_library
    inst = instances.contiguous()
    if inst.dtype != torch.float32 or inst.dim() != 2 or inst.shape[1] != 16:
        raise ValueError("instances must be an (n, 16) float32 tensor")
    n = inst.shape[0]
    v = 4 if shape_type == abi.QUAD else 24
    if out is None:
        verts = torch.empty((n * v, 9), dtype=torch.float32, device=inst.device)
        idx = torch.empty((n * (v + v // 2),), dtype=torch.int32, device=inst.device)
    else:
        verts, idx = out
        if (verts.dtype != torch.float32 or tuple(verts.shape) != (n * v, 9) or not verts.is_contiguous()
                or idx.dtype != torch.int32 or tuple(idx.shape) != (n * (v + v // 2),) or not idx.is_contiguous()):
            raise ValueError(f"out must be contiguous ({n * v}, 9) float32 and ({n * (v + v // 2)},) int32 tensors")
    if n == 0:
        return verts, idx
    st = load_library().pga_tessellate(C.c_void_p(inst.data_ptr()), n, int(shape_type), C.c_void_p(verts.data_ptr()),
                                       C.c_void_p(idx.data_ptr()), int(base_vertex),
                                       C.c_void_p(torch.cuda.current_stream(inst.device).cuda_stream))
    if st != abi.OK:
        raise PGAError(st, "pga_tessellate")
    return verts, idx


def tessellate_terminal(engine, terminal, shape_type, base_vertex=0, out=None):
    """Tessellates a terminal type's instance buffer with the count read on
    the device (pga_tessellate_device_count on the engine's counter), on the
    engine's stream: no host synchronisation after the derivation, and the
    call can be recorded with the derivation in one `engine.capture()`.
    The buffers are sized for the bound capacity; after synchronising, the
    first engine.counts()[terminal] shapes of them hold the mesh."""
    import torch

    from .engine import PGAError
    buf = engine.buffer(terminal)
    if buf is None:
        raise ValueError(f"terminal {terminal} has no bound instance buffer")
    cap = buf.shape[0]
    v = 4 if shape_type == abi.QUAD else 24
    if out is None:
        verts = torch.empty((cap * v, 9), dtype=torch.float32, device=buf.device)
        idx = torch.empty((cap * (v + v // 2),), dtype=torch.int32, device=buf.device)
    else:
        verts, idx = out
        if (verts.dtype != torch.float32 or verts.shape[0] < cap * v or not verts.is_contiguous()
                or idx.dtype != torch.int32 or idx.shape[0] < cap * (v + v // 2) or not idx.is_contiguous()):
            raise ValueError("out buffers are too small for the bound capacity")
    counts = engine.device_counts() + 8 * terminal
    st = engine.L.pga_tessellate_device_count(C.c_void_p(buf.data_ptr()), C.c_void_p(counts), cap, int(shape_type),
                                              C.c_void_p(verts.data_ptr()), C.c_void_p(idx.data_ptr()),
                                              int(base_vertex), C.c_void_p(engine.stream()))
    if st != abi.OK:
        raise PGAError(st, "pga_tessellate_device_count")
    return verts, idx


def write_obj(path, vertices, indices, name="pga"):
    """Writes an OBJ file: v (xyz), vn, vt per vertex, one `f a/a/a b/b/b
    c/c/c` line per triangle (1-based). Accepts numpy arrays or tensors."""
    v = np.asarray(vertices.cpu() if hasattr(vertices, "cpu") else vertices, dtype=np.float32).reshape(-1, 9)
    i = np.asarray(indices.cpu() if hasattr(indices, "cpu") else indices).astype(np.int64).reshape(-1, 3) + 1
    with open(path, "w", encoding="ascii") as f:
        f.write(f"# {len(v)} vertices, {len(i)} triangles\no {name}\n")
        np.savetxt(f, v[:, 0:3], fmt="v %.9g %.9g %.9g")
        np.savetxt(f, v[:, 4:7], fmt="vn %.9g %.9g %.9g")
        np.savetxt(f, v[:, 7:9], fmt="vt %.9g %.9g")
        np.savetxt(f, np.repeat(i, 3, axis=1), fmt="f %d/%d/%d %d/%d/%d %d/%d/%d")


def read_obj(path):
    """(positions (n, 3), normals (n, 3), uvs (n, 2), triangles (m, 3), 0-based)
    of an OBJ file written by write_obj."""
    v, vn, vt, tri = [], [], [], []
    with open(path, encoding="ascii") as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                v.append([float(x) for x in parts[1:4]])
            elif parts[0] == "vn":
                vn.append([float(x) for x in parts[1:4]])
            elif parts[0] == "vt":
                vt.append([float(x) for x in parts[1:3]])
            elif parts[0] == "f":
                tri.append([int(p.split("/")[0]) - 1 for p in parts[1:4]])
    return (np.array(v, dtype=np.float32), np.array(vn, dtype=np.float32), np.array(vt, dtype=np.float32),
            np.array(tri, dtype=np.int64))
