// pga_mesh.hip - generation function 0 (triangle soup) over instance
// records (gfx950).
//
// The reference's GenerationFunction 0 writes, per Box, 24 vertices
// (four per face: position, face normal, uv) and 36 indices, and per Quad
// 4 vertices and 6 indices (ShapeGenerator<Box, false>::run,
// include/pga/core/ShapeGenerator.cuh:376-511; <Quad, false>:670-708, the
// default winding with PGA_CW undefined). A corner is model4 * (+-size/2, 1).
// The instance record holds I = model4 * scale(size) (rows 0..2), and
//   m_ij * (s_j * 0.5) == (m_ij * s_j) * 0.5
// in FP32 (a power-of-two scale commutes with rounding), so a corner is
// I * (+-0.5, 1), with the same operation order as the reference's
// matrix-vector product: ((a x + b y) + c z) + d w. The result equals the
// reference's vertex values; the only exception is the sign of an exact zero.
//
// A workgroup covers a tile of kTileVerts vertices (32 boxes or 192
// quads). Each thread builds whole vertices into LDS, rebuilding its face's
// corners 0..2 for the normal; after one barrier the workgroup copies the
// tile out with 16-byte stores that cover consecutive addresses (the
// vertex records are 36 bytes, so a store straight from the building
// thread would scatter a wave's addresses over 2.3 KB) and writes the
// tile's indices as consecutive dwords. Every output store is nontemporal:
// the mesh is written once and read by a later consumer, and streaming it
// past the caches measured 3.5 -> 6.0 TB/s on 2^20 boxes. Shapes keep the
// order of the instance buffer; the reference's atomics leave it unspecified.
#include <hip/hip_runtime.h>

#include "../../include/pga.h"

namespace pgamesh {

constexpr int kBlock = 256;
constexpr uint32_t kTileVerts = 768;                 // 32 boxes or 192 quads
constexpr uint32_t kTileFloats = kTileVerts * 9u;   // 27,648 bytes of LDS
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void corner(const float4& r0, const float4& r1, const float4& r2, float x, float y, float z,
                                       float o[3]) {
    o[0] = ((r0.x * x + r0.y * y) + r0.z * z) + r0.w * 1.0f;
    o[1] = ((r1.x * x + r1.y * y) + r1.z * z) + r1.w * 1.0f;
    o[2] = ((r2.x * x + r2.y * y) + r2.z * z) + r2.w * 1.0f;
}

// corner signs (x, y, z) of corner c of face f (0..5: top, bottom, left,
// right, front, back; 6: a quad, z = 0), in half-size units
__constant__ int8_t kCorners[7][4][3] = {
    {{-1, 1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, 1}},       // top
    {{-1, -1, 1}, {1, -1, 1}, {1, -1, -1}, {-1, -1, -1}},   // bottom
    {{-1, 1, -1}, {-1, 1, 1}, {-1, -1, 1}, {-1, -1, -1}},   // left
    {{1, 1, 1}, {1, 1, -1}, {1, -1, -1}, {1, -1, 1}},       // right
    {{-1, 1, 1}, {1, 1, 1}, {1, -1, 1}, {-1, -1, 1}},       // front
    {{1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, -1}},   // back
    {{-1, 1, 0}, {1, 1, 0}, {1, -1, 0}, {-1, -1, 0}}};      // quad

__device__ __forceinline__ void face_corner(uint32_t f, uint32_t c, float& x, float& y, float& z) {
    x = 0.5f * (float)kCorners[f][c][0];
    y = 0.5f * (float)kCorners[f][c][1];
    z = 0.5f * (float)kCorners[f][c][2];
}

// vertex k (0..V-1) of a shape with instance rows r0..r2, as nine floats
__device__ __forceinline__ void build_vertex(const float4& r0, const float4& r1, const float4& r2, uint32_t quad,
                                             uint32_t k, float* o) {
    const uint32_t f = quad ? 6u : k / 4u, c = k & 3u;
    float v[3][3], p[3], x, y, z;
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) {
        face_corner(f, q, x, y, z);
        corner(r0, r1, r2, x, y, z, v[q]);
    }
    face_corner(f, c, x, y, z);
    corner(r0, r1, r2, x, y, z, p);
    const float a[3] = {v[0][0] - v[1][0], v[0][1] - v[1][1], v[0][2] - v[1][2]};
    const float b[3] = {v[2][0] - v[1][0], v[2][1] - v[1][1], v[2][2] - v[1][2]};
    const float n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const float len = sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    const float inv = 1.0f / len;
    o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = ((0.0f * x + 0.0f * y) + 0.0f * z) + 1.0f * 1.0f;
    o[4] = n[0] * inv; o[5] = n[1] * inv; o[6] = n[2] * inv;
    o[7] = (c == 0u || c == 3u) ? 1.0f : 0.0f;   // uv0..uv3 = (1,1), (0,1), (0,0), (1,0)
    o[8] = c < 2u ? 1.0f : 0.0f;
}

// vec4: vertices is 16-byte aligned (the tile copy then uses 16-byte stores).
// dcount: when not null, the shape count is min(*dcount, count), read on the
// device (the grid covers `count` shapes; tiles past the count return).
__global__ void __launch_bounds__(kBlock) k_tessellate(const float4* __restrict__ inst, uint64_t count, uint32_t quad,
                                                       float* __restrict__ vert, uint32_t* __restrict__ idx,
                                                       uint32_t base_vertex, uint32_t vec4,
                                                       const unsigned long long* __restrict__ dcount) {
    __shared__ __align__(16) float tile[kTileFloats];
    if (dcount) count = min((uint64_t)*dcount, count);
    const uint32_t V = quad ? 4u : 24u;
    const uint32_t tile_shapes = kTileVerts / V;
    const uint64_t s0 = (uint64_t)blockIdx.x * tile_shapes;
    if (s0 >= count) return;
    const uint32_t ns = (uint32_t)min((uint64_t)tile_shapes, count - s0);
    const uint32_t nv = ns * V;
    for (uint32_t v = threadIdx.x; v < nv; v += kBlock) {
        const uint32_t ls = v / V, k = v - ls * V;
        const float4* r = inst + (s0 + ls) * 4u;
        build_vertex(r[0], r[1], r[2], quad, k, tile + v * 9u);
    }
    __syncthreads();
    const uint64_t g0 = (s0 * V) * 9u;                 // first float of the tile in vertices
    const uint32_t nf = nv * 9u;                       // a multiple of 4, as V is
    if (vec4) {
        f4* out = reinterpret_cast<f4*>(vert + g0);
        const f4* src = reinterpret_cast<const f4*>(tile);
        for (uint32_t q = threadIdx.x; q < nf / 4u; q += kBlock) __builtin_nontemporal_store(src[q], out + q);
    } else {
        for (uint32_t q = threadIdx.x; q < nf; q += kBlock) __builtin_nontemporal_store(tile[q], vert + g0 + q);
    }
    // indices: (0, 1, 2), (0, 2, 3) per face of four vertices
    const uint32_t NI = V + V / 2u;
    const uint64_t i0 = s0 * NI;
    const uint32_t first = base_vertex + (uint32_t)(s0 * V);
    for (uint32_t j = threadIdx.x; j < ns * NI; j += kBlock) {
        const uint32_t face = j / 6u, r = j - face * 6u;
        const uint32_t cc = r == 0u || r == 3u ? 0u : (r == 1u ? 1u : (r == 2u || r == 4u ? 2u : 3u));
        __builtin_nontemporal_store(first + face * 4u + cc, idx + i0 + j);
    }
}

} // namespace pgamesh

namespace {

pga_status launch(const pga_instance* instances, uint64_t count, const uint64_t* device_count, int32_t shape_type,
                  pga_vertex* vertices, uint32_t* indices, uint32_t base_vertex, void* stream) {
    if (count == 0u) return PGA_OK;
    if (!instances || !vertices || !indices) return PGA_ERR_INVALID_ARG;
    if (shape_type != PGA_SHAPE_BOX && shape_type != PGA_SHAPE_QUAD) return PGA_ERR_UNSUPPORTED;
    if (((uintptr_t)instances & 15u) != 0u || ((uintptr_t)vertices & 3u) != 0u || ((uintptr_t)indices & 3u) != 0u ||
        ((uintptr_t)device_count & 7u) != 0u)
        return PGA_ERR_INVALID_ARG;
    const uint64_t V = shape_type == PGA_SHAPE_QUAD ? 4u : 24u;
    if (count > (0xFFFFFFFFull - base_vertex) / V) return PGA_ERR_INVALID_ARG;   // 32-bit vertex indices
    const uint64_t tile_shapes = pgamesh::kTileVerts / V;
    const uint64_t blocks = (count + tile_shapes - 1u) / tile_shapes;
    if (blocks > 0x7FFFFFFFull) return PGA_ERR_INVALID_ARG;
    hipLaunchKernelGGL(pgamesh::k_tessellate, dim3((uint32_t)blocks), dim3(pgamesh::kBlock), 0, (hipStream_t)stream,
                       reinterpret_cast<const float4*>(instances), count, shape_type == PGA_SHAPE_QUAD ? 1u : 0u,
                       reinterpret_cast<float*>(vertices), indices, base_vertex,
                       ((uintptr_t)vertices & 15u) == 0u ? 1u : 0u,
                       reinterpret_cast<const unsigned long long*>(device_count));
    return hipGetLastError() == hipSuccess ? PGA_OK : PGA_ERR_HIP;
}

} // namespace

extern "C" {

// Declared in include/pga.h.
PGA_API pga_status pga_tessellate(const pga_instance* instances, uint64_t count, int32_t shape_type,
                                  pga_vertex* vertices, uint32_t* indices, uint32_t base_vertex, void* stream) {
    return launch(instances, count, nullptr, shape_type, vertices, indices, base_vertex, stream);
}

// Declared in include/pga.h.
PGA_API pga_status pga_tessellate_device_count(const pga_instance* instances, const uint64_t* device_count,
                                               uint64_t capacity, int32_t shape_type, pga_vertex* vertices,
                                               uint32_t* indices, uint32_t base_vertex, void* stream) {
    if (!device_count && capacity) return PGA_ERR_INVALID_ARG;
    return launch(instances, capacity, device_count, shape_type, vertices, indices, base_vertex, stream);
}

} // extern "C"
