// pga_loss.hip - the centre/size MSE of instances against targets and its
// gradient in one pass (gfx950).
//
// For instance i with model rows 0..2 (m, 3x4) the centre is column 3 and
// the size is the column norms of the 3x3 part:
//   c_i = (m[0][3], m[1][3], m[2][3]),  s_ij = |m[:, j]|,  j = 0..2,
//   loss = wc * sum_i |c_i - tc_i|^2 + ws * sum_i |s_i - ts_i|^2,
//   dloss/dm[r][3] = 2 wc (c_ir - tc_ir),
//   dloss/dm[r][j] = 2 ws (s_ij - ts_ij) m[r][j] / s_ij   (0 where s_ij = 0).
// One thread per instance: 48 bytes of the instance and 24 of the targets
// read, the 64-byte gradient row written (column 12..15 zero), the loss
// summed per workgroup in FP64 and added once per workgroup. The reverse
// pass then reads the gradient rows as dL/dinstance.
#include <hip/hip_runtime.h>

#include "../../include/pga.h"

namespace pgaloss {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_center_size_mse(const float4* __restrict__ inst,
                                                             const float* __restrict__ tc,
                                                             const float* __restrict__ ts, uint64_t n, float wc,
                                                             float ws, float4* __restrict__ grad,
                                                             double* __restrict__ loss) {
    __shared__ double part[kBlock / 64];
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    double l = 0.0;
    if (i < n) {
        const float4 r0 = inst[i * 4u], r1 = inst[i * 4u + 1u], r2 = inst[i * 4u + 2u];
        const float m[12] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w};
        float g[12];
        float lc = 0.0f, lsz = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) {           // centre: column 3
            const float d = m[r * 4 + 3] - tc[i * 3u + r];
            lc += d * d;
            g[r * 4 + 3] = 2.0f * wc * d;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {           // size: norm of column j
            const float a = m[j], b = m[4 + j], c = m[8 + j];
            const float s = sqrtf(a * a + b * b + c * c);
            const float d = s - ts[i * 3u + j];
            lsz += d * d;
            const float k = s > 0.0f ? 2.0f * ws * d / s : 0.0f;
            g[j] = k * a;
            g[4 + j] = k * b;
            g[8 + j] = k * c;
        }
        grad[i * 4u] = make_float4(g[0], g[1], g[2], g[3]);
        grad[i * 4u + 1u] = make_float4(g[4], g[5], g[6], g[7]);
        grad[i * 4u + 2u] = make_float4(g[8], g[9], g[10], g[11]);
        grad[i * 4u + 3u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        l = (double)wc * (double)lc + (double)ws * (double)lsz;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) l += __shfl_xor(l, d);
    if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += part[w];
        if (t != 0.0) atomicAdd(loss, t);
    }
}

} // namespace pgaloss

extern "C" {

// Declared in include/pga.h.
PGA_API pga_status pga_center_size_mse(const float* instances, const float* target_centers,
                                       const float* target_sizes, uint64_t count, float weight_centers,
                                       float weight_sizes, float* dl_dinstances, double* loss, void* stream) {
    if (count == 0u) return PGA_OK;
    if (!instances || !target_centers || !target_sizes || !dl_dinstances || !loss) return PGA_ERR_INVALID_ARG;
    if ((((uintptr_t)instances) | ((uintptr_t)dl_dinstances)) & 15u) return PGA_ERR_INVALID_ARG;
    const uint64_t blocks = (count + pgaloss::kBlock - 1u) / pgaloss::kBlock;
    if (blocks > 0x7FFFFFFFull) return PGA_ERR_INVALID_ARG;
    hipLaunchKernelGGL(pgaloss::k_center_size_mse, dim3((uint32_t)blocks), dim3(pgaloss::kBlock), 0,
                       (hipStream_t)stream, reinterpret_cast<const float4*>(instances), target_centers,
                       target_sizes, count, weight_centers, weight_sizes, reinterpret_cast<float4*>(dl_dinstances),
                       loss);
    return hipGetLastError() == hipSuccess ? PGA_OK : PGA_ERR_HIP;
}

} // extern "C"
