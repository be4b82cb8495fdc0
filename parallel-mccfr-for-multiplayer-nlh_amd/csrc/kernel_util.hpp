// Small device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace mccfr {

constexpr int kDeckStride = 52;  // bytes of LDS deck per lane (13 dwords: odd stride, no bank aliasing)

template <int ROW>
__device__ __forceinline__ void load_bf16_row(const uint16_t* p, float (&out)[ROW]) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int v = 0; v < ROW / 8; ++v) {
    const uint4 w = q[v];
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      out[v * 8 + 2 * k] = bf16_to_f32(ws[k] & 0xFFFFu);
      out[v * 8 + 2 * k + 1] = bf16_to_f32(ws[k] >> 16);
    }
  }
}

template <int ROW>
__device__ __forceinline__ void load_f32_row(const float* p, float (&out)[ROW]) {
  const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
  for (int v = 0; v < ROW / 4; ++v) {
    const float4 w = q[v];
    out[4 * v] = w.x;
    out[4 * v + 1] = w.y;
    out[4 * v + 2] = w.z;
    out[4 * v + 3] = w.w;
  }
}

__device__ __forceinline__ uint32_t fnv_path(uint32_t h, uint32_t slot, uint32_t a) {
  h = (h ^ (slot & 0xFFu)) * 16777619u;
  h = (h ^ ((slot >> 8) & 0xFFu)) * 16777619u;
  h = (h ^ ((slot >> 16) & 0xFFu)) * 16777619u;
  h = (h ^ (slot >> 24)) * 16777619u;
  return (h ^ (a & 0xFFu)) * 16777619u;
}

__device__ __forceinline__ int wrap(int p, int N) { return p >= N ? p - N : p; }

inline unsigned grid_for(size_t n, unsigned block, unsigned cap) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (unsigned)(g < cap ? g : cap);
}

}  // namespace mccfr
