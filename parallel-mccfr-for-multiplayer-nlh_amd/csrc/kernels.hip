// gfx950 kernels of the MCCFR hot path.
//
//   (the launch pipeline -- deal, walk, record sort, table update -- is in traverse.hip)
//   k_rm_sweep<ROW>   full-table regret matching (after tables are uploaded).
//   k_normalize<ROW>  average strategy (kernel_normalize_strategy, cfr_gpu.cu:1029-1058).
//   k_clamp<ROW>, k_plane<ROW>  CFR+ clamp of every regret; one plane of the
//                     interleaved regret / strategy-sum rows to or from a dense buffer.
//   (regret and strategy-sum rows are interleaved: slot s at [s * 2 * ROW], S = R + ROW)
//   k_eval7_walk      7-card rank via the rank-walk table (eval7_gpu, cfr_gpu.cu:126-137).
//   k_eval7_arith     7-card rank via flush table + perfect hash of the rank multiset.
#include <hip/hip_runtime.h>

#include "kernel_util.hpp"
#include "kernels.hpp"

namespace mccfr {
namespace {

template <int ROW>
__global__ __launch_bounds__(256) void k_rm_sweep(const float* __restrict__ R, uint16_t* __restrict__ sig,
                                                  uint64_t slots) {
  for (uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < slots;
       slot += (uint64_t)gridDim.x * blockDim.x) {
    float r[ROW];
    load_f32_row<ROW>(R + slot * 2 * ROW, r);
    float sum = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a) sum += fmaxf(0.0f, r[a]);
    uint32_t packed[ROW / 2];
#pragma unroll
    for (int a = 0; a < ROW; a += 2) {
      const float x0 = (sum > 1e-7f) ? fmaxf(0.0f, r[a]) / sum : 0.0f;
      const float x1 = (sum > 1e-7f) ? fmaxf(0.0f, r[a + 1]) / sum : 0.0f;
      packed[a / 2] = (uint32_t)f32_to_bf16(x0) | ((uint32_t)f32_to_bf16(x1) << 16);
    }
#pragma unroll
    for (int v = 0; v < ROW / 8; ++v)
      reinterpret_cast<uint4*>(sig + slot * ROW)[v] =
          make_uint4(packed[4 * v], packed[4 * v + 1], packed[4 * v + 2], packed[4 * v + 3]);
  }
}

template <int ROW>
__global__ __launch_bounds__(256) void k_normalize(const float* __restrict__ S, float* __restrict__ out,
                                                   uint64_t slots) {
  for (uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < slots;
       slot += (uint64_t)gridDim.x * blockDim.x) {
    float s[ROW];
    load_f32_row<ROW>(S + slot * 2 * ROW, s);
    float total = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a) total += s[a];
#pragma unroll
    for (int v = 0; v < ROW / 4; ++v) {
      float4 o;
      o.x = (total > 1e-7f) ? s[4 * v] / total : 0.0f;
      o.y = (total > 1e-7f) ? s[4 * v + 1] / total : 0.0f;
      o.z = (total > 1e-7f) ? s[4 * v + 2] / total : 0.0f;
      o.w = (total > 1e-7f) ? s[4 * v + 3] / total : 0.0f;
      reinterpret_cast<float4*>(out + slot * ROW)[v] = o;
    }
  }
}

// Writes one tenant descriptor (passed by value, so the host copy may change
// as soon as the launch is queued).
__global__ void k_set_tenant(DevGame* dst, DevGame g) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = g;
}

// kernel_clamp_regrets (cfr_gpu.cu:1018-1022) over a whole table.
template <int ROW>
__global__ __launch_bounds__(256) void k_clamp(float* __restrict__ R, uint64_t slots) {
  for (uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < slots;
       slot += (uint64_t)gridDim.x * blockDim.x) {
    float r[ROW];
    load_f32_row<ROW>(R + slot * 2 * ROW, r);
#pragma unroll
    for (int a = 0; a < ROW; ++a) r[a] = fmaxf(0.0f, r[a]);
    store_f32_row<ROW>(R + slot * 2 * ROW, r);
  }
}

// One plane (regrets or strategy sums) of slots [slot0, slot0 + n) between the
// interleaved device rows and a dense [n][ROW] staging buffer.
template <int ROW, bool TO_DENSE>
__global__ __launch_bounds__(256) void k_plane(float* __restrict__ plane, float* __restrict__ dense, uint64_t slot0,
                                               uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    float v[ROW];
    if (TO_DENSE) {
      load_f32_row<ROW>(plane + (slot0 + i) * 2 * ROW, v);
      store_f32_row<ROW>(dense + i * ROW, v);
    } else {
      load_f32_row<ROW>(dense + i * ROW, v);
      store_f32_row<ROW>(plane + (slot0 + i) * 2 * ROW, v);
    }
  }
}

__global__ __launch_bounds__(256) void k_zero(uint4* __restrict__ p, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------------------------
// Seven-card evaluators.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_hand(const uint8_t* cards, size_t stride, size_t i, uint32_t c[7]) {
  const uint8_t* p = cards + i * stride;
  if (stride == 8) {
    const uint2 w = *reinterpret_cast<const uint2*>(p);
    c[0] = w.x & 0xFF; c[1] = (w.x >> 8) & 0xFF; c[2] = (w.x >> 16) & 0xFF; c[3] = w.x >> 24;
    c[4] = w.y & 0xFF; c[5] = (w.y >> 8) & 0xFF; c[6] = (w.y >> 16) & 0xFF;
  } else {
#pragma unroll
    for (int k = 0; k < 7; ++k) c[k] = p[k];
  }
}

__device__ __forceinline__ bool distinct_cards(const uint32_t c[7]) {
  uint64_t m = 0;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    ok = ok && c[k] < 52u;
    m |= 1ull << (c[k] & 63u);
  }
  return ok && __popcll(m) == 7;
}

// kWalkHands hands per lane, walked in lockstep: their dependent gathers are
// independent of each other, so a lane keeps kWalkHands loads in flight
// (one hand per lane leaves the walk latency-bound: 7 dependent gathers).
constexpr int kWalkHands = 4;
__global__ __launch_bounds__(256) void k_eval7_walk(const int32_t* __restrict__ hr,
                                                    const uint8_t* __restrict__ cards, size_t stride,
                                                    uint16_t* __restrict__ out, size_t n) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  for (size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += T * kWalkHands) {
    uint32_t c[kWalkHands][7];
    bool ok[kWalkHands];
    int p[kWalkHands];
#pragma unroll
    for (int j = 0; j < kWalkHands; ++j) {
      const size_t i = i0 + (size_t)j * T;
      ok[j] = false;
      if (i < n) {
        load_hand(cards, stride, i, c[j]);
        ok[j] = distinct_cards(c[j]);  // codes above 51 or repeated cards have no rank
      }
      p[j] = ok[j] ? hr[53 + card_rm(c[j][0])] : 0;
    }
#pragma unroll
    for (int k = 1; k < 6; ++k)
#pragma unroll
      for (int j = 0; j < kWalkHands; ++j) p[j] = ok[j] ? hr[p[j] + card_rm(c[j][k])] : 0;
#pragma unroll
    for (int j = 0; j < kWalkHands; ++j) {
      const size_t i = i0 + (size_t)j * T;
      if (i < n) out[i] = ok[j] ? (uint16_t)decode_packed((uint32_t)hr[p[j] + card_rm(c[j][6])] & 0xFFFFu) : 0;
    }
  }
}

// Arithmetic evaluator (arith_rank7) with its tables staged in LDS.
constexpr int kArithWG = 512;
__global__ __launch_bounds__(kArithWG) void k_eval7_arith(ArithTables t, const uint8_t* __restrict__ cards,
                                                          size_t stride, uint16_t* __restrict__ out, size_t n) {
  __shared__ uint16_t s_nf[kNonFlushClasses + 1];
  __shared__ uint16_t s_fl[8192];
  __shared__ uint16_t s_off[13 * 8 * 5];
  for (int i = threadIdx.x; i < kNonFlushClasses; i += kArithWG) s_nf[i] = t.nonflush[i];
  for (int i = threadIdx.x; i < 8192; i += kArithWG) s_fl[i] = t.flush[i];
  for (int i = threadIdx.x; i < 13 * 8 * 5; i += kArithWG) s_off[i] = t.offsets[i];
  __syncthreads();
  for (size_t i = (size_t)blockIdx.x * kArithWG + threadIdx.x; i < n; i += (size_t)gridDim.x * kArithWG) {
    uint32_t c[7];
    load_hand(cards, stride, i, c);
    if (!distinct_cards(c)) {
      out[i] = 0;
      continue;
    }
    const uint32_t rank = arith_rank7(c, s_nf, s_fl, s_off);
    out[i] = (uint16_t)rank;
  }
}

// Hands of 7 distinct uniform cards: the first 7 positions of a Fisher-Yates
// shuffle driven by the traversal draw stream of id i.
__global__ __launch_bounds__(256) void k_random_hands(uint32_t seed_lo, uint32_t seed_hi, uint32_t tag,
                                                      uint8_t* __restrict__ out, size_t n) {
  __shared__ uint8_t s_deck[kDeckStride * 256];
  uint8_t* dk = s_deck + threadIdx.x * kDeckStride;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    DrawStream rng;
    rng.init(seed_lo, seed_hi, i, tag);
    for (int k = 0; k < 52; ++k) dk[k] = (uint8_t)k;
    for (int k = 0; k < 7; ++k) {
      const uint32_t r = rng.next();
      const int j = k + (int)(((uint64_t)r * (uint64_t)(52 - k)) >> 32);
      const uint8_t a = dk[k], b = dk[j];
      dk[k] = b;
      dk[j] = a;
    }
    uint2 w;
    w.x = (uint32_t)dk[0] | ((uint32_t)dk[1] << 8) | ((uint32_t)dk[2] << 16) | ((uint32_t)dk[3] << 24);
    w.y = (uint32_t)dk[4] | ((uint32_t)dk[5] << 8) | ((uint32_t)dk[6] << 16);
    *reinterpret_cast<uint2*>(out + i * 8) = w;
  }
}

template <int ROW>
__global__ __launch_bounds__(256) void k_randomize(float* __restrict__ R, float* __restrict__ S, uint64_t slots,
                                                   int num_actions, uint32_t seed_lo, uint32_t seed_hi) {
  for (uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < slots;
       slot += (uint64_t)gridDim.x * blockDim.x) {
    DrawStream rng;
    rng.init(seed_lo, seed_hi, slot, 0x72616E64u);
    float r[ROW];
#pragma unroll
    for (int a = 0; a < ROW; ++a) r[a] = a < num_actions ? u01(rng.next()) : 0.0f;
#pragma unroll
    for (int v = 0; v < ROW / 4; ++v) {
      reinterpret_cast<float4*>(R + slot * 2 * ROW)[v] = make_float4(r[4 * v], r[4 * v + 1], r[4 * v + 2], r[4 * v + 3]);
      reinterpret_cast<float4*>(S + slot * 2 * ROW)[v] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

// Bucket of every K-card set (K = 5: flop, K = 6: turn) indexed by colex rank:
// the bucket the reference derives from the rank-walk word after K cards
// (cfr_gpu.cu:114-122, 140-150, 476-489).  The walk state depends only on the
// card set, so walking the sorted cards gives the reference's word.
// CLAMP: the strategy harness's bucket (abstraction.cpp:61-70), min(.., 127),
// instead of the training walk's uint8 truncation; they differ for rank 0 only.
template <int K, bool CLAMP>
__global__ __launch_bounds__(256) void k_build_buckets(const int32_t* __restrict__ hr, uint8_t* __restrict__ out,
                                                       uint32_t count) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < count; idx += gridDim.x * blockDim.x) {
    uint32_t cards[K];
    uint32_t r = idx, top = 52;
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {  // unrank: largest c with C(c, i+1) <= r
      uint32_t cc = top - 1;
      while (true) {
        uint32_t b;
        switch (i) {
          case 0: b = binom<1>(cc); break;
          case 1: b = binom<2>(cc); break;
          case 2: b = binom<3>(cc); break;
          case 3: b = binom<4>(cc); break;
          case 4: b = binom<5>(cc); break;
          default: b = binom<6>(cc); break;
        }
        if (b <= r) {
          r -= b;
          break;
        }
        --cc;
      }
      cards[i] = cc;
      top = cc;
    }
    int p = hr[53 + card_rm(cards[0])];
#pragma unroll
    for (int i = 1; i < K; ++i) p = hr[p + card_rm(cards[i])];
    const uint32_t rank = decode_packed((uint32_t)p & 0xFFFFu);
    out[idx] = (uint8_t)(CLAMP ? eval_bucket_of_rank(rank) : bucket_of_rank(rank));
  }
}

// Random 16-byte row reads (byte-counter calibration for narrow gathers):
// read i takes row splitmix(i) mod rows (rows a power of two) and its sum of
// words is stored as 4 bytes, so every read is used.
__global__ __launch_bounds__(256) void k_gather(const uint4* __restrict__ table, uint64_t row_mask,
                                                uint32_t* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint64_t z = (uint64_t)i + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const uint4 v = table[z & row_mask];
    out[i] = v.x + v.y + v.z + v.w;
  }
}

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}


}  // namespace

void launch_set_tenant(hipStream_t s, DevGame* dst, const DevGame& g) {
  hipLaunchKernelGGL(k_set_tenant, dim3(1), dim3(64), 0, s, dst, g);
}
void launch_zero(hipStream_t s, void* p, size_t bytes) {  // bytes: a multiple of 16
  hipLaunchKernelGGL(k_zero, dim3(grid_for(bytes / 16, 256, 1u << 16)), dim3(256), 0, s, (uint4*)p, bytes / 16);
}
void launch_clamp(hipStream_t s, int row, float* R, uint64_t slots) {
  const unsigned grid = grid_for(slots, 256, 65536);
  if (row == 8)
    hipLaunchKernelGGL(k_clamp<8>, dim3(grid), dim3(256), 0, s, R, slots);
  else
    hipLaunchKernelGGL(k_clamp<16>, dim3(grid), dim3(256), 0, s, R, slots);
}
void launch_plane(hipStream_t s, int row, float* plane, float* dense, uint64_t slot0, uint64_t n, bool to_dense) {
  const unsigned grid = grid_for(n, 256, 65536);
  if (row == 8) {
    if (to_dense) hipLaunchKernelGGL((k_plane<8, true>), dim3(grid), dim3(256), 0, s, plane, dense, slot0, n);
    else hipLaunchKernelGGL((k_plane<8, false>), dim3(grid), dim3(256), 0, s, plane, dense, slot0, n);
  } else {
    if (to_dense) hipLaunchKernelGGL((k_plane<16, true>), dim3(grid), dim3(256), 0, s, plane, dense, slot0, n);
    else hipLaunchKernelGGL((k_plane<16, false>), dim3(grid), dim3(256), 0, s, plane, dense, slot0, n);
  }
}
void launch_rm_sweep(hipStream_t s, int row, const float* R, uint16_t* sig, uint64_t slots) {
  const unsigned grid = grid_for(slots, 256, 65536);
  if (row == 8)
    hipLaunchKernelGGL(k_rm_sweep<8>, dim3(grid), dim3(256), 0, s, R, sig, slots);
  else
    hipLaunchKernelGGL(k_rm_sweep<16>, dim3(grid), dim3(256), 0, s, R, sig, slots);
}
void launch_normalize(hipStream_t s, int row, const float* S, float* out, uint64_t slots) {
  const unsigned grid = grid_for(slots, 256, 65536);
  if (row == 8)
    hipLaunchKernelGGL(k_normalize<8>, dim3(grid), dim3(256), 0, s, S, out, slots);
  else
    hipLaunchKernelGGL(k_normalize<16>, dim3(grid), dim3(256), 0, s, S, out, slots);
}
void launch_eval7_walk(hipStream_t s, const int32_t* hr, const uint8_t* cards, size_t stride, uint16_t* out,
                       size_t n) {
  hipLaunchKernelGGL(k_eval7_walk, dim3(grid_for((n + kWalkHands - 1) / kWalkHands, 256, 1u << 20)), dim3(256), 0, s,
                     hr, cards, stride, out, n);
}
void launch_eval7_arith(hipStream_t s, ArithTables t, const uint8_t* cards, size_t stride, uint16_t* out,
                        size_t n) {
  hipLaunchKernelGGL(k_eval7_arith, dim3(grid_for(n, kArithWG, 2048)), dim3(kArithWG), 0, s, t, cards, stride,
                     out, n);
}
void launch_random_hands(hipStream_t s, uint64_t seed, uint32_t tag, uint8_t* out, size_t n) {
  hipLaunchKernelGGL(k_random_hands, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, (uint32_t)seed,
                     (uint32_t)(seed >> 32), tag, out, n);
}
void launch_randomize(hipStream_t s, int row, int num_actions, float* R, float* S, uint64_t slots, uint64_t seed) {
  const unsigned grid = grid_for(slots, 256, 65536);
  if (row == 8)
    hipLaunchKernelGGL(k_randomize<8>, dim3(grid), dim3(256), 0, s, R, S, slots, num_actions, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
  else
    hipLaunchKernelGGL(k_randomize<16>, dim3(grid), dim3(256), 0, s, R, S, slots, num_actions, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}
void launch_build_buckets(hipStream_t s, const int32_t* hr, BucketTables b, bool clamp) {
  if (clamp) {
    hipLaunchKernelGGL((k_build_buckets<5, true>), dim3(grid_for(kFlopSets, 256, 65536)), dim3(256), 0, s, hr,
                       b.flop5, (uint32_t)kFlopSets);
    hipLaunchKernelGGL((k_build_buckets<6, true>), dim3(grid_for(kTurnSets, 256, 65536)), dim3(256), 0, s, hr,
                       b.turn6, (uint32_t)kTurnSets);
    return;
  }
  hipLaunchKernelGGL((k_build_buckets<5, false>), dim3(grid_for(kFlopSets, 256, 65536)), dim3(256), 0, s, hr,
                     b.flop5, (uint32_t)kFlopSets);
  hipLaunchKernelGGL((k_build_buckets<6, false>), dim3(grid_for(kTurnSets, 256, 65536)), dim3(256), 0, s, hr,
                     b.turn6, (uint32_t)kTurnSets);
}
void launch_copy(hipStream_t s, const uint4* src, uint4* dst, size_t n16) {
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n16, 256, 1u << 16)), dim3(256), 0, s, src, dst, n16);
}
void launch_gather(hipStream_t s, const uint4* table, uint64_t rows, uint32_t* out, size_t n) {
  hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, table, rows - 1, out, n);
}

}  // namespace mccfr
