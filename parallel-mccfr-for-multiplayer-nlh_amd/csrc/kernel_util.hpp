// Small device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace mccfr {

constexpr int kDeckStride = 52;  // bytes of LDS deck per lane (13 dwords: odd stride, no bank aliasing)

// Pointers read out of descriptors (tenant tables) are generic to the
// compiler, which then emits flat accesses; these are global buffers, so the
// accessors below say so (global loads and stores wait on their own counter).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gbl(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gbl(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int ROW>
__device__ __forceinline__ void load_bf16_row(const uint16_t* p, float (&out)[ROW]) {
  const auto q = gbl(reinterpret_cast<const u32x4*>(p));
#pragma unroll
  for (int v = 0; v < ROW / 8; ++v) {
    const u32x4 w = q[v];
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
  const auto q = gbl(reinterpret_cast<const f32x4*>(p));
#pragma unroll
  for (int v = 0; v < ROW / 4; ++v) {
    const f32x4 w = q[v];
    out[4 * v] = w.x;
    out[4 * v + 1] = w.y;
    out[4 * v + 2] = w.z;
    out[4 * v + 3] = w.w;
  }
}

// Orders one wave's LDS writes before its later LDS reads by other lanes (a
// wave's LDS instructions execute in order; this keeps the compiler from
// moving accesses across it).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int ROW>
__device__ __forceinline__ void store_f32_row(float* p, const float (&v)[ROW]) {
  const auto q = gbl(reinterpret_cast<f32x4*>(p));
#pragma unroll
  for (int k = 0; k < ROW / 4; ++k) {
    f32x4 w;
    w.x = v[4 * k];
    w.y = v[4 * k + 1];
    w.z = v[4 * k + 2];
    w.w = v[4 * k + 3];
    q[k] = w;
  }
}

// A BF16 row from ROW / 2 packed pairs (action 2k in the low half).
template <int ROW>
__device__ __forceinline__ void store_bf16_pairs(uint16_t* p, const uint32_t (&pairs)[ROW / 2]) {
  const auto q = gbl(reinterpret_cast<u32x4*>(p));
#pragma unroll
  for (int k = 0; k < ROW / 8; ++k) {
    u32x4 w;
    w.x = pairs[4 * k];
    w.y = pairs[4 * k + 1];
    w.z = pairs[4 * k + 2];
    w.w = pairs[4 * k + 3];
    q[k] = w;
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

namespace mccfr {

// C(n, K) for 0 <= n < 64 and a compile-time K <= 7 (0 when n < K: the
// product then contains the factor n - n).
template <int K>
__host__ __device__ __forceinline__ uint32_t binom(uint32_t n) {
  uint64_t p = 1;
#pragma unroll
  for (int j = 0; j < K; ++j) p *= (uint64_t)(n - (uint32_t)j);
  constexpr uint64_t f = K == 0 ? 1 : K == 1 ? 1 : K == 2 ? 2 : K == 3 ? 6 : K == 4 ? 24 : K == 5 ? 120 : K == 6 ? 720 : 5040;
  return (uint32_t)(p / f);
}

__device__ __forceinline__ void cswap(uint32_t& a, uint32_t& b) {
  const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  a = lo;
  b = hi;
}

// Colex rank of a 5- or 6-card set given unsorted: sum_i C(c_(i), i+1) over the
// ascending order c_(0) < ... < c_(K-1).
__device__ __forceinline__ uint32_t colex5(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e) {
  cswap(a, b); cswap(d, e); cswap(c, e); cswap(c, d); cswap(b, e);
  cswap(a, d); cswap(a, c); cswap(b, d); cswap(b, c);
  return binom<1>(a) + binom<2>(b) + binom<3>(c) + binom<4>(d) + binom<5>(e);
}
__device__ __forceinline__ uint32_t colex6(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e, uint32_t f) {
  cswap(a, b); cswap(c, d); cswap(e, f); cswap(a, c); cswap(b, e); cswap(d, f);
  cswap(a, b); cswap(c, d); cswap(e, f); cswap(b, c); cswap(d, e); cswap(c, d);
  return binom<1>(a) + binom<2>(b) + binom<3>(c) + binom<4>(d) + binom<5>(e) + binom<6>(f);
}

// Colex ranks from a table of binomials, tb[k * 52 + n] = C(n, k) for
// k = 0..6, n < 52 (same values as colex5 / colex6).
__device__ __forceinline__ uint32_t colex5_tb(const uint32_t* tb, uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                              uint32_t e) {
  cswap(a, b); cswap(d, e); cswap(c, e); cswap(c, d); cswap(b, e);
  cswap(a, d); cswap(a, c); cswap(b, d); cswap(b, c);
  return tb[52 + a] + tb[104 + b] + tb[156 + c] + tb[208 + d] + tb[260 + e];
}
__device__ __forceinline__ uint32_t colex6_tb(const uint32_t* tb, uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                              uint32_t e, uint32_t f) {
  cswap(a, b); cswap(c, d); cswap(e, f); cswap(a, c); cswap(b, e); cswap(d, f);
  cswap(a, b); cswap(c, d); cswap(e, f); cswap(b, c); cswap(d, e); cswap(c, d);
  return tb[52 + a] + tb[104 + b] + tb[156 + c] + tb[208 + d] + tb[260 + e] + tb[312 + f];
}

// 7-card rank 1..7462 of distinct suit-major cards by the arithmetic
// evaluator: a suit holding >= 5 cards makes a flush or straight flush, which
// no other hand of seven cards beats, ranked by its 13-bit rank mask; otherwise
// the rank multiset (counts c_0..c_12, sum 7, each <= 4) is ranked
// lexicographically, index = sum_i off[i][k_i][c_i] with k_i the cards left
// before rank i (prefix counts from one multiply of packed 3-bit fields).
//
// The card sums (packed rank counts, packed suit counts, card mask) are exact
// integer sums, so a set of cards shared by several hands (a board) is summed
// once (rank7_add) and each hand adds its own cards before rank7_finish.
struct Rank7Sums {
  uint64_t cnt = 0, cmask = 0;  // 3-bit count per rank; bit per card
  uint32_t scnt = 0;            // 4-bit count per suit
};

__device__ __forceinline__ void rank7_add(Rank7Sums& a, uint32_t card) {
  const uint32_t s = (card * 79u) >> 10;  // card / 13 for card < 52
  const uint32_t r = card - 13u * s;
  a.cnt += 1ull << (3 * r);
  a.scnt += 1u << (4 * s);
  a.cmask |= 1ull << card;
}

__device__ __forceinline__ uint32_t rank7_finish(const Rank7Sums& a, const uint16_t* __restrict__ nonflush,
                                                 const uint16_t* __restrict__ flush, const uint16_t* off) {
  const uint32_t fl = (a.scnt + 0x3333u) & 0x8888u;  // a nibble >= 5
  if (fl) {
    const uint32_t s = (uint32_t)__builtin_ctz(fl) >> 2;
    return flush[(uint32_t)(a.cmask >> (13 * s)) & 0x1FFFu];
  }
  const uint64_t pre = a.cnt * 0x0000001249249249ull;  // field i = c_0 + .. + c_i
  uint32_t idx = 0;
#pragma unroll
  for (int r = 0; r < 13; ++r) {
    const uint32_t before = r == 0 ? 0u : (uint32_t)(pre >> (3 * (r - 1))) & 7u;
    const uint32_t cr = (uint32_t)(a.cnt >> (3 * r)) & 7u;
    idx += off[(r * 8 + (7 - before)) * 5 + cr];
  }
  // keep the sum in 32 bits: otherwise the compiler folds each term into the
  // 64-bit address (13 64-bit adds instead of one)
  asm volatile("" : "+v"(idx));
  return nonflush[idx];
}

__device__ __forceinline__ uint32_t arith_rank7(const uint32_t c[7], const uint16_t* __restrict__ nonflush,
                                                const uint16_t* __restrict__ flush, const uint16_t* off) {
  Rank7Sums a;
#pragma unroll
  for (int k = 0; k < 7; ++k) rank7_add(a, c[k]);
  return rank7_finish(a, nonflush, flush, off);
}

}  // namespace mccfr
