// ntt.hpp -- negacyclic NTT / INTT of one N-point polynomial held in LDS,
// executed cooperatively by one workgroup.
//
// Forward: Cooley-Tukey with bit-reversed twiddles psi^bitrev(k), Harvey lazy
// butterflies (values in [0, 4q)); output slot j = a(psi^(2*bitrev(j)+1)),
// the order SEAL's ntt_negacyclic_harvey produces.  Inverse: Gentleman-Sande
// with psi^-bitrev(k); values stay in [0, 2q); the N^-1 scale is applied by
// the caller when it reads the result out (ntt_inv_final).
//
// Stages whose butterfly span is >= 64 elements go through LDS with one
// barrier per stage; the last six stages (span 32..1) run inside a wavefront
// on registers, exchanging partners with ds_swizzle/permute (__shfl_xor).
#pragma once

#include "arith.hpp"

namespace onion {

constexpr int kNttThreads = 256;

// Butterfly helpers --------------------------------------------------------------
__device__ __forceinline__ void ct_bfly(uint64_t &x, uint64_t &y, uint64_t w, uint64_t ws,
                                        uint64_t q, uint64_t two_q) {
  uint64_t u = x;
  if (u >= two_q) u -= two_q;
  uint64_t v = mul_shoup_lazy(y, w, ws, q);
  x = u + v;
  y = u + two_q - v;
}

__device__ __forceinline__ void gs_bfly(uint64_t &x, uint64_t &y, uint64_t w, uint64_t ws,
                                        uint64_t q, uint64_t two_q) {
  uint64_t u = x, v = y;
  uint64_t s = u + v;
  if (s >= two_q) s -= two_q;
  x = s;
  y = mul_shoup_lazy(u + two_q - v, w, ws, q);
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int mask) {
  int lo = __shfl_xor((int)(uint32_t)v, mask, 64);
  int hi = __shfl_xor((int)(uint32_t)(v >> 32), mask, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Forward NTT of s[0..N) in LDS.  Entry: values < 4q, visible to the whole
// workgroup (caller synchronises).  Exit: values in [0, 4q) in LDS, synchronised.
template <int N>
__device__ __forceinline__ void ntt_fwd_lds(uint64_t *s, const NttTab &T) {
  static_assert(N >= 128 && (N & (N - 1)) == 0, "N must be a power of two >= 128");
  const uint64_t q = T.q, two_q = 2 * q;
  const int tid = threadIdx.x;
  // LDS stages: half = N/2 .. 64
  int groups = 1;
  for (int half = N / 2; half >= 64; half >>= 1, groups <<= 1) {
#pragma unroll
    for (int r = 0; r < N / 2 / kNttThreads; ++r) {
      const int b = tid + r * kNttThreads;
      const int i = b / half, j = b % half;
      const int x = 2 * i * half + j;
      ct_bfly(s[x], s[x + half], T.rp[groups + i], T.rps[groups + i], q, two_q);
    }
    __syncthreads();
  }
  // Register stages: each wave owns contiguous blocks of 64 elements
  // (block = 64 consecutive slots, one per lane); stages half = 32 .. 1.
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int kBlocks = N / 64;
  constexpr int kWaves = kNttThreads / 64;
#pragma unroll
  for (int r = 0; r < kBlocks / kWaves; ++r) {
    const int blk = wave + r * kWaves;
    uint64_t v = s[blk * 64 + lane];
    int g = groups * (blk * 64) / 64;  // first group index of this block at half=32
    // at stage `half`, the group covering this lane is g0 + lane/(2*half)
#pragma unroll
    for (int half = 32, gmul = 1; half >= 1; half >>= 1, gmul <<= 1) {
      const int grp = (groups * gmul) + (blk * 64 + lane) / (2 * half);
      const uint64_t w = T.rp[grp], ws = T.rps[grp];
      const uint64_t other = shfl_xor64(v, half);
      const bool upper = (lane & half) != 0;
      // lower lane holds x, upper lane holds y
      uint64_t x = upper ? other : v, y = upper ? v : other;
      ct_bfly(x, y, w, ws, q, two_q);
      v = upper ? y : x;
      (void)g;
    }
    s[blk * 64 + lane] = v;
  }
  __syncthreads();
}

// Inverse NTT (without the N^-1 scale) of s[0..N) in LDS.  Entry: values in
// [0, 2q), synchronised.  Exit: values in [0, 2q), synchronised.
template <int N>
__device__ __forceinline__ void ntt_inv_lds(uint64_t *s, const NttTab &T) {
  const uint64_t q = T.q, two_q = 2 * q;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int kBlocks = N / 64;
  constexpr int kWaves = kNttThreads / 64;
  // Register stages first: half = 1 .. 32 (span = 2 .. 64)
#pragma unroll
  for (int r = 0; r < kBlocks / kWaves; ++r) {
    const int blk = wave + r * kWaves;
    uint64_t v = s[blk * 64 + lane];
#pragma unroll
    for (int half = 1; half <= 32; half <<= 1) {
      const int groups = N / (2 * half);
      const int grp = groups + (blk * 64 + lane) / (2 * half);
      const uint64_t w = T.irp[grp], ws = T.irps[grp];
      const uint64_t other = shfl_xor64(v, half);
      const bool upper = (lane & half) != 0;
      uint64_t x = upper ? other : v, y = upper ? v : other;
      gs_bfly(x, y, w, ws, q, two_q);
      v = upper ? y : x;
    }
    s[blk * 64 + lane] = v;
  }
  __syncthreads();
  // LDS stages: half = 64 .. N/2
  for (int half = 64; half <= N / 2; half <<= 1) {
    const int groups = N / (2 * half);
#pragma unroll
    for (int r = 0; r < N / 2 / kNttThreads; ++r) {
      const int b = tid + r * kNttThreads;
      const int i = b / half, j = b % half;
      const int x = 2 * i * half + j;
      gs_bfly(s[x], s[x + half], T.irp[groups + i], T.irps[groups + i], q, two_q);
    }
    __syncthreads();
  }
}

// Canonical value of a forward-NTT output in [0, 4q).
__device__ __forceinline__ uint64_t ntt_fwd_final(uint64_t v, uint64_t q) {
  const uint64_t two_q = 2 * q;
  if (v >= two_q) v -= two_q;
  if (v >= q) v -= q;
  return v;
}

// Canonical value of an inverse-NTT output in [0, 2q), scaled by N^-1.
__device__ __forceinline__ uint64_t ntt_inv_final(uint64_t v, const NttTab &T) {
  uint64_t r = mul_shoup_lazy(v, T.n_inv, T.n_inv_s, T.q);
  return r >= T.q ? r - T.q : r;
}

}  // namespace onion
