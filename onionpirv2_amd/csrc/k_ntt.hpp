// k_ntt.hpp -- batched NTT kernels: one workgroup (256 threads) per polynomial.
#pragma once

#include "ntt.hpp"

namespace onion {

constexpr int kN = 2048;  // ring degree served by the device path

// Forward NTT of n polys: out[b] = NTT(in[b]) with canonical output.  The
// input may hold any values < 4q (in == out allowed).  R: pass schedule
// (radix 8 or 4).
template <int R = 8>
__global__ __launch_bounds__(kNttThreads) void k_ntt_fwd(const uint64_t *in, uint64_t *out,
                                                         NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  uint64_t x[8];
  load_natural(x, in + (size_t)blockIdx.x * kN);
  ntt_fwd_sched<R>(x, lds, T);
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = ntt_fwd_final(x[j], T.q);
  store_blocked(out + (size_t)blockIdx.x * kN, x);
}

// Inverse NTT of n polys (canonical input and output).
template <int R = 8>
__global__ __launch_bounds__(kNttThreads) void k_ntt_inv(const uint64_t *in, uint64_t *out,
                                                         NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  uint64_t x[8];
  load_blocked(x, in + (size_t)blockIdx.x * kN);
  ntt_inv_sched<R>(x, lds, T);
  uint64_t *dst = out + (size_t)blockIdx.x * kN;
#pragma unroll
  for (int j = 0; j < 8; ++j) dst[threadIdx.x + 256 * j] = ntt_inv_final(x[j], T);
}

// Issue-rate probe: radix-8 forward passes of the transforms' butterflies
// (Shoup products with the truncated quotient, the lazy additions, one
// reduction stage per pass) on registers only, with no memory, LDS or barrier
// inside the loop.  The VALU instruction rate it reaches at a given number of
// waves per SIMD is the ceiling for the transform-bound kernels' instruction
// mix; their own rate is compared against it.  The values wrap and mean
// nothing; sink keeps them live.
__global__ __launch_bounds__(kNttThreads) void k_valu_probe(NttTab T, uint32_t rounds, uint64_t *sink) {
  const int t = threadIdx.x;
  uint64_t x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = T.rp[(8 * t + j) & (kN - 1)] ^ blockIdx.x;
  Tw8 tw;
  load_tw8(tw, T.rp, T.rps, 256 + t);
#pragma unroll 1
  for (uint32_t r = 0; r < rounds; ++r) fwd8<2>(x, tw, T.q, true);
  uint64_t v = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) v ^= x[j];
  sink[(size_t)blockIdx.x * kNttThreads + t] = v;
}

}  // namespace onion
