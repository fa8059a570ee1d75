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

}  // namespace onion
