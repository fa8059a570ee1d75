// k_ntt.hpp -- batched NTT kernels: one workgroup (256 threads) per polynomial.
#pragma once

#include "ntt.hpp"

namespace onion {

constexpr int kN = 2048;              // ring degree served by the device path
constexpr int kPer = kN / kNttThreads;  // coefficients owned by one thread

// Forward NTT of n polys: out[b] = NTT(in[b]) with canonical output.  The
// input may hold any values < 4q (in == out allowed).
__global__ __launch_bounds__(kNttThreads) void k_ntt_fwd(const uint64_t *in, uint64_t *out,
                                                         NttTab T) {
  __shared__ uint64_t s[kN];
  const uint64_t *src = in + (size_t)blockIdx.x * kN;
  uint64_t *dst = out + (size_t)blockIdx.x * kN;
#pragma unroll
  for (int r = 0; r < kPer; ++r) s[threadIdx.x + r * kNttThreads] = src[threadIdx.x + r * kNttThreads];
  __syncthreads();
  ntt_fwd_lds<kN>(s, T);
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kNttThreads;
    dst[e] = ntt_fwd_final(s[e], T.q);
  }
}

// Inverse NTT of n polys (canonical input and output).
__global__ __launch_bounds__(kNttThreads) void k_ntt_inv(const uint64_t *in, uint64_t *out,
                                                         NttTab T) {
  __shared__ uint64_t s[kN];
  const uint64_t *src = in + (size_t)blockIdx.x * kN;
  uint64_t *dst = out + (size_t)blockIdx.x * kN;
#pragma unroll
  for (int r = 0; r < kPer; ++r) s[threadIdx.x + r * kNttThreads] = src[threadIdx.x + r * kNttThreads];
  __syncthreads();
  ntt_inv_lds<kN>(s, T);
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kNttThreads;
    dst[e] = ntt_inv_final(s[e], T);
  }
}

}  // namespace onion
