// k_tail.hpp -- modulus switch to small_q and response bit-packing.
#pragma once

#include "k_ntt.hpp"

namespace onion {

// PirServer::mod_switch_inplace (src/server.cpp:779-801) computes
//   (uint64) round((long double)x * (long double)((double)q' / (double)Q))
// in x87 extended precision.  Here the double scale is M * 2^-f exactly
// (M < 2^53); the product x*M is rounded to a 64-bit significand (round half
// to even, the x87 default), then rounded half away from zero at bit f.
// Integer-only, so every lane reproduces the host's long double result.
// grid.y = query: ct + query * stride holds the 2N values.
__global__ void k_mod_switch(uint64_t *cts, uint64_t stride, uint32_t n, uint64_t M, uint32_t f) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t *ct = cts + (size_t)blockIdx.y * stride;
  const uint64_t x = ct[i];
  if (x == 0) return;
  unsigned __int128 P = (unsigned __int128)x * M;
  const uint64_t hi = (uint64_t)(P >> 64), lo = (uint64_t)P;
  const int nb = hi ? 128 - __clzll(hi) : 64 - __clzll(lo);
  if (nb > 64) {
    const int sh = nb - 64;
    unsigned __int128 keep = P >> sh;
    const unsigned __int128 rem = P & ((((unsigned __int128)1) << sh) - 1);
    const unsigned __int128 halfv = ((unsigned __int128)1) << (sh - 1);
    if (rem > halfv || (rem == halfv && (keep & 1))) keep += 1;
    P = keep << sh;
  }
  ct[i] = (uint64_t)((P + (((unsigned __int128)1) << (f - 1))) >> f);
}

// PirServer::save_resp_to_stream (src/server.cpp:581-633): every coefficient
// of c0 then c1, masked to `w` bits, packed LSB-first.  Thread t packs
// coefficients 8t..8t+7, which fill exactly bytes [w*t, w*t + w).
__global__ void k_pack_response(const uint64_t *ct, uint8_t *out, uint32_t n_coeffs, uint32_t w) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (8 * t >= n_coeffs) return;
  const uint64_t mask = (w >= 64) ? ~0ULL : ((1ULL << w) - 1);
  unsigned __int128 buf = 0;
  uint32_t bits = 0, nbyte = 0;
  uint8_t *dst = out + (size_t)w * t;
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t idx = 8 * t + k;
    const uint64_t v = idx < n_coeffs ? (ct[idx] & mask) : 0;
    buf |= (unsigned __int128)v << bits;
    bits += w;
    while (bits >= 8) {
      dst[nbyte++] = (uint8_t)buf;
      buf >>= 8;
      bits -= 8;
    }
  }
}

// Plain streaming copy of n 16-byte vectors (calibrates the HBM byte counters).
__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace onion
