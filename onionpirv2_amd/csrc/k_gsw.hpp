// k_gsw.hpp -- RGSW external product pieces
// (GSWEval::external_product / decomp_rlwe_single_mod / decomp_to_ntt,
//  src/gsw_eval.cpp:33-106, 184-252), used by query conversion
// (GSWEval::query_to_gsw, src/gsw_eval.cpp:254-286) and by the other-dimension
// multiplexer (PirServer::other_dim_mux, src/server.cpp:295-334).
//
// Split in two launches so that every digit NTT of every ciphertext runs in
// parallel:
//   k_decomp_ntt  -- one workgroup per (ciphertext, digit): extract the digit
//                    polynomial, NTT it, store it (canonical, NTT form);
//   k_ext_mac_*   -- one workgroup per (ciphertext, output component):
//                    sum_j D_j * G_j[comp] in 128-bit, Barrett mod q, then
//                    either store (conversion, NTT form) or INTT and add into
//                    the mux accumulator (other-dimension fold).
#pragma once

#include "k_ntt.hpp"

namespace onion {

// Digits of ciphertext c: c = src_a[c] (or src_b[c] - src_a[c] mod q when
// src_b is given, the "y - x" of the mux).  Digit j = poly*l + (l-1-p) holds
// (x >> p*B) & (2^B - 1).  D layout: [ct][2l][N].
__global__ __launch_bounds__(kNttThreads) void k_decomp_ntt(const uint64_t *src_a,
                                                            const uint64_t *src_b, uint64_t *D,
                                                            uint32_t l, uint32_t base_log2,
                                                            NttTab T) {
  __shared__ uint64_t s[kN];
  const uint32_t rows = 2 * l;
  const uint64_t ct = blockIdx.x / rows;
  const uint32_t j = blockIdx.x % rows;
  const uint32_t poly = j / l, pp = l - 1 - (j % l);
  const uint32_t sh = pp * base_log2;
  const uint64_t mask = (base_log2 >= 64) ? ~0ULL : ((1ULL << base_log2) - 1);
  const uint64_t *a = src_a + ct * 2 * kN + (size_t)poly * kN;
  const uint64_t *b = src_b ? src_b + ct * 2 * kN + (size_t)poly * kN : nullptr;
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kNttThreads;
    uint64_t x = a[e];
    if (b) x = sub_mod(b[e], x, T.q);
    s[e] = (x >> sh) & mask;
  }
  __syncthreads();
  ntt_fwd_lds<kN>(s, T);
  uint64_t *dst = D + ((size_t)ct * rows + j) * kN;
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kNttThreads;
    dst[e] = ntt_fwd_final(s[e], T.q);
  }
}

// 128-bit accumulation of the 2l digit products for coefficient e.
__device__ __forceinline__ uint64_t ext_mac_one(const uint64_t *Dc, const uint64_t *G, uint32_t rows,
                                                int comp, int e, const Mod &mq) {
  uint64_t lo = 0, hi = 0;
  for (uint32_t j = 0; j < rows; ++j) {
    uint64_t plo, phi;
    mul128(Dc[(size_t)j * kN + e], G[(size_t)j * 2 * kN + (size_t)comp * kN + e], plo, phi);
    add128(lo, hi, plo, phi);
  }
  return barrett128(lo, hi, mq);
}

// Conversion: out row for ciphertext c, component comp (NTT form):
//   out + (c / l) * dim_stride + (l + c % l) * 2N + comp * N
// G: one RGSW of 2l rows x 2N (the client's RGSW(s) key).
__global__ __launch_bounds__(kNttThreads) void k_ext_mac_store(const uint64_t *D, const uint64_t *G,
                                                               uint32_t rows, uint64_t *out,
                                                               uint32_t l_out, Mod mq) {
  const uint64_t ct = blockIdx.x / 2;
  const int comp = blockIdx.x % 2;
  const uint64_t *Dc = D + ct * rows * kN;
  uint64_t *dst = out + (ct / l_out) * (2 * l_out) * 2 * kN + (l_out + ct % l_out) * 2 * kN +
                  (size_t)comp * kN;
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kNttThreads;
    dst[e] = ext_mac_one(Dc, G, rows, comp, e, mq);
  }
}

// Mux fold: x_c[comp] += INTT(sum_j D_j * G_j[comp]), x_c = cts + c*2N.
__global__ __launch_bounds__(kNttThreads) void k_ext_mac_intt_add(const uint64_t *D,
                                                                  const uint64_t *G, uint32_t rows,
                                                                  uint64_t *cts, NttTab T, Mod mq) {
  __shared__ uint64_t s[kN];
  const uint64_t ct = blockIdx.x / 2;
  const int comp = blockIdx.x % 2;
  const uint64_t *Dc = D + ct * rows * kN;
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kNttThreads;
    s[e] = ext_mac_one(Dc, G, rows, comp, e, mq);
  }
  __syncthreads();
  ntt_inv_lds<kN>(s, T);
  uint64_t *x = cts + ct * 2 * kN + (size_t)comp * kN;
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kNttThreads;
    x[e] = add_mod(x[e], ntt_inv_final(s[e], T), T.q);
  }
}

// Conversion rows 0..l-1: NTT of both polys of leaf c into
//   out + (c / l) * dim_stride + (c % l) * 2N + poly * N
__global__ __launch_bounds__(kNttThreads) void k_rgsw_top_rows(const uint64_t *leaves, uint64_t *out,
                                                               uint32_t l, NttTab T) {
  __shared__ uint64_t s[kN];
  const uint64_t ct = blockIdx.x / 2;
  const int poly = blockIdx.x % 2;
  const uint64_t *src = leaves + ct * 2 * kN + (size_t)poly * kN;
  uint64_t *dst = out + (ct / l) * (2 * l) * 2 * kN + (ct % l) * 2 * kN + (size_t)poly * kN;
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

}  // namespace onion
