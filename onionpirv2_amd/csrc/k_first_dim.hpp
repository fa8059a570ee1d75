// k_first_dim.hpp -- first-dimension product over the HBM-resident database
// (PirServer::evaluate_first_dim, src/server.cpp:123-175, whose core is
//  naive_level_mat_mat_128 / mat_mat_128, src/matrix.cpp:79-97, 220-235, and
//  whose epilogue is delay_modulus, src/server.cpp:178-293).
//
// For every NTT slot l: C_l[r][c] = sum_j A_l[r][j] * B_l[j][c], exact and
// wrapping mod 2^128 like the reference's uint128 accumulator, then
// (C mod 2^128) mod q with the same Barrett reduction.
//
// Exact 64x64 products from 30-bit limbs: a = a1*2^30 + a0, b = b1*2^30 + b0
// (all limbs < 2^30), so every 32x32 partial product is < 2^60 and a plain
// 64-bit accumulator takes 16 of them without overflow.  Three partial sums
// (s0 = sum a0*b0, s1 = sum a0*b1 + a1*b0, s2 = sum a1*b1) are folded into a
// wrapping 128-bit accumulator every 8 terms:  W += s0 + s1*2^30 + s2*2^60.
// Each partial product is one v_mad_u64_u32.
//
// Layouts
//   DB   A[l * num_pt + r * fst + j]            (the reference's db_aligned_)
//   Qs   [l][j] x {b0_lo, b0_hi, b1_lo, b1_hi}  (uint32 limbs of the two query
//        polys at slot l; written by k_sel_ntt)
//   out  [r][c][l]                              (NTT-form ciphertexts)
#pragma once

#include "k_ntt.hpp"

namespace onion {

constexpr uint32_t kMask30 = (1u << 30) - 1;
constexpr int kMaxFst = 1024;

__device__ __forceinline__ uint64_t mad32(uint32_t a, uint32_t b, uint64_t c) {
  return (uint64_t)a * b + c;
}

// Selection vector: NTT of poly `c` of leaf j (grid = fst x 2) scattered into
// the slot-major limb layout Qs.
__global__ __launch_bounds__(kNttThreads) void k_sel_ntt(const uint64_t *leaves, uint32_t *Qs,
                                                         uint32_t fst, NttTab T) {
  __shared__ uint64_t s[kN];
  const uint32_t j = blockIdx.x / 2, c = blockIdx.x % 2;
  const uint64_t *src = leaves + (size_t)j * 2 * kN + (size_t)c * kN;
#pragma unroll
  for (int r = 0; r < kPer; ++r) s[threadIdx.x + r * kNttThreads] = src[threadIdx.x + r * kNttThreads];
  __syncthreads();
  ntt_fwd_lds<kN>(s, T);
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kNttThreads;
    const uint64_t v = ntt_fwd_final(s[e], T.q);
    uint2 lim;
    lim.x = (uint32_t)v & kMask30;
    lim.y = (uint32_t)(v >> 30);
    *reinterpret_cast<uint2 *>(Qs + ((size_t)e * fst + j) * 4 + 2 * c) = lim;
  }
}

struct Acc3 {
  uint64_t s0, s1, s2;
};

__device__ __forceinline__ void acc_word(Acc3 &x, uint32_t a0, uint32_t a1, uint32_t b0,
                                         uint32_t b1) {
  x.s0 = mad32(a0, b0, x.s0);
  x.s1 = mad32(a0, b1, x.s1);
  x.s1 = mad32(a1, b0, x.s1);
  x.s2 = mad32(a1, b1, x.s2);
}

__device__ __forceinline__ void fold(uint64_t &lo, uint64_t &hi, Acc3 &x) {
  add128(lo, hi, x.s0, 0);
  add128(lo, hi, x.s1 << 30, x.s1 >> 34);
  add128(lo, hi, x.s2 << 60, x.s2 >> 4);
  x.s0 = x.s1 = x.s2 = 0;
}

__device__ __forceinline__ void split30(uint64_t a, uint32_t &a0, uint32_t &a1) {
  a0 = (uint32_t)a & kMask30;
  a1 = (uint32_t)(a >> 30);
}

// Single query (B = 1).  grid.x = slot, grid.y = row chunk of `rows_per_block`
// rows; 256 threads.  Lane layout per wave: 8 rows x 8 column groups; a lane
// reads 16 B (two adjacent words) per step, so 8 lanes cover 128 contiguous
// bytes of one row.  Requires fst % 16 == 0.
__global__ __launch_bounds__(256) void k_first_dim_b1(const uint64_t *__restrict__ db,
                                                      const uint32_t *__restrict__ Qs,
                                                      uint64_t *__restrict__ out, uint32_t fst,
                                                      uint32_t other, uint64_t num_pt,
                                                      uint32_t rows_per_block, Mod mq) {
  __shared__ uint4 qs[kMaxFst];  // per column j: {b0_lo, b0_hi, b1_lo, b1_hi}
  const uint32_t slot = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rsub = lane >> 3, cg = lane & 7;
  const uint4 *qsrc = reinterpret_cast<const uint4 *>(Qs) + (size_t)slot * fst;
  for (uint32_t j = tid; j < fst; j += 256) qs[j] = qsrc[j];
  __syncthreads();

  const uint64_t *A = db + (size_t)slot * num_pt;
  const uint32_t r_begin = blockIdx.y * rows_per_block;
  const uint32_t r_end = min(other, r_begin + rows_per_block);
  const uint32_t steps = fst / 16;  // 16 words per 8-lane group per step
  for (uint32_t r0 = r_begin; r0 < r_end; r0 += 32) {
    const uint32_t r = r0 + wave * 8 + rsub;
    const bool active = r < r_end;
    const uint4 *row = reinterpret_cast<const uint4 *>(A + (size_t)(active ? r : r_begin) * fst);
    Acc3 x0 = {0, 0, 0}, x1 = {0, 0, 0};
    uint64_t w0lo = 0, w0hi = 0, w1lo = 0, w1hi = 0;
    uint32_t k = 0;
    for (; k < steps; k += 4) {
      uint4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = (k + u < steps) ? row[(k + u) * 8 + cg] : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t j = (k + u) * 16 + cg * 2;
        const uint4 qa = qs[min(j, fst - 1)], qb = qs[min(j + 1, fst - 1)];
        uint32_t a0, a1;
        split30(((uint64_t)v[u].y << 32) | v[u].x, a0, a1);
        acc_word(x0, a0, a1, qa.x, qa.y);
        acc_word(x1, a0, a1, qa.z, qa.w);
        split30(((uint64_t)v[u].w << 32) | v[u].z, a0, a1);
        acc_word(x0, a0, a1, qb.x, qb.y);
        acc_word(x1, a0, a1, qb.z, qb.w);
      }
      fold(w0lo, w0hi, x0);
      fold(w1lo, w1hi, x1);
    }
    // reduce the 8 column groups of this row (lanes differing in bits 0..2)
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
      const uint64_t a = shfl_xor64(w0lo, m), b = shfl_xor64(w0hi, m);
      const uint64_t c = shfl_xor64(w1lo, m), d = shfl_xor64(w1hi, m);
      add128(w0lo, w0hi, a, b);
      add128(w1lo, w1hi, c, d);
    }
    if (active && cg < 2) {
      const uint64_t res = cg == 0 ? barrett128(w0lo, w0hi, mq) : barrett128(w1lo, w1hi, mq);
      out[((size_t)r * 2 + cg) * kN + slot] = res;
    }
  }
}

// Generic shape (any fst): one lane per row, scalar loads.  Used for shapes
// the streaming kernel does not cover.
__global__ __launch_bounds__(256) void k_first_dim_generic(const uint64_t *__restrict__ db,
                                                           const uint32_t *__restrict__ Qs,
                                                           uint64_t *__restrict__ out, uint32_t fst,
                                                           uint32_t other, uint64_t num_pt, Mod mq) {
  const uint32_t slot = blockIdx.x;
  const uint4 *qsrc = reinterpret_cast<const uint4 *>(Qs) + (size_t)slot * fst;
  const uint64_t *A = db + (size_t)slot * num_pt;
  for (uint32_t r = threadIdx.x; r < other; r += blockDim.x) {
    Acc3 x0 = {0, 0, 0}, x1 = {0, 0, 0};
    uint64_t w0lo = 0, w0hi = 0, w1lo = 0, w1hi = 0;
    for (uint32_t j = 0; j < fst; ++j) {
      const uint4 qa = qsrc[j];
      uint32_t a0, a1;
      split30(A[(size_t)r * fst + j], a0, a1);
      acc_word(x0, a0, a1, qa.x, qa.y);
      acc_word(x1, a0, a1, qa.z, qa.w);
      if ((j & 7) == 7) {
        fold(w0lo, w0hi, x0);
        fold(w1lo, w1hi, x1);
      }
    }
    fold(w0lo, w0hi, x0);
    fold(w1lo, w1hi, x1);
    out[((size_t)r * 2 + 0) * kN + slot] = barrett128(w0lo, w0hi, mq);
    out[((size_t)r * 2 + 1) * kN + slot] = barrett128(w1lo, w1hi, mq);
  }
}

}  // namespace onion
