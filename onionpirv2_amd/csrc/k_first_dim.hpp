// k_first_dim.hpp -- first-dimension product over the HBM-resident database
// (PirServer::evaluate_first_dim, src/server.cpp:123-175, whose core is
//  naive_level_mat_mat_128 / mat_mat_128, src/matrix.cpp:79-97, 220-235, and
//  whose epilogue is delay_modulus, src/server.cpp:178-293).
//
// For every NTT slot l: C_l[r][c] = sum_j A_l[r][j] * B_l[j][c], exact and
// wrapping mod 2^128 like the reference's uint128 accumulator, then
// the canonical residue of C mod 2^128 (reduce128, exact like the
// reference's barrett_reduce_128).
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
//   Qs   [l][j][col][2] uint32 limbs, col = 2*query + poly, padded to `cpad`
//        columns (written by k_sel_ntt)
//   out  per query: [r][c][l]                   (NTT-form ciphertexts)
#pragma once

#include "k_ntt.hpp"

namespace onion {

constexpr uint32_t kMask30 = (1u << 30) - 1;
constexpr int kMaxFst = 1024;

__device__ __forceinline__ uint64_t mad32(uint32_t a, uint32_t b, uint64_t c) {
  return (uint64_t)a * b + c;
}

// Selection vectors: NTT of poly c of leaf j of query b (grid = fst*2 x B),
// scattered as limbs into Qs[l][j][2b + c]; NTT_IN: the leaves are in NTT
// form already (slot order, canonical).
template <bool NTT_IN>
__global__ __launch_bounds__(kNttThreads) void k_sel_ntt(const uint64_t *leaves, uint64_t leaf_stride,
                                                         uint32_t *Qs, uint32_t fst, uint32_t cpad,
                                                         NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  const uint32_t j = blockIdx.x / 2, c = blockIdx.x % 2, b = blockIdx.y;
  const uint64_t *src = leaves + (size_t)b * leaf_stride + (size_t)j * 2 * kN + (size_t)c * kN;
  uint64_t x[8];
  if (NTT_IN) {
    load_blocked(x, src);
  } else {
    load_natural(x, src);
    ntt_fwd(x, lds, T);
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int e = 8 * threadIdx.x + r;
    const uint64_t v = NTT_IN ? x[r] : ntt_fwd_final(x[r], T.q);
    uint2 lim;
    lim.x = (uint32_t)v & kMask30;
    lim.y = (uint32_t)(v >> 30);
    *reinterpret_cast<uint2 *>(Qs + (((size_t)e * fst + j) * cpad + 2 * b + c) * 2) = lim;
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

// Four independent partial sums (the two cross products apart), so each of
// the four v_mad_u64_u32 per product accumulates in place; 16 terms fit.
struct Acc4 {
  uint64_t s0, s1a, s1b, s2;
};

__device__ __forceinline__ void acc_word4(Acc4 &x, uint32_t a0, uint32_t a1, uint32_t b0,
                                          uint32_t b1) {
  x.s0 = mad32(a0, b0, x.s0);
  x.s1a = mad32(a0, b1, x.s1a);
  x.s1b = mad32(a1, b0, x.s1b);
  x.s2 = mad32(a1, b1, x.s2);
}

__device__ __forceinline__ void fold4(uint64_t &lo, uint64_t &hi, Acc4 &x) {
  add128(lo, hi, x.s0, 0);
  add128(lo, hi, x.s1a << 30, x.s1a >> 34);
  add128(lo, hi, x.s1b << 30, x.s1b >> 34);
  add128(lo, hi, x.s2 << 60, x.s2 >> 4);
  x.s0 = x.s1a = x.s1b = x.s2 = 0;
}

__device__ __forceinline__ void split30(uint64_t a, uint32_t &a0, uint32_t &a1) {
  a0 = (uint32_t)a & kMask30;
  a1 = (uint32_t)(a >> 30);
}

// 16-byte streaming load; NT marks it nontemporal (the database is read once
// per query, so it should not displace the L2-resident selection vector).
template <bool NT>
__device__ __forceinline__ uint4 stream_load(const uint4 *p) {
  if constexpr (NT) {
    const uint64_t *w = reinterpret_cast<const uint64_t *>(p);
    const uint64_t lo = __builtin_nontemporal_load(w), hi = __builtin_nontemporal_load(w + 1);
    return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
  } else {
    return *p;
  }
}

// Single query (B = 1, Qs with cpad = 2).  grid.x = slot, grid.y = row chunk
// of `rows_per_block` rows; 256 threads.  Lane layout per wave: 8 rows x 8
// column groups; a lane reads 16 B (two adjacent words) per step, so 8 lanes
// cover 128 contiguous bytes of one row; UNROLL steps are in flight per lane.
// Requires fst % 16 == 0.
template <int UNROLL, bool NT>
__global__ __launch_bounds__(256) void k_first_dim_b1(const uint64_t *__restrict__ db,
                                                      const uint32_t *__restrict__ Qs,
                                                      uint64_t *__restrict__ out, uint32_t fst,
                                                      uint32_t other, uint64_t num_pt,
                                                      uint32_t rows_per_block, Mod mq) {
  static_assert(UNROLL == 4 || UNROLL == 8, "fold cadence assumes 4 or 8 steps");
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
    for (uint32_t k = 0; k < steps; k += UNROLL) {
      uint4 v[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
        v[u] = (k + u < steps) ? stream_load<NT>(row + (k + u) * 8 + cg) : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const uint32_t j = (k + u) * 16 + cg * 2;
        const uint4 qa = qs[min(j, fst - 1)], qb = qs[min(j + 1, fst - 1)];
        uint32_t a0, a1;
        split30(((uint64_t)v[u].y << 32) | v[u].x, a0, a1);
        acc_word(x0, a0, a1, qa.x, qa.y);
        acc_word(x1, a0, a1, qa.z, qa.w);
        split30(((uint64_t)v[u].w << 32) | v[u].z, a0, a1);
        acc_word(x0, a0, a1, qb.x, qb.y);
        acc_word(x1, a0, a1, qb.z, qb.w);
        // s1 takes two products per word and a step is two words: fold every 4 steps
        if ((u & 3) == 3) {
          fold(w0lo, w0hi, x0);
          fold(w1lo, w1hi, x1);
        }
      }
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
      const uint64_t res = cg == 0 ? reduce128(w0lo, w0hi, mq) : reduce128(w1lo, w1hi, mq);
      out[((size_t)r * 2 + cg) * kN + slot] = res;
    }
  }
}

// Generic shape (any fst), single query: one lane per row, scalar loads.
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
    out[((size_t)r * 2 + 0) * kN + slot] = reduce128(w0lo, w0hi, mq);
    out[((size_t)r * 2 + 1) * kN + slot] = reduce128(w1lo, w1hi, mq);
  }
}

// Batched: C_l (other x cpad) = A_l (other x fst) * Q_l (fst x cpad) for all
// columns (2 per query) at once, so each database word leaves HBM once per
// batch.  Tiles: a block owns one slot and RBLK rows; per 32-wide k chunk the
// A tile (RBLK x 32, split into limbs) and the Q tile (32 x cpad) are staged in
// LDS; each thread accumulates TR rows x TC columns (rows rg + i*RG, columns
// cg + k*CG, so LDS reads are conflict-free or broadcast).
// Block order is XCD-aware: consecutive blocks of one XCD (block id mod 8)
// walk the row tiles of one slot, so that slot's Q (fst x cpad words) stays in
// that XCD's L2 while its tiles run.
template <int TR, int TC, int CPAD>
__global__ __launch_bounds__(256) void k_first_dim_bat(const uint64_t *__restrict__ db,
                                                       const uint32_t *__restrict__ Qs,
                                                       uint64_t *__restrict__ fd, uint64_t fd_stride,
                                                       uint32_t fst, uint32_t other, uint64_t num_pt,
                                                       uint32_t nq, uint32_t n_tiles, Mod mq) {
  constexpr int CG = CPAD / TC;   // thread columns
  constexpr int RG = 256 / CG;    // thread rows
  constexpr int RBLK = RG * TR;   // rows per block
  constexpr int KC = 32;          // k chunk
  static_assert(CG * RG == 256, "thread grid must be 256");
  __shared__ uint2 As[RBLK][KC + 1];
  __shared__ uint2 Qt[KC][CPAD];

  const uint32_t L = blockIdx.x, xcd = L & 7, seq = L >> 3;
  const uint32_t slot = (seq / n_tiles) * 8 + xcd;
  const uint32_t tile = seq % n_tiles;
  const int tid = threadIdx.x, cg = tid % CG, rg = tid / CG;
  const uint32_t row0 = tile * RBLK;
  const uint64_t *A = db + (size_t)slot * num_pt;
  const uint2 *Q = reinterpret_cast<const uint2 *>(Qs) + (size_t)slot * fst * CPAD;

  Acc4 acc[TR][TC];
  uint64_t wlo[TR][TC], whi[TR][TC];
#pragma unroll
  for (int i = 0; i < TR; ++i)
#pragma unroll
    for (int k = 0; k < TC; ++k) {
      acc[i][k] = Acc4{0, 0, 0, 0};
      wlo[i][k] = whi[i][k] = 0;
    }

  for (uint32_t k0 = 0; k0 < fst; k0 += KC) {
    __syncthreads();
    for (int idx = tid; idx < RBLK * KC; idx += 256) {
      const int rr = idx / KC, kk = idx % KC;
      const uint32_t r = row0 + rr, j = k0 + kk;
      const uint64_t v = (r < other && j < fst) ? __builtin_nontemporal_load(A + (size_t)r * fst + j) : 0;
      As[rr][kk] = make_uint2((uint32_t)v & kMask30, (uint32_t)(v >> 30));
    }
    for (int idx = tid; idx < KC * CPAD; idx += 256) {
      const int kk = idx / CPAD, col = idx % CPAD;
      Qt[kk][col] = (k0 + kk < fst) ? Q[(size_t)(k0 + kk) * CPAD + col] : make_uint2(0, 0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KC; ++kk) {
      uint2 a[TR], b[TC];
#pragma unroll
      for (int i = 0; i < TR; ++i) a[i] = As[rg + i * RG][kk];
#pragma unroll
      for (int k = 0; k < TC; ++k) b[k] = Qt[kk][cg + k * CG];
#pragma unroll
      for (int i = 0; i < TR; ++i)
#pragma unroll
        for (int k = 0; k < TC; ++k) acc_word4(acc[i][k], a[i].x, a[i].y, b[k].x, b[k].y);
      if ((kk & 15) == 15) {
#pragma unroll
        for (int i = 0; i < TR; ++i)
#pragma unroll
          for (int k = 0; k < TC; ++k) fold4(wlo[i][k], whi[i][k], acc[i][k]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < TR; ++i) {
    const uint32_t r = row0 + rg + i * RG;
#pragma unroll
    for (int k = 0; k < TC; ++k) {
      const uint32_t col = cg + k * CG, b = col >> 1, c = col & 1;
      if (r < other && b < nq)
        fd[(size_t)b * fd_stride + ((size_t)r * 2 + c) * kN + slot] = reduce128(wlo[i][k], whi[i][k], mq);
    }
  }
}

}  // namespace onion
