// k_gsw.hpp -- RGSW external product pieces
// (GSWEval::external_product / decomp_rlwe_single_mod / decomp_to_ntt,
//  src/gsw_eval.cpp:33-106, 184-252), used by query conversion
// (GSWEval::query_to_gsw, src/gsw_eval.cpp:254-286) and by the other-dimension
// multiplexer (PirServer::other_dim_mux, src/server.cpp:295-334).
//
// Split in two launches so that every digit NTT of every ciphertext of every
// query in the batch runs in parallel:
//   k_decomp_ntt  -- one workgroup per (ciphertext, digit): extract the digit
//                    polynomial, NTT it, store it (NTT form, below 2^63:
//                    the 128-bit sums reduce it exactly);
//   k_ext_mac_*   -- one workgroup per (ciphertext, output component):
//                    sum_j D_j * G_j[comp] in 128-bit, reduced mod q, then
//                    either store (conversion, NTT form) or INTT and add into
//                    the mux accumulator (other-dimension fold).
// Fused variants run one workgroup per ciphertext (k_mux_fused,
// k_convert_fused); k_mux_fused_ntt folds pairs of NTT-form ciphertexts with
// a level RGSW prepared by k_rgsw_fold_last, and k_cts_to_coeff returns
// NTT-form ciphertexts to coefficient form.
// grid.y is the query index; every buffer has a per-query stride.
#pragma once

#include "k_ntt.hpp"

namespace onion {

// Digits of ciphertext c: x = src_a[c] (or src_b[c] - src_a[c] mod q when
// src_b is given, the "y - x" of the mux).  Digit j = poly*l + (l-1-p) holds
// (x >> p*B) & (2^B - 1).  D layout per query: [ct][2l][N] (NTT form).
// FOLD (a gadget matrix prepared by k_rgsw_fold_last): the last digit's row
// holds NTT(x) itself.
template <bool FOLD = false>
__global__ __launch_bounds__(kNttThreads) void k_decomp_ntt(const uint64_t *src_a,
                                                            const uint64_t *src_b, uint64_t src_stride,
                                                            uint64_t *D, uint64_t d_stride,
                                                            uint32_t l, uint32_t base_log2,
                                                            NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  const uint32_t rows = 2 * l;
  const uint64_t ct = blockIdx.x / rows;
  const uint32_t j = blockIdx.x % rows;
  const uint32_t poly = j / l, pp = l - 1 - (j % l);
  const uint32_t sh = pp * base_log2;
  const uint64_t mask = (1ULL << base_log2) - 1;
  const size_t qoff = (size_t)blockIdx.y * src_stride;
  const uint64_t *a = src_a + qoff + ct * 2 * kN + (size_t)poly * kN;
  const uint64_t *b = src_b ? src_b + qoff + ct * 2 * kN + (size_t)poly * kN : nullptr;
  const bool whole = FOLD && pp == 0;
  uint64_t x[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int e = threadIdx.x + 256 * r;
    uint64_t v = a[e];
    if (b) v = sub_mod(b[e], v, T.q);
    x[r] = whole ? v : (v >> sh) & mask;
  }
  if (whole)
    ntt_fwd(x, lds, T);
  else
    ntt_fwd<true>(x, lds, T);  // digits < 2^base_log2 <= 2^32
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = ntt_fwd_below_2p63(x[r], T.q);  // only ext_mac8 reads D
  store_blocked(D + (size_t)blockIdx.y * d_stride + ((size_t)ct * rows + j) * kN, x);
}

// 128-bit accumulation of the 2l digit products, blocked slots 8t .. 8t+7
__device__ __forceinline__ void ext_mac8(const uint64_t *Dc, const uint64_t *G, uint32_t rows, int comp,
                                         const Mod &mq, uint64_t (&out)[8]) {
  U128 acc[8] = {};
  for (uint32_t j = 0; j < rows; ++j) {
    uint64_t d[8], g[8];
    load_blocked(d, Dc + (size_t)j * kN);
    load_blocked(g, G + (size_t)j * 2 * kN + (size_t)comp * kN);
    mac128x8(acc, d, g);
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) out[r] = reduce128(acc[r], mq);
}

// Conversion: out row for ciphertext c, component comp (NTT form):
//   out + (c / l) * 2l * 2N + (l + c % l) * 2N + comp * N   (per query)
// G: the query's client RGSW(s) key (2 l_key rows x 2N) at g_ptrs[query].
__global__ __launch_bounds__(kNttThreads) void k_ext_mac_store(const uint64_t *D, uint64_t d_stride,
                                                               const uint64_t *const *g_ptrs,
                                                               uint32_t rows, uint64_t *out,
                                                               uint64_t out_stride, uint32_t l_out,
                                                               Mod mq) {
  const uint64_t ct = blockIdx.x / 2;
  const int comp = blockIdx.x % 2;
  const uint64_t *Dc = D + (size_t)blockIdx.y * d_stride + ct * rows * kN;
  uint64_t *dst = out + (size_t)blockIdx.y * out_stride + (ct / l_out) * (2 * l_out) * 2 * kN +
                  (l_out + ct % l_out) * 2 * kN + (size_t)comp * kN;
  uint64_t v[8];
  ext_mac8(Dc, g_ptrs[blockIdx.y], rows, comp, mq, v);
  store_blocked(dst, v);
}

// Mux fold: x_c[comp] += INTT(sum_j D_j * G_j[comp]), x_c = cts + c*2N.
// G: this level's RGSW of the query at g + query * g_stride.
__global__ __launch_bounds__(kNttThreads) void k_ext_mac_intt_add(const uint64_t *D, uint64_t d_stride,
                                                                  const uint64_t *G, uint64_t g_stride,
                                                                  uint32_t rows, uint64_t *cts,
                                                                  uint64_t cts_stride, NttTab T, Mod mq) {
  __shared__ uint64_t lds[kNttLds];
  const uint64_t ct = blockIdx.x / 2;
  const int comp = blockIdx.x % 2;
  const uint64_t *Dc = D + (size_t)blockIdx.y * d_stride + ct * rows * kN;
  uint64_t v[8];
  ext_mac8(Dc, G + (size_t)blockIdx.y * g_stride, rows, comp, mq, v);
  ntt_inv(v, lds, T);
  uint64_t *x = cts + (size_t)blockIdx.y * cts_stride + ct * 2 * kN + (size_t)comp * kN;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int e = threadIdx.x + 256 * r;
    x[e] = add_mod(x[e], ntt_inv_final(v[r], T), T.q);
  }
}

// 128-bit multiply-accumulate of one transformed digit with RGSW row g (both
// components; blocked slots 8t .. 8t+7).
__device__ __forceinline__ void mac_row(U128 (&acc)[2][8], const uint64_t (&d)[8],
                                        const uint64_t *g) {
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    uint64_t gv[8];
    load_blocked(gv, g + (size_t)c * kN);
    mac128x8(acc[c], d, gv);
  }
}

// ---- fused variant: one workgroup per (ciphertext, query) ---------------------------
// All 2l digit NTTs of one ciphertext run back to back in the workgroup and
// both output components accumulate in registers, so digits never leave the
// CU.  Chosen when a level has enough ciphertexts to fill the device.
// Result: res[comp][r] = blocked slot 8t + r of sum_j D_j * G_j[comp] mod q.
// FOLD: G was prepared by k_rgsw_fold_last, so the last digit's row takes
// NTT(x) itself.  NTT_IN (conversion of NTT-form leaves, ys unused): xs is in
// NTT form; each polynomial is copied to top + poly * N as it is (the RGSW's
// top row), returns to coefficient form in registers for its digits, and is
// read again for the last row.  PAIR: lds holds 2 kNttLds words and the
// digits are transformed two at a time (ntt_fwd2).
template <bool FOLD = false, bool NTT_IN = false, bool PAIR = false>
__device__ __forceinline__ void ext_fused_core(const uint64_t *xs, const uint64_t *ys, const uint64_t *G,
                                               uint32_t l, uint32_t base_log2, uint64_t *lds,
                                               const NttTab &T, const Mod &mq, uint64_t (&res)[2][8],
                                               uint64_t *top = nullptr) {
  static_assert(FOLD || !NTT_IN, "NTT-form input needs the folded gadget matrix");
  const uint64_t mask = (1ULL << base_log2) - 1;
  U128 acc[2][8] = {};
#pragma unroll 1
  for (uint32_t poly = 0; poly < 2; ++poly) {
    const uint64_t *g_last = G + (size_t)(poly * l + l - 1) * 2 * kN;
    uint64_t v[8];
    if (NTT_IN) {
      load_blocked(v, xs + (size_t)poly * kN);  // canonical, slot order
      store_blocked(top + (size_t)poly * kN, v);
      ntt_inv(v, lds, T);
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = ntt_inv_final(v[r], T);  // coefficient t + 256 r
    } else {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int e = threadIdx.x + 256 * r;
        const uint64_t xa = xs[(size_t)poly * kN + e];
        v[r] = ys ? sub_mod(ys[(size_t)poly * kN + e], xa, T.q) : xa;
      }
    }
    const uint32_t ndig = FOLD ? l - 1 : l;  // digits transformed
    uint32_t kk = 0;
    if (PAIR) {
#pragma unroll 1
      for (; kk + 1 < ndig; kk += 2) {  // digits kk, kk + 1
        const uint32_t sh = (l - 1 - kk) * base_log2;
        uint64_t d[8], e[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          d[r] = (v[r] >> sh) & mask;
          e[r] = (v[r] >> (sh - base_log2)) & mask;
        }
        ntt_fwd2<true>(d, e, lds, T);  // digits < 2^base_log2 <= 2^32
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          d[r] = ntt_fwd_below_2p63(d[r], T.q);
          e[r] = ntt_fwd_below_2p63(e[r], T.q);
        }
        mac_row(acc, d, G + (size_t)(poly * l + kk) * 2 * kN);
        mac_row(acc, e, G + (size_t)(poly * l + kk + 1) * 2 * kN);
      }
    }
#pragma unroll 1
    for (; kk < ndig; ++kk) {
      const uint32_t sh = (l - 1 - kk) * base_log2;
      uint64_t d[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) d[r] = (v[r] >> sh) & mask;
      ntt_fwd<true>(d, lds, T);  // digits < 2^base_log2 <= 2^32
#pragma unroll
      for (int r = 0; r < 8; ++r) d[r] = ntt_fwd_below_2p63(d[r], T.q);
      mac_row(acc, d, G + (size_t)(poly * l + kk) * 2 * kN);
    }
    if (NTT_IN) {  // the last row takes NTT(x): the leaf as read
      load_blocked(v, xs + (size_t)poly * kN);
      mac_row(acc, v, g_last);
    } else if (FOLD) {
      ntt_fwd(v, lds, T);
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = ntt_fwd_below_2p63(v[r], T.q);
      mac_row(acc, v, g_last);
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int r = 0; r < 8; ++r) res[c][r] = reduce128(acc[c][r], mq);
}

// Mux fold, fused: x_c += INTT(ext(RGSW, y_c - x_c)) for ciphertext blockIdx.x
// of query blockIdx.y (x = cts + c*2N, y = cts + (c + half)*2N).
__global__ __launch_bounds__(kNttThreads) void k_mux_fused(uint64_t *cts, uint64_t cts_stride, uint64_t half,
                                                           const uint64_t *G, uint64_t g_stride, uint32_t l,
                                                           uint32_t base_log2, NttTab T, Mod mq) {
  __shared__ uint64_t lds[kNttLds];
  __shared__ uint64_t park[kN];  // component 1 waits here during component 0's INTT
  uint64_t *x = cts + (size_t)blockIdx.y * cts_stride + (size_t)blockIdx.x * 2 * kN;
  const uint64_t *y = x + half * 2 * kN;
  uint64_t res[2][8];
  ext_fused_core(x, y, G + (size_t)blockIdx.y * g_stride, l, base_log2, lds, T, mq, res);
#pragma unroll
  for (int r = 0; r < 8; ++r) park[8 * threadIdx.x + r] = res[1][r];
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {
    uint64_t v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = c == 0 ? res[0][r] : park[8 * threadIdx.x + r];
    ntt_inv(v, lds, T);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = threadIdx.x + 256 * r;
      x[(size_t)c * kN + e] = add_mod(x[(size_t)c * kN + e], ntt_inv_final(v[r], T), T.q);
    }
  }
}

// An RGSW for the NTT-form multiplexer: within each polynomial's block of l
// rows, row j < l - 1 becomes G_j - 2^((l-1-j)B) G_{l-1} (mod q), in place
// (row l - 1, which the others read, is left as it is).  Then, with
// C = NTT(c) and c = sum_j d_j 2^((l-1-j)B) exactly,
//   sum_j NTT(d_j) G_j = sum_{j<l-1} NTT(d_j) G'_j + C G_{l-1}  (mod q),
// because NTT(d_{l-1}) = C - sum_{j<l-1} 2^((l-1-j)B) NTT(d_j): the last
// digit needs no transform.  grid = (4 N / 256, nq): one thread per
// (polynomial block, component, slot); G: the RGSW of query 0, per-query
// stride g_stride.
__global__ __launch_bounds__(256) void k_rgsw_fold_last(uint64_t *G, uint64_t g_stride, uint32_t l,
                                                        uint32_t base_log2, Mod mq) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;  // (poly, comp, e) over 2 * 2 * N
  const uint32_t poly = i / (2 * kN), rest = i % (2 * kN);
  uint64_t *blk = G + (size_t)blockIdx.y * g_stride + (size_t)poly * l * 2 * kN + rest;
  const uint64_t last = blk[(size_t)(l - 1) * 2 * kN];
  for (uint32_t j = 0; j + 1 < l; ++j) {
    const uint64_t pw = 1ULL << ((l - 1 - j) * base_log2);  // (l - 1) B < 60: below q
    blk[(size_t)j * 2 * kN] = sub_mod(blk[(size_t)j * 2 * kN], mulmod_barrett(last, pw, mq), mq.q);
  }
}

// Mux fold on NTT-form ciphertexts (slot order):
//   X_c += ext(RGSW, y - x)_c,   y - x = INTT(Y - X) (canonical coefficients),
// the same sum as the reference's x + INTT(ext(RGSW, y - x)), kept in NTT
// form.  G is the level's RGSW after k_rgsw_fold_last, so of each
// polynomial's l digits only d_0 .. d_{l-2} are transformed and the last
// row multiplies C = Y - X itself.  Per pair: two inverse and 2(l - 1)
// forward transforms, against 2l forward and two inverse in coefficient form.
//
// FIRST = false: pair (blockIdx.x, blockIdx.x + half) of query blockIdx.y in
// cts (per-query stride cts_stride), result in place.
// FIRST = true (the first level, straight from the matrix-core product's
// NTT-form outputs out2[row][slot][column], column = 2 query + poly, ncols
// columns): pair (r, r + half) of query b, result in cts row r; with Q = 2,
// of queries b and b + 1 (an even batch).  Grid half * nq / Q, block index
// mod 8 = XCD: the workgroups that read one 128-byte line of out2 run on one
// XCD one after another, so the line comes from HBM once and the 16-byte
// loads at the 1 KiB slot stride hit that L2; with 8 | nq, an XCD also keeps
// to 8 queries over many row pairs, so their level RGSWs stay in its L2.
template <bool FIRST, int Q = 1, int R = 8>
__global__ __launch_bounds__(kNttThreads) void k_mux_fused_ntt(uint64_t *cts, uint64_t cts_stride, uint64_t half,
                                                               const uint64_t *out2, uint32_t ncols, uint32_t nq,
                                                               const uint64_t *G, uint64_t g_stride, uint32_t l,
                                                               uint32_t base_log2, NttTab T, Mod mq) {
  static_assert(Q == 1 || (FIRST && Q == 2), "query pairs only at the first level");
  // radix 8: digit transforms run two at a time (ntt_fwd2), each with half
  // of lds; the inverse transforms use the first half
  __shared__ uint64_t lds[(R == 8 ? 2 : 1) * kNttLds];
  __shared__ uint64_t coef[kN];  // y - x of the current polynomial; slots t + 256 j of thread t
  __shared__ uint64_t ymx[kN];   // Y - X of the current polynomial; slots 8t .. 8t+7 of thread t
  const int t = threadIdx.x;
  const uint32_t nqw = nq / Q;  // workgroups per row pair
  uint32_t r, b;                // row pair (r, r + half); first query b
  if (!FIRST) {
    r = blockIdx.x;
    b = blockIdx.y;
  } else if (Q == 2 && half % 8 == 0 && nqw % 4 == 0) {
    // units u = (group g of 4 workgroups = 8 queries, row pair r), r fastest;
    // XCD blockIdx % 8 takes the units u with u % 8 = its index, each unit's
    // 4 workgroups one after another: they share every 128-byte line of out2
    // they read (16 columns), and the group's 8 RGSWs stay in that XCD's L2
    // over its row pairs
    const uint32_t k = blockIdx.x >> 3, u = (k >> 2) * 8 + (blockIdx.x & 7);
    r = u % half;
    b = Q * (4 * (u / half) + (k & 3));
  } else if (half % 8 == 0) {
    const uint32_t k = blockIdx.x >> 3;
    r = 8 * (k / nqw) + (blockIdx.x & 7);
    b = Q * (k % nqw);
  } else {
    r = blockIdx.x / nqw;
    b = Q * (blockIdx.x % nqw);
  }
  const uint64_t *ox = out2 + (size_t)r * kN * ncols + 2 * b;  // FIRST: X of query b at ox[slot * ncols + poly]
  const uint64_t *oy = ox + (size_t)half * kN * ncols;
  // FIRST: Y - X of query b + i, parked in its row r + half, which no level
  // after the first reads (the levels fold rows below half)
  auto dpark = [&](uint32_t i) { return cts + (size_t)(b + i) * cts_stride + (size_t)(r + half) * 2 * kN; };
  const uint64_t q = T.q, mask = (1ULL << base_log2) - 1;
  // FIRST: Y - X (canonical NTT form) into the parks, and X into its
  // destination row for the final sum, instead of reading out2 at the slot
  // stride again
  if (FIRST && Q == 1) {
    uint64_t *x = cts + (size_t)b * cts_stride + (size_t)r * 2 * kN;
    uint64_t xs[2][8], ds[2][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const size_t o = (size_t)(8 * t + j) * ncols;
      const uint4 u = *reinterpret_cast<const uint4 *>(ox + o);
      const uint4 w = *reinterpret_cast<const uint4 *>(oy + o);
      xs[0][j] = ((uint64_t)u.y << 32) | u.x;
      xs[1][j] = ((uint64_t)u.w << 32) | u.z;
      ds[0][j] = sub_mod(((uint64_t)w.y << 32) | w.x, xs[0][j], q);
      ds[1][j] = sub_mod(((uint64_t)w.w << 32) | w.z, xs[1][j], q);
    }
    store_blocked(x, xs[0]);
    store_blocked(x + kN, xs[1]);
    store_blocked(dpark(0), ds[0]);  // read back below by this thread only
    store_blocked(dpark(0) + kN, ds[1]);
  } else if (FIRST) {
    // lanes 2i and 2i + 1 read queries b and b + 1 of one slot: the 32
    // contiguous bytes of each row, so a wave load touches 32 lines, not 64
    const uint32_t h = t & 1;
    uint64_t *xh = cts + (size_t)(b + h) * cts_stride + (size_t)r * 2 * kN;
    uint64_t *dp = dpark(h);
#pragma unroll 4
    for (int i = 0; i < kN / 128; ++i) {
      const uint32_t sl = 128 * i + (t >> 1);
      const size_t o = (size_t)sl * ncols + 2 * h;
      const uint4 u = *reinterpret_cast<const uint4 *>(ox + o);
      const uint4 w = *reinterpret_cast<const uint4 *>(oy + o);
      const uint64_t x0 = ((uint64_t)u.y << 32) | u.x, x1 = ((uint64_t)u.w << 32) | u.z;
      xh[sl] = x0;
      xh[kN + sl] = x1;
      dp[sl] = sub_mod(((uint64_t)w.y << 32) | w.x, x0, q);
      dp[kN + sl] = sub_mod(((uint64_t)w.w << 32) | w.z, x1, q);
    }
    __syncthreads();  // the parks are read below in other threads' slot order
  }
#pragma unroll 1
  for (int qi = 0; qi < Q; ++qi) {
    uint64_t *x = cts + (size_t)(b + qi) * cts_stride + (size_t)r * 2 * kN;
    const uint64_t *Gb = G + (size_t)(b + qi) * g_stride;
    // this thread's slots 8t .. 8t+7 of Y - X
    auto y_minus_x = [&](uint32_t poly, uint64_t(&v)[8]) {
      if (FIRST) {
        load_blocked(v, dpark(qi) + (size_t)poly * kN);
      } else {
        uint64_t u[8];
        load_blocked(u, x + (size_t)poly * kN);
        load_blocked(v, x + half * 2 * kN + (size_t)poly * kN);
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = sub_mod(v[j], u[j], q);
      }
    };
    U128 acc[2][8] = {};
#pragma unroll 1
    for (uint32_t poly = 0; poly < 2; ++poly) {
      {
        uint64_t v[8];
        y_minus_x(poly, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) ymx[8 * t + j] = v[j];  // read back by this thread only
        ntt_inv_sched<R>(v, lds, T);
#pragma unroll
        for (int j = 0; j < 8; ++j) coef[t + 256 * j] = ntt_inv_final(v[j], T);
      }
      const uint64_t *grow = Gb + (size_t)(poly * l) * 2 * kN;  // row kk at grow + kk * 2N
      uint32_t kk = 0;
      if (R == 8) {
#pragma unroll 1
        for (; kk + 2 < l; kk += 2) {  // digits kk, kk + 1 (of 0 .. l - 2) together
          const uint32_t sh = (l - 1 - kk) * base_log2;
          uint64_t d[8], e[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const uint64_t c = coef[t + 256 * j];
            d[j] = (c >> sh) & mask;
            e[j] = (c >> (sh - base_log2)) & mask;
          }
          ntt_fwd2<true>(d, e, lds, T);  // digits < 2^base_log2
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            d[j] = ntt_fwd_below_2p63(d[j], q);
            e[j] = ntt_fwd_below_2p63(e[j], q);
          }
          mac_row(acc, d, grow + (size_t)kk * 2 * kN);
          mac_row(acc, e, grow + (size_t)(kk + 1) * 2 * kN);
        }
      }
#pragma unroll 1
      for (; kk + 1 < l; ++kk) {
        const uint32_t sh = (l - 1 - kk) * base_log2;
        uint64_t d[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) d[j] = (coef[t + 256 * j] >> sh) & mask;
        ntt_fwd_sched<R, true>(d, lds, T);  // digits < 2^base_log2
#pragma unroll
        for (int j = 0; j < 8; ++j) d[j] = ntt_fwd_below_2p63(d[j], q);
        mac_row(acc, d, grow + (size_t)kk * 2 * kN);
      }
      uint64_t d[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) d[j] = ymx[8 * t + j];
      mac_row(acc, d, grow + (size_t)(l - 1) * 2 * kN);
    }
    // X + E into cts row r.  The fence keeps these loads after the loop:
    // hoisted, they would stay live through it.
    asm volatile("" ::: "memory");
    uint64_t w[2][8];
    load_blocked(w[0], x);
    load_blocked(w[1], x + kN);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
      for (int j = 0; j < 8; ++j) w[c][j] = add_mod(w[c][j], reduce128(acc[c][j], mq), q);
      store_blocked(x + (size_t)c * kN, w[c]);
    }
  }
}

// NTT-form ciphertexts (slot order) back to canonical coefficient form, in
// place: polynomial blockIdx.x of query blockIdx.y.  Every thread has read its
// slots before the first barrier inside ntt_inv, and writes after the last.
__global__ __launch_bounds__(kNttThreads) void k_cts_to_coeff(uint64_t *cts, uint64_t cts_stride, NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  uint64_t *p = cts + (size_t)blockIdx.y * cts_stride + (size_t)blockIdx.x * kN;
  uint64_t x[8];
  load_blocked(x, p);
  ntt_inv(x, lds, T);
#pragma unroll
  for (int j = 0; j < 8; ++j) p[threadIdx.x + 256 * j] = ntt_inv_final(x[j], T);
}

// Conversion bottom rows, fused: row l + c%l of RGSW c/l = ext(RGSW(s) key, leaf c).
// FOLD: the key was prepared by k_rgsw_fold_last.  NTT_IN: the leaves are in
// NTT form, and this kernel also writes the top row c%l (the leaf itself).
template <bool FOLD, bool NTT_IN>
__global__ __launch_bounds__(kNttThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_convert_fused(const uint64_t *leaves, uint64_t leaf_stride,
                                                               const uint64_t *const *g_ptrs, uint32_t lk,
                                                               uint32_t base_log2_key, uint64_t *out,
                                                               uint64_t out_stride, uint32_t l_out, NttTab T,
                                                               Mod mq) {
  __shared__ uint64_t lds[2 * kNttLds];  // digit transforms in pairs
  const uint64_t ct = blockIdx.x;
  const uint64_t *src = leaves + (size_t)blockIdx.y * leaf_stride + ct * 2 * kN;
  uint64_t *rgsw = out + (size_t)blockIdx.y * out_stride + (ct / l_out) * (2 * l_out) * 2 * kN;
  uint64_t *dst = rgsw + (l_out + ct % l_out) * 2 * kN;
  uint64_t res[2][8];
  ext_fused_core<FOLD, NTT_IN, true>(src, nullptr, g_ptrs[blockIdx.y], lk, base_log2_key, lds, T, mq, res,
                               rgsw + (ct % l_out) * 2 * kN);
  store_blocked(dst, res[0]);
  store_blocked(dst + kN, res[1]);
}

// Conversion rows 0..l-1: NTT of both polys of leaf c into
//   out + (c / l) * 2l * 2N + (c % l) * 2N + poly * N   (per query);
// NTT_IN: the leaves are in NTT form already (a copy).
template <bool NTT_IN>
__global__ __launch_bounds__(kNttThreads) void k_rgsw_top_rows(const uint64_t *leaves, uint64_t leaf_stride,
                                                               uint64_t *out, uint64_t out_stride,
                                                               uint32_t l, NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  const uint64_t ct = blockIdx.x / 2;
  const int poly = blockIdx.x % 2;
  const uint64_t *src = leaves + (size_t)blockIdx.y * leaf_stride + ct * 2 * kN + (size_t)poly * kN;
  uint64_t *dst = out + (size_t)blockIdx.y * out_stride + (ct / l) * (2 * l) * 2 * kN + (ct % l) * 2 * kN +
                  (size_t)poly * kN;
  uint64_t x[8];
  if (NTT_IN) {
    load_blocked(x, src);
  } else {
    load_natural(x, src);
    ntt_fwd(x, lds, T);
#pragma unroll
    for (int r = 0; r < 8; ++r) x[r] = ntt_fwd_final(x[r], T.q);
  }
  store_blocked(dst, x);
}

}  // namespace onion
