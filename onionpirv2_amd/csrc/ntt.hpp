// ntt.hpp -- negacyclic NTT / INTT of one 2048-point polynomial, computed by
// one workgroup of 256 threads with 8 coefficients per thread in registers.
//
// Forward: Cooley-Tukey with bit-reversed twiddles psi^bitrev(k) and lazy
// butterflies: the twiddle product (mul_shoup_4q, a Shoup product with a
// truncated quotient) is in [0, 4q) for any 64-bit input, so each stage grows
// the bound on the values by 4q.  The moduli are q = 2^60 - c with c < 2^28,
// so 8q = 2^63 - 8c: the input half of a butterfly loses 8q when its top bit
// is set (reduce_2p63: a bit test and a multiply-add) in stages 3, 5, 7 and 9
// (of 0..10), which leaves it below 2^63, and the values stay below
// 2^63 + 8q < 2^64 throughout (ntt_fwd_final brings them to [0, q)).  Output slot
// e = a(psi^(2*bitrev(e)+1)), the order SEAL's ntt_negacyclic_harvey
// produces.  Inverse: Gentleman-Sande with psi^-bitrev(k); the sum half is
// brought below 2^62 + 12c by reduce_2p62 (4q = 2^62 - 4c) and the difference
// half is a twiddle product, below 4q; the N^-1 scale is applied by the caller
// as it reads the result out (ntt_inv_final).
//
// The 11 radix-2 stages run as four register passes (3 + 3 + 3 + 2 stages).
// In a pass with top span H a thread owns the 8 coefficients b + j*H/4,
// j = 0..7, which form one radix-8 butterfly: stage H pairs (j, j+4) under
// twiddle index m, stage H/2 pairs (j, j+2) under 2m + j/4, stage H/4 pairs
// (j, j+1) under 4m + j/2.  Between passes the coefficients are exchanged
// through LDS (index padded by i/16 against bank conflicts); each pass loads
// its 7 twiddles (and Shoup companions) at once.
//
// Register layouts (t = threadIdx.x):
//   "natural"   x[j] = coefficient t + 256*j   (coalesced global access)
//   "blocked"   x[j] = slot 8*t + j            (64 contiguous bytes per thread)
// ntt_fwd: natural -> blocked.  ntt_inv: blocked -> natural.
#pragma once

#include "arith.hpp"

namespace onion {

constexpr int kNttThreads = 256;
constexpr int kNttLds = 2048 + 2048 / 16;  // padded LDS words per transform

__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }

// v - 8q when bit 63 of v is set, else v (q = 2^60 - c, c < 2^28, so
// 8q = 2^63 - 8c and the subtraction is clearing the bit and adding 8c):
// below 2^63 for every v < 2^63 + 8q, and congruent to v mod q.
__device__ __forceinline__ uint64_t reduce_2p63(uint64_t v, uint64_t q) {
  const uint32_t c8 = (uint32_t)((1ULL << 63) - 8 * q);
  return (uint64_t)(uint32_t)(v >> 63) * c8 + (v & 0x7FFFFFFFFFFFFFFFULL);
}

// Forward butterfly on values below 2^63 + 8q; CORRECT first brings x below
// 2^63 (reduce_2p63).  Outputs are below (bound of u) + 4q.
template <bool CORRECT>
__device__ __forceinline__ void ct_bfly(uint64_t &x, uint64_t &y, uint64_t w, uint64_t ws, uint64_t q,
                                        uint64_t two_q) {
  const uint64_t u = CORRECT ? reduce_2p63(x, q) : x;
  const uint64_t v = mul_shoup_4q(y, w, ws, q);
  x = u + v;
  y = u + 2 * two_q - v;
}

// The first forward stage for inputs below 2^32 (digit polynomials): u needs
// no range correction and the Shoup quotient floor(y ws / 2^64) needs only the
// two products of y with the halves of ws (exact: y ws_hi + floor(y ws_lo / 2^32)
// < 2^64), so the result is the same as ct_bfly's with fewer multiplies.
__device__ __forceinline__ void ct_bfly_small(uint64_t &x, uint64_t &y, uint64_t w, uint64_t ws, uint64_t q,
                                              uint64_t two_q) {
  const uint64_t u = x;
  const uint32_t y32 = (uint32_t)y;
  const uint64_t t = (uint64_t)y32 * (ws >> 32) + (((uint64_t)y32 * (uint32_t)ws) >> 32);
  const uint64_t v = mul_sub_lo(y32, w, t >> 32, q);
  x = u + v;
  y = u + two_q - v;
}

// (v mod 2^62) + (v >> 62) 4c: congruent to v mod q (4q = 2^62 - 4c) and
// below 2^62 + 12c for every 64-bit v.
__device__ __forceinline__ uint64_t reduce_2p62(uint64_t v, uint64_t q) {
  const uint32_t c4 = (uint32_t)((1ULL << 62) - 4 * q);
  return (uint64_t)(uint32_t)(v >> 62) * c4 + (v & ((1ULL << 62) - 1));
}

// Inverse butterfly on values below 8q; outputs below 2^62 + 12c (< 8q).
__device__ __forceinline__ void gs_bfly(uint64_t &x, uint64_t &y, uint64_t w, uint64_t ws,
                                        uint64_t q, uint64_t two_q) {
  const uint64_t u = x, v = y;
  x = reduce_2p62(u + v, q);
  y = mul_shoup_4q(u + 4 * two_q - v, w, ws, q);
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int mask) {
  const int lo = __shfl_xor((int)(uint32_t)v, mask, 64);
  const int hi = __shfl_xor((int)(uint32_t)(v >> 32), mask, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Twiddles of one radix-8 pass rooted at index m: w[0] = tab[m],
// w[1..2] = tab[2m..2m+1], w[3..6] = tab[4m..4m+3].
struct Tw8 {
  uint64_t w[7], s[7];
};

__device__ __forceinline__ void load_tw8(Tw8 &t, const uint64_t *tab, const uint64_t *tabs, int m) {
  // keep each pass's twiddle loads in that pass: without this fence the
  // scheduler hoists all four passes' loads to the top and the live twiddles
  // alone take ~110 VGPRs in kernels that run several transforms
  __builtin_amdgcn_sched_barrier(0);
  t.w[0] = tab[m];
  t.s[0] = tabs[m];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    t.w[1 + i] = tab[2 * m + i];
    t.s[1 + i] = tabs[2 * m + i];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    t.w[3 + i] = tab[4 * m + i];
    t.s[3 + i] = tabs[4 * m + i];
  }
}

// forward radix-8 (three CT stages); `first` = false skips the span-H stage;
// SMALL: every input is below 2^32 (the span-H stage uses ct_bfly_small);
// bit s of CMASK reduces the inputs of stage s (0 = span H, 1 = H/2, 2 = H/4)
template <int CMASK, bool SMALL = false>
__device__ __forceinline__ void fwd8(uint64_t (&x)[8], const Tw8 &t, uint64_t q, bool first) {
  const uint64_t two_q = 2 * q;
  if (first) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (SMALL)
        ct_bfly_small(x[j], x[j + 4], t.w[0], t.s[0], q, two_q);
      else
        ct_bfly<(CMASK & 1) != 0>(x[j], x[j + 4], t.w[0], t.s[0], q, two_q);
    }
  }
#pragma unroll
  for (int j = 0; j < 8; j = (j & 1) ? j + 3 : j + 1)  // j = 0, 1, 4, 5
    ct_bfly<(CMASK & 2) != 0>(x[j], x[j + 2], t.w[1 + (j >> 2)], t.s[1 + (j >> 2)], q, two_q);
#pragma unroll
  for (int j = 0; j < 8; j += 2)
    ct_bfly<(CMASK & 4) != 0>(x[j], x[j + 1], t.w[3 + (j >> 1)], t.s[3 + (j >> 1)], q, two_q);
}

// inverse radix-8 (three GS stages, reverse order); `last` = false skips the span-H stage
__device__ __forceinline__ void inv8(uint64_t (&x)[8], const Tw8 &t, uint64_t q, bool last) {
  const uint64_t two_q = 2 * q;
#pragma unroll
  for (int j = 0; j < 8; j += 2) gs_bfly(x[j], x[j + 1], t.w[3 + (j >> 1)], t.s[3 + (j >> 1)], q, two_q);
#pragma unroll
  for (int j = 0; j < 8; j = (j & 1) ? j + 3 : j + 1)
    gs_bfly(x[j], x[j + 2], t.w[1 + (j >> 2)], t.s[1 + (j >> 2)], q, two_q);
  if (last) {
#pragma unroll
    for (int j = 0; j < 4; ++j) gs_bfly(x[j], x[j + 4], t.w[0], t.s[0], q, two_q);
  }
}

// base coefficient of thread t in the pass whose element stride is 2^s0
__device__ __forceinline__ int pass_base(int t, int s0) {
  return (t & ((1 << s0) - 1)) | ((t >> s0) << (s0 + 3));
}

// Synchronisation of an exchange that stays inside one wave: thread t's
// coefficients in passes B, C and D all lie in [512 w, 512 w + 512) for its
// wave w = t / 64, so in the B <-> C and C <-> D exchanges every value a wave
// reads was written by the same wave, and a wave's LDS operations complete in
// issue order.  The fences only stop the compiler from moving LDS accesses
// across the exchange; no wave waits for another.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void lds_put(uint64_t *lds, const uint64_t (&x)[8], int b, int stride) {
#pragma unroll
  for (int j = 0; j < 8; ++j) lds[lds_pad(b + j * stride)] = x[j];
}
__device__ __forceinline__ void lds_get(const uint64_t *lds, uint64_t (&x)[8], int b, int stride) {
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = lds[lds_pad(b + j * stride)];
}

// Forward NTT.  Entry: natural layout, values < 4q (SMALL: < 2^32, as digit
// polynomials are).  Exit: blocked layout, values below 2^63 + 8q.
// Bounds by stage (R = 2^63): 8q 12q 16q | R+4q R+8q R+4q | R+8q R+4q R+8q |
// R+4q R+8q, the reductions being in stages 3 and 5 (pass B, spans 128 and
// 32), stage 7 (pass C, span 8) and stage 9 (pass D, span 2).  Uses
// lds[kNttLds]; begins with a barrier so the buffer may have been read just
// before the call.  Only the A -> B exchange crosses waves (two workgroup
// barriers); B -> C and C -> D synchronise per wave.  On return other waves
// may still be reading their part of the buffer: the next writer of it
// starts with a barrier (every transform does).
template <bool SMALL = false>
__device__ __forceinline__ void ntt_fwd(uint64_t (&x)[8], uint64_t *lds, const NttTab &T) {
  const int t = threadIdx.x;
  const uint64_t q = T.q;
  Tw8 tw;
  // pass A: spans 1024, 512, 256 (stride 256, b = t): twiddle root m = 1
  load_tw8(tw, T.rp, T.rps, 1);
  fwd8<0, SMALL>(x, tw, q, true);
  // pass B: spans 128, 64, 32 (stride 32): m = 8 + t/32
  const int bB = pass_base(t, 5);
  load_tw8(tw, T.rp, T.rps, 8 + (t >> 5));
  __syncthreads();
  lds_put(lds, x, t, 256);
  __syncthreads();
  lds_get(lds, x, bB, 32);
  fwd8<5>(x, tw, q, true);
  // pass C: spans 16, 8, 4 (stride 4): m = 64 + t/4
  const int bC = pass_base(t, 2);
  load_tw8(tw, T.rp, T.rps, 64 + (t >> 2));
  wave_sync();
  lds_put(lds, x, bB, 32);
  wave_sync();
  lds_get(lds, x, bC, 4);
  fwd8<2>(x, tw, q, true);
  // pass D: spans 2, 1 on slots 8t .. 8t+7: m = 256 + t
  load_tw8(tw, T.rp, T.rps, 256 + t);
  wave_sync();
  lds_put(lds, x, bC, 4);
  wave_sync();
  lds_get(lds, x, 8 * t, 1);
  fwd8<2>(x, tw, q, false);
}

// Two forward transforms at once: the same passes, twiddle loads and
// barriers serve both, and each pass has twice the independent butterflies.
// x uses lds2[0, kNttLds), y lds2[kNttLds, 2 kNttLds).  Same entry and exit
// conditions as ntt_fwd.
template <bool SMALL = false>
__device__ __forceinline__ void ntt_fwd2(uint64_t (&x)[8], uint64_t (&y)[8], uint64_t *lds2, const NttTab &T) {
  const int t = threadIdx.x;
  const uint64_t q = T.q;
  uint64_t *l0 = lds2, *l1 = lds2 + kNttLds;
  Tw8 tw;
  load_tw8(tw, T.rp, T.rps, 1);
  fwd8<0, SMALL>(x, tw, q, true);
  fwd8<0, SMALL>(y, tw, q, true);
  const int bB = pass_base(t, 5);
  load_tw8(tw, T.rp, T.rps, 8 + (t >> 5));
  __syncthreads();
  lds_put(l0, x, t, 256);
  lds_put(l1, y, t, 256);
  __syncthreads();
  lds_get(l0, x, bB, 32);
  lds_get(l1, y, bB, 32);
  fwd8<5>(x, tw, q, true);
  fwd8<5>(y, tw, q, true);
  const int bC = pass_base(t, 2);
  load_tw8(tw, T.rp, T.rps, 64 + (t >> 2));
  wave_sync();
  lds_put(l0, x, bB, 32);
  lds_put(l1, y, bB, 32);
  wave_sync();
  lds_get(l0, x, bC, 4);
  lds_get(l1, y, bC, 4);
  fwd8<2>(x, tw, q, true);
  fwd8<2>(y, tw, q, true);
  load_tw8(tw, T.rp, T.rps, 256 + t);
  wave_sync();
  lds_put(l0, x, bC, 4);
  lds_put(l1, y, bC, 4);
  wave_sync();
  lds_get(l0, x, 8 * t, 1);
  lds_get(l1, y, 8 * t, 1);
  fwd8<2>(x, tw, q, false);
  fwd8<2>(y, tw, q, false);
}

// Inverse NTT without the N^-1 scale.  Entry: blocked layout, values < 8q.
// Exit: natural layout, values below 2^62 + 12c.  Begins with a barrier.
// D -> C and C -> B synchronise per wave; B -> A crosses waves (barrier).
__device__ __forceinline__ void ntt_inv(uint64_t (&x)[8], uint64_t *lds, const NttTab &T) {
  const int t = threadIdx.x;
  const uint64_t q = T.q;
  Tw8 tw;
  // pass D: spans 1, 2 on slots 8t .. 8t+7
  load_tw8(tw, T.irp, T.irps, 256 + t);
  inv8(x, tw, q, false);
  // pass C: spans 4, 8, 16 (stride 4)
  const int bC = pass_base(t, 2);
  load_tw8(tw, T.irp, T.irps, 64 + (t >> 2));
  __syncthreads();
  lds_put(lds, x, 8 * t, 1);
  wave_sync();
  lds_get(lds, x, bC, 4);
  inv8(x, tw, q, true);
  // pass B: spans 32, 64, 128 (stride 32)
  const int bB = pass_base(t, 5);
  load_tw8(tw, T.irp, T.irps, 8 + (t >> 5));
  wave_sync();
  lds_put(lds, x, bC, 4);
  wave_sync();
  lds_get(lds, x, bB, 32);
  inv8(x, tw, q, true);
  // pass A: spans 256, 512, 1024 (stride 256)
  load_tw8(tw, T.irp, T.irps, 1);
  wave_sync();
  lds_put(lds, x, bB, 32);
  __syncthreads();
  lds_get(lds, x, t, 256);
  inv8(x, tw, q, true);
}

// ---- radix-4 schedule: 2 + 2 + 2 + 2 + 3 stages -------------------------------
// The same butterflies, stage order, twiddles and lazy-reduction points as
// ntt_fwd / ntt_inv (the results are bit-identical), in five register passes
// and four LDS exchanges instead of four and three.  In a radix-4 pass with
// element stride 2^s0 (top span H = 2^(s0+1)) a thread owns two groups,
// g = 0, 1 with group index t + 256 g, each the four coefficients
// b, b + H/2, b + H, b + 3H/2 (x[4g + k] = b + k 2^s0): stage H pairs
// (k, k + 2) under twiddle m = N/(2H) + b/(2H), stage H/2 pairs (0, 1) under
// 2m and (2, 3) under 2m + 1.

// base coefficient of group gi in the radix-4 pass whose element stride is 2^s0
__device__ __forceinline__ int r4_base(int gi, int s0) {
  return (gi & ((1 << s0) - 1)) | ((gi >> s0) << (s0 + 2));
}

struct Tw4 {
  uint64_t w[6], s[6];  // group g: w[3g] span H, w[3g + 1 .. 3g + 2] span H/2
};

__device__ __forceinline__ void load_tw4(Tw4 &tw, const uint64_t *tab, const uint64_t *tabs, int t, int s0) {
  __builtin_amdgcn_sched_barrier(0);  // as in load_tw8
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int m = (1 << (9 - s0)) + (r4_base(t + 256 * g, s0) >> (s0 + 2));
    tw.w[3 * g] = tab[m];
    tw.s[3 * g] = tabs[m];
    tw.w[3 * g + 1] = tab[2 * m];
    tw.s[3 * g + 1] = tabs[2 * m];
    tw.w[3 * g + 2] = tab[2 * m + 1];
    tw.s[3 * g + 2] = tabs[2 * m + 1];
  }
}

// forward radix-4 on both groups; bit s of CMASK reduces the inputs of stage
// s (0 = span H, 1 = H/2); SMALL: inputs below 2^32 (span-H stage)
template <int CMASK, bool SMALL = false>
__device__ __forceinline__ void fwd4(uint64_t (&x)[8], const Tw4 &tw, uint64_t q) {
  const uint64_t two_q = 2 * q;
#pragma unroll
  for (int g = 0; g < 2; ++g) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (SMALL)
        ct_bfly_small(x[4 * g + k], x[4 * g + k + 2], tw.w[3 * g], tw.s[3 * g], q, two_q);
      else
        ct_bfly<(CMASK & 1) != 0>(x[4 * g + k], x[4 * g + k + 2], tw.w[3 * g], tw.s[3 * g], q, two_q);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
      ct_bfly<(CMASK & 2) != 0>(x[4 * g + 2 * k], x[4 * g + 2 * k + 1], tw.w[3 * g + 1 + k], tw.s[3 * g + 1 + k],
                                q, two_q);
  }
}

// inverse radix-4 on both groups (span H/2, then H)
__device__ __forceinline__ void inv4(uint64_t (&x)[8], const Tw4 &tw, uint64_t q) {
  const uint64_t two_q = 2 * q;
#pragma unroll
  for (int g = 0; g < 2; ++g) {
#pragma unroll
    for (int k = 0; k < 2; ++k)
      gs_bfly(x[4 * g + 2 * k], x[4 * g + 2 * k + 1], tw.w[3 * g + 1 + k], tw.s[3 * g + 1 + k], q, two_q);
#pragma unroll
    for (int k = 0; k < 2; ++k) gs_bfly(x[4 * g + k], x[4 * g + k + 2], tw.w[3 * g], tw.s[3 * g], q, two_q);
  }
}

__device__ __forceinline__ void lds_put4(uint64_t *lds, const uint64_t (&x)[8], int t, int s0) {
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int k = 0; k < 4; ++k) lds[lds_pad(r4_base(t + 256 * g, s0) + (k << s0))] = x[4 * g + k];
}
__device__ __forceinline__ void lds_get4(const uint64_t *lds, uint64_t (&x)[8], int t, int s0) {
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int k = 0; k < 4; ++k) x[4 * g + k] = lds[lds_pad(r4_base(t + 256 * g, s0) + (k << s0))];
}

// ntt_fwd in the radix-4 schedule (same entry and exit layouts and bounds).
// The first pass's groups are coefficients t + 256 (g + 2k), a register
// permutation of the natural layout.
template <bool SMALL = false>
__device__ __forceinline__ void ntt_fwd_r4(uint64_t (&x)[8], uint64_t *lds, const NttTab &T) {
  const int t = threadIdx.x;
  const uint64_t q = T.q;
  uint64_t y[8];
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int k = 0; k < 4; ++k) y[4 * g + k] = x[g + 2 * k];
  Tw4 tw;
  load_tw4(tw, T.rp, T.rps, t, 9);  // spans 1024, 512 (stages 0, 1)
  fwd4<0, SMALL>(y, tw, q);
  load_tw4(tw, T.rp, T.rps, t, 7);  // spans 256, 128 (stages 2, 3: reduce before 3)
  __syncthreads();
  lds_put4(lds, y, t, 9);
  __syncthreads();
  lds_get4(lds, y, t, 7);
  fwd4<2>(y, tw, q);
  load_tw4(tw, T.rp, T.rps, t, 5);  // spans 64, 32 (stages 4, 5: reduce before 5)
  __syncthreads();
  lds_put4(lds, y, t, 7);
  __syncthreads();
  lds_get4(lds, y, t, 5);
  fwd4<2>(y, tw, q);
  load_tw4(tw, T.rp, T.rps, t, 3);  // spans 16, 8 (stages 6, 7: reduce before 7)
  __syncthreads();
  lds_put4(lds, y, t, 5);
  __syncthreads();
  lds_get4(lds, y, t, 3);
  fwd4<2>(y, tw, q);
  Tw8 tw8;  // spans 4, 2, 1 on slots 8t .. 8t+7 (stages 8, 9, 10: reduce before 9)
  load_tw8(tw8, T.rp, T.rps, 256 + t);
  __syncthreads();
  lds_put4(lds, y, t, 3);
  __syncthreads();
  lds_get(lds, x, 8 * t, 1);
  fwd8<2>(x, tw8, q, true);
}

// ntt_inv in the radix-4 schedule (same entry and exit layouts and bounds).
__device__ __forceinline__ void ntt_inv_r4(uint64_t (&x)[8], uint64_t *lds, const NttTab &T) {
  const int t = threadIdx.x;
  const uint64_t q = T.q;
  {
    Tw8 tw8;  // spans 1, 2, 4 on slots 8t .. 8t+7
    load_tw8(tw8, T.irp, T.irps, 256 + t);
    inv8(x, tw8, q, true);
  }
  Tw4 tw;
  load_tw4(tw, T.irp, T.irps, t, 3);  // spans 8, 16
  __syncthreads();
  lds_put(lds, x, 8 * t, 1);
  __syncthreads();
  lds_get4(lds, x, t, 3);
  inv4(x, tw, q);
#pragma unroll
  for (int s0 = 5; s0 <= 9; s0 += 2) {  // spans 32, 64 | 128, 256 | 512, 1024
    load_tw4(tw, T.irp, T.irps, t, s0);
    __syncthreads();
    lds_put4(lds, x, t, s0 - 2);
    __syncthreads();
    lds_get4(lds, x, t, s0);
    inv4(x, tw, q);
  }
  uint64_t y[8];
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int k = 0; k < 4; ++k) y[g + 2 * k] = x[4 * g + k];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = y[j];
}

// Transforms in the schedule R (8: radix-8 passes, 4: radix-4 passes).
template <int R, bool SMALL = false>
__device__ __forceinline__ void ntt_fwd_sched(uint64_t (&x)[8], uint64_t *lds, const NttTab &T) {
  static_assert(R == 8 || R == 4, "radix 8 or 4");
  if constexpr (R == 4)
    ntt_fwd_r4<SMALL>(x, lds, T);
  else
    ntt_fwd<SMALL>(x, lds, T);
}
template <int R>
__device__ __forceinline__ void ntt_inv_sched(uint64_t (&x)[8], uint64_t *lds, const NttTab &T) {
  static_assert(R == 8 || R == 4, "radix 8 or 4");
  if constexpr (R == 4)
    ntt_inv_r4(x, lds, T);
  else
    ntt_inv(x, lds, T);
}

// Canonical residue of any 64-bit value (q = 2^60 - c, c < 2^28).
__device__ __forceinline__ uint64_t canonical_2p60(uint64_t v, uint64_t q) {
  // 2^60 = c mod q: v = 2^60 v_hi + v_lo = v_lo + c v_hi < 2^60 + 15c < 2q
  const uint32_t c = (uint32_t)((1ULL << 60) - q);
  const uint64_t r = (uint64_t)(uint32_t)(v >> 60) * c + (v & ((1ULL << 60) - 1));
  return r >= q ? r - q : r;
}

// Canonical value of a forward-NTT output.
__device__ __forceinline__ uint64_t ntt_fwd_final(uint64_t v, uint64_t q) { return canonical_2p60(v, q); }

// A forward-NTT output (below 2^63 + 8q) brought below 2^63 by one
// conditional subtraction of 8q, enough for the 128-bit multiply-accumulate of
// external products (mac128x4 with a canonical second operand), whose 128-bit
// reduction at the end (reduce128) returns the canonical residue.  (reduce_2p63 does the
// same in fewer instructions, but there the compiler spills k_convert_fused.)
__device__ __forceinline__ uint64_t ntt_fwd_below_2p63(uint64_t v, uint64_t q) { return v >= 8 * q ? v - 8 * q : v; }

// Canonical value of an inverse-NTT output, not scaled (for inputs that carry
// the N^-1 factor already).
__device__ __forceinline__ uint64_t ntt_inv_canonical(uint64_t v, uint64_t q) { return canonical_2p60(v, q); }

// Canonical value of an inverse-NTT output, scaled by N^-1 (the Shoup product
// takes any input and lands in [0, 2q)).
__device__ __forceinline__ uint64_t ntt_inv_final(uint64_t v, const NttTab &T) {
  const uint64_t r = mul_shoup_lazy(v, T.n_inv, T.n_inv_s, T.q);
  return r >= T.q ? r - T.q : r;
}

// natural-layout global load / blocked-layout global store helpers
__device__ __forceinline__ void load_natural(uint64_t (&x)[8], const uint64_t *src) {
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = src[threadIdx.x + 256 * j];
}
__device__ __forceinline__ void load_blocked(uint64_t (&x)[8], const uint64_t *src) {
  const uint4 *p = reinterpret_cast<const uint4 *>(src + 8 * threadIdx.x);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint4 v = p[i];
    x[2 * i] = ((uint64_t)v.y << 32) | v.x;
    x[2 * i + 1] = ((uint64_t)v.w << 32) | v.z;
  }
}
__device__ __forceinline__ void store_blocked(uint64_t *dst, const uint64_t (&x)[8]) {
  uint4 *p = reinterpret_cast<uint4 *>(dst + 8 * threadIdx.x);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    p[i] = make_uint4((uint32_t)x[2 * i], (uint32_t)(x[2 * i] >> 32), (uint32_t)x[2 * i + 1],
                      (uint32_t)(x[2 * i + 1] >> 32));
}

}  // namespace onion
