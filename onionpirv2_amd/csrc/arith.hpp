// arith.hpp -- exact 64-bit modular arithmetic on CDNA4 (gfx950).
//
// Every value is an unsigned residue below a 60-bit prime.  64x64->128
// products are built from 32-bit multiplies (v_mad_u64_u32); nothing here
// touches floating point or the matrix cores.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace onion {

// Modulus with the two-word Barrett ratio floor(2^128 / q).
struct Mod {
  uint64_t q;
  uint64_t r_lo, r_hi;
};

// NTT tables for one prime, resident in device memory.
//   rp[k]  = psi^bitrev(k),  rps  = Shoup companions floor(rp * 2^64 / q)
//   irp[k] = psi^-bitrev(k), irps = Shoup companions
struct NttTab {
  const uint64_t *rp, *rps, *irp, *irps;
  uint64_t q, n_inv, n_inv_s;
};

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// x w - h q mod 2^64, computed as x w + h (2^64 - q): two 32x32->64
// multiply-adds with 64-bit addends for the low word and four 32-bit
// multiplies for the high word, where the subtraction would take a borrow
// chain (two instructions and a wait state).  The empty asm keeps the compiler
// from refolding the high-word sum into a 64-bit multiply-add.
__device__ __forceinline__ uint64_t mul_sub_lo(uint64_t x, uint64_t w, uint64_t h, uint64_t q) {
  const uint64_t nq = 0 - q;
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32), w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32);
  const uint32_t h0 = (uint32_t)h, h1 = (uint32_t)(h >> 32), n0 = (uint32_t)nq, n1 = (uint32_t)(nq >> 32);
  const uint64_t r = (uint64_t)h0 * n0 + (uint64_t)x0 * w0;
  uint32_t c = x0 * w1 + x1 * w0 + h0 * n1 + h1 * n0;
  asm("" : "+v"(c));
  return r + ((uint64_t)c << 32);
}

// x * w mod q in [0, 2q) for any 64-bit x (Shoup; q < 2^62)
__device__ __forceinline__ uint64_t mul_shoup_lazy(uint64_t x, uint64_t w, uint64_t ws,
                                                   uint64_t q) {
  return mul_sub_lo(x, w, mulhi64(x, ws), q);
}

// x * w mod q in [0, 4q) for any 64-bit x (q < 2^62): Shoup with a truncated
// quotient.  floor(x ws / 2^64) = x1 s1 + floor((x1 s0 + x0 s1 + hi(x0 s0)) / 2^32)
// (x = x1 2^32 + x0, ws = s1 2^32 + s0); x1 s1 + hi(x1 s0) + hi(x0 s1) is at
// most 2 below it, so the remainder is below 4q instead of 2q.  Three fewer
// multiply-adds than the exact quotient, and no zero-extended addends.  For
// the transforms' butterflies, whose bounds allow for it.
__device__ __forceinline__ uint64_t mul_shoup_4q(uint64_t x, uint64_t w, uint64_t ws, uint64_t q) {
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32), s0 = (uint32_t)ws, s1 = (uint32_t)(ws >> 32);
  const uint64_t qh = (uint64_t)x1 * s1 + __umulhi(x1, s0) + __umulhi(x0, s1);
  return mul_sub_lo(x, w, qh, q);
}

__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b, uint64_t q) {
  uint64_t s = a + b;
  return s >= q ? s - q : s;
}
__device__ __forceinline__ uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t q) {
  return a >= b ? a - b : a + q - b;
}

// Barrett reduction of the 128-bit value hi*2^64 + lo, the same algorithm as
// SEAL's barrett_reduce_128: quotient estimate from the two-word ratio with
// the lowest partial product dropped, one conditional subtraction.
__device__ __forceinline__ uint64_t barrett128(uint64_t lo, uint64_t hi, const Mod &m) {
  const uint64_t p0h = mulhi64(lo, m.r_lo);
  const uint64_t p1l = lo * m.r_hi, p1h = mulhi64(lo, m.r_hi);
  const uint64_t p2l = hi * m.r_lo, p2h = mulhi64(hi, m.r_lo);
  uint64_t mid = p0h + p1l;
  uint64_t c = mid < p0h;
  mid += p2l;
  c += mid < p2l;
  const uint64_t qhat = hi * m.r_hi + p1h + p2h + c;
  const uint64_t r = lo - qhat * m.q;
  return r >= m.q ? r - m.q : r;
}

// Full product a*b as (lo, hi).
__device__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
  lo = a * b;
  hi = mulhi64(a, b);
}

__device__ __forceinline__ void add128(uint64_t &lo, uint64_t &hi, uint64_t blo, uint64_t bhi) {
  lo += blo;
  hi += bhi + (lo < blo);
}

// Unsigned 128-bit sum as four 32-bit words, least significant first.
struct U128 {
  uint32_t w[4];
};

// s[k] += a[k] * b[k] for k < 4, with a[k] < 2^63 and b[k] < 2^60.  With
// a = a1 2^32 + a0 and b = b1 2^32 + b0 the product is ll + m 2^32 + hh 2^64
// (ll = a0 b0, hh = a1 b1, m = a0 b1 + a1 b0 < 2^63 + 2^60): one carry chain
// adds ll + hh 2^64 to the words, a second adds m 2^32.  The four sums' chains
// are interleaved, each with its own SGPR pair for the carry, so a carry is
// read three instructions after the VALU write of it (two wait states are
// required).  Eleven instructions per product; the compiler's lowering of the
// same sums (64-bit adds, compares and selects for the carries) takes about 18.
__device__ __forceinline__ void mac128x4(U128 *s, const uint64_t *a, const uint64_t *b) {
  uint32_t p[4][6];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t a0 = (uint32_t)a[k], a1 = (uint32_t)(a[k] >> 32), b0 = (uint32_t)b[k], b1 = (uint32_t)(b[k] >> 32);
    const uint64_t ll = (uint64_t)a0 * b0, hh = (uint64_t)a1 * b1;
    const uint64_t m = (uint64_t)a0 * b1 + (uint64_t)a1 * b0;
    p[k][0] = (uint32_t)ll;
    p[k][1] = (uint32_t)(ll >> 32);
    p[k][2] = (uint32_t)hh;
    p[k][3] = (uint32_t)(hh >> 32);
    p[k][4] = (uint32_t)m;
    p[k][5] = (uint32_t)(m >> 32);
  }
  uint64_t c0, c1, c2, c3;
  asm("v_add_co_u32 %0, %16, %0, %20\n\t"
      "v_add_co_u32 %4, %17, %4, %26\n\t"
      "v_add_co_u32 %8, %18, %8, %32\n\t"
      "v_add_co_u32 %12, %19, %12, %38\n\t"
      "v_addc_co_u32 %1, %16, %1, %21, %16\n\t"
      "v_addc_co_u32 %5, %17, %5, %27, %17\n\t"
      "v_addc_co_u32 %9, %18, %9, %33, %18\n\t"
      "v_addc_co_u32 %13, %19, %13, %39, %19\n\t"
      "v_addc_co_u32 %2, %16, %2, %22, %16\n\t"
      "v_addc_co_u32 %6, %17, %6, %28, %17\n\t"
      "v_addc_co_u32 %10, %18, %10, %34, %18\n\t"
      "v_addc_co_u32 %14, %19, %14, %40, %19\n\t"
      "v_addc_co_u32 %3, %16, %3, %23, %16\n\t"
      "v_addc_co_u32 %7, %17, %7, %29, %17\n\t"
      "v_addc_co_u32 %11, %18, %11, %35, %18\n\t"
      "v_addc_co_u32 %15, %19, %15, %41, %19\n\t"
      "v_add_co_u32 %1, %16, %1, %24\n\t"
      "v_add_co_u32 %5, %17, %5, %30\n\t"
      "v_add_co_u32 %9, %18, %9, %36\n\t"
      "v_add_co_u32 %13, %19, %13, %42\n\t"
      "v_addc_co_u32 %2, %16, %2, %25, %16\n\t"
      "v_addc_co_u32 %6, %17, %6, %31, %17\n\t"
      "v_addc_co_u32 %10, %18, %10, %37, %18\n\t"
      "v_addc_co_u32 %14, %19, %14, %43, %19\n\t"
      "v_addc_co_u32 %3, %16, %3, 0, %16\n\t"
      "v_addc_co_u32 %7, %17, %7, 0, %17\n\t"
      "v_addc_co_u32 %11, %18, %11, 0, %18\n\t"
      "v_addc_co_u32 %15, %19, %15, 0, %19"
      : "+v"(s[0].w[0]), "+v"(s[0].w[1]), "+v"(s[0].w[2]), "+v"(s[0].w[3]), "+v"(s[1].w[0]), "+v"(s[1].w[1]),
        "+v"(s[1].w[2]), "+v"(s[1].w[3]), "+v"(s[2].w[0]), "+v"(s[2].w[1]), "+v"(s[2].w[2]), "+v"(s[2].w[3]),
        "+v"(s[3].w[0]), "+v"(s[3].w[1]), "+v"(s[3].w[2]), "+v"(s[3].w[3]), "=&s"(c0), "=&s"(c1), "=&s"(c2),
        "=&s"(c3)
      : "v"(p[0][0]), "v"(p[0][1]), "v"(p[0][2]), "v"(p[0][3]), "v"(p[0][4]), "v"(p[0][5]), "v"(p[1][0]),
        "v"(p[1][1]), "v"(p[1][2]), "v"(p[1][3]), "v"(p[1][4]), "v"(p[1][5]), "v"(p[2][0]), "v"(p[2][1]),
        "v"(p[2][2]), "v"(p[2][3]), "v"(p[2][4]), "v"(p[2][5]), "v"(p[3][0]), "v"(p[3][1]), "v"(p[3][2]),
        "v"(p[3][3]), "v"(p[3][4]), "v"(p[3][5]));
}

__device__ __forceinline__ void mac128x8(U128 (&s)[8], const uint64_t (&a)[8], const uint64_t (&b)[8]) {
  __builtin_amdgcn_sched_barrier(0);  // one group's products live at a time
  mac128x4(s, a, b);
  __builtin_amdgcn_sched_barrier(0);
  mac128x4(s + 4, a + 4, b + 4);
}

__device__ __forceinline__ uint64_t barrett128(const U128 &s, const Mod &m) {
  return barrett128(((uint64_t)s.w[1] << 32) | s.w[0], ((uint64_t)s.w[3] << 32) | s.w[2], m);
}

// Canonical residue of any 128-bit value w3:w2:w1:w0 modulo q = 2^60 - c with
// c < 2^20 (the same result as barrett128, since both are exact).  With
// 2^60 = c mod q and H = floor(V / 2^60) = h0 + 2^32 h1 + 2^64 h2:
//   V = L + c h0 + 2^32 c h1 + 2^64 c h2 (mod q),  L = V mod 2^60;
//   A = L + c h0 < 2^61;
//   B = c h1 + floor(A / 2^32) + 2^32 c h2 < 2^52 + 2^29 + 2^56 (h2 < 16);
//   T = 2^32 B + (A mod 2^32) = 2^60 T1 + T0 with T1 = floor(B / 2^28) < 2^29;
//   R = T0 + c T1 < 2^60 + 2^49 < 2q, and one conditional subtraction.
// About 15 instructions where the Barrett quotient takes about 50.
__device__ __forceinline__ uint64_t reduce128(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint64_t q) {
  const uint32_t c = (uint32_t)((1ULL << 60) - q);
  const uint32_t h0 = __builtin_amdgcn_alignbit(w2, w1, 28);
  const uint32_t h1 = __builtin_amdgcn_alignbit(w3, w2, 28);
  const uint32_t h2 = w3 >> 28;
  const uint64_t A = (uint64_t)h0 * c + (((uint64_t)(w1 & 0x0FFFFFFFu) << 32) | w0);
  const uint64_t B = (uint64_t)h1 * c + (A >> 32) + ((uint64_t)(h2 * c) << 32);
  const uint64_t T0 = ((B & 0x0FFFFFFFu) << 32) | (uint32_t)A;
  const uint32_t T1 = __builtin_amdgcn_alignbit((uint32_t)(B >> 32), (uint32_t)B, 28);
  const uint64_t R = (uint64_t)T1 * c + T0;
  return R >= q ? R - q : R;
}
__device__ __forceinline__ uint64_t reduce128(uint64_t lo, uint64_t hi, const Mod &m) {
  return reduce128((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32), m.q);
}
__device__ __forceinline__ uint64_t reduce128(const U128 &s, const Mod &m) {
  return reduce128(s.w[0], s.w[1], s.w[2], s.w[3], m.q);
}

__device__ __forceinline__ uint64_t mulmod_barrett(uint64_t a, uint64_t b, const Mod &m) {
  uint64_t lo, hi;
  mul128(a, b, lo, hi);
  return barrett128(lo, hi, m);
}

// Counter-based generator shared by the database ingest and the seeded client
// (SplitMix64 finaliser over seed, stream and counter).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xbf58476d1ce4e5b9ULL;
  z ^= z >> 27;
  z *= 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return z;
}
__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed * 0x9e3779b97f4a7c15ULL + mix64(stream ^ 0x632be59bd9b4e019ULL));
}
__host__ __device__ __forceinline__ uint64_t rand64_keyed(uint64_t key, uint64_t ctr) {
  return mix64(key + (ctr + 1) * 0x9e3779b97f4a7c15ULL);
}

}  // namespace onion
