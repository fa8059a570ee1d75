// k_crt.hpp -- the batched first-dimension product in a residue number
// system on the int8 matrix cores, exact.
//
// The product is the reference's mat_mat_128 (src/matrix.cpp:220-235): for
// every NTT slot, C[r][c] = sum_j A[r][j] * B[j][c] wrapping mod 2^128, then
// reduced mod q (delay_modulus, src/server.cpp:178-293).  With every operand
// below q < 2^60 and fst <= 1024, the exact sum S is below 1024 q^2 < 2^130.
//
// Residues.  The 17 moduli m_i below are pairwise coprime and at most 256;
// their product M (about 2^132.95) exceeds 2^130, so S is fixed by its
// residues mod the m_i (Chinese remainder theorem).  Every operand enters as
// its balanced residues (signed bytes in [-128, 127]); one
// v_mfma_i32_32x32x32_i8 per modulus and 32-wide step of j accumulates
// s_i = sum_j rA_i[j] rB_i[j] (|s_i| <= fst * 2^14 <= 2^24, exact in int32),
// and s_i == S (mod m_i).  That is 17 matrix products per 64x64-bit product,
// against 64 for signed base-256 digits.  From the s_i, with
// u_i = s_i + OFF_i (OFF_i a multiple of m_i: 0 <= u_i < 2^26) and
// w_i = (M / m_i)^-1 mod m_i:
//   sum_i u_i w_i M / m_i = S + V M,  V = floor(sum_i u_i w_i / m_i)
// (each term is S mod M times 1 mod m_i and 0 mod the other moduli), and
// since S / M < 0.13, V = floor(sum_i u_i F_i / 2^32 + 1/2) with F_i =
// round(2^32 w_i / m_i) (error below 0.14).  Then
//   S mod 2^128 = (sum_i u_i K_i - V M) mod 2^128,  K_i = w_i M / m_i mod 2^128,
// reduced mod q by the same Barrett reduction as every other path: the
// results are bit-identical.
//
// Layouts (chunk order as in k_first_dim_mfma.hpp, with 17 residues in place
// of 8 digits; a chunk is 16 bytes of one residue of 16 consecutive values of
// j of one row or column, lane l = 32 g + row or column within the block):
//   chunk (tile, s, i, l)  at  16 * (((tile * steps + s) * 17 + i) * 64 + l)
//   A  the database's residue copy (k_db_residues): tile = slot * n_rb + row
//      block, rows padded to a multiple of 32
//   B  the selection vectors (k_sel_residues): tile = slot * CB + column
//      block, columns padded to 32 CB
//   output out2[r][slot][c] (uint64, canonical mod q), as the digit kernels
#pragma once

#include "k_first_dim_mfma.hpp"

namespace onion {

constexpr int kRes = 17;
constexpr uint32_t kResMod[kRes] = {256, 255, 253, 251, 247, 241, 239, 233, 229,
                                    227, 223, 217, 211, 199, 197, 193, 191};
constexpr uint32_t kResStepBytes = kRes * 1024;  // one tile step's 17 chunks x 64 lanes

// Reconstruction constants, fixed by the moduli (computed at compile time).
struct CrtTab {
  uint32_t off[kRes];     // m_i ceil(2^24 / m_i)
  uint32_t f[kRes];       // round(2^32 w_i / m_i)
  uint32_t k[kRes][4];    // w_i M / m_i mod 2^128, 32-bit limbs, least significant first
  uint32_t neg_m[4];      // -M mod 2^128
};

__host__ __device__ constexpr CrtTab make_crt_tab() {
  using u128 = unsigned __int128;
  CrtTab T{};
  u128 mm[kRes] = {};
  for (int i = 0; i < kRes; ++i) {
    mm[i] = 1;
    for (int k = 0; k < kRes; ++k)
      if (k != i) mm[i] *= kResMod[k];  // below 2^126: no wrap
  }
  const u128 neg = (u128)0 - mm[0] * kResMod[0];  // -M mod 2^128
  for (int k = 0; k < 4; ++k) T.neg_m[k] = (uint32_t)(neg >> (32 * k));
  for (int i = 0; i < kRes; ++i) {
    const uint32_t m = kResMod[i];
    const uint32_t mmod = (uint32_t)(mm[i] % m);
    uint32_t w = 1;
    while ((uint64_t)w * mmod % m != 1) ++w;
    T.off[i] = m * (((1u << 24) + m - 1) / m);
    T.f[i] = (uint32_t)((((uint64_t)w << 32) + m / 2) / m);
    const u128 kk = mm[i] * w;  // mod 2^128
    for (int k = 0; k < 4; ++k) T.k[i][k] = (uint32_t)(kk >> (32 * k));
  }
  return T;
}

constexpr CrtTab kCrt = make_crt_tab();
static_assert(kCrt.off[0] == (1u << 24) && kCrt.f[0] == 0x1f000000u, "residue constants");

__host__ __device__ constexpr uint32_t res_p20(uint32_t m) { return (uint32_t)((1u << 20) % m); }
__host__ __device__ constexpr uint32_t res_p40(uint32_t m) {
  return (uint32_t)(((uint64_t)res_p20(m) * res_p20(m)) % m);
}
__host__ __device__ constexpr uint32_t res_magic(uint32_t m) {
  return (uint32_t)(((uint64_t(1) << 32) + m - 1) / m);
}

// Byte offset of chunk (tile, s, i, l).
__device__ __forceinline__ size_t res_chunk_offset(size_t tile, uint32_t steps, uint32_t s, uint32_t i,
                                                   uint32_t l) {
  return (((tile * steps + s) * kRes + i) * 64 + l) * 16;
}

// A value x < 2^60 split in 20-bit pieces for its residues.
struct Split20 {
  uint32_t a, b, c;  // x = a 2^40 + b 2^20 + c
};
__device__ __forceinline__ Split20 split20(uint64_t x) {
  return Split20{(uint32_t)(x >> 40), (uint32_t)(x >> 20) & 0xFFFFFu, (uint32_t)x & 0xFFFFFu};
}

// Balanced residue of x mod kResMod[I], as a byte (two's complement).
// t = a (2^40 mod m) + b (2^20 mod m) + c + floor(m / 2) < 2^29; the quotient
// from ceil(2^32 / m) is floor(t / m) or one more, fixed by one conditional
// add; the residue of t, minus floor(m / 2), lies in [-128, 127].
template <int I>
__device__ __forceinline__ uint32_t res_byte(const Split20 &x) {
  constexpr uint32_t m = kResMod[I];
  if constexpr (m == 256) {
    return x.c & 0xFFu;  // x mod 256 read as a signed byte is the balanced residue
  } else {
    const uint32_t t = x.a * res_p40(m) + x.b * res_p20(m) + x.c + (m >> 1);
    const uint32_t qq = __umulhi(t, res_magic(m));
    int32_t r = (int32_t)(t - qq * m);
    r += (r >> 31) & (int32_t)m;
    return (uint32_t)(r - (int32_t)(m >> 1)) & 0xFFu;
  }
}

// The chunk of residue I of 16 values (bytes in value order).
template <int I>
__device__ __forceinline__ uint4 res_chunk(const Split20 (&x)[16]) {
  uint32_t w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    w[k] = res_byte<I>(x[4 * k]) | (res_byte<I>(x[4 * k + 1]) << 8) | (res_byte<I>(x[4 * k + 2]) << 16) |
           (res_byte<I>(x[4 * k + 3]) << 24);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// Store the 17 residue chunks of 16 values at dst + 1024 i.
template <int I = 0>
__device__ __forceinline__ void store_res_chunks(uint8_t *dst, const Split20 (&x)[16]) {
  *reinterpret_cast<uint4 *>(dst + (size_t)I * 1024) = res_chunk<I>(x);
  if constexpr (I + 1 < kRes) store_res_chunks<I + 1>(dst, x);
}

// The database's residue copy: records rec_lo .. rec_hi - 1 of every slot, a
// record being the 16 words j = 16 jp .. 16 jp + 15 of one row (record =
// row * fst / 16 + jp; fst % 32 == 0).  Each becomes the 17 chunks (tile
// slot * n_rb + row / 32, step jp / 2, lane 32 (jp % 2) + row % 32).
// grid = (ceil((rec_hi - rec_lo) / 256), N).
__global__ __launch_bounds__(256) void k_db_residues(const uint64_t *__restrict__ db, uint8_t *__restrict__ dbr,
                                                     uint32_t fst, uint32_t other, uint64_t rec_lo,
                                                     uint64_t rec_hi) {
  const uint64_t r = rec_lo + (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= rec_hi) return;
  const uint32_t per_row = fst / 16, n_rb = (other + 31) / 32;
  const uint32_t row = (uint32_t)(r / per_row), jp = (uint32_t)(r % per_row);
  const size_t slot = blockIdx.y;
  const uint4 *src = reinterpret_cast<const uint4 *>(db + slot * fst * other + 16 * r);
  Split20 x[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint4 v = src[k];
    x[2 * k] = split20(((uint64_t)v.y << 32) | v.x);
    x[2 * k + 1] = split20(((uint64_t)v.w << 32) | v.z);
  }
  store_res_chunks(dbr + res_chunk_offset(slot * n_rb + row / 32, fst / 32, jp / 2, 0, 32 * (jp % 2) + row % 32), x);
}

// Selection vectors as B residue chunks: 8 slots x one 32-column block x one
// step per workgroup, as k_sel_chunks.  CANON: the leaves are NTT-form
// canonical residues (k_expand_ntt); otherwise k_sel_digits has transformed
// them in place into signed-digit form, which is undone here.
// grid = (N / 8, NCP / 32, fst / 32).
template <bool CANON>
__global__ __launch_bounds__(256) void k_sel_residues(const uint64_t *__restrict__ leaves, uint64_t leaf_stride,
                                                      uint8_t *__restrict__ Br, uint32_t fst, uint32_t ncols) {
  __shared__ uint64_t buf[8][32][33];  // [slot][column][j]
  const uint32_t L = blockIdx.x, sgi = (L >> 4) * 16 + (L & 7) * 2 + ((L >> 3) & 1);
  const uint32_t slot0 = 8 * sgi, cb = blockIdx.y, s = blockIdx.z, ncb = gridDim.y;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t piece = threadIdx.x + 256 * k, li = piece / 32, jj = piece % 32, col = 32 * cb + li;
    if (col < ncols) {  // padded columns stay unread; their outputs are never stored
      const uint4 *src = reinterpret_cast<const uint4 *>(leaves + (size_t)(col / 2) * leaf_stride +
                                                         (size_t)(2 * (32 * s + jj) + col % 2) * kN + slot0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint4 v = src[q];
        const uint64_t a = ((uint64_t)v.y << 32) | v.x, b = ((uint64_t)v.w << 32) | v.z;
        buf[2 * q][li][jj] = CANON ? a : (a ^ kDigitBias) - kDigitBias;
        buf[2 * q + 1][li][jj] = CANON ? b : (b ^ kDigitBias) - kDigitBias;
      }
    }
  }
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < 2; ++k) {
    const uint32_t pair = threadIdx.x + 256 * k, sg = pair / 64, l = pair % 64, li = l % 32, g = l / 32;
    Split20 x[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) x[t] = split20(buf[sg][li][16 * g + t]);
    store_res_chunks(Br + res_chunk_offset((size_t)(slot0 + sg) * ncb + cb, fst / 32, s, 0, l), x);
  }
}

// ---- the product ----------------------------------------------------------------------

// Operands of one step: per residue group k (residues v + 8k) RB A chunks
// (CB B chunks), then residue 16's A (B) chunk.
template <int CB>
struct CrtA {
  static constexpr int RB = 4 / CB, N = 2 * RB + 1;
  uint4 c[N];
};
template <int CB>
struct CrtB {
  static constexpr int N = 2 * CB + 1;
  uint4 c[N];
};

// S mod 2^128 of one output from its 17 residue sums (see the header), then
// mod q.
__device__ __forceinline__ uint64_t crt_rebuild(const int32_t (&sum)[kRes], const Mod &mq) {
  uint64_t L0 = 0, L1 = 0, L2 = 0, L3 = 0, F = 0;
#pragma unroll
  for (int i = 0; i < kRes; ++i) {
    const uint32_t u = (uint32_t)(sum[i] + (int32_t)kCrt.off[i]);
    L0 += (uint64_t)u * kCrt.k[i][0];
    L1 += (uint64_t)u * kCrt.k[i][1];
    L2 += (uint64_t)u * kCrt.k[i][2];
    L3 += (uint64_t)u * kCrt.k[i][3];
    F += (uint64_t)u * kCrt.f[i];
  }
  const uint32_t v = (uint32_t)((F + 0x80000000ull) >> 32);
  L0 += (uint64_t)v * kCrt.neg_m[0];
  L1 += (uint64_t)v * kCrt.neg_m[1] + (L0 >> 32);
  L2 += (uint64_t)v * kCrt.neg_m[2] + (L1 >> 32);
  L3 += (uint64_t)v * kCrt.neg_m[3] + (L2 >> 32);
  const uint64_t lo = (L0 & 0xFFFFFFFFull) | (L1 << 32), hi = (L2 & 0xFFFFFFFFull) | (L3 << 32);
  return barrett128(lo, hi, mq);
}

// A block's sums of residue `res` from accumulator tuple `a` into rec[res][o]
// (o = 32 row + column within the block).
__device__ __forceinline__ void crt_park(int32_t *rec, const v16i_t &a, uint32_t res, uint32_t g, uint32_t li) {
#pragma unroll
  for (int e = 0; e < 16; ++e) rec[res * 1024 + 32 * (8 * (e >> 2) + 4 * g + (e & 3)) + li] = a[e];
}

// Persistent workgroups of 8 waves (2 per SIMD; one workgroup per compute
// unit, G a multiple of 8 so each keeps its XCD) walk tiles of RB x CB blocks
// of 32 x 32 outputs (RB * CB = 4) of one slot, as k_first_dim_mfma_tiles.
// Wave v owns residues v and v + 8 on all four blocks, and waves 0..3 also
// residue 16 on block v: the two waves of a SIMD (v and v + 4) issue 17
// matrix products per step together, and a wave holds at most 9 accumulator
// tuples.  Its operands are its own (only residue 16's A or B chunk is also
// read by another wave).  The database's chunks (HBM) are loaded D steps
// ahead into a ring of D register buffers, the selection's (mostly from L2)
// one step ahead; each chunk is reloaded as soon as its last product of the
// step has issued.  The loads form one stream over all of a workgroup's
// tiles; steps per tile must be a multiple of D.
// At a tile's end, block by block, the waves park their residue sums in LDS
// (int32, residue-major) and each lane rebuilds two outputs from them (S mod
// 2^128, Barrett) and stores them, while the next tile's first loads are in
// flight.
template <int CB, int D>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_first_dim_crt(
    const uint8_t *__restrict__ dbr, const uint8_t *__restrict__ Br, uint64_t *__restrict__ out2, uint32_t fst,
    uint32_t other, uint32_t ncols, uint32_t n_rt, Mod mq) {
  constexpr int RB = 4 / CB;
  __shared__ int32_t rec[kRes * 1024];
  const uint32_t v = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t t = threadIdx.x, l = t & 63, g = l >> 5, li = l & 31;
  const bool has16 = v < 4;  // residue 16 on block v
  const uint32_t vrb = (v & 3) / CB, vcb = (v & 3) % CB;
  const uint32_t n_rb = (other + 31) / 32, steps = fst / 32, total = kN * n_rt, G = gridDim.x;
  if (blockIdx.x >= total) return;
  // tile L: slot (L / 8 / n_rt) * 8 + L % 8, row tile (L / 8) % n_rt
  auto slot_of = [&](uint32_t L) { return ((L >> 3) / n_rt) * 8 + (L & 7); };
  auto rt_of = [&](uint32_t L) { return (L >> 3) % n_rt; };

  // two loaders over the same (tile, step) sequence, A ahead of B; past the
  // last tile each repeats its last step
  struct Pos {
    uint32_t lt, ls;
  };
  Pos pa{blockIdx.x, 0}, pb{blockIdx.x, 0};
  const uint8_t *la[RB], *la16, *lb[CB], *lb16;
  auto set_a = [&]() {
    const uint32_t slot = slot_of(pa.lt), rt = rt_of(pa.lt);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
      la[rb] = dbr + res_chunk_offset((size_t)slot * n_rb + min(rt * RB + rb, n_rb - 1), steps, 0, 0, 0);
    la16 = dbr + res_chunk_offset((size_t)slot * n_rb + min(rt * RB + vrb, n_rb - 1), steps, 0, 0, 0);
  };
  auto set_b = [&]() {
    const uint32_t slot = slot_of(pb.lt);
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) lb[cb] = Br + res_chunk_offset((size_t)slot * CB + cb, steps, 0, 0, 0);
    lb16 = Br + res_chunk_offset((size_t)slot * CB + vcb, steps, 0, 0, 0);
  };
  auto advance = [&](Pos &p, bool a_side) {
    if (++p.ls == steps) {
      if (p.lt + G < total) {
        p.ls = 0;
        p.lt += G;
        if (a_side)
          set_a();
        else
          set_b();
      } else {
        p.ls = steps - 1;
      }
    }
  };
  auto load_a = [&](CrtA<CB> &r, int k) {
    const size_t so = (size_t)pa.ls * kResStepBytes + l * 16;
    if (k < 2) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
        r.c[k * RB + rb] = *reinterpret_cast<const uint4 *>(la[rb] + so + (size_t)(v + 8 * k) * 1024);
    } else if (has16) {
      r.c[2 * RB] = *reinterpret_cast<const uint4 *>(la16 + so + 16 * 1024);
    }
  };
  auto load_b = [&](CrtB<CB> &r, int k) {
    const size_t so = (size_t)pb.ls * kResStepBytes + l * 16;
    if (k < 2) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
        r.c[k * CB + cb] = *reinterpret_cast<const uint4 *>(lb[cb] + so + (size_t)(v + 8 * k) * 1024);
    } else if (has16) {
      r.c[2 * CB] = *reinterpret_cast<const uint4 *>(lb16 + so + 16 * 1024);
    }
  };
  set_a();
  set_b();
  CrtA<CB> A[D];
  CrtB<CB> B;
#pragma unroll
  for (int d = 0; d < D; ++d) {
#pragma unroll
    for (int k = 0; k < 3; ++k) load_a(A[d], k);
    advance(pa, true);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) load_b(B, k);
  advance(pb, false);

  v16i_t acc[2][4], acc16;
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[k][b] = (v16i_t){0};
  acc16 = (v16i_t){0};

  for (uint32_t ct = blockIdx.x; ct < total; ct += G) {
    for (uint32_t s = 0; s < steps; s += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
#pragma unroll
          for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
              v16i_t &a = acc[k][rb * CB + cb];
              a = __builtin_amdgcn_mfma_i32_32x32x32_i8(as_operand(A[d].c[k * RB + rb]), as_operand(B.c[k * CB + cb]),
                                                        a, 0, 0, 0);
            }
          load_a(A[d], k);
          load_b(B, k);
        }
        if (has16)
          acc16 = __builtin_amdgcn_mfma_i32_32x32x32_i8(as_operand(A[d].c[2 * RB]), as_operand(B.c[2 * CB]), acc16,
                                                        0, 0, 0);
        load_a(A[d], 2);
        load_b(B, 2);
        advance(pa, true);
        advance(pb, false);
      }
    }
    // one block at a time: park its residue sums, then rebuild its outputs
    // (two per lane: o = t and t + 512, row o / 32, column o % 32)
    const uint32_t slot = slot_of(ct), row0 = rt_of(ct) * (32 * RB);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      __syncthreads();  // the previous block's sums are read
      crt_park(rec, acc[0][b], v, g, li);
      crt_park(rec, acc[1][b], v + 8, g, li);
      if (v == (uint32_t)b) crt_park(rec, acc16, 16, g, li);
      __syncthreads();  // the block's sums are parked
      const uint32_t rb = b / CB, cb = b % CB;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t o = t + 512 * h;
        int32_t sum[kRes];
#pragma unroll
        for (int i = 0; i < kRes; ++i) sum[i] = rec[i * 1024 + o];
        const uint64_t val = crt_rebuild(sum, mq);
        const uint32_t row = row0 + rb * 32 + (o >> 5), col = cb * 32 + (o & 31);
        if (row < other && col < ncols) out2[((size_t)row * kN + slot) * ncols + col] = val;
      }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[k][b] = (v16i_t){0};
    acc16 = (v16i_t){0};
  }
}

}  // namespace onion
