// k_first_dim_mfma.hpp -- the batched first-dimension product on the int8
// matrix cores, exact.
//
// The product is the reference's mat_mat_128 (src/matrix.cpp:220-235): for
// every NTT slot l, C_l[r][c] = sum_j A_l[r][j] * B_l[j][c] as an integer
// wrapping mod 2^128, then reduced mod q (delay_modulus, src/server.cpp:
// 178-293).  Every operand is < q < 2^60.
//
// Signed base-256 digits.  A value a < 2^60 is written as
//   a = sum_{i<8} d_i 256^i,  d_i in [-128, 127],
// whose two's-complement bytes are those of V = (a + K) ^ K with
// K = 0x8080808080808080 (a + K cannot overflow, and byte i of a + K is
// d_i + 128).  Then
//   C = sum_{m<15} 256^m S_m,   S_m = sum_{i+k=m} sum_j dA_i[j] dB_k[j],
// and each S_m is a sum of at most 8 * fst products of magnitude <= 2^14, so
// it is exact in int32 for fst <= 2^14.  One v_mfma_i32_32x32x32_i8 adds a
// 32x32 tile of dA_i * dB_k over 32 values of j; a wave issues the 64 (i, k)
// digit products of each 32-wide step of j into 15 diagonal accumulators.
// The epilogue forms C mod 2^128 from the 15 signed sums and applies the same
// Barrett reduction as the 64-bit path, so results are bit-identical.
//
// Operand layout.  MFMA lane l holds 16 bytes of row (l % 32) of the A tile
// and 16 bytes of column (l % 32) of the B tile, both in the same k order;
// since the sum runs over all 32 k, any assignment of j to (lane group, byte)
// shared by A and B gives the same result.  Here byte b of lane group
// g = l / 32 is j = 16 g + b of the 32-wide step, so an operand holds one
// digit of 16 consecutive values of j.
//   B planes  Bp[l][j/16][col][128 B], col = 2*query + poly, padded to NCP
//             columns (NCP = 32, 64 or 128; written by k_sel_planes).  Byte
//             16 k + b of a record is digit k of value j = 16 (j/16) + b, so
//             the 16 bytes at 16 k are the operand for digit k as they stand.
//   A         each lane reads the 16 consecutive database words of its row
//             and turns them into digit planes in registers
//   output    out2[l][r][c] (uint64, canonical mod q, c < 2*nq), turned into
//             coefficient-form ciphertexts by k_ntt_inv_gather.
#pragma once

#include "k_first_dim.hpp"

namespace onion {

typedef int v4i_t __attribute__((ext_vector_type(4)));
typedef int v16i_t __attribute__((ext_vector_type(16)));

constexpr uint64_t kDigitBias = 0x8080808080808080ULL;

__device__ __forceinline__ uint64_t signed_digits(uint64_t a) { return (a + kDigitBias) ^ kDigitBias; }

// 4x4 byte transpose: y_i byte t = x_t byte i.
__device__ __forceinline__ void transpose4(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3, uint32_t &y0,
                                           uint32_t &y1, uint32_t &y2, uint32_t &y3) {
  // v_perm_b32(S0, S1, sel): selector 0-3 picks bytes of S1, 4-7 bytes of S0
  const uint32_t p0 = __builtin_amdgcn_perm(x1, x0, 0x06020400u);  // x0.0 x1.0 x0.2 x1.2
  const uint32_t p1 = __builtin_amdgcn_perm(x1, x0, 0x07030501u);  // x0.1 x1.1 x0.3 x1.3
  const uint32_t p2 = __builtin_amdgcn_perm(x3, x2, 0x06020400u);
  const uint32_t p3 = __builtin_amdgcn_perm(x3, x2, 0x07030501u);
  y0 = __builtin_amdgcn_perm(p2, p0, 0x05040100u);
  y2 = __builtin_amdgcn_perm(p2, p0, 0x07060302u);
  y1 = __builtin_amdgcn_perm(p3, p1, 0x05040100u);
  y3 = __builtin_amdgcn_perm(p3, p1, 0x07060302u);
}

// Eight values (already in signed-digit form) -> one 64-byte plane block:
// out[2i + h] = digit i of values 4h .. 4h+3.
__device__ __forceinline__ void planes8(const uint64_t (&w)[8], uint32_t (&out)[16]) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint64_t *v = w + 4 * h;
    transpose4((uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2], (uint32_t)v[3], out[h], out[2 + h], out[4 + h],
               out[6 + h]);
    transpose4((uint32_t)(v[0] >> 32), (uint32_t)(v[1] >> 32), (uint32_t)(v[2] >> 32), (uint32_t)(v[3] >> 32),
               out[8 + h], out[10 + h], out[12 + h], out[14 + h]);
  }
}

// Selection vectors as digit planes: block (j-block jb of 8 leaves, poly c)
// of query b NTTs its 8 polynomials and writes, for each slot, its half
// (bytes 16 k + 8 (jb % 2) .. +7 for digit k) of the 128-byte record
// Bp[slot][jb / 2][2b + c].  grid = (fst/8 * 2, nq); fst % 16 == 0.
__global__ __launch_bounds__(kNttThreads) void k_sel_planes(const uint64_t *leaves, uint64_t leaf_stride,
                                                            uint8_t *Bp, uint32_t fst, uint32_t ncp, NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  const uint32_t jb = blockIdx.x / 2, c = blockIdx.x % 2, b = blockIdx.y;
  uint64_t xs[8][8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    uint64_t x[8];
    load_natural(x, leaves + (size_t)b * leaf_stride + (size_t)(jb * 8 + j) * 2 * kN + (size_t)c * kN);
    ntt_fwd(x, lds, T);
#pragma unroll
    for (int r = 0; r < 8; ++r) xs[j][r] = signed_digits(ntt_fwd_final(x[r], T.q));
  }
  const uint32_t col = 2 * b + c, nr = fst / 16;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const uint32_t slot = 8 * threadIdx.x + r;
    uint64_t w[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) w[j] = xs[j][r];
    uint32_t o[16];
    planes8(w, o);
    uint2 *dst = reinterpret_cast<uint2 *>(Bp + (((size_t)slot * nr + jb / 2) * ncp + col) * 128 + 8 * (jb % 2));
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[2 * k] = make_uint2(o[2 * k], o[2 * k + 1]);
  }
}

// A operand of digit i: plane i of values j = 16g .. 16g+7 (p0), then of
// 16g+8 .. 16g+15 (p1).
__device__ __forceinline__ v4i_t operand(const uint32_t (&p0)[16], const uint32_t (&p1)[16], int i) {
  v4i_t v;
  v[0] = (int)p0[2 * i];
  v[1] = (int)p0[2 * i + 1];
  v[2] = (int)p1[2 * i];
  v[3] = (int)p1[2 * i + 1];
  return v;
}

// B operand of digit k: the 16 bytes at 16 k of the lane's record.
__device__ __forceinline__ v4i_t b_operand(const uint4 &v) {
  v4i_t r;
  r[0] = (int)v.x;
  r[1] = (int)v.y;
  r[2] = (int)v.z;
  r[3] = (int)v.w;
  return r;
}

// Raw operands of one 32-wide step of j for this lane: the 16 database words
// j0+16g .. +15 of its row, the plane record j/16 = 2s+g of its column.  Rows
// past the matrix read row 0 instead (arow points there); their outputs are
// not stored, and no other output depends on them.
struct StepRaw {
  uint4 a[8];
  uint4 b[8];
};

__device__ __forceinline__ const uint4 *a_step(const uint64_t *arow, uint32_t s, uint32_t g) {
  return reinterpret_cast<const uint4 *>(arow + 32 * s + 16 * g);
}
__device__ __forceinline__ const uint4 *b_step(const uint8_t *bcol, uint32_t s, uint32_t g, uint32_t ncp) {
  return reinterpret_cast<const uint4 *>(bcol + (size_t)(2 * s + g) * ncp * 128);
}

__device__ __forceinline__ void load_step(StepRaw &R, const uint64_t *arow, const uint8_t *bcol, uint32_t s,
                                          uint32_t g, uint32_t ncp) {
  const uint4 *a = a_step(arow, s, g), *b = b_step(bcol, s, g, ncp);
#pragma unroll
  for (int q = 0; q < 8; ++q) R.a[q] = a[q];
#pragma unroll
  for (int q = 0; q < 8; ++q) R.b[q] = b[q];
}

__device__ __forceinline__ void a_planes(const uint4 *a, uint32_t (&out)[16]) {
  uint64_t w[8];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    w[2 * q] = signed_digits(((uint64_t)a[q].y << 32) | a[q].x);
    w[2 * q + 1] = signed_digits(((uint64_t)a[q].w << 32) | a[q].z);
  }
  planes8(w, out);
}

// C mod 2^128 from the 15 diagonal sums of one output, then mod q.
__device__ __forceinline__ uint64_t diagonals_mod_q(const v16i_t (&acc)[15], int e, const Mod &mq) {
  int64_t P0 = 0, P1 = 0, P2 = 0;  // weights 2^0, 2^40, 2^80; |P| < 2^61
#pragma unroll
  for (int d = 0; d < 5; ++d) {
    P0 += (int64_t)acc[d][e] * ((int64_t)1 << (8 * d));
    P1 += (int64_t)acc[d + 5][e] * ((int64_t)1 << (8 * d));
    P2 += (int64_t)acc[d + 10][e] * ((int64_t)1 << (8 * d));
  }
  uint64_t lo = (uint64_t)P0, hi = (uint64_t)(P0 >> 63);
  add128(lo, hi, (uint64_t)P1 << 40, (uint64_t)(P1 >> 24));
  hi += (uint64_t)P2 << 16;
  return barrett128(lo, hi, mq);
}

// First dimension for 2..64 queries.  CB = column blocks of 32 (NCP = 32 CB
// padded columns); the workgroup's four waves cover RB = 4 / CB row blocks of
// 32 rows x CB column blocks.  grid.x = N * n_rt with n_rt row tiles of
// 32 RB rows; all row tiles of a slot run on one XCD (block id mod 8), so the
// slot's plane blocks are read from that XCD's L2.  Requires fst % 32 == 0.
template <int CB>
__global__ __launch_bounds__(256) void k_first_dim_mfma(const uint64_t *__restrict__ db,
                                                        const uint8_t *__restrict__ Bp,
                                                        uint64_t *__restrict__ out2, uint32_t fst,
                                                        uint32_t other, uint64_t num_pt, uint32_t ncols,
                                                        uint32_t n_rt, Mod mq) {
  constexpr int RB = 4 / CB, NCP = 32 * CB;
  const uint32_t L = blockIdx.x, xcd = L & 7, seq = L >> 3;
  const uint32_t slot = (seq / n_rt) * 8 + xcd, rt = seq % n_rt;
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 5, li = l & 31;
  const uint32_t rb = w / CB, cb = w % CB;
  const uint32_t row0 = rt * (32 * RB) + rb * 32;
  const uint32_t row = row0 + li, col = cb * 32 + li;
  const uint64_t *arow = db + (size_t)slot * num_pt + (size_t)(row < other ? row : 0) * fst;
  const uint8_t *bcol = Bp + (size_t)slot * (fst / 16) * NCP * 128 + (size_t)col * 128;

  v16i_t acc[15];
#pragma unroll
  for (int d = 0; d < 15; ++d) acc[d] = (v16i_t){0};

  // one raw buffer: the step's words become planes, then the buffer takes the
  // next step's loads while the 64 products run
  const uint32_t steps = fst / 32;
  StepRaw raw;
  load_step(raw, arow, bcol, 0, g, NCP);
  for (uint32_t s = 0; s < steps; ++s) {
    uint32_t pa0[16], pa1[16];
    a_planes(raw.a, pa0);
    a_planes(raw.a + 4, pa1);
    v4i_t B[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) B[k] = b_operand(raw.b[k]);
    if (s + 1 < steps) load_step(raw, arow, bcol, s + 1, g, NCP);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const v4i_t A = operand(pa0, pa1, i);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        acc[i + k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B[k], acc[i + k], 0, 0, 0);
    }
  }

  // output (m, n) of register e: n = lane % 32, m = 8 (e / 4) + 4 g + e % 4
  if (col < ncols) {
    uint64_t *dst = out2 + (size_t)slot * other * ncols + col;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const uint32_t m = row0 + 8 * (e / 4) + 4 * g + (e % 4);
      const uint64_t v = diagonals_mod_q(acc, e, mq);
      if (m < other) dst[(size_t)m * ncols] = v;
    }
  }
}

// ---- pre-converted database --------------------------------------------------------------
//
// The database as digit planes: dbp has the byte size and indexing of db, and
// the 16 words j = 16 jp .. 16 jp + 15 of (slot, row) hold instead the record
// of that pair of j-blocks in the B-plane format (byte 16 k + b = digit k of
// value 16 jp + b).  The database is static between writes, so converting it
// once per write takes the whole A conversion out of the first dimension: a
// lane's 8 A operands of a step are its 128-byte record as loaded.

// Records rec_lo .. rec_hi - 1 of every slot (a record is 16 words of one
// row; num_pt % 16 == 0).  grid = (ceil((rec_hi - rec_lo) / 256), N).
__global__ __launch_bounds__(256) void k_db_planes(const uint64_t *__restrict__ db, uint8_t *__restrict__ dbp,
                                                   uint64_t num_pt, uint64_t rec_lo, uint64_t rec_hi) {
  const uint64_t r = rec_lo + (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= rec_hi) return;
  const size_t w0 = (size_t)blockIdx.y * num_pt + 16 * r;
  const uint4 *src = reinterpret_cast<const uint4 *>(db + w0);
  uint4 a[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) a[q] = src[q];
  uint32_t p0[16], p1[16];
  a_planes(a, p0);
  a_planes(a + 4, p1);
  uint4 *dst = reinterpret_cast<uint4 *>(dbp + 8 * w0);
#pragma unroll
  for (int k = 0; k < 8; ++k) dst[k] = make_uint4(p0[2 * k], p0[2 * k + 1], p1[2 * k], p1[2 * k + 1]);
}

// The products of one step from a lane's A and B records.
__device__ __forceinline__ void step_products(v16i_t (&acc)[15], const StepRaw &R) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const v4i_t A = b_operand(R.a[i]);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      acc[i + k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, b_operand(R.b[k]), acc[i + k], 0, 0, 0);
  }
}

__device__ __forceinline__ void load_records(StepRaw &R, const uint8_t *arec, const uint8_t *brec, uint32_t s,
                                             uint32_t ncp) {
  const uint4 *a = reinterpret_cast<const uint4 *>(arec + (size_t)s * 256);
  const uint4 *b = reinterpret_cast<const uint4 *>(brec + (size_t)s * 2 * ncp * 128);
#pragma unroll
  for (int q = 0; q < 8; ++q) R.a[q] = a[q];
#pragma unroll
  for (int q = 0; q < 8; ++q) R.b[q] = b[q];
}

// First dimension for 2..64 queries from the pre-converted database: the tiles
// of k_first_dim_mfma, with three register buffers so each step's records are
// loaded two steps before their products.
template <int CB>
__global__ __launch_bounds__(256) void k_first_dim_mfma_pre(const uint8_t *__restrict__ dbp,
                                                            const uint8_t *__restrict__ Bp,
                                                            uint64_t *__restrict__ out2, uint32_t fst,
                                                            uint32_t other, uint64_t num_pt, uint32_t ncols,
                                                            uint32_t n_rt, Mod mq) {
  constexpr int RB = 4 / CB, NCP = 32 * CB;
  const uint32_t L = blockIdx.x, xcd = L & 7, seq = L >> 3;
  const uint32_t slot = (seq / n_rt) * 8 + xcd, rt = seq % n_rt;
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 5, li = l & 31;
  const uint32_t rb = w / CB, cb = w % CB;
  const uint32_t row0 = rt * (32 * RB) + rb * 32;
  const uint32_t row = row0 + li, col = cb * 32 + li;
  // record 2s + g of the row (rows past the matrix read row 0; not stored)
  const uint8_t *arec = dbp + ((size_t)slot * num_pt + (size_t)(row < other ? row : 0) * fst) * 8 + (size_t)g * 128;
  const uint8_t *brec = Bp + ((size_t)slot * (fst / 16) + g) * NCP * 128 + (size_t)col * 128;

  v16i_t acc[15];
#pragma unroll
  for (int d = 0; d < 15; ++d) acc[d] = (v16i_t){0};

  const uint32_t steps = fst / 32, body = steps - steps % 3;
  StepRaw R[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) load_records(R[d], arec, brec, min((uint32_t)d, steps - 1), NCP);
  // groups of three steps; a buffer is refilled (with step s + d + 3, or
  // again the last step once past the end) as soon as its products issue
  for (uint32_t s = 0; s < body; s += 3) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      step_products(acc, R[d]);
      load_records(R[d], arec, brec, min(s + d + 3, steps - 1), NCP);
    }
  }
#pragma unroll
  for (int d = 0; d < 2; ++d)
    if (body + d < steps) step_products(acc, R[d]);

  if (col < ncols) {
    uint64_t *dst = out2 + (size_t)slot * other * ncols + col;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const uint32_t m = row0 + 8 * (e / 4) + 4 * g + (e % 4);
      const uint64_t v = diagonals_mod_q(acc, e, mq);
      if (m < other) dst[(size_t)m * ncols] = v;
    }
  }
}

// ---- LDS-pipelined variant ------------------------------------------------------------
//
// Same tiles and arithmetic as k_first_dim_mfma, but the operands travel
// global -> LDS by direct-to-LDS loads (global_load_lds_dwordx4), three steps
// deep, so the matrix cores are not stalled by HBM latency and no VGPRs hold
// loads in flight.  A step's operands are 16-byte "pieces": piece p of an A
// row block holds, for reading lane l, the two database words
// j = 32 s + 16 g + 2 p, +1 of row (l % 32), and piece p of a B column block
// holds bytes 16 p .. 16 p + 15 (the operand for digit p) of plane record
// 2 s + g of column (l % 32); g = l / 32.  A piece is 1 KiB in LDS with lane
// l's 16 bytes at 16 l, so both the DMA write and the reading ds_read_b128
// are contiguous.
// The RB row blocks and CB column blocks of a step (8 (RB + CB) pieces) are
// loaded by the four waves together and shared through LDS.

constexpr int kMfmaStages = 3;

__device__ __forceinline__ void lds_dma16(const void *g, void *lds) {
  __builtin_amdgcn_global_load_lds(g, (void __attribute__((address_space(3))) *)lds, 16, 0, 0);
}

// s_waitcnt vmcnt(n), other counters untouched (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt(0x0F70 | (N & 15) | ((N >> 4) << 14));
}

template <int CB>
struct MfmaPipe {
  static constexpr int RB = 4 / CB;
  static constexpr int PIECES = 8 * (RB + CB);  // per step
  static constexpr int PER_WAVE = PIECES / 4;   // loads per wave per step
  static constexpr int STAGE_BYTES = PIECES * 1024;
};

// Issue this wave's share of the loads of step s into LDS stage `stage`.
template <int CB>
__device__ __forceinline__ void issue_step(uint8_t *lds, int stage, uint32_t s, const uint64_t *a_base,
                                           uint32_t fst, uint32_t row_base, uint32_t other, const uint8_t *b_base,
                                           uint32_t w, uint32_t l) {
  using P = MfmaPipe<CB>;
  const uint32_t g = l >> 5, li = l & 31;
#pragma unroll
  for (int i = 0; i < P::PER_WAVE; ++i) {
    const uint32_t piece = w + 4 * i;  // wave-uniform
    uint8_t *dst = lds + stage * P::STAGE_BYTES + piece * 1024;
    const uint32_t p = piece & 7;
    const uint32_t blk = piece >> 3;
    if (blk < (uint32_t)P::RB) {
      uint32_t row = row_base + blk * 32 + li;
      row = row < other ? row : 0;  // masked rows read row 0; their outputs are not stored
      lds_dma16(a_base + (size_t)row * fst + 32 * s + 16 * g + 2 * p, dst);
    } else {
      const uint32_t cb = blk - P::RB, col = cb * 32 + li;
      lds_dma16(b_base + ((size_t)(2 * s + g) * (32 * CB) + col) * 128 + 16 * p, dst);
    }
  }
}

typedef unsigned v4u_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 to_uint4(v4u_t v) { return make_uint4(v[0], v[1], v[2], v[3]); }

// The 16 pieces (8 of A at byte address a, 8 of B at b, 1 KiB apart) of one
// step from LDS.  Written as one asm block that waits for its own reads: the
// compiler, which cannot tell these reads from the direct-to-LDS loads still
// in flight for later steps, would otherwise wait for all of them.
__device__ __forceinline__ void lds_read_step(uint4 (&ra)[8], uint4 (&rb)[8], uint32_t a, uint32_t b) {
  v4u_t x0, x1, x2, x3, x4, x5, x6, x7, y0, y1, y2, y3, y4, y5, y6, y7;
  asm volatile(
      "ds_read_b128 %0, %16\n"
      "ds_read_b128 %1, %16 offset:1024\n"
      "ds_read_b128 %2, %16 offset:2048\n"
      "ds_read_b128 %3, %16 offset:3072\n"
      "ds_read_b128 %4, %16 offset:4096\n"
      "ds_read_b128 %5, %16 offset:5120\n"
      "ds_read_b128 %6, %16 offset:6144\n"
      "ds_read_b128 %7, %16 offset:7168\n"
      "ds_read_b128 %8, %17\n"
      "ds_read_b128 %9, %17 offset:1024\n"
      "ds_read_b128 %10, %17 offset:2048\n"
      "ds_read_b128 %11, %17 offset:3072\n"
      "ds_read_b128 %12, %17 offset:4096\n"
      "ds_read_b128 %13, %17 offset:5120\n"
      "ds_read_b128 %14, %17 offset:6144\n"
      "ds_read_b128 %15, %17 offset:7168\n"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3), "=&v"(x4), "=&v"(x5), "=&v"(x6), "=&v"(x7), "=&v"(y0),
        "=&v"(y1), "=&v"(y2), "=&v"(y3), "=&v"(y4), "=&v"(y5), "=&v"(y6), "=&v"(y7)
      : "v"(a), "v"(b)
      : "memory");
  ra[0] = to_uint4(x0); ra[1] = to_uint4(x1); ra[2] = to_uint4(x2); ra[3] = to_uint4(x3);
  ra[4] = to_uint4(x4); ra[5] = to_uint4(x5); ra[6] = to_uint4(x6); ra[7] = to_uint4(x7);
  rb[0] = to_uint4(y0); rb[1] = to_uint4(y1); rb[2] = to_uint4(y2); rb[3] = to_uint4(y3);
  rb[4] = to_uint4(y4); rb[5] = to_uint4(y5); rb[6] = to_uint4(y6); rb[7] = to_uint4(y7);
}

// PRE: db is the digit-plane copy (k_db_planes; same indexing), so a piece of
// A is already the operand of one digit.
template <int CB, bool PRE>
__global__ __launch_bounds__(256) void k_first_dim_mfma_lds(const uint64_t *__restrict__ db,
                                                            const uint8_t *__restrict__ Bp,
                                                            uint64_t *__restrict__ out2, uint32_t fst,
                                                            uint32_t other, uint64_t num_pt, uint32_t ncols,
                                                            uint32_t n_rt, Mod mq) {
  using P = MfmaPipe<CB>;
  constexpr int RB = P::RB, NCP = 32 * CB;
  __shared__ __attribute__((aligned(16))) uint8_t lds[kMfmaStages * P::STAGE_BYTES];
  const uint32_t L = blockIdx.x, xcd = L & 7, seq = L >> 3;
  const uint32_t slot = (seq / n_rt) * 8 + xcd, rt = seq % n_rt;
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 5, li = l & 31;
  const uint32_t rb = w / CB, cb = w % CB;
  const uint32_t row_base = rt * (32 * RB);
  const uint64_t *a_base = db + (size_t)slot * num_pt;
  const uint8_t *b_base = Bp + (size_t)slot * (fst / 16) * NCP * 128;

  v16i_t acc[15];
#pragma unroll
  for (int d = 0; d < 15; ++d) acc[d] = (v16i_t){0};

  // LDS byte offset of the staging area (generic -> local address space)
  const uint32_t lds_off =
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) uint8_t *)lds + l * 16;
  const uint32_t steps = fst / 32;
  issue_step<CB>(lds, 0, 0, a_base, fst, row_base, other, b_base, w, l);
  if (steps > 1) issue_step<CB>(lds, 1, 1, a_base, fst, row_base, other, b_base, w, l);
  for (uint32_t s = 0; s < steps; ++s) {
    // this wave's loads of step s are done once at most one later step is in flight
    if (s + 1 < steps)
      wait_vmcnt<P::PER_WAVE>();
    else
      wait_vmcnt<0>();
    // every wave's pieces of step s landed, and every wave has read step s - 1
    // (its reads wait for completion), so stage (s + 2) % 3 is free.  A plain
    // s_barrier: the loads of later steps stay in flight across it.
    asm volatile("s_barrier" ::: "memory");
    if (s + 2 < steps) issue_step<CB>(lds, (s + 2) % kMfmaStages, s + 2, a_base, fst, row_base, other, b_base, w, l);
    const uint32_t st = lds_off + (s % kMfmaStages) * P::STAGE_BYTES;
    uint4 ra[8], rbv[8];
    lds_read_step(ra, rbv, st + rb * 8 * 1024, st + (RB + cb) * 8 * 1024);
    uint32_t pa0[16], pa1[16];
    if (!PRE) {
      a_planes(ra, pa0);
      a_planes(ra + 4, pa1);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const v4i_t A = PRE ? b_operand(ra[i]) : operand(pa0, pa1, i);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        acc[i + k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, b_operand(rbv[k]), acc[i + k], 0, 0, 0);
    }
  }

  const uint32_t row0 = row_base + rb * 32, col = cb * 32 + li;
  if (col < ncols) {
    uint64_t *dst = out2 + (size_t)slot * other * ncols + col;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const uint32_t m = row0 + 8 * (e / 4) + 4 * g + (e % 4);
      const uint64_t v = diagonals_mod_q(acc, e, mq);
      if (m < other) dst[(size_t)m * ncols] = v;
    }
  }
}

// Inverse NTT of the first-dimension outputs gathered from out2[slot][r][c]
// into fd (query b = c / 2, poly c % 2 of ciphertext r, coefficient form).
// One workgroup per (r, c); when other % 8 == 0 the rows are dealt to XCDs
// (r = xcd mod 8) with c fastest, so the workgroups that share a cache line
// of out2 run on one XCD at about the same time.
__global__ __launch_bounds__(kNttThreads) void k_ntt_inv_gather(const uint64_t *__restrict__ out2,
                                                                uint32_t other, uint32_t ncols, uint64_t *fd,
                                                                uint64_t fd_stride, NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  uint32_t r, c;
  if (other % 8 == 0) {
    const uint32_t xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    r = 8 * (k / ncols) + xcd;
    c = k % ncols;
  } else {
    r = blockIdx.x / ncols;
    c = blockIdx.x % ncols;
  }
  const size_t sstride = (size_t)other * ncols;
  const uint64_t *src = out2 + (size_t)r * ncols + c;
  uint64_t x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = src[(size_t)(8 * threadIdx.x + j) * sstride];
  ntt_inv(x, lds, T);
  uint64_t *dst = fd + (size_t)(c >> 1) * fd_stride + ((size_t)r * 2 + (c & 1)) * kN;
#pragma unroll
  for (int j = 0; j < 8; ++j) dst[threadIdx.x + 256 * j] = ntt_inv_final(x[j], T);
}

}  // namespace onion
