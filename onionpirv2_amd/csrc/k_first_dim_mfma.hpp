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
// The epilogue forms C mod 2^128 from the 15 signed sums and reduces it mod q
// exactly as the 64-bit path does, so results are bit-identical.
//
// Operand layout.  MFMA lane l holds 16 bytes of row (l % 32) of the A tile
// and 16 bytes of column (l % 32) of the B tile, both in the same k order;
// since the sum runs over all 32 k, any assignment of j to (lane group, byte)
// shared by A and B gives the same result.  Here byte b of lane group
// g = l / 32 is j = 16 g + b of the 32-wide step, so an operand ("chunk")
// holds one digit of 16 consecutive values of j of one row or column.
//
// Chunk order.  The plane buffers keep chunks in the order the matrix cores
// consume them:
//   chunk (tile, s, i, l)  at  16 * (((tile * steps + s) * 8 + i) * 64 + l)
// for step s (steps = fst / 32), digit i and lane l = 32 g + (row or column
// within the 32-block), where tile = slot * n_blocks + block.  A wave reads
// digit i of a step as one contiguous 1 KiB load that lands in its operand
// registers as it stands (8 cache lines per load rather than 64).
//   B planes  written by k_sel_chunks; 32-column blocks of col = 2*query +
//             poly, padded to NCP columns (NCP = 32, 64 or 128)
//   A planes  the database's digit-plane copy (k_db_planes), 32-row blocks,
//             rows padded to a multiple of 32
//   A words   without the copy each lane reads the 16 database words of its
//             row and converts them in registers
//   output    out2[r][l][c] (uint64, canonical mod q, c < 2*nq): a tile's
//             32 columns of one (row, slot) are 256 contiguous bytes, and one
//             row's 2048 slots are one contiguous 2 MiB run (1 KiB per slot at
//             128 columns), read by the first multiplexer level
//             (k_mux_fused_ntt<true>, in k_gsw.hpp) or by k_ntt_inv_gather,
//             which transforms the rows into coefficient-form ciphertexts.
#pragma once

#include "k_first_dim.hpp"

namespace onion {

typedef int v4i_t __attribute__((ext_vector_type(4)));
typedef int v16i_t __attribute__((ext_vector_type(16)));

constexpr uint64_t kDigitBias = 0x8080808080808080ULL;
constexpr uint32_t kStepBytes = 8 * 1024;  // one tile's 8 chunks x 64 lanes

__device__ __forceinline__ uint64_t signed_digits(uint64_t a) { return (a + kDigitBias) ^ kDigitBias; }

// Byte offset of chunk (tile, s, i, l).
__device__ __forceinline__ size_t chunk_offset(size_t tile, uint32_t steps, uint32_t s, uint32_t i, uint32_t l) {
  return (((tile * steps + s) * 8 + i) * 64 + l) * 16;
}

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

// Eight values (already in signed-digit form) -> their 8 digit planes of 8
// bytes: out[2i + h] = digit i of values 4h .. 4h+3.
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

// Eight raw words (four uint4) -> their digit planes.
__device__ __forceinline__ void a_planes(const uint4 *a, uint32_t (&out)[16]) {
  uint64_t w[8];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    w[2 * q] = signed_digits(((uint64_t)a[q].y << 32) | a[q].x);
    w[2 * q + 1] = signed_digits(((uint64_t)a[q].w << 32) | a[q].z);
  }
  planes8(w, out);
}

// Selection vectors as B chunks, in two launches so that both write
// contiguous lines (a single pass owning 8 leaves of one column writes 8-byte
// pieces 1 KiB apart, and measured 4.3x its bytes in write traffic).
//   k_sel_digits  NTT of each leaf polynomial, in place, in signed-digit
//                 form (slot order).  grid = (fst * 2, nq).  Skipped when
//                 the leaves are in NTT form already (k_expand_ntt):
//                 k_sel_chunks<true> then takes canonical residues.
//   k_sel_chunks  8 slots x one 32-column block x one step per workgroup:
//                 reads each (column, j)'s 8 slots (64 contiguous bytes) into
//                 LDS, then each thread turns 16 values of j of one column and
//                 slot into that lane's 8 chunks; a wave stores one digit of
//                 one slot as 1 KiB.  grid = (N / 8, NCP / 32, fst / 32).
// Leaf j, poly c of query b is at leaves + b * leaf_stride + (2j + c) * N.
__global__ __launch_bounds__(kNttThreads) void k_sel_digits(uint64_t *leaves, uint64_t leaf_stride, NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  uint64_t *p = leaves + (size_t)blockIdx.y * leaf_stride + (size_t)blockIdx.x * kN;
  uint64_t x[8];
  load_natural(x, p);
  ntt_fwd(x, lds, T);  // every thread has read its inputs before the first barrier inside
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = signed_digits(ntt_fwd_final(x[r], T.q));
  store_blocked(p, x);
}

template <bool CANON>
__global__ __launch_bounds__(256) void k_sel_chunks(const uint64_t *__restrict__ leaves, uint64_t leaf_stride,
                                                    uint8_t *__restrict__ Bp, uint32_t fst, uint32_t ncols) {
  __shared__ uint64_t buf[8][32][33];  // [slot][column][j], rows padded against bank conflicts
  // slot groups 2m and 2m + 1 read the two halves of the same 128-byte lines:
  // blocks L and L + 8 (same XCD, dispatched together) take them
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
        buf[2 * q][li][jj] = CANON ? signed_digits(a) : a;
        buf[2 * q + 1][li][jj] = CANON ? signed_digits(b) : b;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint32_t pair = threadIdx.x + 256 * k, sg = pair / 64, l = pair % 64, li = l % 32, g = l / 32;
    uint64_t w0[8], w1[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      w0[t] = buf[sg][li][16 * g + t];
      w1[t] = buf[sg][li][16 * g + 8 + t];
    }
    uint32_t p0[16], p1[16];
    planes8(w0, p0);
    planes8(w1, p1);
    uint8_t *dst = Bp + chunk_offset((size_t)(slot0 + sg) * ncb + cb, fst / 32, s, 0, l);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      *reinterpret_cast<uint4 *>(dst + (size_t)i * 1024) = make_uint4(p0[2 * i], p0[2 * i + 1], p1[2 * i], p1[2 * i + 1]);
  }
}

// The database's digit-plane copy: records rec_lo .. rec_hi - 1 of every
// slot, a record being the 16 words j = 16 jp .. 16 jp + 15 of one row
// (record = row * fst / 16 + jp; num_pt = fst * other, fst % 32 == 0).  Each
// becomes the 8 chunks (tile slot * n_rb + row / 32, step jp / 2, lane
// 32 (jp % 2) + row % 32).  grid = (ceil((rec_hi - rec_lo) / 256), N).
__global__ __launch_bounds__(256) void k_db_planes(const uint64_t *__restrict__ db, uint8_t *__restrict__ dbp,
                                                   uint32_t fst, uint32_t other, uint64_t rec_lo,
                                                   uint64_t rec_hi) {
  const uint64_t r = rec_lo + (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= rec_hi) return;
  const uint32_t per_row = fst / 16, n_rb = (other + 31) / 32;
  const uint32_t row = (uint32_t)(r / per_row), jp = (uint32_t)(r % per_row);
  const size_t slot = blockIdx.y;
  const uint4 *src = reinterpret_cast<const uint4 *>(db + slot * fst * other + 16 * r);
  uint4 a[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) a[q] = src[q];
  uint32_t p0[16], p1[16];
  a_planes(a, p0);
  a_planes(a + 4, p1);
  uint8_t *dst = dbp + chunk_offset(slot * n_rb + row / 32, fst / 32, jp / 2, 0, 32 * (jp % 2) + row % 32);
#pragma unroll
  for (int k = 0; k < 8; ++k)
    *reinterpret_cast<uint4 *>(dst + (size_t)k * 1024) = make_uint4(p0[2 * k], p0[2 * k + 1], p1[2 * k], p1[2 * k + 1]);
}

// Operand of digit i from planes: plane i of values j = 16g .. 16g+7 (p0),
// then of 16g+8 .. 16g+15 (p1).
__device__ __forceinline__ v4i_t operand(const uint32_t (&p0)[16], const uint32_t (&p1)[16], int i) {
  v4i_t v;
  v[0] = (int)p0[2 * i];
  v[1] = (int)p0[2 * i + 1];
  v[2] = (int)p1[2 * i];
  v[3] = (int)p1[2 * i + 1];
  return v;
}

// Operand from a chunk as loaded.
__device__ __forceinline__ v4i_t as_operand(const uint4 &v) {
  v4i_t r;
  r[0] = (int)v.x;
  r[1] = (int)v.y;
  r[2] = (int)v.z;
  r[3] = (int)v.w;
  return r;
}

// The 64 products of one step from B chunks and A chunks (step_products) or
// A database words converted in registers (step_products_words).
__device__ __forceinline__ void step_products(v16i_t (&acc)[15], const uint4 (&A)[8], const uint4 (&B)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const v4i_t a = as_operand(A[i]);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      acc[i + k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, as_operand(B[k]), acc[i + k], 0, 0, 0);
  }
}
__device__ __forceinline__ void step_products_words(v16i_t (&acc)[15], const uint4 (&A)[8], const uint4 (&B)[8]) {
  uint32_t pa0[16], pa1[16];
  a_planes(A, pa0);
  a_planes(A + 4, pa1);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const v4i_t a = operand(pa0, pa1, i);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      acc[i + k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, as_operand(B[k]), acc[i + k], 0, 0, 0);
  }
}

// The weights 2^{8d}, d = 0..3, held in registers the optimizer cannot see
// through, so S * 2^{8d} + P stays one v_mad_i64_i32 (a product by a known
// power of two would become a sign extension, a 64-bit shift and an add).
struct DigitWeights {
  int32_t w[4];
  __device__ __forceinline__ DigitWeights() {
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      w[d] = 1 << (8 * d);
      asm volatile("" : "+v"(w[d]));
    }
  }
};

// C mod 2^128 from the 15 diagonal sums of one output, then mod q.
__device__ __forceinline__ uint64_t diagonals_mod_q(const v16i_t (&acc)[15], int e, const Mod &mq,
                                                    const DigitWeights &W) {
  // P_k = sum_{d<4} S_{4k+d} 2^{8d}: |S| < 2^31, so |P_k| < 2^56
  int64_t P[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    P[k] = acc[4 * k][e];
#pragma unroll
    for (int d = 1; d < 4; ++d)
      if (4 * k + d < 15) P[k] = (int64_t)acc[4 * k + d][e] * (int64_t)W.w[d] + P[k];
  }
  // C = P0 + P1 2^32 + P2 2^64 + P3 2^96 mod 2^128
  uint64_t lo = (uint64_t)P[0], hi = (uint64_t)(P[0] >> 63);
  add128(lo, hi, (uint64_t)P[1] << 32, (uint64_t)(P[1] >> 32));
  hi += (uint64_t)P[2] + ((uint64_t)P[3] << 32);
  return reduce128(lo, hi, mq);
}

// Common tile geometry.  CB = column blocks of 32 (NCP = 32 CB padded
// columns); the workgroup's four waves cover RB = 4 / CB row blocks of 32
// rows x CB column blocks.  grid.x = N * n_rt with n_rt row tiles of 32 RB
// rows; all row tiles of a slot run on one XCD (block id mod 8), so the
// slot's B chunks are read from that XCD's L2.  Requires fst % 32 == 0.
template <int CB>
struct Tile {
  static constexpr int RB = 4 / CB, NCP = 32 * CB;
  uint32_t slot, rt, w, l, g, li, rb, cb, row0, rblk, n_rb, steps;
  __device__ __forceinline__ Tile(uint32_t fst, uint32_t other, uint32_t n_rt) {
    const uint32_t L = blockIdx.x, xcd = L & 7, seq = L >> 3;
    slot = (seq / n_rt) * 8 + xcd;
    rt = seq % n_rt;
    w = threadIdx.x >> 6;
    l = threadIdx.x & 63;
    g = l >> 5;
    li = l & 31;
    rb = w / CB;
    cb = w % CB;
    row0 = rt * (32 * RB) + rb * 32;
    n_rb = (other + 31) / 32;
    rblk = min(row0 / 32, n_rb - 1);  // blocks wholly past the matrix read the last one; not stored
    steps = fst / 32;
  }
  __device__ __forceinline__ const uint8_t *b_base(const uint8_t *Bp) const {
    return Bp + chunk_offset((size_t)slot * CB + cb, steps, 0, 0, l);
  }
  __device__ __forceinline__ void store(const v16i_t (&acc)[15], uint64_t *out2, uint32_t other, uint32_t ncols,
                                        const Mod &mq) const {
    const uint32_t col = cb * 32 + li;
    if (col >= ncols) return;
    uint64_t *dst = out2 + (size_t)slot * ncols + col;
    const DigitWeights W;
    // output (m, n) of register e: n = lane % 32, m = 8 (e / 4) + 4 g + e % 4
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const uint32_t m = row0 + 8 * (e / 4) + 4 * g + (e % 4);
      const uint64_t v = diagonals_mod_q(acc, e, mq, W);
      if (m < other) dst[(size_t)m * kN * ncols] = v;
    }
  }
};

// Operands of one step for this lane: A (chunks, or the 16 database words of
// the row) and the B chunks.
struct StepRaw {
  uint4 a[8];
  uint4 b[8];
};

__device__ __forceinline__ void load_chunks(uint4 (&x)[8], const uint8_t *base, uint32_t s) {
  const uint8_t *p = base + (size_t)s * kStepBytes;
#pragma unroll
  for (int i = 0; i < 8; ++i) x[i] = *reinterpret_cast<const uint4 *>(p + i * 1024);
}
__device__ __forceinline__ void load_words(uint4 (&x)[8], const uint64_t *arow, uint32_t s, uint32_t g) {
  const uint4 *p = reinterpret_cast<const uint4 *>(arow + 32 * s + 16 * g);
#pragma unroll
  for (int q = 0; q < 8; ++q) x[q] = p[q];
}

// First dimension for 2..64 queries from the database words: one register
// buffer; the step's words become planes, then the buffer takes the next
// step's loads while the 64 products run.
template <int CB>
__global__ __launch_bounds__(256) void k_first_dim_mfma(const uint64_t *__restrict__ db,
                                                        const uint8_t *__restrict__ Bp,
                                                        uint64_t *__restrict__ out2, uint32_t fst,
                                                        uint32_t other, uint32_t ncols, uint32_t n_rt, Mod mq) {
  const Tile<CB> T(fst, other, n_rt);
  const uint32_t row = T.row0 + T.li;
  // rows past the matrix read row 0; their outputs are not stored
  const uint64_t *arow = db + (size_t)T.slot * fst * other + (size_t)(row < other ? row : 0) * fst;
  const uint8_t *bb = T.b_base(Bp);
  v16i_t acc[15];
#pragma unroll
  for (int d = 0; d < 15; ++d) acc[d] = (v16i_t){0};
  StepRaw R;
  load_words(R.a, arow, 0, T.g);
  load_chunks(R.b, bb, 0);
  for (uint32_t s = 0; s < T.steps; ++s) {
    uint32_t pa0[16], pa1[16];
    a_planes(R.a, pa0);
    a_planes(R.a + 4, pa1);
    uint4 B[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) B[k] = R.b[k];
    if (s + 1 < T.steps) {
      load_words(R.a, arow, s + 1, T.g);
      load_chunks(R.b, bb, s + 1);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const v4i_t a = operand(pa0, pa1, i);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        acc[i + k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, as_operand(B[k]), acc[i + k], 0, 0, 0);
    }
  }
  T.store(acc, out2, other, ncols, mq);
}

// First dimension for 2..64 queries from the digit-plane copy: every operand
// is a contiguous 1 KiB load; three register buffers, so each step's chunks
// are loaded two steps before their products.
template <int CB>
__global__ __launch_bounds__(256) void k_first_dim_mfma_pre(const uint8_t *__restrict__ dbp,
                                                            const uint8_t *__restrict__ Bp,
                                                            uint64_t *__restrict__ out2, uint32_t fst,
                                                            uint32_t other, uint32_t ncols, uint32_t n_rt, Mod mq) {
  const Tile<CB> T(fst, other, n_rt);
  const uint8_t *ab = dbp + chunk_offset((size_t)T.slot * T.n_rb + T.rblk, T.steps, 0, 0, T.l);
  const uint8_t *bb = T.b_base(Bp);
  v16i_t acc[15];
#pragma unroll
  for (int d = 0; d < 15; ++d) acc[d] = (v16i_t){0};
  const uint32_t steps = T.steps, body = steps - steps % 3;
  StepRaw R[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    load_chunks(R[d].a, ab, min((uint32_t)d, steps - 1));
    load_chunks(R[d].b, bb, min((uint32_t)d, steps - 1));
  }
  // groups of three steps; a buffer is refilled (with step s + d + 3, or
  // again the last step once past the end) as soon as its products issue
  for (uint32_t s = 0; s < body; s += 3) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      step_products(acc, R[d].a, R[d].b);
      const uint32_t n = min(s + d + 3, steps - 1);
      load_chunks(R[d].a, ab, n);
      load_chunks(R[d].b, bb, n);
    }
  }
#pragma unroll
  for (int d = 0; d < 2; ++d)
    if (body + d < steps) step_products(acc, R[d].a, R[d].b);
  T.store(acc, out2, other, ncols, mq);
}

// ---- Persistent variant ----------------------------------------------------------------
//
// k_first_dim_mfma_pre spends about a fifth of a tile in its epilogue with the
// matrix cores idle (one wave per SIMD: no other wave fills the gap), and each
// new workgroup waits for its first loads.  Here a workgroup walks tiles
// blockIdx.x, blockIdx.x + G, ... (G = gridDim.x, a multiple of 8, so every
// tile of a workgroup keeps its XCD and the slot-to-XCD assignment of the
// plain variant).  Two register buffers form one stream of steps over all of
// a workgroup's tiles (a step's loads issue as soon as the products of the
// step two back have), so the loads of a tile's first steps are in flight
// while the previous tile ends.  At a tile's end each wave forms C mod 2^128
// of its 16 outputs from the accumulators (the serial part) and parks them in
// LDS; during the next tile's steps 0..15 it reduces one parked output per
// step (Barrett, the store), VALU work that issues between the matrix
// products.  Outputs outside the matrix are dropped by the buffer store's
// range check (offset 2^31 against a 2^31-byte range, or a range of 0 while
// nothing is parked), so every step is branch-free.  Requires fst % 64 == 0
// (an even number of steps); a third buffer would not fit beside the
// accumulators and the epilogue.

constexpr int kStageWords = 16 * 64;  // per wave: 16 outputs x 64 lanes of (lo, hi)

__device__ __forceinline__ void diagonals_c128(const v16i_t (&acc)[15], int e, const DigitWeights &W, uint64_t &lo,
                                               uint64_t &hi) {
  int64_t P[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    P[k] = acc[4 * k][e];
#pragma unroll
    for (int d = 1; d < 4; ++d)
      if (4 * k + d < 15) P[k] = (int64_t)acc[4 * k + d][e] * (int64_t)W.w[d] + P[k];
  }
  lo = (uint64_t)P[0];
  hi = (uint64_t)(P[0] >> 63);
  add128(lo, hi, (uint64_t)P[1] << 32, (uint64_t)(P[1] >> 32));
  hi += (uint64_t)P[2] + ((uint64_t)P[3] << 32);
}

typedef unsigned v2u_t __attribute__((ext_vector_type(2)));

// A step's 8 chunks from a wave-uniform base (scalar address) plus the lane's
// byte offset.
__device__ __forceinline__ void load_chunks_at(uint4 (&x)[8], const uint8_t *base, uint32_t s, uint32_t lane_byte) {
  const uint8_t *p = base + (size_t)s * kStepBytes;
#pragma unroll
  for (int i = 0; i < 8; ++i) x[i] = *reinterpret_cast<const uint4 *>(p + lane_byte + i * 1024);
}

// Where a wave's outputs go: row m = row0 + 8 (e / 4) + 4 g + e % 4 of slot
// `slot`, columns cb * 32 + (lane % 32).
struct OutTile {
  uint32_t slot, row0;
};

template <int CB>
__global__ __launch_bounds__(256) void k_first_dim_mfma_tiles(const uint8_t *__restrict__ dbp,
                                                              const uint8_t *__restrict__ Bp,
                                                              uint64_t *__restrict__ out2, uint32_t fst,
                                                              uint32_t other, uint32_t ncols, uint32_t n_rt, Mod mq) {
  constexpr int RB = 4 / CB;
  __shared__ uint4 stage[4 * kStageWords];
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t l = threadIdx.x & 63, g = l >> 5, li = l & 31;
  const uint32_t rb = w / CB, cb = w % CB;
  const uint32_t n_rb = (other + 31) / 32, steps = fst / 32, total = kN * n_rt, G = gridDim.x;
  if (blockIdx.x >= total) return;
  uint4 *st = stage + w * kStageWords + l;
  // tile L: slot (L / 8 / n_rt) * 8 + L % 8, row tile (L / 8) % n_rt
  auto geo = [&](uint32_t L) {
    const uint32_t seq = L >> 3;
    return OutTile{(seq / n_rt) * 8 + (L & 7), (seq % n_rt) * (32 * RB) + rb * 32};
  };
  // loader: tile lt, step ls; past the last tile it repeats the last step
  uint32_t lt = blockIdx.x, ls = 0;
  const uint8_t *la, *lb;
  auto set_bases = [&]() {
    const OutTile o = geo(lt);
    la = dbp + chunk_offset((size_t)o.slot * n_rb + min(o.row0 / 32, n_rb - 1), steps, 0, 0, 0);
    lb = Bp + chunk_offset((size_t)o.slot * CB + cb, steps, 0, 0, 0);
  };
  auto load_next = [&](StepRaw &r) {
    load_chunks_at(r.a, la, ls, l * 16);
    load_chunks_at(r.b, lb, ls, l * 16);
    if (++ls == steps) {
      if (lt + G < total) {
        ls = 0;
        lt += G;
        set_bases();
      } else {
        ls = steps - 1;
      }
    }
  };
  set_bases();
  StepRaw R[2];
  load_next(R[0]);
  load_next(R[1]);

  // output e of the parked tile `prev`, stored only when `live`
  const uint32_t lane_off = (4 * g * kN * ncols + li) * 8;
  const bool col_ok = cb * 32 + li < ncols;
  auto put = [&](const OutTile &prev, uint32_t e, bool live) {
    const uint4 c = st[e * 64];
    // the Barrett form here, not reduce128: with the shorter reduction the
    // stores issue earlier in each step and the kernel read 1.4x the HBM bytes
    // (FETCH_SIZE 6.8e6 against 4.8e6 KiB per batch of 64 at 512x1024) in the
    // same time; the matrix cores set the pace either way
    const uint64_t v = barrett128((uint64_t)c.x | (uint64_t)c.y << 32, (uint64_t)c.z | (uint64_t)c.w << 32, mq);
    const uint32_t urow = prev.row0 + 8 * (e >> 2) + (e & 3);
    uint64_t *base = out2 + ((size_t)urow * kN + prev.slot) * ncols + cb * 32;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(base, 0, live ? 0x80000000u : 0u, 0x00020000);
    const uint32_t off = col_ok && urow + 4 * g < other ? lane_off : 0x80000000u;
    v2u_t vv;
    vv[0] = (uint32_t)v;
    vv[1] = (uint32_t)(v >> 32);
    __builtin_amdgcn_raw_buffer_store_b64(vv, rs, off, 0, 0);
  };

  const DigitWeights W;
  v16i_t acc[15];
#pragma unroll
  for (int d = 0; d < 15; ++d) acc[d] = (v16i_t){0};
  OutTile prev{0, 0};
  bool have_prev = false;
  for (uint32_t ct = blockIdx.x; ct < total; ct += G) {
    for (uint32_t s = 0; s < steps; s += 2) {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        step_products(acc, R[d].a, R[d].b);
        put(prev, min(s + d, 15u), have_prev && s + d < 16);
        load_next(R[d]);
        // spread the step's VALU work evenly between its 64 products (grouped
        // after a few of them it would stall the matrix core)
#pragma unroll
        for (int i = 0; i < 64; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
          __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // two VALU
        }
      }
    }
    if (have_prev)
      for (uint32_t e = steps; e < 16; ++e) put(prev, e, true);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      uint64_t lo, hi;
      diagonals_c128(acc, e, W, lo, hi);
      st[e * 64] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
    }
#pragma unroll
    for (int k = 0; k < 15; ++k) acc[k] = (v16i_t){0};
    prev = geo(ct);
    have_prev = true;
  }
  for (uint32_t e = 0; e < 16; ++e) put(prev, e, true);
}

// ---- LDS-pipelined variant ------------------------------------------------------------
//
// Same tiles and arithmetic, but the operands travel global -> LDS by
// direct-to-LDS loads (global_load_lds_dwordx4), three steps deep, so no VGPRs
// hold loads in flight, and the RB row blocks and CB column blocks of a step
// (8 (RB + CB) "pieces" of 1 KiB) are loaded once by the four waves together.
// Piece p of a block holds, at 16 l, reading lane l's operand for digit p
// (chunks), or for A without the plane copy the database words
// j = 32 s + 16 g + 2 p, +1 of row (l % 32).  Both the DMA write and the
// reading ds_read_b128 are contiguous.

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
// a_base: the tile's first A chunk (PRE) or the slot's database words.
template <int CB, bool PRE>
__device__ __forceinline__ void issue_step(uint8_t *lds, int stage, uint32_t s, const Tile<CB> &T,
                                           const uint8_t *a_base, uint32_t fst, uint32_t other,
                                           const uint8_t *b_base) {
  using P = MfmaPipe<CB>;
#pragma unroll
  for (int i = 0; i < P::PER_WAVE; ++i) {
    const uint32_t piece = T.w + 4 * i;  // wave-uniform
    uint8_t *dst = lds + stage * P::STAGE_BYTES + piece * 1024;
    const uint32_t p = piece & 7;
    const uint32_t blk = piece >> 3;
    if (blk < (uint32_t)P::RB) {
      if (PRE) {
        const uint32_t rblk = min(T.rt * P::RB + blk, T.n_rb - 1);
        lds_dma16(a_base + chunk_offset(rblk, T.steps, s, p, T.l), dst);
      } else {
        uint32_t row = T.rt * (32 * P::RB) + blk * 32 + T.li;
        row = row < other ? row : 0;  // masked rows read row 0; their outputs are not stored
        lds_dma16(reinterpret_cast<const uint64_t *>(a_base) + (size_t)row * fst + 32 * s + 16 * T.g + 2 * p, dst);
      }
    } else {
      lds_dma16(b_base + chunk_offset(blk - P::RB, T.steps, s, p, T.l), dst);
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

// PRE: A from the digit-plane copy dbp (chunks); otherwise from the database
// words db.
template <int CB, bool PRE>
__global__ __launch_bounds__(256) void k_first_dim_mfma_lds(const uint64_t *__restrict__ db,
                                                            const uint8_t *__restrict__ dbp,
                                                            const uint8_t *__restrict__ Bp,
                                                            uint64_t *__restrict__ out2, uint32_t fst,
                                                            uint32_t other, uint32_t ncols, uint32_t n_rt, Mod mq) {
  using P = MfmaPipe<CB>;
  constexpr int RB = P::RB;
  __shared__ __attribute__((aligned(16))) uint8_t lds[kMfmaStages * P::STAGE_BYTES];
  const Tile<CB> T(fst, other, n_rt);
  const uint8_t *a_base = PRE ? dbp + chunk_offset((size_t)T.slot * T.n_rb, T.steps, 0, 0, 0)
                              : reinterpret_cast<const uint8_t *>(db + (size_t)T.slot * fst * other);
  const uint8_t *b_base = Bp + chunk_offset((size_t)T.slot * CB, T.steps, 0, 0, 0);

  v16i_t acc[15];
#pragma unroll
  for (int d = 0; d < 15; ++d) acc[d] = (v16i_t){0};

  // LDS byte offset of the staging area (generic -> local address space)
  const uint32_t lds_off =
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) uint8_t *)lds + T.l * 16;
  const uint32_t steps = T.steps;
  issue_step<CB, PRE>(lds, 0, 0, T, a_base, fst, other, b_base);
  if (steps > 1) issue_step<CB, PRE>(lds, 1, 1, T, a_base, fst, other, b_base);
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
    if (s + 2 < steps) issue_step<CB, PRE>(lds, (s + 2) % kMfmaStages, s + 2, T, a_base, fst, other, b_base);
    const uint32_t st = lds_off + (s % kMfmaStages) * P::STAGE_BYTES;
    uint4 ra[8], rbv[8];
    lds_read_step(ra, rbv, st + T.rb * 8 * 1024, st + (RB + T.cb) * 8 * 1024);
    if (PRE)
      step_products(acc, ra, rbv);
    else
      step_products_words(acc, ra, rbv);
  }
  T.store(acc, out2, other, ncols, mq);
}

// Inverse NTTs of the first-dimension outputs gathered from out2[r][slot][c]
// into fd (query b = c / 2, poly c % 2 of ciphertext r, coefficient form),
// for batches whose first multiplexer level does not read out2 itself.
// One workgroup per (row, group of G adjacent columns): a thread reads the G
// columns of each of its 8 slots as one contiguous 8G-byte piece (slots are
// 1 KiB apart at 128 columns) and runs the G transforms one after another.
// Rows are dealt to XCDs (r = xcd mod 8, column groups fastest) when the row
// count is a multiple of 8, so the workgroups reading one row's 2 MiB share an
// L2.
// Column K of the group, then the following ones (a recursion on a constant
// K keeps x in registers; a loop around the transform is not unrolled).
template <int G, int K = 0>
__device__ __forceinline__ void gather_transforms(uint64_t (&x)[G][8], uint64_t *lds, const NttTab &T, uint64_t *fd,
                                                  uint64_t fd_stride, uint32_t r, uint32_t c0) {
  ntt_inv(x[K], lds, T);
  const uint32_t c = c0 + K;
  uint64_t *dst = fd + (size_t)(c >> 1) * fd_stride + ((size_t)r * 2 + (c & 1)) * kN;
#pragma unroll
  for (int j = 0; j < 8; ++j) dst[threadIdx.x + 256 * j] = ntt_inv_final(x[K][j], T);
  if constexpr (K + 1 < G) gather_transforms<G, K + 1>(x, lds, T, fd, fd_stride, r, c0);
}

template <int G>
__global__ __launch_bounds__(kNttThreads) void k_ntt_inv_gather(const uint64_t *__restrict__ out2, uint32_t other,
                                                                uint32_t ncols, uint64_t *fd, uint64_t fd_stride,
                                                                NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  const uint32_t ngroups = ncols / G;
  uint32_t r, cg;
  if (other % 8 == 0) {
    const uint32_t xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    r = 8 * (k / ngroups) + xcd;
    cg = k % ngroups;
  } else {
    r = blockIdx.x / ngroups;
    cg = blockIdx.x % ngroups;
  }
  const uint64_t *src = out2 + (size_t)r * kN * ncols + cg * G;
  uint64_t x[G][8];
#pragma unroll
  for (int k = 0; k < G; k += 2) {  // column pairs: one 16-byte load per slot
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint4 v = *reinterpret_cast<const uint4 *>(src + (size_t)(8 * threadIdx.x + j) * ncols + k);
      x[k][j] = ((uint64_t)v.y << 32) | v.x;
      x[k + 1][j] = ((uint64_t)v.w << 32) | v.z;
    }
  }
  gather_transforms<G>(x, lds, T, fd, fd_stride, r, cg * G);
}

}  // namespace onion
