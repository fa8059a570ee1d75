// k_db.hpp -- database ingest on the device: seeded generation (or uploaded
// plaintext coefficients), plain lift, NTT, and the slot-major realignment
// (PirServer::gen_data / push_database_chunk / preprocess_ntt / realign_db,
//  src/server.cpp:47-78, 635-748).
#pragma once

#include "k_ntt.hpp"

namespace onion {

struct LiftConst {
  uint64_t t;    // plaintext modulus
  uint64_t thr;  // (t + 1) / 2
  uint64_t inc;  // q - t
};

// SEAL transform_to_ntt_inplace(Plaintext): v >= (t+1)/2 becomes v + (q - t)
__device__ __forceinline__ uint64_t plain_lift(uint64_t v, const LiftConst &L) {
  return v >= L.thr ? v + L.inc : v;
}

// NTT of the lifted plaintext (natural layout in x), then
// db[slot * num_pt + pt] for every slot.
__device__ __forceinline__ void ntt_and_scatter(uint64_t (&x)[8], uint64_t *lds, uint64_t *db,
                                                uint64_t pt, uint64_t num_pt, const NttTab &T) {
  ntt_fwd(x, lds, T);
#pragma unroll
  for (int r = 0; r < 8; ++r) db[(size_t)(8 * threadIdx.x + r) * num_pt + pt] = ntt_fwd_final(x[r], T.q);
}

// Byte b of the seeded entry of plaintext pt: an 8-byte big-endian entry
// index (utils::writeIdxToEntry, src/utils.cpp:98-104), then random bytes.
__device__ __forceinline__ uint32_t entry_byte(uint64_t key, uint64_t pt, uint32_t b) {
  return b < 8 ? (uint32_t)((pt >> (8 * (7 - b))) & 0xFF)
               : (uint32_t)((rand64_keyed(key, pt * kN + b / 8) >> (8 * (b % 8))) & 0xFF);
}

// Seeded generation, one workgroup per plaintext pt_lo + blockIdx.x.
// mode 0: coefficient e = floor(r * t / 2^64), r = rand64(key, pt*N + e)
// mode 1: the entry bytes above packed bpc bits per coefficient, LSB first
//         (push_database_chunk, src/server.cpp:635-701); bpc <= 24.
__global__ __launch_bounds__(kNttThreads) void k_db_generate(uint64_t *db, uint64_t pt_lo,
                                                             uint64_t num_pt, uint64_t key,
                                                             int mode, uint32_t bpc, LiftConst L,
                                                             NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  const uint64_t pt = pt_lo + blockIdx.x;
  const uint32_t entry_bytes = bpc * kN / 8;
  uint64_t x[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const uint32_t e = threadIdx.x + 256 * r;
    uint64_t v;
    if (mode == 0) {
      v = mulhi64(rand64_keyed(key, pt * kN + e), L.t);
    } else {
      const uint32_t bit = e * bpc, b0 = bit / 8;
      uint32_t window = 0;
      for (uint32_t h = 0; h < 4 && b0 + h < entry_bytes; ++h) window |= entry_byte(key, pt, b0 + h) << (8 * h);
      v = (window >> (bit % 8)) & ((1u << bpc) - 1);
    }
    x[r] = plain_lift(v, L);
  }
  ntt_and_scatter(x, lds, db, pt, num_pt, T);
}

// Uploaded plaintexts: coeffs[i*N + e] (< t) for pt = pt_lo + i.
__global__ __launch_bounds__(kNttThreads) void k_db_plain(const uint64_t *coeffs, uint64_t *db,
                                                          uint64_t pt_lo, uint64_t num_pt,
                                                          LiftConst L, NttTab T) {
  __shared__ uint64_t lds[kNttLds];
  const uint64_t pt = pt_lo + blockIdx.x;
  uint64_t x[8];
  load_natural(x, coeffs + (size_t)blockIdx.x * kN);
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = plain_lift(x[r], L);
  ntt_and_scatter(x, lds, db, pt, num_pt, T);
}

}  // namespace onion
