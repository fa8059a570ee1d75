// params.hpp -- derivation of the scheme parameters from onion_params, by the
// reference's rules (PirParams::PirParams, src/pir.cpp:9-95;
// src/includes/pir.h:59-68): data prime and special prime from
// CoeffModulus::Create(N, {60, 60}) (the list is handed out from its back, so
// the data prime is the smaller one), small_q from Create(N, {27, 60})[0],
// t = smallest prime above 2^(plain_bits-1), base_log2 = ceil(60 / l),
// dims = [fst, 2, 2, ...], h = ceil(log2(fst + l*(ndim-1))).
#pragma once

#include <cmath>
#include <string>

#include "../../include/onionpir_gpu.h"
#include "host_math.hpp"

namespace onion {

inline uint32_t ceil_log2_u64(uint64_t x) {
  uint32_t r = 0;
  while ((uint64_t(1) << r) < x) ++r;
  return r;
}

// Returns ONION_OK or ONION_EINVAL with a message in err.
inline int derive_params(const onion_params &P, onion_param_info &I, std::string &err,
                         uint32_t required_n = 0) {
  if (P.poly_degree < 16 || (P.poly_degree & (P.poly_degree - 1))) {
    err = "poly_degree must be a power of two >= 16";
    return ONION_EINVAL;
  }
  if (required_n && P.poly_degree != required_n) {
    err = "the device path serves poly_degree " + std::to_string(required_n) + " only";
    return ONION_EINVAL;
  }
  if (P.plain_bits < 2 || P.plain_bits > 30) {
    err = "plain_bits out of range";
    return ONION_EINVAL;
  }
  // an external product sums 2l products of a transformed digit (below 2^63)
  // and a gadget value (below q < 2^60) in 128 bits: 32 of them stay below
  // 2^128, so l and l_key are at most 16
  if (P.gsw_l == 0 || P.gsw_l_key == 0 || P.gsw_l > 16 || P.gsw_l_key > 16) {
    err = "gsw_l / gsw_l_key out of range (1..16)";
    return ONION_EINVAL;
  }
  if (P.fst_dim == 0 || P.other_dim == 0 || (P.other_dim & (P.other_dim - 1))) {
    err = "fst_dim must be > 0 and other_dim a power of two";
    return ONION_EINVAL;
  }
  const uint64_t N = P.poly_degree;
  auto big = host::primes_down(2 * N, 60, 2);
  auto sq = host::primes_down(2 * N, 27, 1);
  if (big.size() != 2 || sq.size() != 1) {
    err = "no NTT-friendly primes for this ring degree";
    return ONION_EINVAL;
  }
  I = onion_param_info{};
  I.q = big[1];
  I.p = big[0];
  I.small_q = sq[0];
  I.t = host::smallest_prime_above_pow2((int)P.plain_bits - 1);
  const uint32_t qbits = 64 - __builtin_clzll(I.q);
  I.base_log2 = (qbits + P.gsw_l - 1) / P.gsw_l;
  I.base_log2_key = (qbits + P.gsw_l_key - 1) / P.gsw_l_key;
  I.ndim = 1 + ceil_log2_u64(P.other_dim);
  I.useful_cnt = P.fst_dim + P.gsw_l * (I.ndim - 1);
  I.expan_height = ceil_log2_u64(I.useful_cnt);
  if ((uint64_t(1) << I.expan_height) > N) {
    err = "expansion height above log2(N): too many first-dimension slots";
    return ONION_EINVAL;
  }
  I.num_pt = (uint64_t)P.fst_dim * P.other_dim;
  I.galois_key_words = I.expan_height * 4 * (uint32_t)N;
  I.rgsw_key_words = 2 * P.gsw_l_key * 2 * (uint32_t)N;
  const uint32_t wbits = (uint32_t)std::ceil(std::log2((double)I.small_q));
  I.response_bytes = (uint32_t)((2 * N * wbits + 7) / 8);
  I.poly_degree = (uint32_t)N;
  return ONION_OK;
}

inline uint32_t response_bits(uint64_t small_q) {
  return (uint32_t)std::ceil(std::log2((double)small_q));
}

}  // namespace onion
