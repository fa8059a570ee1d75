// host_math.hpp -- host-side number theory for parameter derivation and NTT
// tables: Miller-Rabin, SEAL's prime selection rules, minimal primitive
// 2N-th roots, bit-reversed twiddle tables with Shoup companions.
#pragma once

#include <stdint.h>

#include <vector>

namespace onion {
namespace host {

typedef unsigned __int128 u128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

inline uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  a %= m;
  for (; e; e >>= 1) {
    if (e & 1) r = mulmod(r, a, m);
    a = mulmod(a, a, m);
  }
  return r;
}

inline uint64_t invmod(uint64_t a, uint64_t m) {
  // Fermat is enough: every modulus inverted here is prime
  return powmod(a % m, m - 2, m);
}

inline bool is_prime(uint64_t n) {
  if (n < 2) return false;
  static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t p : small) {
    if (n == p) return true;
    if (n % p == 0) return false;
  }
  uint64_t d = n - 1;
  int s = 0;
  while (!(d & 1)) { d >>= 1; ++s; }
  for (uint64_t a : small) {
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool composite = true;
    for (int r = 1; r < s && composite; ++r) {
      x = mulmod(x, x, n);
      if (x == n - 1) composite = false;
    }
    if (composite) return false;
  }
  return true;  // the first twelve primes as bases decide all n < 3.3e24
}

// primes = 1 mod factor with exactly `bits` bits, largest first
inline std::vector<uint64_t> primes_down(uint64_t factor, int bits, size_t count) {
  std::vector<uint64_t> out;
  uint64_t v = (uint64_t(1) << bits) - factor + 1, lo = uint64_t(1) << (bits - 1);
  while (out.size() < count && v > lo) {
    if (is_prime(v)) out.push_back(v);
    v -= factor;
  }
  return out;
}

inline uint64_t smallest_prime_above_pow2(int bits_minus_one) {
  uint64_t c = uint64_t(1) << bits_minus_one;
  do {
    ++c;
    c |= 1;
  } while (!is_prime(c));
  return c;
}

inline uint64_t bitrev(uint64_t x, int logn) {
  uint64_t y = 0;
  for (int i = 0; i < logn; ++i, x >>= 1) y = (y << 1) | (x & 1);
  return y;
}

inline uint64_t shoup(uint64_t w, uint64_t m) { return (uint64_t)(((u128)w << 64) / m); }

// smallest primitive 2N-th root of unity mod m
inline uint64_t min_primitive_root(uint64_t m, uint64_t N) {
  const uint64_t deg = 2 * N;
  uint64_t g = 0;
  for (uint64_t x = 2; x < m; ++x) {
    uint64_t r = powmod(x, (m - 1) / deg, m);
    if (powmod(r, N, m) == m - 1) { g = r; break; }
  }
  uint64_t best = g, cur = g, g2 = mulmod(g, g, m);
  for (uint64_t i = 1; i < N; ++i) {
    cur = mulmod(cur, g2, m);  // odd powers g^(2i+1)
    if (cur < best) best = cur;
  }
  return best;
}

struct Tables {
  uint64_t q = 0, psi = 0, n_inv = 0, n_inv_s = 0;
  std::vector<uint64_t> rp, rps, irp, irps;
};

inline Tables make_tables(uint64_t q, uint32_t N) {
  Tables T;
  int logn = 0;
  while ((1u << logn) < N) ++logn;
  T.q = q;
  T.psi = min_primitive_root(q, N);
  const uint64_t psi_inv = invmod(T.psi, q);
  T.rp.resize(N); T.rps.resize(N); T.irp.resize(N); T.irps.resize(N);
  uint64_t pw = 1, ipw = 1;
  for (uint32_t i = 0; i < N; ++i) {
    const uint64_t k = bitrev(i, logn);
    T.rp[k] = pw;
    T.irp[k] = ipw;
    pw = mulmod(pw, T.psi, q);
    ipw = mulmod(ipw, psi_inv, q);
  }
  for (uint32_t k = 0; k < N; ++k) {
    T.rps[k] = shoup(T.rp[k], q);
    T.irps[k] = shoup(T.irp[k], q);
  }
  T.n_inv = invmod(N, q);
  T.n_inv_s = shoup(T.n_inv, q);
  return T;
}

inline void barrett_ratio(uint64_t q, uint64_t &lo, uint64_t &hi) {
  const u128 r = (~(u128)0) / q;  // floor(2^128 / q) for odd q
  lo = (uint64_t)r;
  hi = (uint64_t)(r >> 64);
}

// The double scale (double)q_small / (double)q_big of the modulus switch as
// M * 2^-f with M < 2^53.
inline void mod_switch_scale(uint64_t q_big, uint64_t q_small, uint64_t &M, uint32_t &f) {
  const double s = (double)q_small / (double)q_big;
  int e = 0;
  double fr = __builtin_frexp(s, &e);
  M = (uint64_t)__builtin_ldexp(fr, 53);
  f = (uint32_t)(53 - e);
  while (M && !(M & 1) && f > 1) {
    M >>= 1;
    --f;
  }
}

}  // namespace host
}  // namespace onion
