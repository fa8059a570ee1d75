/*
 * oracle_client.c -- seeded test client for the OnionPIRv2 CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference client (src/client.cpp) draws its secret key, errors and
 * seeded c1 from SEAL's PRNGs, which are unavailable.  This client keeps the
 * reference's message layout exactly (query packing, Galois elements, RGSW
 * gadgets, decryption modulus) and draws all randomness from the counter-based
 * generator orc_rand64, so that every key, query and database is reproducible
 * from seeds.  The product's own client (onionpirv2_amd) follows the same
 * sampling rules; tests check the two bit for bit.
 *
 * Sampling rules (stream ids are part of the rule):
 *   secret key  s_i  = {0, 1, -1}[floor(r * 3 / 2^64)],   r = rand64(seed, 1, i)
 *   error       e_i  = popcnt(r & M21) - popcnt((r >> 21) & M21),
 *                      r = rand64(seed, kind_e<<56 | nonce, i)
 *   uniform     a_i  = (r_2i * 2^64 + r_2i+1) mod m,
 *                      r_j = rand64(seed, kind_a<<56 | limb<<48 | nonce, j)
 * Queries and test encryptions: kind_a = 2, kind_e = 3, the caller's nonce.
 * Key material: kind_a = 4, kind_e = 5; Galois key of level a has nonce
 * 0x10000 + a, RGSW key row r nonce 0x20000 + r.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef orc_u128 u128;

#define K_SK 1ULL
#define K_UNI 2ULL
#define K_ERR 3ULL
#define K_KEY_UNI 4ULL
#define K_KEY_ERR 5ULL

static inline uint64_t add_mod(uint64_t a, uint64_t b, uint64_t m) {
  uint64_t s = a + b;
  return s >= m ? s - m : s;
}

static void sample_uniform(uint64_t seed, uint64_t kind, uint64_t nonce, uint64_t limb, uint64_t m,
                           uint32_t N, uint64_t *out) {
  const uint64_t st = (kind << 56) | (limb << 48) | nonce;
  for (uint32_t i = 0; i < N; ++i) {
    uint64_t hi = orc_rand64(seed, st, 2 * (uint64_t)i), lo = orc_rand64(seed, st, 2 * (uint64_t)i + 1);
    out[i] = (uint64_t)((((u128)hi << 64) | lo) % m);
  }
}

static void sample_error(uint64_t seed, uint64_t kind, uint64_t nonce, uint32_t N, int64_t *out) {
  const uint64_t st = (kind << 56) | nonce, M21 = (1ULL << 21) - 1;
  for (uint32_t i = 0; i < N; ++i) {
    uint64_t r = orc_rand64(seed, st, i);
    out[i] = (int64_t)__builtin_popcountll(r & M21) - (int64_t)__builtin_popcountll((r >> 21) & M21);
  }
}

static inline uint64_t signed_to_mod(int64_t v, uint64_t m) {
  return v >= 0 ? (uint64_t)v % m : m - ((uint64_t)(-v) % m);
}

/* negacyclic product mod the NTT modulus `which` (0 = q, 1 = p, 2 = small_q) */
static void poly_mul(const orc_ctx *c, int which, const uint64_t *a, const uint64_t *b,
                     uint64_t *out) {
  const uint32_t N = c->N;
  const uint64_t m = which == 0 ? c->q : which == 1 ? c->p : c->small_q;
  uint64_t *x = orc_alloc(sizeof(uint64_t) * N), *y = orc_alloc(sizeof(uint64_t) * N);
  memcpy(x, a, sizeof(uint64_t) * N);
  memcpy(y, b, sizeof(uint64_t) * N);
  orc_ntt_fwd(c, which, x);
  orc_ntt_fwd(c, which, y);
  for (uint32_t i = 0; i < N; ++i) out[i] = orc_mulmod(x[i], y[i], m);
  orc_ntt_inv(c, which, out);
  free(x);
  free(y);
}

/* secret key as residues mod q: {0, 1, q-1} (KeyGenerator::secret_key) */
void orc_client_keygen(const orc_ctx *c, uint64_t seed, uint64_t *sk) {
  for (uint32_t i = 0; i < c->N; ++i) {
    uint64_t r = orc_rand64(seed, K_SK, i);
    uint64_t v = (uint64_t)(((u128)r * 3) >> 64);
    sk[i] = v == 2 ? c->q - 1 : v;
  }
}

static void sk_mod(const orc_ctx *c, const uint64_t *sk, uint64_t m, uint64_t *out) {
  for (uint32_t i = 0; i < c->N; ++i) out[i] = sk[i] > 1 ? m - 1 : sk[i];
}

/* symmetric encryption of zero mod `which`: (c0, c1) = (-a*s + e, a), a from
 * the uniform stream of kind `uni` */
static void enc_zero_mod(const orc_ctx *c, int which, uint64_t seed, uint64_t uni, uint64_t nonce,
                         const uint64_t *sk, const int64_t *e, uint64_t *c0, uint64_t *c1) {
  const uint32_t N = c->N;
  const uint64_t m = which == 0 ? c->q : c->p;
  uint64_t *s = orc_alloc(sizeof(uint64_t) * N), *as = orc_alloc(sizeof(uint64_t) * N);
  sk_mod(c, sk, m, s);
  sample_uniform(seed, uni, nonce, (uint64_t)which, m, N, c1);
  poly_mul(c, which, c1, s, as);
  for (uint32_t i = 0; i < N; ++i) c0[i] = add_mod(as[i] ? m - as[i] : 0, signed_to_mod(e[i], m), m);
  free(s);
  free(as);
}

static void encrypt_zero_kind(const orc_ctx *c, uint64_t seed, int key, uint64_t nonce, const uint64_t *sk,
                              uint64_t *ct) {
  int64_t *e = orc_alloc(sizeof(int64_t) * c->N);
  sample_error(seed, key ? K_KEY_ERR : K_ERR, nonce, c->N, e);
  enc_zero_mod(c, 0, seed, key ? K_KEY_UNI : K_UNI, nonce, sk, e, ct, ct + c->N);
  free(e);
}

void orc_client_encrypt_zero(const orc_ctx *c, uint64_t seed, uint64_t nonce, const uint64_t *sk,
                             uint64_t *ct) {
  encrypt_zero_kind(c, seed, 0, nonce, sk, ct);
}

/* BFV symmetric encryption: c0 += round(q * m / t) */
void orc_client_encrypt(const orc_ctx *c, uint64_t seed, uint64_t nonce, const uint64_t *sk,
                        const uint64_t *msg, uint64_t *ct) {
  orc_client_encrypt_zero(c, seed, nonce, sk, ct);
  for (uint32_t i = 0; i < c->N; ++i) {
    uint64_t m = msg[i] % c->t;
    uint64_t d = (uint64_t)(((u128)c->q * m + c->t / 2) / c->t);
    ct[i] = add_mod(ct[i], d % c->q, c->q);
  }
}

/* PirClient::create_galois_keys (src/client.cpp:227-241): elements
 * 1 + (N >> a), a < h.  One key-switching key per element
 * (KeyGenerator::generate_one_kswitch_key): over limbs (q, p),
 * comp0 = -a*s + e (+ (p mod q) * sigma(s) in the q limb), comp1 = a, NTT form.
 * gk layout: [a][comp][limb][N]. */
void orc_client_galois_keys(const orc_ctx *c, uint64_t seed, const uint64_t *sk, uint64_t *gk) {
  const uint32_t N = c->N;
  int64_t *e = orc_alloc(sizeof(int64_t) * N);
  uint64_t *gs = orc_alloc(sizeof(uint64_t) * N);
  uint64_t *c0 = orc_alloc(sizeof(uint64_t) * N), *c1 = orc_alloc(sizeof(uint64_t) * N);
  const uint64_t p_mod_q = c->p % c->q;
  for (uint32_t a = 0; a < c->h; ++a) {
    const uint32_t g = 1 + (N >> a);
    const uint64_t nonce = 0x10000 + a;
    uint64_t *key = gk + (size_t)a * 4 * N;
    orc_galois_coeff(c, sk, g, c->q, gs);
    sample_error(seed, K_KEY_ERR, nonce, N, e);
    for (int limb = 0; limb < 2; ++limb) {
      enc_zero_mod(c, limb, seed, K_KEY_UNI, nonce, sk, e, c0, c1);
      const uint64_t m = limb == 0 ? c->q : c->p;
      if (limb == 0)
        for (uint32_t i = 0; i < N; ++i) c0[i] = add_mod(c0[i], orc_mulmod(gs[i], p_mod_q, m), m);
      orc_ntt_fwd(c, limb, c0);
      orc_ntt_fwd(c, limb, c1);
      memcpy(key + (size_t)(0 * 2 + limb) * N, c0, sizeof(uint64_t) * N);
      memcpy(key + (size_t)(1 * 2 + limb) * N, c1, sizeof(uint64_t) * N);
    }
  }
  free(e); free(gs); free(c0); free(c1);
}

/* PirClient::generate_gsw_from_key (src/client.cpp:17-31) over
 * GSWEval::plain_to_half_gsw (src/gsw_eval.cpp:339-405): RGSW(s) with the key
 * gadget (l_key, B_key).  Rows k < l_key: Enc(0) with s*g_k added to c0; rows
 * l_key + k: Enc(0) whose phase gains s*g_k*s (here: s*g_k added to c1).
 * Coefficient form, rk layout [2*l_key][c0|c1][N]. */
void orc_client_rgsw_key(const orc_ctx *c, uint64_t seed, const uint64_t *sk, uint64_t *rk) {
  const uint32_t N = c->N, L = c->l_key;
  uint64_t *g = orc_alloc(sizeof(uint64_t) * L);
  orc_gadget(L, c->base_log2_key, c->q, g);
  for (uint32_t half = 0; half < 2; ++half)
    for (uint32_t k = 0; k < L; ++k) {
      uint64_t *row = rk + (size_t)(half * L + k) * 2 * N;
      encrypt_zero_kind(c, seed, 1, 0x20000 + half * L + k, sk, row);
      uint64_t *dst = row + (size_t)half * N;
      for (uint32_t i = 0; i < N; ++i) dst[i] = add_mod(dst[i], orc_mulmod(sk[i], g[k], c->q), c->q);
    }
  free(g);
}

/* RGSW(bit) with the data gadget, NTT form (GSWEval::plain_to_gsw,
 * src/gsw_eval.cpp:288-337, for tests of the external product) */
void orc_client_rgsw(const orc_ctx *c, uint64_t seed, uint64_t nonce, const uint64_t *sk,
                     uint64_t bit, uint64_t *rgsw_ntt) {
  const uint32_t N = c->N, L = c->l;
  uint64_t *g = orc_alloc(sizeof(uint64_t) * L);
  orc_gadget(L, c->base_log2, c->q, g);
  for (uint32_t half = 0; half < 2; ++half)
    for (uint32_t k = 0; k < L; ++k) {
      uint64_t *row = rgsw_ntt + (size_t)(half * L + k) * 2 * N;
      orc_client_encrypt_zero(c, seed, nonce * 64 + half * L + k, sk, row);
      if (bit) {
        uint64_t *dst = row + (size_t)half * N;
        dst[0] = add_mod(dst[0], orc_mulmod(bit, g[k], c->q), c->q);
      }
      orc_ntt_fwd(c, 0, row);
      orc_ntt_fwd(c, 0, row + N);
    }
  free(g);
}

/* PirClient::get_query_indices (src/client.cpp:38-53): [col, MSB-first bits] */
void orc_client_query_indices(const orc_ctx *c, uint64_t pt_index, uint64_t *idx) {
  uint64_t col = pt_index % c->fst, row = pt_index / c->fst, remain = c->num_pt / c->fst;
  idx[0] = col;
  for (uint32_t i = 1; i < c->ndim; ++i) {
    remain /= 2;
    idx[i] = row / remain;
    row %= remain;
  }
}

/* PirClient::fast_generate_query + add_gsw_to_query (src/client.cpp:130-209) */
void orc_client_query(const orc_ctx *c, uint64_t seed, uint64_t nonce, const uint64_t *sk,
                      uint64_t pt_index, uint64_t *ct) {
  const uint32_t N = c->N;
  uint64_t *idx = orc_alloc(sizeof(uint64_t) * c->ndim);
  uint64_t *msg = calloc(N, sizeof(uint64_t));
  orc_client_query_indices(c, pt_index, idx);
  const uint64_t w = (uint64_t)1 << c->h;
  msg[orc_bit_reverse(idx[0], c->h)] = orc_invmod(w % c->t, c->t);
  orc_client_encrypt(c, seed, nonce, sk, msg, ct);
  uint64_t *g = orc_alloc(sizeof(uint64_t) * c->l);
  orc_gadget(c->l, c->base_log2, c->q, g);
  const uint64_t inv_q = orc_invmod(w % c->q, c->q);
  for (uint32_t i = 1; i < c->ndim; ++i) {
    if (idx[i] != 1) continue;
    for (uint32_t k = 0; k < c->l; ++k) {
      uint64_t pos = orc_bit_reverse(c->fst + (uint64_t)(i - 1) * c->l + k, c->h);
      ct[pos] = add_mod(ct[pos], orc_mulmod(g[k], inv_q, c->q), c->q);
    }
  }
  free(idx); free(msg); free(g);
}

/* phase c0 + c1*s mod q */
void orc_client_phase(const orc_ctx *c, const uint64_t *sk, const uint64_t *ct, uint64_t *phase) {
  poly_mul(c, 0, ct + c->N, sk, phase);
  for (uint32_t i = 0; i < c->N; ++i) phase[i] = add_mod(phase[i], ct[i] % c->q, c->q);
}

/* BFV decryption mod q: m = round(t * phase / q) mod t */
void orc_client_decrypt(const orc_ctx *c, const uint64_t *sk, const uint64_t *ct, uint64_t *pt) {
  orc_client_phase(c, sk, ct, pt);
  for (uint32_t i = 0; i < c->N; ++i)
    pt[i] = (uint64_t)((((u128)pt[i] * c->t) + c->q / 2) / c->q) % c->t;
}

/* PirClient::decrypt_mod_q (src/client.cpp:458-479) with the switched secret
 * key of sk_mod_switch (src/client.cpp:482-512): -1 becomes small_q - 1. */
void orc_client_decrypt_small(const orc_ctx *c, const uint64_t *sk, const uint64_t *resp,
                              uint64_t *pt) {
  const uint32_t N = c->N;
  const uint64_t m = c->small_q;
  uint64_t *s = orc_alloc(sizeof(uint64_t) * N), *c1 = orc_alloc(sizeof(uint64_t) * N);
  sk_mod(c, sk, m, s);
  for (uint32_t i = 0; i < N; ++i) c1[i] = resp[N + i] % m;
  poly_mul(c, 2, c1, s, pt);
  for (uint32_t i = 0; i < N; ++i) {
    uint64_t ph = add_mod(pt[i], resp[i] % m, m);
    pt[i] = (uint64_t)((((u128)ph * c->t) + m / 2) / m) % c->t;
  }
  free(s);
  free(c1);
}

/* PirClient::get_entry_from_plaintext (src/client.cpp:255-287) for the single
 * entry of a plaintext (entry index 0 within it) */
void orc_entry_from_plaintext(const orc_ctx *c, const uint64_t *pt, uint8_t *entry) {
  const uint32_t bpc = c->plain_bits - 1;
  const size_t entry_size = (size_t)bpc * c->N / 8;
  size_t ci = 0, n = 0;
  u128 buf = pt[0];
  uint32_t off = bpc;
  while (n < entry_size) {
    if (off >= 8) {
      entry[n++] = (uint8_t)(buf & 0xFF);
      buf >>= 8;
      off -= 8;
    } else {
      ci += 1;
      buf |= (u128)pt[ci] << off;
      off += bpc;
    }
  }
}
