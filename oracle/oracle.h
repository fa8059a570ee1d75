/*
 * oracle.h -- CPU restatement of the OnionPIRv2 server answer path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product (onionpirv2_amd/) never links, loads or calls it.
 *
 * Every function restates one piece of the reference (paths relative to the
 * reference repository root) or of the SEAL 4.1 routine the reference calls
 * at that line.  SEAL itself is not available, so its routines are restated
 * from their published algorithms; which of them are pinned by which test is
 * listed in tests/test_oracle.py.
 *
 * Conventions
 *   - N-coefficient polynomials are arrays of N uint64.
 *   - A ciphertext ("ct") is 2*N uint64: c0 then c1 (one data limb, rns=1).
 *   - Galois keys: h levels x 2 components x 2 limbs (q then p) x N, NTT form.
 *   - RGSW rows: row-major [row][c0|c1][N].
 *   - The database is kept in the reference's slot-major layout
 *     db[slot * num_pt + row * fst + col] (src/server.cpp:722-748).
 */
#ifndef ONIONPIR_ORACLE_H
#define ONIONPIR_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

/* Zeroed allocation that stops the process when memory runs out (the oracle
 * is test infrastructure: a failed allocation is a fault of the test host). */
static inline void *orc_alloc(size_t n) {
  void *p = calloc(n ? n : 1, 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
    abort();
  }
  return p;
}

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned __int128 orc_u128;

typedef struct {
  uint64_t value;
  uint64_t ratio_lo, ratio_hi; /* floor(2^128 / value), two words */
} orc_mod;

typedef struct {
  uint32_t N, logN;
  orc_mod mod;
  uint64_t psi;                 /* minimal primitive 2N-th root of unity */
  uint64_t *rp, *rp_shoup;      /* rp[k] = psi^bitrev(k)                 */
  uint64_t *irp, *irp_shoup;    /* irp[k] = psi^-bitrev(k)               */
  uint64_t n_inv, n_inv_shoup;
} orc_ntt_tables;

typedef struct {
  /* ring and moduli */
  uint32_t N, logN;
  uint64_t q, p, t, small_q;
  uint32_t plain_bits;
  uint32_t l, base_log2, l_key, base_log2_key;
  /* database shape: dims = [fst, 2, 2, ...] */
  uint32_t fst, other, ndim, h, u;
  uint64_t num_pt;
  orc_mod mq, mp, msq;
  orc_ntt_tables ntt_q, ntt_p, ntt_sq;
} orc_ctx;

/* per-stage wall-clock seconds of one orc_answer call; names follow the
 * reference's TimerLogger sections (src/includes/logging.h:27-65) */
typedef struct {
  double expand, convert, first_dim, first_dim_core, other_dim, mod_switch, total;
} orc_times;

/* ---- parameters and arithmetic ------------------------------------------------ */
orc_ctx *orc_create(uint32_t N, uint32_t plain_bits, uint32_t l, uint32_t l_key,
                    uint32_t fst, uint32_t other);
void orc_destroy(orc_ctx *c);
int orc_is_prime(uint64_t n);
size_t orc_get_primes(uint64_t factor, uint32_t bits, size_t count, uint64_t *out);
uint64_t orc_generate_prime(uint32_t bit_width);
void orc_shape_from_entries(uint64_t num_entries, uint64_t entries_per_pt, uint64_t max_fst,
                            uint64_t *fst, uint64_t *other);
uint64_t orc_barrett128(const orc_mod *m, uint64_t lo, uint64_t hi);
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t m);
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t m);
uint64_t orc_invmod(uint64_t a, uint64_t m);
void orc_gadget(uint32_t l, uint32_t base_log2, uint64_t q, uint64_t *out);
uint64_t orc_bit_reverse(uint64_t x, uint32_t logn);
uint64_t orc_rand64(uint64_t seed, uint64_t stream, uint64_t ctr);

/* ---- NTT ------------------------------------------------------------------------ */
/* which: 0 = q, 1 = p, 2 = small_q */
void orc_ntt_fwd(const orc_ctx *c, int which, uint64_t *a);
void orc_ntt_inv(const orc_ctx *c, int which, uint64_t *a);
void orc_ntt_fwd_naive(const orc_ctx *c, int which, const uint64_t *a, uint64_t *out);
uint64_t orc_ntt_psi(const orc_ctx *c, int which);

/* ---- HE primitives ---------------------------------------------------------------- */
void orc_galois_coeff(const orc_ctx *c, const uint64_t *in, uint32_t g, uint64_t m, uint64_t *out);
void orc_shift_neg(const orc_ctx *c, const uint64_t *in, uint64_t k, uint64_t *out);
void orc_keyswitch_add(const orc_ctx *c, const uint64_t *target, const uint64_t *key, uint64_t *ct);
void orc_apply_galois(const orc_ctx *c, const uint64_t *ct, uint32_t g, const uint64_t *key,
                      uint64_t *out);
void orc_decompose(const orc_ctx *c, const uint64_t *ct, uint32_t l, uint32_t base_log2,
                   uint64_t *digits);
void orc_external_product(const orc_ctx *c, const uint64_t *rgsw, uint32_t l, uint32_t base_log2,
                          const uint64_t *ct, uint64_t *out_ntt);

/* ---- server stages ---------------------------------------------------------------- */
void orc_expand(const orc_ctx *c, const uint64_t *query, const uint64_t *galois_keys,
                uint64_t *leaves);
void orc_convert(const orc_ctx *c, const uint64_t *leaves, const uint64_t *rgsw_key_ntt,
                 uint64_t *rgsw_out);
void orc_rgsw_key_to_ntt(const orc_ctx *c, const uint64_t *rgsw_key_coeff, uint64_t *out);
void orc_first_dim_raw(const orc_ctx *c, const uint64_t *db, const uint64_t *sel_ntt,
                       orc_u128 *inter);
void orc_first_dim(const orc_ctx *c, const uint64_t *db, const uint64_t *sel_coeff,
                   uint64_t *out, orc_times *tm);
void orc_mux_level(const orc_ctx *c, uint64_t *cts, uint64_t n_cts, const uint64_t *rgsw);
void orc_mod_switch(const orc_ctx *c, uint64_t *ct);
uint64_t orc_mod_switch_one(uint64_t x, uint64_t q_big, uint64_t q_small);
uint64_t orc_mod_switch_one_int(uint64_t x, uint64_t q_big, uint64_t q_small);
size_t orc_pack_response(const orc_ctx *c, const uint64_t *ct, uint8_t *out);
int orc_unpack_response(const orc_ctx *c, const uint8_t *in, size_t len, uint64_t *ct);
int orc_answer(const orc_ctx *c, const uint64_t *db, const uint64_t *galois_keys,
               const uint64_t *rgsw_key_ntt, const uint64_t *query, uint64_t *resp,
               orc_times *tm);

/* ---- database ---------------------------------------------------------------------- */
void orc_db_plaintext(const orc_ctx *c, uint64_t seed, int mode, uint64_t pt, uint64_t *coeffs);
void orc_db_build(const orc_ctx *c, uint64_t seed, int mode, uint64_t *db);
void orc_db_build_range(const orc_ctx *c, uint64_t seed, int mode, uint64_t pt_lo, uint64_t pt_hi,
                        uint64_t *db);
void orc_plain_to_ntt(const orc_ctx *c, const uint64_t *coeffs, uint64_t *out);

/* ---- seeded test client ------------------------------------------------------------- */
void orc_client_keygen(const orc_ctx *c, uint64_t seed, uint64_t *sk);
void orc_client_encrypt_zero(const orc_ctx *c, uint64_t seed, uint64_t nonce, const uint64_t *sk,
                             uint64_t *ct);
void orc_client_encrypt(const orc_ctx *c, uint64_t seed, uint64_t nonce, const uint64_t *sk,
                        const uint64_t *msg, uint64_t *ct);
void orc_client_galois_keys(const orc_ctx *c, uint64_t seed, const uint64_t *sk, uint64_t *gk);
void orc_client_rgsw_key(const orc_ctx *c, uint64_t seed, const uint64_t *sk, uint64_t *rk);
void orc_client_rgsw(const orc_ctx *c, uint64_t seed, uint64_t nonce, const uint64_t *sk,
                     uint64_t bit, uint64_t *rgsw_ntt);
void orc_client_query_indices(const orc_ctx *c, uint64_t pt_index, uint64_t *idx);
void orc_client_query(const orc_ctx *c, uint64_t seed, uint64_t nonce, const uint64_t *sk,
                      uint64_t pt_index, uint64_t *ct);
void orc_client_phase(const orc_ctx *c, const uint64_t *sk, const uint64_t *ct, uint64_t *phase);
void orc_client_decrypt(const orc_ctx *c, const uint64_t *sk, const uint64_t *ct, uint64_t *pt);
void orc_client_decrypt_small(const orc_ctx *c, const uint64_t *sk, const uint64_t *resp,
                              uint64_t *pt);
void orc_entry_from_plaintext(const orc_ctx *c, const uint64_t *pt, uint8_t *entry);

#ifdef __cplusplus
}
#endif
#endif
