/*
 * onionpir_client.h -- C ABI of the PIR client (CPU).
 *
 * The reference client (src/client.cpp, class PirClient) runs on the CPU and
 * draws its keys and errors from SEAL's PRNGs.  This client keeps the
 * reference's message layout (query packing, Galois elements, RGSW gadgets,
 * decryption after the switch to small_q) and draws every random word as
 * word j of a stream, rand(stream, j), from one of two generators:
 *   secure  (onion_client_create_secure): ChaCha20 under a 256-bit key from
 *           getrandom(); the stream id is the 64-bit nonce and j/8 the block
 *           counter (word j = 32-bit words 2(j%8), 2(j%8)+1 of the block).
 *   seeded  (onion_client_create): a SplitMix64-style counter generator keyed
 *           by a 64-bit seed, so keys and queries are reproducible.  NOT
 *           SECURE: 64 bits of key and no cryptographic generator.  For tests
 *           and benchmarks only.
 * Sampling rules (stream ids carry a kind in their top byte):
 *   secret key  s_i = {0, 1, -1}[floor(r * 3 / 2^64)],  r = rand(1, i)
 *   error       e_i = popcnt(r & M21) - popcnt((r >> 21) & M21),
 *               r = rand(kind_e<<56 | nonce, i)
 *   uniform     a_i = (r_2i * 2^64 + r_2i+1) mod m,
 *               r_j = rand(kind_a<<56 | limb<<48 | nonce, j)
 * Queries use kind_a = 2, kind_e = 3 and the caller's nonce (< 2^48, never
 * repeated by one client: a repeated nonce repeats a and e, and the
 * difference of two such queries reveals both indices).  Key material uses
 * kind_a = 4, kind_e = 5: Galois level a has nonce 0x10000 + a, RGSW key row
 * r nonce 0x20000 + r, so no query nonce reaches a key stream.  A client
 * object is used by one thread at a time.  Formats match
 * include/onionpir_gpu.h.
 */
#ifndef ONIONPIR_CLIENT_H
#define ONIONPIR_CLIENT_H

#include <stddef.h>
#include <stdint.h>

#include "onionpir_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct onion_client onion_client;

const char *onion_client_last_error(void);

/* PirClient::PirClient (src/client.cpp:10-15): randomness from ChaCha20
 * keyed by 256 bits of operating-system entropy */
int onion_client_create_secure(const onion_params *params, uint64_t client_id, onion_client **out);
/* PirClient::PirClient with an explicit seed: reproducible and NOT SECURE
 * (tests and benchmarks) */
int onion_client_create(const onion_params *params, uint64_t client_id, uint64_t seed,
                        onion_client **out);
int onion_client_destroy(onion_client *c);
int onion_client_id(const onion_client *c, uint64_t *id);

/* PirClient::create_galois_keys (src/client.cpp:227-241) */
int onion_client_galois_keys(onion_client *c, uint64_t *out, size_t n_words);
/* PirClient::generate_gsw_from_key (src/client.cpp:17-31) */
int onion_client_rgsw_key(onion_client *c, uint64_t *out, size_t n_words);
/* PirClient::fast_generate_query (src/client.cpp:130-164), plaintext index.
 * ONION_EINVAL for a nonce >= 2^48 or one this client has used before. */
int onion_client_query(onion_client *c, uint64_t pt_index, uint64_t nonce, uint64_t *out_ct);
/* PirClient::load_resp_from_stream (src/client.cpp:364-412); ONION_EINVAL if short */
int onion_client_unpack_response(const onion_client *c, const uint8_t *data, size_t len,
                                 uint64_t *resp);
/* PirClient::decrypt_reply / decrypt_mod_q (src/client.cpp:243-246, 458-479) */
int onion_client_decrypt(const onion_client *c, const uint64_t *resp, uint64_t *out_pt);
/* PirClient::get_entry_from_plaintext (src/client.cpp:255-287) */
int onion_client_entry(const onion_client *c, uint64_t entry_index, const uint64_t *pt,
                       uint8_t *entry, size_t cap, size_t *entry_size);

/* ChaCha20 block function of the secure generator (no reference
 * counterpart; exported for its known-answer test): state words 12-15 are
 * ctr_nonce[0..3] (64-bit block counter, then 64-bit nonce, low words first). */
void onion_chacha20_block(const uint32_t key[8], const uint32_t ctr_nonce[4], uint32_t out[16]);

#ifdef __cplusplus
}
#endif
#endif
