/*
 * onionpir_client.h -- C ABI of the seeded PIR client (CPU).
 *
 * The reference client (src/client.cpp, class PirClient) runs on the CPU and
 * draws its keys and errors from SEAL's PRNGs.  This client keeps the
 * reference's message layout (query packing, Galois elements, RGSW gadgets,
 * decryption after the switch to small_q) and draws all randomness from a
 * counter-based generator keyed by (seed, stream, counter), so keys and
 * queries are reproducible:
 *   secret key  s_i = {0, 1, -1}[floor(r * 3 / 2^64)],  r = rand(seed, 1, i)
 *   error       e_i = popcnt(r & M21) - popcnt((r >> 21) & M21),
 *               r = rand(seed, 3<<56 | nonce, i)
 *   uniform     a_i = (r_2i * 2^64 + r_2i+1) mod m,
 *               r_j = rand(seed, 2<<56 | limb<<48 | nonce, j)
 * Nonces: query = caller's nonce; Galois level a = 0x10000 + a; RGSW key
 * row r = 0x20000 + r.  Formats match include/onionpir_gpu.h.
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

/* PirClient::PirClient (src/client.cpp:10-15) with an explicit id and seed */
int onion_client_create(const onion_params *params, uint64_t client_id, uint64_t seed,
                        onion_client **out);
int onion_client_destroy(onion_client *c);
int onion_client_id(const onion_client *c, uint64_t *id);

/* PirClient::create_galois_keys (src/client.cpp:227-241) */
int onion_client_galois_keys(onion_client *c, uint64_t *out, size_t n_words);
/* PirClient::generate_gsw_from_key (src/client.cpp:17-31) */
int onion_client_rgsw_key(onion_client *c, uint64_t *out, size_t n_words);
/* PirClient::fast_generate_query (src/client.cpp:130-164), plaintext index */
int onion_client_query(onion_client *c, uint64_t pt_index, uint64_t nonce, uint64_t *out_ct);
/* PirClient::load_resp_from_stream (src/client.cpp:364-412); ONION_EINVAL if short */
int onion_client_unpack_response(const onion_client *c, const uint8_t *data, size_t len,
                                 uint64_t *resp);
/* PirClient::decrypt_reply / decrypt_mod_q (src/client.cpp:243-246, 458-479) */
int onion_client_decrypt(const onion_client *c, const uint64_t *resp, uint64_t *out_pt);
/* PirClient::get_entry_from_plaintext (src/client.cpp:255-287) */
int onion_client_entry(const onion_client *c, uint64_t entry_index, const uint64_t *pt,
                       uint8_t *entry, size_t cap, size_t *entry_size);

#ifdef __cplusplus
}
#endif
#endif
