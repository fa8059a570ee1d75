// onionpir.hpp -- C++ API of the MI355X OnionPIR v2 server path, mirroring
// the reference's PirParams / PirServer / PirClient (src/includes/pir.h,
// server.h, client.h) so that callers of the reference switch by changing the
// namespace.  SEAL types are replaced by project-owned ones (Ciphertext,
// Plaintext, Entry); the server path runs on the GPU through the C ABI of
// onionpir_gpu.h, the client on the CPU through onionpir_client.h.
//
// Differences a caller sees:
//   - PirParams is configured at run time (the reference's is compile-time);
//     the default constructor reproduces the reference's active preset
//     (src/includes/database_constants.h:215-223).
//   - PirClient takes an explicit id and seed (the reference draws them from
//     rand() and SEAL's PRNG); the one-argument constructor still draws them.
//   - Missing client keys raise std::out_of_range for the RGSW key too (the
//     reference default-inserts an empty key there, src/server.cpp:543).
//   - push_database_chunk rejects oversized entries (the reference builds the
//     exception object without throwing it, src/server.cpp:644-646).
//   - make_queries answers a batch in one pipelined pass.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iosfwd>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "onionpir_client.h"
#include "onionpir_gpu.h"

namespace onionpir {

typedef std::vector<uint8_t> Entry;  // src/includes/pir.h:173
typedef std::vector<uint64_t> Plaintext;  // N coefficients < t

// Two polynomials (c0 then c1) of N coefficients, coefficient form.
struct Ciphertext {
  std::vector<uint64_t> data;
  Ciphertext() = default;
  explicit Ciphertext(size_t poly_degree) : data(2 * poly_degree, 0) {}
  size_t poly_degree() const { return data.size() / 2; }
  uint64_t *c0() { return data.data(); }
  uint64_t *c1() { return data.data() + poly_degree(); }
  const uint64_t *c0() const { return data.data(); }
  const uint64_t *c1() const { return data.data() + poly_degree(); }
  // Wire form: 8-byte tag, 8-byte word count, then the words (little endian).
  size_t save(std::ostream &out) const;
  void load(std::istream &in);
};

class PirParams {
 public:
  PirParams();  // the reference's active preset: 2^16 entries, MaxFstDimSz 256, PlainMod 16
  PirParams(size_t fst_dim, size_t other_dim, size_t plain_bits = 16, size_t gsw_l = 5,
            size_t gsw_l_key = 10, size_t poly_degree = 2048);
  // PirParams::PirParams shape rule (src/pir.cpp:67-94) for a number of entries
  static PirParams from_entries(size_t num_entries, size_t max_fst_dim, size_t plain_bits = 16,
                                size_t gsw_l = 5, size_t gsw_l_key = 10, size_t poly_degree = 2048);

  size_t get_num_entries_per_plaintext() const;
  size_t get_num_bits_per_coeff() const;
  size_t get_num_bits_per_plaintext() const;
  double get_DBSize_MB() const;
  double get_physical_storage_MB() const;
  size_t get_num_entries() const { return get_num_pt() * get_num_entries_per_plaintext(); }
  size_t get_num_pt() const { return info_.num_pt; }
  size_t get_entry_size() const;
  std::vector<size_t> get_dims() const;
  size_t get_l() const { return p_.gsw_l; }
  size_t get_l_key() const { return p_.gsw_l_key; }
  size_t get_small_q() const { return info_.small_q; }
  size_t get_base_log2() const { return info_.base_log2; }
  size_t get_base_log2_key() const { return info_.base_log2_key; }
  size_t get_fst_dim_sz() const { return p_.fst_dim; }
  size_t get_other_dim_sz() const { return p_.other_dim; }
  size_t get_rns_mod_cnt() const { return 1; }
  size_t get_coeff_val_cnt() const { return p_.poly_degree; }
  uint64_t get_plain_mod() const { return info_.t; }
  uint64_t get_coeff_modulus() const { return info_.q; }
  uint64_t get_special_modulus() const { return info_.p; }
  size_t get_expan_height() const { return info_.expan_height; }
  size_t get_poly_degree() const { return p_.poly_degree; }
  void print_params() const;

  const onion_params &c_params() const { return p_; }
  const onion_param_info &info() const { return info_; }

 private:
  void derive();
  onion_params p_{};
  onion_param_info info_{};
};

class PirServer {
 public:
  explicit PirServer(const PirParams &pir_params);
  ~PirServer();
  PirServer(const PirServer &) = delete;
  PirServer &operator=(const PirServer &) = delete;

  // Seeded entries generated on the device (src/server.cpp:47-78)
  void gen_data();
  void gen_data(uint64_t seed);
  // src/server.cpp:635-701; the chunk is uploaded, lifted and NTT'd at once
  void push_database_chunk(std::vector<Entry> &chunk_entry, size_t chunk_idx);
  // Throws std::logic_error unless every plaintext has been written
  void finalize_database() const;
  // src/server.cpp:520-579
  Ciphertext make_query(size_t client_id, std::stringstream &query_stream);
  // Batched make_query: one pipelined pass for all queries
  std::vector<Ciphertext> make_queries(const std::vector<size_t> &client_ids,
                                       std::vector<std::stringstream> &query_streams);
  // src/server.cpp:581-633
  size_t save_resp_to_stream(const Ciphertext &response, std::stringstream &resp_stream);
  void set_client_galois_key(size_t client_id, std::stringstream &galois_stream);
  void set_client_gsw_key(size_t client_id, std::stringstream &gsw_stream);
  // src/server.cpp:503-517 (testing only: reveals the index)
  Entry direct_get_entry(size_t entry_index) const;

  onion_stage_times last_stage_times() const;
  onion_server *handle() { return srv_; }

 private:
  PirParams params_;
  onion_server *srv_ = nullptr;
  bool seeded_ = false;
  uint64_t seed_ = 0;
  std::map<size_t, Entry> pushed_;  // entries given by push_database_chunk
};

class PirClient {
 public:
  // id from rand(); keys, errors and uniform polynomials from ChaCha20 keyed by
  // 256 bits of operating-system entropy
  explicit PirClient(const PirParams &pir_params);
  // Reproducible client from a 64-bit seed: NOT SECURE, for tests and benchmarks
  PirClient(const PirParams &pir_params, size_t client_id, uint64_t seed);
  ~PirClient();
  PirClient(const PirClient &) = delete;
  PirClient &operator=(const PirClient &) = delete;

  // src/client.cpp:130-164 (entry index; one entry per plaintext)
  Ciphertext fast_generate_query(size_t entry_index);
  static size_t write_query_to_stream(const Ciphertext &query, std::stringstream &data_stream);
  static size_t write_gsw_to_stream(const std::vector<Ciphertext> &gsw, std::stringstream &gsw_stream);
  size_t create_galois_keys(std::stringstream &galois_key_stream);
  std::vector<Ciphertext> generate_gsw_from_key();
  Ciphertext load_resp_from_stream(std::stringstream &resp_stream);
  Plaintext decrypt_reply(const Ciphertext &reply);
  Entry get_entry_from_plaintext(size_t entry_index, const Plaintext &plaintext) const;
  size_t get_client_id() const { return id_; }

 private:
  PirParams params_;
  onion_client *cl_ = nullptr;
  size_t id_ = 0;
  uint64_t nonce_ = 0;
};

}  // namespace onionpir
