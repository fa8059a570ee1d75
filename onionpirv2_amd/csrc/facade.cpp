// facade.cpp -- the C++ API of include/onionpir.hpp over the C ABI.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>

#include "../../include/onionpir.hpp"

namespace onionpir {

namespace {

constexpr uint64_t kTagCt = 0x3154435249504f4fULL;  // "OOPIRCT1"
constexpr uint64_t kTagGk = 0x314b475249504f4fULL;  // "OOPIRGK1"

void throw_gpu(int rc) {
  if (rc == ONION_OK) return;
  const std::string msg = onion_gpu_last_error();
  if (rc == ONION_EINVAL) throw std::invalid_argument(msg);
  if (rc == ONION_ENOKEY) throw std::out_of_range(msg);
  throw std::runtime_error(msg);
}

void throw_client(int rc) {
  if (rc == ONION_OK) return;
  const std::string msg = onion_client_last_error();
  if (rc == ONION_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

void put_u64(std::ostream &out, uint64_t v) { out.write(reinterpret_cast<const char *>(&v), 8); }

uint64_t get_u64(std::istream &in) {
  uint64_t v = 0;
  in.read(reinterpret_cast<char *>(&v), 8);
  if (in.gcount() != 8) throw std::runtime_error("unexpected end of stream");
  return v;
}

std::vector<uint64_t> read_words(std::istream &in, uint64_t tag) {
  if (get_u64(in) != tag) throw std::invalid_argument("stream does not hold the expected object");
  const uint64_t n = get_u64(in);
  if (n > (uint64_t(1) << 28)) throw std::invalid_argument("object too large");
  std::vector<uint64_t> w(n);
  in.read(reinterpret_cast<char *>(w.data()), (std::streamsize)(n * 8));
  if ((uint64_t)in.gcount() != n * 8) throw std::runtime_error("unexpected end of stream");
  return w;
}

size_t write_words(std::ostream &out, uint64_t tag, const uint64_t *w, size_t n) {
  put_u64(out, tag);
  put_u64(out, n);
  out.write(reinterpret_cast<const char *>(w), (std::streamsize)(n * 8));
  return 16 + n * 8;
}

uint64_t next_pow2(uint64_t n) {
  uint64_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

// counter-based generator of the seeded database (same rule as the device)
uint64_t mix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xbf58476d1ce4e5b9ULL;
  z ^= z >> 27;
  z *= 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return z;
}

}  // namespace

// ---- Ciphertext ---------------------------------------------------------------------

size_t Ciphertext::save(std::ostream &out) const { return write_words(out, kTagCt, data.data(), data.size()); }

void Ciphertext::load(std::istream &in) { data = read_words(in, kTagCt); }

// ---- PirParams ----------------------------------------------------------------------

PirParams::PirParams() : PirParams(256, 256, 16, 5, 10, 2048) {}

PirParams::PirParams(size_t fst_dim, size_t other_dim, size_t plain_bits, size_t gsw_l, size_t gsw_l_key,
                     size_t poly_degree) {
  p_.poly_degree = (uint32_t)poly_degree;
  p_.plain_bits = (uint32_t)plain_bits;
  p_.gsw_l = (uint32_t)gsw_l;
  p_.gsw_l_key = (uint32_t)gsw_l_key;
  p_.fst_dim = (uint32_t)fst_dim;
  p_.other_dim = (uint32_t)other_dim;
  derive();
}

PirParams PirParams::from_entries(size_t num_entries, size_t max_fst_dim, size_t plain_bits, size_t gsw_l,
                                  size_t gsw_l_key, size_t poly_degree) {
  // one entry per plaintext (EntrySize = 0); other = ceil(nextpow2(n) / max_fst),
  // fst = ceil(n / other)
  if (num_entries == 0 || max_fst_dim == 0) throw std::invalid_argument("empty database");
  const uint64_t other = (next_pow2(num_entries) + max_fst_dim - 1) / max_fst_dim;
  const uint64_t fst = (num_entries + other - 1) / other;
  return PirParams(fst, other, plain_bits, gsw_l, gsw_l_key, poly_degree);
}

void PirParams::derive() { throw_gpu(onion_params_derive(&p_, &info_)); }

size_t PirParams::get_num_bits_per_coeff() const { return p_.plain_bits - 1; }
size_t PirParams::get_num_bits_per_plaintext() const { return get_num_bits_per_coeff() * p_.poly_degree; }
size_t PirParams::get_entry_size() const { return get_num_bits_per_plaintext() / 8; }
size_t PirParams::get_num_entries_per_plaintext() const {
  return get_num_bits_per_plaintext() / (get_entry_size() * 8);
}
double PirParams::get_DBSize_MB() const { return (double)get_num_entries() * get_entry_size() / 1024 / 1024; }
double PirParams::get_physical_storage_MB() const {
  return (double)get_coeff_val_cnt() * get_num_pt() * 8 / 1024 / 1024;
}
std::vector<size_t> PirParams::get_dims() const {
  std::vector<size_t> d{p_.fst_dim};
  for (uint32_t i = 1; i < info_.ndim; ++i) d.push_back(2);
  return d;
}

void PirParams::print_params() const {
  std::printf("==============================================================\n");
  std::printf("                       PIR PARAMETERS                         \n");
  std::printf("==============================================================\n");
  std::printf("  Database size (MB)              = %.2f\n", get_DBSize_MB());
  std::printf("  Physical storage (MB)           = %.2f\n", get_physical_storage_MB());
  std::printf("  entry_size_                     = %zu B\n", get_entry_size());
  std::printf("  num_pt_                         = %zu\n", get_num_pt());
  std::printf("  l_ / l_key_                     = %zu / %zu\n", get_l(), get_l_key());
  std::printf("  base_log2_ / base_log2_key_     = %zu / %zu\n", get_base_log2(), get_base_log2_key());
  std::printf("  dimensions_                     = [ %u x 2^%u ]\n", p_.fst_dim, info_.ndim - 1);
  std::printf("  poly_modulus_degree             = %u\n", p_.poly_degree);
  std::printf("  coeff_modulus (q, special p)    = %llu, %llu\n", (unsigned long long)info_.q,
              (unsigned long long)info_.p);
  std::printf("  plain_modulus t / small_q       = %llu / %llu\n", (unsigned long long)info_.t,
              (unsigned long long)info_.small_q);
  std::printf("==============================================================\n");
}

// ---- PirServer -------------------------------------------------------------------------

PirServer::PirServer(const PirParams &pir_params) : params_(pir_params) {
  throw_gpu(onion_gpu_create(&params_.c_params(), &srv_));
}

PirServer::~PirServer() { onion_gpu_destroy(srv_); }

void PirServer::gen_data() {
  std::random_device rd;
  gen_data(((uint64_t)rd() << 32) | rd());
}

void PirServer::gen_data(uint64_t seed) {
  throw_gpu(onion_gpu_db_generate(srv_, seed, 1));
  seeded_ = true;
  seed_ = seed;
  pushed_.clear();
}

void PirServer::push_database_chunk(std::vector<Entry> &chunk_entry, size_t chunk_idx) {
  const size_t esize = params_.get_entry_size(), bpc = params_.get_num_bits_per_coeff();
  const size_t N = params_.get_poly_degree(), fst = params_.get_fst_dim_sz();
  if (chunk_entry.size() > fst) throw std::invalid_argument("chunk holds more than fst_dim entries");
  std::vector<uint64_t> coeffs(chunk_entry.size() * N, 0);
  for (size_t i = 0; i < chunk_entry.size(); ++i) {
    Entry &e = chunk_entry[i];
    if (e.size() > esize) throw std::invalid_argument("Entry size is too large");
    if (!e.empty()) e.resize(esize, 0);
    // pack bpc bits per coefficient, LSB first
    unsigned __int128 buf = 0;
    size_t off = 0, idx = 0;
    for (uint8_t b : e) {
      buf |= (unsigned __int128)b << off;
      off += 8;
      while (off >= bpc) {
        coeffs[i * N + idx++] = (uint64_t)(buf & ((uint64_t(1) << bpc) - 1));
        buf >>= bpc;
        off -= bpc;
      }
    }
    if (off > 0 && idx < N) coeffs[i * N + idx] = (uint64_t)(buf & ((uint64_t(1) << bpc) - 1));
    pushed_[chunk_idx * fst + i] = e;
  }
  throw_gpu(onion_gpu_db_upload_plain(srv_, chunk_idx * fst, chunk_entry.size(), coeffs.data()));
}

void PirServer::finalize_database() const {
  // a zero-length answer batch checks that the database is complete
  const int rc = onion_gpu_answer_batch(srv_, 0, nullptr, nullptr, nullptr);
  if (rc == ONION_ESTATE) throw std::logic_error(onion_gpu_last_error());
  throw_gpu(rc);
}

Ciphertext PirServer::make_query(size_t client_id, std::stringstream &query_stream) {
  Ciphertext q;
  q.load(query_stream);
  if (q.poly_degree() != params_.get_poly_degree()) throw std::invalid_argument("query has the wrong size");
  Ciphertext r(params_.get_poly_degree());
  throw_gpu(onion_gpu_answer(srv_, client_id, q.data.data(), r.data.data()));
  return r;
}

std::vector<Ciphertext> PirServer::make_queries(const std::vector<size_t> &client_ids,
                                                std::vector<std::stringstream> &query_streams) {
  if (client_ids.size() != query_streams.size()) throw std::invalid_argument("ids and streams differ in count");
  const size_t n = client_ids.size(), N = params_.get_poly_degree();
  std::vector<uint64_t> ids(client_ids.begin(), client_ids.end()), qs(n * 2 * N), rs(n * 2 * N);
  for (size_t i = 0; i < n; ++i) {
    Ciphertext q;
    q.load(query_streams[i]);
    if (q.poly_degree() != N) throw std::invalid_argument("query has the wrong size");
    std::memcpy(qs.data() + i * 2 * N, q.data.data(), 2 * N * 8);
  }
  throw_gpu(onion_gpu_answer_batch(srv_, (uint32_t)n, ids.data(), qs.data(), rs.data()));
  std::vector<Ciphertext> out(n, Ciphertext(N));
  for (size_t i = 0; i < n; ++i) std::memcpy(out[i].data.data(), rs.data() + i * 2 * N, 2 * N * 8);
  return out;
}

size_t PirServer::save_resp_to_stream(const Ciphertext &response, std::stringstream &resp_stream) {
  std::vector<uint8_t> buf(params_.info().response_bytes);
  size_t n = 0;
  throw_gpu(onion_gpu_pack_response(srv_, response.data.data(), buf.data(), buf.size(), &n));
  resp_stream.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)n);
  return n;
}

void PirServer::set_client_galois_key(size_t client_id, std::stringstream &galois_stream) {
  const std::vector<uint64_t> w = read_words(galois_stream, kTagGk);
  throw_gpu(onion_gpu_set_galois_key(srv_, client_id, w.data(), w.size()));
}

void PirServer::set_client_gsw_key(size_t client_id, std::stringstream &gsw_stream) {
  const size_t rows = 2 * params_.get_l_key(), N = params_.get_poly_degree();
  std::vector<uint64_t> flat;
  flat.reserve(rows * 2 * N);
  for (size_t r = 0; r < rows; ++r) {
    Ciphertext ct;
    ct.load(gsw_stream);
    if (ct.poly_degree() != N) throw std::invalid_argument("GSW row has the wrong size");
    flat.insert(flat.end(), ct.data.begin(), ct.data.end());
  }
  throw_gpu(onion_gpu_set_rgsw_key(srv_, client_id, flat.data(), flat.size()));
}

Entry PirServer::direct_get_entry(size_t entry_index) const {
  auto it = pushed_.find(entry_index);
  if (it != pushed_.end()) return it->second;
  if (!seeded_ || entry_index >= params_.get_num_entries())
    throw std::invalid_argument("entry was never stored");
  // regenerate the seeded entry: big-endian index, then generator bytes
  const size_t esize = params_.get_entry_size(), N = params_.get_poly_degree();
  const uint64_t key = mix64(seed_ * 0x9e3779b97f4a7c15ULL + mix64(0x4442ULL ^ 0x632be59bd9b4e019ULL));
  Entry e(esize);
  for (size_t b = 0; b < esize; ++b) {
    if (b < 8) {
      e[b] = (uint8_t)(entry_index >> (8 * (7 - b)));
    } else {
      const uint64_t r = mix64(key + (entry_index * N + b / 8 + 1) * 0x9e3779b97f4a7c15ULL);
      e[b] = (uint8_t)(r >> (8 * (b % 8)));
    }
  }
  return e;
}

onion_stage_times PirServer::last_stage_times() const {
  onion_stage_times t{};
  throw_gpu(onion_gpu_last_stage_times(srv_, &t));
  return t;
}

// ---- PirClient -----------------------------------------------------------------------

PirClient::PirClient(const PirParams &pir_params) : params_(pir_params), id_((size_t)std::rand()) {
  throw_client(onion_client_create_secure(&params_.c_params(), id_, &cl_));
}

PirClient::PirClient(const PirParams &pir_params, size_t client_id, uint64_t seed)
    : params_(pir_params), id_(client_id) {
  throw_client(onion_client_create(&params_.c_params(), client_id, seed, &cl_));
}

PirClient::~PirClient() { onion_client_destroy(cl_); }

Ciphertext PirClient::fast_generate_query(size_t entry_index) {
  const size_t pt = entry_index / params_.get_num_entries_per_plaintext();
  Ciphertext q(params_.get_poly_degree());
  throw_client(onion_client_query(cl_, pt, nonce_++, q.data.data()));
  return q;
}

size_t PirClient::write_query_to_stream(const Ciphertext &query, std::stringstream &data_stream) {
  return query.save(data_stream);
}

size_t PirClient::write_gsw_to_stream(const std::vector<Ciphertext> &gsw, std::stringstream &gsw_stream) {
  size_t total = 0;
  for (const Ciphertext &ct : gsw) total += ct.save(gsw_stream);
  return total;
}

size_t PirClient::create_galois_keys(std::stringstream &galois_key_stream) {
  std::vector<uint64_t> w(params_.info().galois_key_words);
  throw_client(onion_client_galois_keys(cl_, w.data(), w.size()));
  return write_words(galois_key_stream, kTagGk, w.data(), w.size());
}

std::vector<Ciphertext> PirClient::generate_gsw_from_key() {
  const size_t N = params_.get_poly_degree(), rows = 2 * params_.get_l_key();
  std::vector<uint64_t> w(params_.info().rgsw_key_words);
  throw_client(onion_client_rgsw_key(cl_, w.data(), w.size()));
  std::vector<Ciphertext> out(rows, Ciphertext(N));
  for (size_t r = 0; r < rows; ++r) std::memcpy(out[r].data.data(), w.data() + r * 2 * N, 2 * N * 8);
  return out;
}

Ciphertext PirClient::load_resp_from_stream(std::stringstream &resp_stream) {
  const size_t n = params_.info().response_bytes;
  std::vector<uint8_t> buf(n);
  resp_stream.read(reinterpret_cast<char *>(buf.data()), (std::streamsize)n);
  const size_t got = (size_t)resp_stream.gcount();
  Ciphertext r(params_.get_poly_degree());
  const int rc = onion_client_unpack_response(cl_, buf.data(), got, r.data.data());
  if (rc != ONION_OK) throw std::runtime_error(onion_client_last_error());
  return r;
}

Plaintext PirClient::decrypt_reply(const Ciphertext &reply) {
  Plaintext pt(params_.get_poly_degree());
  throw_client(onion_client_decrypt(cl_, reply.data.data(), pt.data()));
  return pt;
}

Entry PirClient::get_entry_from_plaintext(size_t entry_index, const Plaintext &plaintext) const {
  Entry e(params_.get_entry_size());
  size_t n = 0;
  throw_client(onion_client_entry(cl_, entry_index, plaintext.data(), e.data(), e.size(), &n));
  e.resize(n);
  return e;
}

}  // namespace onionpir
