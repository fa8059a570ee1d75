// client.cpp -- seeded PIR client (CPU), the C ABI of include/onionpir_client.h.
//
// Restates the reference client's message construction (src/client.cpp):
// Galois keys for elements 1 + (N >> a) (create_galois_keys, :227-241), the
// RGSW(s) key with the key gadget (generate_gsw_from_key, :17-31 over
// GSWEval::plain_to_half_gsw, src/gsw_eval.cpp:339-405), the packed query
// (fast_generate_query + add_gsw_to_query, :130-209), decryption after the
// switch to small_q (decrypt_mod_q + sk_mod_switch, :458-512), response
// unpacking (:364-412) and entry extraction (:255-287).  Randomness comes from
// one of the two generators described in the header: ChaCha20 keyed from the
// operating system (onion_client_create_secure), or the seeded counter-based
// generator (onion_client_create; reproducible, not secure).
#include <sys/random.h>

#include <cerrno>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/onionpir_client.h"
#include "arith.hpp"
#include "host_math.hpp"
#include "params.hpp"

using namespace onion;
using host::u128;

namespace {

thread_local std::string g_cerr;

int cfail(int code, const std::string &m) {
  g_cerr = m;
  return code;
}

// Stream kinds (top byte of a stream id).  Key material draws from kinds of
// its own, so no query nonce can reach a key's uniform or error stream.
constexpr uint64_t kSk = 1, kUni = 2, kErr = 3, kKeyUni = 4, kKeyErr = 5;
constexpr uint64_t kNonceLimit = uint64_t(1) << 48;  // nonce bits of a stream id

// ChaCha20 block function (D. J. Bernstein, "ChaCha, a variant of Salsa20",
// 2008; the quarter round and state layout of RFC 8439 section 2.3) with a
// 64-bit block counter in words 12-13 and a 64-bit nonce in words 14-15.
inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
inline void quarter(uint32_t *x, int a, int b, int c, int d) {
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16);
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12);
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);
}

}  // namespace

extern "C" void onion_chacha20_block(const uint32_t key[8], const uint32_t ctr_nonce[4], uint32_t out[16]) {
  uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
  for (int i = 0; i < 8; ++i) in[4 + i] = key[i];
  for (int i = 0; i < 4; ++i) in[12 + i] = ctr_nonce[i];
  uint32_t x[16];
  memcpy(x, in, sizeof x);
  for (int r = 0; r < 10; ++r) {
    quarter(x, 0, 4, 8, 12); quarter(x, 1, 5, 9, 13); quarter(x, 2, 6, 10, 14); quarter(x, 3, 7, 11, 15);
    quarter(x, 0, 5, 10, 15); quarter(x, 1, 6, 11, 12); quarter(x, 2, 7, 8, 13); quarter(x, 3, 4, 9, 14);
  }
  for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

namespace {

inline uint64_t addm(uint64_t a, uint64_t b, uint64_t m) {
  uint64_t s = a + b;
  return s >= m ? s - m : s;
}

// Harvey/Shoup negacyclic NTT on the host (bit-reversed output order).
struct HostNtt {
  host::Tables T;
  uint32_t N = 0;
  void fwd(uint64_t *a) const {
    const uint64_t q = T.q, two_q = 2 * q;
    for (uint32_t groups = 1, half = N >> 1; groups < N; groups <<= 1, half >>= 1)
      for (uint32_t i = 0; i < groups; ++i) {
        const uint64_t w = T.rp[groups + i], ws = T.rps[groups + i];
        uint64_t *x = a + 2 * (size_t)i * half, *y = x + half;
        for (uint32_t j = 0; j < half; ++j) {
          uint64_t u = x[j] >= two_q ? x[j] - two_q : x[j];
          uint64_t v = y[j] * w - (uint64_t)(((u128)y[j] * ws) >> 64) * q;
          x[j] = u + v;
          y[j] = u + two_q - v;
        }
      }
    for (uint32_t j = 0; j < N; ++j) {
      uint64_t v = a[j];
      if (v >= two_q) v -= two_q;
      if (v >= q) v -= q;
      a[j] = v;
    }
  }
  void inv(uint64_t *a) const {
    const uint64_t q = T.q, two_q = 2 * q;
    for (uint32_t span = N, half = 1; span > 1; span >>= 1, half <<= 1) {
      const uint32_t groups = span >> 1;
      for (uint32_t i = 0; i < groups; ++i) {
        const uint64_t w = T.irp[groups + i], ws = T.irps[groups + i];
        uint64_t *x = a + 2 * (size_t)i * half, *y = x + half;
        for (uint32_t j = 0; j < half; ++j) {
          uint64_t u = x[j], v = y[j], s = u + v;
          x[j] = s >= two_q ? s - two_q : s;
          uint64_t d = u + two_q - v;
          y[j] = d * w - (uint64_t)(((u128)d * ws) >> 64) * q;
        }
      }
    }
    for (uint32_t j = 0; j < N; ++j) {
      uint64_t v = a[j] * T.n_inv - (uint64_t)(((u128)a[j] * T.n_inv_s) >> 64) * q;
      a[j] = v >= q ? v - q : v;
    }
  }
  // negacyclic a*b mod q
  void mul(const uint64_t *a, const uint64_t *b, uint64_t *out) const {
    std::vector<uint64_t> x(a, a + N), y(b, b + N);
    fwd(x.data());
    fwd(y.data());
    for (uint32_t i = 0; i < N; ++i) out[i] = host::mulmod(x[i], y[i], T.q);
    inv(out);
  }
};

}  // namespace

struct onion_client {
  onion_params P{};
  onion_param_info I{};
  uint64_t id = 0, seed = 0;
  bool secure = false;
  uint32_t chacha_key[8] = {};
  uint32_t N = 0;
  HostNtt nq, np, nsq;  // q, p, small_q
  std::vector<uint64_t> sk;  // residues mod q: {0, 1, q-1}
  std::unordered_set<uint64_t> used_nonces;  // query nonces already spent
};

namespace {

// n words of the stream `stream`: ChaCha20 blocks 0, 1, ... with the stream id
// as nonce (secure), or rand64_keyed(stream_key(seed, stream), i) (seeded).
void fill_rand(const onion_client *c, uint64_t stream, uint32_t n, uint64_t *out) {
  if (!c->secure) {
    const uint64_t key = stream_key(c->seed, stream);
    for (uint32_t i = 0; i < n; ++i) out[i] = rand64_keyed(key, i);
    return;
  }
  uint32_t blk[16];
  for (uint32_t i = 0; i < n; i += 8) {
    const uint64_t ctr = i / 8;
    const uint32_t cn[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
    onion_chacha20_block(c->chacha_key, cn, blk);
    for (uint32_t k = 0; k < 8 && i + k < n; ++k) out[i + k] = blk[2 * k] | ((uint64_t)blk[2 * k + 1] << 32);
  }
}

void sample_uniform(const onion_client *c, uint64_t kind, uint64_t nonce, uint64_t limb, uint64_t m,
                    uint64_t *out) {
  std::vector<uint64_t> r(2 * (size_t)c->N);
  fill_rand(c, (kind << 56) | (limb << 48) | nonce, 2 * c->N, r.data());
  for (uint32_t i = 0; i < c->N; ++i) out[i] = (uint64_t)((((u128)r[2 * i] << 64) | r[2 * i + 1]) % m);
}

void sample_error(const onion_client *c, uint64_t kind, uint64_t nonce, int64_t *out) {
  const uint64_t m21 = (1ULL << 21) - 1;
  std::vector<uint64_t> r(c->N);
  fill_rand(c, (kind << 56) | nonce, c->N, r.data());
  for (uint32_t i = 0; i < c->N; ++i)
    out[i] = (int64_t)__builtin_popcountll(r[i] & m21) - (int64_t)__builtin_popcountll((r[i] >> 21) & m21);
}

inline uint64_t to_mod(int64_t v, uint64_t m) {
  return v >= 0 ? (uint64_t)v % m : m - ((uint64_t)(-v) % m);
}

// (c0, c1) = (-a*s + e, a) mod the modulus of `ntt` (limb index `limb`); a
// from the uniform stream of kind `uni`
void enc_zero(const onion_client *c, const HostNtt &ntt, uint64_t uni, uint64_t limb, uint64_t nonce,
              const int64_t *e, uint64_t *c0, uint64_t *c1) {
  const uint32_t N = c->N;
  const uint64_t m = ntt.T.q;
  std::vector<uint64_t> s(N), as(N);
  for (uint32_t i = 0; i < N; ++i) s[i] = c->sk[i] > 1 ? m - 1 : c->sk[i];
  sample_uniform(c, uni, nonce, limb, m, c1);
  ntt.mul(c1, s.data(), as.data());
  for (uint32_t i = 0; i < N; ++i) c0[i] = addm(as[i] ? m - as[i] : 0, to_mod(e[i], m), m);
}

// encryption of zero mod q: key rows (key = true) or queries
void enc_zero_q(const onion_client *c, bool key, uint64_t nonce, uint64_t *ct) {
  std::vector<int64_t> e(c->N);
  sample_error(c, key ? kKeyErr : kErr, nonce, e.data());
  enc_zero(c, c->nq, key ? kKeyUni : kUni, 0, nonce, e.data(), ct, ct + c->N);
}

void gadget(uint32_t l, uint32_t base_log2, uint64_t q, uint64_t *out) {
  u128 pw = 1;
  for (int j = (int)l - 1; j >= 0; --j) {
    out[j] = (uint64_t)pw;
    pw = (pw << base_log2) % q;
  }
}

// sigma_g on coefficients mod m (SEAL GaloisTool::apply_galois)
void galois(const uint64_t *in, uint32_t N, uint64_t g, uint64_t m, uint64_t *out) {
  uint64_t raw = 0;
  const uint32_t logn = ceil_log2_u64(N);
  for (uint32_t i = 0; i < N; ++i, raw += g) {
    uint64_t v = in[i];
    if ((raw >> logn) & 1) v = v ? m - v : 0;
    out[raw & (N - 1)] = v;
  }
}

}  // namespace

extern "C" {

const char *onion_client_last_error(void) { return g_cerr.c_str(); }

namespace {

int create_client(const onion_params *params, uint64_t client_id, uint64_t seed, bool secure,
                  onion_client **out) {
  if (!params || !out) return cfail(ONION_EINVAL, "null argument");
  std::string err;
  onion_param_info I;
  if (derive_params(*params, I, err) != ONION_OK) return cfail(ONION_EINVAL, err);
  onion_client *c = new onion_client();
  c->P = *params;
  c->I = I;
  c->id = client_id;
  c->seed = seed;
  c->secure = secure;
  if (secure) {
    // 256 bits from the operating system's generator
    uint8_t *k = reinterpret_cast<uint8_t *>(c->chacha_key);
    size_t got = 0;
    while (got < sizeof c->chacha_key) {
      const ssize_t r = getrandom(k + got, sizeof c->chacha_key - got, 0);
      if (r < 0 && errno == EINTR) continue;
      if (r <= 0) {
        delete c;
        return cfail(ONION_EINVAL, std::string("getrandom failed: ") + strerror(errno));
      }
      got += (size_t)r;
    }
  }
  c->N = params->poly_degree;
  c->nq.T = host::make_tables(I.q, c->N);
  c->np.T = host::make_tables(I.p, c->N);
  c->nsq.T = host::make_tables(I.small_q, c->N);
  c->nq.N = c->np.N = c->nsq.N = c->N;
  // KeyGenerator::secret_key: ternary, stored as residues mod q
  c->sk.resize(c->N);
  std::vector<uint64_t> r(c->N);
  fill_rand(c, kSk, c->N, r.data());
  for (uint32_t i = 0; i < c->N; ++i) {
    const uint64_t v = (uint64_t)(((u128)r[i] * 3) >> 64);
    c->sk[i] = v == 2 ? I.q - 1 : v;
  }
  *out = c;
  return ONION_OK;
}

}  // namespace

int onion_client_create(const onion_params *params, uint64_t client_id, uint64_t seed, onion_client **out) {
  return create_client(params, client_id, seed, false, out);
}

int onion_client_create_secure(const onion_params *params, uint64_t client_id, onion_client **out) {
  return create_client(params, client_id, 0, true, out);
}

int onion_client_destroy(onion_client *c) {
  delete c;
  return ONION_OK;
}

int onion_client_id(const onion_client *c, uint64_t *id) {
  if (!c || !id) return cfail(ONION_EINVAL, "null argument");
  *id = c->id;
  return ONION_OK;
}

int onion_client_galois_keys(onion_client *c, uint64_t *out, size_t n) {
  if (!c || !out) return cfail(ONION_EINVAL, "null argument");
  if (n != c->I.galois_key_words) return cfail(ONION_EINVAL, "wrong Galois key buffer size");
  const uint32_t N = c->N;
  const uint64_t q = c->I.q, p = c->I.p, p_mod_q = p % q;
  std::vector<uint64_t> gs(N), c0(N), c1(N);
  std::vector<int64_t> e(N);
  for (uint32_t a = 0; a < c->I.expan_height; ++a) {
    const uint64_t g = 1 + (N >> a), nonce = 0x10000 + a;
    uint64_t *key = out + (size_t)a * 4 * N;
    galois(c->sk.data(), N, g, q, gs.data());
    sample_error(c, kKeyErr, nonce, e.data());
    for (uint64_t limb = 0; limb < 2; ++limb) {
      const HostNtt &ntt = limb == 0 ? c->nq : c->np;
      const uint64_t m = ntt.T.q;
      enc_zero(c, ntt, kKeyUni, limb, nonce, e.data(), c0.data(), c1.data());
      // KeyGenerator::generate_one_kswitch_key: + (p mod q) * sigma(s) in the q limb
      if (limb == 0)
        for (uint32_t i = 0; i < N; ++i) c0[i] = addm(c0[i], host::mulmod(gs[i], p_mod_q, m), m);
      ntt.fwd(c0.data());
      ntt.fwd(c1.data());
      memcpy(key + (0 * 2 + limb) * N, c0.data(), N * 8);
      memcpy(key + (1 * 2 + limb) * N, c1.data(), N * 8);
    }
  }
  return ONION_OK;
}

int onion_client_rgsw_key(onion_client *c, uint64_t *out, size_t n) {
  if (!c || !out) return cfail(ONION_EINVAL, "null argument");
  if (n != c->I.rgsw_key_words) return cfail(ONION_EINVAL, "wrong RGSW key buffer size");
  const uint32_t N = c->N, L = c->P.gsw_l_key;
  const uint64_t q = c->I.q;
  std::vector<uint64_t> g(L);
  gadget(L, c->I.base_log2_key, q, g.data());
  for (uint32_t half = 0; half < 2; ++half)
    for (uint32_t k = 0; k < L; ++k) {
      uint64_t *row = out + (size_t)(half * L + k) * 2 * N;
      enc_zero_q(c, true, 0x20000 + half * L + k, row);
      uint64_t *dst = row + (size_t)half * N;  // s*g_k into c0 (top rows) or c1 (bottom rows)
      for (uint32_t i = 0; i < N; ++i) dst[i] = addm(dst[i], host::mulmod(c->sk[i], g[k], q), q);
    }
  return ONION_OK;
}

int onion_client_query(onion_client *c, uint64_t pt_index, uint64_t nonce, uint64_t *ct) {
  if (!c || !ct) return cfail(ONION_EINVAL, "null argument");
  if (pt_index >= c->I.num_pt) return cfail(ONION_EINVAL, "plaintext index beyond the database");
  // a query's uniform a and error e come from its nonce: a repeated nonce would
  // repeat them, and two such queries' difference would reveal both indices
  if (nonce >= kNonceLimit) return cfail(ONION_EINVAL, "query nonce must be below 2^48");
  if (!c->used_nonces.insert(nonce).second) return cfail(ONION_EINVAL, "query nonce already used by this client");
  const uint32_t h = c->I.expan_height, fst = c->P.fst_dim, l = c->P.gsw_l;
  const uint64_t q = c->I.q, t = c->I.t, w = uint64_t(1) << h;
  // PirClient::get_query_indices: [col, MSB-first bits of the row]
  std::vector<uint64_t> idx(c->I.ndim);
  uint64_t row = pt_index / fst, remain = c->I.num_pt / fst;
  idx[0] = pt_index % fst;
  for (uint32_t i = 1; i < c->I.ndim; ++i) {
    remain /= 2;
    idx[i] = row / remain;
    row %= remain;
  }
  // plain_query[bitrev(col, h)] = (2^h)^-1 mod t, BFV-encrypted
  enc_zero_q(c, false, nonce, ct);
  const uint64_t pos0 = host::bitrev(idx[0], h);
  const uint64_t m = host::invmod(w % t, t);
  const uint64_t delta_m = (uint64_t)(((u128)q * m + t / 2) / t);
  ct[pos0] = addm(ct[pos0], delta_m % q, q);
  // add_gsw_to_query: c0[bitrev(fst + (i-1)l + k, h)] += g_k * (2^h)^-1 mod q
  std::vector<uint64_t> g(l);
  gadget(l, c->I.base_log2, q, g.data());
  const uint64_t inv_q = host::invmod(w % q, q);
  for (uint32_t i = 1; i < c->I.ndim; ++i) {
    if (idx[i] != 1) continue;
    for (uint32_t k = 0; k < l; ++k) {
      const uint64_t pos = host::bitrev(fst + (uint64_t)(i - 1) * l + k, h);
      ct[pos] = addm(ct[pos], host::mulmod(g[k], inv_q, q), q);
    }
  }
  return ONION_OK;
}

int onion_client_unpack_response(const onion_client *c, const uint8_t *data, size_t len, uint64_t *resp) {
  if (!c || !data || !resp) return cfail(ONION_EINVAL, "null argument");
  const uint32_t w = response_bits(c->I.small_q);
  size_t bit = 0;
  for (size_t i = 0; i < 2 * (size_t)c->N; ++i) {
    uint64_t v = 0;
    for (uint32_t j = 0; j < w; ++j, ++bit) {
      if (bit / 8 >= len) return cfail(ONION_EINVAL, "unexpected end of response stream");
      v |= (uint64_t)((data[bit / 8] >> (bit % 8)) & 1) << j;
    }
    resp[i] = v;
  }
  return ONION_OK;
}

int onion_client_decrypt(const onion_client *c, const uint64_t *resp, uint64_t *pt) {
  if (!c || !resp || !pt) return cfail(ONION_EINVAL, "null argument");
  const uint32_t N = c->N;
  const uint64_t m = c->I.small_q, t = c->I.t;
  // sk_mod_switch: -1 becomes small_q - 1
  std::vector<uint64_t> s(N), c1(N), prod(N);
  for (uint32_t i = 0; i < N; ++i) {
    s[i] = c->sk[i] > 1 ? m - 1 : c->sk[i];
    c1[i] = resp[N + i] % m;
  }
  c->nsq.mul(c1.data(), s.data(), prod.data());
  for (uint32_t i = 0; i < N; ++i) {
    const uint64_t ph = addm(prod[i], resp[i] % m, m);
    pt[i] = (uint64_t)((((u128)ph * t) + m / 2) / m) % t;
  }
  return ONION_OK;
}

int onion_client_entry(const onion_client *c, uint64_t entry_index, const uint64_t *pt, uint8_t *entry,
                       size_t cap, size_t *entry_size) {
  if (!c || !pt || !entry) return cfail(ONION_EINVAL, "null argument");
  const uint32_t bpc = c->P.plain_bits - 1;
  const size_t esize = (size_t)bpc * c->N / 8;  // one entry per plaintext (EntrySize = 0)
  if (cap < esize) return cfail(ONION_EINVAL, "entry buffer too small");
  const size_t per_pt = ((size_t)bpc * c->N) / (esize * 8);
  const size_t start_bits = (entry_index % per_pt) * esize * 8;
  size_t ci = start_bits / bpc;
  u128 buf = pt[ci] >> (start_bits % bpc);
  uint32_t off = bpc - (uint32_t)(start_bits % bpc);
  size_t n = 0;
  while (n < esize) {
    if (off >= 8) {
      entry[n++] = (uint8_t)buf;
      buf >>= 8;
      off -= 8;
    } else {
      ++ci;
      buf |= (u128)pt[ci] << off;
      off += bpc;
    }
  }
  if (entry_size) *entry_size = esize;
  return ONION_OK;
}

}  // extern "C"
