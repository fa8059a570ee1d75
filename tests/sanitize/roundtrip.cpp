// End-to-end run of the host-side code under AddressSanitizer and
// UndefinedBehaviorSanitizer, on the CPU: the product client
// (onionpirv2_amd/csrc/client.cpp) makes keys and queries in both generator
// modes, the oracle (oracle/*.c) answers them, and the client unpacks the
// packed response, decrypts it and must recover the record.  Also drives the
// client's argument checks (used and oversized nonces, wrong buffer sizes,
// short responses).
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
#include "../../oracle/oracle.h"
}
#include "onionpir_client.h"

#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
      return 1;                                                            \
    }                                                                      \
  } while (0)

int main() {
  const uint32_t N = 2048, fst = 16, other = 4, plain_bits = 16, l = 5, l_key = 10;
  const uint64_t db_seed = 7;
  orc_ctx *c = orc_create(N, plain_bits, l, l_key, fst, other);
  CHECK(c != nullptr);
  const onion_params P{N, plain_bits, l, l_key, fst, other};
  std::vector<uint64_t> db((size_t)c->num_pt * N);
  orc_db_build(c, db_seed, 0, db.data());
  const size_t gk_words = (size_t)c->h * 4 * N, rk_words = (size_t)2 * l_key * 2 * N;
  for (int secure = 0; secure < 2; ++secure) {
    onion_client *cl = nullptr;
    CHECK((secure ? onion_client_create_secure(&P, 3, &cl) : onion_client_create(&P, 3, 99, &cl)) == ONION_OK);
    std::vector<uint64_t> gk(gk_words), rk(rk_words), rkn(rk_words);
    CHECK(onion_client_galois_keys(cl, gk.data(), gk_words - 1) == ONION_EINVAL);
    CHECK(onion_client_galois_keys(cl, gk.data(), gk_words) == ONION_OK);
    CHECK(onion_client_rgsw_key(cl, rk.data(), rk_words) == ONION_OK);
    orc_rgsw_key_to_ntt(c, rk.data(), rkn.data());
    uint64_t nonce = 1;
    for (uint64_t idx : {uint64_t(0), c->num_pt - 1, uint64_t(5)}) {
      std::vector<uint64_t> q(2 * N), resp(2 * N), back(2 * N), pt(N), want(N);
      CHECK(onion_client_query(cl, idx, nonce, q.data()) == ONION_OK);
      CHECK(onion_client_query(cl, idx, nonce, q.data()) == ONION_EINVAL);  // used nonce
      ++nonce;
      orc_times tm;
      CHECK(orc_answer(c, db.data(), gk.data(), rkn.data(), q.data(), resp.data(), &tm) == 0);
      std::vector<uint8_t> wire(2 * N * 8);
      const size_t len = orc_pack_response(c, resp.data(), wire.data());
      CHECK(len > 0 && len <= wire.size());
      CHECK(onion_client_unpack_response(cl, wire.data(), len - 1, back.data()) == ONION_EINVAL);
      CHECK(onion_client_unpack_response(cl, wire.data(), len, back.data()) == ONION_OK);
      CHECK(std::memcmp(back.data(), resp.data(), 2 * N * 8) == 0);
      CHECK(onion_client_decrypt(cl, back.data(), pt.data()) == ONION_OK);
      orc_db_plaintext(c, db_seed, 0, idx, want.data());
      CHECK(std::memcmp(pt.data(), want.data(), N * 8) == 0);
    }
    std::vector<uint64_t> q(2 * N);
    CHECK(onion_client_query(cl, 0, uint64_t(1) << 48, q.data()) == ONION_EINVAL);
    CHECK(onion_client_query(cl, c->num_pt, 100, q.data()) == ONION_EINVAL);
    CHECK(onion_client_destroy(cl) == ONION_OK);
  }
  orc_destroy(c);
  std::printf("sanitized round trip ok\n");
  return 0;
}
