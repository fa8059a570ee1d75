// test_pir.cpp -- the reference's end-to-end test (PirTest::test_pir,
// src/tests.cpp:40-153) written against the onionpir C++ API: build a
// database, then for each run make a fresh client, ship its keys, query a
// random entry, answer on the GPU, decrypt, and compare with
// direct_get_entry.  Exit status 0 only when every run succeeds.
//
//   test_pir [fst_dim other_dim [runs [plain_bits]]]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <vector>

#include "onionpir.hpp"

using namespace onionpir;

int main(int argc, char **argv) {
  const size_t fst = argc > 2 ? std::strtoul(argv[1], nullptr, 10) : 256;
  const size_t other = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 32;
  const int runs = argc > 3 ? std::atoi(argv[3]) : 13;
  const size_t plain_bits = argc > 4 ? std::strtoul(argv[4], nullptr, 10) : 16;
  const int warmup = 3;

  PirParams pir_params(fst, other, plain_bits);
  pir_params.print_params();
  PirServer server(pir_params);
  server.gen_data(20240601);
  server.finalize_database();

  size_t galois_key_size = 0, gsw_key_size = 0, query_size = 0, resp_size = 0;
  int success = 0;
  double server_ms = 0;
  std::srand(7);
  for (int i = 0; i < runs; ++i) {
    PirClient client(pir_params, 100 + i, 9000 + i);
    const size_t client_id = client.get_client_id();
    std::stringstream galois_key_stream, gsw_stream, query_stream, resp_stream;
    galois_key_size = client.create_galois_keys(galois_key_stream);
    gsw_key_size = client.write_gsw_to_stream(client.generate_gsw_from_key(), gsw_stream);
    server.set_client_galois_key(client_id, galois_key_stream);
    server.set_client_gsw_key(client_id, gsw_stream);

    const size_t query_index = (size_t)std::rand() % pir_params.get_num_entries();
    Ciphertext query = client.fast_generate_query(query_index);
    query_size = PirClient::write_query_to_stream(query, query_stream);

    const auto t0 = std::chrono::steady_clock::now();
    Ciphertext response = server.make_query(client_id, query_stream);
    const auto t1 = std::chrono::steady_clock::now();
    if (i >= warmup) server_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();

    resp_size = server.save_resp_to_stream(response, resp_stream);
    Ciphertext reconstructed = client.load_resp_from_stream(resp_stream);
    Plaintext decrypted = client.decrypt_reply(reconstructed);
    Entry got = client.get_entry_from_plaintext(query_index, decrypted);
    Entry want = server.direct_get_entry(query_index);
    const bool ok = got == want;
    success += ok;
    std::printf("run %d: entry %zu %s\n", i + 1, query_index, ok ? "Success!" : "Failure!");
  }
  const int measured = runs > warmup ? runs - warmup : 0;
  const double avg_ms = measured ? server_ms / measured : 0;
  std::printf("Success rate: %d/%d\n", success, runs);
  std::printf("galois key size: %zu bytes\n", galois_key_size);
  std::printf("gsw key size: %zu bytes\n", gsw_key_size);
  std::printf("query size: %zu bytes\n", query_size);
  std::printf("response size: %zu bytes\n", resp_size);
  if (measured)
    std::printf("Server throughput: %.1f MB/s (%.3f ms per query)\n", pir_params.get_DBSize_MB() / (avg_ms / 1000),
                avg_ms);
  return success == runs ? 0 : 1;
}
