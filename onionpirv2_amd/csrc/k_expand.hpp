// k_expand.hpp -- one level of the query-expansion tree
// (PirServer::fast_expand_qry, src/server.cpp:382-428).
//
// The tree lives in a device buffer of 2w ciphertexts (heap slots, root = 1).
// At level a (k = 2^a) the active nodes are i = k .. k + n_active - 1; one
// workgroup handles one node end to end:
//   c'   = Subs(c_i, N/k + 1)  -- Galois permutation of c0 and c1, then key
//          switching of sigma(c1) with the special prime p (NTT mod q and mod
//          p, product with the level's key, INTT, rounding division by p)
//   c_2i = c_i + c',   c_2i+1 = (c_i - c') * x^-k.
// Batched: blockIdx.y selects the query; each query has its own tree and key.
#pragma once

#include "k_ntt.hpp"

namespace onion {

struct KsConst {
  uint64_t half_p;        // floor(p/2)
  uint64_t half_p_mod_q;  // floor(p/2) mod q
  uint64_t p_inv;         // p^-1 mod q
  uint64_t p_inv_s;       // Shoup companion of p_inv
};

// Galois permutation sigma_g of a coefficient-form polynomial (SEAL
// GaloisTool::apply_galois): out[(e*g) mod N] = +/- in[e], negated when
// floor(e*g/N) is odd.  Written as a gather: out[j] = in[i] when
// i = j*g^-1 mod 2N < N, else -in[i - N].
__device__ __forceinline__ uint64_t galois_gather(const uint64_t *in, int j, uint32_t ginv,
                                                  uint64_t q) {
  const uint32_t i = (uint32_t)(((uint64_t)j * ginv) & (2 * kN - 1));
  if (i < (uint32_t)kN) return in[i];
  const uint64_t v = in[i - kN];
  return v ? q - v : 0;
}

// trees: query b's tree at trees + b * tree_stride (2w cts, slot s at s*2N)
// keys:  query b's Galois keys at key_ptrs[b]; this level's key at + level*4N,
//        laid out [comp][limb q,p][N], NTT form
__global__ __launch_bounds__(kNttThreads) void k_expand_level(uint64_t *trees, uint64_t tree_stride,
                                                              const uint64_t *const *key_ptrs,
                                                              uint32_t level, uint32_t ginv,
                                                              NttTab Tq, NttTab Tp, Mod mq, Mod mp,
                                                              KsConst ks) {
  __shared__ uint64_t lds[kNttLds];
  const int tid = threadIdx.x;
  const uint32_t k = 1u << level;
  const uint64_t node = (uint64_t)k + blockIdx.x;
  const uint64_t q = Tq.q, p = Tp.q;
  uint64_t *tree = trees + (size_t)blockIdx.y * tree_stride;
  const uint64_t *key = key_ptrs[blockIdx.y] + (size_t)level * 4 * kN;
  const uint64_t *c0 = tree + node * 2 * kN, *c1 = c0 + kN;

  // sigma(c1), NTT mod q and mod p (q < p: the same residues serve both)
  uint64_t tq[8], tp[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) tq[j] = tp[j] = galois_gather(c1, tid + 256 * j, ginv, q);
  ntt_fwd(tq, lds, Tq);
  ntt_fwd(tp, lds, Tp);

  // Component comp of c' is sigma(c0) + KS_0 (comp 0) or KS_1 (comp 1); the
  // children's polynomial comp is written as soon as it is known.
  uint64_t *left = tree + (2 * node) * 2 * kN;       // c_{2i}
  uint64_t *right = tree + (2 * node + 1) * 2 * kN;  // c_{2i+1}
#pragma unroll 1
  for (int comp = 0; comp < 2; ++comp) {
    const uint64_t *kq = key + (size_t)(2 * comp) * kN, *kp = kq + kN;
    uint64_t aq[8], ap[8], kv[8];
    load_blocked(kv, kq);
#pragma unroll
    for (int j = 0; j < 8; ++j) aq[j] = mulmod_barrett(tq[j], kv[j], mq);
    load_blocked(kv, kp);
#pragma unroll
    for (int j = 0; j < 8; ++j) ap[j] = mulmod_barrett(tp[j], kv[j], mp);
    ntt_inv(aq, lds, Tq);
    ntt_inv(ap, lds, Tp);
    const uint64_t *cin = comp == 0 ? c0 : c1;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = tid + 256 * j;
      const uint64_t xq = ntt_inv_final(aq[j], Tq);
      const uint64_t xp = ntt_inv_final(ap[j], Tp);
      uint64_t last = xp + ks.half_p;  // (X_p + floor(p/2)) mod p
      if (last >= p) last -= p;
      const uint64_t lq = last >= q ? last - q : last;  // p < 2q
      const uint64_t diff = sub_mod(add_mod(xq, ks.half_p_mod_q, q), lq, q);
      uint64_t ksv = mul_shoup_lazy(diff, ks.p_inv, ks.p_inv_s, q);
      if (ksv >= q) ksv -= q;
      const uint64_t cp = comp == 0 ? add_mod(galois_gather(c0, e, ginv, q), ksv, q) : ksv;
      const uint64_t a = cin[e];
      left[(size_t)comp * kN + e] = add_mod(a, cp, q);
      uint64_t d = sub_mod(a, cp, q);
      // multiply by x^-k: slot e moves to e - k; wrapped slots are negated
      int dst = e - (int)k;
      if (dst < 0) {
        dst += kN;
        d = d ? q - d : 0;
      }
      right[(size_t)comp * kN + dst] = d;
    }
  }
}

}  // namespace onion
