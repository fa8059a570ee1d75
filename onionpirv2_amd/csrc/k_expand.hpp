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

// tree:  2w cts, slot s at tree + s*2N (c0 then c1), coefficient form
// key:   this level's key [comp][limb q,p][N], NTT form
__global__ __launch_bounds__(kNttThreads) void k_expand_level(uint64_t *tree, const uint64_t *key,
                                                              uint32_t level, uint32_t ginv,
                                                              NttTab Tq, NttTab Tp, Mod mq, Mod mp,
                                                              KsConst ks) {
  __shared__ uint64_t A[kN], B[kN];
  const int tid = threadIdx.x;
  const uint32_t k = 1u << level;
  const uint64_t node = (uint64_t)k + blockIdx.x;
  const uint64_t q = Tq.q, p = Tp.q;
  const uint64_t *c0 = tree + node * 2 * kN, *c1 = c0 + kN;

  // sigma(c1) into both buffers; q < p so the same residues serve both limbs
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = tid + r * kNttThreads;
    const uint64_t v = galois_gather(c1, e, ginv, q);
    A[e] = v;
    B[e] = v;
  }
  __syncthreads();
  ntt_fwd_lds<kN>(A, Tq);
  ntt_fwd_lds<kN>(B, Tp);
  uint64_t tq[kPer], tp[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = tid + r * kNttThreads;
    tq[r] = ntt_fwd_final(A[e], q);
    tp[r] = ntt_fwd_final(B[e], p);
  }
  __syncthreads();

  // Component comp of c' is sigma(c0) + KS_0 (comp 0) or KS_1 (comp 1); the
  // children's polynomial comp is written as soon as it is known.
  uint64_t *left = tree + (2 * node) * 2 * kN;       // c_{2i}
  uint64_t *right = tree + (2 * node + 1) * 2 * kN;  // c_{2i+1}
#pragma unroll 1
  for (int comp = 0; comp < 2; ++comp) {
    const uint64_t *kq = key + (size_t)(2 * comp) * kN, *kp = kq + kN;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int e = tid + r * kNttThreads;
      A[e] = mulmod_barrett(tq[r], kq[e], mq);
      B[e] = mulmod_barrett(tp[r], kp[e], mp);
    }
    __syncthreads();
    ntt_inv_lds<kN>(A, Tq);
    ntt_inv_lds<kN>(B, Tp);
    const uint64_t *cin = comp == 0 ? c0 : c1;
#pragma unroll 2
    for (int r = 0; r < kPer; ++r) {
      const int e = tid + r * kNttThreads;
      const uint64_t xq = ntt_inv_final(A[e], Tq);
      const uint64_t xp = ntt_inv_final(B[e], Tp);
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
    __syncthreads();
  }
}

}  // namespace onion
