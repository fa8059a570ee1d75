// k_expand.hpp -- one level of the query-expansion tree
// (PirServer::fast_expand_qry, src/server.cpp:382-428).
//
// The tree lives in a device buffer of 2w ciphertexts (heap slots, root = 1).
// At level a (k = 2^a) the active nodes are i = k .. k + n_active - 1.  Node i
// computes c' = Subs(c_i, N/k + 1): the Galois permutation of c0 and c1, then
// key switching of sigma(c1) with the special prime p, and the children
// c_2i = c_i + c',  c_2i+1 = (c_i - c') * x^-k.  Narrow levels take two
// launches:
//   k_expand_ks      -- one workgroup per (node, component, limb q|p):
//                       NTT(sigma(c1)) * key[component][limb], INTT -> X
//   k_expand_combine -- one workgroup per node: the rounding division by p
//                       (SEAL switch_key_inplace), sigma(c0), the children.
// Splitting the key switch four ways keeps each workgroup at two transforms
// deep, so a level's latency is two transforms, not six.  Wide levels take
// one launch, k_expand_node (one workgroup per node, six transforms deep),
// which keeps X out of memory and transforms sigma(c1) once per limb for both
// components; from the first wide level on, k_expand_ntt writes NTT-form
// children (the form every reader of the leaves transforms them to).
// Batched: blockIdx.y selects the query; each query has its own tree, key and
// X scratch.
#pragma once

#include "k_ntt.hpp"

namespace onion {

// Device words per Galois-key level: the key (2 components x 2 limbs x N),
// then the Shoup companions of its words.
constexpr size_t kKeyLevelWords = 8 * (size_t)kN;

struct KsConst {
  uint64_t half_p;        // floor(p/2)
  uint64_t half_p_mod_q;  // floor(p/2) mod q
  uint64_t p_inv;         // p^-1 mod q
  uint64_t p_inv_s;       // Shoup companion of p_inv
  uint64_t np_inv;        // N p^-1 mod q (k_expand_ntt)
  uint64_t np_inv_s;      // Shoup companion of np_inv
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

// trees: query b's tree at trees + b*tree_stride (2w cts, slot s at s*2N)
// keys:  query b's Galois keys at key_ptrs[b]; this level's block of 8N at
//        + level*8N: the key [comp][limb q,p][N], NTT form, each limb already
//        scaled by N^-1 (so the inverse transforms of the key products need
//        no scaling of their own), then the Shoup companion of every word at
//        + 4N (k_key_prep at upload), so a key product is a Shoup product
// xbuf:  query b's X at xbuf + b*x_stride, [node - first][comp][limb][N]
__global__ __launch_bounds__(kNttThreads) void k_expand_ks(const uint64_t *trees, uint64_t tree_stride,
                                                           const uint64_t *const *key_ptrs, uint32_t level,
                                                           uint32_t ginv, uint64_t *xbuf, uint64_t x_stride,
                                                           NttTab Tq, NttTab Tp, Mod mq, Mod mp) {
  __shared__ uint64_t lds[kNttLds];
  const uint32_t nl = blockIdx.x >> 2, comp = (blockIdx.x >> 1) & 1, limb = blockIdx.x & 1;
  const uint64_t node = (1ull << level) + nl;
  const NttTab &T = limb ? Tp : Tq;
  const uint64_t *c1 = trees + (size_t)blockIdx.y * tree_stride + node * 2 * kN + kN;
  const uint64_t *key = key_ptrs[blockIdx.y] + (size_t)level * kKeyLevelWords + (size_t)(2 * comp + limb) * kN;
  uint64_t x[8];
  // sigma(c1); its residues < q < p serve both limbs unchanged
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = galois_gather(c1, threadIdx.x + 256 * j, ginv, Tq.q);
  ntt_fwd(x, lds, T);
  uint64_t kv[8], kvs[8];
  load_blocked(kv, key);
  load_blocked(kvs, key + 4 * kN);
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = mul_shoup_lazy(x[j], kv[j], kvs[j], T.q);  // < 2q: ntt_inv's input bound
  ntt_inv(x, lds, T);
  uint64_t *dst = xbuf + (size_t)blockIdx.y * x_stride + ((size_t)(nl * 2 + comp) * 2 + limb) * kN;
#pragma unroll
  for (int j = 0; j < 8; ++j) dst[threadIdx.x + 256 * j] = ntt_inv_canonical(x[j], T.q);
}

// The rounding division of the key switch and the children for one
// coefficient e of component comp (shared by k_expand_combine and
// k_expand_node).
__device__ __forceinline__ void expand_child(uint64_t *left, uint64_t *right, const uint64_t *c0, int comp, int e,
                                             uint64_t xq, uint64_t xp, uint32_t k, uint32_t ginv, uint64_t q,
                                             uint64_t p, const KsConst &ks) {
  // SEAL switch_key_inplace: floor((X + floor(p/2)) / p) mod q, with
  // X = CRT(X_q, X_p):  (X_q - ((X_p + floor(p/2)) mod p) + floor(p/2)) * p^-1
  uint64_t last = xp + ks.half_p;
  if (last >= p) last -= p;
  const uint64_t lq = last >= q ? last - q : last;  // p < 2q
  const uint64_t diff = sub_mod(add_mod(xq, ks.half_p_mod_q, q), lq, q);
  uint64_t ksv = mul_shoup_lazy(diff, ks.p_inv, ks.p_inv_s, q);
  if (ksv >= q) ksv -= q;
  // c'_0 = sigma(c0) + KS_0, c'_1 = KS_1 (c1 was cleared before the key switch)
  const uint64_t cp = comp == 0 ? add_mod(galois_gather(c0, e, ginv, q), ksv, q) : ksv;
  const uint64_t a = c0[(size_t)comp * kN + e];
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

__global__ __launch_bounds__(256) void k_expand_combine(uint64_t *trees, uint64_t tree_stride,
                                                        const uint64_t *xbuf, uint64_t x_stride,
                                                        uint32_t level, uint32_t ginv, uint64_t q,
                                                        uint64_t p, KsConst ks) {
  const uint32_t nl = blockIdx.x, k = 1u << level;
  const uint64_t node = (uint64_t)k + nl;
  uint64_t *tree = trees + (size_t)blockIdx.y * tree_stride;
  const uint64_t *c0 = tree + node * 2 * kN;
  uint64_t *left = tree + (2 * node) * 2 * kN;       // c_{2i}
  uint64_t *right = tree + (2 * node + 1) * 2 * kN;  // c_{2i+1}
  const uint64_t *X = xbuf + (size_t)blockIdx.y * x_stride + (size_t)nl * 4 * kN;
  for (int comp = 0; comp < 2; ++comp) {
    for (int j = 0; j < 8; ++j) {
      const int e = threadIdx.x + 256 * j;
      expand_child(left, right, c0, comp, e, X[(size_t)(2 * comp) * kN + e], X[(size_t)(2 * comp + 1) * kN + e], k,
                   ginv, q, p, ks);
    }
  }
}

// One component of a node's key switch and children (k_expand_node).
__device__ __forceinline__ void expand_node_comp(int comp, const uint64_t (&pq)[kN], const uint64_t (&pp)[kN],
                                                 const uint64_t *key, uint64_t *lds, uint64_t *left, uint64_t *right,
                                                 const uint64_t *c0, uint32_t k, uint32_t ginv, const NttTab &Tq,
                                                 const NttTab &Tp, const Mod &mq, const Mod &mp, const KsConst &ks) {
  const int t = threadIdx.x;
  uint64_t x[8], xq[8], kv[8], kvs[8];
  load_blocked(kv, key + (size_t)(2 * comp) * kN);
  load_blocked(kvs, key + (size_t)(4 + 2 * comp) * kN);
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = mul_shoup_lazy(pq[8 * t + j], kv[j], kvs[j], Tq.q);
  ntt_inv(x, lds, Tq);
#pragma unroll
  for (int j = 0; j < 8; ++j) xq[j] = ntt_inv_canonical(x[j], Tq.q);
  load_blocked(kv, key + (size_t)(2 * comp + 1) * kN);
  load_blocked(kvs, key + (size_t)(5 + 2 * comp) * kN);
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = mul_shoup_lazy(pp[8 * t + j], kv[j], kvs[j], Tp.q);
  ntt_inv(x, lds, Tp);
#pragma unroll
  for (int j = 0; j < 8; ++j)
    expand_child(left, right, c0, comp, t + 256 * j, xq[j], ntt_inv_canonical(x[j], Tp.q), k, ginv, Tq.q, Tp.q,
                 ks);
}

// One level in one launch, one workgroup per node: both components share the
// forward transforms of sigma(c1) (limbs q and p), parked in LDS in the
// blocked order each thread wrote them, so a node costs 2 forward + 4 inverse
// transforms (a workgroup per (node, component) would repeat the forward pair:
// 4 + 4, measured 4.8 ms against 3.9 ms per batch of 64).  The two components run one after the other as straight
// code with a compiler fence between them: as a loop, the optimizer hoists the
// inverse transforms' twiddle loads out of it and they stay live throughout.
__global__ __launch_bounds__(kNttThreads) void k_expand_node(uint64_t *trees, uint64_t tree_stride,
                                                             const uint64_t *const *key_ptrs, uint32_t level,
                                                             uint32_t ginv, NttTab Tq, NttTab Tp, Mod mq, Mod mp,
                                                             KsConst ks) {
  __shared__ uint64_t lds[kNttLds];
  __shared__ uint64_t park[2][kN];  // NTT_q and NTT_p of sigma(c1); each thread reads back its own slots
  const uint32_t nl = blockIdx.x, k = 1u << level;
  const uint64_t node = (uint64_t)k + nl;
  uint64_t *tree = trees + (size_t)blockIdx.y * tree_stride;
  const uint64_t *c0 = tree + node * 2 * kN;
  const uint64_t *c1 = c0 + kN;
  const uint64_t *key = key_ptrs[blockIdx.y] + (size_t)level * kKeyLevelWords;
  const int t = threadIdx.x;
  uint64_t x[8];
  // sigma(c1); its residues < q < p serve both limbs unchanged
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = galois_gather(c1, t + 256 * j, ginv, Tq.q);
  ntt_fwd(x, lds, Tq);
#pragma unroll
  for (int j = 0; j < 8; ++j) park[0][8 * t + j] = x[j];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = galois_gather(c1, t + 256 * j, ginv, Tq.q);
  ntt_fwd(x, lds, Tp);
#pragma unroll
  for (int j = 0; j < 8; ++j) park[1][8 * t + j] = x[j];
  // children: every thread reads its own coefficients of this node and writes
  // its own coefficients of the children (other tree slots), so no barrier is
  // needed between reading c_i and writing c_2i, c_2i+1
  uint64_t *left = tree + (2 * node) * 2 * kN, *right = tree + (2 * node + 1) * 2 * kN;
  expand_node_comp(0, park[0], park[1], key, lds, left, right, c0, k, ginv, Tq, Tp, mq, mp, ks);
  asm volatile("" ::: "memory");
  expand_node_comp(1, park[0], park[1], key, lds, left, right, c0, k, ginv, Tq, Tp, mq, mp, ks);
}

// slot of NTT(a) holding sigma_g(a)'s slot s: slot s holds a(psi^o(s)) with
// o(s) = 2 bitrev(s) + 1, and sigma_g(a)(psi^o) = a(psi^(g o)), so the source
// is the slot with o = g o(s) mod 2N (odd: its (o - 1) / 2 is o >> 1)
__device__ __forceinline__ uint32_t galois_slot(uint32_t s, uint32_t g) {
  const uint32_t o = 2 * (__brev(s) >> 21) + 1;  // 11-bit reversal (N = 2048)
  return __brev(((g * o) & (2 * kN - 1)) >> 1) >> 21;
}

// One component of a node with NTT-form children (k_expand_ntt).  P: the
// thread's slots of NTT_p(sigma(c1)); Q[0], Q[1]: NTT_q(c0), NTT_q(c1).
__device__ __forceinline__ void ntt_node_comp(int comp, const uint64_t (&P)[8], const uint64_t (&Q)[2][kN],
                                              const uint64_t *key, const uint64_t *xk, uint32_t g, uint64_t *lds,
                                              uint64_t *left, uint64_t *right, const NttTab &Tq, const NttTab &Tp,
                                              const Mod &mq, const Mod &mp, const KsConst &ks) {
  const int t = threadIdx.x;
  const uint64_t q = Tq.q, p = Tp.q;
  uint64_t x[8], kv[8], ksh[8];
  // the p-limb product in coefficient form: the rounding term of the key switch
  load_blocked(kv, key + (size_t)(2 * comp + 1) * kN);
  load_blocked(ksh, key + (size_t)(5 + 2 * comp) * kN);
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = mul_shoup_lazy(P[j], kv[j], ksh[j], p);  // < 2p: ntt_inv's input bound
  ntt_inv(x, lds, Tp);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    // expand_child's division, rearranged: KS = (X_q - corr) p^-1 with
    // corr = ((X_p + floor(p/2)) mod p) mod q - floor(p/2) mod q, taken
    // here times N^-1 (the q-limb key carries N^-1)
    uint64_t last = ntt_inv_canonical(x[j], p) + ks.half_p;
    if (last >= p) last -= p;
    const uint64_t lq = last >= q ? last - q : last;  // p < 2q
    x[j] = mul_shoup_lazy(sub_mod(lq, ks.half_p_mod_q, q), Tq.n_inv, Tq.n_inv_s, q);  // < 2q
  }
  ntt_fwd(x, lds, Tq);
  // slot pairs 8t + j, 8t + j + 1: 16-byte loads and stores
  const uint4 *kq = reinterpret_cast<const uint4 *>(key + (size_t)(2 * comp) * kN + 8 * t);
  const uint4 *kqs = reinterpret_cast<const uint4 *>(key + (size_t)(4 + 2 * comp) * kN + 8 * t);
  uint4 *lo = reinterpret_cast<uint4 *>(left + (size_t)comp * kN + 8 * t);
  uint4 *ro = reinterpret_cast<uint4 *>(right + (size_t)comp * kN + 8 * t);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint4 kp = kq[i], kps = kqs[i];
    uint64_t lv[2], rv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = 2 * i + h;
      const uint32_t s = 8 * t + j, src = galois_slot(s, g);
      const uint64_t kj = h ? ((uint64_t)kp.w << 32) | kp.z : ((uint64_t)kp.y << 32) | kp.x;
      const uint64_t kjs = h ? ((uint64_t)kps.w << 32) | kps.z : ((uint64_t)kps.y << 32) | kps.x;
      // N^-1 NTT(X_q) = NTT_q(sigma(c1)) * key_q (as stored), in [0, 2q): the
      // subtraction below leaves it below 2q, and the Shoup product after it
      // takes any input
      const uint64_t xq = mul_shoup_lazy(Q[1][src], kj, kjs, q);
      uint64_t ksv = mul_shoup_lazy(sub_mod(xq, ntt_fwd_final(x[j], q), q), ks.np_inv, ks.np_inv_s, q);
      if (ksv >= q) ksv -= q;
      // c'_0 = sigma(c0) + KS_0, c'_1 = KS_1
      const uint64_t cp = comp == 0 ? add_mod(Q[0][src], ksv, q) : ksv;
      const uint64_t a = Q[comp][s];
      lv[h] = add_mod(a, cp, q);
      const uint64_t r = mul_shoup_lazy(sub_mod(a, cp, q), xk[s], xk[kN + s], q);  // times x^-k
      rv[h] = r >= q ? r - q : r;
    }
    lo[i] = make_uint4((uint32_t)lv[0], (uint32_t)(lv[0] >> 32), (uint32_t)lv[1], (uint32_t)(lv[1] >> 32));
    ro[i] = make_uint4((uint32_t)rv[0], (uint32_t)(rv[0] >> 32), (uint32_t)rv[1], (uint32_t)(rv[1] >> 32));
  }
}

// An expansion level with NTT-form children (slot order, canonical), one
// workgroup per node: the form the selection vectors and the conversion's top
// rows take, so their forward transforms disappear.  In NTT form sigma_g is a
// slot permutation (galois_slot) and x^-k the factor psi^(-k o(s)) of slot s
// (xk: the level's N factors, then N Shoup companions).  The q-limb key
// products need no transform; each component's p-limb product returns to
// coefficient form for the rounding division and its correction is
// transformed forward (2 + 2).  The p-limb product needs NTT_p(sigma(c1)):
//   NTT_IN = false (coefficient-form node): c0, c1 forward mod q and sigma(c1)
//     forward mod p, 3 + 4 = 7 transforms (k_expand_node takes 6, and the
//     children's next readers 4 more);
//   NTT_IN = true (the level above wrote NTT form): sigma(c1) from c1's
//     slots, back to coefficient form mod q and forward mod p, 2 + 4 = 6.
// The node's residues in coefficient form go unchanged into the rounding, so
// the children equal k_expand_node's, transformed.
template <bool NTT_IN>
__global__ __launch_bounds__(kNttThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_expand_ntt(
    uint64_t *trees, uint64_t tree_stride, const uint64_t *const *key_ptrs, uint32_t level, uint32_t g,
    uint32_t ginv, const uint64_t *xk, NttTab Tq, NttTab Tp, Mod mq, Mod mp, KsConst ks) {
  __shared__ uint64_t lds[kNttLds];
  __shared__ uint64_t Q[2][kN];  // NTT_q(c0), NTT_q(c1), canonical; read at permuted slots
  const uint32_t nl = blockIdx.x, k = 1u << level;
  const uint64_t node = (uint64_t)k + nl;
  uint64_t *tree = trees + (size_t)blockIdx.y * tree_stride;
  const uint64_t *c = tree + node * 2 * kN;
  const uint64_t *key = key_ptrs[blockIdx.y] + (size_t)level * kKeyLevelWords;
  const int t = threadIdx.x;
  uint64_t x[8], P[8];
  if (NTT_IN) {
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      load_blocked(x, c + (size_t)comp * kN);
#pragma unroll
      for (int j = 0; j < 8; ++j) Q[comp][8 * t + j] = x[j];
    }
    __syncthreads();  // Q[1] is read at permuted slots
    // sigma(c1) in NTT form, then its coefficients (< q < p serve mod p unchanged)
#pragma unroll
    for (int j = 0; j < 8; ++j) P[j] = Q[1][galois_slot(8 * t + j, g)];
    ntt_inv(P, lds, Tq);
#pragma unroll
    for (int j = 0; j < 8; ++j) P[j] = ntt_inv_final(P[j], Tq);
  } else {
#pragma unroll 1
    for (int comp = 0; comp < 2; ++comp) {
      load_natural(x, c + (size_t)comp * kN);
      ntt_fwd(x, lds, Tq);
#pragma unroll
      for (int j = 0; j < 8; ++j) Q[comp][8 * t + j] = ntt_fwd_final(x[j], Tq.q);
    }
    // sigma(c1) mod p; its residues < q < p serve unchanged
#pragma unroll
    for (int j = 0; j < 8; ++j) P[j] = galois_gather(c + kN, t + 256 * j, ginv, Tq.q);
  }
  ntt_fwd(P, lds, Tp);  // its barriers also order the Q writes before the reads below
  // the children are other tree slots: every read of c precedes these writes
  uint64_t *left = tree + (2 * node) * 2 * kN, *right = tree + (2 * node + 1) * 2 * kN;
  ntt_node_comp(0, P, Q, key, xk, g, lds, left, right, Tq, Tp, mq, mp, ks);
  asm volatile("" ::: "memory");
  ntt_node_comp(1, P, Q, key, xk, g, lds, left, right, Tq, Tp, mq, mp, ks);
}

// Shoup companion floor(w 2^64 / m) of w < m, from the Barrett ratio
// R = floor(2^128 / m) of m: floor(w R / 2^64) = w R_hi + floor(w R_lo / 2^64)
// is the companion or one less (w R / 2^64 falls short of w 2^64 / m by
// w (2^128 mod m) / (m 2^64) < 1), and the remainder w 2^64 - c m, below 2m
// and so equal to -c m mod 2^64, tells which.
__device__ __forceinline__ uint64_t shoup_companion(uint64_t w, const Mod &m) {
  uint64_t c = w * m.r_hi + mulhi64(w, m.r_lo);
  if ((uint64_t)0 - c * m.q >= m.q) ++c;
  return c;
}

// Galois keys as stored: every value of limb q (p) times N^-1 mod q (mod p),
// then its Shoup companion.  Layout [level][key | companions][comp][limb][N]
// (kKeyLevelWords per level); the uploaded keys are in the first half of each
// level.  grid = levels * 4N / 256.
__global__ __launch_bounds__(256) void k_key_prep(uint64_t *keys, uint64_t n_words, uint64_t nq_inv, uint64_t np_inv,
                                                  Mod mq, Mod mp) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;  // index among the levels' 4N key words
  if (i >= n_words) return;
  const uint64_t level = i / (4 * kN), o = i % (4 * kN);
  uint64_t *w = keys + level * kKeyLevelWords + o;
  const bool limb_p = (o / kN) & 1;
  const Mod &m = limb_p ? mp : mq;
  const uint64_t v = mulmod_barrett(*w, limb_p ? np_inv : nq_inv, m);
  w[0] = v;
  w[4 * kN] = shoup_companion(v, m);
}

}  // namespace onion
