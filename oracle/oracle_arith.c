/*
 * oracle_arith.c -- parameters, modular arithmetic and the negacyclic NTT of
 * the OnionPIRv2 CPU oracle.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates:
 *   - SEAL 4.1 CoeffModulus::Create / get_primes (called at src/pir.cpp:22-24,
 *     src/pir.cpp:40-41): primes = 1 mod 2N taken downward from 2^bits, the
 *     list consumed from its back, so the data prime is the smaller one.
 *   - utils::generate_prime (src/utils.cpp:82-95).
 *   - PirParams shape rules (src/pir.cpp:53-54, 67-94; src/includes/pir.h:66-68).
 *   - utils::gsw_gadget (src/utils.cpp:61-74), utils::bit_reverse
 *     (src/includes/utils.h:110-118).
 *   - SEAL barrett_reduce_128 (used at src/server.cpp:222, src/gsw_eval.cpp:100).
 *   - SEAL ntt_negacyclic_harvey / inverse_ntt_negacyclic_harvey (used at
 *     src/gsw_eval.cpp:24,28,248 and inside transform_to/from_ntt_inplace):
 *     Harvey butterflies with Shoup-precomputed twiddles, minimal primitive
 *     2N-th root, output in bit-reversed order.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef orc_u128 u128;

/* ---------------------------------------------------------------------------------- */
/* scalar modular arithmetic                                                          */
/* ---------------------------------------------------------------------------------- */

uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((u128)a * b) % m); }

uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  a %= m;
  while (e) {
    if (e & 1) r = orc_mulmod(r, a, m);
    a = orc_mulmod(a, a, m);
    e >>= 1;
  }
  return r;
}

/* inverse by the extended Euclidean algorithm; 0 when not invertible */
uint64_t orc_invmod(uint64_t a, uint64_t m) {
  __int128 t = 0, newt = 1, r = m, newr = a % m;
  while (newr != 0) {
    __int128 qq = r / newr, tmp;
    tmp = t - qq * newt; t = newt; newt = tmp;
    tmp = r - qq * newr; r = newr; newr = tmp;
  }
  if (r != 1) return 0;
  if (t < 0) t += m;
  return (uint64_t)t;
}

/* deterministic Miller-Rabin; these seven bases decide every n < 2^64 */
int orc_is_prime(uint64_t n) {
  static const uint64_t bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
  if (n < 2) return 0;
  for (uint64_t sp = 2; sp < 64; ++sp) {
    if (n == sp) {
      for (uint64_t d = 2; d * d <= sp; ++d)
        if (sp % d == 0) return 0;
      return 1;
    }
  }
  if ((n & 1) == 0) return 0;
  uint64_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) { d >>= 1; ++s; }
  for (int i = 0; i < 7; ++i) {
    uint64_t a = bases[i] % n;
    if (a == 0) continue;
    uint64_t x = orc_powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    int comp = 1;
    for (int r = 1; r < s; ++r) {
      x = orc_mulmod(x, x, n);
      if (x == n - 1) { comp = 0; break; }
    }
    if (comp) return 0;
  }
  return 1;
}

/* SEAL get_primes: `count` primes = 1 mod factor, scanning down from 2^bits */
size_t orc_get_primes(uint64_t factor, uint32_t bits, size_t count, uint64_t *out) {
  uint64_t v = ((uint64_t)1 << bits) - factor + 1;
  uint64_t lo = (uint64_t)1 << (bits - 1);
  size_t n = 0;
  while (n < count && v > lo) {
    if (orc_is_prime(v)) out[n++] = v;
    v -= factor;
  }
  return n;
}

/* utils::generate_prime: smallest prime above 2^(bit_width-1) (src/utils.cpp:82-95) */
uint64_t orc_generate_prime(uint32_t bit_width) {
  uint64_t cand = (uint64_t)1 << (bit_width - 1);
  do {
    cand++;
    cand |= 1;
  } while (!orc_is_prime(cand));
  return cand;
}

/* PirParams ctor, src/pir.cpp:67-73 */
void orc_shape_from_entries(uint64_t num_entries, uint64_t entries_per_pt, uint64_t max_fst,
                            uint64_t *fst, uint64_t *other) {
  uint64_t need = (num_entries + entries_per_pt - 1) / entries_per_pt;
  uint64_t p2 = 1;
  while (p2 < need) p2 <<= 1;
  uint64_t o = (p2 + max_fst - 1) / max_fst;
  *other = o;
  *fst = (need + o - 1) / o;
}

/* SEAL barrett_reduce_128: x mod m for any 128-bit x = hi*2^64 + lo, using the
 * two-word ratio floor(2^128/m); one conditional subtraction finishes it. */
uint64_t orc_barrett128(const orc_mod *m, uint64_t lo, uint64_t hi) {
  u128 p0 = (u128)lo * m->ratio_lo;
  u128 p1 = (u128)lo * m->ratio_hi;
  u128 p2 = (u128)hi * m->ratio_lo;
  /* middle column: high word of lo*r_lo + low words of lo*r_hi and hi*r_lo */
  u128 mid = (p0 >> 64) + (uint64_t)p1 + (uint64_t)p2;
  /* estimate of floor(x * ratio / 2^128), mod 2^64 is all that is needed */
  uint64_t qhat = hi * m->ratio_hi + (uint64_t)(p1 >> 64) + (uint64_t)(p2 >> 64) +
                  (uint64_t)(mid >> 64);
  uint64_t r = lo - qhat * m->value;
  return r >= m->value ? r - m->value : r;
}

static void mod_init(orc_mod *m, uint64_t v) {
  m->value = v;
  u128 r = (~(u128)0) / v; /* == floor(2^128/v) for odd v */
  m->ratio_lo = (uint64_t)r;
  m->ratio_hi = (uint64_t)(r >> 64);
}

void orc_gadget(uint32_t l, uint32_t base_log2, uint64_t q, uint64_t *out) {
  /* utils::gsw_gadget: out[l-1] = 1, out[j] = 2^((l-1-j)*B) mod q */
  u128 pw = 1;
  for (int j = (int)l - 1; j >= 0; --j) {
    out[j] = (uint64_t)pw;
    pw = (pw << base_log2) % q;
  }
}

uint64_t orc_bit_reverse(uint64_t x, uint32_t logn) {
  uint64_t y = 0;
  for (uint32_t i = 0; i < logn; ++i) {
    y = (y << 1) | (x & 1);
    x >>= 1;
  }
  return y;
}

/* counter-based generator shared by the database and the seeded client:
 * SplitMix64 finaliser over (seed, stream, counter) */
static inline uint64_t mix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xbf58476d1ce4e5b9ULL;
  z ^= z >> 27;
  z *= 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return z;
}
uint64_t orc_rand64(uint64_t seed, uint64_t stream, uint64_t ctr) {
  uint64_t key = mix64(seed * 0x9e3779b97f4a7c15ULL + mix64(stream ^ 0x632be59bd9b4e019ULL));
  return mix64(key + (ctr + 1) * 0x9e3779b97f4a7c15ULL);
}

/* ---------------------------------------------------------------------------------- */
/* NTT tables and transforms                                                          */
/* ---------------------------------------------------------------------------------- */

static inline uint64_t shoup_pre(uint64_t w, uint64_t m) {
  return (uint64_t)(((u128)w << 64) / m);
}
/* x*w mod m, lazy: result in [0, 2m) for any 64-bit x (m < 2^62) */
static inline uint64_t mul_shoup_lazy(uint64_t x, uint64_t w, uint64_t ws, uint64_t m) {
  uint64_t qh = (uint64_t)(((u128)x * ws) >> 64);
  return x * w - qh * m;
}

/* SEAL try_minimal_primitive_root(2N): smallest among the odd powers of one
 * primitive 2N-th root, i.e. the smallest primitive 2N-th root overall */
static uint64_t minimal_root(uint64_t m, uint32_t N) {
  uint64_t deg = 2 * (uint64_t)N;
  uint64_t root = 0;
  for (uint64_t x = 2; x < m; ++x) {
    uint64_t r = orc_powmod(x, (m - 1) / deg, m);
    if (orc_powmod(r, N, m) == m - 1) { root = r; break; }
  }
  uint64_t sq = orc_mulmod(root, root, m), cur = root, best = root;
  for (uint64_t i = 0; i < deg; i += 2) {
    if (cur < best) best = cur;
    cur = orc_mulmod(cur, sq, m);
  }
  return best;
}

static void ntt_init(orc_ntt_tables *T, uint64_t m, uint32_t N, uint32_t logN) {
  T->N = N;
  T->logN = logN;
  mod_init(&T->mod, m);
  T->psi = minimal_root(m, N);
  T->rp = orc_alloc(sizeof(uint64_t) * N);
  T->rp_shoup = orc_alloc(sizeof(uint64_t) * N);
  T->irp = orc_alloc(sizeof(uint64_t) * N);
  T->irp_shoup = orc_alloc(sizeof(uint64_t) * N);
  uint64_t psi_inv = orc_invmod(T->psi, m);
  uint64_t pw = 1, ipw = 1;
  for (uint32_t i = 0; i < N; ++i) {
    uint64_t k = orc_bit_reverse(i, logN);
    T->rp[k] = pw;
    T->irp[k] = ipw;
    pw = orc_mulmod(pw, T->psi, m);
    ipw = orc_mulmod(ipw, psi_inv, m);
  }
  for (uint32_t k = 0; k < N; ++k) {
    T->rp_shoup[k] = shoup_pre(T->rp[k], m);
    T->irp_shoup[k] = shoup_pre(T->irp[k], m);
  }
  T->n_inv = orc_invmod(N, m);
  T->n_inv_shoup = shoup_pre(T->n_inv, m);
}

static void ntt_free(orc_ntt_tables *T) {
  free(T->rp); free(T->rp_shoup); free(T->irp); free(T->irp_shoup);
}

static const orc_ntt_tables *tables(const orc_ctx *c, int which) {
  return which == 0 ? &c->ntt_q : which == 1 ? &c->ntt_p : &c->ntt_sq;
}

uint64_t orc_ntt_psi(const orc_ctx *c, int which) { return tables(c, which)->psi; }

/* Forward negacyclic NTT, Cooley-Tukey with bit-reversed twiddles, Harvey lazy
 * butterflies (values kept in [0, 4m)), fully reduced at the end.  Output
 * slot j holds a(psi^(2*bitrev(j)+1)). */
void orc_ntt_fwd(const orc_ctx *c, int which, uint64_t *a) {
  const orc_ntt_tables *T = tables(c, which);
  const uint64_t m = T->mod.value, two_m = 2 * m;
  const uint32_t N = T->N;
  for (uint32_t groups = 1, half = N >> 1; groups < N; groups <<= 1, half >>= 1) {
    for (uint32_t i = 0; i < groups; ++i) {
      const uint64_t w = T->rp[groups + i], ws = T->rp_shoup[groups + i];
      uint64_t *x = a + 2 * (size_t)i * half, *y = x + half;
      for (uint32_t j = 0; j < half; ++j) {
        uint64_t u = x[j];
        if (u >= two_m) u -= two_m;
        uint64_t v = mul_shoup_lazy(y[j], w, ws, m);
        x[j] = u + v;
        y[j] = u + two_m - v;
      }
    }
  }
  for (uint32_t j = 0; j < N; ++j) {
    uint64_t v = a[j];
    if (v >= two_m) v -= two_m;
    if (v >= m) v -= m;
    a[j] = v;
  }
}

/* Inverse negacyclic NTT, Gentleman-Sande with the inverse twiddles, scaled by
 * N^-1; input in bit-reversed order and canonical, output canonical. */
void orc_ntt_inv(const orc_ctx *c, int which, uint64_t *a) {
  const orc_ntt_tables *T = tables(c, which);
  const uint64_t m = T->mod.value, two_m = 2 * m;
  const uint32_t N = T->N;
  for (uint32_t span = N, half = 1; span > 1; span >>= 1, half <<= 1) {
    const uint32_t groups = span >> 1;
    for (uint32_t i = 0; i < groups; ++i) {
      const uint64_t w = T->irp[groups + i], ws = T->irp_shoup[groups + i];
      uint64_t *x = a + 2 * (size_t)i * half, *y = x + half;
      for (uint32_t j = 0; j < half; ++j) {
        uint64_t u = x[j], v = y[j];
        uint64_t s = u + v;
        if (s >= two_m) s -= two_m;
        x[j] = s;
        y[j] = mul_shoup_lazy(u + two_m - v, w, ws, m);
      }
    }
  }
  for (uint32_t j = 0; j < N; ++j) {
    uint64_t v = mul_shoup_lazy(a[j], T->n_inv, T->n_inv_shoup, m);
    if (v >= m) v -= m;
    a[j] = v;
  }
}

/* O(N^2) definition of the forward transform, for tests */
void orc_ntt_fwd_naive(const orc_ctx *c, int which, const uint64_t *a, uint64_t *out) {
  const orc_ntt_tables *T = tables(c, which);
  const uint64_t m = T->mod.value;
  for (uint32_t j = 0; j < T->N; ++j) {
    uint64_t e = 2 * orc_bit_reverse(j, T->logN) + 1;
    uint64_t x = orc_powmod(T->psi, e, m), pw = 1, acc = 0;
    for (uint32_t i = 0; i < T->N; ++i) {
      acc = (uint64_t)(((u128)acc + (u128)(a[i] % m) * pw) % m);
      pw = orc_mulmod(pw, x, m);
    }
    out[j] = acc;
  }
}

/* ---------------------------------------------------------------------------------- */
/* context                                                                            */
/* ---------------------------------------------------------------------------------- */

static uint32_t ceil_log2(uint64_t x) {
  uint32_t r = 0;
  while (((uint64_t)1 << r) < x) ++r;
  return r;
}

orc_ctx *orc_create(uint32_t N, uint32_t plain_bits, uint32_t l, uint32_t l_key, uint32_t fst,
                    uint32_t other) {
  if (N < 8 || (N & (N - 1)) || fst == 0 || other == 0 || (other & (other - 1))) return NULL;
  orc_ctx *c = calloc(1, sizeof(orc_ctx));
  c->N = N;
  c->logN = ceil_log2(N);
  uint64_t pr[2];
  if (orc_get_primes(2 * (uint64_t)N, 60, 2, pr) != 2) { free(c); return NULL; }
  /* Create() hands out the list from its back: data prime first */
  c->q = pr[1];
  c->p = pr[0];
  uint64_t sq;
  if (orc_get_primes(2 * (uint64_t)N, 27, 1, &sq) != 1) { free(c); return NULL; }
  c->small_q = sq;
  c->plain_bits = plain_bits;
  c->t = orc_generate_prime(plain_bits);
  c->l = l;
  c->l_key = l_key;
  /* base_log2 = ceil(log q / l) with log q = bit count of the data prime */
  uint32_t bits = 64 - (uint32_t)__builtin_clzll(c->q);
  c->base_log2 = (bits + l - 1) / l;
  c->base_log2_key = (bits + l_key - 1) / l_key;
  c->fst = fst;
  c->other = other;
  c->ndim = 1 + ceil_log2(other);
  c->u = fst + l * (c->ndim - 1);
  c->h = ceil_log2(c->u);
  c->num_pt = (uint64_t)fst * other;
  if (c->h > c->logN) { free(c); return NULL; }
  mod_init(&c->mq, c->q);
  mod_init(&c->mp, c->p);
  mod_init(&c->msq, c->small_q);
  ntt_init(&c->ntt_q, c->q, N, c->logN);
  ntt_init(&c->ntt_p, c->p, N, c->logN);
  ntt_init(&c->ntt_sq, c->small_q, N, c->logN);
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  ntt_free(&c->ntt_q);
  ntt_free(&c->ntt_p);
  ntt_free(&c->ntt_sq);
  free(c);
}
