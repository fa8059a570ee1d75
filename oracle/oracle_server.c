/*
 * oracle_server.c -- HE primitives and server stages of the OnionPIRv2 CPU
 * oracle.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Each function cites the reference lines it restates.  The arithmetic is the
 * reference's: exact uint64/uint128 modular arithmetic, uint128 accumulation
 * that wraps mod 2^128 in the first dimension, Barrett reduction at the end.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef orc_u128 u128;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static inline uint64_t add_mod(uint64_t a, uint64_t b, uint64_t m) {
  uint64_t s = a + b;
  return s >= m ? s - m : s;
}
static inline uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t m) {
  return a >= b ? a - b : a + m - b;
}

/* ---------------------------------------------------------------------------------- */
/* primitives                                                                         */
/* ---------------------------------------------------------------------------------- */

/* SEAL GaloisTool::apply_galois (coefficient form), reached from
 * Evaluator::apply_galois_inplace at src/server.cpp:409-411:
 * out[(i*g) mod N] = +/- in[i], negated when floor(i*g/N) is odd, -0 = 0. */
void orc_galois_coeff(const orc_ctx *c, const uint64_t *in, uint32_t g, uint64_t m,
                      uint64_t *out) {
  const uint64_t mask = c->N - 1;
  uint64_t raw = 0;
  for (uint32_t i = 0; i < c->N; ++i, raw += g) {
    uint64_t v = in[i];
    if ((raw >> c->logN) & 1) v = v ? m - v : 0;
    out[raw & mask] = v;
  }
}

/* utils::negacyclic_shift_poly_coeffmod with shift = (size_t)(-k)
 * (src/utils.cpp:6-26, called through shift_polynomial at src/server.cpp:422):
 * multiplication by x^-k in Z_q[x]/(x^N+1). */
void orc_shift_neg(const orc_ctx *c, const uint64_t *in, uint64_t k, uint64_t *out) {
  const uint64_t N = c->N, mask = N - 1;
  if (k % (2 * N) == 0) {
    memcpy(out, in, sizeof(uint64_t) * N);
    return;
  }
  uint64_t raw = (uint64_t)0 - k;
  for (uint64_t i = 0; i < N; ++i, ++raw) {
    uint64_t v = in[i];
    if ((raw & N) && v) v = c->q - v;
    out[raw & mask] = v;
  }
}

/* SEAL Evaluator::switch_key_inplace for one decomposition limb (rns=1) with
 * the special prime p (called from apply_galois_inplace, src/server.cpp:409).
 * target: the permuted c1 in coefficient form (values < q).
 * key: [comp 0: q-limb N, p-limb N][comp 1: q-limb N, p-limb N], NTT form.
 * ct[k] += ((X_q - ((X_p + floor(p/2)) mod p) + floor(p/2)) * p^-1) mod q,
 * X_m = INTT_m(NTT_m(target) * key[k][m]); this equals
 * floor((X + floor(p/2)) / p) mod q for X = CRT(X_q, X_p) in [0, qp). */
void orc_keyswitch_add(const orc_ctx *c, const uint64_t *target, const uint64_t *key,
                       uint64_t *ct) {
  const uint32_t N = c->N;
  const uint64_t q = c->q, p = c->p, half_p = p >> 1;
  const uint64_t half_p_mod_q = half_p % q;
  const uint64_t p_inv_q = orc_invmod(p % q, q);
  uint64_t *tq = orc_alloc(sizeof(uint64_t) * N), *tp = orc_alloc(sizeof(uint64_t) * N);
  uint64_t *xq = orc_alloc(sizeof(uint64_t) * N), *xp = orc_alloc(sizeof(uint64_t) * N);
  memcpy(tq, target, sizeof(uint64_t) * N);
  memcpy(tp, target, sizeof(uint64_t) * N); /* q < p: no reduction needed */
  orc_ntt_fwd(c, 0, tq);
  orc_ntt_fwd(c, 1, tp);
  for (int k = 0; k < 2; ++k) {
    const uint64_t *kq = key + (size_t)(2 * k) * N, *kp = key + (size_t)(2 * k + 1) * N;
    for (uint32_t i = 0; i < N; ++i) {
      u128 a = (u128)tq[i] * kq[i], b = (u128)tp[i] * kp[i];
      xq[i] = orc_barrett128(&c->mq, (uint64_t)a, (uint64_t)(a >> 64));
      xp[i] = orc_barrett128(&c->mp, (uint64_t)b, (uint64_t)(b >> 64));
    }
    orc_ntt_inv(c, 0, xq);
    orc_ntt_inv(c, 1, xp);
    uint64_t *dst = ct + (size_t)k * N;
    for (uint32_t i = 0; i < N; ++i) {
      uint64_t last = (xp[i] + half_p) % p;  /* rounding offset, mod p */
      uint64_t lq = last % q;                /* p > q: reduce into Z_q */
      uint64_t diff = sub_mod(add_mod(xq[i], half_p_mod_q, q), lq, q);
      dst[i] = add_mod(dst[i], orc_mulmod(diff, p_inv_q, q), q);
    }
  }
  free(tq); free(tp); free(xq); free(xp);
}

/* Evaluator::apply_galois_inplace on a coefficient-form BFV ciphertext:
 * permute both polynomials, keep sigma(c0), zero c1, key-switch sigma(c1). */
void orc_apply_galois(const orc_ctx *c, const uint64_t *ct, uint32_t g, const uint64_t *key,
                      uint64_t *out) {
  const uint32_t N = c->N;
  uint64_t *tmp = orc_alloc(sizeof(uint64_t) * N);
  orc_galois_coeff(c, ct, g, c->q, out);
  orc_galois_coeff(c, ct + N, g, c->q, tmp);
  memset(out + N, 0, sizeof(uint64_t) * N);
  orc_keyswitch_add(c, tmp, key, out);
  free(tmp);
}

/* GSWEval::decomp_rlwe_single_mod (src/gsw_eval.cpp:184-224): for poly in
 * {0,1}, for p = l-1 .. 0: digit = (x >> p*B) & (2^B - 1), stored at
 * index poly*l + (l-1-p). */
void orc_decompose(const orc_ctx *c, const uint64_t *ct, uint32_t l, uint32_t base_log2,
                   uint64_t *digits) {
  const uint32_t N = c->N;
  const uint64_t mask = ((uint64_t)1 << base_log2) - 1;
  for (uint32_t poly = 0; poly < 2; ++poly) {
    for (uint32_t pp = l; pp-- > 0;) {
      uint64_t *d = digits + (size_t)(poly * l + (l - 1 - pp)) * N;
      const uint32_t sh = pp * base_log2;
      for (uint32_t i = 0; i < N; ++i) d[i] = (ct[(size_t)poly * N + i] >> sh) & mask;
    }
  }
}

/* GSWEval::external_product (src/gsw_eval.cpp:33-106): decompose, NTT the 2l
 * digits, R_k = sum_j D_j * G_j[k] as uint128, Barrett-reduce; NTT-form out. */
void orc_external_product(const orc_ctx *c, const uint64_t *rgsw, uint32_t l, uint32_t base_log2,
                          const uint64_t *ct, uint64_t *out_ntt) {
  const uint32_t N = c->N, rows = 2 * l;
  uint64_t *dig = orc_alloc(sizeof(uint64_t) * (size_t)rows * N);
  u128 *acc = calloc(2 * (size_t)N, sizeof(u128));
  orc_decompose(c, ct, l, base_log2, dig);
  for (uint32_t j = 0; j < rows; ++j) orc_ntt_fwd(c, 0, dig + (size_t)j * N);
  for (uint32_t k = 0; k < 2; ++k) {
    for (uint32_t j = 0; j < rows; ++j) {
      const uint64_t *g = rgsw + (size_t)j * 2 * N + (size_t)k * N;
      const uint64_t *d = dig + (size_t)j * N;
      u128 *a = acc + (size_t)k * N;
      for (uint32_t i = 0; i < N; ++i) a[i] += (u128)d[i] * g[i];
    }
  }
  for (size_t i = 0; i < 2 * (size_t)N; ++i)
    out_ntt[i] = orc_barrett128(&c->mq, (uint64_t)acc[i], (uint64_t)(acc[i] >> 64));
  free(dig);
  free(acc);
}

/* ---------------------------------------------------------------------------------- */
/* server stages                                                                      */
/* ---------------------------------------------------------------------------------- */

/* PirServer::fast_expand_qry (src/server.cpp:382-428).  Heap slots 1..w-1 in
 * level order; a node whose left-most leaf is >= u is skipped.  Node i at
 * level k = 2^floor(log2 i): c' = Subs(c_i, N/k + 1); c_{2i} = c_i + c';
 * c_{2i+1} = (c_i - c') * x^-k.  Returns leaves w .. w+u-1 (u cts).
 * galois_keys: level a = log2(k) at offset a * 4N. */
void orc_expand(const orc_ctx *c, const uint64_t *query, const uint64_t *galois_keys,
                uint64_t *leaves) {
  const uint32_t N = c->N;
  const size_t ct_sz = 2 * (size_t)N;
  const uint64_t w = (uint64_t)1 << c->h, q = c->q;
  uint64_t *cts = calloc(2 * w * ct_sz, sizeof(uint64_t));
  uint64_t *cp = orc_alloc(sizeof(uint64_t) * ct_sz);
  memcpy(cts + ct_sz, query, sizeof(uint64_t) * ct_sz);
  for (uint64_t i = 1; i < w; ++i) {
    uint32_t lvl = 63 - (uint32_t)__builtin_clzll(i);
    uint64_t k = (uint64_t)1 << lvl;
    uint64_t left_leaf = i * w / k - w;
    if (left_leaf >= c->u) continue;
    uint64_t *ci = cts + i * ct_sz;
    orc_apply_galois(c, ci, (uint32_t)(N / k + 1), galois_keys + (size_t)lvl * 4 * N, cp);
    uint64_t *c2 = cts + 2 * i * ct_sz;
    for (size_t j = 0; j < ct_sz; ++j) c2[j] = add_mod(ci[j], cp[j], q);
    for (size_t j = 0; j < ct_sz; ++j) ci[j] = sub_mod(ci[j], cp[j], q);
    uint64_t *c3 = cts + (2 * i + 1) * ct_sz;
    orc_shift_neg(c, ci, k, c3);
    orc_shift_neg(c, ci + N, k, c3 + N);
  }
  memcpy(leaves, cts + w * ct_sz, sizeof(uint64_t) * ct_sz * c->u);
  free(cts);
  free(cp);
}

/* GSWEval::gsw_ntt_negacyclic_harvey on the client RGSW key
 * (PirServer::set_client_gsw_key, src/server.cpp:486-500): 2*l_key rows,
 * coefficient form in, NTT form out. */
void orc_rgsw_key_to_ntt(const orc_ctx *c, const uint64_t *rgsw_key_coeff, uint64_t *out) {
  const size_t n = (size_t)2 * c->l_key * 2 * c->N;
  memcpy(out, rgsw_key_coeff, sizeof(uint64_t) * n);
  for (size_t r = 0; r < (size_t)2 * c->l_key * 2; ++r) orc_ntt_fwd(c, 0, out + r * c->N);
}

/* make_query conversion loop (src/server.cpp:533-546) over
 * GSWEval::query_to_gsw (src/gsw_eval.cpp:254-286): for further dim d,
 * rows 0..l-1 = NTT(leaf[fst + d*l + k]), rows l..2l-1 =
 * external_product(RGSW_key (l_key, B_key), leaf).  rgsw_out holds
 * (ndim-1) RGSWs of 2l rows x 2N. */
void orc_convert(const orc_ctx *c, const uint64_t *leaves, const uint64_t *rgsw_key_ntt,
                 uint64_t *rgsw_out) {
  const uint32_t N = c->N, l = c->l;
  const size_t ct_sz = 2 * (size_t)N;
  for (uint32_t d = 0; d + 1 < c->ndim; ++d) {
    uint64_t *g = rgsw_out + (size_t)d * 2 * l * ct_sz;
    for (uint32_t k = 0; k < l; ++k) {
      const uint64_t *leaf = leaves + (size_t)(c->fst + d * l + k) * ct_sz;
      uint64_t *row = g + (size_t)k * ct_sz;
      memcpy(row, leaf, sizeof(uint64_t) * ct_sz);
      orc_ntt_fwd(c, 0, row);
      orc_ntt_fwd(c, 0, row + N);
      orc_external_product(c, rgsw_key_ntt, c->l_key, c->base_log2_key, leaf,
                           g + (size_t)(l + k) * ct_sz);
    }
  }
}

/* naive_level_mat_mat_128 / mat_mat_128 (src/matrix.cpp:79-97, 220-235) over
 * the prepared query (prep_query, src/server.cpp:80-118).  sel_ntt holds the
 * fst selection cts in NTT form.  For each slot: C[r][c] = sum_j A[r][j] *
 * B[j][c] in uint128 arithmetic (wrapping mod 2^128), plain stores.
 * inter: [slot][row][2]. */
void orc_first_dim_raw(const orc_ctx *c, const uint64_t *db, const uint64_t *sel_ntt,
                       orc_u128 *inter) {
  const uint32_t N = c->N, fst = c->fst, other = c->other;
  const size_t ct_sz = 2 * (size_t)N;
  uint64_t *qd = orc_alloc(sizeof(uint64_t) * (size_t)fst * ct_sz);
  /* prep_query: query_data[slot*(2*fst) + j*2 + poly] */
  for (uint32_t j = 0; j < fst; ++j)
    for (uint32_t s = 0; s < N; ++s) {
      qd[(size_t)s * 2 * fst + 2 * j] = sel_ntt[j * ct_sz + s];
      qd[(size_t)s * 2 * fst + 2 * j + 1] = sel_ntt[j * ct_sz + N + s];
    }
  for (uint32_t s = 0; s < N; ++s) {
    const uint64_t *A = db + (size_t)s * c->num_pt;
    const uint64_t *B = qd + (size_t)s * 2 * fst;
    u128 *C = inter + (size_t)s * 2 * other;
    for (uint32_t r = 0; r < other; ++r) {
      u128 t0 = 0, t1 = 0;
      const uint64_t *Ar = A + (size_t)r * fst;
      for (uint32_t k = 0; k < fst; ++k) {
        t0 += Ar[k] * (u128)B[2 * k];
        t1 += Ar[k] * (u128)B[2 * k + 1];
      }
      C[2 * r] = t0;
      C[2 * r + 1] = t1;
    }
  }
  free(qd);
}

/* PirServer::evaluate_first_dim (src/server.cpp:123-175) with delay_modulus
 * (src/server.cpp:178-293): NTT the fst selection cts, multiply, Barrett each
 * uint128 (x mod 2^128 mod q), INTT.  out: `other` cts in coefficient form. */
void orc_first_dim(const orc_ctx *c, const uint64_t *db, const uint64_t *sel_coeff,
                   uint64_t *out, orc_times *tm) {
  const uint32_t N = c->N, fst = c->fst, other = c->other;
  const size_t ct_sz = 2 * (size_t)N;
  uint64_t *sel = orc_alloc(sizeof(uint64_t) * fst * ct_sz);
  memcpy(sel, sel_coeff, sizeof(uint64_t) * fst * ct_sz);
  for (size_t j = 0; j < 2 * (size_t)fst; ++j) orc_ntt_fwd(c, 0, sel + j * N);
  u128 *inter = orc_alloc(sizeof(u128) * (size_t)N * 2 * other);
  double t0 = now_s();
  orc_first_dim_raw(c, db, sel, inter);
  if (tm) tm->first_dim_core = now_s() - t0;
  for (uint32_t r = 0; r < other; ++r) {
    for (uint32_t poly = 0; poly < 2; ++poly) {
      uint64_t *dst = out + (size_t)r * ct_sz + (size_t)poly * N;
      for (uint32_t s = 0; s < N; ++s) {
        u128 x = inter[(size_t)s * 2 * other + 2 * r + poly];
        dst[s] = orc_barrett128(&c->mq, (uint64_t)x, (uint64_t)(x >> 64));
      }
      orc_ntt_inv(c, 0, dst);
    }
  }
  free(inter);
  free(sel);
}

/* PirServer::other_dim_mux (src/server.cpp:295-334): for i < n/2 with
 * x = cts[i], y = cts[i + n/2]: y = y - x; y = ext(RGSW, y); y = INTT(y);
 * x = x + y.  The first n/2 cts hold the result. */
void orc_mux_level(const orc_ctx *c, uint64_t *cts, uint64_t n_cts, const uint64_t *rgsw) {
  const uint32_t N = c->N;
  const size_t ct_sz = 2 * (size_t)N;
  const uint64_t q = c->q, half = n_cts / 2;
  uint64_t *y = orc_alloc(sizeof(uint64_t) * ct_sz);
  for (uint64_t i = 0; i < half; ++i) {
    uint64_t *x = cts + i * ct_sz, *yy = cts + (i + half) * ct_sz;
    for (size_t j = 0; j < ct_sz; ++j) y[j] = sub_mod(yy[j], x[j], q);
    orc_external_product(c, rgsw, c->l, c->base_log2, y, y);
    orc_ntt_inv(c, 0, y);
    orc_ntt_inv(c, 0, y + N);
    for (size_t j = 0; j < ct_sz; ++j) x[j] = add_mod(x[j], y[j], q);
  }
  free(y);
}

/* PirServer::mod_switch_inplace (src/server.cpp:779-801), exactly as the
 * reference computes it: scale = (double)q' / (double)Q widened to long
 * double, x * scale in x87 extended precision, std::round, cast. */
uint64_t orc_mod_switch_one(uint64_t x, uint64_t q_big, uint64_t q_small) {
  const long double scale = (double)q_small / (double)q_big;
  return (uint64_t)roundl((long double)x * scale);
}

/* The same value computed with integers only (the form the device kernel
 * uses): scale = M * 2^E exactly (M < 2^53); the product x*M is rounded to a
 * 64-bit significand (round half to even, the x87 default), then
 * round-half-away-from-zero at the binary point 2^-E. */
uint64_t orc_mod_switch_one_int(uint64_t x, uint64_t q_big, uint64_t q_small) {
  if (x == 0) return 0;
  const double s = (double)q_small / (double)q_big;
  int e2;
  double fr = frexp(s, &e2);                 /* s = fr * 2^e2, fr in [0.5, 1) */
  uint64_t M = (uint64_t)ldexp(fr, 53);      /* exact 53-bit significand */
  int E = e2 - 53;                           /* s = M * 2^E */
  u128 P = (u128)x * M;
  int nb = 128 - (P >> 64 ? __builtin_clzll((uint64_t)(P >> 64)) + 0
                           : 64 + __builtin_clzll((uint64_t)P));
  if (nb > 64) {
    int sh = nb - 64;
    u128 keep = P >> sh, rem = P & ((((u128)1) << sh) - 1), halfv = ((u128)1) << (sh - 1);
    if (rem > halfv || (rem == halfv && (keep & 1))) keep += 1;
    P = keep << sh;
  }
  /* value = P * 2^E with E < 0: round half away from zero */
  int f = -E;
  if (f <= 0) return (uint64_t)(P << (-f));
  if (f >= 128) return 0;
  u128 r = (P + (((u128)1) << (f - 1))) >> f;
  return (uint64_t)r;
}

void orc_mod_switch(const orc_ctx *c, uint64_t *ct) {
  for (size_t i = 0; i < 2 * (size_t)c->N; ++i) ct[i] = orc_mod_switch_one(ct[i], c->q, c->small_q);
}

/* PirServer::save_resp_to_stream (src/server.cpp:581-633): every coefficient
 * of c0 then c1 masked to ceil(log2 small_q) bits, packed LSB-first. */
size_t orc_pack_response(const orc_ctx *c, const uint64_t *ct, uint8_t *out) {
  const uint32_t w = (uint32_t)ceil(log2((double)c->small_q));
  const uint64_t mask = ((uint64_t)1 << w) - 1;
  size_t nbytes = 0;
  uint32_t filled = 0;
  uint8_t buf = 0;
  for (size_t i = 0; i < 2 * (size_t)c->N; ++i) {
    uint64_t v = ct[i] & mask;
    uint32_t done = 0;
    while (done < w) {
      uint32_t room = 8 - filled, take = (w - done) < room ? (w - done) : room;
      buf |= (uint8_t)(((v >> done) & ((1u << take) - 1)) << filled);
      filled += take;
      done += take;
      if (filled == 8) { out[nbytes++] = buf; buf = 0; filled = 0; }
    }
  }
  if (filled) out[nbytes++] = buf;
  return nbytes;
}

/* PirClient::load_resp_from_stream (src/client.cpp:364-412); -1 when short */
int orc_unpack_response(const orc_ctx *c, const uint8_t *in, size_t len, uint64_t *ct) {
  const uint32_t w = (uint32_t)ceil(log2((double)c->small_q));
  size_t bit = 0;
  for (size_t i = 0; i < 2 * (size_t)c->N; ++i) {
    uint64_t v = 0;
    for (uint32_t j = 0; j < w; ++j, ++bit) {
      if (bit / 8 >= len) return -1;
      v |= (uint64_t)((in[bit / 8] >> (bit % 8)) & 1) << j;
    }
    ct[i] = v;
  }
  return 0;
}

/* PirServer::make_query (src/server.cpp:520-579) minus stream parsing:
 * expand -> convert -> first dim -> other dims -> mod switch.  resp: 2N. */
int orc_answer(const orc_ctx *c, const uint64_t *db, const uint64_t *galois_keys,
               const uint64_t *rgsw_key_ntt, const uint64_t *query, uint64_t *resp,
               orc_times *tm) {
  const uint32_t N = c->N;
  const size_t ct_sz = 2 * (size_t)N;
  orc_times local;
  if (!tm) tm = &local;
  memset(tm, 0, sizeof(*tm));
  double t_start = now_s(), t;
  uint64_t *leaves = orc_alloc(sizeof(uint64_t) * ct_sz * c->u);
  t = now_s();
  orc_expand(c, query, galois_keys, leaves);
  tm->expand = now_s() - t;
  uint64_t *rgsw = orc_alloc(sizeof(uint64_t) * ((size_t)c->ndim - 1 + 1) * 2 * c->l * ct_sz);
  t = now_s();
  orc_convert(c, leaves, rgsw_key_ntt, rgsw);
  tm->convert = now_s() - t;
  uint64_t *res = orc_alloc(sizeof(uint64_t) * ct_sz * c->other);
  t = now_s();
  orc_first_dim(c, db, leaves, res, tm);
  tm->first_dim = now_s() - t;
  t = now_s();
  uint64_t n = c->other;
  for (uint32_t d = 0; d + 1 < c->ndim; ++d) {
    orc_mux_level(c, res, n, rgsw + (size_t)d * 2 * c->l * ct_sz);
    n /= 2;
  }
  tm->other_dim = now_s() - t;
  t = now_s();
  orc_mod_switch(c, res);
  tm->mod_switch = now_s() - t;
  memcpy(resp, res, sizeof(uint64_t) * ct_sz);
  tm->total = now_s() - t_start;
  free(leaves);
  free(rgsw);
  free(res);
  return 0;
}

/* ---------------------------------------------------------------------------------- */
/* database                                                                           */
/* ---------------------------------------------------------------------------------- */

#define ORC_STREAM_DB 0x4442ULL

/* Seeded plaintext for database slot `pt`.
 * mode 0: coefficients uniform in [0, t): floor(r * t / 2^64), r = rand64.
 * mode 1: entry bytes -- an 8-byte big-endian entry index followed by random
 *         bytes (utils::generate_entry, src/utils.cpp:160-172), packed 16 bits
 *         per coefficient LSB-first (push_database_chunk, server.cpp:635-701). */
void orc_db_plaintext(const orc_ctx *c, uint64_t seed, int mode, uint64_t pt, uint64_t *coeffs) {
  const uint32_t N = c->N;
  if (mode == 0) {
    for (uint32_t i = 0; i < N; ++i) {
      uint64_t r = orc_rand64(seed, ORC_STREAM_DB, pt * N + i);
      coeffs[i] = (uint64_t)(((u128)r * c->t) >> 64);
    }
    return;
  }
  /* bits per coefficient = bit_count(t) - 1 (src/pir.cpp:97-99) */
  const uint32_t bpc = c->plain_bits - 1;
  const size_t entry_bytes = (size_t)bpc * N / 8;
  uint8_t *e = orc_alloc(entry_bytes);
  for (size_t b = 0; b < entry_bytes; b += 8) {
    uint64_t r = orc_rand64(seed, ORC_STREAM_DB, pt * N + b / 8);
    for (int k = 0; k < 8 && b + k < entry_bytes; ++k) e[b + k] = (uint8_t)(r >> (8 * k));
  }
  for (int k = 0; k < 8; ++k) e[7 - k] = (uint8_t)(pt >> (8 * k));
  u128 buf = 0;
  uint32_t off = 0, idx = 0;
  const u128 cmask = (((u128)1) << bpc) - 1;
  for (size_t b = 0; b < entry_bytes; ++b) {
    buf += (u128)e[b] << off;
    off += 8;
    while (off >= bpc) { coeffs[idx++] = (uint64_t)(buf & cmask); buf >>= bpc; off -= bpc; }
  }
  if (off > 0) coeffs[idx++] = (uint64_t)(buf & cmask);
  for (; idx < N; ++idx) coeffs[idx] = 0;
  free(e);
}

/* SEAL transform_to_ntt_inplace(Plaintext, parms_id) as used by
 * PirServer::preprocess_ntt (src/server.cpp:703-711): lift each coefficient
 * v >= (t+1)/2 to v + (q - t), then NTT mod q. */
void orc_plain_to_ntt(const orc_ctx *c, const uint64_t *coeffs, uint64_t *out) {
  const uint64_t thr = (c->t + 1) >> 1, inc = c->q - c->t;
  for (uint32_t i = 0; i < c->N; ++i) out[i] = coeffs[i] >= thr ? coeffs[i] + inc : coeffs[i];
  orc_ntt_fwd(c, 0, out);
}

/* gen_data -> preprocess_ntt -> realign_db (src/server.cpp:47-78, 703-748) for
 * plaintexts [pt_lo, pt_hi): db[slot * num_pt + pt] */
void orc_db_build_range(const orc_ctx *c, uint64_t seed, int mode, uint64_t pt_lo, uint64_t pt_hi,
                        uint64_t *db) {
  const uint32_t N = c->N;
  uint64_t *co = orc_alloc(sizeof(uint64_t) * N), *nt = orc_alloc(sizeof(uint64_t) * N);
  for (uint64_t pt = pt_lo; pt < pt_hi; ++pt) {
    orc_db_plaintext(c, seed, mode, pt, co);
    orc_plain_to_ntt(c, co, nt);
    for (uint32_t s = 0; s < N; ++s) db[(size_t)s * c->num_pt + pt] = nt[s];
  }
  free(co);
  free(nt);
}

void orc_db_build(const orc_ctx *c, uint64_t seed, int mode, uint64_t *db) {
  orc_db_build_range(c, seed, mode, 0, c->num_pt, db);
}
