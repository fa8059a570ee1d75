/*
 * onionpir_gpu.h -- C ABI of the MI355X OnionPIR v2 server answer path.
 *
 * Plain integers and pointers only; no exceptions cross this boundary.  Every
 * function returns ONION_OK (0) or a negative ONION_E* code, and
 * onion_gpu_last_error() returns the message of the calling thread's last
 * failure.  Host buffers are borrowed for the duration of a call; the server
 * handle owns all device memory.  Batches on different lanes of one handle run
 * concurrently; calls on one lane are serialised; database, key, tuning and
 * lane changes wait for the batches in flight.
 *
 * Data formats (N = ring degree, one data limb):
 *   ciphertext      2*N uint64, c0 then c1, coefficient form, values < q
 *   Galois keys     h levels x 2 components x 2 limbs (q, p) x N uint64, NTT
 *                   form; level a belongs to Galois element 1 + (N >> a)
 *   RGSW key        2*l_key rows x [c0|c1] x N uint64, coefficient form
 *   response        2*N uint64 after the switch to small_q (values <= small_q),
 *                   or its bit-packed wire form (ceil(log2 small_q) bits per
 *                   coefficient, LSB-first)
 *   plaintexts      N uint64 coefficients < t per database entry slot
 *
 * Reference interface each entry point replaces (paths relative to the
 * reference repository root) is given beside it.
 */
#ifndef ONIONPIR_GPU_H
#define ONIONPIR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ONION_OK 0
#define ONION_EINVAL (-1)   /* bad argument or shape (std::invalid_argument)       */
#define ONION_ENOKEY (-2)   /* unknown client id (std::out_of_range from map::at)   */
#define ONION_ENOMEM (-3)   /* device memory budget exceeded                        */
#define ONION_EHIP (-4)     /* HIP runtime failure                                  */
#define ONION_ESTATE (-5)   /* call out of order (e.g. no database yet)             */

typedef struct onion_params {
  uint32_t poly_degree; /* N; the device path serves N = 2048                    */
  uint32_t plain_bits;  /* PlainMod: t = smallest prime above 2^(plain_bits-1)   */
  uint32_t gsw_l;       /* l of the data RGSW (GSW_L), 1..16                      */
  uint32_t gsw_l_key;   /* l of the key RGSW (GSW_L_KEY), 1..16                   */
  uint32_t fst_dim;     /* first dimension size                                   */
  uint32_t other_dim;   /* product of the further dimensions (power of two)      */
} onion_params;

/* Values derived from onion_params by the reference's rules. */
typedef struct onion_param_info {
  uint64_t q;        /* data prime (CoeffModulus::Create(N, {60, 60})[0])        */
  uint64_t p;        /* special key-switching prime ([1])                          */
  uint64_t t;        /* plaintext modulus                                          */
  uint64_t small_q;  /* response modulus (CoeffModulus::Create(N, {27, 60})[0])   */
  uint32_t base_log2, base_log2_key;
  uint32_t ndim;     /* 1 + log2(other_dim)                                        */
  uint32_t expan_height; /* h = ceil(log2(fst + l*(ndim-1)))                       */
  uint32_t useful_cnt;   /* u = fst + l*(ndim-1)                                   */
  uint64_t num_pt;
  uint32_t galois_key_words;  /* h * 4 * N                                         */
  uint32_t rgsw_key_words;    /* 2 * l_key * 2 * N                                 */
  uint32_t response_bytes;    /* packed response size                              */
  uint32_t poly_degree;       /* N (a ciphertext is 2N words)                      */
} onion_param_info;

/* Per-stage device time of the last answer, milliseconds, section names of the
 * reference's TimerLogger (src/includes/logging.h:27-65). */
typedef struct onion_stage_times {
  double expand, convert, first_dim, other_dim, mod_switch, total;
  double first_dim_core; /* the database-streaming kernel alone                   */
} onion_stage_times;

/* Launch-shape choices of the device path (no reference counterpart: the
 * reference fixes its loop shapes at compile time).  Every choice gives
 * bit-identical results; only speed differs. */
typedef struct onion_tuning {
  uint32_t fuse_min;     /* external products (conversion, mux levels) run one
                            workgroup per ciphertext, and an expansion level
                            one launch of one workgroup per node, when a
                            stage has at least this many ciphertexts across
                            the batch; below it every transform gets its own
                            workgroup.  A fused last expansion level writes
                            its leaves in NTT form, and the first dimension
                            and the conversion then skip their transforms   */
  uint32_t fd_rows_b1;   /* rows per workgroup of the single-query first-
                            dimension kernel (multiple of 32, 32..1024)       */
  uint32_t fd_tile_rows; /* thread rows of the batched first-dimension tile
                            (TR, 2 or 4)                                      */
  uint32_t fd_b1_variant; /* single-query first-dimension loads: bit 0 = 8
                             (else 4) 16-byte loads in flight per lane, bit 1
                             = nontemporal loads                             */
  uint32_t use_graph;     /* 1: each lane captures its per-batch pipeline in a
                             hipGraph once per batch size and replays it; 2
                             (the default): single queries only (one launch
                             instead of ~45 keeps the synchronous call's tail
                             short; batches of varying size launch eagerly);
                             0: never.  A replayed call's stage times hold
                             only the total.                                 */
  uint32_t fd_mfma_min;   /* batches of at least this many queries (and >= 2)
                             take the first dimension on the int8 matrix
                             cores (exact signed-digit products); 0 = never  */
  uint32_t fd_mfma_variant; /* 0: operands loaded to registers ahead of their
                               step; 1: direct-to-LDS loads three steps
                               ahead, shared by the workgroup's waves; 2: one
                               persistent workgroup per compute unit that
                               reduces each tile's outputs between the next
                               tile's products (needs the plane copy and
                               fst % 64 == 0, else 0 runs)                 */
  uint32_t db_planes;     /* 1: keep a second device copy of the database in
                             signed-digit planes (as many bytes again as the
                             database), rebuilt on every database write, so
                             the matrix-core first dimension reads operands
                             as stored; 0: free it and convert inside the
                             kernel.  Only for first dimensions that are a
                             multiple of 32.  The server starts with 1 when
                             the copy takes at most a third of the free
                             device memory, else with 0.                    */
  uint32_t ntt_radix;     /* pass schedule of the transforms in the fused
                             multiplexer levels (and onion_gpu_tap_ntt):
                             8 = radix-8 register passes (3 + 3 + 3 + 2
                             stages, three LDS exchanges per transform);
                             4 = radix-4 passes (2 + 2 + 2 + 2 + 3 stages,
                             four exchanges).  Bit-identical results.       */
  uint32_t batch_split;   /* a call with more queries than one pass takes
                             (64) runs them in chunks: 0 = chunks of 64 and
                             the rest (100 = 64 + 36); 1 = equal chunks
                             (100 = 50 + 50).  The serving loop splits a
                             backlog above a database's batch cap the same
                             way.                                           */
} onion_tuning;

typedef struct onion_server onion_server;

const char *onion_gpu_last_error(void);

/* Number of visible HIP devices (0 when none); no reference counterpart. */
int onion_gpu_device_count(int *count);

/* PirParams::PirParams (src/pir.cpp:29-95) and getters (src/includes/pir.h:
 * 29-68): derive moduli, gadget bases and shape without touching a device. */
int onion_params_derive(const onion_params *params, onion_param_info *out);

/* PirParams::PirParams + PirServer::PirServer (src/pir.cpp:29-95,
 * src/server.cpp:25-39): validate the shape, derive the moduli, allocate the
 * NTT-form database and the per-query scratch on the current HIP device. */
int onion_gpu_create(const onion_params *params, onion_server **out);
int onion_gpu_destroy(onion_server *srv);
int onion_gpu_param_info(const onion_server *srv, onion_param_info *out);

/* PirServer::gen_data (src/server.cpp:47-78): seeded database generated on the
 * device, lifted, NTT'd and realigned.  mode 0: coefficients uniform in
 * [0, t); mode 1: reference-style entries (8-byte big-endian index + random
 * bytes, 16 bits per coefficient). */
int onion_gpu_db_generate(onion_server *srv, uint64_t seed, int mode);

/* PirServer::push_database_chunk + preprocess_ntt + realign_db
 * (src/server.cpp:635-748): plaintexts [pt_lo, pt_lo + n_pt), coefficients
 * < t, N per plaintext.  The database is usable once every plaintext was
 * written (by this call or onion_gpu_db_generate). */
int onion_gpu_db_upload_plain(onion_server *srv, uint64_t pt_lo, uint64_t n_pt,
                              const uint64_t *coeffs);

/* Read back NTT-form database words db[slot * num_pt + pt] for one plaintext
 * (debug/parity tap). */
int onion_gpu_db_read_ntt(onion_server *srv, uint64_t pt, uint64_t *out_n);

/* PirServer::set_client_galois_key (src/server.cpp:480-484).  n_words = levels
 * x 4N with expan_height <= levels <= log2 N: a client may upload more levels
 * than this database needs, so that one upload serves every database that
 * shares the key store (onion_gpu_share_keys).  The device holds each key
 * word with its Shoup companion: 2 x n_words x 8 bytes per client. */
int onion_gpu_set_galois_key(onion_server *srv, uint64_t client_id, const uint64_t *keys,
                             size_t n_words);
/* PirServer::set_client_gsw_key (src/server.cpp:486-500): coefficient-form
 * rows, NTT'd on the device. */
int onion_gpu_set_rgsw_key(onion_server *srv, uint64_t client_id, const uint64_t *rows,
                           size_t n_words);
int onion_gpu_drop_client(onion_server *srv, uint64_t client_id);
/* Make dst use src's device key store (no reference counterpart: the reference
 * keeps keys per PirServer).  Both servers then see every key uploaded through
 * either; a key upload waits for the batches in flight on all of them.  dst
 * must hold no keys yet and use the same moduli and key gadget.  The store
 * lives until its last server is destroyed. */
int onion_gpu_share_keys(onion_server *dst, onion_server *src);

/* PirServer::make_query (src/server.cpp:520-579): one query ciphertext in,
 * the mod-switched response ciphertext out (2*N uint64). */
int onion_gpu_answer(onion_server *srv, uint64_t client_id, const uint64_t *query,
                     uint64_t *resp);

/* make_query for n queries of (possibly) different clients in one batched
 * pipeline; queries and resps hold n ciphertexts back to back. */
int onion_gpu_answer_batch(onion_server *srv, uint32_t n, const uint64_t *client_ids,
                           const uint64_t *queries, uint64_t *resps);

/* PirServer::save_resp_to_stream (src/server.cpp:581-633): bit-pack a
 * response ciphertext on the device; *n_bytes receives the size. */
int onion_gpu_pack_response(onion_server *srv, const uint64_t *resp, uint8_t *out,
                            size_t out_cap, size_t *n_bytes);

int onion_gpu_last_stage_times(const onion_server *srv, onion_stage_times *out);

int onion_gpu_get_tuning(const onion_server *srv, onion_tuning *out);
int onion_gpu_set_tuning(onion_server *srv, const onion_tuning *tuning);
/* Pick the tuning on the live database (no reference counterpart): times the
 * whole pipeline for each of the n_sizes batch sizes (1..256; above 64 in
 * chunks as tune.batch_split cuts them) with a synthetic client, host clock
 * from the first launch to completion, median of `reps` runs, and walks each
 * choice in turn keeping what lowers the sum over the sizes of each size's
 * time relative to the starting tuning's (so every size weighs alike).
 * Applies the result and writes it to *chosen (may be NULL).  Waits for
 * batches in flight. */
int onion_gpu_autotune(onion_server *srv, const uint32_t *batch_sizes, uint32_t n_sizes, uint32_t reps,
                       onion_tuning *chosen);

/* Concurrent pipelines (no reference counterpart: the reference server is
 * single-threaded).  A server owns n_lanes lanes; each lane has its own HIP
 * stream and scratch, so batches on different lanes overlap on the device.
 * Lane 0 serves every other entry point of this header.  Key and database
 * changes wait for batches in flight on any lane. */
int onion_gpu_set_lanes(onion_server *srv, uint32_t n_lanes);
int onion_gpu_answer_batch_lane(onion_server *srv, uint32_t lane, uint32_t n, const uint64_t *client_ids,
                                const uint64_t *queries, uint64_t *resps, onion_stage_times *times);
/* onion_gpu_answer_batch without allocating (no reference counterpart; for
 * callers that cannot allocate device memory: tight serving loops and
 * harnesses that forbid hipMalloc).  The per-query scratch comes from a
 * 256 MiB static device array that every GPU the library runs on holds from
 * its first use until the process ends, instead of from lane buffers; the
 * pipeline runs eagerly (never as a captured graph, whose instantiation
 * allocates).  n queries must fit the array: ONION_EINVAL otherwise (at
 * 512x1024 one query, about 195 MiB).  Calls on one GPU share the array and
 * run one at a time within the process; the array keeps the last call's
 * contents, which no call reads.  Results are identical to
 * onion_gpu_answer_batch. */
int onion_gpu_answer_batch_noalloc(onion_server *srv, uint32_t n, const uint64_t *client_ids,
                                   const uint64_t *queries, uint64_t *resps);
/* Device bytes of per-query scratch a lane allocates for each query of its
 * largest batch (a 64-query lane's scratch divided by 64). */
int onion_gpu_scratch_bytes_per_query(const onion_server *srv, uint64_t *bytes);
/* Device bytes of the scratch one lane allocates for batches of up to cap
 * queries (1..64; the unit of admission by device-memory budget). */
int onion_gpu_scratch_bytes(const onion_server *srv, uint32_t cap, uint64_t *bytes);
/* Allocate lane `lane`'s scratch for batches of up to n_queries (capped at 64)
 * now rather than on first use; fails with ONION_ENOMEM when it does not fit. */
int onion_gpu_reserve(onion_server *srv, uint32_t lane, uint32_t n_queries);

/* ---- stage taps: each runs one stage of the device path on host-given
 * inputs, for parity tests against the CPU oracle -------------------------- */

/* fast_expand_qry (src/server.cpp:382-428): u leaves, coefficient form. */
int onion_gpu_tap_expand(onion_server *srv, uint64_t client_id, const uint64_t *query,
                         uint64_t *leaves);
/* query_to_gsw for every further dimension (src/server.cpp:533-546):
 * (ndim-1) RGSWs x 2l rows x 2N, NTT form, from the u leaves. */
int onion_gpu_tap_convert(onion_server *srv, uint64_t client_id, const uint64_t *leaves,
                          uint64_t *rgsw);
/* evaluate_first_dim (src/server.cpp:123-175): fst selection cts in
 * coefficient form -> other cts in coefficient form. */
int onion_gpu_tap_first_dim(onion_server *srv, const uint64_t *sel, uint64_t *out);
/* evaluate_first_dim for nq (1..64) queries at once through the batched
 * kernels: nq x fst selection cts in, nq x other cts out (coefficient form). */
int onion_gpu_tap_first_dim_batch(onion_server *srv, uint32_t nq, const uint64_t *sel, uint64_t *out);
/* other_dim_mux (src/server.cpp:295-334): n cts and one RGSW (2l rows x 2N,
 * NTT form) -> n/2 cts. */
int onion_gpu_tap_mux(onion_server *srv, const uint64_t *cts, uint32_t n, const uint64_t *rgsw,
                      uint64_t *out);
/* SEAL NTT / INTT used on the path: which = 0 (q) or 1 (p). */
int onion_gpu_tap_ntt(onion_server *srv, int which, int inverse, uint64_t *polys, uint32_t n);
/* apply_galois_inplace with the client's key of level a (src/server.cpp:409). */
int onion_gpu_tap_galois(onion_server *srv, uint64_t client_id, uint32_t level,
                         const uint64_t *ct, uint64_t *out);
/* mod_switch_inplace (src/server.cpp:779-801) on 2N values. */
int onion_gpu_tap_mod_switch(onion_server *srv, const uint64_t *ct, uint64_t *out);
/* Plain device copy of n_bytes (multiple of 16) from the database to scratch;
 * no reference counterpart: calibrates profiler byte counters.  *copied gets
 * the size actually copied (clamped to the scratch size). */
int onion_gpu_tap_copy(onion_server *srv, uint64_t n_bytes, uint64_t *copied);
/* VALU issue-rate probe: `blocks` workgroups of 256 threads run `rounds`
 * radix-8 passes of the transforms' butterflies on registers (no memory
 * traffic, LDS or barriers); *ms gets the kernel time.  No reference
 * counterpart: it measures the instruction-rate ceiling that the
 * transform-bound kernels are compared against (blocks = 256 w gives w
 * waves per SIMD). */
int onion_gpu_tap_valu_probe(onion_server *srv, uint32_t blocks, uint32_t rounds, double *ms);

#ifdef __cplusplus
}
#endif
#endif
