/*
 * onionpir_serve.h -- serving loop over one or more device-resident databases.
 *
 * No reference counterpart: the reference answers one query per call on one
 * thread (PirServer::make_query, src/server.cpp:520-579, driven by
 * PirTest::test_pir, src/tests.cpp:77-101).  This layer keeps that per-query
 * result -- every response is bit-identical to onion_gpu_answer on the same
 * server, client and query -- and adds what a server answering many clients
 * needs:
 *
 *   - several databases (onion_server handles of any shape) served together;
 *   - one pool of worker threads for all databases, each driving a free lane
 *     of the server it picks (its own HIP stream and scratch), so batches
 *     overlap on the GPU; a free worker serves the database whose oldest
 *     waiting query is the oldest of all, so no database waits behind
 *     another's backlog;
 *   - continuous batching: a worker takes every waiting query of the
 *     database it picks, up to the database's batch cap, as one batch;
 *   - admission by device-memory budget: a database's batch cap is the largest
 *     that fits its lanes' scratch into what remains of the budget, and the
 *     scratch is allocated when the database is added;
 *   - latency statistics (submit to response ready): p50, p99, max, queries/s
 *     (quantiles, mean and max over each database's most recent 2^20
 *     completions).
 *
 * Thread safety: submit, wait and stats may be called from any threads.  A
 * query whose client has no keys on its database fails alone (ONION_ENOKEY);
 * the rest of its batch is answered.  Keys are managed on each onion_server
 * directly (onion_gpu_set_galois_key / set_rgsw_key), also while serving.
 */
#ifndef ONIONPIR_SERVE_H
#define ONIONPIR_SERVE_H

#include <stdint.h>

#include "onionpir_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ONION_ETIMEDOUT (-6) /* onion_serve_wait: the response is not ready yet */

typedef struct onion_serve_config {
  uint32_t max_batch;      /* largest batch one lane runs (1..64)                    */
  uint32_t lanes_per_db;   /* concurrent pipelines per database (1..16)              */
  uint64_t scratch_budget; /* device bytes of lane scratch over all databases;
                              0 = no limit beyond max_batch                        */
  uint32_t linger_us;      /* a database with fewer than its cap waiting is left
                              up to this long after its oldest arrival for more
                              (0 = launch at once)                                */
  uint32_t workers;        /* batches in flight over all databases (1..64); 0 =
                              one per lane of every database added             */
} onion_serve_config;

typedef struct onion_serve_stats {
  uint64_t submitted, completed, failed, batches;
  double mean_batch;           /* queries per batch                                */
  double p50_ms, p99_ms, max_ms, mean_ms; /* submit -> response ready            */
  double span_s;               /* first submit -> last completion                  */
  double queries_per_s;        /* completed / span_s                               */
  uint32_t batch_cap;          /* admitted batch cap (per database; 0 for all)     */
} onion_serve_stats;

typedef struct onion_serve onion_serve;

const char *onion_serve_last_error(void);

int onion_serve_create(const onion_serve_config *cfg, onion_serve **out);
/* Stops the workers (batches in flight finish) and fails queries still
 * queued.  Call it after every wait has returned.  Servers are not destroyed. */
int onion_serve_destroy(onion_serve *sv);

/* Serve `srv` (not owned; it must outlive sv and hold a complete database).
 * Sets its lane count to lanes_per_db and reserves the scratch of every lane
 * for the largest batch cap (<= max_batch) whose lanes fit the remaining
 * budget; ONION_ENOMEM when not even one query per lane fits.  *db_id
 * numbers the database for submit; *batch_cap gets the admitted cap. */
int onion_serve_add_db(onion_serve *sv, onion_server *srv, uint32_t *db_id, uint32_t *batch_cap);

/* Queue one query (2*N uint64, copied) of client_id against database db_id. */
int onion_serve_submit(onion_serve *sv, uint32_t db_id, uint64_t client_id, const uint64_t *query,
                       uint64_t *ticket);

/* Wait up to timeout_ms (UINT32_MAX: no limit; 0: poll) for a ticket.  On
 * ONION_OK the response (2*N uint64) is in resp and the ticket is retired; a
 * failed query returns its error code and retires the ticket too;
 * ONION_ETIMEDOUT leaves it pending. */
int onion_serve_wait(onion_serve *sv, uint64_t ticket, uint64_t *resp, uint32_t timeout_ms);

/* Retire a ticket without its response (a pending query still runs; its
 * result is dropped).  ONION_EINVAL for an unknown ticket. */
int onion_serve_discard(onion_serve *sv, uint64_t ticket);

/* Statistics of database db_id, or over all databases for db_id = UINT32_MAX. */
int onion_serve_stats_get(onion_serve *sv, uint32_t db_id, onion_serve_stats *out);
int onion_serve_reset_stats(onion_serve *sv);

#ifdef __cplusplus
}
#endif
#endif
