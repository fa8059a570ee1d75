// serve.cpp -- serving loop of include/onionpir_serve.h: per-database request
// queues, one pool of worker threads for all databases doing continuous
// batching on the lanes of the database each picks (oldest waiting query
// first), admission by device-memory budget, latency statistics.
//
// The device work is entirely onion_gpu_answer_batch_lane (server.hip); this
// file is host scheduling only.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/onionpir_serve.h"

namespace {

using Clock = std::chrono::steady_clock;

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct Request {
  uint64_t ticket = 0;
  uint64_t client = 0;
  std::vector<uint64_t> query, resp;
  int status = 1;  // 1 pending, 0 done, < 0 error code
  std::string err;
  Clock::time_point t_submit, t_done;
};

// Latencies kept per database for the quantiles: the most recent kLatWindow
// completions (a ring), so a long-running loop holds at most 8 MiB each.
constexpr size_t kLatWindow = size_t(1) << 20;

struct Stats {
  uint64_t submitted = 0, completed = 0, failed = 0, batches = 0, batched = 0;
  std::vector<double> lat_ms;
  size_t lat_next = 0;  // next ring slot once lat_ms holds kLatWindow values
  bool any = false;
  Clock::time_point first, last;

  void add_latency(double ms) {
    if (lat_ms.size() < kLatWindow) {
      lat_ms.push_back(ms);
    } else {
      lat_ms[lat_next] = ms;
      lat_next = (lat_next + 1) % kLatWindow;
    }
  }
};

struct Db {
  onion_server *srv = nullptr;
  uint32_t cap = 1;
  size_t ct_words = 0;  // 2N
  std::deque<std::shared_ptr<Request>> queue;
  std::vector<uint8_t> lane_busy;
  Stats st;
};

}  // namespace

struct onion_serve {
  onion_serve_config cfg{};
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::vector<std::unique_ptr<Db>> dbs;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> tickets;
  std::vector<std::thread> workers;
  uint64_t next_ticket = 1;
  uint64_t budget_used = 0;
  bool stop = false;
};

namespace {

// The database whose oldest waiting query is the oldest of all, among those
// with a free lane (nullptr when none); *lane gets that free lane.  The
// caller holds sv->m.
Db *pick(onion_serve *sv, uint32_t *lane) {
  Db *best = nullptr;
  for (auto &D : sv->dbs) {
    if (D->queue.empty()) continue;
    auto it = std::find(D->lane_busy.begin(), D->lane_busy.end(), 0);
    if (it == D->lane_busy.end()) continue;
    if (!best || D->queue.front()->t_submit < best->queue.front()->t_submit) {
      best = D.get();
      *lane = (uint32_t)(it - D->lane_busy.begin());
    }
  }
  return best;
}

// A worker of the pool: take the picked database's waiting queries (up to
// its cap) as one batch, answer it on the free lane, publish the responses.
void worker(onion_serve *sv) {
  std::vector<uint64_t> ids, qbuf, rbuf;
  std::vector<std::shared_ptr<Request>> batch;
  const auto linger = std::chrono::microseconds(sv->cfg.linger_us);
  for (;;) {
    batch.clear();
    Db *D = nullptr;
    uint32_t lane = 0;
    {
      std::unique_lock<std::mutex> g(sv->m);
      for (;;) {
        if (sv->stop) return;
        D = pick(sv, &lane);
        if (!D) {
          sv->cv_work.wait(g);
          continue;
        }
        if (sv->cfg.linger_us && D->queue.size() < D->cap) {
          const auto deadline = D->queue.front()->t_submit + linger;
          if (Clock::now() < deadline) {
            sv->cv_work.wait_until(g, deadline);
            continue;  // pick again: more queries, another database, or the deadline
          }
        }
        break;
      }
      // a backlog above the cap: the cap, or with batch_split = 1 an equal
      // share of it (100 waiting at cap 64: 50 now, 50 next)
      size_t n = std::min<size_t>(D->cap, D->queue.size());
      onion_tuning tn;
      if (D->queue.size() > D->cap && onion_gpu_get_tuning(D->srv, &tn) == ONION_OK && tn.batch_split == 1) {
        const size_t k = (D->queue.size() + D->cap - 1) / D->cap;
        n = (D->queue.size() + k - 1) / k;
      }
      for (size_t i = 0; i < n; ++i) {
        batch.push_back(std::move(D->queue.front()));
        D->queue.pop_front();
      }
      D->lane_busy[lane] = 1;
    }
    const size_t ct = D->ct_words;
    const uint32_t n = (uint32_t)batch.size();
    ids.resize(n);
    qbuf.resize(n * ct);
    rbuf.resize(n * ct);
    for (uint32_t i = 0; i < n; ++i) {
      ids[i] = batch[i]->client;
      memcpy(qbuf.data() + i * ct, batch[i]->query.data(), ct * 8);
    }
    std::vector<int> rc(n, ONION_OK);
    std::vector<std::string> msg(n);
    int brc = onion_gpu_answer_batch_lane(D->srv, lane, n, ids.data(), qbuf.data(), rbuf.data(), nullptr);
    if (brc != ONION_OK) {
      // isolate the failing queries: answer each alone
      for (uint32_t i = 0; i < n; ++i) {
        rc[i] = onion_gpu_answer_batch_lane(D->srv, lane, 1, ids.data() + i, qbuf.data() + i * ct,
                                            rbuf.data() + i * ct, nullptr);
        if (rc[i] != ONION_OK) msg[i] = onion_gpu_last_error();
      }
    }
    const auto now = Clock::now();
    {
      std::lock_guard<std::mutex> g(sv->m);
      D->lane_busy[lane] = 0;
      D->st.batches += 1;
      D->st.batched += n;
      for (uint32_t i = 0; i < n; ++i) {
        Request &r = *batch[i];
        r.t_done = now;
        r.status = rc[i];
        if (rc[i] == ONION_OK) {
          r.resp.assign(rbuf.begin() + i * ct, rbuf.begin() + (i + 1) * ct);
          D->st.completed += 1;
          D->st.add_latency(std::chrono::duration<double, std::milli>(now - r.t_submit).count());
        } else {
          r.err = msg[i];
          D->st.failed += 1;
        }
        D->st.last = now;
      }
    }
    sv->cv_done.notify_all();
    sv->cv_work.notify_all();  // a lane is free again
  }
}

void merge(const Stats &a, Stats &into) {
  into.submitted += a.submitted;
  into.completed += a.completed;
  into.failed += a.failed;
  into.batches += a.batches;
  into.batched += a.batched;
  into.lat_ms.insert(into.lat_ms.end(), a.lat_ms.begin(), a.lat_ms.end());
  if (a.any) {
    into.first = into.any ? std::min(into.first, a.first) : a.first;
    into.last = into.any ? std::max(into.last, a.last) : a.last;
    into.any = true;
  }
}

double quantile(std::vector<double> &v, double q) {
  if (v.empty()) return 0.0;
  const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
  std::nth_element(v.begin(), v.begin() + k, v.end());
  return v[k];
}

}  // namespace

extern "C" {

const char *onion_serve_last_error(void) { return g_err.c_str(); }

int onion_serve_create(const onion_serve_config *cfg, onion_serve **out) {
  if (!cfg || !out) return fail(ONION_EINVAL, "null argument");
  if (cfg->max_batch < 1 || cfg->max_batch > 64) return fail(ONION_EINVAL, "max_batch must be in [1, 64]");
  if (cfg->lanes_per_db < 1 || cfg->lanes_per_db > 16) return fail(ONION_EINVAL, "lanes_per_db must be in [1, 16]");
  if (cfg->workers > 64) return fail(ONION_EINVAL, "workers must be in [0, 64]");
  onion_serve *sv = new onion_serve();
  sv->cfg = *cfg;
  for (uint32_t k = 0; k < cfg->workers; ++k) sv->workers.emplace_back(worker, sv);
  *out = sv;
  return ONION_OK;
}

int onion_serve_destroy(onion_serve *sv) {
  if (!sv) return ONION_OK;
  {
    std::lock_guard<std::mutex> g(sv->m);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  for (auto &t : sv->workers) t.join();
  {
    // queries still queued fail
    std::lock_guard<std::mutex> g(sv->m);
    for (auto &D : sv->dbs)
      for (auto &r : D->queue) {
        r->status = ONION_ESTATE;
        r->err = "serving loop stopped";
      }
  }
  sv->cv_done.notify_all();
  delete sv;
  return ONION_OK;
}

int onion_serve_add_db(onion_serve *sv, onion_server *srv, uint32_t *db_id, uint32_t *batch_cap) {
  if (!sv || !srv) return fail(ONION_EINVAL, "null argument");
  onion_param_info info;
  if (onion_gpu_param_info(srv, &info) != ONION_OK) return fail(ONION_EINVAL, onion_gpu_last_error());
  const uint32_t lanes = sv->cfg.lanes_per_db;
  std::vector<uint64_t> bytes(sv->cfg.max_batch + 1, 0);  // lane scratch by cap
  for (uint32_t c = 1; c <= sv->cfg.max_batch; ++c)
    if (onion_gpu_scratch_bytes(srv, c, &bytes[c]) != ONION_OK) return fail(ONION_EINVAL, onion_gpu_last_error());
  uint32_t cap = sv->cfg.max_batch;
  uint64_t reserved = 0;
  {
    // admission: claim the budget under the lock, allocate outside it
    std::lock_guard<std::mutex> g(sv->m);
    if (sv->stop) return fail(ONION_ESTATE, "serving loop stopped");
    for (auto &D : sv->dbs)
      if (D->srv == srv) return fail(ONION_EINVAL, "server already added");
    if (sv->cfg.scratch_budget) {
      const uint64_t left = sv->cfg.scratch_budget > sv->budget_used ? sv->cfg.scratch_budget - sv->budget_used : 0;
      while (cap > 0 && lanes * bytes[cap] > left) --cap;
      if (cap == 0)
        return fail(ONION_ENOMEM, "scratch budget: " + std::to_string(left) + " bytes left, " +
                                      std::to_string(lanes) + " lanes x " + std::to_string(bytes[1]) +
                                      " bytes for one query needed");
    }
    reserved = (uint64_t)lanes * bytes[cap];
    sv->budget_used += reserved;
  }
  int rc = onion_gpu_set_lanes(srv, lanes);
  for (uint32_t k = 0; rc == ONION_OK && k < lanes; ++k) rc = onion_gpu_reserve(srv, k, cap);
  std::lock_guard<std::mutex> g(sv->m);
  bool dup = false;
  for (auto &D : sv->dbs) dup |= D->srv == srv;
  if (rc != ONION_OK || dup || sv->stop) {
    sv->budget_used -= reserved;
    if (rc != ONION_OK) return fail(rc, onion_gpu_last_error());
    return dup ? fail(ONION_EINVAL, "server already added") : fail(ONION_ESTATE, "serving loop stopped");
  }
  auto D = std::make_unique<Db>();
  D->srv = srv;
  D->cap = cap;
  D->lane_busy.assign(lanes, 0);
  D->ct_words = (size_t)2 * info.poly_degree;
  sv->dbs.push_back(std::move(D));
  const uint32_t id = (uint32_t)sv->dbs.size() - 1;
  if (sv->cfg.workers == 0)  // one worker per lane of every database
    for (uint32_t k = 0; k < lanes; ++k) sv->workers.emplace_back(worker, sv);
  if (db_id) *db_id = id;
  if (batch_cap) *batch_cap = cap;
  return ONION_OK;
}

int onion_serve_submit(onion_serve *sv, uint32_t db_id, uint64_t client_id, const uint64_t *query,
                       uint64_t *ticket) {
  if (!sv || !query || !ticket) return fail(ONION_EINVAL, "null argument");
  auto r = std::make_shared<Request>();
  r->client = client_id;
  {
    std::lock_guard<std::mutex> g(sv->m);
    if (sv->stop) return fail(ONION_ESTATE, "serving loop stopped");
    if (db_id >= sv->dbs.size()) return fail(ONION_EINVAL, "unknown database " + std::to_string(db_id));
    Db *D = sv->dbs[db_id].get();
    r->query.assign(query, query + D->ct_words);
    r->ticket = sv->next_ticket++;
    r->t_submit = Clock::now();
    if (!D->st.any) {
      D->st.any = true;
      D->st.first = r->t_submit;
    }
    D->st.submitted += 1;
    sv->tickets[r->ticket] = r;
    D->queue.push_back(r);
    *ticket = r->ticket;
  }
  sv->cv_work.notify_all();
  return ONION_OK;
}

int onion_serve_wait(onion_serve *sv, uint64_t ticket, uint64_t *resp, uint32_t timeout_ms) {
  if (!sv || !resp) return fail(ONION_EINVAL, "null argument");
  std::unique_lock<std::mutex> g(sv->m);
  auto it = sv->tickets.find(ticket);
  if (it == sv->tickets.end()) return fail(ONION_EINVAL, "unknown ticket " + std::to_string(ticket));
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->status != 1; };
  if (timeout_ms == UINT32_MAX)
    sv->cv_done.wait(g, ready);
  else if (!sv->cv_done.wait_for(g, std::chrono::milliseconds(timeout_ms), ready))
    return fail(ONION_ETIMEDOUT, "ticket " + std::to_string(ticket) + " not ready");
  sv->tickets.erase(ticket);
  if (r->status != ONION_OK) return fail(r->status, r->err);
  memcpy(resp, r->resp.data(), r->resp.size() * 8);
  return ONION_OK;
}

int onion_serve_discard(onion_serve *sv, uint64_t ticket) {
  if (!sv) return fail(ONION_EINVAL, "null argument");
  std::lock_guard<std::mutex> g(sv->m);
  if (sv->tickets.erase(ticket) == 0) return fail(ONION_EINVAL, "unknown ticket " + std::to_string(ticket));
  return ONION_OK;
}

int onion_serve_stats_get(onion_serve *sv, uint32_t db_id, onion_serve_stats *out) {
  if (!sv || !out) return fail(ONION_EINVAL, "null argument");
  Stats s;
  uint32_t cap = 0;
  {
    std::lock_guard<std::mutex> g(sv->m);
    if (db_id == UINT32_MAX) {
      for (auto &D : sv->dbs) merge(D->st, s);
    } else {
      if (db_id >= sv->dbs.size()) return fail(ONION_EINVAL, "unknown database " + std::to_string(db_id));
      merge(sv->dbs[db_id]->st, s);
      cap = sv->dbs[db_id]->cap;
    }
  }
  onion_serve_stats o{};
  o.submitted = s.submitted;
  o.completed = s.completed;
  o.failed = s.failed;
  o.batches = s.batches;
  o.mean_batch = s.batches ? (double)s.batched / (double)s.batches : 0.0;
  double sum = 0;
  for (double x : s.lat_ms) sum += x;
  o.mean_ms = s.lat_ms.empty() ? 0.0 : sum / (double)s.lat_ms.size();
  o.max_ms = s.lat_ms.empty() ? 0.0 : *std::max_element(s.lat_ms.begin(), s.lat_ms.end());
  o.p50_ms = quantile(s.lat_ms, 0.50);
  o.p99_ms = quantile(s.lat_ms, 0.99);
  o.span_s = s.any && s.completed ? std::chrono::duration<double>(s.last - s.first).count() : 0.0;
  o.queries_per_s = o.span_s > 0 ? (double)s.completed / o.span_s : 0.0;
  o.batch_cap = cap;
  *out = o;
  return ONION_OK;
}

int onion_serve_reset_stats(onion_serve *sv) {
  if (!sv) return fail(ONION_EINVAL, "null argument");
  std::lock_guard<std::mutex> g(sv->m);
  for (auto &D : sv->dbs) D->st = Stats{};
  return ONION_OK;
}

}  // extern "C"
