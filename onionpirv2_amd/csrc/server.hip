// server.hip -- the C ABI of include/onionpir_gpu.h: device-resident
// database, per-client key store, and the answer pipeline
//   expand -> convert -> first dimension -> other dimensions -> mod switch
// (PirServer::make_query, src/server.cpp:520-579) as a chain of HIP kernels.
// Every stage is batched: grid.y indexes the query, each query has its own
// scratch slice and its own client keys, and the first dimension reads the
// database once for the whole batch.
//
// Concurrency: a server owns one or more lanes.  A lane is a pipeline -- its
// own HIP stream, events, per-query scratch and (optionally) the hipGraphs of
// its pipeline, one per batch size -- so batches on different lanes overlap on
// the device while they share the database, the tables and the key store.
// A writer-preferring gate orders the two kinds of call: batches hold it
// shared, database / key / tuning / lane changes hold it exclusively.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/onionpir_gpu.h"
#include "host_math.hpp"
#include "k_db.hpp"
#include "k_expand.hpp"
#include "k_first_dim.hpp"
#include "k_first_dim_mfma.hpp"
#include "k_gsw.hpp"
#include "k_ntt.hpp"
#include "k_tail.hpp"
#include "params.hpp"

using namespace onion;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) return fail(ONION_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct ClientKeys {
  uint64_t *gk = nullptr;  // levels x kKeyLevelWords: NTT-form key, then Shoup companions
  uint64_t *rk = nullptr;  // 2 l_key x 2N, NTT form
  uint32_t gk_levels = 0;
};

constexpr uint32_t kMaxBatch = 64;  // queries per pipeline pass
constexpr uint32_t kMaxLanes = 16;

// Shared/exclusive gate in which a waiting writer goes ahead of readers that
// arrive after it, so a stream of overlapping batches cannot starve a key
// upload.
class Gate {
  std::mutex m_;
  std::condition_variable cv_;
  int readers_ = 0, writers_waiting_ = 0;
  bool writer_ = false;

 public:
  void lock_shared() {
    std::unique_lock<std::mutex> g(m_);
    cv_.wait(g, [&] { return !writer_ && writers_waiting_ == 0; });
    ++readers_;
  }
  void unlock_shared() {
    std::lock_guard<std::mutex> g(m_);
    if (--readers_ == 0) cv_.notify_all();
  }
  void lock() {
    std::unique_lock<std::mutex> g(m_);
    ++writers_waiting_;
    cv_.wait(g, [&] { return !writer_ && readers_ == 0; });
    --writers_waiting_;
    writer_ = true;
  }
  void unlock() {
    std::lock_guard<std::mutex> g(m_);
    writer_ = false;
    cv_.notify_all();
  }
};

void dfree_any(void *p) {
  if (p) (void)hipFree(p);
}

}  // namespace

// Device key store: the keys of every client, shared by all servers (databases)
// that use it.  A client's Galois key of level a belongs to Galois element
// 1 + (N >> a) whatever the database shape, so one upload with enough levels
// serves databases of every expansion height up to that many levels.  Batches
// hold the store's gate shared while their kernels may read keys; uploads and
// drops hold it exclusively.
struct KeyStore {
  std::map<uint64_t, ClientKeys> clients;
  Gate gate;
  uint64_t q = 0, p = 0;
  uint32_t l_key = 0;
  ~KeyStore() {
    for (auto &kv : clients) {
      dfree_any(kv.second.gk);
      dfree_any(kv.second.rk);
    }
  }
};

// One pipeline: stream, events, per-query scratch for `cap` queries, the
// device arrays of per-query key pointers, and the graphs captured for it.
struct Lane {
  // one stream: no stage waits on another stream's event
  hipStream_t st = nullptr;
  hipEvent_t ev[8]{};
  uint32_t cap = 0;
  uint64_t *trees = nullptr, *rgsw = nullptr, *fd = nullptr, *digits = nullptr;
  uint32_t *Qs = nullptr;  // N x fst x cpad x 2 uint32 (cpad for `cap` queries)
  const uint64_t **gk_ptrs = nullptr, **rk_ptrs = nullptr;  // device [cap]
  std::vector<const uint64_t *> h_gk, h_rk;
  std::map<uint32_t, hipGraphExec_t> graphs;  // by batch size
  uint64_t graph_epoch = 0;                   // tuning epoch the graphs were captured under
  std::mutex mu;                              // one batch at a time per lane
  bool borrowed = false;  // buffers are carved from the static scratch, never freed
  // pinned host staging of `cap` queries and responses (2N words each), so the
  // copies run as direct DMA instead of through the runtime's pageable path;
  // none on a borrowed lane, which allocates nothing
  uint64_t *hq = nullptr, *hr = nullptr;
  // a call returned (on an error) after queueing copies that read or write
  // hq / hr and before synchronising: the next call waits for the stream first
  bool in_flight = false;
};

struct onion_server {
  onion_params P{};
  onion_param_info I{};
  int device = 0;
  uint32_t n_cu = 0;  // compute units of the device
  uint64_t w = 0;  // 2^h
  // device NTT tables: q then p, each rp rps irp irps (4N words); then per
  // expansion level a, x^-k (k = 2^a) as slot factors and their Shoup
  // companions (2N words each)
  uint64_t *tabs = nullptr;
  const uint64_t *xk = nullptr;
  NttTab Tq{}, Tp{};
  Mod mq{}, mp{};
  KsConst ks{};
  LiftConst lift{};
  // the stored RGSW(s) keys are folded (k_rgsw_fold_last) for the conversion;
  // fixed by q, p and l_key, which every server sharing a key store agrees on
  bool conv_fold = false;
  std::vector<uint32_t> gal, ginv;  // each level's Galois element N/2^a + 1, its inverse mod 2N
  uint64_t ms_M = 0;
  uint32_t ms_f = 0;
  // database and its signed-digit plane copy (tune.db_planes)
  uint64_t *db = nullptr;
  uint8_t *dbp = nullptr;
  std::vector<uint8_t> pt_written;
  uint64_t n_written = 0;
  // client keys (possibly shared with other servers)
  std::shared_ptr<KeyStore> keys;
  // per-query scratch strides in words
  uint64_t tree_words = 0, rgsw_words = 0, fd_words = 0, digit_words = 0;
  std::vector<std::unique_ptr<Lane>> lanes;  // lane 0 also serves taps and ingest
  std::unique_ptr<Lane> slane;               // lane of onion_gpu_answer_batch_noalloc
  uint64_t *io = nullptr;                    // staging for taps and ingest (lane 0)
  size_t io_words = 0;
  uint8_t *pack = nullptr;
  // measured on MI355X at 512x1024: 8 nontemporal 16-byte loads in flight per
  // lane and 256-row blocks stream the database at 94% of a plain read kernel;
  // fuse_min 512 runs a single query's first multiplexer level (512 pairs)
  // fused on the first dimension's NTT-form outputs, 1.4% faster than split
  // (batches of 64 take the same time either way)
  onion_tuning tune{512, 256, 4, 3, 2, 2, 2, 1, 8, 0};
  uint64_t tune_epoch = 1;
  mutable Gate gate;  // tuning reads take it shared (onion_gpu_get_tuning)
  mutable std::mutex last_mu;
  onion_stage_times last{};
};

namespace {

using Shared = std::shared_lock<Gate>;
using Exclusive = std::unique_lock<Gate>;

template <typename T>
int dalloc(T **p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? ONION_ENOMEM : ONION_EHIP,
                std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                    " bytes failed: " + hipGetErrorString(e));
  }
  return ONION_OK;
}

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

// Pinned (page-locked) host memory, allocated as such (hipHostMalloc).
template <typename T>
int halloc(T **p, size_t count) {
  hipError_t e = hipHostMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(ONION_ENOMEM, std::string("hipHostMalloc of ") + std::to_string(count * sizeof(T)) +
                                  " bytes failed: " + hipGetErrorString(e));
  }
  return ONION_OK;
}

uint32_t cpad_for(uint32_t nq) {
  if (nq <= 1) return 2;
  if (nq <= 16) return 32;
  if (nq <= 32) return 64;
  return 128;
}

// uint32 words of a lane's selection buffer for `cap` queries: limbs for one
// query, digit chunks for batches (8 bytes per slot, j and column either way).
uint64_t qs_words(const onion_server *s, uint32_t cap) {
  return (uint64_t)kN * s->P.fst_dim * cpad_for(cap) * 2;
}

void drop_graphs(Lane *L) {
  for (auto &kv : L->graphs) (void)hipGraphExecDestroy(kv.second);
  L->graphs.clear();
}

void free_scratch(Lane *L) {
  drop_graphs(L);
  for (uint64_t *h : {L->hq, L->hr})
    if (h) (void)hipHostFree(h);
  L->hq = L->hr = nullptr;
  void *bufs[] = {L->trees, L->rgsw, L->fd, L->digits, L->Qs, L->gk_ptrs, L->rk_ptrs};
  if (!L->borrowed)
    for (void *b : bufs) dfree(b);
  L->trees = L->rgsw = L->fd = L->digits = nullptr;
  L->Qs = nullptr;
  L->gk_ptrs = L->rk_ptrs = nullptr;
  L->cap = 0;
}

// Scratch of lane L for `cap` queries (kept when large enough).
int ensure_scratch(const onion_server *s, Lane *L, uint32_t cap) {
  if (L->cap >= cap) return ONION_OK;
  (void)hipStreamSynchronize(L->st);
  free_scratch(L);
  int rc;
  if ((rc = dalloc(&L->trees, cap * s->tree_words)) || (rc = dalloc(&L->rgsw, cap * s->rgsw_words)) ||
      (rc = dalloc(&L->fd, cap * s->fd_words)) || (rc = dalloc(&L->digits, cap * s->digit_words)) ||
      (rc = dalloc(&L->Qs, qs_words(s, cap))) || (rc = dalloc(&L->gk_ptrs, cap)) ||
      (rc = dalloc(&L->rk_ptrs, cap)) || (rc = halloc(&L->hq, (size_t)cap * 2 * kN)) ||
      (rc = halloc(&L->hr, (size_t)cap * 2 * kN))) {
    std::string msg = g_err;
    free_scratch(L);
    return fail(rc, msg + " (batch scratch for " + std::to_string(cap) + " queries)");
  }
  L->cap = cap;
  L->h_gk.assign(cap, nullptr);
  L->h_rk.assign(cap, nullptr);
  return ONION_OK;
}

void destroy_lane(Lane *L) {
  if (L->st) (void)hipStreamSynchronize(L->st);
  free_scratch(L);
  for (auto &e : L->ev)
    if (e) (void)hipEventDestroy(e);
  if (L->st) (void)hipStreamDestroy(L->st);
  L->st = nullptr;
}

int make_lane(std::unique_ptr<Lane> &out) {
  auto L = std::make_unique<Lane>();
  bool ok = hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking) == hipSuccess;
  for (auto &e : L->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  if (!ok) {
    destroy_lane(L.get());
    return fail(ONION_EHIP, "lane stream or event creation failed");
  }
  out = std::move(L);
  return ONION_OK;
}

// Device bytes ensure_scratch allocates for a lane of `cap` queries.
uint64_t scratch_bytes(const onion_server *s, uint32_t cap) {
  const uint64_t words = s->tree_words + s->rgsw_words + s->fd_words + s->digit_words;
  return cap * (words * 8 + 2 * sizeof(void *)) + qs_words(s, cap) * 4;
}

// Device bytes of lane scratch per query of a kMaxBatch-query lane (the
// selection tile Qs is shared by the batch and counted pro rata).
uint64_t scratch_bytes_per_query(const onion_server *s) {
  return (scratch_bytes(s, kMaxBatch) + kMaxBatch - 1) / kMaxBatch;
}

// ---- the static scratch of onion_gpu_answer_batch_noalloc ---------------------------
//
// A file-scope device array: every GPU the library runs on holds its own copy
// from the library's first use on that GPU until the process ends.  Host code
// never passes `scratch` itself (in host code the name is a stand-in, not a
// device address): onion_gpu_create resolves the array's address on its
// device once with hipGetSymbolAddress and the no-allocation entry point uses
// that address.  Nothing in it needs to start at zero: the pipeline writes
// every word before it reads it, as it does in freshly allocated lane scratch.
constexpr uint64_t kStaticScratchBytes = uint64_t(256) << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {

std::mutex g_static_mu[kMaxDevices];      // one no-allocation call at a time per device
void *g_static_ptr[kMaxDevices] = {};     // scratch's address on each device
std::string g_static_err[kMaxDevices];    // why it could not be resolved
bool g_static_tried[kMaxDevices] = {};
std::mutex g_static_init_mu;

// Resolve scratch's address on device `dev` (current) once per device; a
// failure is kept for the no-allocation entry point to report and does not
// stop the server's other entry points.
void static_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_static_init_mu);
  if (g_static_tried[dev]) return;
  g_static_tried[dev] = true;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_static_err[dev] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    return;
  }
  g_static_ptr[dev] = p;
}

// Bytes of static scratch a batch of n queries takes (each buffer 256-byte
// aligned): the lane buffers of ensure_scratch for cap = n.
uint64_t static_bytes(const onion_server *s, uint32_t n) {
  auto al = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
  return al(n * s->tree_words * 8) + al(n * s->rgsw_words * 8) + al(n * s->fd_words * 8) +
         al(n * s->digit_words * 8) + al(qs_words(s, n) * 4) +
         2 * al(n * sizeof(void *));
}

// Point the static lane's buffers at the static scratch for n queries.
void carve_static(const onion_server *s, Lane *L, uint32_t n, uint8_t *base) {
  auto take = [&](uint64_t bytes) {
    uint8_t *p = base;
    base += (bytes + 255) & ~uint64_t(255);
    return p;
  };
  L->trees = (uint64_t *)take(n * s->tree_words * 8);
  L->rgsw = (uint64_t *)take(n * s->rgsw_words * 8);
  L->fd = (uint64_t *)take(n * s->fd_words * 8);
  L->digits = (uint64_t *)take(n * s->digit_words * 8);
  L->Qs = (uint32_t *)take(qs_words(s, n) * 4);
  L->gk_ptrs = (const uint64_t **)take(n * sizeof(void *));
  L->rk_ptrs = (const uint64_t **)take(n * sizeof(void *));
  L->cap = n;
}

// The caller holds the key store's gate (shared or exclusive).
int lookup(const onion_server *s, uint64_t id, ClientKeys &out, bool need_rk) {
  auto it = s->keys->clients.find(id);
  if (it == s->keys->clients.end() || !it->second.gk)
    return fail(ONION_ENOKEY, "no Galois keys for client " + std::to_string(id));
  if (it->second.gk_levels < s->I.expan_height)
    return fail(ONION_ENOKEY, "Galois keys of client " + std::to_string(id) + " cover " +
                                  std::to_string(it->second.gk_levels) + " levels; this database needs " +
                                  std::to_string(s->I.expan_height));
  if (need_rk && !it->second.rk)
    return fail(ONION_ENOKEY, "no RGSW key for client " + std::to_string(id));
  out = it->second;
  return ONION_OK;
}

// Resolve the keys of n clients into lane L's device pointer arrays.
int bind_keys(const onion_server *s, Lane *L, uint32_t n, const uint64_t *ids, bool need_rk) {
  for (uint32_t i = 0; i < n; ++i) {
    ClientKeys ck;
    int rc = lookup(s, ids[i], ck, need_rk);
    if (rc) return rc;
    L->h_gk[i] = ck.gk;
    L->h_rk[i] = ck.rk;
  }
  HIPCHK(hipMemcpyAsync(L->gk_ptrs, L->h_gk.data(), n * sizeof(void *), hipMemcpyHostToDevice, L->st));
  HIPCHK(hipMemcpyAsync(L->rk_ptrs, L->h_rk.data(), n * sizeof(void *), hipMemcpyHostToDevice, L->st));
  return ONION_OK;
}

// Bytes of the digit-plane copy: rows padded to a multiple of 32.
uint64_t planes_bytes(const onion_server *s) {
  return (uint64_t)kN * ((s->P.other_dim + 31) / 32 * 32) * s->P.fst_dim * 8;
}

// Rebuild the plane chunks that cover plaintexts pt_lo .. pt_lo + n - 1 of
// every slot from the database words (stream st; the caller holds the gate
// exclusively).  The unit is a record of 16 plaintexts of one row (fst % 32
// == 0 whenever the copy exists, so records never straddle rows).
void planes_update(const onion_server *s, hipStream_t st, uint64_t pt_lo, uint64_t n) {
  if (!s->dbp || !n) return;
  const uint64_t lo = pt_lo / 16, hi = (pt_lo + n + 15) / 16;
  k_db_planes<<<dim3((uint32_t)((hi - lo + 255) / 256), kN), 256, 0, st>>>(s->db, s->dbp, s->P.fst_dim,
                                                                           s->P.other_dim, lo, hi);
}

// Allocate (and fill) or free the plane copy to match tune.db_planes.  The
// caller holds the gate exclusively, so no batch is in flight.
int sync_planes(onion_server *s) {
  const bool want = s->tune.db_planes && s->P.fst_dim % 32 == 0;
  if (want == (s->dbp != nullptr)) return ONION_OK;
  hipStream_t st = s->lanes[0]->st;
  if (!want) {
    HIPCHK(hipStreamSynchronize(st));
    dfree(s->dbp);
    s->dbp = nullptr;
    return ONION_OK;
  }
  int rc;
  if ((rc = dalloc(&s->dbp, planes_bytes(s)))) return rc;
  if (s->n_written) {  // plaintexts written later convert as they arrive
    planes_update(s, st, 0, s->I.num_pt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
  }
  return ONION_OK;
}

int check_db(const onion_server *s) {
  if (!s->db || s->n_written != s->I.num_pt)
    return fail(ONION_ESTATE, "database incomplete: " + std::to_string(s->n_written) + " of " +
                                  std::to_string(s->I.num_pt) + " plaintexts written");
  return ONION_OK;
}

// ---- pipeline stages (nq queries; query b's ciphertext at tree b, slot 1) ------------

// Nodes 2^a .. 2^a + n_active - 1 of every query's tree; X scratch in digits.
// Levels with at least tune.fuse_min nodes across the batch run fused and,
// with ntt_out, write NTT-form children (k_expand_ntt); ntt_in: the nodes
// are in NTT form.  Returns whether the children are in NTT form.
bool expand_level(const onion_server *s, Lane *L, uint32_t nq, uint32_t a, uint32_t n_active, bool ntt_in,
                  bool ntt_out) {
  const uint64_t k = uint64_t(1) << a;
  const bool fused = (uint64_t)n_active * nq >= s->tune.fuse_min;
  if (ntt_in && !(fused && ntt_out)) {  // back to coefficient form (the fused levels are the lower ones)
    k_cts_to_coeff<<<dim3(2 * n_active, nq), kNttThreads, 0, L->st>>>(L->trees + k * 2 * kN, s->tree_words, s->Tq);
    ntt_in = false;
  }
  if (fused && ntt_out) {
    const uint64_t *xk = s->xk + 2 * (size_t)a * kN;
    if (ntt_in)
      k_expand_ntt<true><<<dim3(n_active, nq), kNttThreads, 0, L->st>>>(
          L->trees, s->tree_words, L->gk_ptrs, a, s->gal[a], s->ginv[a], xk, s->Tq, s->Tp, s->mq, s->mp, s->ks);
    else
      k_expand_ntt<false><<<dim3(n_active, nq), kNttThreads, 0, L->st>>>(
          L->trees, s->tree_words, L->gk_ptrs, a, s->gal[a], s->ginv[a], xk, s->Tq, s->Tp, s->mq, s->mp, s->ks);
    return true;
  }
  if (fused) {
    k_expand_node<<<dim3(n_active, nq), kNttThreads, 0, L->st>>>(L->trees, s->tree_words, L->gk_ptrs, a,
                                                                 s->ginv[a], s->Tq, s->Tp, s->mq, s->mp, s->ks);
    return false;
  }
  k_expand_ks<<<dim3(n_active * 4, nq), kNttThreads, 0, L->st>>>(
      L->trees, s->tree_words, L->gk_ptrs, a, s->ginv[a], L->digits, s->digit_words, s->Tq, s->Tp, s->mq, s->mp);
  k_expand_combine<<<dim3(n_active, nq), 256, 0, L->st>>>(L->trees, s->tree_words, L->digits, s->digit_words, a,
                                                          s->ginv[a], s->I.q, s->I.p, s->ks);
  return false;
}

// The expansion tree of each query; returns whether the leaves are in NTT
// form.  From the first fused level on, the tree stays in NTT form: that
// level pays one transform per node more than k_expand_node, every level
// after it none, and the last level saves one against writing NTT-form
// leaves from coefficient form.
bool run_expand(const onion_server *s, Lane *L, uint32_t nq) {
  const uint32_t h = s->I.expan_height, u = s->I.useful_cnt;
  bool ntt = false;
  for (uint32_t a = 0; a < h; ++a) {
    // active nodes at level a: those whose left-most leaf is < u
    const uint64_t span = s->w >> a;  // leaves under one node
    const uint32_t n_active = (uint32_t)((u + span - 1) / span);
    ntt = expand_level(s, L, nq, a, n_active, ntt, true);
  }
  return ntt;
}

// k_convert_fused<FOLD, NTT_IN> for nq queries, on stream st.
template <bool FOLD, bool NTT_IN>
void launch_convert_fused(const onion_server *s, Lane *L, uint32_t nq, uint32_t n_ct, const uint64_t *conv,
                          uint64_t conv_stride, hipStream_t st) {
  k_convert_fused<FOLD, NTT_IN><<<dim3(n_ct, nq), kNttThreads, 0, st>>>(
      conv, conv_stride, L->rk_ptrs, s->P.gsw_l_key, s->I.base_log2_key, L->rgsw, s->rgsw_words, s->P.gsw_l,
      s->Tq, s->mq);
}

// Conversion of the (ndim-1)*l RGSW leaves starting at conv (stride per query).
// NTT-form leaves (leaves_ntt) are the top rows as they are; the fused kernel
// copies them and returns them to coefficient form in registers, the split
// kernels after k_cts_to_coeff has done so in place.
void run_convert(const onion_server *s, Lane *L, uint32_t nq, uint64_t *conv, uint64_t conv_stride,
                 bool leaves_ntt = false) {
  hipStream_t st = L->st;
  const uint32_t nd = s->I.ndim - 1, l = s->P.gsw_l, lk = s->P.gsw_l_key;
  if (nd == 0) return;
  const uint32_t n_ct = nd * l;
  const bool fused = (uint64_t)n_ct * nq >= s->tune.fuse_min;
  if (fused && leaves_ntt && s->conv_fold) {
    launch_convert_fused<true, true>(s, L, nq, n_ct, conv, conv_stride, st);
    return;
  }
  if (leaves_ntt) {
    k_rgsw_top_rows<true><<<dim3(n_ct * 2, nq), kNttThreads, 0, st>>>(conv, conv_stride, L->rgsw,
                                                                         s->rgsw_words, l, s->Tq);
    k_cts_to_coeff<<<dim3(n_ct * 2, nq), kNttThreads, 0, st>>>(conv, conv_stride, s->Tq);
  } else {
    k_rgsw_top_rows<false><<<dim3(n_ct * 2, nq), kNttThreads, 0, st>>>(conv, conv_stride, L->rgsw,
                                                                          s->rgsw_words, l, s->Tq);
  }
  if (fused) {
    if (s->conv_fold)
      launch_convert_fused<true, false>(s, L, nq, n_ct, conv, conv_stride, st);
    else
      launch_convert_fused<false, false>(s, L, nq, n_ct, conv, conv_stride, st);
    return;
  }
  const dim3 dgrid(n_ct * 2 * lk, nq);
  if (s->conv_fold)
    k_decomp_ntt<true><<<dgrid, kNttThreads, 0, st>>>(conv, nullptr, conv_stride, L->digits, s->digit_words, lk,
                                                         s->I.base_log2_key, s->Tq);
  else
    k_decomp_ntt<false><<<dgrid, kNttThreads, 0, st>>>(conv, nullptr, conv_stride, L->digits, s->digit_words,
                                                          lk, s->I.base_log2_key, s->Tq);
  k_ext_mac_store<<<dim3(n_ct * 2, nq), kNttThreads, 0, st>>>(L->digits, s->digit_words, L->rk_ptrs, 2 * lk,
                                                                 L->rgsw, s->rgsw_words, l, s->mq);
}

template <int TR, int CPAD>
void launch_fd_bat_tr(const onion_server *s, Lane *L, uint32_t nq) {
  constexpr int RBLK = (256 / (CPAD / 4)) * TR;
  const uint32_t n_tiles = (s->P.other_dim + RBLK - 1) / RBLK;
  k_first_dim_bat<TR, 4, CPAD><<<kN * n_tiles, 256, 0, L->st>>>(
      s->db, L->Qs, L->fd, s->fd_words, s->P.fst_dim, s->P.other_dim, s->I.num_pt, nq, n_tiles, s->mq);
}

template <int CPAD>
void launch_fd_bat(const onion_server *s, Lane *L, uint32_t nq) {
  if (s->tune.fd_tile_rows == 2)
    launch_fd_bat_tr<2, CPAD>(s, L, nq);
  else
    launch_fd_bat_tr<4, CPAD>(s, L, nq);
}

// The int8-MFMA product serves batches of at least tune.fd_mfma_min queries
// when fst is a multiple of 32 and the lane's digit scratch holds the
// first-dimension outputs (other x N x 2nq words).
bool use_mfma(const onion_server *s, const Lane *L, uint32_t nq) {
  return nq >= 2 && s->tune.fd_mfma_min != 0 && nq >= s->tune.fd_mfma_min && s->P.fst_dim % 32 == 0 &&
         (uint64_t)L->cap * s->digit_words >= (uint64_t)kN * s->P.other_dim * 2 * nq;
}

// The product kernel for CB column blocks: from the plane copy when the server
// keeps one, else converting the database words in the kernel.
template <int CB>
void launch_mfma(const onion_server *s, Lane *L, const uint8_t *Bp, dim3 grid, uint32_t n_rt, uint32_t ncols) {
  const uint32_t fst = s->P.fst_dim, other = s->P.other_dim;
  const bool lds = s->tune.fd_mfma_variant == 1;
  if (s->dbp && s->tune.fd_mfma_variant == 2 && fst % 64 == 0) {
    // persistent: one workgroup per compute unit (a multiple of 8, at most the tile count)
    const uint32_t total = grid.x, G = std::max(8u, std::min(total, s->n_cu) / 8 * 8);
    k_first_dim_mfma_tiles<CB><<<G, 256, 0, L->st>>>(s->dbp, Bp, L->digits, fst, other, ncols, n_rt, s->mq);
  } else if (s->dbp) {
    if (lds)
      k_first_dim_mfma_lds<CB, true><<<grid, 256, 0, L->st>>>(s->db, s->dbp, Bp, L->digits, fst, other, ncols,
                                                             n_rt, s->mq);
    else
      k_first_dim_mfma_pre<CB><<<grid, 256, 0, L->st>>>(s->dbp, Bp, L->digits, fst, other, ncols, n_rt, s->mq);
  } else if (lds) {
    k_first_dim_mfma_lds<CB, false><<<grid, 256, 0, L->st>>>(s->db, s->dbp, Bp, L->digits, fst, other, ncols,
                                                              n_rt, s->mq);
  } else {
    k_first_dim_mfma<CB><<<grid, 256, 0, L->st>>>(s->db, Bp, L->digits, fst, other, ncols, n_rt, s->mq);
  }
}

// Inverse transforms of the outputs out2 into L->fd (coefficient form).
// Column groups of 8, 4 or 2 (ncols = 2 nq is even).
void launch_gather(const onion_server *s, Lane *L, uint32_t nq) {
  const uint32_t other = s->P.other_dim, ncols = 2 * nq;
  const uint32_t G = ncols % 8 == 0 ? 8 : ncols % 4 == 0 ? 4 : 2;
  const uint32_t grid = other * (ncols / G);
  if (G == 8)
    k_ntt_inv_gather<8><<<grid, kNttThreads, 0, L->st>>>(L->digits, other, ncols, L->fd, s->fd_words, s->Tq);
  else if (G == 4)
    k_ntt_inv_gather<4><<<grid, kNttThreads, 0, L->st>>>(L->digits, other, ncols, L->fd, s->fd_words, s->Tq);
  else
    k_ntt_inv_gather<2><<<grid, kNttThreads, 0, L->st>>>(L->digits, other, ncols, L->fd, s->fd_words, s->Tq);
}

// keep_ntt: leave the outputs in out2 (NTT form) for
// run_mux_first instead of transforming them into L->fd.  sel_ntt: the
// selection vectors are in NTT form already.
void run_first_dim_mfma(const onion_server *s, Lane *L, uint32_t nq, uint64_t *sel, uint64_t sel_stride,
                        bool time_core, bool keep_ntt, bool sel_ntt) {
  const uint32_t fst = s->P.fst_dim, other = s->P.other_dim, ncp = cpad_for(nq), ncols = 2 * nq;
  uint8_t *Bp = reinterpret_cast<uint8_t *>(L->Qs);
  const dim3 chunk_grid(kN / 8, ncp / 32, fst / 32);
  if (sel_ntt) {
    k_sel_chunks<true><<<chunk_grid, 256, 0, L->st>>>(sel, sel_stride, Bp, fst, ncols);
  } else {
    // the leaves are transformed in place: the selection vectors have no other reader
    k_sel_digits<<<dim3(fst * 2, nq), kNttThreads, 0, L->st>>>(sel, sel_stride, s->Tq);
    k_sel_chunks<false><<<chunk_grid, 256, 0, L->st>>>(sel, sel_stride, Bp, fst, ncols);
  }
  if (time_core) (void)hipEventRecord(L->ev[6], L->st);
  const uint32_t rb = 128 / ncp;  // row blocks of 32 per workgroup
  const uint32_t n_rt = (other + 32 * rb - 1) / (32 * rb);
  const dim3 grid(kN * n_rt);
  if (ncp == 32)
    launch_mfma<1>(s, L, Bp, grid, n_rt, ncols);
  else if (ncp == 64)
    launch_mfma<2>(s, L, Bp, grid, n_rt, ncols);
  else
    launch_mfma<4>(s, L, Bp, grid, n_rt, ncols);
  if (time_core) (void)hipEventRecord(L->ev[7], L->st);
  if (!keep_ntt) launch_gather(s, L, nq);
}

// First dimension: fst selection cts per query (coefficient form, or NTT
// form with sel_ntt; stride sel_stride) -> L->fd (other cts per query,
// coefficient form).
void run_first_dim(const onion_server *s, Lane *L, uint32_t nq, uint64_t *sel, uint64_t sel_stride,
                   bool time_core, bool keep_ntt = false, bool sel_ntt = false) {
  if (use_mfma(s, L, nq)) {
    run_first_dim_mfma(s, L, nq, sel, sel_stride, time_core, keep_ntt, sel_ntt);
    return;
  }
  const uint32_t fst = s->P.fst_dim, other = s->P.other_dim, cpad = cpad_for(nq);
  const uint64_t num_pt = s->I.num_pt;
  if (sel_ntt)
    k_sel_ntt<true><<<dim3(fst * 2, nq), kNttThreads, 0, L->st>>>(sel, sel_stride, L->Qs, fst, cpad, s->Tq);
  else
    k_sel_ntt<false><<<dim3(fst * 2, nq), kNttThreads, 0, L->st>>>(sel, sel_stride, L->Qs, fst, cpad, s->Tq);
  if (time_core) (void)hipEventRecord(L->ev[6], L->st);
  if (nq == 1) {
    if (fst % 16 == 0) {
      const uint32_t rpb = std::max(32u, std::min(s->tune.fd_rows_b1, (other + 31) / 32 * 32));
      const dim3 grid(kN, (other + rpb - 1) / rpb);
      switch (s->tune.fd_b1_variant & 3) {
        case 0: k_first_dim_b1<4, false><<<grid, 256, 0, L->st>>>(s->db, L->Qs, L->fd, fst, other, num_pt, rpb, s->mq); break;
        case 1: k_first_dim_b1<8, false><<<grid, 256, 0, L->st>>>(s->db, L->Qs, L->fd, fst, other, num_pt, rpb, s->mq); break;
        case 2: k_first_dim_b1<4, true><<<grid, 256, 0, L->st>>>(s->db, L->Qs, L->fd, fst, other, num_pt, rpb, s->mq); break;
        default: k_first_dim_b1<8, true><<<grid, 256, 0, L->st>>>(s->db, L->Qs, L->fd, fst, other, num_pt, rpb, s->mq); break;
      }
    } else {
      k_first_dim_generic<<<kN, 256, 0, L->st>>>(s->db, L->Qs, L->fd, fst, other, num_pt, s->mq);
    }
  } else if (cpad == 32) {
    launch_fd_bat<32>(s, L, nq);
  } else if (cpad == 64) {
    launch_fd_bat<64>(s, L, nq);
  } else {
    launch_fd_bat<128>(s, L, nq);
  }
  if (time_core) (void)hipEventRecord(L->ev[7], L->st);
  if (!keep_ntt) k_ntt_inv<8><<<other * 2 * nq, kNttThreads, 0, L->st>>>(L->fd, L->fd, s->Tq);
}

// n NTT-form cts per query in fd back to coefficient form.
void cts_to_coeff(const onion_server *s, Lane *L, uint32_t nq, uint64_t n) {
  k_cts_to_coeff<<<dim3((uint32_t)(2 * n), nq), kNttThreads, 0, L->st>>>(L->fd, s->fd_words, s->Tq);
}

// k_mux_fused_ntt transforms digits of at most 32 bits (ntt_fwd<true>) and
// needs at least two of them; k_rgsw_fold_last's gadget powers
// 2^((l-1-j)B) stay below 2^60 > q.
bool mux_ntt_ok(const onion_server *s) {
  const uint32_t l = s->P.gsw_l, b = s->I.base_log2;
  return b <= 32 && l >= 2 && (uint64_t)(l - 1) * b < 60;
}

// The level's RGSW `g` (per query) folded for k_mux_fused_ntt.
void fold_last(const onion_server *s, Lane *L, uint32_t nq, uint64_t *g) {
  k_rgsw_fold_last<<<dim3(4 * kN / 256, nq), 256, 0, L->st>>>(g, s->rgsw_words, s->P.gsw_l, s->I.base_log2, s->mq);
}

// One multiplexer level over n cts per query in fd, with RGSW `g` (per-query
// stride).  *ntt_form: the cts are in NTT form (slot order); a wide level
// keeps them so, a narrow one (the split kernels) first brings them back to
// coefficient form.
void run_mux_level(const onion_server *s, Lane *L, uint32_t nq, uint64_t n, uint64_t *g,
                   bool *ntt_form = nullptr) {
  const uint32_t l = s->P.gsw_l;
  const uint64_t half = n / 2;
  const bool in_ntt = ntt_form && *ntt_form;
  if (half * nq >= s->tune.fuse_min) {
    if (in_ntt) {
      fold_last(s, L, nq, g);
      if (s->tune.ntt_radix == 4)
        k_mux_fused_ntt<false, 1, 4><<<dim3((uint32_t)half, nq), kNttThreads, 0, L->st>>>(
            L->fd, s->fd_words, half, nullptr, 0, nq, g, s->rgsw_words, l, s->I.base_log2, s->Tq, s->mq);
      else
        k_mux_fused_ntt<false><<<dim3((uint32_t)half, nq), kNttThreads, 0, L->st>>>(
            L->fd, s->fd_words, half, nullptr, 0, nq, g, s->rgsw_words, l, s->I.base_log2, s->Tq, s->mq);
    }
    else
      k_mux_fused<<<dim3((uint32_t)half, nq), kNttThreads, 0, L->st>>>(L->fd, s->fd_words, half, g, s->rgsw_words,
                                                                      l, s->I.base_log2, s->Tq, s->mq);
    return;
  }
  if (in_ntt) {
    cts_to_coeff(s, L, nq, n);
    *ntt_form = false;
  }
  k_decomp_ntt<false><<<dim3((uint32_t)(half * 2 * l), nq), kNttThreads, 0, L->st>>>(
      L->fd, L->fd + half * 2 * kN, s->fd_words, L->digits, s->digit_words, l, s->I.base_log2, s->Tq);
  k_ext_mac_intt_add<<<dim3((uint32_t)(half * 2), nq), kNttThreads, 0, L->st>>>(
      L->digits, s->digit_words, g, s->rgsw_words, 2 * l, L->fd, s->fd_words, s->Tq, s->mq);
}

// The first multiplexer level takes the batched first dimension's NTT-form
// outputs directly (k_mux_fused_ntt<true> reads both rows of a pair from
// out2: two inverse transforms per pair instead of the reference's six) and
// leaves NTT-form cts for the wide levels after it, when the matrix-core
// product runs and the level is wide enough for the fused external product.
bool mux_first_fused(const onion_server *s, const Lane *L, uint32_t nq) {
  return s->I.ndim >= 2 && use_mfma(s, L, nq) && mux_ntt_ok(s) &&
         (uint64_t)(s->P.other_dim / 2) * nq >= s->tune.fuse_min;
}

// The VALU first-dimension kernels (one query, or batches off the matrix
// cores) also leave NTT-form outputs in fd: a wide first level folds them as
// they are (k_mux_fused_ntt<false>) instead of transforming them all back.
bool first_dim_keeps_ntt(const onion_server *s, const Lane *L, uint32_t nq) {
  return s->I.ndim >= 2 && !use_mfma(s, L, nq) && mux_ntt_ok(s) &&
         (uint64_t)(s->P.other_dim / 2) * nq >= s->tune.fuse_min;
}

// An even batch runs one workgroup per row pair and two queries (reading the
// 32 bytes of both per slot and row); an odd one, one per query.
void run_mux_first(const onion_server *s, Lane *L, uint32_t nq) {
  const uint32_t half = s->P.other_dim / 2;
  fold_last(s, L, nq, L->rgsw);
  const bool r4 = s->tune.ntt_radix == 4;
  if (nq % 2 == 0 && r4)
    k_mux_fused_ntt<true, 2, 4><<<half * nq / 2, kNttThreads, 0, L->st>>>(
        L->fd, s->fd_words, half, L->digits, 2 * nq, nq, L->rgsw, s->rgsw_words, s->P.gsw_l, s->I.base_log2, s->Tq,
        s->mq);
  else if (nq % 2 == 0)
    k_mux_fused_ntt<true, 2><<<half * nq / 2, kNttThreads, 0, L->st>>>(
        L->fd, s->fd_words, half, L->digits, 2 * nq, nq, L->rgsw, s->rgsw_words, s->P.gsw_l, s->I.base_log2, s->Tq,
        s->mq);
  else if (r4)
    k_mux_fused_ntt<true, 1, 4><<<half * nq, kNttThreads, 0, L->st>>>(
        L->fd, s->fd_words, half, L->digits, 2 * nq, nq, L->rgsw, s->rgsw_words, s->P.gsw_l, s->I.base_log2, s->Tq,
        s->mq);
  else
    k_mux_fused_ntt<true, 1><<<half * nq, kNttThreads, 0, L->st>>>(
        L->fd, s->fd_words, half, L->digits, 2 * nq, nq, L->rgsw, s->rgsw_words, s->P.gsw_l, s->I.base_log2, s->Tq,
        s->mq);
}

void run_mod_switch(const onion_server *s, Lane *L, uint32_t nq, uint64_t *cts, uint64_t stride) {
  k_mod_switch<<<dim3((2 * kN + 255) / 256, nq), 256, 0, L->st>>>(cts, stride, 2 * kN, s->ms_M, s->ms_f);
}

// Full answer for nq queries already in tree slot 1; responses in fd[b][0].
// With `timed`, events ev[0..5] bracket the stages and ev[6..7] the
// database-streaming kernel (never inside a graph capture).
void answer_device(const onion_server *s, Lane *L, uint32_t nq, bool timed) {
  const uint32_t fst = s->P.fst_dim, l = s->P.gsw_l;
  uint64_t *leaves = L->trees + s->w * 2 * kN;
  if (timed) (void)hipEventRecord(L->ev[0], L->st);
  const bool leaves_ntt = run_expand(s, L, nq);
  if (timed) (void)hipEventRecord(L->ev[1], L->st);
  // Every stage runs on the lane's one stream: no event orders one stream
  // after another, inside a captured graph or out of it.  (A batch's
  // conversion on a second stream beside the selection stage, forked and
  // joined by events, saved about 0.15 ms per batch of 64.)
  run_convert(s, L, nq, leaves + (size_t)fst * 2 * kN, s->tree_words, leaves_ntt);
  if (timed) (void)hipEventRecord(L->ev[2], L->st);
  const bool first_fused = mux_first_fused(s, L, nq), fd_ntt = first_dim_keeps_ntt(s, L, nq);
  run_first_dim(s, L, nq, leaves, s->tree_words, timed, first_fused || fd_ntt, leaves_ntt);
  if (timed) (void)hipEventRecord(L->ev[3], L->st);
  uint64_t n = s->P.other_dim;
  uint32_t d = 0;
  bool ntt_form = fd_ntt;
  if (first_fused) {
    run_mux_first(s, L, nq);
    n /= 2;
    d = 1;
    ntt_form = true;
  }
  for (; d + 1 < s->I.ndim; ++d) {
    run_mux_level(s, L, nq, n, L->rgsw + (size_t)d * 2 * l * 2 * kN, &ntt_form);
    n /= 2;
  }
  if (ntt_form) cts_to_coeff(s, L, nq, n);
  if (timed) (void)hipEventRecord(L->ev[4], L->st);
  run_mod_switch(s, L, nq, L->fd, s->fd_words);
  if (timed) (void)hipEventRecord(L->ev[5], L->st);
}

// The lane's graph of the pipeline for nq queries, captured on first use.
int graph_for(const onion_server *s, Lane *L, uint32_t nq, hipGraphExec_t *out) {
  if (L->graph_epoch != s->tune_epoch) {
    drop_graphs(L);
    L->graph_epoch = s->tune_epoch;
  }
  auto it = L->graphs.find(nq);
  if (it != L->graphs.end()) {
    *out = it->second;
    return ONION_OK;
  }
  HIPCHK(hipStreamBeginCapture(L->st, hipStreamCaptureModeThreadLocal));
  answer_device(s, L, nq, false);
  hipGraph_t g = nullptr;
  const hipError_t launch_err = hipGetLastError();
  const hipError_t end_err = hipStreamEndCapture(L->st, &g);
  if (launch_err != hipSuccess || end_err != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(ONION_EHIP, std::string("pipeline capture failed: ") +
                                hipGetErrorString(launch_err != hipSuccess ? launch_err : end_err));
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t inst_err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (inst_err != hipSuccess) return fail(ONION_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(inst_err));
  L->graphs[nq] = ex;
  *out = ex;
  return ONION_OK;
}

int collect_times(onion_server *s, Lane *L, bool staged, onion_stage_times *out) {
  HIPCHK(hipEventSynchronize(L->ev[5]));
  onion_stage_times t{};
  float tot = 0;
  HIPCHK(hipEventElapsedTime(&tot, L->ev[0], L->ev[5]));
  t.total = tot;
  if (staged) {
    float ms[5], core = 0;
    for (int i = 0; i < 5; ++i) HIPCHK(hipEventElapsedTime(&ms[i], L->ev[i], L->ev[i + 1]));
    HIPCHK(hipEventElapsedTime(&core, L->ev[6], L->ev[7]));
    t.expand = ms[0];
    t.convert = ms[1];
    t.first_dim = ms[2];
    t.other_dim = ms[3];
    t.mod_switch = ms[4];
    t.first_dim_core = core;
  }
  if (out) *out = t;
  std::lock_guard<std::mutex> g(s->last_mu);
  s->last = t;
  return ONION_OK;
}

// Whether a call of n queries replays a captured graph (tune.use_graph: 0
// never, 1 always, 2 single queries only).
bool graph_for_size(const onion_server *s, uint32_t n) {
  return s->tune.use_graph == 1 || (s->tune.use_graph == 2 && n == 1);
}

// Answer n (<= L->cap) queries on lane L: upload, run, download.
// With eager_only, the pipeline is never captured (hipGraphInstantiate
// allocates).
int answer_chunk(onion_server *s, Lane *L, uint32_t n, const uint64_t *ids, const uint64_t *queries,
                 uint64_t *resps, onion_stage_times *times, bool eager_only = false) {
  int rc;
  if (L->in_flight) {
    HIPCHK(hipStreamSynchronize(L->st));
    L->in_flight = false;
  }
  if ((rc = bind_keys(s, L, n, ids, s->I.ndim > 1))) return rc;
  const size_t qbytes = (size_t)n * 2 * kN * 8;
  const uint64_t *qsrc = queries;
  if (L->hq) {  // no copy of an earlier call reads it any more (each synchronised)
    std::memcpy(L->hq, queries, qbytes);
    qsrc = L->hq;
  }
  L->in_flight = true;
  HIPCHK(hipMemcpy2DAsync(L->trees + 2 * kN, s->tree_words * 8, qsrc, 2 * kN * 8, 2 * kN * 8, n,
                          hipMemcpyHostToDevice, L->st));
  const bool graph = graph_for_size(s, n) && !eager_only;
  if (graph) {
    hipGraphExec_t ex;
    if ((rc = graph_for(s, L, n, &ex))) return rc;
    HIPCHK(hipEventRecord(L->ev[0], L->st));
    HIPCHK(hipGraphLaunch(ex, L->st));
    HIPCHK(hipEventRecord(L->ev[5], L->st));
  } else {
    answer_device(s, L, n, true);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpy2DAsync(L->hr ? L->hr : resps, 2 * kN * 8, L->fd, s->fd_words * 8, 2 * kN * 8, n,
                          hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  L->in_flight = false;
  if (L->hr) std::memcpy(resps, L->hr, qbytes);
  return collect_times(s, L, !graph, times);
}

// Size of the next chunk of n waiting queries, at most cap per pass: the cap
// itself (then the rest), or with batch_split = 1 equal chunks.
uint32_t next_chunk(uint32_t split, uint32_t n, uint32_t cap) {
  if (n <= cap) return n;
  if (split == 1) {
    const uint32_t k = (n + cap - 1) / cap;
    return (n + k - 1) / k;
  }
  return cap;
}

// n queries on lane L in chunks of at most kMaxBatch (tune.batch_split); the
// caller holds the gate shared and L->mu.
int answer_on_lane(onion_server *s, Lane *L, uint32_t n, const uint64_t *ids, const uint64_t *queries,
                   uint64_t *resps, onion_stage_times *times) {
  int rc;
  if ((rc = check_db(s))) return rc;
  if (n == 0) return ONION_OK;
  if ((rc = ensure_scratch(s, L, std::min(n, kMaxBatch)))) return rc;
  onion_stage_times sum{};
  for (uint32_t off = 0, m = 0; off < n; off += m) {
    m = next_chunk(s->tune.batch_split, n - off, kMaxBatch);
    onion_stage_times t{};
    if ((rc = answer_chunk(s, L, m, ids + off, queries + (size_t)off * 2 * kN, resps + (size_t)off * 2 * kN, &t)))
      return rc;
    sum.expand += t.expand;
    sum.convert += t.convert;
    sum.first_dim += t.first_dim;
    sum.other_dim += t.other_dim;
    sum.mod_switch += t.mod_switch;
    sum.total += t.total;
    sum.first_dim_core += t.first_dim_core;
  }
  if (times) *times = sum;
  return ONION_OK;
}

// ---- autotuning -------------------------------------------------------------------

// Device time of one pass of the pipeline for n queries already bound on lane
// L (minimum over reps, after one untimed pass that also captures the graph).
int time_pipeline(onion_server *s, Lane *L, uint32_t n, uint32_t reps, double *ms) {
  std::vector<double> runs;
  for (uint32_t r = 0; r <= reps; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    if (graph_for_size(s, n)) {
      hipGraphExec_t ex;
      int rc = graph_for(s, L, n, &ex);
      if (rc) return rc;
      HIPCHK(hipEventRecord(L->ev[0], L->st));
      HIPCHK(hipGraphLaunch(ex, L->st));
      HIPCHK(hipEventRecord(L->ev[5], L->st));
    } else {
      HIPCHK(hipEventRecord(L->ev[0], L->st));
      answer_device(s, L, n, false);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(L->ev[5], L->st));
    }
    HIPCHK(hipEventSynchronize(L->ev[5]));
    const std::chrono::duration<double, std::milli> t = std::chrono::steady_clock::now() - t0;
    if (r > 0) runs.push_back(t.count());
  }
  std::sort(runs.begin(), runs.end());
  *ms = runs[runs.size() / 2];
  return ONION_OK;
}

// A call of n queries (any n) as answer_on_lane runs it: chunk by chunk.
int time_call(onion_server *s, Lane *L, uint32_t n, uint32_t reps, double *ms) {
  *ms = 0;
  for (uint32_t off = 0, m = 0; off < n; off += m) {
    m = next_chunk(s->tune.batch_split, n - off, kMaxBatch);
    double t;
    int rc = time_pipeline(s, L, m, reps, &t);
    if (rc) return rc;
    *ms += t;
  }
  return ONION_OK;
}

// Coordinate descent over the launch-shape choices on the live database with
// a synthetic client (uniform key words and queries: the work does not depend
// on the values).  Score = sum over the batch sizes of the call's time
// (host clock from the first launch to completion, median of the runs: the
// launch cost counts, as it does for a caller waiting on the call)
// relative to the starting choice's, so every size counts alike (in plain ms
// per query a single query would outweigh a batch of 64 and decide choices
// that only move the batch).  The caller holds the gate exclusively.
int autotune_impl(onion_server *s, const uint32_t *sizes, uint32_t n_sizes, uint32_t reps, onion_tuning *out) {
  Lane *L = s->lanes[0].get();
  uint32_t maxn = 1;
  bool has1 = false, hasb = false, has_small = false, has_big = false;
  for (uint32_t i = 0; i < n_sizes; ++i) {
    maxn = std::max(maxn, std::min(sizes[i], kMaxBatch));
    has1 |= sizes[i] == 1;
    hasb |= sizes[i] >= 2;
    has_small |= sizes[i] >= 2 && sizes[i] < 16;
    has_big |= sizes[i] > kMaxBatch && sizes[i] % kMaxBatch != 0;
  }
  int rc;
  if ((rc = ensure_scratch(s, L, maxn))) return rc;
  uint64_t *gk = nullptr, *rk = nullptr;
  // Galois keys as the store holds them: key and Shoup companions per level
  const size_t gk_dev_words = 2 * (size_t)s->I.galois_key_words;
  if ((rc = dalloc(&gk, gk_dev_words)) || (rc = dalloc(&rk, s->I.rgsw_key_words))) {
    dfree(gk);
    return rc;
  }
  struct Cleanup {
    uint64_t *a, *b;
    ~Cleanup() { dfree(a); dfree(b); }
  } cleanup{gk, rk};
  {
    std::vector<uint64_t> h(std::max<size_t>(gk_dev_words, s->I.rgsw_key_words));
    uint64_t key = stream_key(0x5475ULL, 1);
    for (size_t i = 0; i < h.size(); ++i) h[i] = rand64_keyed(key, i) % s->I.q;
    HIPCHK(hipMemcpy(gk, h.data(), gk_dev_words * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(rk, h.data(), (size_t)s->I.rgsw_key_words * 8, hipMemcpyHostToDevice));
    std::vector<uint64_t> q(2 * kN);
    for (uint32_t b = 0; b < maxn; ++b) {
      for (size_t i = 0; i < q.size(); ++i) q[i] = rand64_keyed(key, (uint64_t)(b + 1) << 32 | i) % s->I.q;
      HIPCHK(hipMemcpy(L->trees + (size_t)b * s->tree_words + 2 * kN, q.data(), 2 * kN * 8, hipMemcpyHostToDevice));
      L->h_gk[b] = gk;
      L->h_rk[b] = rk;
    }
    HIPCHK(hipMemcpy(L->gk_ptrs, L->h_gk.data(), maxn * sizeof(void *), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L->rk_ptrs, L->h_rk.data(), maxn * sizeof(void *), hipMemcpyHostToDevice));
  }
  std::vector<double> base(n_sizes, 0.0);
  auto score = [&](const onion_tuning &t, double *sc) -> int {
    s->tune = t;
    ++s->tune_epoch;
    double total = 0;
    for (uint32_t i = 0; i < n_sizes; ++i) {
      double ms;
      int r = time_call(s, L, sizes[i], reps, &ms);
      if (r) return r;
      if (base[i] == 0) base[i] = std::max(ms, 1e-6);
      total += ms / base[i];
    }
    *sc = total;
    return ONION_OK;
  };
  onion_tuning best = s->tune;
  double best_sc;
  {
    // untimed passes first: a device coming out of idle clocks up during the
    // first runs, which would flatter whichever choice is timed after them
    double ms;
    for (uint32_t w = 0; w < 3; ++w)
      if ((rc = time_pipeline(s, L, maxn, 1, &ms))) return rc;
  }
  if ((rc = score(best, &best_sc))) return rc;
  struct Axis {
    uint32_t onion_tuning::*field;
    std::vector<uint32_t> values;
    bool active;
  };
  const std::vector<Axis> axes = {
      {&onion_tuning::fd_mfma_variant, {0, 1, 2}, hasb},
      {&onion_tuning::fd_mfma_min, {2, 4, 8, 16, 0}, has_small},
      {&onion_tuning::fd_tile_rows, {2, 4}, hasb},
      {&onion_tuning::fd_b1_variant, {0, 1, 2, 3}, has1},
      {&onion_tuning::fd_rows_b1, {64, 128, 256, 512}, has1},
      {&onion_tuning::fuse_min, {1, 256, 512, 1024, 4096, 1u << 30}, true},
      {&onion_tuning::ntt_radix, {8, 4}, true},
      {&onion_tuning::use_graph, {0, 1, 2}, true},
      {&onion_tuning::batch_split, {0, 1}, has_big},
  };
  for (const Axis &ax : axes) {
    if (!ax.active) continue;
    for (uint32_t v : ax.values) {
      if (best.*ax.field == v) continue;
      onion_tuning t = best;
      t.*ax.field = v;
      double sc;
      if ((rc = score(t, &sc))) {
        s->tune = best;
        ++s->tune_epoch;
        return rc;
      }
      if (sc < 0.99 * best_sc) {  // 1% hysteresis against timing noise
        best = t;
        best_sc = sc;
      }
    }
  }
  s->tune = best;
  ++s->tune_epoch;
  if (out) *out = best;
  return ONION_OK;
}

}  // namespace

// =====================================================================================
// C ABI
// =====================================================================================

extern "C" {

const char *onion_gpu_last_error(void) { return g_err.c_str(); }

int onion_gpu_device_count(int *count) {
  if (!count) return fail(ONION_EINVAL, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return ONION_OK;
}

int onion_params_derive(const onion_params *params, onion_param_info *out) {
  if (!params || !out) return fail(ONION_EINVAL, "null argument");
  std::string err;
  if (derive_params(*params, *out, err) != ONION_OK) return fail(ONION_EINVAL, err);
  return ONION_OK;
}

int onion_gpu_create(const onion_params *params, onion_server **out) {
  if (!params || !out) return fail(ONION_EINVAL, "null argument");
  const onion_params P = *params;
  onion_param_info info;
  std::string err;
  if (derive_params(P, info, err, kN) != ONION_OK) return fail(ONION_EINVAL, err);
  if (P.fst_dim > (uint32_t)kMaxFst) return fail(ONION_EINVAL, "fst_dim above 1024");
  // the transforms' reductions take moduli 2^60 - c with c < 2^28 (reduce_2p63,
  // ntt_fwd_final) and the 128-bit sums' reduction c < 2^20 (reduce128); the
  // N = 2048 primes are 2^60 - 98303 and 2^60 - 16383
  for (uint64_t m : {info.q, info.p})
    if (m >= (1ULL << 60) || (1ULL << 60) - m >= (1ULL << 20))
      return fail(ONION_EINVAL, "moduli must be 2^60 - c with c < 2^20");

  onion_server *s = new onion_server();
  s->P = P;
  s->I = info;
  const onion_param_info &I = s->I;
  const uint64_t N = kN;
  s->w = uint64_t(1) << I.expan_height;

  int rc;
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return fail(ONION_EHIP, "no HIP device");
  }
  {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess || cu <= 0) {
      delete s;
      return fail(ONION_EHIP, "cannot read the device's compute unit count");
    }
    s->n_cu = (uint32_t)cu;
  }
  s->lanes.emplace_back();
  if ((rc = make_lane(s->lanes[0])) || (rc = make_lane(s->slane))) {
    std::string msg = g_err;
    onion_gpu_destroy(s);
    return fail(rc, msg);
  }
  static_scratch_init(s->device);
  s->slane->borrowed = true;
  s->slane->h_gk.assign(kMaxBatch, nullptr);
  s->slane->h_rk.assign(kMaxBatch, nullptr);

  // NTT tables and constants
  host::Tables tq = host::make_tables(I.q, kN), tp = host::make_tables(I.p, kN);
  const size_t tab_words = (8 + 2 * (size_t)I.expan_height) * N;
  if ((rc = dalloc(&s->tabs, tab_words))) { onion_gpu_destroy(s); return rc; }
  std::vector<uint64_t> h(tab_words);
  const std::vector<uint64_t> *src[8] = {&tq.rp, &tq.rps, &tq.irp, &tq.irps, &tp.rp, &tp.rps, &tp.irp, &tp.irps};
  for (int i = 0; i < 8; ++i) memcpy(h.data() + i * N, src[i]->data(), N * 8);
  for (uint32_t a = 0; a < I.expan_height; ++a) {
    // x^-k in NTT form, k = 2^a: slot s (holding a(psi^o), o = 2 bitrev(s) + 1) times psi^(-k o)
    const uint64_t base = host::invmod(host::powmod(tq.psi, uint64_t(1) << a, I.q), I.q);
    uint64_t *f = h.data() + (8 + 2 * (size_t)a) * N;
    for (uint32_t e = 0; e < N; ++e) {
      f[e] = host::powmod(base, 2 * host::bitrev(e, 11) + 1, I.q);
      f[N + e] = host::shoup(f[e], I.q);
    }
  }
  s->xk = s->tabs + 8 * N;
  if (hipMemcpy(s->tabs, h.data(), tab_words * 8, hipMemcpyHostToDevice) != hipSuccess) {
    onion_gpu_destroy(s);
    return fail(ONION_EHIP, "table upload failed");
  }
  s->Tq = NttTab{s->tabs, s->tabs + N, s->tabs + 2 * N, s->tabs + 3 * N, I.q, tq.n_inv, tq.n_inv_s};
  s->Tp = NttTab{s->tabs + 4 * N, s->tabs + 5 * N, s->tabs + 6 * N, s->tabs + 7 * N, I.p, tp.n_inv, tp.n_inv_s};
  s->mq.q = I.q;
  host::barrett_ratio(I.q, s->mq.r_lo, s->mq.r_hi);
  s->mp.q = I.p;
  host::barrett_ratio(I.p, s->mp.r_lo, s->mp.r_hi);
  s->ks.half_p = I.p >> 1;
  s->ks.half_p_mod_q = s->ks.half_p % I.q;
  s->ks.p_inv = host::invmod(I.p % I.q, I.q);
  s->ks.p_inv_s = host::shoup(s->ks.p_inv, I.q);
  s->ks.np_inv = host::mulmod(N % I.q, s->ks.p_inv, I.q);
  s->ks.np_inv_s = host::shoup(s->ks.np_inv, I.q);
  s->lift = LiftConst{I.t, (I.t + 1) / 2, I.q - I.t};
  // k_rgsw_fold_last's gadget powers 2^((l_key-1)B) stay below q, and the
  // digits below 2^32 (the digit transforms' first stage)
  s->conv_fold = P.gsw_l_key >= 2 && I.base_log2_key <= 32 && (uint64_t)(P.gsw_l_key - 1) * I.base_log2_key < 60;
  for (uint32_t a = 0; a < I.expan_height; ++a) {
    const uint64_t g = N / (uint64_t(1) << a) + 1;
    // g is odd, so it is invertible mod 2N
    uint64_t gi = 1;
    for (uint64_t x = 1; x < 2 * N; x += 2)
      if ((x * g) % (2 * N) == 1) { gi = x; break; }
    s->gal.push_back((uint32_t)g);
    s->ginv.push_back((uint32_t)gi);
  }
  host::mod_switch_scale(I.q, I.small_q, s->ms_M, s->ms_f);

  // per-query scratch sizes
  const uint64_t ct = 2 * N;
  const uint32_t nd = I.ndim - 1;
  s->tree_words = 2 * s->w * ct;
  s->rgsw_words = (uint64_t)std::max(1u, nd) * 2 * P.gsw_l * ct;
  s->fd_words = (uint64_t)P.other_dim * ct;
  const uint64_t conv_digits = (uint64_t)nd * P.gsw_l * 2 * P.gsw_l_key * N;
  const uint64_t mux_digits = (uint64_t)std::max(1u, P.other_dim / 2) * 2 * P.gsw_l * N;
  const uint64_t expand_x = (uint64_t)((I.useful_cnt + 1) / 2) * 4 * N;  // widest level's X
  s->digit_words = std::max({conv_digits, mux_digits, expand_x});
  s->io_words = std::max<uint64_t>({(uint64_t)I.useful_cnt * ct, (uint64_t)P.other_dim * ct,
                                    (uint64_t)std::max(1u, nd) * 2 * P.gsw_l * ct, 64 * ct,
                                    (uint64_t)P.fst_dim * ct});
  if ((rc = dalloc(&s->db, I.num_pt * N)) || (rc = dalloc(&s->io, s->io_words)) ||
      (rc = dalloc(&s->pack, (uint64_t)I.response_bytes + 64)) || (rc = ensure_scratch(s, s->lanes[0].get(), 1))) {
    std::string msg = g_err;
    onion_gpu_destroy(s);
    return fail(rc, msg);
  }
  s->pt_written.assign(I.num_pt, 0);
  // the plane copy when it takes at most a third of the free device memory
  // (the rest stays for batch scratch and other databases); without it the
  // products convert in-kernel.
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && (uint64_t)free_b / 3 < planes_bytes(s))
    s->tune.db_planes = 0;
  if ((rc = sync_planes(s)) == ONION_ENOMEM) {
    s->tune.db_planes = 0;
  } else if (rc) {
    std::string msg = g_err;
    onion_gpu_destroy(s);
    return fail(rc, msg);
  }
  s->keys = std::make_shared<KeyStore>();
  s->keys->q = I.q;
  s->keys->p = I.p;
  s->keys->l_key = P.gsw_l_key;
  *out = s;
  return ONION_OK;
}

int onion_gpu_destroy(onion_server *s) {
  if (!s) return ONION_OK;
  {
    Exclusive x(s->gate);
    for (auto &L : s->lanes)
      if (L) destroy_lane(L.get());
    s->lanes.clear();
    if (s->slane) destroy_lane(s->slane.get());
    s->slane.reset();
    s->keys.reset();  // the store goes with its last server
    void *bufs[] = {s->tabs, s->db, s->dbp, s->io, s->pack};
    for (void *b : bufs) dfree(b);
  }
  delete s;
  return ONION_OK;
}

int onion_gpu_param_info(const onion_server *s, onion_param_info *out) {
  if (!s || !out) return fail(ONION_EINVAL, "null argument");
  *out = s->I;
  return ONION_OK;
}

int onion_gpu_set_lanes(onion_server *s, uint32_t n) {
  if (!s) return fail(ONION_EINVAL, "null server");
  if (n < 1 || n > kMaxLanes) return fail(ONION_EINVAL, "n_lanes must be in [1, 16]");
  Exclusive x(s->gate);
  while (s->lanes.size() > n) {
    destroy_lane(s->lanes.back().get());
    s->lanes.pop_back();
  }
  while (s->lanes.size() < n) {
    std::unique_ptr<Lane> L;
    int rc = make_lane(L);
    if (rc) return rc;
    s->lanes.push_back(std::move(L));
  }
  return ONION_OK;
}

int onion_gpu_scratch_bytes_per_query(const onion_server *s, uint64_t *bytes) {
  if (!s || !bytes) return fail(ONION_EINVAL, "null argument");
  *bytes = scratch_bytes_per_query(s);
  return ONION_OK;
}

int onion_gpu_scratch_bytes(const onion_server *s, uint32_t cap, uint64_t *bytes) {
  if (!s || !bytes) return fail(ONION_EINVAL, "null argument");
  if (cap < 1 || cap > kMaxBatch) return fail(ONION_EINVAL, "cap must be in [1, 64]");
  *bytes = scratch_bytes(s, cap);
  return ONION_OK;
}

int onion_gpu_reserve(onion_server *s, uint32_t lane, uint32_t n) {
  if (!s) return fail(ONION_EINVAL, "null server");
  Shared g(s->gate);
  if (lane >= s->lanes.size())
    return fail(ONION_EINVAL, "lane " + std::to_string(lane) + " of " + std::to_string(s->lanes.size()));
  Lane *L = s->lanes[lane].get();
  std::lock_guard<std::mutex> lk(L->mu);
  return ensure_scratch(s, L, std::max(1u, std::min(n, kMaxBatch)));
}

int onion_gpu_db_generate(onion_server *s, uint64_t seed, int mode) {
  if (!s) return fail(ONION_EINVAL, "null server");
  if (mode != 0 && mode != 1) return fail(ONION_EINVAL, "mode must be 0 or 1");
  if (mode == 1 && s->P.plain_bits > 25)
    return fail(ONION_EINVAL, "entry mode packs at most 24 bits per coefficient");
  Exclusive x(s->gate);
  hipStream_t st = s->lanes[0]->st;
  const uint64_t key = stream_key(seed, 0x4442ULL);
  const uint64_t chunk = 65535;
  for (uint64_t lo = 0; lo < s->I.num_pt; lo += chunk) {
    const uint64_t n = std::min(chunk, s->I.num_pt - lo);
    k_db_generate<<<(uint32_t)n, kNttThreads, 0, st>>>(s->db, lo, s->I.num_pt, key, mode, s->P.plain_bits - 1,
                                                       s->lift, s->Tq);
  }
  planes_update(s, st, 0, s->I.num_pt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  std::fill(s->pt_written.begin(), s->pt_written.end(), 1);
  s->n_written = s->I.num_pt;
  return ONION_OK;
}

int onion_gpu_db_upload_plain(onion_server *s, uint64_t pt_lo, uint64_t n_pt, const uint64_t *coeffs) {
  if (!s || (!coeffs && n_pt)) return fail(ONION_EINVAL, "null argument");
  if (pt_lo > s->I.num_pt || n_pt > s->I.num_pt - pt_lo)
    return fail(ONION_EINVAL, "plaintext range beyond the database");
  for (uint64_t i = 0; i < n_pt * kN; ++i)
    if (coeffs[i] >= s->I.t) return fail(ONION_EINVAL, "plaintext coefficient >= t");
  Exclusive x(s->gate);
  hipStream_t st = s->lanes[0]->st;
  const uint64_t chunk = std::max<uint64_t>(1, s->io_words / kN);
  for (uint64_t off = 0; off < n_pt; off += chunk) {
    const uint64_t n = std::min(chunk, n_pt - off);
    HIPCHK(hipMemcpyAsync(s->io, coeffs + off * kN, n * kN * 8, hipMemcpyHostToDevice, st));
    k_db_plain<<<(uint32_t)n, kNttThreads, 0, st>>>(s->io, s->db, pt_lo + off, s->I.num_pt, s->lift, s->Tq);
    planes_update(s, st, pt_lo + off, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    for (uint64_t i = 0; i < n; ++i) {
      uint8_t &f = s->pt_written[pt_lo + off + i];
      if (!f) { f = 1; ++s->n_written; }
    }
  }
  return ONION_OK;
}

int onion_gpu_db_read_ntt(onion_server *s, uint64_t pt, uint64_t *out) {
  if (!s || !out || pt >= s->I.num_pt) return fail(ONION_EINVAL, "bad argument");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  HIPCHK(hipMemcpy2DAsync(out, 8, s->db + pt, s->I.num_pt * 8, 8, kN, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  return ONION_OK;
}

int onion_gpu_set_galois_key(onion_server *s, uint64_t id, const uint64_t *keys, size_t n) {
  if (!s || !keys) return fail(ONION_EINVAL, "null argument");
  const size_t level_words = 4 * (size_t)kN, max_levels = 11;  // log2 N
  if (n % level_words || n / level_words < s->I.expan_height || n / level_words > max_levels)
    return fail(ONION_EINVAL, "Galois key size " + std::to_string(n) + " is not " +
                                  std::to_string(level_words) + " x levels with " +
                                  std::to_string(s->I.expan_height) + " <= levels <= 11");
  Shared g(s->gate);
  std::shared_ptr<KeyStore> ks = s->keys;
  Exclusive kx(ks->gate);  // no batch of any server sharing the store is in flight
  hipStream_t st = s->lanes[0]->st;
  ClientKeys &ck = ks->clients[id];
  const size_t levels = n / level_words;
  if (ck.gk && ck.gk_levels < levels) {
    dfree(ck.gk);
    ck.gk = nullptr;
    ck.gk_levels = 0;
  }
  if (!ck.gk) {
    int rc = dalloc(&ck.gk, levels * kKeyLevelWords);
    if (rc) {
      if (!ck.rk) ks->clients.erase(id);
      return rc;
    }
  }
  ck.gk_levels = (uint32_t)levels;
  // each level's 4N words into the first half of its kKeyLevelWords block,
  // stored times N^-1 per limb (the expansion's inverse transforms then need
  // no scaling), with their Shoup companions in the second half
  if (n) {  // a database without expansion levels takes an empty key
    HIPCHK(hipMemcpy2DAsync(ck.gk, kKeyLevelWords * 8, keys, level_words * 8, level_words * 8, levels,
                            hipMemcpyHostToDevice, st));
    k_key_prep<<<(uint32_t)((n + 255) / 256), 256, 0, st>>>(ck.gk, n, s->Tq.n_inv, s->Tp.n_inv, s->mq, s->mp);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));
  return ONION_OK;
}

int onion_gpu_set_rgsw_key(onion_server *s, uint64_t id, const uint64_t *rows, size_t n) {
  if (!s || !rows) return fail(ONION_EINVAL, "null argument");
  if (n != s->I.rgsw_key_words)
    return fail(ONION_EINVAL, "RGSW key size " + std::to_string(n) + " != " +
                                  std::to_string(s->I.rgsw_key_words));
  Shared g(s->gate);
  std::shared_ptr<KeyStore> ks = s->keys;
  Exclusive kx(ks->gate);
  hipStream_t st = s->lanes[0]->st;
  ClientKeys &ck = ks->clients[id];
  if (!ck.rk) {
    int rc = dalloc(&ck.rk, n);
    if (rc) {
      if (!ck.gk) ks->clients.erase(id);
      return rc;
    }
  }
  HIPCHK(hipMemcpyAsync(ck.rk, rows, n * 8, hipMemcpyHostToDevice, st));
  // GSWEval::gsw_ntt_negacyclic_harvey: every half-row to NTT form
  k_ntt_fwd<8><<<(uint32_t)(n / kN), kNttThreads, 0, st>>>(ck.rk, ck.rk, s->Tq);
  if (s->conv_fold)  // stored folded: the conversion multiplies the last row by the leaf itself
    k_rgsw_fold_last<<<dim3(4 * kN / 256, 1), 256, 0, st>>>(ck.rk, 0, s->P.gsw_l_key, s->I.base_log2_key, s->mq);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return ONION_OK;
}

int onion_gpu_drop_client(onion_server *s, uint64_t id) {
  if (!s) return fail(ONION_EINVAL, "null server");
  Shared g(s->gate);
  std::shared_ptr<KeyStore> ks = s->keys;
  Exclusive kx(ks->gate);
  auto it = ks->clients.find(id);
  if (it == ks->clients.end()) return fail(ONION_ENOKEY, "unknown client " + std::to_string(id));
  dfree(it->second.gk);
  dfree(it->second.rk);
  ks->clients.erase(it);
  return ONION_OK;
}

int onion_gpu_share_keys(onion_server *dst, onion_server *src) {
  if (!dst || !src) return fail(ONION_EINVAL, "null argument");
  if (dst == src) return ONION_OK;
  static std::mutex share_mu;  // one sharing change at a time: no lock-order cycle between servers
  std::lock_guard<std::mutex> sm(share_mu);
  Exclusive x(dst->gate);
  Shared g(src->gate);
  if (dst->device != src->device) return fail(ONION_EINVAL, "servers on different devices");
  if (dst->I.q != src->I.q || dst->I.p != src->I.p || dst->P.gsw_l_key != src->P.gsw_l_key)
    return fail(ONION_EINVAL, "servers differ in moduli or key gadget");
  if (dst->keys == src->keys) return ONION_OK;
  {
    Shared kg(dst->keys->gate);
    if (!dst->keys->clients.empty())
      return fail(ONION_EINVAL, "this database already holds client keys; drop them first");
  }
  dst->keys = src->keys;
  return ONION_OK;
}

int onion_gpu_answer(onion_server *s, uint64_t id, const uint64_t *query, uint64_t *resp) {
  if (!s || !query || !resp) return fail(ONION_EINVAL, "null argument");
  Shared g(s->gate);
  Shared kg(s->keys->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  return answer_on_lane(s, L, 1, &id, query, resp, nullptr);
}

int onion_gpu_answer_batch(onion_server *s, uint32_t n, const uint64_t *ids, const uint64_t *queries,
                           uint64_t *resps) {
  if (!s || (n && (!ids || !queries || !resps))) return fail(ONION_EINVAL, "null argument");
  Shared g(s->gate);
  Shared kg(s->keys->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  return answer_on_lane(s, L, n, ids, queries, resps, nullptr);
}

int onion_gpu_answer_batch_lane(onion_server *s, uint32_t lane, uint32_t n, const uint64_t *ids,
                                const uint64_t *queries, uint64_t *resps, onion_stage_times *times) {
  if (!s || (n && (!ids || !queries || !resps))) return fail(ONION_EINVAL, "null argument");
  Shared g(s->gate);
  if (lane >= s->lanes.size())
    return fail(ONION_EINVAL, "lane " + std::to_string(lane) + " of " + std::to_string(s->lanes.size()));
  Shared kg(s->keys->gate);
  Lane *L = s->lanes[lane].get();
  std::lock_guard<std::mutex> lk(L->mu);
  return answer_on_lane(s, L, n, ids, queries, resps, times);
}

int onion_gpu_answer_batch_noalloc(onion_server *s, uint32_t n, const uint64_t *ids, const uint64_t *queries,
                                   uint64_t *resps) {
  if (!s || !ids || !queries || !resps) return fail(ONION_EINVAL, "null argument");
  if (n == 0 || n > kMaxBatch) return fail(ONION_EINVAL, "n must be in [1, 64]");
  const uint64_t need = static_bytes(s, n);
  if (need > kStaticScratchBytes) {
    uint32_t fit = 0;
    while (fit < kMaxBatch && static_bytes(s, fit + 1) <= kStaticScratchBytes) ++fit;
    return fail(ONION_EINVAL, std::to_string(n) + " queries need " + std::to_string(need) +
                                  " bytes of scratch; the static scratch holds " +
                                  std::to_string(kStaticScratchBytes) + " (" + std::to_string(fit) +
                                  " queries of this database)");
  }
  Shared g(s->gate);
  Shared kg(s->keys->gate);
  int rc;
  if ((rc = check_db(s))) return rc;
  if (s->device < 0 || s->device >= kMaxDevices) return fail(ONION_EHIP, "device index beyond the static scratch table");
  void *base = g_static_ptr[s->device];
  if (!base) return fail(ONION_EHIP, g_static_err[s->device]);
  std::lock_guard<std::mutex> dev(g_static_mu[s->device]);
  Lane *L = s->slane.get();
  std::lock_guard<std::mutex> lk(L->mu);
  carve_static(s, L, n, (uint8_t *)base);
  return answer_chunk(s, L, n, ids, queries, resps, nullptr, true);
}

int onion_gpu_pack_response(onion_server *s, const uint64_t *resp, uint8_t *out, size_t cap,
                            size_t *n_bytes) {
  if (!s || !resp || !out) return fail(ONION_EINVAL, "null argument");
  if (cap < s->I.response_bytes) return fail(ONION_EINVAL, "output buffer too small");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  const uint32_t wbits = response_bits(s->I.small_q);
  HIPCHK(hipMemcpyAsync(s->io, resp, 2 * kN * 8, hipMemcpyHostToDevice, L->st));
  k_pack_response<<<(2 * kN / 8 + 255) / 256, 256, 0, L->st>>>(s->io, s->pack, 2 * kN, wbits);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->pack, s->I.response_bytes, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  if (n_bytes) *n_bytes = s->I.response_bytes;
  return ONION_OK;
}

int onion_gpu_get_tuning(const onion_server *s, onion_tuning *out) {
  if (!s || !out) return fail(ONION_EINVAL, "null argument");
  Shared g(s->gate);  // set_tuning and the autotuner write it exclusively
  *out = s->tune;
  return ONION_OK;
}

int onion_gpu_set_tuning(onion_server *s, const onion_tuning *t) {
  if (!s || !t) return fail(ONION_EINVAL, "null argument");
  if (t->fd_rows_b1 < 32 || t->fd_rows_b1 > 1024 || t->fd_rows_b1 % 32)
    return fail(ONION_EINVAL, "fd_rows_b1 must be a multiple of 32 in [32, 1024]");
  if (t->fd_tile_rows != 2 && t->fd_tile_rows != 4) return fail(ONION_EINVAL, "fd_tile_rows must be 2 or 4");
  if (t->fd_b1_variant > 3) return fail(ONION_EINVAL, "fd_b1_variant must be 0..3");
  if (t->use_graph > 2) return fail(ONION_EINVAL, "use_graph must be 0, 1 or 2");
  if (t->fd_mfma_variant > 2) return fail(ONION_EINVAL, "fd_mfma_variant must be 0, 1 or 2");
  if (t->db_planes > 1) return fail(ONION_EINVAL, "db_planes must be 0 or 1");
  if (t->ntt_radix != 8 && t->ntt_radix != 4) return fail(ONION_EINVAL, "ntt_radix must be 8 or 4");
  if (t->batch_split > 1) return fail(ONION_EINVAL, "batch_split must be 0 or 1");
  Exclusive x(s->gate);
  const uint32_t had = s->tune.db_planes;
  s->tune = *t;
  s->tune.fuse_min = std::max(1u, t->fuse_min);
  ++s->tune_epoch;  // graphs captured before hold the old kernels and buffers
  int rc;
  if ((rc = sync_planes(s))) {
    s->tune.db_planes = had && s->dbp ? 1 : 0;
    return rc;
  }
  return ONION_OK;
}

int onion_gpu_autotune(onion_server *s, const uint32_t *sizes, uint32_t n_sizes, uint32_t reps,
                       onion_tuning *chosen) {
  if (!s || !sizes) return fail(ONION_EINVAL, "null argument");
  if (n_sizes < 1 || n_sizes > 16) return fail(ONION_EINVAL, "n_sizes must be in [1, 16]");
  for (uint32_t i = 0; i < n_sizes; ++i)
    if (sizes[i] < 1 || sizes[i] > 4 * kMaxBatch) return fail(ONION_EINVAL, "batch sizes must be in [1, 256]");
  if (reps < 1 || reps > 20) return fail(ONION_EINVAL, "reps must be in [1, 20]");
  Exclusive x(s->gate);
  int rc;
  if ((rc = check_db(s))) return rc;
  return autotune_impl(s, sizes, n_sizes, reps, chosen);
}

int onion_gpu_last_stage_times(const onion_server *s, onion_stage_times *out) {
  if (!s || !out) return fail(ONION_EINVAL, "null argument");
  std::lock_guard<std::mutex> g(s->last_mu);
  *out = s->last;
  return ONION_OK;
}

// ---- taps (lane 0) ------------------------------------------------------------------

int onion_gpu_tap_expand(onion_server *s, uint64_t id, const uint64_t *query, uint64_t *leaves) {
  if (!s || !query || !leaves) return fail(ONION_EINVAL, "null argument");
  Shared g(s->gate);
  Shared kg(s->keys->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  int rc;
  if ((rc = ensure_scratch(s, L, 1)) || (rc = bind_keys(s, L, 1, &id, false))) return rc;
  HIPCHK(hipMemcpyAsync(L->trees + 2 * kN, query, 2 * kN * 8, hipMemcpyHostToDevice, L->st));
  if (run_expand(s, L, 1))  // leaves in coefficient form, as the reference's expansion returns them
    k_cts_to_coeff<<<dim3(2 * s->I.useful_cnt, 1), kNttThreads, 0, L->st>>>(L->trees + s->w * 2 * kN, 0, s->Tq);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(leaves, L->trees + s->w * 2 * kN, (size_t)s->I.useful_cnt * 2 * kN * 8,
                        hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  return ONION_OK;
}

int onion_gpu_tap_convert(onion_server *s, uint64_t id, const uint64_t *leaves, uint64_t *rgsw) {
  if (!s || !leaves || !rgsw) return fail(ONION_EINVAL, "null argument");
  Shared g(s->gate);
  Shared kg(s->keys->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  int rc;
  if ((rc = ensure_scratch(s, L, 1)) || (rc = bind_keys(s, L, 1, &id, true))) return rc;
  const uint32_t nd = s->I.ndim - 1;
  const size_t n_ct = (size_t)nd * s->P.gsw_l;
  if (nd == 0) return ONION_OK;
  HIPCHK(hipMemcpyAsync(s->io, leaves + (size_t)s->P.fst_dim * 2 * kN, n_ct * 2 * kN * 8,
                        hipMemcpyHostToDevice, L->st));
  run_convert(s, L, 1, s->io, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(rgsw, L->rgsw, (size_t)nd * 2 * s->P.gsw_l * 2 * kN * 8, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  return ONION_OK;
}

int onion_gpu_tap_first_dim(onion_server *s, const uint64_t *sel, uint64_t *out) {
  if (!s || !sel || !out) return fail(ONION_EINVAL, "null argument");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  int rc;
  if ((rc = check_db(s)) || (rc = ensure_scratch(s, L, 1))) return rc;
  HIPCHK(hipMemcpyAsync(s->io, sel, (size_t)s->P.fst_dim * 2 * kN * 8, hipMemcpyHostToDevice, L->st));
  run_first_dim(s, L, 1, s->io, 0, false);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, L->fd, (size_t)s->P.other_dim * 2 * kN * 8, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  return ONION_OK;
}

int onion_gpu_tap_first_dim_batch(onion_server *s, uint32_t nq, const uint64_t *sel, uint64_t *out) {
  if (!s || !sel || !out) return fail(ONION_EINVAL, "null argument");
  if (nq < 1 || nq > kMaxBatch) return fail(ONION_EINVAL, "nq must be in [1, 64]");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  int rc;
  if ((rc = check_db(s)) || (rc = ensure_scratch(s, L, nq))) return rc;
  const size_t sel_bytes = (size_t)s->P.fst_dim * 2 * kN * 8, out_bytes = (size_t)s->P.other_dim * 2 * kN * 8;
  HIPCHK(hipMemcpy2DAsync(L->trees, s->tree_words * 8, sel, sel_bytes, sel_bytes, nq, hipMemcpyHostToDevice, L->st));
  run_first_dim(s, L, nq, L->trees, s->tree_words, false);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy2DAsync(out, out_bytes, L->fd, s->fd_words * 8, out_bytes, nq, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  return ONION_OK;
}

int onion_gpu_tap_mux(onion_server *s, const uint64_t *cts, uint32_t n, const uint64_t *rgsw, uint64_t *out) {
  if (!s || !cts || !rgsw || !out) return fail(ONION_EINVAL, "null argument");
  if (n < 2 || (n & 1) || n > s->P.other_dim) return fail(ONION_EINVAL, "n must be even and <= other_dim");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  int rc;
  if ((rc = ensure_scratch(s, L, 1))) return rc;
  HIPCHK(hipMemcpyAsync(L->fd, cts, (size_t)n * 2 * kN * 8, hipMemcpyHostToDevice, L->st));
  HIPCHK(hipMemcpyAsync(L->rgsw, rgsw, (size_t)2 * s->P.gsw_l * 2 * kN * 8, hipMemcpyHostToDevice, L->st));
  run_mux_level(s, L, 1, n, L->rgsw);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, L->fd, (size_t)(n / 2) * 2 * kN * 8, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  return ONION_OK;
}

int onion_gpu_tap_ntt(onion_server *s, int which, int inverse, uint64_t *polys, uint32_t n) {
  if (!s || !polys) return fail(ONION_EINVAL, "null argument");
  if (which != 0 && which != 1) return fail(ONION_EINVAL, "which must be 0 (q) or 1 (p)");
  if ((size_t)n * kN > s->io_words) return fail(ONION_EINVAL, "too many polynomials");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  const NttTab &T = which == 0 ? s->Tq : s->Tp;
  HIPCHK(hipMemcpyAsync(s->io, polys, (size_t)n * kN * 8, hipMemcpyHostToDevice, L->st));
  const bool r4 = s->tune.ntt_radix == 4;
  if (inverse && r4)
    k_ntt_inv<4><<<n, kNttThreads, 0, L->st>>>(s->io, s->io, T);
  else if (inverse)
    k_ntt_inv<8><<<n, kNttThreads, 0, L->st>>>(s->io, s->io, T);
  else if (r4)
    k_ntt_fwd<4><<<n, kNttThreads, 0, L->st>>>(s->io, s->io, T);
  else
    k_ntt_fwd<8><<<n, kNttThreads, 0, L->st>>>(s->io, s->io, T);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(polys, s->io, (size_t)n * kN * 8, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  return ONION_OK;
}

int onion_gpu_tap_galois(onion_server *s, uint64_t id, uint32_t level, const uint64_t *ct, uint64_t *out) {
  if (!s || !ct || !out) return fail(ONION_EINVAL, "null argument");
  if (level >= s->I.expan_height) return fail(ONION_EINVAL, "level >= expansion height");
  Shared g(s->gate);
  Shared kg(s->keys->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  int rc;
  if ((rc = ensure_scratch(s, L, 1)) || (rc = bind_keys(s, L, 1, &id, false))) return rc;
  // place ct at node 2^level; its left child is ct + Subs(ct), so Subs(ct) = left - ct
  const uint64_t node = uint64_t(1) << level;
  HIPCHK(hipMemcpyAsync(L->trees + node * 2 * kN, ct, 2 * kN * 8, hipMemcpyHostToDevice, L->st));
  expand_level(s, L, 1, level, 1, false, false);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, L->trees + 2 * node * 2 * kN, 2 * kN * 8, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  const uint64_t q = s->I.q;
  for (size_t i = 0; i < 2 * (size_t)kN; ++i) out[i] = out[i] >= ct[i] ? out[i] - ct[i] : out[i] + q - ct[i];
  return ONION_OK;
}

int onion_gpu_tap_copy(onion_server *s, uint64_t n_bytes, uint64_t *copied) {
  if (!s) return fail(ONION_EINVAL, "null server");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  int rc;
  if ((rc = ensure_scratch(s, L, 1))) return rc;
  const uint64_t n = std::min<uint64_t>({n_bytes, L->cap * s->digit_words * 8, s->I.num_pt * kN * 8}) / 16;
  if (n) {
    k_copy<<<1024, 256, 0, L->st>>>(reinterpret_cast<const uint4 *>(s->db), reinterpret_cast<uint4 *>(L->digits), n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(L->st));
  }
  if (copied) *copied = n * 16;
  return ONION_OK;
}

int onion_gpu_tap_valu_probe(onion_server *s, uint32_t blocks, uint32_t rounds, double *ms) {
  if (!s || !blocks || !rounds) return fail(ONION_EINVAL, "null server or empty probe");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  int rc;
  if ((rc = ensure_scratch(s, L, 1))) return rc;
  if ((uint64_t)blocks * kNttThreads > L->cap * s->digit_words) return fail(ONION_EINVAL, "too many probe blocks");
  HIPCHK(hipEventRecord(L->ev[6], L->st));
  k_valu_probe<<<blocks, kNttThreads, 0, L->st>>>(s->Tq, rounds, L->digits);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(L->ev[7], L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  float f = 0;
  HIPCHK(hipEventElapsedTime(&f, L->ev[6], L->ev[7]));
  if (ms) *ms = f;
  return ONION_OK;
}

int onion_gpu_tap_mod_switch(onion_server *s, const uint64_t *ct, uint64_t *out) {
  if (!s || !ct || !out) return fail(ONION_EINVAL, "null argument");
  Shared g(s->gate);
  Lane *L = s->lanes[0].get();
  std::lock_guard<std::mutex> lk(L->mu);
  HIPCHK(hipMemcpyAsync(s->io, ct, 2 * kN * 8, hipMemcpyHostToDevice, L->st));
  run_mod_switch(s, L, 1, s->io, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->io, 2 * kN * 8, hipMemcpyDeviceToHost, L->st));
  HIPCHK(hipStreamSynchronize(L->st));
  return ONION_OK;
}

}  // extern "C"
