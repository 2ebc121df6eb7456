// Serving loop (declared in include/tc.h, tc_server_*).
//
// tc_server_submit copies each request once, into a slab of a pinned host
// pool (as large as the device arena): [offsets | ids | counts].  Large
// requests are copied by several threads.  One worker thread per server
// then runs the loop
//   retire  slots whose completion event has fired: each request's counts
//           are taken from its slab and its total from the slot, the slab
//           and the batch's arena block are released, waiters are woken
//   admit   the next batch (first in first out, host/admission.h): the
//           longest prefix of the queue whose block fits the arena of
//           hbm_budget bytes
//   launch  on a free slot: the table and every request's slab regions are
//           uploaded straight into the block's layout by DMA copies on the
//           slot's stream (no host packing), the batch is checked and
//           counted by the batched counter with its scratch in the rest of
//           the same block, the counts go back by DMA into the slabs, then
//           an event is recorded
// so requests that arrive while batches run are admitted as blocks free up
// and batches on different slots overlap on the device.  Every byte of
// device memory a batch uses is in its block, so hbm_budget bounds them all,
// and nothing is allocated on the device after tc_server_create.
// A request's offsets are checked on the host when it is submitted (from 0
// to nnz, never decreasing: O(nv)); the rest of the simple-graph contract
// (ids in range, rows strictly increasing, no self-loops, every arc with its
// reverse) is checked on the device, per graph, before the batch is counted
// (k_validate_batch).  A graph that breaks it has its ids rewritten to
// self-loops of their own rows, so the count reads none of its rows and no
// other graph of the batch is touched, and its request ends with TC_EGRAPH.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host/admission.h"
#include "tc.h"
#include "tc_device.h"

// tc_capi.hip: the batched count with the scratch at a given address.
namespace tci {
int batch_count_at(tc_ctx* c, void* base, size_t base_bytes, int32_t G, const int64_t* d_tab, int64_t nv_total,
                   int64_t nnz_total, const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat,
                   int64_t* d_totals, void* stream);
}

namespace {

using clk = std::chrono::steady_clock;

// A request's slab in the pinned pool: offsets, ids and counts, each region
// kArenaAlign aligned.
struct Slab {
  size_t off, ind, out, bytes;
};
inline Slab slab_layout(int64_t nv, int64_t nnz) {
  Slab b;
  b.off = 0;
  b.ind = tcs::arena_round(4 * ((size_t)nv + 1));
  b.out = b.ind + tcs::arena_round(4 * (size_t)nnz);
  b.bytes = b.out + tcs::arena_round(4 * (size_t)(nv > 0 ? nv : 1));
  return b;
}

struct Request {
  int64_t id = 0;
  int32_t nv = 0;
  int64_t nnz = 0;
  size_t slab = SIZE_MAX;  // offset of its slab in the pinned pool, until retired
  bool done = false;
  int err = TC_OK;
  std::vector<int32_t> out;
  int64_t total = 0;
  clk::time_point t_submit, t_done;
};

struct Slot {
  tc_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  int64_t* h_tab = nullptr;  // pinned: the batch's table, then its totals
  bool busy = false;
  bool empty = false;  // a batch without arcs: nothing launched, every count 0
  std::vector<Request*> batch;
  std::vector<int64_t> vbase;  // per request of the batch
  size_t blk_off = 0, blk_bytes = 0;
  tcs::BatchLayout L{};
  int err = TC_OK;
};

}  // namespace

struct tc_server {
  tc_server_cfg cfg{};
  char* d_arena = nullptr;
  char* h_pool = nullptr;  // pinned slabs of the requests not yet retired
  std::unique_ptr<tcs::Admission> adm;
  std::unique_ptr<tcs::Arena> pool;
  std::unique_ptr<tcs::CopyPool> copier;  // threads for large submits
  std::vector<Slot> slots;

  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_pool;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_id = 0;
  bool stop = false;
  tc_server_stats st{};
  std::thread worker;

  void run();
  int launch(Slot& sl);
  void retire(Slot& sl);
  void release_all();
};

// Queues the batch's upload (table and slabs, by DMA into the block's
// layout), device check, count and download on the slot's stream.  Returns
// TC_OK or a TC_E* code (nothing then pending on the stream that refers to
// the block except what the caller synchronises).
int tc_server::launch(Slot& sl) {
  const int32_t G = (int32_t)sl.batch.size();
  int64_t nv_total = 0, nnz_total = 0;
  sl.vbase.assign(G, 0);
  std::vector<int32_t> nv(G);
  std::vector<int64_t> nnz(G);
  for (int32_t g = 0; g < G; ++g) {
    sl.vbase[g] = nv_total;
    nv[g] = sl.batch[g]->nv;
    nnz[g] = sl.batch[g]->nnz;
    nv_total += nv[g];
    nnz_total += nnz[g];
  }
  if (nv_total >= INT32_MAX || nnz_total > INT32_MAX) return TC_ERANGE;
  sl.L = tcs::batch_layout(G, nv_total, nnz_total);
  if (sl.L.bytes > sl.blk_bytes) return TC_EINVAL;  // the admission bound guarantees this never happens
  hipStream_t s = sl.stream;
  sl.empty = nnz_total == 0;
  if (sl.empty) {  // no arcs: every count and total is zero (set at retire)
    return hipEventRecord(sl.done, s) == hipSuccess ? TC_OK : TC_EHIP;
  }
  char* d = d_arena + sl.blk_off;
  tcs::write_tab(G, nv.data(), nnz.data(), sl.h_tab);
  if (hipMemcpyAsync(d + sl.L.tab, sl.h_tab, 16 * ((size_t)G + 1), hipMemcpyHostToDevice, s) != hipSuccess)
    return TC_EHIP;
  for (const tcs::UploadCopy& c : tcs::upload_plan(G, nv.data(), nnz.data(), sl.L)) {
    const Request* r = sl.batch[c.g];
    const Slab b = slab_layout(r->nv, r->nnz);
    const char* src = h_pool + r->slab + (c.part ? b.ind : b.off);
    if (hipMemcpyAsync(d + c.dst, src, c.bytes, hipMemcpyHostToDevice, s) != hipSuccess) return TC_EHIP;
  }
  const int64_t* d_tab = reinterpret_cast<const int64_t*>(d + sl.L.tab);
  int32_t* d_off = reinterpret_cast<int32_t*>(d + sl.L.off);
  int32_t* d_ind = reinterpret_cast<int32_t*>(d + sl.L.ind);
  int32_t* d_out = reinterpret_cast<int32_t*>(d + sl.L.out);
  int64_t* d_tot = reinterpret_cast<int64_t*>(d + sl.L.tot);
  // graphs that break the contract: total -1, ids made harmless
  if (tcd::launch_validate_batch(d_tab, G, d_off, d_ind, nnz_total, d_tot, tc_device_cus(sl.ctx), s) != hipSuccess)
    return TC_EHIP;
  const int rc = tci::batch_count_at(sl.ctx, d + sl.L.scr, sl.L.bytes - sl.L.scr, G, d_tab, nv_total, nnz_total,
                                     d_off, d_ind, d_out, nullptr, s);
  if (rc != TC_OK) return rc;
  if (tcd::launch_graph_totals(d_tab, G, d_out, d_tot, s, true, nv_total) != hipSuccess) return TC_EHIP;
  // counts into each slab, totals over the slot's table
  for (int32_t g = 0; g < G; ++g) {
    const Request* r = sl.batch[g];
    if (!r->nv) continue;
    if (hipMemcpyAsync(h_pool + r->slab + slab_layout(r->nv, r->nnz).out, d_out + sl.vbase[g], 4 * (size_t)r->nv,
                       hipMemcpyDeviceToHost, s) != hipSuccess)
      return TC_EHIP;
  }
  if (hipMemcpyAsync(sl.h_tab, d_tot, 8 * (size_t)G, hipMemcpyDeviceToHost, s) != hipSuccess) return TC_EHIP;
  if (hipEventRecord(sl.done, s) != hipSuccess) return TC_EHIP;
  return TC_OK;
}

// Hands results (or sl.err) to the batch's requests and frees its block and
// their slabs.
void tc_server::retire(Slot& sl) {
  const auto now = clk::now();
  const bool empty = sl.empty;
  // counts out of the slabs before the lock (they are the requests' own)
  for (size_t g = 0; g < sl.batch.size(); ++g) {
    Request* r = sl.batch[g];
    const int64_t tot = empty ? 0 : sl.h_tab[g];
    if (sl.err == TC_OK && tot >= 0) {
      if (empty) {
        r->out.assign((size_t)r->nv, 0);
      } else {
        const int32_t* o = reinterpret_cast<const int32_t*>(h_pool + r->slab + slab_layout(r->nv, r->nnz).out);
        r->out.assign(o, o + r->nv);
      }
      r->total = tot;
    }
  }
  std::lock_guard<std::mutex> lk(mu);
  for (size_t g = 0; g < sl.batch.size(); ++g) {
    Request* r = sl.batch[g];
    const int64_t tot = empty ? 0 : sl.h_tab[g];
    if (sl.err != TC_OK) {
      r->err = sl.err;
    } else if (tot < 0) {  // set aside by the device check
      r->err = TC_EGRAPH;
      st.rejected++;
    }
    pool->release(r->slab, slab_layout(r->nv, r->nnz).bytes);
    r->slab = SIZE_MAX;
    r->done = true;
    r->t_done = now;
    st.completed++;
    st.in_flight--;
  }
  adm->arena().release(sl.blk_off, sl.blk_bytes);
  sl.batch.clear();
  sl.busy = false;
  sl.err = TC_OK;
  st.bytes_in_use = (int64_t)adm->arena().used();
  cv_done.notify_all();
  cv_pool.notify_all();
}

void tc_server::run() {
  (void)hipSetDevice(cfg.device);
  std::vector<int64_t> ids;
  for (;;) {
    bool progressed = false;
    for (auto& sl : slots) {
      if (!sl.busy) continue;
      const hipError_t e = hipEventQuery(sl.done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) sl.err = TC_EHIP;
      retire(sl);
      progressed = true;
    }
    for (auto& sl : slots) {
      if (sl.busy) continue;
      {
        std::lock_guard<std::mutex> lk(mu);
        size_t bytes = 0;
        const size_t off = adm->admit_batch(cfg.max_batch_graphs, ids, bytes);
        st.waiting = (int64_t)adm->waiting();
        if (off == SIZE_MAX) break;
        sl.blk_off = off;
        sl.blk_bytes = bytes;
        sl.batch.clear();
        for (int64_t id : ids) sl.batch.push_back(reqs[id].get());
        st.batches++;
        st.in_flight += (int64_t)ids.size();
        st.bytes_in_use = (int64_t)adm->arena().used();
        st.peak_bytes = (int64_t)adm->arena().peak();
      }
      sl.busy = true;
      sl.err = launch(sl);
      if (sl.err != TC_OK) {
        // Whatever was queued may still read the block: drain before freeing.
        (void)hipStreamSynchronize(sl.stream);
        retire(sl);
      }
      progressed = true;
    }
    if (progressed) continue;
    std::unique_lock<std::mutex> lk(mu);
    bool busy_any = false;
    for (auto& sl : slots) busy_any |= sl.busy;
    if (stop && adm->waiting() == 0 && !busy_any) break;
    if (busy_any) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(50));
    } else if (adm->waiting() == 0) {
      cv_work.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

void tc_server::release_all() {
  (void)hipSetDevice(cfg.device);
  for (auto& sl : slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.ctx) tc_ctx_destroy(sl.ctx);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.h_tab) (void)hipHostFree(sl.h_tab);
  }
  if (d_arena) (void)hipFree(d_arena);
  if (h_pool) (void)hipHostFree(h_pool);
}

extern "C" {

int tc_server_create(const tc_server_cfg* cfg, tc_server** out) {
  if (!cfg || !out || cfg->hbm_budget < (int64_t)tcs::kArenaAlign || cfg->slots < 1 || cfg->slots > 8 ||
      cfg->max_batch_graphs < 1)
    return TC_EINVAL;
  *out = nullptr;
  if (hipSetDevice(cfg->device) != hipSuccess) return TC_EHIP;
  auto s = std::make_unique<tc_server>();
  s->cfg = *cfg;
  const size_t budget = (size_t)cfg->hbm_budget;
  s->slots.resize(cfg->slots);
  int rc = TC_OK;
  if (hipMalloc(reinterpret_cast<void**>(&s->d_arena), budget) != hipSuccess) rc = TC_ENOMEM;
  if (rc == TC_OK && hipHostMalloc(reinterpret_cast<void**>(&s->h_pool), budget, hipHostMallocDefault) != hipSuccess)
    rc = TC_ENOMEM;
  for (auto& sl : s->slots) {
    if (rc != TC_OK) break;
    if (tc_ctx_create(cfg->device, &sl.ctx) != TC_OK ||
        hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      rc = TC_EHIP;
    else if (hipHostMalloc(reinterpret_cast<void**>(&sl.h_tab), 16 * ((size_t)cfg->max_batch_graphs + 1),
                           hipHostMallocDefault) != hipSuccess)
      rc = TC_ENOMEM;
  }
  if (rc != TC_OK) {
    s->release_all();
    return rc;
  }
  s->adm = std::make_unique<tcs::Admission>(budget);
  s->pool = std::make_unique<tcs::Arena>(budget);
  s->copier = std::make_unique<tcs::CopyPool>(7);
  tc_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->run(); });
  *out = raw;
  return TC_OK;
}

int tc_server_destroy(tc_server* s) {
  if (!s) return TC_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_pool.notify_all();
  if (s->worker.joinable()) s->worker.join();
  s->release_all();
  delete s;
  return TC_OK;
}

int64_t tc_server_submit(tc_server* s, int32_t nv, int64_t nnz, const int32_t* h_off, const int32_t* h_ind) {
  if (!s || nv < 0 || nnz < 0 || nnz > INT32_MAX || !h_off || (nnz > 0 && !h_ind)) return TC_EINVAL;
  // offsets from 0 to nnz, never decreasing: every row inside the request's
  // ids (the device checks the rest before counting)
  bool offsets_ok = h_off[0] == 0 && h_off[nv] == nnz;
  for (int32_t v = 0; offsets_ok && v < nv; ++v) offsets_ok = h_off[v] <= h_off[v + 1];
  if (!offsets_ok) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->st.rejected++;
    return TC_EGRAPH;
  }
  const size_t share = tcs::request_share(nv, nnz);
  const Slab b = slab_layout(nv, nnz);
  auto r = std::make_unique<Request>();
  r->nv = nv;
  r->nnz = nnz;
  {
    // a slab in the pinned pool; while it is full (requests queued or in
    // flight hold it), wait for the worker to retire some
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->adm->fits_ever(share) || b.bytes > s->pool->capacity()) {
      s->st.rejected++;
      return TC_ENOMEM;
    }
    for (;;) {
      if (s->stop) return TC_EINVAL;
      r->slab = s->pool->alloc(b.bytes);
      if (r->slab != SIZE_MAX) break;
      s->cv_pool.wait(lk);
    }
  }
  // the one host copy of the request (the copy pool's threads when large)
  char* slab = s->h_pool + r->slab;
  s->copier->run({{slab + b.off, reinterpret_cast<const char*>(h_off), 4 * ((size_t)nv + 1)},
                  {slab + b.ind, reinterpret_cast<const char*>(h_ind), 4 * (size_t)nnz}});
  r->t_submit = clk::now();
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      s->pool->release(r->slab, b.bytes);
      s->cv_pool.notify_all();
      return TC_EINVAL;
    }
    id = s->next_id++;
    r->id = id;
    s->reqs[id] = std::move(r);
    s->adm->push(id, share);
    s->st.submitted++;
    s->st.waiting = (int64_t)s->adm->waiting();
  }
  s->cv_work.notify_all();
  return id;
}

int tc_server_wait(tc_server* s, int64_t id, int32_t* h_out, int64_t* h_total, int32_t timeout_ms,
                   double* latency_ms) {
  if (!s) return TC_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return TC_EINVAL;
  Request* r = it->second.get();
  auto pred = [&] { return r->done; };
  if (timeout_ms < 0) s->cv_done.wait(lk, pred);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) return 1;
  const int err = r->err;
  if (err == TC_OK) {
    if (h_out && r->nv) std::memcpy(h_out, r->out.data(), 4 * (size_t)r->nv);
    if (h_total) *h_total = r->total;
  }
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  s->reqs.erase(it);
  return err;
}

int tc_server_cancel(tc_server* s, int64_t id) {
  if (!s) return TC_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return TC_EINVAL;
  Request* r = it->second.get();
  if (r->done || !s->adm->remove(id)) return 1;  // admitted already: it completes as usual
  r->done = true;
  r->err = TC_ECANCELED;
  r->t_done = clk::now();
  s->pool->release(r->slab, slab_layout(r->nv, r->nnz).bytes);
  r->slab = SIZE_MAX;
  s->cv_pool.notify_all();
  s->st.canceled++;
  s->st.waiting = (int64_t)s->adm->waiting();
  s->cv_done.notify_all();
  return TC_OK;
}

int tc_server_get_stats(tc_server* s, tc_server_stats* st) {
  if (!s || !st) return TC_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->st;
  return TC_OK;
}

}  // extern "C"
