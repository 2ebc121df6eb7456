// Serving loop (declared in include/tc.h, tc_server_*).
//
// One worker thread per server runs the loop
//   retire  slots whose completion event has fired: per-vertex counts and
//           totals are handed to their requests, the batch's arena block is
//           released, waiters are woken
//   admit   the next batch (first in first out, host/admission.h): the
//           longest prefix of the queue whose block fits the arena of
//           hbm_budget bytes
//   launch  on a free slot: the batch is packed into the arena block's
//           pinned host mirror in the batched counter's layout, uploaded with
//           one copy on the slot's stream, counted by the batched counter
//           with its scratch in the rest of the same block and its counts
//           copied back with one copy, then an event is recorded
// so requests that arrive while batches run are admitted as blocks free up
// and batches on different slots overlap on the device.  Every byte of
// device memory a batch uses is in its block, so hbm_budget bounds them all,
// and nothing is allocated on the device after tc_server_create.
// Requests are checked on the host when submitted (a simple undirected CSR:
// ids in range, rows strictly increasing, no self-loops, every arc with its
// reverse), so one malformed request can never touch another's block.
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
#include "host/csr_check.h"
#include "tc.h"

// tc_capi.hip: the batched count with the scratch at a given address.
namespace tci {
int batch_count_at(tc_ctx* c, void* base, size_t base_bytes, int32_t G, const int64_t* d_tab, int64_t nv_total,
                   int64_t nnz_total, const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat,
                   int64_t* d_totals, void* stream);
}

namespace {

using clk = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  int32_t nv = 0;
  int64_t nnz = 0;
  std::vector<int32_t> off, ind;  // host copy until packed
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
  bool busy = false;
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
  char* h_mirror = nullptr;
  std::unique_ptr<tcs::Admission> adm;
  std::vector<Slot> slots;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
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

// Packs sl.batch into its block and queues upload, count and download on the
// slot's stream.  Returns TC_OK or a TC_E* code (nothing then pending on the
// stream that refers to the block except what the caller synchronises).
int tc_server::launch(Slot& sl) {
  const int32_t G = (int32_t)sl.batch.size();
  int64_t nv_total = 0, nnz_total = 0;
  sl.vbase.assign(G, 0);
  for (int32_t g = 0; g < G; ++g) {
    sl.vbase[g] = nv_total;
    nv_total += sl.batch[g]->nv;
    nnz_total += sl.batch[g]->nnz;
  }
  if (nv_total >= INT32_MAX || nnz_total > INT32_MAX) return TC_ERANGE;
  std::vector<int32_t> nv(G);
  std::vector<int64_t> nnz(G);
  std::vector<const int32_t*> offp(G), indp(G);
  for (int32_t g = 0; g < G; ++g) {
    nv[g] = sl.batch[g]->nv;
    nnz[g] = sl.batch[g]->nnz;
    offp[g] = sl.batch[g]->off.data();
    indp[g] = sl.batch[g]->ind.data();
  }
  if (tcs::batch_layout(G, nv_total, nnz_total).bytes > sl.blk_bytes)
    return TC_EINVAL;  // the admission bound guarantees this never happens
  char* h = h_mirror + sl.blk_off;
  char* d = d_arena + sl.blk_off;
  sl.L = tcs::pack_block(G, nv.data(), nnz.data(), offp.data(), indp.data(), h);
  for (Request* r : sl.batch) {
    std::vector<int32_t>().swap(r->off);
    std::vector<int32_t>().swap(r->ind);
  }
  hipStream_t s = sl.stream;
  if (nnz_total == 0) {  // no arcs: every count and total is zero
    std::memset(h + sl.L.out, 0, sl.L.res_end - sl.L.out);
    return hipEventRecord(sl.done, s) == hipSuccess ? TC_OK : TC_EHIP;
  }
  if (hipMemcpyAsync(d, h, sl.L.upload_end, hipMemcpyHostToDevice, s) != hipSuccess) return TC_EHIP;
  const int rc = tci::batch_count_at(sl.ctx, d + sl.L.scr, sl.L.bytes - sl.L.scr, G,
                                     reinterpret_cast<const int64_t*>(d + sl.L.tab), nv_total, nnz_total,
                                     reinterpret_cast<const int32_t*>(d + sl.L.off),
                                     reinterpret_cast<const int32_t*>(d + sl.L.ind),
                                     reinterpret_cast<int32_t*>(d + sl.L.out),
                                     reinterpret_cast<int64_t*>(d + sl.L.tot), s);
  if (rc != TC_OK) return rc;
  if (hipMemcpyAsync(h + sl.L.out, d + sl.L.out, sl.L.res_end - sl.L.out, hipMemcpyDeviceToHost, s) != hipSuccess)
    return TC_EHIP;
  if (hipEventRecord(sl.done, s) != hipSuccess) return TC_EHIP;
  return TC_OK;
}

// Hands results (or sl.err) to the batch's requests and frees its block.
void tc_server::retire(Slot& sl) {
  const auto now = clk::now();
  const char* h = h_mirror + sl.blk_off;
  const int32_t* out = reinterpret_cast<const int32_t*>(h + sl.L.out);
  const int64_t* tot = reinterpret_cast<const int64_t*>(h + sl.L.tot);
  std::lock_guard<std::mutex> lk(mu);
  for (size_t g = 0; g < sl.batch.size(); ++g) {
    Request* r = sl.batch[g];
    if (sl.err == TC_OK) {
      r->out.assign(out + sl.vbase[g], out + sl.vbase[g] + r->nv);
      r->total = tot[g];
    } else {
      r->err = sl.err;
    }
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
  }
  if (d_arena) (void)hipFree(d_arena);
  if (h_mirror) (void)hipHostFree(h_mirror);
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
  if (rc == TC_OK && hipHostMalloc(reinterpret_cast<void**>(&s->h_mirror), budget, hipHostMallocDefault) != hipSuccess)
    rc = TC_ENOMEM;
  for (auto& sl : s->slots) {
    if (rc != TC_OK) break;
    if (tc_ctx_create(cfg->device, &sl.ctx) != TC_OK ||
        hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      rc = TC_EHIP;
  }
  if (rc != TC_OK) {
    s->release_all();
    return rc;
  }
  s->adm = std::make_unique<tcs::Admission>(budget);
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
  if (s->worker.joinable()) s->worker.join();
  s->release_all();
  delete s;
  return TC_OK;
}

int64_t tc_server_submit(tc_server* s, int32_t nv, int64_t nnz, const int32_t* h_off, const int32_t* h_ind) {
  if (!s || nv < 0 || nnz < 0 || nnz > INT32_MAX || !h_off || (nnz > 0 && !h_ind)) return TC_EINVAL;
  if (h_off[0] != 0 || h_off[nv] != nnz) return TC_EINVAL;
  if (tcs::csr_fault(nv, nnz, h_off, h_ind)) {  // checked before anything is queued or launched
    std::lock_guard<std::mutex> lk(s->mu);
    s->st.rejected++;
    return TC_EGRAPH;
  }
  const size_t share = tcs::request_share(nv, nnz);
  auto r = std::make_unique<Request>();
  r->nv = nv;
  r->nnz = nnz;
  r->off.assign(h_off, h_off + (size_t)nv + 1);
  r->ind.assign(h_ind, h_ind + nnz);
  r->t_submit = clk::now();
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return TC_EINVAL;
    if (!s->adm->fits_ever(share)) {
      s->st.rejected++;
      return TC_ENOMEM;
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
  std::vector<int32_t>().swap(r->off);
  std::vector<int32_t>().swap(r->ind);
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
