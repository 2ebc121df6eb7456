// Device-memory admission for the serving loop: a first-fit arena over a
// fixed byte budget, a FIFO of waiting requests, and the layout of one batch
// block.  Header-only and free of HIP so the policy can be exercised on the
// CPU with a fake budget.
//
// A batch of G requests occupies ONE arena block laid out as the batched
// counter's packed CSR (tab rows {v_base, arc_base}, concatenated offsets
// and ids, then the per-vertex counts and per-graph totals), so the block
// can be handed to the batched entry point as it is, followed by the
// scratch of the batch's count: the arena bounds all the device memory of
// the batches in flight.
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <system_error>
#include <thread>
#include <vector>

namespace tcs {

constexpr size_t kArenaAlign = 256;
inline size_t arena_round(size_t x) { return (x + kArenaAlign - 1) & ~(kArenaAlign - 1); }

// First-fit allocator over [0, capacity): free blocks kept by offset and
// coalesced on release.  Offsets are multiples of kArenaAlign.
class Arena {
 public:
  explicit Arena(size_t capacity) : cap_(capacity - capacity % kArenaAlign) {
    if (cap_) free_[0] = cap_;
  }
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t largest_free() const {
    size_t m = 0;
    for (const auto& f : free_) m = f.second > m ? f.second : m;
    return m;
  }

  // Offset of a block of at least `bytes`, or SIZE_MAX when nothing fits.
  size_t alloc(size_t bytes) {
    bytes = arena_round(bytes ? bytes : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes) free_[off + bytes] = len - bytes;
      used_ += bytes;
      if (used_ > peak_) peak_ = used_;
      return off;
    }
    return SIZE_MAX;
  }

  void release(size_t off, size_t bytes) {
    bytes = arena_round(bytes ? bytes : 1);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  size_t cap_;
  size_t used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;  // offset -> length
};

// Device scratch of a batched count of nv vertices and nnz arcs on the
// oriented scheme (tc_capi.hip's layout): 32 bytes per vertex (widened
// offsets, graph bases, N+ bounds, degrees), 4 per arc (the orientation)
// and fewer than 1/32 per arc more (piece, hub-row and long-row lists), and
// a fixed part (the piece list's slack of 65,536 entries, the scalars and
// ticket counters, 4,352 bytes, the rounding of ten regions) within
// kScratchFixed.
constexpr size_t kScratchFixed = 1060 * kArenaAlign;
inline size_t scratch_share(int64_t nv, int64_t nnz) {
  return 32 * (size_t)nv + 4 * (size_t)nnz + ((size_t)nnz + 31) / 32;
}
inline size_t scratch_bound(int64_t nv, int64_t nnz) { return kScratchFixed + scratch_share(nv, nnz); }

// Byte offsets of the regions of one batch block, each kArenaAlign aligned:
//   tab   int64[2(G+1)]     rows {v_base, arc_base}, row G = totals
//   off   int32[nv + G]     graph g's offsets at v_base[g] + g
//   ind   int32[nnz]        graph g's ids at arc_base[g]
//   out   int32[max(nv,1)]  per-vertex counts
//   tot   int64[G]          per-graph totals
//   scr   scratch_bound(nv, nnz) bytes of the count's scratch
// [0, upload_end) goes host -> device, [out, res_end) device -> host.
struct BatchLayout {
  size_t tab, off, ind, out, tot, res_end, scr, upload_end, bytes;
};
inline BatchLayout batch_layout(int64_t G, int64_t nv, int64_t nnz) {
  BatchLayout L;
  size_t p = 0;
  L.tab = p;
  p += arena_round(16 * (size_t)(G + 1));
  L.off = p;
  p += arena_round(4 * (size_t)(nv + G));
  L.ind = p;
  L.upload_end = p + 4 * (size_t)nnz;
  p += arena_round(4 * (size_t)nnz);
  L.out = p;
  p += arena_round(4 * (size_t)(nv > 0 ? nv : 1));
  L.tot = p;
  p += arena_round(8 * (size_t)G);
  L.res_end = p;
  L.scr = p;
  p += arena_round(scratch_bound(nv, nnz));
  L.bytes = p;
  return L;
}

// Table rows of a batch: tab[2g] = v_base, tab[2g+1] = arc_base, row G the
// totals.
inline void write_tab(int32_t G, const int32_t* nv, const int64_t* nnz, int64_t* tab) {
  int64_t v = 0, a = 0;
  for (int32_t g = 0; g < G; ++g) {
    tab[2 * g] = v;
    tab[2 * g + 1] = a;
    v += nv[g];
    a += nnz[g];
  }
  tab[2 * G] = v;
  tab[2 * G + 1] = a;
}

// Where each graph's offsets (part 0: 4(nv+1) bytes, to block byte
// L.off + 4(v_base + g)) and ids (part 1: 4 nnz bytes, to L.ind +
// 4 arc_base) land in a block laid out by batch_layout: the one placement
// both the host packer (pack_block) and the serving loop's uploads follow.
struct UploadCopy {
  int32_t g;
  int part;
  size_t dst, bytes;
};
inline std::vector<UploadCopy> upload_plan(int32_t G, const int32_t* nv, const int64_t* nnz, const BatchLayout& L) {
  std::vector<UploadCopy> plan;
  plan.reserve(2 * (size_t)G);
  int64_t v = 0, a = 0;
  for (int32_t g = 0; g < G; ++g) {
    plan.push_back({g, 0, L.off + 4 * (size_t)(v + g), 4 * ((size_t)nv[g] + 1)});
    if (nnz[g]) plan.push_back({g, 1, L.ind + 4 * (size_t)a, 4 * (size_t)nnz[g]});
    v += nv[g];
    a += nnz[g];
  }
  return plan;
}

// n copies (dst[i], src[i], bytes[i]) split by bytes over up to `threads`
// threads (one below 16 MiB in all); every byte lands where the serial
// copies would put it.
struct CopyTask {
  char* dst;
  const char* src;
  size_t n;
};
inline void copy_tasks(const std::vector<CopyTask>& in, int threads = 8) {
  size_t total = 0;
  for (const CopyTask& t : in) total += t.n;
  const int nt = total < (size_t(16) << 20) || threads < 2 ? 1 : threads;
  if (nt == 1) {
    for (const CopyTask& t : in) std::memcpy(t.dst, t.src, t.n);
    return;
  }
  // tasks cut into pieces of at most `cut` bytes
  const size_t cut = (total + (size_t)nt - 1) / (size_t)nt;
  std::vector<CopyTask> tasks;
  for (const CopyTask& t : in)
    for (size_t o = 0; o < t.n; o += cut) tasks.push_back({t.dst + o, t.src + o, t.n - o < cut ? t.n - o : cut});
  // thread k copies the tasks whose bytes fall in its share, in order
  std::vector<size_t> start(tasks.size() + 1, 0);
  for (size_t i = 0; i < tasks.size(); ++i) start[i + 1] = start[i] + tasks[i].n;
  std::vector<std::thread> pool;
  pool.reserve((size_t)nt);
  size_t i0 = 0;
  try {
    for (int k = 0; k < nt && i0 < tasks.size(); ++k) {
      const size_t hi = (total * (size_t)(k + 1)) / (size_t)nt;
      size_t i1 = i0;
      while (i1 < tasks.size() && (start[i1] < hi || k == nt - 1)) ++i1;
      pool.emplace_back([&tasks, i0, i1] {
        for (size_t i = i0; i < i1; ++i) std::memcpy(tasks[i].dst, tasks[i].src, tasks[i].n);
      });
      i0 = i1;
    }
  } catch (const std::system_error&) {
    // a thread could not start: the tasks not yet handed out are copied below
  }
  for (size_t i = i0; i < tasks.size(); ++i) std::memcpy(tasks[i].dst, tasks[i].src, tasks[i].n);
  for (auto& t : pool) t.join();
}

// A fixed set of threads for large copies (the serving loop's submits), so
// a copy is split without starting threads for every call.  run() cuts the
// copies into pieces of 1 MiB that the pool's threads and the caller take
// in turn, and returns when all are copied.  Each call is a job object of
// its own, held by every thread working on it, so a thread still looking at
// a finished job only finds its piece counter exhausted.
class CopyPool {
 public:
  explicit CopyPool(int threads) {
    for (int k = 0; k < threads; ++k) {
      try {
        th_.emplace_back([this] { loop(); });
      } catch (const std::system_error&) {
        break;  // fewer helpers; the callers copy what they do not
      }
    }
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void run(const std::vector<CopyTask>& in) {
    size_t total = 0;
    for (const CopyTask& t : in) total += t.n;
    if (th_.empty() || total < (size_t(2) << 20)) {
      for (const CopyTask& t : in) std::memcpy(t.dst, t.src, t.n);
      return;
    }
    constexpr size_t kCut = size_t(1) << 20;
    auto job = std::make_shared<Job>();
    for (const CopyTask& t : in)
      for (size_t o = 0; o < t.n; o += kCut)
        job->tasks.push_back({t.dst + o, t.src + o, t.n - o < kCut ? t.n - o : kCut});
    job->left = job->tasks.size();
    {
      std::lock_guard<std::mutex> lk(mu_);
      cur_ = job;
      ++gen_;
    }
    cv_.notify_all();
    work(*job);
    {
      std::unique_lock<std::mutex> lk(job->m);
      job->cv.wait(lk, [&] { return job->left == 0; });
    }
    std::lock_guard<std::mutex> lk(mu_);
    if (cur_ == job) cur_.reset();
  }

 private:
  struct Job {
    std::vector<CopyTask> tasks;
    std::atomic<size_t> next{0};
    size_t left = 0;  // pieces not yet copied (under m)
    std::mutex m;
    std::condition_variable cv;
  };
  // Copies pieces of the job until none is left; the last one wakes the caller.
  static void work(Job& j) {
    size_t mine = 0;
    for (size_t i; (i = j.next.fetch_add(1)) < j.tasks.size();) {
      std::memcpy(j.tasks[i].dst, j.tasks[i].src, j.tasks[i].n);
      ++mine;
    }
    if (!mine) return;
    std::lock_guard<std::mutex> lk(j.m);
    j.left -= mine;
    if (j.left == 0) j.cv.notify_all();
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      std::shared_ptr<Job> j;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || (gen_ != seen && cur_); });
        if (stop_) return;
        seen = gen_;
        j = cur_;
      }
      work(*j);
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::shared_ptr<Job> cur_;  // the newest job (under mu_)
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Writes graphs 0..G-1 (nv[g] vertices, nnz[g] arcs, offsets off[g][0..nv],
// ids ind[g][0..nnz)) into the [0, upload_end) part of a host block laid
// out by batch_layout(G, sum nv, sum nnz); returns that layout.  The table
// is written first, then the copies of upload_plan (threaded when large).
inline BatchLayout pack_block(int32_t G, const int32_t* nv, const int64_t* nnz, const int32_t* const* off,
                              const int32_t* const* ind, char* blk, int threads = 8) {
  int64_t v = 0, a = 0;
  for (int32_t g = 0; g < G; ++g) {
    v += nv[g];
    a += nnz[g];
  }
  const BatchLayout L = batch_layout(G, v, a);
  write_tab(G, nv, nnz, reinterpret_cast<int64_t*>(blk + L.tab));
  std::vector<CopyTask> tasks;
  for (const UploadCopy& c : upload_plan(G, nv, nnz, L))
    tasks.push_back({blk + c.dst, reinterpret_cast<const char*>(c.part ? ind[c.g] : off[c.g]), c.bytes});
  copy_tasks(tasks, threads);
  return L;
}

// Batch block size as a sum over requests: each request contributes its
// share (a tab row, nv+1 offsets, nnz ids, nv counts, one total, and its
// scratch_share) and the block adds kBatchFixed, which covers the tab's
// closing row and the rounding of the six regions (each < kArenaAlign), and
// kScratchFixed.  So batch_layout(G, sum nv, sum nnz).bytes <=
// batch_bytes(sum of shares).
constexpr size_t kBatchFixed = 7 * kArenaAlign;
constexpr size_t kBlockFixed = kBatchFixed + kScratchFixed;
inline size_t request_share(int64_t nv, int64_t nnz) {
  return 28 + 8 * (size_t)nv + 4 * (size_t)nnz + scratch_share(nv, nnz);
}
inline size_t batch_bytes(size_t share_sum, size_t fixed = kBlockFixed) { return arena_round(fixed + share_sum); }
// Arena bytes of a request batched alone.
inline size_t request_bytes(int64_t nv, int64_t nnz) { return batch_bytes(request_share(nv, nnz)); }

// FIFO admission of batches: the next batch is the longest prefix of the
// queue (at most max_n requests) whose block fits the largest free arena
// block.  Later requests never overtake the head, so a large request cannot
// starve; when the head alone does not fit nothing is admitted.
struct Pending {
  int64_t id;
  size_t share;
};

class Admission {
 public:
  // fixed: the bytes a block adds to its requests' shares (kBlockFixed for
  // the serving loop's blocks)
  explicit Admission(size_t budget, size_t fixed = kBlockFixed) : arena_(budget), fixed_(fixed) {}
  Arena& arena() { return arena_; }
  // False when the request can never fit (larger than the whole budget).
  bool fits_ever(size_t share) const { return batch_bytes(share, fixed_) <= arena_.capacity(); }
  // Takes request id out of the queue; false when it is not waiting.
  bool remove(int64_t id) {
    for (auto it = q_.begin(); it != q_.end(); ++it)
      if (it->id == id) {
        q_.erase(it);
        return true;
      }
    return false;
  }
  void push(int64_t id, size_t share) { q_.push_back(Pending{id, share}); }
  size_t waiting() const { return q_.size(); }
  // Admit one batch: fills ids and returns the block offset and size, or
  // SIZE_MAX (ids empty) when the head does not fit or nothing waits.
  size_t admit_batch(int max_n, std::vector<int64_t>& ids, size_t& bytes) {
    ids.clear();
    bytes = 0;
    const size_t room = arena_.largest_free();
    size_t sum = 0;
    int n = 0;
    for (const Pending& p : q_) {
      if (n >= max_n || batch_bytes(sum + p.share, fixed_) > room) break;
      sum += p.share;
      ++n;
    }
    if (n == 0) return SIZE_MAX;
    bytes = batch_bytes(sum, fixed_);
    const size_t off = arena_.alloc(bytes);
    if (off == SIZE_MAX) return SIZE_MAX;  // not reached: room >= bytes
    for (int i = 0; i < n; ++i) {
      ids.push_back(q_.front().id);
      q_.pop_front();
    }
    return off;
  }

 private:
  Arena arena_;
  size_t fixed_;
  std::deque<Pending> q_;
};

}  // namespace tcs
