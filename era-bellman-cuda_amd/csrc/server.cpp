// Request loop with continuous batching (extension API of include/bellman-amd.h;
// no counterpart in the reference).
//
// Clients submit MSM / NTT / PN jobs whose buffers already live on the device
// (bases, twiddles and power tables stay resident between jobs).  A loop
// thread owns `slots` streams.  Whenever a slot is free it admits the longest
// FIFO prefix of waiting jobs whose summed footprint fits the HBM budget left
// by the running batches (at most an even share of the waiting jobs per free
// slot) and launches it as one batch on that slot's stream:
// MSMs and NTTs as batched launches, PN jobs one after another.  Batches on
// different slots run concurrently, so jobs that arrive while others are on the
// GPU start as soon as a slot and memory free up.  Completion is observed by
// polling each batch's event.
//
// Memory: the server allocates its whole budget once, as one workspace, and
// gives each batch a region of it sized by a sizing capture of the batch (the
// launch sequence recorded against placeholder addresses, never run).  No
// device allocation happens while batches run: growing a memory pool can wait
// for work queued on other streams, which would serialise the slots.  A batch
// larger than the workspace runs alone with memory from a pool.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.h"
#include "common.h"
#include "ntt.h"
#include "pn.h"

namespace bc {

namespace {

uint64_t now_ns() { return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

constexpr unsigned SERVER_MAX_PN = 8; // PN jobs per batch

struct request {
  uint64_t ticket = 0;
  bc_amd_job_kind kind = BC_AMD_JOB_MSM;
  msm_job msm{};
  ntt_job ntt{};
  bc_amd_pn_job pn{};
  size_t footprint = 0;
  bc_amd_job_stats stats{};
  bool done = false;
  bool claimed = false; // a wait has taken the ticket
};

struct batch {
  unsigned slot = 0;
  std::vector<request *> reqs;
  size_t bytes = 0;         // summed footprints (admission)
  size_t region = 0;        // offset of the workspace region
  size_t region_bytes = 0;  // exact size; 0 when it took no region (no temporaries, or memory from the pool)
};

// First-fit allocator of workspace regions (byte offsets, 256-byte granules).
class region_heap {
public:
  void reset(size_t size) {
    free_.clear();
    if (size)
      free_[0] = size;
  }
  bool take(size_t bytes, size_t *offset) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes)
        continue;
      *offset = it->first;
      const size_t at = it->first + bytes, rest = it->second - bytes;
      free_.erase(it);
      if (rest)
        free_[at] = rest;
      return true;
    }
    return false;
  }
  void give(size_t offset, size_t bytes) {
    bytes = round(bytes);
    auto next = free_.lower_bound(offset);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == offset) {
        offset = prev->first;
        bytes += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && offset + bytes == next->first) {
      bytes += next->second;
      free_.erase(next);
    }
    free_[offset] = bytes;
  }

private:
  static size_t round(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~size_t(255); }
  std::map<size_t, size_t> free_;
};

hipError_t job_footprint(const bc_amd_job &job, request *r) {
  r->kind = job.kind;
  switch (job.kind) {
  case BC_AMD_JOB_MSM:
    BC_TRY(resolve_msm(job.msm, &r->msm));
    r->footprint = msm_footprint(r->msm);
    return hipSuccess;
  case BC_AMD_JOB_NTT: {
    ntt_job n;
    BC_TRY(resolve_ntt(job.ntt, &n));
    r->ntt = n;
    r->footprint = ntt_footprint(n);
    return hipSuccess;
  }
  case BC_AMD_JOB_PN:
    if ((job.pn.columns_count != 3 && job.pn.columns_count != 4) || job.pn.log_rows_count > 28 || !job.pn.indexes || !job.pn.scalars || !job.pn.target)
      return hipErrorInvalidValue;
    r->pn = job.pn;
    r->footprint = pn_footprint(job.pn.columns_count, job.pn.log_rows_count);
    return hipSuccess;
  }
  return hipErrorInvalidValue;
}

class server {
public:
  server(int device, size_t budget, unsigned slots) : device_(device), budget_(budget), slots_(slots) {}

  hipError_t start() {
    BC_TRY(hipMalloc(&workspace_, budget_));
    heap_.reset(budget_);
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device_;
    BC_TRY(hipMemPoolCreate(&pool_, &props));
    streams_.resize(slots_, nullptr);
    events_.resize(slots_, nullptr);
    busy_.assign(slots_, false);
    for (unsigned i = 0; i < slots_; i++) {
      BC_TRY(hipStreamCreateWithFlags(&streams_[i], hipStreamNonBlocking));
      BC_TRY(hipEventCreateWithFlags(&events_[i], hipEventDisableTiming));
    }
    thread_ = std::thread([this] { loop(); });
    return hipSuccess;
  }

  ~server() {
    if (thread_.joinable()) {
      {
        std::lock_guard<std::mutex> lk(m_);
        stop_ = true;
      }
      cv_work_.notify_all();
      thread_.join();
    }
    for (hipStream_t s : streams_)
      if (s)
        (void)hipStreamDestroy(s);
    for (hipEvent_t e : events_)
      if (e)
        (void)hipEventDestroy(e);
    if (pool_)
      (void)hipMemPoolDestroy(pool_);
    if (workspace_)
      (void)hipFree(workspace_);
    for (auto &kv : all_)
      delete kv.second;
  }

  hipError_t submit(const bc_amd_job &job, uint64_t *ticket) {
    request *r = new request;
    const hipError_t e = job_footprint(job, r);
    if (e != hipSuccess) {
      delete r;
      return e;
    }
    r->stats.submit_ns = now_ns();
    {
      std::lock_guard<std::mutex> lk(m_);
      if (stop_) {
        delete r;
        return hipErrorInvalidValue;
      }
      r->ticket = next_ticket_++;
      all_[r->ticket] = r;
      pending_.push_back(r);
      *ticket = r->ticket;
    }
    cv_work_.notify_all();
    return hipSuccess;
  }

  // One waiter per ticket: the first wait claims it, and a second wait on the
  // same ticket (concurrent or later) fails instead of sharing the request
  // the first one frees.
  hipError_t wait(uint64_t ticket, bc_amd_job_stats *stats) {
    std::unique_lock<std::mutex> lk(m_);
    auto it = all_.find(ticket);
    if (it == all_.end() || it->second->claimed)
      return hipErrorInvalidValue;
    request *r = it->second;
    r->claimed = true;
    cv_done_.wait(lk, [&] { return r->done; });
    if (stats)
      *stats = r->stats;
    all_.erase(ticket);
    delete r;
    return hipSuccess;
  }

  struct work {
    std::vector<msm_job> msm;
    std::vector<ntt_job> ntt;
    std::vector<const bc_amd_pn_job *> pn;
  };

  static work collect(const batch &b) {
    work w;
    for (const request *r : b.reqs) {
      if (r->kind == BC_AMD_JOB_MSM)
        w.msm.push_back(r->msm);
      else if (r->kind == BC_AMD_JOB_NTT)
        w.ntt.push_back(r->ntt);
      else
        w.pn.push_back(&r->pn);
    }
    return w;
  }

  static hipError_t record(const work &w, arena &ar) {
    BC_TRY(msm_run(w.msm, ar));
    BC_TRY(ntt_run(w.ntt, ar));
    for (const bc_amd_pn_job *p : w.pn)
      BC_TRY(pn_generate_device(p->indexes, static_cast<const fe *>(p->scalars), static_cast<fe *>(p->target), p->columns_count, p->log_rows_count, ar));
    return hipSuccess;
  }

  // Exact workspace bytes of a batch, by a sizing capture.  Twiddle tables are
  // grown first, outside the capture.
  static hipError_t size_of(const work &w, hipStream_t s, size_t *bytes) {
    unsigned max_log = 0;
    for (const ntt_job &j : w.ntt)
      max_log = std::max(max_log, j.log_n);
    if (max_log)
      BC_TRY(ntt_prepare(max_log, s, nullptr));
    return measure_scratch(s, [&](arena &ar) { return record(w, ar); }, bytes);
  }

private:
  // Longest FIFO prefix that fits the free memory and the per-launch limits,
  // and takes no more than its share of the waiting jobs when several slots
  // are free (so that the free slots' batches run side by side).
  bool take_batch(batch &b, unsigned free_slots) {
    size_t bytes = 0;
    unsigned n_msm = 0, n_ntt = 0, n_pn = 0;
    uint64_t entries = 0;
    const size_t share = (pending_.size() + free_slots - 1) / free_slots;
    while (!pending_.empty() && b.reqs.size() < share) {
      request *r = pending_.front();
      const bool alone = inflight_.empty() && b.reqs.empty();
      if (!alone && used_ + bytes + r->footprint > budget_)
        break;
      if (r->kind == BC_AMD_JOB_MSM) {
        const uint64_t e = (uint64_t)((254 - 1) / r->msm.c + 1) << r->msm.log_n;
        if (n_msm == MSM_MAX_JOBS || (n_msm && entries + e >= (1ull << 31)))
          break;
        entries += e;
        n_msm++;
      } else if (r->kind == BC_AMD_JOB_NTT) {
        if (n_ntt == NTT_MAX_JOBS)
          break;
        n_ntt++;
      } else {
        if (n_pn == SERVER_MAX_PN)
          break;
        n_pn++;
      }
      pending_.pop_front();
      bytes += r->footprint;
      b.reqs.push_back(r);
    }
    b.bytes = bytes;
    return !b.reqs.empty();
  }

  // Sizes the batch, places it in the workspace and launches it on its slot's
  // stream.  When the region does not fit while other batches hold parts of
  // the workspace, nothing is launched and *deferred is set; a batch that does
  // not fit the empty workspace runs alone from the pool.
  hipError_t place_and_launch(batch &b, bool *deferred) {
    *deferred = false;
    const work w = collect(b);
    hipStream_t s = streams_[b.slot];
    size_t need = 0;
    BC_TRY(size_of(w, s, &need));
    // a batch with no temporaries (plain NTTs) takes no region
    b.region = 0;
    const bool placed = need == 0 || heap_.take(need, &b.region);
    if (!placed && !inflight_.empty()) {
      *deferred = true;
      return hipSuccess;
    }
    const uint64_t t = now_ns();
    for (request *r : b.reqs) {
      r->stats.start_ns = t;
      r->stats.batch_jobs = (unsigned)b.reqs.size();
      r->stats.slot = b.slot;
    }
    b.region_bytes = placed ? need : 0; // returned to the heap when the batch completes
    arena ar = placed ? arena::bump_mode(static_cast<char *>(workspace_) + b.region, need, s) : arena::pool_mode(pool_, s);
    BC_TRY(record(w, ar));
    BC_TRY(hipEventRecord(events_[b.slot], s));
    const uint64_t enqueued = now_ns();
    for (request *r : b.reqs)
      r->stats.enqueued_ns = enqueued;
    return hipSuccess;
  }

  void finish(batch &b, hipError_t status) {
    const uint64_t t = now_ns();
    for (request *r : b.reqs) {
      r->stats.done_ns = t;
      r->stats.status = as_bc(status);
      r->done = true;
    }
    if (b.region_bytes)
      heap_.give(b.region, b.region_bytes);
    used_ -= b.bytes;
    busy_[b.slot] = false;
  }

  void loop() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(m_);
    bool deferred = false; // a batch waits for workspace that a running batch holds
    for (;;) {
      bool progress = false;
      for (size_t i = 0; i < inflight_.size();) {
        const hipError_t q = hipEventQuery(events_[inflight_[i].slot]);
        if (q == hipErrorNotReady) {
          i++;
          continue;
        }
        if (q != hipSuccess)
          (void)hipStreamSynchronize(streams_[inflight_[i].slot]);
        finish(inflight_[i], q);
        inflight_.erase(inflight_.begin() + i);
        progress = true;
      }
      if (progress) {
        deferred = false;
        cv_done_.notify_all();
      }
      int slot = -1;
      unsigned free_slots = 0;
      for (unsigned i = 0; i < slots_; i++)
        if (!busy_[i]) {
          if (slot < 0)
            slot = (int)i;
          free_slots++;
        }
      if (slot >= 0 && !pending_.empty() && !deferred) {
        batch b;
        b.slot = (unsigned)slot;
        if (take_batch(b, free_slots)) {
          used_ += b.bytes;
          busy_[b.slot] = true;
          lk.unlock();
          bool later = false;
          hipError_t e = place_and_launch(b, &later);
          if (e != hipSuccess)
            (void)hipStreamSynchronize(streams_[b.slot]); // work already enqueued must end before its memory is reused
          lk.lock();
          if (later) {
            for (auto it = b.reqs.rbegin(); it != b.reqs.rend(); ++it)
              pending_.push_front(*it);
            used_ -= b.bytes;
            busy_[b.slot] = false;
            deferred = true;
          } else if (e != hipSuccess) {
            finish(b, e);
            cv_done_.notify_all();
          } else {
            inflight_.push_back(std::move(b));
          }
          continue;
        }
      }
      if (stop_ && pending_.empty() && inflight_.empty())
        break;
      if (inflight_.empty())
        cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
      else
        cv_work_.wait_for(lk, std::chrono::microseconds(50));
    }
  }

  int device_;
  size_t budget_;
  unsigned slots_;
  size_t used_ = 0;
  void *workspace_ = nullptr;
  region_heap heap_;
  hipMemPool_t pool_ = nullptr;
  std::vector<hipStream_t> streams_;
  std::vector<hipEvent_t> events_;
  std::vector<bool> busy_;
  std::mutex m_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<request *> pending_;
  std::unordered_map<uint64_t, request *> all_;
  std::vector<batch> inflight_;
  uint64_t next_ticket_ = 1;
  bool stop_ = false;
  std::thread thread_;
};

} // namespace

} // namespace bc

using namespace bc;

extern "C" {

bc_error bc_amd_job_footprint(const bc_amd_job *job, size_t *bytes) {
  if (!job || !bytes)
    return bc_error_invalid_value;
  request r;
  BC_TRY_BC(job_footprint(*job, &r));
  *bytes = r.footprint;
  return bc_success;
}

bc_error bc_amd_job_workspace_bytes(const bc_amd_job *job, size_t *bytes) {
  if (!job || !bytes)
    return bc_error_invalid_value;
  request r;
  BC_TRY_BC(job_footprint(*job, &r));
  batch b;
  b.reqs.push_back(&r);
  const auto w = server::collect(b);
  hipStream_t s = nullptr;
  BC_TRY_BC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const hipError_t e = server::size_of(w, s, bytes);
  (void)hipStreamDestroy(s);
  return as_bc(e);
}

bc_error bc_amd_server_create(bc_amd_server *server_out, size_t hbm_budget_bytes, unsigned max_in_flight) {
  if (!server_out || max_in_flight == 0 || max_in_flight > 16 || hbm_budget_bytes == 0)
    return bc_error_invalid_value;
  int device = 0;
  BC_TRY_BC(hipGetDevice(&device));
  server *s = new server(device, hbm_budget_bytes, max_in_flight);
  const hipError_t e = s->start();
  if (e != hipSuccess) {
    delete s;
    server_out->handle = nullptr;
    return as_bc(e);
  }
  server_out->handle = s;
  return bc_success;
}

bc_error bc_amd_server_submit(bc_amd_server srv, const bc_amd_job *job, unsigned long long *ticket) {
  if (!srv.handle || !job || !ticket)
    return bc_error_invalid_value;
  uint64_t t = 0;
  BC_TRY_BC(static_cast<server *>(srv.handle)->submit(*job, &t));
  *ticket = t;
  return bc_success;
}

bc_error bc_amd_server_wait(bc_amd_server srv, unsigned long long ticket, bc_amd_job_stats *stats) {
  if (!srv.handle)
    return bc_error_invalid_value;
  return as_bc(static_cast<server *>(srv.handle)->wait(ticket, stats));
}

bc_error bc_amd_server_destroy(bc_amd_server srv) {
  if (!srv.handle)
    return bc_error_invalid_value;
  delete static_cast<server *>(srv.handle);
  return bc_success;
}

} // extern "C"
