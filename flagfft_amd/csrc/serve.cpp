// serve.cpp — FFT serving loop: HBM-budget admission, resident plans and
// tables, continuous batching into persistent mixed-length kernel launches.
//
// No counterpart in the reference (its runtime executes one plan per call).
// A dispatcher thread owns every device call after creation:
//   reap    completed batches (event queries) -> requests done, workspaces freed
//   admit   pending requests in arrival order while output + workspace fit the arena
//   launch  admitted requests as one batch per free stream slot: single-pass
//           requests become tiles of one k_queue launch, the rest run their
//           plan's kernels; one event per group marks completion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "exec.hpp"
#include "flagfft.h"
#include "flagfft_ext.h"
#include "planner.hpp"

namespace flagfft {
namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
constexpr int kMaxJobs = 2048;                     // queue jobs per persistent launch
constexpr int64_t kQueueMaxBytes = int64_t{32} << 20;   // larger requests with run-time-radix passes launch on their own
// the same for plans of register-resident passes: larger requests run faster
// on their own plans' kernels (wider tiles than the queue's 256-thread
// instances) and, in a batch, would hold its small requests back (measured on
// the mixed-length trace: 64 MiB gave 2.9 % more transforms/s and a 10 % lower
// p99 than 1 GiB)
constexpr int64_t kQueueMaxBytesFast = int64_t{64} << 20;
constexpr int kMaxQueueSteps = 6;                       // queue launches per batch (passes of the longest plan)
// queue launches per batch: per step, the run-time-radix kernel and one per
// register-resident variant present (at most 10)
constexpr int kMaxQueueLaunches = kMaxQueueSteps * 11;
constexpr size_t kStagingBytes = 256 * kMaxQueueLaunches + sizeof(QueueJob) * kMaxJobs * kMaxQueueSteps;
constexpr int64_t kNominalBatch = 512;             // batch the cached plans are shaped for
constexpr double kLargeRequestBytes = 64.0 * (1 << 20);  // direct requests above this stay in arrival order
// ... on lane 0 unless this many requests wait (pending admission or
// dispatch): then they take the less loaded lane too.  Two large requests at
// once stretch each other, which costs latency when the server keeps up, but
// with a backlog the overlap of their kernels' tails serves more per second.
// Measured on the mixed-length trace: with 16 the batched server serves
// 4.8-5.7 % more transforms/s than per-request plain launches (0.7-2.8 % with
// large requests on lane 0 only) at the same p99 (32.8 ms against 36.5);
// always spreading them gives about the same capacity but a p99 of 34-36 ms.
constexpr int kLargeBacklogRequests = 16;
// BATCHED admission: a request that does not fit the arena yet lets later
// ones of at most kBypassMaxBytes pass, while those hold at most 1 /
// kBypassShareDiv of the budget (so the waiting one fits once the requests
// admitted before it finish)
constexpr size_t kBypassMaxBytes = size_t{256} << 20;
constexpr size_t kBypassShareDiv = 8;
constexpr int64_t kQueueNominal = int64_t{1} << 20;  // queue templates: full tiles (C * tpu ~ 256 threads)

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device allocation.
class Arena {
 public:
  explicit Arena(size_t bytes) : cap_(align_up(bytes)) {
    hip_check(hipMalloc(&base_, cap_), "server arena");
    free_[0] = cap_;
  }
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  bool alloc(size_t bytes, size_t &off) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      off = it->first;
      const size_t rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
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
  char *base() const { return static_cast<char *>(base_); }
  size_t capacity() const { return cap_; }
  size_t peak() const { return peak_; }
  void reset_peak() { peak_ = used_; }

 private:
  void *base_ = nullptr;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

// How one (type, n, direction) is served.
struct ServePlan {
  int type = 0;
  int64_t n = 0;
  bool dbl = false;
  int64_t in_bytes = 0, out_bytes = 0;      // per transform
  std::unique_ptr<Executable> plain;        // specialised kernels, caller workspaces
  std::vector<size_t> ws_per_transform;     // plain: workspace bytes per transform
  // The plan as persistent-queue jobs: every launch of it has a queue kind
  // (run-time-radix bodies, register-resident passes of 256 threads, real
  // split passes); step s of a batch's requests runs in the batch's s-th
  // queue launch.  Null when some launch has no queue form.
  std::unique_ptr<Executable> queue;
  std::vector<int> queue_kinds, queue_variants;  // per step
  std::vector<size_t> queue_ws_per_transform;
  bool queue_fast = false;  // no run-time-radix pass over more than one tile column (see kQueueMaxBytesFast)
};

struct Request {
  int64_t id = 0;
  ServePlan *plan = nullptr;
  int64_t batch = 0;
  const void *in = nullptr;
  size_t out_off = 0, out_size = 0;
  std::vector<size_t> ws_off, ws_size;
  Clock::time_point arrival, done_at;
  enum State { Pending, Admitted, InFlight, Done } state = Pending;
  bool queued = false;
  // arena bytes (output, workspaces) held by a request admitted ahead of an
  // earlier one that did not fit yet
  size_t bypass_out = 0, bypass_ws = 0;
};

struct Group {  // requests completed by one event
  hipEvent_t ev = nullptr;
  std::vector<Request *> reqs;
  bool queue = false;  // a persistent-kernel launch (the slot's job staging is in use until it completes)
  double bytes = 0.0;  // algorithmic bytes, counted in the slot's load
};

struct Slot {
  hipStream_t stream = nullptr;
  bool busy = false;  // job staging in use by a queue launch
  double load = 0.0;  // algorithmic bytes launched and not yet completed
  std::vector<Group> groups;
  // pinned staging of the batch's queue launches: each launch's counter block
  // (256 bytes), then its jobs, one launch after the other
  QueueJob *host_jobs = nullptr;
  void *dev_jobs = nullptr;
};

bool queue_kind_of(const Step &s, int &kind, int &variant) {
  variant = 0;
  switch (s.kind) {
    case Step::Col:
      kind = kQueueCol;
      variant = col_queue_variant(s.leaf_id);
      return variant >= 0 && s.shape.block.x == static_cast<unsigned>(kQueueThreads) &&
             s.shape.lds_bytes <= 80 * 1024;
    case Step::RealPost:
    case Step::RealPre:
      kind = s.kind == Step::RealPost ? kQueueRealPost : kQueueRealPre;
      return s.shape.block.x == static_cast<unsigned>(kQueueThreads) && s.shape.lds_bytes == 0;
    case Step::Pass: kind = kQueuePass; break;
    case Step::Rader: kind = kQueueRader; break;
    case Step::Bluestein: kind = kQueueBluestein; break;
    case Step::R2CEven: kind = kQueueR2CEven; break;
    case Step::C2REven: kind = kQueueC2REven; break;
    default: return false;
  }
  return s.ept == (s.dbl ? 8 : 16) && s.shape.block.x <= static_cast<unsigned>(kQueueThreads) &&
         s.shape.lds_bytes <= 64 * 1024;
}

}  // namespace

class Server {
 public:
  Server(size_t budget, int max_inflight, int mode)
      : arena_(budget),
        mode_(mode),
        nslots_(mode == FLAGFFTX_SERVE_PLAIN ? 1 : std::max(1, max_inflight) + 1),
        ndirect_(mode == FLAGFFTX_SERVE_PLAIN ? 1 : 2) {
    // FLAGFFT_SERVE_QUEUE_MAX_MIB: the per-request size limit of queued plans
    // with register-resident passes (MiB of input + output; default 64)
    if (const char *e = std::getenv("FLAGFFT_SERVE_QUEUE_MAX_MIB")) {
      const long long mib = std::atoll(e);
      if (mib >= 0) queue_max_fast_ = static_cast<int64_t>(std::min<long long>(mib, 1LL << 30)) << 20;
    }
    // FLAGFFT_SERVE_LARGE_BACKLOG: waiting requests past which large ones
    // may take the other direct lane (default 16)
    if (const char *e = std::getenv("FLAGFFT_SERVE_LARGE_BACKLOG")) {
      const int v = std::atoi(e);
      if (v >= 0) large_backlog_requests_ = v;
    }
    hip_check(hipGetDevice(&device_), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device_), "hipGetDeviceProperties");
    ncu_ = prop.multiProcessorCount;
    slots_.resize(static_cast<size_t>(nslots_));
    const size_t job_bytes = kStagingBytes;
    for (Slot &s : slots_) {
      hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "server stream");
      if (mode_ == FLAGFFTX_SERVE_BATCHED) {
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&s.host_jobs), job_bytes, hipHostMallocDefault),
                  "pinned job staging");
        hip_check(hipMalloc(&s.dev_jobs, job_bytes), "job buffer");
      }
    }
    worker_ = std::thread([this] { loop(); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
    for (Slot &s : slots_) {
      for (Group &g : s.groups)
        if (g.ev) (void)hipEventDestroy(g.ev);
      if (s.stream) (void)hipStreamDestroy(s.stream);
      if (s.host_jobs) (void)hipHostFree(s.host_jobs);
      if (s.dev_jobs) (void)hipFree(s.dev_jobs);
    }
    for (hipEvent_t e : free_events_) (void)hipEventDestroy(e);
  }

  flagfftResult prepare(int type, int64_t n, int dir, ServePlan **out) {
    const auto key = std::make_tuple(type, n, dir);
    std::lock_guard<std::mutex> lk(plans_mu_);
    auto it = plans_.find(key);
    if (it != plans_.end()) {
      if (out) *out = it->second.get();
      return FLAGFFT_SUCCESS;
    }
    const bool real_fwd = type == FLAGFFT_R2C || type == FLAGFFT_D2Z;
    const bool real_inv = type == FLAGFFT_C2R || type == FLAGFFT_Z2D;
    const bool dbl = type == FLAGFFT_Z2Z || type == FLAGFFT_D2Z || type == FLAGFFT_Z2D;
    if (!(real_fwd || real_inv || type == FLAGFFT_C2C || type == FLAGFFT_Z2Z)) return FLAGFFT_INVALID_TYPE;
    const int64_t h = n / 2 + 1;
    const int64_t in_len = real_inv ? h : n, out_len = real_fwd ? h : n;
    PlanDesc d;
    d.rank = 1;
    d.n0 = n;
    d.type = type;
    d.batch = kNominalBatch;
    d.prec = dbl ? Prec::F64 : Prec::F32;
    d.real_fwd = real_fwd;
    d.real_inv = real_inv;
    const bool inverse = real_inv || dir == FLAGFFT_INVERSE;
    auto sp = std::make_unique<ServePlan>();
    sp->type = type;
    sp->n = n;
    sp->dbl = dbl;
    const int eb = dbl ? 16 : 8;
    sp->in_bytes = real_inv ? h * eb : (real_fwd ? n * eb / 2 : n * eb);
    sp->out_bytes = real_fwd ? h * eb : (real_inv ? n * eb / 2 : n * eb);
    try {
      PlanOptions o;
      PlanTree t = plan_tree(d, o);
      sp->plain = build_executable(t, inverse, o, in_len, out_len, false);
      for (size_t b : sp->plain->ws_bytes) sp->ws_per_transform.push_back(b / static_cast<size_t>(kNominalBatch));
      table_bytes_ += tables_of(*sp->plain);
      if (mode_ == FLAGFFTX_SERVE_BATCHED) {
        PlanOptions g;
        g.generic_kernels = true;
        g.queue_form = true;
        g.max_threads = kQueueThreads;
        g.ept = dbl ? 8 : 16;
        // single-pass queue jobs hold a whole sequence in one 256-thread tile;
        // longer lengths run as chains of queue passes
        g.leaf_max_f32 = int64_t{kQueueThreads} * 16;
        g.leaf_max_f64 = int64_t{kQueueThreads} * 8;
        try {
          PlanDesc dg = d;
          dg.batch = kQueueNominal;
          PlanTree tg = plan_tree(dg, g);
          auto ex = build_executable(tg, inverse, g, in_len, out_len, false);
          bool ok = !ex->steps.empty() && ex->steps.size() <= static_cast<size_t>(kMaxQueueSteps);
          bool fast = true;
          std::vector<int> kinds, variants;
          const bool explain = std::getenv("FLAGFFT_SERVE_EXPLAIN") != nullptr;
          for (const Step &s : ex->steps) {
            int kind = -1, variant = 0;
            const bool q = queue_kind_of(s, kind, variant);
            if (!q && explain)
              std::fprintf(stderr, "[flagfft] type %d n=%lld: no queue form for launch '%s' (%u threads, %zu LDS bytes)\n",
                           type, static_cast<long long>(n), s.label.c_str(), s.shape.block.x, s.shape.lds_bytes);
            ok = ok && q;
            kinds.push_back(kind);
            variants.push_back(variant);
            // a run-time-radix pass of a multi-pass chain runs at a fraction of the
            // register-resident rate: such plans stay on the small-request threshold
            if (s.kind == Step::Pass && ex->steps.size() > 1) fast = false;
          }
          if (ok) {
            sp->queue_kinds = kinds;
            sp->queue_variants = variants;
            for (size_t b : ex->ws_bytes) sp->queue_ws_per_transform.push_back(b / static_cast<size_t>(kQueueNominal));
            sp->queue_fast = fast;
            table_bytes_ += tables_of(*ex);
            sp->queue = std::move(ex);
          }
        } catch (const std::exception &ex) {
          // no queue form (e.g. more than 256 threads per transform): plan launches only
          if (std::getenv("FLAGFFT_SERVE_EXPLAIN") != nullptr)
            std::fprintf(stderr, "[flagfft] type %d n=%lld: no queue plan (%s)\n", type, static_cast<long long>(n),
                         ex.what());
        }
      }
    } catch (const HipError &e) {
      std::fprintf(stderr, "[flagfft] server plan failed: %s\n", e.what());
      return e.code == hipErrorOutOfMemory ? FLAGFFT_ALLOC_FAILED : FLAGFFT_SETUP_FAILED;
    } catch (const std::bad_alloc &) {
      return FLAGFFT_ALLOC_FAILED;
    } catch (const std::exception &e) {
      std::fprintf(stderr, "[flagfft] server plan failed: %s\n", e.what());
      return FLAGFFT_SETUP_FAILED;
    }
    if (out) *out = sp.get();
    plans_[key] = std::move(sp);
    return FLAGFFT_SUCCESS;
  }

  flagfftResult submit(int type, int64_t n, int64_t batch, int dir, const void *in, int64_t &ticket) {
    if (type == FLAGFFT_C2C || type == FLAGFFT_Z2Z) {
      if (dir != FLAGFFT_FORWARD && dir != FLAGFFT_INVERSE) return FLAGFFT_INVALID_VALUE;
    } else {
      dir = 0;
    }
    ServePlan *sp = nullptr;
    flagfftResult r = prepare(type, n, dir, &sp);
    if (r != FLAGFFT_SUCCESS) return r;
    auto req = std::make_unique<Request>();
    req->plan = sp;
    req->batch = batch;
    req->in = in;
    req->out_size = static_cast<size_t>(batch * sp->out_bytes);
    size_t need = align_up(req->out_size);
    req->queued = sp->queue &&
                  batch * (sp->in_bytes + sp->out_bytes) <= (sp->queue_fast ? queue_max_fast_ : kQueueMaxBytes);
    for (size_t w : (req->queued ? sp->queue_ws_per_transform : sp->ws_per_transform)) {
      req->ws_size.push_back(w * static_cast<size_t>(batch));
      need += align_up(req->ws_size.back());
    }
    if (need > arena_.capacity()) return FLAGFFT_ALLOC_FAILED;
    req->arrival = Clock::now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      req->id = ++next_id_;
      ticket = req->id;
      pending_.push_back(req.get());
      reqs_[req->id] = std::move(req);
    }
    cv_work_.notify_all();
    return FLAGFFT_SUCCESS;
  }

  flagfftResult poll(int64_t ticket, bool wait, int &done, void *&out, double &latency_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return FLAGFFT_INVALID_VALUE;
    Request *r = it->second.get();
    if (wait) cv_done_.wait(lk, [&] { return r->state == Request::Done || failed_; });
    if (failed_ && r->state != Request::Done) return FLAGFFT_EXEC_FAILED;
    done = r->state == Request::Done ? 1 : 0;
    out = done ? arena_.base() + r->out_off : nullptr;
    latency_ms = done ? std::chrono::duration<double, std::milli>(r->done_at - r->arrival).count() : 0.0;
    return FLAGFFT_SUCCESS;
  }

  flagfftResult release(int64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return FLAGFFT_INVALID_VALUE;
    if (it->second->state != Request::Done) return FLAGFFT_INVALID_VALUE;
    arena_.release(it->second->out_off, it->second->out_size);
    bypass_held_ -= it->second->bypass_out;
    reqs_.erase(it);
    cv_work_.notify_all();
    return FLAGFFT_SUCCESS;
  }

  void stats(flagfftxServerStats &s) {
    std::lock_guard<std::mutex> lk(mu_);
    std::memset(&s, 0, sizeof(s));
    s.completed = static_cast<long long>(lat_ms_.size());
    s.transforms = st_transforms_;
    s.bytes = st_bytes_;
    if (!lat_ms_.empty()) {
      std::vector<double> v = lat_ms_;
      std::sort(v.begin(), v.end());
      auto pct = [&](double p) {
        const size_t i = static_cast<size_t>(p * static_cast<double>(v.size() - 1) + 0.5);
        return v[std::min(i, v.size() - 1)];
      };
      s.p50_ms = pct(0.50);
      s.p99_ms = pct(0.99);
      s.max_ms = v.back();
      double sum = 0;
      for (double x : v) sum += x;
      s.mean_ms = sum / static_cast<double>(v.size());
      s.span_s = std::chrono::duration<double>(st_last_done_ - st_first_arrival_).count();
    }
    s.batches = st_batches_;
    s.queue_launches = st_queue_launches_;
    s.queued_requests = st_queued_;
    s.arena_peak_bytes = static_cast<double>(arena_.peak());
    s.table_bytes = static_cast<double>(table_bytes_);
  }

  void reset_stats() {
    std::lock_guard<std::mutex> lk(mu_);
    lat_ms_.clear();
    st_transforms_ = 0;
    st_bytes_ = 0;
    st_batches_ = st_queue_launches_ = st_queued_ = 0;
    st_have_first_ = false;
    arena_.reset_peak();
  }

  // Waits for everything submitted so far.
  flagfftResult drain() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [&] { return failed_ || (pending_.empty() && admitted_.empty() && inflight_ == 0); });
    return failed_ ? FLAGFFT_EXEC_FAILED : FLAGFFT_SUCCESS;
  }

 private:
  static size_t tables_of(const Executable &ex) {
    size_t b = 0;
    for (const auto &kv : ex.tables) b += kv.second->bytes;
    return b;
  }

  // Completed groups -> requests done (mu_ held).
  bool reap_locked() {
    bool any = false;
    const Clock::time_point now = Clock::now();
    for (Slot &s : slots_) {
      size_t k = 0;
      for (; k < s.groups.size(); ++k) {
        const hipError_t q = hipEventQuery(s.groups[k].ev);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) {
          std::fprintf(stderr, "[flagfft] server batch failed: %s\n", hipGetErrorString(q));
          failed_ = true;
        }
        for (Request *r : s.groups[k].reqs) {
          r->state = Request::Done;
          r->done_at = now;
          for (size_t i = 0; i < r->ws_off.size(); ++i) arena_.release(r->ws_off[i], r->ws_size[i]);
          r->ws_off.clear();
          bypass_held_ -= r->bypass_ws;
          r->bypass_ws = 0;
          lat_ms_.push_back(std::chrono::duration<double, std::milli>(now - r->arrival).count());
          st_transforms_ += r->batch;
          st_bytes_ += static_cast<double>(r->batch * (r->plan->in_bytes + r->plan->out_bytes));
          if (!st_have_first_ || r->arrival < st_first_arrival_) st_first_arrival_ = r->arrival;
          st_have_first_ = true;
          st_last_done_ = now;
          --inflight_;
        }
        free_events_.push_back(s.groups[k].ev);
        s.load -= s.groups[k].bytes;
        if (s.groups[k].queue) s.busy = false;
        any = true;
      }
      if (k) s.groups.erase(s.groups.begin(), s.groups.begin() + static_cast<std::ptrdiff_t>(k));
      if (s.groups.empty()) s.load = 0.0;
    }
    return any;
  }

  // Pending -> admitted while memory lasts (mu_ held): in arrival order; in
  // BATCHED mode a request that does not fit yet lets smaller later ones pass
  // (kBypassMaxBytes, kBypassShareDiv), PLAIN mode stops at it.
  bool admit_locked() {
    bool any = false;
    bool blocked = false;
    for (auto it = pending_.begin(); it != pending_.end();) {
      Request *r = *it;
      size_t need_ws = 0;
      for (size_t w : r->ws_size) need_ws += align_up(w);
      const size_t need = align_up(r->out_size) + need_ws;
      if (blocked && (need > kBypassMaxBytes || bypass_held_ + need > arena_.capacity() / kBypassShareDiv)) {
        ++it;
        continue;
      }
      if (!alloc_request(r)) {
        if (mode_ != FLAGFFTX_SERVE_BATCHED) break;
        blocked = true;
        ++it;
        continue;
      }
      if (blocked) {
        r->bypass_out = align_up(r->out_size);
        r->bypass_ws = need_ws;
        bypass_held_ += need;
      }
      r->state = Request::Admitted;
      it = pending_.erase(it);
      admitted_.push_back(r);
      any = true;
    }
    return any;
  }

  // The request's output and workspaces from the arena, all or nothing.
  bool alloc_request(Request *r) {
    size_t out_off = 0;
    if (!arena_.alloc(r->out_size, out_off)) return false;
    std::vector<size_t> offs;
    for (size_t w : r->ws_size) {
      size_t o = 0;
      if (!arena_.alloc(w, o)) {
        for (size_t i = 0; i < offs.size(); ++i) arena_.release(offs[i], r->ws_size[i]);
        arena_.release(out_off, r->out_size);
        return false;
      }
      offs.push_back(o);
    }
    r->out_off = out_off;
    r->ws_off = offs;
    return true;
  }

  void loop() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      bool progress = reap_locked();
      progress = admit_locked() || progress;
      if (progress) cv_done_.notify_all();
      if (failed_) {
        cv_done_.notify_all();
        if (stop_) break;
      }
      if (stop_ && pending_.empty() && admitted_.empty() && inflight_ == 0) break;
      // Requests that run their own plan's kernels go to a direct lane
      // (slots [0, ndirect_)) as soon as they are admitted: large ones in
      // arrival order on lane 0, smaller ones on the less loaded lane, so
      // they are not held behind a large one; single-pass requests wait for a batch slot
      // (slots [ndirect_, nslots_), never behind a direct request) whose
      // pinned job staging is idle and leave together in one
      // persistent-kernel launch.  PLAIN mode has one slot: every request
      // runs its own plan on it, in arrival order.
      std::vector<std::vector<Request *>> direct(slots_.size());
      std::vector<Request *> queued;
      Slot *batch_slot = nullptr;
      bool any_direct = false;
      if (!failed_ && !admitted_.empty()) {
        const size_t first_batch = slots_.size() > static_cast<size_t>(ndirect_) ? static_cast<size_t>(ndirect_) : 0;
        for (size_t si = first_batch; si < slots_.size(); ++si)
          if (!slots_[si].busy && (batch_slot == nullptr || slots_[si].load < batch_slot->load))
            batch_slot = &slots_[si];
        for (auto it = admitted_.begin(); it != admitted_.end();) {
          Request *r = *it;
          const bool q = r->queued && mode_ == FLAGFFTX_SERVE_BATCHED;
          if (!q) {
            // large requests keep arrival order on lane 0 (running two at once
            // only stretches both) unless a backlog waits; the others take
            // the less loaded lane, so they pass the large ones
            size_t best = 0;
            const size_t nd = std::min(slots_.size(), static_cast<size_t>(ndirect_));
            if (request_bytes(r) <= kLargeRequestBytes ||
                pending_.size() + admitted_.size() >= static_cast<size_t>(large_backlog_requests_))
              for (size_t si = 1; si < nd; ++si)
                if (slots_[si].load < slots_[best].load) best = si;
            slots_[best].load += request_bytes(r);
            direct[best].push_back(r);
            any_direct = true;
            it = admitted_.erase(it);
          } else if (batch_slot != nullptr && queued.size() < static_cast<size_t>(kMaxJobs)) {
            queued.push_back(r);
            it = admitted_.erase(it);
          } else {
            ++it;
          }
        }
      }
      if (any_direct || !queued.empty()) {
        for (auto &d : direct)
          for (Request *r : d) r->state = Request::InFlight;
        for (Request *r : queued) {
          r->state = Request::InFlight;
          batch_slot->load += request_bytes(r);
        }
        for (auto &d : direct) inflight_ += static_cast<int64_t>(d.size());
        inflight_ += static_cast<int64_t>(queued.size());
        if (!queued.empty()) batch_slot->busy = true;
        lk.unlock();
        std::vector<std::vector<Group>> groups(slots_.size());
        std::vector<Group> batch_groups;
        hipError_t e = hipSuccess;
        for (size_t si = 0; si < slots_.size() && e == hipSuccess; ++si)
          if (!direct[si].empty()) e = launch(slots_[si], direct[si], groups[si]);
        if (e == hipSuccess && !queued.empty()) e = launch(*batch_slot, queued, batch_groups);
        lk.lock();
        for (size_t si = 0; si < slots_.size(); ++si)
          for (Group &g : groups[si]) slots_[si].groups.push_back(std::move(g));
        if (!queued.empty())
          for (Group &g : batch_groups) batch_slot->groups.push_back(std::move(g));
        if (e != hipSuccess) {
          std::fprintf(stderr, "[flagfft] server launch failed: %s\n", hipGetErrorString(e));
          failed_ = true;
        }
        continue;
      }
      // nothing to launch: wait for a submission, a release, or poll in-flight work
      if (inflight_ > 0)
        cv_work_.wait_for(lk, std::chrono::microseconds(20));
      else
        cv_work_.wait_for(lk, std::chrono::milliseconds(5));
    }
    cv_done_.notify_all();
  }

  // Enqueue one batch on the slot's stream (no lock held).
  hipError_t launch(Slot &slot, const std::vector<Request *> &batch, std::vector<Group> &groups) {
    hipError_t e = hipSuccess;
    std::vector<Request *> queued, direct;
    for (Request *r : batch) (r->queued && mode_ == FLAGFFTX_SERVE_BATCHED ? queued : direct).push_back(r);
    char *base = arena_.base();
    if (!queued.empty()) {
      // every request's plan as queue jobs, largest tiles first; launch s runs
      // step s of every request that has one (stream order keeps a request's
      // passes in sequence), one event after the last launch
      struct Item {
        Request *r;
        std::vector<Step> st;
        double cost;
      };
      std::vector<Item> items;
      items.reserve(queued.size());
      size_t nsteps = 0;
      for (Request *r : queued) {
        const Executable &q = *r->plan->queue;
        Item it{r, {}, 0.0};
        for (const Step &t : q.steps) it.st.push_back(t.rebatched(r->batch, q.built_batch));
        it.cost = static_cast<double>(r->plan->n) * (r->plan->dbl ? 2.0 : 1.0) * it.st[0].shape.block.x;
        nsteps = std::max(nsteps, it.st.size());
        items.push_back(std::move(it));
      }
      std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
      size_t cursor = 0;  // staging bytes used by the batch's earlier launches
      int64_t launches = 0;
      for (size_t s = 0; s < nsteps; ++s) {
        // one launch per kernel class present in this step: the run-time-radix
        // queue kernel (class -1) and each register-resident variant
        std::vector<int> classes;
        for (const Item &it : items)
          if (s < it.st.size()) {
            const int c = it.r->plan->queue_kinds[s] == kQueueCol ? it.r->plan->queue_variants[s] : -1;
            if (std::find(classes.begin(), classes.end(), c) == classes.end()) classes.push_back(c);
          }
        for (int cls : classes) {
          char *hreg = reinterpret_cast<char *>(slot.host_jobs) + cursor;
          char *dreg = static_cast<char *>(slot.dev_jobs) + cursor;
          uint32_t *counter = reinterpret_cast<uint32_t *>(hreg);
          QueueJob *jobs = reinterpret_cast<QueueJob *>(hreg + kAlign);
          int32_t total = 0, njobs = 0;
          size_t lds = 0;
          for (const Item &it : items) {
            if (s >= it.st.size()) continue;
            Request *r = it.r;
            const int c = r->plan->queue_kinds[s] == kQueueCol ? r->plan->queue_variants[s] : -1;
            if (c != cls) continue;
            const Step &st = it.st[s];
            auto buf = [&](int id) -> void * {
              if (id == kBufIn) return const_cast<void *>(r->in);
              if (id == kBufOut) return base + r->out_off;
              return base + r->ws_off[static_cast<size_t>(id - kBufWs0)];
            };
            QueueJob &j = jobs[njobs++];
            std::memset(&j, 0, sizeof(j));
            j.kind = r->plan->queue_kinds[s];
            j.variant = r->plan->queue_variants[s];
            j.dbl = r->plan->dbl ? 1 : 0;
            j.first = total;
            j.nblocks = static_cast<int32_t>(st.shape.grid.x);
            unsigned char abuf[Step::kArgBytes];
            const size_t nb = st.pack_args(buf(st.src), buf(st.dst), abuf, st.aux >= 0 ? buf(st.aux) : nullptr);
            std::memcpy(&j.pa, abuf, nb);  // the union's members share one address
            total += j.nblocks;
            lds = std::max(lds, st.shape.lds_bytes);
          }
          *counter = 0;
          const size_t bytes = kAlign + sizeof(QueueJob) * static_cast<size_t>(njobs);
          cursor += (bytes + kAlign - 1) / kAlign * kAlign;
          e = hipMemcpyAsync(dreg, hreg, bytes, hipMemcpyHostToDevice, slot.stream);
          if (e != hipSuccess) return e;
          const void *kern = cls < 0 ? queue_kernel() : queue_col_kernel(cls);
          if (kern == nullptr) return hipErrorInvalidDeviceFunction;
          QueueArgs qa;
          qa.jobs = reinterpret_cast<const QueueJob *>(dreg + kAlign);
          qa.njobs = njobs;
          qa.total = total;
          qa.counter = reinterpret_cast<uint32_t *>(dreg);
          int per_cu = 0;
          if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kQueueThreads, lds) != hipSuccess ||
              per_cu < 1)
            per_cu = 1;
          const int grid = std::max(1, std::min(total, per_cu * ncu_));
          void *kp[] = {&qa};
          e = allow_lds(kern, lds);
          if (e == hipSuccess)
            e = hipLaunchKernel(kern, dim3(static_cast<unsigned>(grid)), dim3(kQueueThreads), kp, lds, slot.stream);
          if (e != hipSuccess) return e;
          ++launches;
        }
      }
      Group g;
      g.ev = take_event_unlocked();
      g.queue = true;
      for (const Item &it : items) {
        g.reqs.push_back(it.r);
        g.bytes += request_bytes(it.r);
      }
      e = hipEventRecord(g.ev, slot.stream);
      groups.push_back(std::move(g));
      if (e != hipSuccess) return e;
      std::lock_guard<std::mutex> lk(mu_);
      st_queue_launches_ += launches;
      st_queued_ += static_cast<int64_t>(items.size());
    }
    for (Request *r : direct) {
      std::vector<void *> ws;
      for (size_t o : r->ws_off) ws.push_back(base + o);
      e = r->plan->plain->run_batch(r->in, base + r->out_off, ws.data(), r->batch, slot.stream);
      if (e != hipSuccess) return e;
      Group g;
      g.ev = take_event_unlocked();
      g.reqs.push_back(r);
      g.bytes = request_bytes(r);
      e = hipEventRecord(g.ev, slot.stream);
      groups.push_back(std::move(g));
      if (e != hipSuccess) return e;
    }
    std::lock_guard<std::mutex> lk(mu_);
    ++st_batches_;
    return hipSuccess;
  }

  static double request_bytes(const Request *r) {
    return static_cast<double>(r->batch) * static_cast<double>(r->plan->in_bytes + r->plan->out_bytes);
  }

  hipEvent_t take_event_unlocked() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (!free_events_.empty()) {
        hipEvent_t e = free_events_.back();
        free_events_.pop_back();
        return e;
      }
    }
    hipEvent_t e = nullptr;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    return e;
  }

  Arena arena_;
  int mode_;
  int nslots_;
  int ndirect_;  // slots [0, ndirect_) run requests with their own plans; the rest take queue batches
  int64_t queue_max_fast_ = kQueueMaxBytesFast;
  int large_backlog_requests_ = kLargeBacklogRequests;
  int device_ = 0;
  int ncu_ = 1;
  std::vector<Slot> slots_;
  std::thread worker_;

  std::mutex plans_mu_;
  std::map<std::tuple<int, int64_t, int>, std::unique_ptr<ServePlan>> plans_;
  size_t table_bytes_ = 0;

  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  bool stop_ = false, failed_ = false;
  int64_t next_id_ = 0;
  std::deque<Request *> pending_, admitted_;
  size_t bypass_held_ = 0;  // arena bytes held by requests admitted ahead of an earlier one
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs_;
  int64_t inflight_ = 0;
  std::vector<hipEvent_t> free_events_;

  std::vector<double> lat_ms_;
  int64_t st_transforms_ = 0;
  double st_bytes_ = 0;
  int64_t st_batches_ = 0, st_queue_launches_ = 0, st_queued_ = 0;
  bool st_have_first_ = false;
  Clock::time_point st_first_arrival_, st_last_done_;
};

}  // namespace flagfft

struct flagfftxServer_st {
  std::unique_ptr<flagfft::Server> impl;
};

extern "C" flagfftResult flagfftxServerCreate(flagfftxServer *server, size_t hbm_budget_bytes, int max_inflight,
                                              int mode) {
  if (server == nullptr) return FLAGFFT_INVALID_VALUE;
  *server = nullptr;
  if (hbm_budget_bytes == 0 || max_inflight < 1 || (mode != FLAGFFTX_SERVE_BATCHED && mode != FLAGFFTX_SERVE_PLAIN))
    return FLAGFFT_INVALID_VALUE;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return FLAGFFT_INVALID_DEVICE;
  }
  try {
    auto s = std::make_unique<flagfftxServer_st>();
    s->impl = std::make_unique<flagfft::Server>(hbm_budget_bytes, max_inflight, mode);
    *server = s.release();
    return FLAGFFT_SUCCESS;
  } catch (const flagfft::HipError &e) {
    std::fprintf(stderr, "[flagfft] server setup failed: %s\n", e.what());
    return e.code == hipErrorOutOfMemory ? FLAGFFT_ALLOC_FAILED : FLAGFFT_SETUP_FAILED;
  } catch (const std::exception &) {
    return FLAGFFT_SETUP_FAILED;
  }
}

extern "C" flagfftResult flagfftxServerPrepare(flagfftxServer server, flagfftType type, int nx, int direction) {
  if (server == nullptr) return FLAGFFT_INVALID_PLAN;
  if (nx <= 0) return FLAGFFT_INVALID_SIZE;
  const int dir = (type == FLAGFFT_C2C || type == FLAGFFT_Z2Z) ? direction : 0;
  if ((type == FLAGFFT_C2C || type == FLAGFFT_Z2Z) && dir != FLAGFFT_FORWARD && dir != FLAGFFT_INVERSE)
    return FLAGFFT_INVALID_VALUE;
  return server->impl->prepare(static_cast<int>(type), nx, dir, nullptr);
}

extern "C" flagfftResult flagfftxServerSubmit(flagfftxServer server, flagfftType type, int nx, int batch,
                                              int direction, const void *idata, long long *ticket) {
  if (server == nullptr) return FLAGFFT_INVALID_PLAN;
  if (idata == nullptr || ticket == nullptr || batch <= 0) return FLAGFFT_INVALID_VALUE;
  if (nx <= 0) return FLAGFFT_INVALID_SIZE;
  // a request's plan is its prepared plan re-batched: unit and workgroup
  // counts are at most its element count and are held in 32 bits
  if (static_cast<int64_t>(nx) * batch > INT32_MAX) return FLAGFFT_INVALID_SIZE;
  int64_t t = 0;
  flagfftResult r = server->impl->submit(static_cast<int>(type), nx, batch, direction, idata, t);
  *ticket = t;
  return r;
}

extern "C" flagfftResult flagfftxServerPoll(flagfftxServer server, long long ticket, int *done, void **odata,
                                            double *latency_ms) {
  if (server == nullptr) return FLAGFFT_INVALID_PLAN;
  int d = 0;
  void *o = nullptr;
  double l = 0;
  flagfftResult r = server->impl->poll(ticket, false, d, o, l);
  if (done) *done = d;
  if (odata) *odata = o;
  if (latency_ms) *latency_ms = l;
  return r;
}

extern "C" flagfftResult flagfftxServerWait(flagfftxServer server, long long ticket, void **odata,
                                            double *latency_ms) {
  if (server == nullptr) return FLAGFFT_INVALID_PLAN;
  int d = 0;
  void *o = nullptr;
  double l = 0;
  flagfftResult r = server->impl->poll(ticket, true, d, o, l);
  if (odata) *odata = o;
  if (latency_ms) *latency_ms = l;
  return r;
}

extern "C" flagfftResult flagfftxServerRelease(flagfftxServer server, long long ticket) {
  if (server == nullptr) return FLAGFFT_INVALID_PLAN;
  return server->impl->release(ticket);
}

extern "C" flagfftResult flagfftxServerGetStats(flagfftxServer server, flagfftxServerStats *stats) {
  if (server == nullptr) return FLAGFFT_INVALID_PLAN;
  if (stats == nullptr) return FLAGFFT_INVALID_VALUE;
  server->impl->stats(*stats);
  return FLAGFFT_SUCCESS;
}

extern "C" flagfftResult flagfftxServerResetStats(flagfftxServer server) {
  if (server == nullptr) return FLAGFFT_INVALID_PLAN;
  server->impl->reset_stats();
  return FLAGFFT_SUCCESS;
}

extern "C" flagfftResult flagfftxServerDestroy(flagfftxServer server) {
  if (server == nullptr) return FLAGFFT_INVALID_PLAN;
  flagfftResult r = server->impl->drain();
  delete server;
  return r;
}
