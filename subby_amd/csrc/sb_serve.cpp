// Continuous-batching server: native admission loop over batch launchers.
//
// The reference's batched driver (scripts/batch_eigensub_maf.py:423-698)
// buckets blocks by tree size, pads columns, runs jax.vmap(Counts) per bucket
// within a memory budget and copies features back.  Here each of several
// lanes (a batch launcher on its own HIP stream, with pinned and device
// buffers) has its own thread: when its lane is free it admits queued
// requests first come first served (one shared queue) while the bytes in
// flight (alignments + outputs of every batch on the GPU) stay within the
// HBM budget, packs their alignments for one upload, runs them as one
// persistent launch (replayed hipGraph), downloads all outputs with one copy
// and, when the lane's event has fired, copies each request's results into
// its caller's buffers.  Host copies and model uploads of a batch are spread
// over the lane's small thread pool; lanes run concurrently, and only the
// batch submission itself takes the context lock.  Every column is computed
// independently of the rest of the batch, so a request's results are bitwise
// those of the request run alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/subby_amd.h"

extern "C" int sb_internal_error(int code, const char* msg);
extern "C" int sb_internal_device(sb_context* ctx);
extern "C" int sb_internal_model_irr(sb_context* ctx, int model_id);
extern "C" int sb_internal_pad_alphabet(int A);
extern "C" void* sb_internal_model_prepare(int A, const double* eigenvalues, const double* eigenvectors,
                                           const double* pi, int* rc);
extern "C" int sb_internal_model_commit(sb_context* ctx, int n, void* const* handles, int* ids);
extern "C" int sb_internal_model_drop_many(sb_context* ctx, int n, const int* ids);

namespace {

using Clock = std::chrono::steady_clock;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

double seconds(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

// Fixed pool of helper threads; run(n, fn) calls fn(i) for i < n on the pool
// and the calling thread and returns when all calls are done.
class Pool {
 public:
  explicit Pool(int n) {
    for (int i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      quit_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void run(int n, const std::function<void(int)>& fn) {
    if (n <= 0) return;
    if (th_.empty() || n == 1) {
      for (int i = 0; i < n; ++i) fn(i);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      active_ = int(th_.size());
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> g(mu_);
    done_cv_.wait(g, [this] { return active_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work() {
    for (;;) {
      const int i = next_.fetch_add(1);
      if (i >= n_) break;
      (*fn_)(i);
    }
  }
  void loop() {
    unsigned seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return gen_ != seen; });
        seen = gen_;
        if (quit_) return;
      }
      work();
      std::lock_guard<std::mutex> g(mu_);
      if (--active_ == 0) done_cv_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* fn_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, active_ = 0;
  unsigned gen_ = 0;
  bool quit_ = false;
};

// A resident copy of a model that requests gave inline, keyed by content:
// requests with the same eigensystem share it, and it is dropped (its arena
// record then reused once the launches that read it have finished) when more
// than kInlineModels are resident and no queued launch uses it.
struct InlineModel {
  int id = -1;
  int refs = 0;                 // admitted requests not yet completed
  uint64_t hash = 0;
  std::vector<double> key;      // A, eigenvalues, eigenvectors, pi
};
const size_t kInlineModels = 4096;

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  return h;
}

void inline_key(const sb_request& r, std::vector<double>* key, uint64_t* hash) {
  const size_t A = size_t(r.A);
  key->resize(1 + 2 * A + A * A);
  (*key)[0] = double(A);
  std::memcpy(key->data() + 1, r.eigenvalues, A * 8);
  std::memcpy(key->data() + 1 + A, r.eigenvectors, A * A * 8);
  std::memcpy(key->data() + 1 + A + A * A, r.pi, A * 8);
  *hash = fnv1a(1469598103934665603ull, key->data(), key->size() * 8);
}

struct Item {
  sb_request r;
  InlineModel* im = nullptr;        // inline model this request holds a reference to
  int A = 0;
  int kind = 0;                     // 2 x padded alphabet + irreversible: one kind per launch
  size_t in_bytes = 0, out_bytes = 0;
  size_t in_off = 0, out_off = 0;   // within the lane buffers
  double* done_slot = nullptr;      // replay: completion time (seconds since server start)
  sb_request_timing* tslot = nullptr;  // timed replay: where the latency went
};

struct Lane {
  sb_batch* batch = nullptr;
  std::thread th;
  std::unique_ptr<Pool> pool;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  hipEvent_t tev[4] = {};           // timed batches: before upload, after upload, kernels, download
  bool timed = false;               // this batch records tev
  double t_admit = 0, t_enq = 0, t_gpu = 0, plan_ms = 0, models_ms = 0;   // seconds since server start; ms
  long long cols = 0;
  int index = 0;
  char* h_in = nullptr;
  char* d_in = nullptr;
  char* h_out = nullptr;
  char* d_out = nullptr;
  size_t in_cap = 0, out_cap = 0;
  std::vector<Item> items;
  size_t bytes = 0;
  bool busy = false;
};

size_t out_layout(const Item& it, size_t* oc, size_t* ol, size_t* orr) {
  const size_t C = size_t(it.r.C), A = size_t(it.A);
  *oc = 0;
  *ol = align256(A * A * C * 4);
  *orr = *ol + align256(C * 8);
  return *orr + align256(A * C * 4);
}

}  // namespace

struct sb_server {
  sb_context* ctx = nullptr;
  int device = 0;
  std::vector<Lane> lanes;
  size_t budget = 0;
  int max_blocks = 0;
  long long max_cols = 0;
  std::mutex mu;
  std::condition_variable cv;       // lane threads: new requests / stop
  std::condition_variable idle_cv;  // drain: everything completed
  std::deque<Item> queue;
  std::vector<std::pair<long long, double>> done;
  // inline models, least recently released first (guarded by mmu)
  std::mutex mmu;
  std::list<InlineModel> inline_models;
  std::unordered_multimap<uint64_t, std::list<InlineModel>::iterator> inline_index;
  size_t in_flight = 0;
  long long pending = 0;            // submitted, not completed
  int err = 0;
  std::string errmsg;
  bool stop = false;
  Clock::time_point t0;
};

namespace {

void set_error(sb_server* s, int code, const std::string& msg) {
  std::lock_guard<std::mutex> g(s->mu);
  if (!s->err) {
    s->err = code;
    s->errmsg = msg;
  }
}

int hip_check(sb_server* s, hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  set_error(s, SB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return SB_ERR_HIP;
}

// Grow a lane buffer pair (pinned host + device); the lane is idle here.
int grow(sb_server* s, char** h, char** d, size_t* cap, size_t want) {
  if (want <= *cap) return 0;
  want = std::max(want, *cap * 3 / 2);
  want = (want + 4095) & ~size_t(4095);
  if (*h) (void)hipHostFree(*h);
  if (*d) (void)hipFree(*d);
  *h = nullptr;
  *d = nullptr;
  *cap = 0;
  if (hipHostMalloc(reinterpret_cast<void**>(h), want, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(d), want) != hipSuccess) {
    (void)hipGetLastError();
    set_error(s, SB_ERR_OOM, "server lane buffer allocation failed");
    return SB_ERR_OOM;
  }
  *cap = want;
  return 0;
}

// Give back the inline-model references of a lane's finished requests.
void release_models(sb_server* s, Lane& L) {
  std::lock_guard<std::mutex> g(s->mmu);
  for (Item& it : L.items) {
    InlineModel* im = it.im;
    it.im = nullptr;
    if (!im || --im->refs > 0) continue;
    // most recently released last: eviction takes from the front
    for (auto p = s->inline_index.equal_range(im->hash); p.first != p.second; ++p.first)
      if (&*p.first->second == im) {
        s->inline_models.splice(s->inline_models.end(), s->inline_models, p.first->second);
        break;
      }
  }
}

// Mark a lane's requests finished (after an error the outputs are undefined).
void finish(sb_server* s, Lane& L) {
  release_models(s, L);
  const double now = seconds(s->t0, Clock::now());
  float ms[3] = {0.f, 0.f, 0.f};
  if (L.timed)
    for (int k = 0; k < 3; ++k)
      if (hipEventElapsedTime(&ms[k], L.tev[k], L.tev[k + 1]) != hipSuccess) {
        (void)hipGetLastError();
        ms[k] = -1.f;
      }
  std::lock_guard<std::mutex> g(s->mu);
  for (Item& it : L.items) {
    if (it.done_slot) *it.done_slot = now;
    if (sb_request_timing* t = it.tslot) {
      t->admitted_s = L.t_admit;
      t->enqueued_s = L.t_enq;
      t->gpu_done_s = L.t_gpu;
      t->done_s = now;
      t->upload_ms = ms[0];
      t->kernel_ms = ms[1];
      t->download_ms = ms[2];
      t->plan_ms = L.plan_ms;
      t->models_ms = L.models_ms;
      t->batch_columns = L.cols;
      t->batch_requests = int(L.items.size());
      t->lane = L.index;
    }
    s->done.emplace_back(it.r.tag, now);
  }
  s->in_flight -= L.bytes;
  s->pending -= (long long)L.items.size();
  L.items.clear();
  L.bytes = 0;
  L.busy = false;
  s->idle_cv.notify_all();
}

// Give every inline-model request of the lane a resident model id, holding
// a reference until the request completes; idle inline models beyond
// kInlineModels are evicted first.
int resolve_models(sb_server* s, Lane& L) {
  const int n = int(L.items.size());
  std::vector<std::vector<double>> keys(n);
  std::vector<uint64_t> hashes(n, 0);
  std::vector<int> miss, follow(n, -1);   // follow[k]: earlier miss of this batch with k's content
  std::vector<int> evicted;
  L.pool->run(n, [&](int k) {
    const Item& it = L.items[k];
    if (it.r.model_id < 0) inline_key(it.r, &keys[k], &hashes[k]);
  });
  {
    std::lock_guard<std::mutex> g(s->mmu);
    std::unordered_multimap<uint64_t, int> fresh;
    for (int k = 0; k < n; ++k) {
      Item& it = L.items[k];
      if (it.r.model_id >= 0) continue;
      bool found = false;
      for (auto p = s->inline_index.equal_range(hashes[k]); p.first != p.second && !found; ++p.first) {
        InlineModel& im = *p.first->second;
        if (im.key == keys[k]) {
          ++im.refs;
          it.im = &im;
          it.r.model_id = im.id;
          found = true;
        }
      }
      for (auto p = fresh.equal_range(hashes[k]); p.first != p.second && !found; ++p.first)
        if (keys[p.first->second] == keys[k]) {
          follow[k] = p.first->second;
          found = true;
        }
      if (!found) {
        fresh.emplace(hashes[k], k);
        miss.push_back(k);
      }
    }
    // make room: evict idle inline models, least recently released first, so
    // that the new ones can take over their arena records
    for (auto q = s->inline_models.begin();
         q != s->inline_models.end() && s->inline_models.size() + miss.size() > kInlineModels;) {
      if (q->refs > 0) {
        ++q;
        continue;
      }
      evicted.push_back(q->id);
      for (auto p = s->inline_index.equal_range(q->hash); p.first != p.second; ++p.first)
        if (p.first->second == q) {
          s->inline_index.erase(p.first);
          break;
        }
      q = s->inline_models.erase(q);
    }
  }
  // evicted records first, so that the new models can reuse them; host math
  // of the new models on the pool, then one context lock for all of them
  if (!evicted.empty()) (void)sb_internal_model_drop_many(s->ctx, int(evicted.size()), evicted.data());
  std::atomic<int> mrc{0};
  std::vector<void*> handles(miss.size(), nullptr);
  std::vector<int> ids(miss.size(), -1);
  L.pool->run(int(miss.size()), [&](int q) {
    const Item& it = L.items[miss[q]];
    int e = 0;
    handles[q] = sb_internal_model_prepare(it.A, it.r.eigenvalues, it.r.eigenvectors, it.r.pi, &e);
    if (e) {
      set_error(s, e, std::string("request model: ") + sb_last_error());
      mrc.store(e);
    }
  });
  if (!miss.empty()) {
    const int e = sb_internal_model_commit(s->ctx, int(miss.size()), handles.data(), ids.data());
    if (e) {
      set_error(s, e, std::string("request model: ") + sb_last_error());
      mrc.store(e);
    }
  }
  for (size_t q = 0; q < miss.size(); ++q) L.items[miss[q]].r.model_id = ids[q];
  std::lock_guard<std::mutex> g(s->mmu);
  for (int k : miss) {
    Item& it = L.items[k];
    if (it.r.model_id < 0) continue;
    InlineModel im;
    im.id = it.r.model_id;
    im.refs = 1;
    im.hash = hashes[k];
    im.key.swap(keys[k]);
    s->inline_models.push_back(std::move(im));
    auto pos = std::prev(s->inline_models.end());
    s->inline_index.emplace(pos->hash, pos);
    it.im = &*pos;
  }
  for (int k = 0; k < n; ++k)
    if (follow[k] >= 0) {
      Item& it = L.items[k];
      it.im = L.items[follow[k]].im;
      it.r.model_id = L.items[follow[k]].r.model_id;
      if (it.im) ++it.im->refs;
    }
  return mrc.load();
}

void launch(sb_server* s, Lane& L, std::vector<Item>& items, size_t bytes) {
  L.items.swap(items);
  L.bytes = bytes;
  L.busy = true;
  const int n = int(L.items.size());
  L.timed = n > 0 && L.items[0].tslot != nullptr;
  L.t_admit = seconds(s->t0, Clock::now());
  L.t_enq = L.t_gpu = L.plan_ms = L.models_ms = 0;
  L.cols = 0;
  size_t in_total = 0, out_total = 0;
  for (Item& it : L.items) {
    L.cols += it.r.C;
    it.in_off = in_total;
    in_total += it.in_bytes;
    it.out_off = out_total;
    out_total += it.out_bytes;
  }
  int rc = grow(s, &L.h_in, &L.d_in, &L.in_cap, in_total);
  if (!rc) rc = grow(s, &L.h_out, &L.d_out, &L.out_cap, out_total);
  // models given inline: a resident copy with the same content is shared;
  // the others go to the resident arena now (their host math runs on the
  // pool; the context lock covers only the arena and the model table)
  std::vector<sb_block> blocks(n);
  if (!rc) {
    const Clock::time_point m0 = Clock::now();
    rc = resolve_models(s, L);
    L.models_ms = 1e3 * seconds(m0, Clock::now());
  }
  if (!rc) {
    L.pool->run(n, [&](int k) {
      const Item& it = L.items[k];
      std::memcpy(L.h_in + it.in_off, it.r.alignment, size_t(it.r.R) * size_t(it.r.C));
    });
    if (L.timed) rc = hip_check(s, hipEventRecord(L.tev[0], L.stream), "timing event");
    if (!rc)
      rc = hip_check(s, hipMemcpyAsync(L.d_in, L.h_in, in_total, hipMemcpyHostToDevice, L.stream),
                     "upload");
    if (!rc && L.timed) rc = hip_check(s, hipEventRecord(L.tev[1], L.stream), "timing event");
  }
  if (!rc) {
    for (int k = 0; k < n; ++k) {
      const Item& it = L.items[k];
      size_t oc, ol, orr;
      out_layout(it, &oc, &ol, &orr);
      char* base = L.d_out + it.out_off;
      sb_block& b = blocks[k];
      std::memset(&b, 0, sizeof b);
      b.model_id = it.r.model_id;
      b.R = it.r.R;
      b.C = it.r.C;
      b.d_alignment = reinterpret_cast<const int8_t*>(L.d_in + it.in_off);
      b.parent_index = it.r.parent_index;
      b.distances = it.r.distances;
      b.want = it.r.want;
      b.mask_mode = it.r.mask_mode;
      b.d_counts = base + oc;
      b.d_loglike = reinterpret_cast<double*>(base + ol);
      b.d_root = base + orr;
    }
    const Clock::time_point p0 = Clock::now();
    const int src = sb_batch_submit(L.batch, n, blocks.data(), SB_F32, 1);
    L.plan_ms = 1e3 * seconds(p0, Clock::now());
    if (src) {
      set_error(s, src, std::string("batch: ") + sb_last_error());
      rc = src;
    }
  }
  if (!rc && L.timed) rc = hip_check(s, hipEventRecord(L.tev[2], L.stream), "timing event");
  if (!rc)
    rc = hip_check(s, hipMemcpyAsync(L.h_out, L.d_out, out_total, hipMemcpyDeviceToHost, L.stream),
                   "download");
  if (!rc && L.timed) rc = hip_check(s, hipEventRecord(L.tev[3], L.stream), "timing event");
  if (!rc) rc = hip_check(s, hipEventRecord(L.ev, L.stream), "event");
  L.t_enq = seconds(s->t0, Clock::now());
  if (rc) L.timed = false;
  if (rc) {
    (void)hipStreamSynchronize(L.stream);
    finish(s, L);
  }
}

void complete(sb_server* s, Lane& L) {
  const int n = int(L.items.size());
  L.pool->run(n, [&](int k) {
    const Item& it = L.items[k];
    size_t oc, ol, orr;
    out_layout(it, &oc, &ol, &orr);
    const char* base = L.h_out + it.out_off;
    const size_t C = size_t(it.r.C), A = size_t(it.A);
    if (it.r.counts && (it.r.want & SB_WANT_COUNTS)) std::memcpy(it.r.counts, base + oc, A * A * C * 4);
    if (it.r.loglike && (it.r.want & SB_WANT_LOGLIKE)) std::memcpy(it.r.loglike, base + ol, C * 8);
    if (it.r.root && (it.r.want & SB_WANT_ROOT)) std::memcpy(it.r.root, base + orr, A * C * 4);
  });
  finish(s, L);
}

// One lane's thread: admit, launch, wait for the lane's event, complete.
void run_lane(sb_server* s, Lane& L) {
  (void)hipSetDevice(s->device);
  std::vector<Item> take;
  for (;;) {
    take.clear();
    size_t bytes = 0;
    {
      std::unique_lock<std::mutex> g(s->mu);
      for (;;) {
        long long cols = 0;
        // One launch runs one kind (padded alphabet, reversibility): the oldest
        // request sets it, and requests of that kind are taken from a bounded
        // window of the queue in order; the others keep their places (the
        // oldest of them leads the next admission, so none waits forever).
        auto q = s->queue.begin();
        int scanned = 0;
        const int window = 4 * s->max_blocks;
        while (q != s->queue.end() && int(take.size()) < s->max_blocks && scanned < window) {
          ++scanned;
          const Item& it = *q;
          if (!take.empty() && it.kind != take[0].kind) {
            ++q;
            continue;
          }
          const size_t b = it.in_bytes + it.out_bytes;
          if (!take.empty() && (cols + it.r.C > s->max_cols || s->in_flight + bytes + b > s->budget))
            break;
          if (take.empty() && s->in_flight > 0 && s->in_flight + b > s->budget) break;
          bytes += b;
          cols += it.r.C;
          take.push_back(it);
          q = s->queue.erase(q);   // `it` is gone from here on
        }
        if (!take.empty()) break;
        if (s->stop && s->queue.empty()) return;
        // nothing admissible: wait for arrivals or for another lane to free budget
        s->cv.wait_for(g, std::chrono::milliseconds(1));
      }
      s->in_flight += bytes;
    }
    launch(s, L, take, bytes);
    if (!L.busy) continue;            // failed launch: already finished
    for (;;) {
      const hipError_t q = hipEventQuery(L.ev);
      if (q == hipSuccess) {
        L.t_gpu = seconds(s->t0, Clock::now());
        complete(s, L);
        break;
      }
      if (q != hipErrorNotReady) {
        hip_check(s, q, "lane event");
        (void)hipStreamSynchronize(L.stream);
        finish(s, L);
        break;
      }
      std::this_thread::yield();
    }
    s->cv.notify_all();               // budget freed
  }
}

int validate(const sb_request& r, int* A) {
  char buf[160];
  if (r.R < 1 || r.C < 0 || !r.parent_index || !r.distances || (r.C > 0 && !r.alignment)) {
    std::snprintf(buf, sizeof buf, "request tag %lld: bad sizes or null tree/alignment", r.tag);
    return sb_internal_error(SB_ERR_ARG, buf);
  }
  if (r.mask_mode != SB_MASK_NONE && r.mask_mode != SB_MASK_AUTO) {
    std::snprintf(buf, sizeof buf, "request tag %lld: mask_mode must be none or auto", r.tag);
    return sb_internal_error(SB_ERR_ARG, buf);
  }
  if (r.model_id < 0) {
    if (r.A < 2 || r.A > 64 || !r.eigenvalues || !r.eigenvectors || !r.pi) {
      std::snprintf(buf, sizeof buf, "request tag %lld: bad inline model", r.tag);
      return sb_internal_error(SB_ERR_MODEL, buf);
    }
  } else if (r.A < 2 || r.A > 64) {
    std::snprintf(buf, sizeof buf, "request tag %lld: A (the model's alphabet) must be set", r.tag);
    return sb_internal_error(SB_ERR_ARG, buf);
  }
  *A = r.A;
  return SB_OK;
}

int push(sb_server* s, int n, const sb_request* reqs, double* done_slots,
         sb_request_timing* tslots) {
  std::vector<Item> items(n);
  for (int i = 0; i < n; ++i) {
    Item& it = items[i];
    it.r = reqs[i];
    const int rc = validate(it.r, &it.A);
    if (rc) return rc;
    int irr = 0;   // models given inline are reversible
    if (it.r.model_id >= 0) {
      irr = sb_internal_model_irr(s->ctx, it.r.model_id);
      if (irr < 0) {
        char buf[128];
        std::snprintf(buf, sizeof buf, "request tag %lld: unknown model id %d", it.r.tag, it.r.model_id);
        return sb_internal_error(SB_ERR_MODEL, buf);
      }
    }
    it.kind = 2 * sb_internal_pad_alphabet(it.A) + irr;
    it.in_bytes = align256(size_t(it.r.R) * size_t(it.r.C));
    size_t oc, ol, orr;
    it.out_bytes = out_layout(it, &oc, &ol, &orr);
    it.done_slot = done_slots ? done_slots + i : nullptr;
    it.tslot = tslots ? tslots + i : nullptr;
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->err) return sb_internal_error(s->err, s->errmsg.c_str());
    for (Item& it : items) s->queue.push_back(it);
    s->pending += n;
  }
  s->cv.notify_all();
  return SB_OK;
}

// Submit request i at arrival_s[i] seconds after the call, wait for all;
// completion times (and, with `timing`, the breakdown) relative to the call.
int replay(sb_server* s, int n, sb_request* requests, const double* arrival_s, double* done_s,
           sb_request_timing* timing) {
  for (int i = 0; i < n; ++i) requests[i].tag = i;
  if (timing) std::memset(timing, 0, sizeof(sb_request_timing) * size_t(n));
  const Clock::time_point start = Clock::now();
  const double base = seconds(s->t0, start);
  int i = 0;
  while (i < n) {
    const double now = seconds(start, Clock::now());
    if (arrival_s[i] > now) {
      const double wait = arrival_s[i] - now;
      if (wait > 3e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 2e-4));
      else std::this_thread::yield();
      continue;
    }
    int j = i;
    while (j < n && arrival_s[j] <= now) ++j;   // everything that has arrived
    const int rc = push(s, j - i, requests + i, done_s ? done_s + i : nullptr,
                        timing ? timing + i : nullptr);
    if (rc) return rc;
    i = j;
  }
  const int rc = sb_server_drain(s);
  for (int k = 0; k < n; ++k) {
    if (done_s) done_s[k] -= base;
    if (timing) {
      sb_request_timing& t = timing[k];
      t.admitted_s -= base;
      t.enqueued_s -= base;
      t.gpu_done_s -= base;
      t.done_s -= base;
    }
  }
  {
    // replay completions are reported through done_s / timing, not poll
    std::lock_guard<std::mutex> g(s->mu);
    s->done.clear();
  }
  return rc;
}

}  // namespace

extern "C" {

int sb_server_create(sb_context* ctx, int lanes, long long hbm_budget, int max_blocks,
                     long long max_cols, sb_server** out) {
  if (!ctx || !out || lanes < 1 || lanes > 16 || hbm_budget <= 0 || max_blocks < 1 || max_cols < 1)
    return sb_internal_error(SB_ERR_ARG, "sb_server_create: bad argument");
  sb_server* s = new sb_server();
  s->ctx = ctx;
  s->device = sb_internal_device(ctx);
  s->budget = size_t(hbm_budget);
  s->max_blocks = max_blocks;
  s->max_cols = max_cols;
  s->t0 = Clock::now();
  if (hipSetDevice(s->device) != hipSuccess) {
    delete s;
    return sb_internal_error(SB_ERR_HIP, "sb_server_create: hipSetDevice failed");
  }
  s->lanes.resize(lanes);
  for (Lane& L : s->lanes) {
    L.index = int(&L - s->lanes.data());
    bool ok = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&L.ev, hipEventDisableTiming) == hipSuccess;
    for (hipEvent_t& e : L.tev) ok = ok && hipEventCreate(&e) == hipSuccess;
    if (!ok) {
      sb_server_destroy(s);
      return sb_internal_error(SB_ERR_HIP, "sb_server_create: stream/event creation failed");
    }
    const int rc = sb_batch_create(ctx, L.stream, &L.batch);
    if (rc) {
      std::string msg = sb_last_error();
      sb_server_destroy(s);
      return sb_internal_error(rc, msg.c_str());
    }
  }
  // host threads: one per lane plus its copy helpers, about 16 in all
  const int hw = int(std::max(2u, std::thread::hardware_concurrency()));
  const int helpers = std::max(1, std::min(16, hw) / lanes - 1);
  for (Lane& L : s->lanes) {
    L.pool.reset(new Pool(helpers));
    L.th = std::thread(run_lane, s, std::ref(L));
  }
  *out = s;
  return SB_OK;
}

int sb_server_destroy(sb_server* s) {
  if (!s) return SB_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (Lane& L : s->lanes)
    if (L.th.joinable()) L.th.join();
  for (Lane& L : s->lanes) {
    L.pool.reset();
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    if (L.batch) sb_batch_destroy(L.batch);
    if (L.h_in) (void)hipHostFree(L.h_in);
    if (L.h_out) (void)hipHostFree(L.h_out);
    if (L.d_in) (void)hipFree(L.d_in);
    if (L.d_out) (void)hipFree(L.d_out);
    if (L.ev) (void)hipEventDestroy(L.ev);
    for (hipEvent_t e : L.tev)
      if (e) (void)hipEventDestroy(e);
    if (L.stream) (void)hipStreamDestroy(L.stream);
  }
  for (const InlineModel& im : s->inline_models) (void)sb_model_drop(s->ctx, im.id);
  const int err = s->err;
  std::string msg = s->errmsg;
  delete s;
  return err ? sb_internal_error(err, msg.c_str()) : SB_OK;
}

int sb_server_submit(sb_server* s, int n, const sb_request* requests) {
  if (!s || n < 0 || (n > 0 && !requests)) return sb_internal_error(SB_ERR_ARG, "sb_server_submit: bad argument");
  return push(s, n, requests, nullptr, nullptr);
}

int sb_server_poll(sb_server* s, int max, long long* tags, double* done_s, int* n_done) {
  if (!s || !n_done || max < 0 || (max > 0 && (!tags || !done_s)))
    return sb_internal_error(SB_ERR_ARG, "sb_server_poll: bad argument");
  std::lock_guard<std::mutex> g(s->mu);
  const int k = int(std::min<size_t>(size_t(max), s->done.size()));
  for (int i = 0; i < k; ++i) {
    tags[i] = s->done[i].first;
    done_s[i] = s->done[i].second;
  }
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n_done = k;
  if (s->err) return sb_internal_error(s->err, s->errmsg.c_str());
  return SB_OK;
}

int sb_server_drain(sb_server* s) {
  if (!s) return sb_internal_error(SB_ERR_ARG, "sb_server_drain: null server");
  std::unique_lock<std::mutex> g(s->mu);
  s->idle_cv.wait(g, [s] { return s->pending == 0; });
  if (s->err) return sb_internal_error(s->err, s->errmsg.c_str());
  return SB_OK;
}

int sb_server_replay(sb_server* s, int n, sb_request* requests, const double* arrival_s,
                     double* done_s) {
  if (!s || n < 0 || (n > 0 && (!requests || !arrival_s || !done_s)))
    return sb_internal_error(SB_ERR_ARG, "sb_server_replay: bad argument");
  return replay(s, n, requests, arrival_s, done_s, nullptr);
}

int sb_server_replay_timed(sb_server* s, int n, sb_request* requests, const double* arrival_s,
                           sb_request_timing* timing) {
  if (!s || n < 0 || (n > 0 && (!requests || !arrival_s || !timing)))
    return sb_internal_error(SB_ERR_ARG, "sb_server_replay_timed: bad argument");
  return replay(s, n, requests, arrival_s, nullptr, timing);
}

}  // extern "C"
