// Host side of the C ABI (include/subby_amd.h): context, resident model
// table, tree planner and the batch launcher.
//
// The planner replaces the per-call Python/JAX preprocessing of the reference
// (children_of, subby/jax/_utils.py:46-75; compute_sub_matrices,
// diagonalize.py:141-172; compute_J, eigensub.py:7-50): tree tables are built
// here in C++, and M(t), J(t) are evaluated in FP64 on the device by
// sb_branch_params_kernel from the resident eigensystem.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/subby_amd.h"
#include "sb_plan.h"

extern "C" hipError_t sb_launch_fused_small(int dtype, int AP, bool f81, bool grad, bool irr,
                                            bool rows, bool split, dim3 grid, dim3 block, int lds,
                                            const SbLaunch* L, hipStream_t s);
extern "C" hipError_t sb_launch_gather_cols(const void* src, void* dst, int esz, int rows, int C,
                                            const int* const* cols, hipStream_t s);
extern "C" hipError_t sb_launch_scatter_cols(const void* src, void* dst, int esz, int rows, int C,
                                             const int* const* cols, hipStream_t s);
extern "C" hipError_t sb_launch_prior_swap(double* ll, void* root, int dtype, int A, int C,
                                           const int* const* cols, const double* w, const int* swap,
                                           hipStream_t s);
extern "C" hipError_t sb_launch_mixture_softmax(const double* ll, double* post, int K, int C,
                                                const double* logw, hipStream_t s);
extern "C" hipError_t sb_launch_large(int dtype, int AP, bool f81, bool grad, int out,
                                      int nwaves, const SbLaunch* L, hipStream_t s);
extern "C" long long sb_large_wave_stride(int dtype, int AP, int nint_max, bool f81, int out);
extern "C" int sb_large_tile_cols(int dtype, int AP, bool f81, int out);
extern "C" int sb_fused_small_occupancy(int dtype, int AP, bool f81, bool grad, bool irr,
                                        bool rows, bool split, int threads, int lds);
extern "C" int sb_large_occupancy(int dtype, int AP, bool f81, bool grad, int out);
extern "C" int sb_f81w_tile_cols();
extern "C" int sb_f81w_waves_per_group();
extern "C" long long sb_f81w_wave_stride(int AP, int nint_max);
extern "C" int sb_f81w_occupancy(int AP);
extern "C" hipError_t sb_launch_f81w(int AP, int nwaves, const SbLaunch* L, hipStream_t s);
extern "C" int sb_post_occupancy(int dtype, int AP, bool f81, bool irr);
extern "C" hipError_t sb_launch_post(int dtype, int AP, bool f81, bool irr, int nwaves,
                                     const SbLaunch* L, hipStream_t s);
extern "C" int sb_gen2_tile_cols();
extern "C" long long sb_gen2_tile_bytes(int AP, int rmax);
extern "C" hipError_t sb_launch_gen2_pass(int AP, int tile0, int slot0, int ntiles, int rmax, int branch, int pass,
                                          const SbLaunch* L, hipStream_t s);
extern "C" int sb_gen2_pass1_per_cu(int AP);
extern "C" hipError_t sb_launch_gen2(int AP, int tile0, int ntiles, int rmax, int branch, const SbLaunch* L,
                                     hipStream_t s);

// Stage memory of the two-pass general path: tiles are processed in rounds
// that fit this many bytes.
static const size_t kGen2StageBudget = size_t(8) << 30;
extern "C" hipError_t sb_launch_branch_params(int dtype, bool irr, int nwg, const SbLaunch* L,
                                              hipStream_t s);
extern "C" hipError_t sb_launch_reset(const SbLaunch* L, hipStream_t s);
extern "C" hipError_t sb_launch_scatter_runs(const char* stage, const SbScatterRun* runs, int nruns,
                                             hipStream_t s);
extern "C" int sb_internal_noalloc_scratch(int device, void** ptr);
extern "C" hipError_t sb_launch_grad_reduce(int nwg, const SbLaunch* L, hipStream_t s);
struct SbPlainBuffers {
  void* p = nullptr;
  size_t cap = 0;
};
extern "C" hipError_t sb_plain_run(const int8_t* aln, const int32_t* parent, int R, int C,
                                   const float* d_bt, const float* d_model, int mask_mode,
                                   float* counts, float* root, double* ll, SbPlainBuffers* buf,
                                   hipStream_t s);

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define SB_HIP(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail(SB_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

int pad_alphabet(int A) {
  static const int sizes[] = {4, 8, 16, 20, 32, 64};
  for (int s : sizes)
    if (A <= s) return s;
  return -1;
}

bool is_small(int AP) { return AP <= 8; }

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// A launch that may read a model's records: submission `seq` of the launcher
// with serial number `serial`.
struct Fence {
  uint64_t serial, seq;
};

struct Model {
  bool alive = false;
  int A = 0, AP = 0;
  bool f81s = false;  // F81-structured: M(t) = e I + (1 - e) 1 pi^T
  bool irr = false;   // irreversible (complex eigensystem) model
  std::vector<double> mu, V, pi;   // host copies (plain path)
  double* d64 = nullptr;  // FP64 master
  void* rec[2] = {nullptr, nullptr};  // float / double records
  char* base = nullptr;   // arena record (FP64 master + float + double records)
  size_t bytes = 0;
  // the last submission of every launcher that referenced the model
  mutable std::vector<Fence> uses;
};

// Arena record of a dropped model; reused by a later sb_model_put of the same
// size once every launch in `uses` has finished.
struct Retired {
  char* base;
  size_t bytes;
  std::vector<Fence> uses;
};

}  // namespace

struct ModelRun {  // one pending host->device copy of model records
  char* dst;
  size_t src, len;
};

struct sb_context {
  int device = 0;
  int num_cus = 256;
  std::vector<Model> models;
  std::vector<int> free_ids;        // ids of dropped models, reused by sb_model_put
  std::mutex mu;
  // Model arena: records of every model live in large device chunks that never
  // move.  sb_model_put stages the records in pinned memory (a rotation of
  // kPend buffers); the next submission on any stream flushes them before its
  // kernels (a few merged copies, or one copy plus a scatter kernel), and
  // every submission waits for the last flush.
  std::vector<void*> arena_chunks;
  char* arena_cur = nullptr;
  size_t arena_off = 0, arena_cap = 0;
  size_t arena_reserved = 0;        // bytes handed out from chunks (live + retired records)
  // dropped records by size, oldest first, waiting for their launches to finish
  std::unordered_map<size_t, std::list<Retired>> retired;
  size_t n_retired = 0;
  std::vector<sb_batch*> live;      // launchers of this context (fences name them by serial)
  uint64_t next_serial = 0;
  static constexpr int kPend = 4;    // staging buffers in rotation
  char* pend[kPend] = {};
  size_t pend_cap[kPend] = {};
  size_t pend_len = 0;
  int pend_k = 0;
  bool pend_busy[kPend] = {};
  hipEvent_t pend_ev[kPend] = {};
  char* dstage[kPend] = {};          // device copies of pend[] for the scattered upload
  size_t dstage_cap[kPend] = {};
  std::vector<ModelRun> runs;
  hipEvent_t model_ev = nullptr;
  bool model_ev_valid = false;
  sb_batch* internal = nullptr;
  // launcher of sb_counts_noalloc: every device buffer lies in the static
  // no-allocation scratch array (null if its set-up failed: noalloc_err says why)
  sb_batch* noalloc = nullptr;
  std::string noalloc_err;
  SbPlainBuffers plain;
  DevBuf plain_bt;
  std::vector<float> plain_bt_host;
  std::vector<double> plain_ll_scratch;
};

struct GraphKey {
  int dtype, AP, f81, grad, large, nwaves, wpg, irr, rows, lds, split;
  bool operator==(const GraphKey& o) const {
    return dtype == o.dtype && AP == o.AP && f81 == o.f81 && grad == o.grad && large == o.large &&
           nwaves == o.nwaves && wpg == o.wpg && irr == o.irr && rows == o.rows && lds == o.lds &&
           split == o.split;
  }
};

struct sb_batch {
  sb_context* ctx = nullptr;
  hipStream_t stream = nullptr;
  // submissions: seq counts them; ring_ev[k] was recorded after submission
  // ring_seq[k] (the newest submissions, one event each)
  static const int kRing = 8;
  uint64_t serial = 0, seq = 0;
  hipEvent_t ring_ev[kRing] = {};
  uint64_t ring_seq[kRing] = {};
  SbLaunch* header = nullptr;       // start of the plan buffer (fixed until it grows)
  // The plan is staged in two pinned host buffers (growable launchers):
  // submission k builds it in host[k % 2] while the copy of submission
  // k - 1's plan may still be pending, and uploads it in order, unless the
  // device plan already holds the same bytes (the same batch launched again;
  // plans up to kPlanCache bytes are compared).  (An upload on a second
  // stream, overlapping the previous launch, cost the serving lanes more than
  // it saved: their extra streams share the process's four hardware queues.)
  // The no-allocation launcher keeps one staging buffer.
  static const size_t kPlanCache = 256 << 10;
  DevBuf plan, bt, scratch, part, stage;
  // calls with several models: gathered alignment / mask, staged outputs,
  // column maps (sb_counts_multi, sb_inside_outside_multi, sb_mixture_posterior)
  DevBuf xaln, xmask, xout, xcols;
  std::vector<char> xhost;          // host side of xcols; rewritten after xev
  hipEvent_t xev = nullptr;
  bool xpending = false;
  int* counter = nullptr;
  char* host[2] = {};               // pinned staging: header + plan
  size_t host_cap[2] = {};
  hipEvent_t staged[2] = {};        // host[s] may be rewritten after this
  bool staged_pending[2] = {};
  int pslot = 0;                    // staging buffer of the next submission
  std::vector<char> uploaded;       // the bytes the device plan holds (small plans)
  hipEvent_t done = nullptr;        // last submission finished (its ring event, or done_own)
  hipEvent_t done_own = nullptr;
  int waves_per_group = 4;
  int waves_per_cu = 0;             // 0: as many as the kernel's occupancy allows
  int gen2_round = 0;               // two-pass path: tiles per round (0: as the stage budget allows)
  int split = 1;                    // split walks: 0 never, 1 wherever a tree names a subtree
  int last_split = 0;               // the last submission walked its tiles with two waves
  int split_min = 8;                // trees with fewer internal nodes walk with one wave
  // fixed buffers (the no-allocation launcher): never grown, never freed here
  bool fixed = false;
  int g2_ntiles = 0, g2_round = 0;  // two-pass general path: tiles, tiles per round
  int g2_rmax = 1;                  // and the largest tree of the launch
  // second stream of the two-pass path: pass 2 of a round's first tiles beside
  // the tail of its pass 1 (created on first use)
  hipStream_t aux = nullptr;
  hipEvent_t g2_ev[2] = {};
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
  // kernel timing (eager submissions only)
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> tev;  // recorded pairs
  std::vector<std::pair<hipEvent_t, hipEvent_t>> tfree;
};

namespace {

// Grow a device buffer; waits for the launcher's stream before freeing the old one.
int ensure(sb_batch* b, DevBuf& buf, size_t bytes) {
  if (bytes <= buf.cap) return SB_OK;
  if (b->fixed)
    return fail(SB_ERR_ARG, "the call needs %zu bytes of a buffer the no-allocation scratch holds %zu of "
                "(scratch %llu MiB per GPU)", bytes, buf.cap, (unsigned long long)(SB_NOALLOC_BYTES >> 20));
  SB_HIP(hipStreamSynchronize(b->stream));
  size_t want = std::max(bytes, buf.cap * 3 / 2);
  if (buf.p) SB_HIP(hipFree(buf.p));
  buf.p = nullptr;
  buf.cap = 0;
  want = (want + 4095) & ~size_t(4095);
  if (hipMalloc(&buf.p, want) != hipSuccess) {
    (void)hipGetLastError();
    buf.p = nullptr;
    return fail(SB_ERR_OOM, "device allocation of %zu bytes failed", want);
  }
  buf.cap = want;
  return SB_OK;
}

int ensure_host(sb_batch* b, int sl, size_t bytes) {
  if (b->staged_pending[sl]) {
    SB_HIP(hipEventSynchronize(b->staged[sl]));
    b->staged_pending[sl] = false;
  }
  if (bytes <= b->host_cap[sl]) return SB_OK;
  if (b->fixed)
    return fail(SB_ERR_ARG, "the call's launch plan needs %zu bytes; the no-allocation plan area holds %zu",
                bytes, b->host_cap[sl]);
  size_t want = std::max(bytes, b->host_cap[sl] * 3 / 2);
  if (b->host[sl]) SB_HIP(hipHostFree(b->host[sl]));
  b->host[sl] = nullptr;
  want = (want + 4095) & ~size_t(4095);
  if (hipHostMalloc(reinterpret_cast<void**>(&b->host[sl]), want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    b->host[sl] = nullptr;
    b->host_cap[sl] = 0;
    return fail(SB_ERR_OOM, "pinned allocation of %zu bytes failed", want);
  }
  b->host_cap[sl] = want;
  return SB_OK;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct TreePlan {
  int nint = 0;
  std::vector<int> recs;  // nint * 8
  bool split = false;     // a subtree for a second wave is named (root record: r[6], r[7])
};

int plan_tree(const int32_t* parent, int R, TreePlan& tp, int split_min = 8) {
  if (R < 1) return fail(SB_ERR_TREE, "tree has %d nodes", R);
  if (parent[0] != -1) return fail(SB_ERR_TREE, "parentIndex[0] must be -1, got %d", parent[0]);
  std::vector<int> first(R, -1), second(R, -1);
  for (int n = 1; n < R; ++n) {
    const int p = parent[n];
    if (p < 0 || p >= n)
      return fail(SB_ERR_TREE, "parentIndex[%d]=%d is not a preorder parent (need 0 <= p < n)", n, p);
    if (first[p] < 0) first[p] = n;
    else if (second[p] < 0) second[p] = n;
    else return fail(SB_ERR_TREE, "Node %d has more than 2 children; tree must be binary", p);
  }
  for (int n = 0; n < R; ++n)
    if (first[n] >= 0 && second[n] < 0)
      return fail(SB_ERR_TREE, "Node %d has 1 child; expected 0 or 2", n);
  // Records (internal nodes) in depth-first preorder, first child's subtree
  // first: the walks rely on an internal first child's record following its
  // parent's.  parentIndex only guarantees parents before children
  // (docs/architecture.md:39-48), and the reference's own balanced trees
  // (parentIndex[i] = (i - 1) // 2, benchmarks/run_benchmarks.py:119-128)
  // are breadth-first; node indices, alignment rows and outputs keep the
  // caller's numbering, only the walk order changes.
  std::vector<int> slot(R, -1), stack;
  int nint = 0;
  stack.push_back(0);
  while (!stack.empty()) {
    const int n = stack.back();
    stack.pop_back();
    if (first[n] < 0) continue;
    slot[n] = nint++;
    stack.push_back(second[n]);
    stack.push_back(first[n]);
  }
  tp.nint = nint;
  tp.recs.assign(size_t(nint) * 8, 0);
  for (int n = 0; n < R; ++n) {
    if (slot[n] < 0) continue;
    int* r = &tp.recs[size_t(slot[n]) * 8];
    r[0] = n;
    r[1] = first[n];
    r[2] = second[n];
    r[4] = slot[first[n]];
    r[5] = slot[second[n]];
  }
  // Split walks (small-alphabet kernels, two waves per tile): the second wave
  // walks the subtree of records [s0, s1) while the first walks the rest, in
  // s0 + max(nint - s1, s1 - s0) records of latency.  The subtree chosen
  // minimises that; a tree that gains less than 15% keeps one wave
  // (s0 = s1 = 0; random 64-leaf trees gain 16-49%, 36% at the median, a
  // caterpillar nothing).  The one-wave kernels ignore it.
  if (nint >= std::max(split_min, 8)) {
    std::vector<int> sz(nint, 1);
    for (int i = nint - 1; i >= 0; --i) {
      const int* r = &tp.recs[size_t(i) * 8];
      if (r[4] >= 0) sz[i] += sz[r[4]];
      if (r[5] >= 0) sz[i] += sz[r[5]];
    }
    int best = nint, bs = 0;
    for (int v = 1; v < nint; ++v) {
      const int cost = v + std::max(nint - v - sz[v], sz[v]);
      if (cost < best) {
        best = cost;
        bs = v;
      }
    }
    if (20 * best <= 17 * nint) {
      tp.recs[6] = bs;
      tp.recs[7] = bs + sz[bs];
      tp.split = true;
    }
  }
  return SB_OK;
}

Model* get_model(sb_context* ctx, int id) {
  if (id < 0 || id >= int(ctx->models.size()) || !ctx->models[id].alive) return nullptr;
  return &ctx->models[id];
}

// True when submission f.seq of launcher f.serial has finished (or the
// launcher is gone: batch_free synchronises its stream first).  A launcher's
// submissions finish in order, so the event of any later one also answers.
bool fence_done(sb_context* ctx, const Fence& f) {
  const sb_batch* b = nullptr;
  for (const sb_batch* x : ctx->live)
    if (x->serial == f.serial) b = x;
  if (!b) return true;
  int best = -1;
  for (int k = 0; k < sb_batch::kRing; ++k)
    if (b->ring_ev[k] && b->ring_seq[k] >= f.seq && (best < 0 || b->ring_seq[k] < b->ring_seq[best])) best = k;
  if (best < 0) return false;   // no event recorded for it yet
  const hipError_t e = hipEventQuery(b->ring_ev[best]);
  if (e == hipSuccess) return true;
  if (e != hipErrorNotReady) (void)hipGetLastError();
  return false;
}

// Note that submission `seq` of launcher b may read model m.
void note_use(const Model* m, const sb_batch* b, uint64_t seq) {
  for (Fence& f : m->uses)
    if (f.serial == b->serial) {
      f.seq = seq;
      return;
    }
  m->uses.push_back({b->serial, seq});
}

// Reserve `bytes` (multiple of 256) in the model arena: a finished retired
// record of the same size if there is one, else fresh chunk space.
int arena_alloc(sb_context* ctx, size_t bytes, char** out) {
  auto bucket = ctx->retired.find(bytes);
  if (bucket != ctx->retired.end()) {
    std::list<Retired>& l = bucket->second;
    int checked = 0;
    for (auto it = l.begin(); it != l.end() && checked < 8; ++it, ++checked) {
      bool done = true;
      for (const Fence& f : it->uses) done = done && fence_done(ctx, f);
      if (!done) continue;
      *out = it->base;
      l.erase(it);
      --ctx->n_retired;
      return SB_OK;
    }
  }
  ctx->arena_reserved += bytes;
  if (!ctx->arena_cur || ctx->arena_off + bytes > ctx->arena_cap) {
    const size_t chunk = std::max<size_t>(64u << 20, bytes);
    void* p = nullptr;
    SB_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(&p, chunk) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SB_ERR_OOM, "model arena allocation of %zu bytes failed", chunk);
    }
    ctx->arena_chunks.push_back(p);
    ctx->arena_cur = static_cast<char*>(p);
    ctx->arena_off = 0;
    ctx->arena_cap = chunk;
  }
  *out = ctx->arena_cur + ctx->arena_off;
  ctx->arena_off += bytes;
  return SB_OK;
}

// Make pinned staging buffer k hold `need` bytes, keeping its first `keep`
// (buffer k is not in use by the GPU here).
int grow_pend(sb_context* ctx, int k, size_t need, size_t keep) {
  if (need <= ctx->pend_cap[k]) return SB_OK;
  size_t want = std::max<size_t>({need, ctx->pend_cap[k] * 2, size_t(1) << 20});
  char* nb = nullptr;
  SB_HIP(hipSetDevice(ctx->device));
  if (hipHostMalloc(reinterpret_cast<void**>(&nb), want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return fail(SB_ERR_OOM, "pinned model staging allocation failed");
  }
  if (ctx->pend[k]) {
    std::memcpy(nb, ctx->pend[k], keep);
    SB_HIP(hipHostFree(ctx->pend[k]));
  }
  ctx->pend[k] = nb;
  ctx->pend_cap[k] = want;
  return SB_OK;
}

// Append bytes destined for device address dst to the pending model staging.
int stage_model_bytes(sb_context* ctx, char* dst, const void* src, size_t len) {
  int k = ctx->pend_k;
  if (ctx->pend_busy[k]) {
    SB_HIP(hipEventSynchronize(ctx->pend_ev[k]));
    ctx->pend_busy[k] = false;
  }
  const int rc = grow_pend(ctx, k, ctx->pend_len + len, ctx->pend_len);
  if (rc) return rc;
  std::memcpy(ctx->pend[k] + ctx->pend_len, src, len);
  if (!ctx->runs.empty() && ctx->runs.back().dst + ctx->runs.back().len == dst &&
      ctx->runs.back().src + ctx->runs.back().len == ctx->pend_len)
    ctx->runs.back().len += len;
  else
    ctx->runs.push_back({dst, ctx->pend_len, len});
  ctx->pend_len += len;
  return SB_OK;
}

// Enqueue pending model uploads on stream s and make s wait for all uploads.
// A few runs are copied directly; more (new models scattered over reused
// arena records) go up as one copy of the staging buffer plus its run table,
// and one kernel moves each run into place.  With `may_alloc` false (the
// no-allocation launcher) only direct copies are used.
int flush_models(sb_context* ctx, hipStream_t s, bool may_alloc = true) {
  if (ctx->model_ev_valid) SB_HIP(hipStreamWaitEvent(s, ctx->model_ev, 0));
  if (!ctx->runs.empty()) {
    const int k = ctx->pend_k;
    const size_t nruns = ctx->runs.size();
    const size_t toff = align_up(ctx->pend_len, 256);
    const size_t total = toff + nruns * sizeof(SbScatterRun);
    bool scatter = may_alloc && nruns > 8;
    if (scatter && grow_pend(ctx, k, total, ctx->pend_len)) {
      (void)sb_last_error();
      scatter = false;
    }
    if (scatter && total > ctx->dstage_cap[k]) {
      // buffer k's last scatter has finished: stage_model_bytes waited for pend_ev[k]
      if (ctx->dstage[k]) SB_HIP(hipFree(ctx->dstage[k]));
      ctx->dstage[k] = nullptr;
      ctx->dstage_cap[k] = 0;
      const size_t want = std::max<size_t>(total * 2, size_t(1) << 20);
      if (hipMalloc(reinterpret_cast<void**>(&ctx->dstage[k]), want) == hipSuccess) {
        ctx->dstage_cap[k] = want;
      } else {
        (void)hipGetLastError();
        ctx->dstage[k] = nullptr;
        scatter = false;
      }
    }
    if (scatter) {
      SbScatterRun* t = reinterpret_cast<SbScatterRun*>(ctx->pend[k] + toff);
      for (size_t i = 0; i < nruns; ++i) {
        const ModelRun& r = ctx->runs[i];
        t[i] = {reinterpret_cast<unsigned long long>(r.dst), r.src, r.len};
      }
      SB_HIP(hipMemcpyAsync(ctx->dstage[k], ctx->pend[k], total, hipMemcpyHostToDevice, s));
      SB_HIP(sb_launch_scatter_runs(ctx->dstage[k], reinterpret_cast<const SbScatterRun*>(ctx->dstage[k] + toff),
                                    int(nruns), s));
    } else {
      for (const ModelRun& r : ctx->runs)
        SB_HIP(hipMemcpyAsync(r.dst, ctx->pend[k] + r.src, r.len, hipMemcpyHostToDevice, s));
    }
    if (!ctx->pend_ev[k]) SB_HIP(hipEventCreateWithFlags(&ctx->pend_ev[k], hipEventDisableTiming));
    if (!ctx->model_ev) SB_HIP(hipEventCreateWithFlags(&ctx->model_ev, hipEventDisableTiming));
    SB_HIP(hipEventRecord(ctx->pend_ev[k], s));
    SB_HIP(hipEventRecord(ctx->model_ev, s));
    ctx->model_ev_valid = true;
    ctx->pend_busy[k] = true;
    ctx->runs.clear();
    ctx->pend_len = 0;
    ctx->pend_k = (k + 1) % sb_context::kPend;
  }
  return SB_OK;
}

struct BlockIn {
  sb_block blk;
  int flags;  // SB_F_*
  const double* g;
  double* dist_grad;
  double* rate_grad;
  int n_rate;
  const double* rate_w_host;  // [n_rate][A*A]
  void* bcounts;              // per-branch outputs (sb_inside_outside)
  void* npost;
  void* bpost;
  const Model* const* rows;   // [R] per-branch models (row 0: the root distribution) or null
  const Model* root_model;    // root distribution of the passes if not the model's own (posteriors)
};

int get_graph(sb_batch* b, const GraphKey& key, hipGraphExec_t* out, int nwg_params, dim3 grid,
              dim3 block, int nwaves, int nwg_reduce);

int enqueue(sb_batch* b, const GraphKey& key, dim3 grid, dim3 block, int nwaves, int nwg_params,
            int nwg_reduce) {
  hipStream_t s = b->stream;
  SB_HIP(sb_launch_branch_params(key.dtype, key.irr != 0, nwg_params, b->header, s));  // also zeroes the tile counter
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (b->timing) {
    if (!b->tfree.empty()) {
      ev = b->tfree.back();
      b->tfree.pop_back();
    } else {
      SB_HIP(hipEventCreate(&ev.first));
      SB_HIP(hipEventCreate(&ev.second));
    }
    SB_HIP(hipEventRecord(ev.first, s));
  }
  if (key.large == 4) {
    SB_HIP(sb_launch_post(key.dtype, key.AP, key.f81, key.irr != 0, nwaves, b->header, s));
  } else if (key.large == 3 || key.large == 8) {
    const int branch = key.large == 8 ? 1 : 0, rmax = b->g2_rmax;
    const int wave1 = b->ctx->num_cus * sb_gen2_pass1_per_cu(key.AP);   // pass-1 workgroups held at once
    for (int t0 = 0; t0 < b->g2_ntiles; t0 += b->g2_round) {
      const int n = std::min(b->g2_round, b->g2_ntiles - t0);
      const int k = n / wave1 * wave1;   // tiles of whole pass-1 waves
      if (b->fixed || k == 0 || k == n) {
        SB_HIP(sb_launch_gen2(key.AP, t0, n, rmax, branch, b->header, s));
        continue;
      }
      // The last n - k pass-1 workgroups leave most CUs idle: pass 2 of the
      // first k tiles runs beside them on the second stream.  The round ends
      // when both streams are done (the next round reuses the stage slots).
      if (!b->aux) {
        SB_HIP(hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking));
        for (hipEvent_t& e : b->g2_ev) SB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      }
      SB_HIP(sb_launch_gen2_pass(key.AP, t0, 0, k, rmax, branch, 1, b->header, s));
      SB_HIP(hipEventRecord(b->g2_ev[0], s));
      SB_HIP(sb_launch_gen2_pass(key.AP, t0 + k, k, n - k, rmax, branch, 1, b->header, s));
      SB_HIP(hipStreamWaitEvent(b->aux, b->g2_ev[0], 0));
      SB_HIP(sb_launch_gen2_pass(key.AP, t0, 0, k, rmax, branch, 2, b->header, b->aux));
      SB_HIP(hipEventRecord(b->g2_ev[1], b->aux));
      SB_HIP(sb_launch_gen2_pass(key.AP, t0 + k, k, n - k, rmax, branch, 2, b->header, s));
      SB_HIP(hipStreamWaitEvent(s, b->g2_ev[1], 0));
    }
  } else if (key.large == 2)
    SB_HIP(sb_launch_f81w(key.AP, nwaves, b->header, s));
  else if (key.large)   // 1 counts, 5 posteriors, 6 per-branch counts, 7 irreversible counts
    SB_HIP(sb_launch_large(key.dtype, key.AP, key.f81, key.grad,
                           key.large == 5 ? 1 : key.large == 6 ? 2 : key.large == 7 ? 3 : 0, nwaves,
                           b->header, s));
  else
    SB_HIP(sb_launch_fused_small(key.dtype, key.AP, key.f81, key.grad, key.irr, key.rows != 0,
                                 key.split != 0, grid, block, key.lds, b->header, s));
  if (b->timing) {
    SB_HIP(hipEventRecord(ev.second, s));
    b->tev.push_back(ev);
  }
  if (key.grad) SB_HIP(sb_launch_grad_reduce(nwg_reduce, b->header, s));
  return SB_OK;
}

int get_graph(sb_batch* b, const GraphKey& key, hipGraphExec_t* out, int nwg_params, dim3 grid,
              dim3 block, int nwaves, int nwg_reduce) {
  for (auto& g : b->graphs)
    if (g.first == key) {
      *out = g.second;
      return SB_OK;
    }
  hipStream_t cap = nullptr;
  SB_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipStream_t saved = b->stream;
  const bool saved_timing = b->timing;
  b->stream = cap;
  b->timing = false;
  SB_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  int rc = enqueue(b, key, grid, block, nwaves, nwg_params, nwg_reduce);
  b->timing = saved_timing;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(cap, &graph);
  b->stream = saved;
  (void)hipStreamDestroy(cap);
  if (rc != SB_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(SB_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(SB_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
  b->graphs.push_back({key, exec});
  *out = exec;
  return SB_OK;
}

// Plan, upload and launch a list of blocks on the launcher's stream.
int submit(sb_batch* b, const std::vector<BlockIn>& in_, int dtype, int use_graph) {
  sb_context* ctx = b->ctx;
  if (dtype != SB_F32 && dtype != SB_F64) return fail(SB_ERR_ARG, "dtype must be SB_F32 or SB_F64");
  const size_t esz = dtype == SB_F32 ? 4 : 8;

  const std::vector<BlockIn>& in = in_;
  const int nb = int(in.size());
  if (nb == 0) return SB_OK;

  // ---- validate and plan every block
  int AP = -1;
  bool f81 = false, grad = false;
  std::vector<TreePlan> plans(nb);
  std::vector<const Model*> mods(nb);
  int nint_max = 0, R_max = 1;
  size_t nsched = 0, njobs = 0, ntiles = 0, nrate = 0, nrowm = 0;
  bool any_rows = false;
  for (int i = 0; i < nb; ++i) {
    const sb_block& k = in[i].blk;
    const Model* m = get_model(ctx, k.model_id);
    if (!m) return fail(SB_ERR_MODEL, "block %d: unknown model id %d", i, k.model_id);
    if (AP < 0) AP = m->AP;
    if (m->AP != AP)
      return fail(SB_ERR_UNSUPPORTED, "block %d: padded alphabet %d differs from %d in one batch", i,
                  m->AP, AP);
    if (k.R < 1 || k.C < 0) return fail(SB_ERR_ARG, "block %d: bad sizes R=%d C=%d", i, k.R, k.C);
    if (!k.parent_index || !k.distances) return fail(SB_ERR_ARG, "block %d: null tree arrays", i);
    if (k.C > 0 && !k.d_alignment) return fail(SB_ERR_ARG, "block %d: null alignment", i);
    for (int n = 0; n < k.R; ++n)
      if (!std::isfinite(k.distances[n]))
        return fail(SB_ERR_ARG, "block %d: distanceToParent[%d] is not finite", i, n);
    const int mm = k.mask_mode;
    if (mm < SB_MASK_NONE || mm > SB_MASK_EXPLICIT) return fail(SB_ERR_ARG, "block %d: bad mask_mode", i);
    if (mm == SB_MASK_EXPLICIT && !k.d_mask && k.C > 0)
      return fail(SB_ERR_ARG, "block %d: explicit mask without mask buffer", i);
    if (in[i].flags & SB_F_GRAD) grad = true;
    if (k.f81_fast) f81 = true;
    int rc = plan_tree(k.parent_index, k.R, plans[i], b->split_min);
    if (rc != SB_OK) {
      std::string msg = g_err;
      return fail(rc, "block %d: %s", i, msg.c_str());
    }
    mods[i] = m;
    if (in[i].rows) {
      for (int n = 0; n < k.R; ++n) {
        const Model* r = in[i].rows[n];
        if (!r) return fail(SB_ERR_MODEL, "block %d: missing model of branch %d", i, n);
        if (r->AP != AP || r->irr != m->irr)
          return fail(SB_ERR_UNSUPPORTED, "block %d: per-branch models must share the alphabet and "
                      "reversibility", i);
      }
      nrowm += k.R;
      any_rows = true;
    }
    if (const Model* rm = in[i].root_model) {
      if (rm->AP != AP || rm->irr != m->irr)
        return fail(SB_ERR_UNSUPPORTED, "block %d: the root model must share the alphabet and reversibility", i);
      if (in[i].flags & (SB_F_COUNTS | SB_F_BCOUNTS | SB_F_GRAD | SB_F_ROOT) || in[i].rows)
        return fail(SB_ERR_UNSUPPORTED, "block %d: a separate root distribution is for posteriors only", i);
    }
    nint_max = std::max(nint_max, plans[i].nint);
    R_max = std::max(R_max, k.R);
    nsched += plans[i].recs.size();
    njobs += k.R;
    (void)0;
    nrate += size_t(in[i].n_rate) * AP * AP;
  }
  for (int i = 0; i < nb; ++i) {
    if ((in[i].flags & SB_F_GRAD) == 0 && grad)
      return fail(SB_ERR_UNSUPPORTED, "gradient and non-gradient blocks cannot share a batch");
    if (bool(in[i].blk.f81_fast) != f81)
      return fail(SB_ERR_UNSUPPORTED, "F81 and eigensub blocks cannot share a batch");
  }
  if (grad && f81) return fail(SB_ERR_UNSUPPORTED, "gradient uses the eigensub path; f81_fast must be 0");
  const bool large = !is_small(AP);
  // per-branch outputs (branch counts, posteriors): their own small-alphabet
  // kernel; above A = 8 the large-alphabet kernel's posterior or per-branch
  // counts variant, one kind of output per launch
  // (above 8, Counts with per-branch models also run the per-branch counts
  // variant, summing the branches' back-transformed counts in the output)
  bool post = false, lpost = false, lbc = false, rows_cnt = false;
  for (int i = 0; i < nb; ++i) {
    if (in[i].flags & (SB_F_BCOUNTS | SB_F_NPOST | SB_F_BPOST)) post = true;
    if (in[i].rows && (in[i].flags & SB_F_COUNTS)) rows_cnt = true;
  }
  if (post && grad) return fail(SB_ERR_UNSUPPORTED, "per-branch outputs cannot be combined with the gradient");
  if ((post || rows_cnt) && large) {
    for (int i = 0; i < nb; ++i) {
      const int fl = in[i].flags;
      const bool cnt = (fl & SB_F_COUNTS) != 0;
      const bool p = (fl & (SB_F_NPOST | SB_F_BPOST)) != 0, bcn = (fl & SB_F_BCOUNTS) != 0 || (cnt && in[i].rows);
      if ((cnt && !in[i].rows) || (p && bcn) || (cnt && (fl & SB_F_BCOUNTS)))
        return fail(SB_ERR_UNSUPPORTED, "above an alphabet of 8, counts, branch counts and posteriors are "
                    "launched apart");
      lpost = lpost || p;
      lbc = lbc || bcn;
    }
    if (lpost && lbc)
      return fail(SB_ERR_UNSUPPORTED, "above an alphabet of 8, branch counts and posteriors are launched apart");
    post = false;
    if (lpost) f81 = false;   // posteriors do not depend on the counts path
  }
  const bool irr = mods[0]->irr;
  for (int i = 1; i < nb; ++i)
    if (mods[i]->irr != irr)
      return fail(SB_ERR_UNSUPPORTED, "irreversible and reversible models cannot share a batch");
  if (irr && f81) return fail(SB_ERR_UNSUPPORTED, "F81 fast path is reversible-only");
  // irreversible models above 8: counts, branch counts and gradients in the
  // large-alphabet kernel's complex variant
  bool lirr = false;
  if (irr && large) {
    for (int i = 0; i < nb; ++i)
      if (in[i].flags & (SB_F_COUNTS | SB_F_BCOUNTS | SB_F_GRAD)) lirr = true;
    if (lirr && lpost)
      return fail(SB_ERR_UNSUPPORTED, "above an alphabet of 8, counts and posteriors are launched apart");
    lbc = false;
  }
  // per-branch models: counts are back-transformed branch by branch (fused
  // ROWS variant; above 8 the large-alphabet per-branch counts variant)
  bool rowsk = false;
  if (any_rows) {
    if (grad || f81)
      return fail(SB_ERR_UNSUPPORTED, "per-branch models need no gradient and no F81 path");
    rowsk = rows_cnt && !large;
  }
  // F81 fast path in FP32 on F81-structured models: MFMA-accumulating kernel
  bool f81w = large && !lbc && f81 && dtype == SB_F32 && !grad;
  for (int i = 0; i < nb && f81w; ++i) f81w = mods[i]->f81s;
  // general models in FP32 (Counts, log-likelihood, root): two-pass matrix-core path
  // (not on the no-allocation launcher: its stage area would not fit); with
  // branch counts (one model per block) its second pass writes them per branch
  const bool gen2b = !b->fixed && large && lbc && !irr && !any_rows && !f81 && dtype == SB_F32 && !grad;
  const bool gen2 = gen2b ||
                    (!b->fixed && large && !lpost && !lbc && !irr && !any_rows && !f81 && dtype == SB_F32 && !grad);
  const int tile_cols = gen2    ? sb_gen2_tile_cols()
                        : f81w  ? sb_f81w_tile_cols()
                        : large ? sb_large_tile_cols(dtype, AP, f81, lpost ? 1 : lbc ? 2 : lirr ? 3 : 0)
                                : SB_WAVE;
  for (int i = 0; i < nb; ++i) ntiles += (size_t(in[i].blk.C) + tile_cols - 1) / tile_cols;

  // ---- host staging layout: header | blocks | sched | tiles | jobs | rate weights
  size_t off = 0;
  const size_t o_hdr = off;
  off = align_up(off + sizeof(SbLaunch), 256);
  const size_t o_blocks = off;
  off = align_up(off + sizeof(SbBlockDesc) * nb, 256);
  const size_t o_sched = off;
  off = align_up(off + sizeof(int) * nsched, 256);
  const size_t o_tiles = off;
  off = align_up(off + sizeof(int2) * ntiles, 256);
  const size_t o_jobs = off;
  off = align_up(off + sizeof(SbBranchJob) * njobs, 256);
  const size_t o_rate = off;
  off = align_up(off + sizeof(double) * nrate, 256);
  const size_t o_rowm = off;
  off = align_up(off + sizeof(void*) * nrowm, 256);
  const size_t total = off;

  const int sl = b->fixed ? 0 : b->pslot;   // this submission's staging buffer
  int rc = ensure_host(b, sl, total);
  if (rc) return rc;
  {
    void* before = b->plan.p;
    rc = ensure(b, b->plan, total);
    if (rc) return rc;
    if (b->plan.p != before) {
      // captured graphs embed the header address: drop them
      for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.second);
      b->graphs.clear();
      b->uploaded.clear();
      b->header = static_cast<SbLaunch*>(b->plan.p);
    }
  }
  const int bts = sb_bt_stride_k(AP, irr);
  rc = ensure(b, b->bt, std::max<size_t>(1, njobs) * bts * esz);
  if (rc) return rc;

  // Split walks: the small-alphabet kinds that can split (single model,
  // reversible or F81, four states) walk each tile with two waves when a tree
  // of the launch names a subtree for the second wave (tiles of the others
  // leave it idle).  Up to half the walk's latency, and twice the waves to
  // hide the latency of each (3,125 tiles of the 127-node workload: 0.243 ms
  // against 0.280 ms).  Whether a tile splits depends on its tree alone, so
  // a block's results do not depend on the rest of the batch.
  bool split = false;
  if (!large && !post && !irr && !rowsk && AP == 4 && b->split > 0)
    for (int i = 0; i < nb; ++i) split = split || plans[i].split;
  // persistent grid: resident waves per CU from the kernel's occupancy unless
  // the launcher was given an explicit value; never more waves than tiles
  b->last_split = split ? 1 : 0;
  const int wpg = (large || post) ? 1 : split ? 2 : b->waves_per_group;
  const int lds_tok = 0;   // (no dynamic LDS in the small-alphabet kernels)
  int wpc = b->waves_per_cu;
  if (wpc <= 0) {
    const int occ = post    ? sb_post_occupancy(dtype, AP, f81, irr)
                    : f81w  ? sb_f81w_occupancy(AP)
                    : large ? sb_large_occupancy(dtype, AP, f81, grad, lpost ? 1 : lbc ? 2 : lirr ? 3 : 0)
                            : sb_fused_small_occupancy(dtype, AP, f81, grad, irr, rowsk, split, 64 * wpg, lds_tok);
    wpc = occ > 0 ? occ * wpg : 8;
  }
  int nwaves = ctx->num_cus * wpc;
  {
    // small batches launch fewer waves (a power of two >= the tile count, so
    // few distinct grids and hence few captured graphs).  The small-alphabet
    // kernels (one wave walks one tile) launch at most the waves per CU their
    // tiles need, rounded up: idle waves beside them only let the dispatcher
    // stack four walking waves on some SIMDs while others hold two (3,125
    // tiles of the 127-node workload: 0.275 ms at 13 waves per CU, 0.289 ms
    // at 16)
    const long long wpt = f81w ? sb_f81w_waves_per_group() : split ? 2 : 1;   // waves per tile
    long long want = 64;
    while (want < (long long)ntiles * wpt) want *= 2;
    if (!large)
      want = std::min(want, ((long long)ntiles * wpt + ctx->num_cus - 1) / ctx->num_cus * ctx->num_cus);
    if (want < nwaves) nwaves = int(want);
  }
  nwaves = (nwaves + wpg - 1) / wpg * wpg;
  if (f81w) {
    const int g = sb_f81w_waves_per_group();
    nwaves = (nwaves + g - 1) / g * g;
  }
  long long wave_stride;
  if (f81w) wave_stride = sb_f81w_wave_stride(AP, nint_max);
  else if (large) wave_stride = sb_large_wave_stride(dtype, AP, nint_max, f81, lpost ? 1 : lbc ? 2 : lirr ? 3 : 0);
  else wave_stride = (long long)nint_max * (2 * AP * esz * SB_WAVE + 2 * 4 * SB_WAVE);
  wave_stride = (long long)align_up(size_t(std::max<long long>(wave_stride, 256)), 256);
  if (b->fixed && !gen2) {
    // no-allocation launcher: as many resident waves as its scratch slots hold
    const int g = f81w ? sb_f81w_waves_per_group() : wpg;
    const long long fit = (long long)(b->scratch.cap / size_t(wave_stride)) / g * g;
    if (fit < g)
      return fail(SB_ERR_ARG, "one wave's scratch slot (%lld bytes for %d internal nodes) exceeds the "
                  "no-allocation scratch area (%zu bytes)", wave_stride, nint_max, b->scratch.cap);
    if (nwaves > fit) nwaves = int(fit);
  }
  if (!gen2) {
    rc = ensure(b, b->scratch, size_t(wave_stride) * nwaves);
    if (rc) return rc;
  }
  long long g2_tile_bytes = 0;
  if (gen2) {
    g2_tile_bytes = sb_gen2_tile_bytes(AP, R_max);
    size_t cap = std::max<size_t>(1, kGen2StageBudget / size_t(g2_tile_bytes));
    if (b->gen2_round > 0) cap = std::min(cap, size_t(b->gen2_round));   // tuned round size
    const size_t nt = std::max<size_t>(ntiles, 1);
    const size_t rounds = (nt + cap - 1) / cap;
    b->g2_ntiles = int(ntiles);
    b->g2_round = int((nt + rounds - 1) / rounds);   // equal rounds
    b->g2_rmax = R_max;
    rc = ensure(b, b->stage, size_t(b->g2_round) * size_t(g2_tile_bytes));
    if (rc) return rc;
  }
  const int grad_stride = R_max;
  if (grad) {
    rc = ensure(b, b->part, std::max<size_t>(1, ntiles) * (grad_stride + SB_MAX_RATE) * sizeof(double));
    if (rc) return rc;
  }

  char* h = b->host[sl];
  char* dplan = static_cast<char*>(b->plan.p);
  auto dev = [&](size_t o) { return dplan + o; };
  SbBlockDesc* hb = reinterpret_cast<SbBlockDesc*>(h + o_blocks);
  int* hs = reinterpret_cast<int*>(h + o_sched);
  int2* ht = reinterpret_cast<int2*>(h + o_tiles);
  SbBranchJob* hj = reinterpret_cast<SbBranchJob*>(h + o_jobs);
  double* hr = reinterpret_cast<double*>(h + o_rate);
  const void** hm = reinterpret_cast<const void**>(h + o_rowm);

  size_t ps = 0, pj = 0, pr = 0, tb = 0, pm = 0;
  std::vector<std::pair<int, int>> order;  // (nint, block) for tile ordering
  for (int i = 0; i < nb; ++i) {
    const sb_block& k = in[i].blk;
    const Model* m = mods[i];
    SbBlockDesc& d = hb[i];
    std::memset(&d, 0, sizeof d);
    d.aln = k.d_alignment;
    d.mask = k.d_mask;
    d.counts = k.d_counts;
    d.root = k.d_root;
    d.loglike = k.d_loglike;
    d.sched = reinterpret_cast<const int*>(dev(o_sched) + ps * sizeof(int));
    d.bt = static_cast<char*>(b->bt.p) + pj * bts * esz;
    d.model = m->rec[dtype];
    d.model64 = m->d64;
    d.g = in[i].g;
    d.dist_grad = in[i].dist_grad;
    d.rate_grad = in[i].rate_grad;
    d.n_rate = in[i].n_rate;
    d.bcounts = in[i].bcounts;
    d.npost = in[i].npost;
    d.bpost = in[i].bpost;
    if (in[i].root_model) {
      d.root_model = in[i].root_model->rec[dtype];
      d.root_model64 = in[i].root_model->d64;
    }
    d.rate_w = reinterpret_cast<const double*>(dev(o_rate) + pr * sizeof(double));
    d.R = k.R;
    d.C = k.C;
    d.nint = plans[i].nint;
    int fl = in[i].flags;
    if (k.f81_fast) fl |= SB_F_F81;
    if (m->irr) fl |= SB_F_IRR;
    fl |= (k.mask_mode & 3) << SB_F_MASK_SHIFT;
    d.flags = fl;
    d.A = m->A;
    d.tile_base = int(tb);
    d.ntiles = int((size_t(k.C) + tile_cols - 1) / tile_cols);
    std::memcpy(hs + ps, plans[i].recs.data(), plans[i].recs.size() * sizeof(int));
    ps += plans[i].recs.size();
    for (int n = 0; n < k.R; ++n) {
      hj[pj + n].model64 = in[i].rows ? in[i].rows[n]->d64 : m->d64;
      hj[pj + n].t = k.distances[n];
    }
    pj += k.R;
    if (in[i].rows) {
      d.row_models = reinterpret_cast<const void* const*>(dev(o_rowm) + pm * sizeof(void*));
      for (int n = 0; n < k.R; ++n) hm[pm + n] = in[i].rows[n]->rec[dtype];
      pm += k.R;
    }
    for (int q = 0; q < in[i].n_rate; ++q)
      for (int a = 0; a < AP; ++a)
        for (int c = 0; c < AP; ++c)
          hr[pr + size_t(q) * AP * AP + a * AP + c] =
              (a < m->A && c < m->A) ? in[i].rate_w_host[size_t(q) * m->A * m->A + a * m->A + c] : 0.0;
    pr += size_t(in[i].n_rate) * AP * AP;
    tb += d.ntiles;
    order.push_back({plans[i].nint, i});
  }
  // tiles: most expensive trees first (longest-processing-time order)
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<int, int>& x, const std::pair<int, int>& y) { return x.first > y.first; });
  size_t pt = 0;
  for (auto& o : order) {
    const int i = o.second;
    for (int t = 0; t < hb[i].ntiles; ++t) ht[pt++] = make_int2(i, t);
  }

  SbLaunch* L = reinterpret_cast<SbLaunch*>(h + o_hdr);
  std::memset(L, 0, sizeof *L);
  L->blocks = reinterpret_cast<const SbBlockDesc*>(dev(o_blocks));
  L->tiles = reinterpret_cast<const int2*>(dev(o_tiles));
  L->jobs = reinterpret_cast<const SbBranchJob*>(dev(o_jobs));
  L->bt = b->bt.p;
  L->counter = b->counter;
  L->scratch = static_cast<char*>(b->scratch.p);
  L->grad_part = grad ? static_cast<double*>(b->part.p) : nullptr;
  L->rate_part = grad ? static_cast<double*>(b->part.p) + ntiles * grad_stride : nullptr;
  L->stage = b->stage.p;
  L->stage_stride = g2_tile_bytes;
  L->rmax = R_max;
  L->wave_stride = wave_stride;
  L->nblocks = nb;
  L->ntiles = int(ntiles);
  L->njobs = int(njobs);
  L->nint_max = nint_max;
  L->grad_stride = grad_stride;
  L->AP = AP;
  L->tile_cols = tile_cols;

  // every model this launch reads is fenced by its submission number
  const uint64_t seq = b->seq + 1;
  for (int i = 0; i < nb; ++i) {
    note_use(mods[i], b, seq);
    if (in[i].rows)
      for (int n = 0; n < in[i].blk.R; ++n) note_use(in[i].rows[n], b, seq);
    if (in[i].root_model) note_use(in[i].root_model, b, seq);
  }

  hipStream_t s = b->stream;
  rc = flush_models(ctx, s, !b->fixed);
  if (rc) return rc;
  {
    // (a small plan equal to the one the device holds, e.g. the same batch
    // launched again, is not sent again: no copy between the launches)
    std::vector<char>& up = b->uploaded;
    const bool cached = !b->fixed && total <= sb_batch::kPlanCache;
    if (!cached || up.size() != total || std::memcmp(up.data(), h, total) != 0) {
      SB_HIP(hipMemcpyAsync(b->plan.p, h, total, hipMemcpyHostToDevice, s));
      SB_HIP(hipEventRecord(b->staged[sl], s));
      b->staged_pending[sl] = true;
      if (cached) up.assign(h, h + total);
      else up.clear();
    }
  }

  GraphKey key{dtype, AP, f81 ? 1 : 0, grad ? 1 : 0,
               gen2b ? 8 : lpost ? 5 : lbc ? 6 : lirr ? 7 : post ? 4 : gen2 ? 3 : f81w ? 2 : large ? 1 : 0,
               nwaves, wpg, irr ? 1 : 0, rowsk ? 1 : 0, lds_tok, split ? 1 : 0};
  const dim3 block(64 * wpg);
  const dim3 grid(nwaves / wpg);
  const int nwg_params = 1024;
  const int nwg_reduce = 64;
  if (use_graph && !gen2) {   // the two-pass path launches per round, uncaptured
    hipGraphExec_t exec = nullptr;
    rc = get_graph(b, key, &exec, nwg_params, grid, block, nwaves, nwg_reduce);
    if (rc) return rc;
    SB_HIP(hipGraphLaunch(exec, s));
  } else {
    rc = enqueue(b, key, grid, block, nwaves, nwg_params, nwg_reduce);
    if (rc) return rc;
  }
  {
    // one event marks the submission: the ring's, which is also "last
    // submission finished" (each queued event record costs the GPU time
    // between launches)
    const int k = int(seq % sb_batch::kRing);
    SB_HIP(hipEventRecord(b->ring_ev[k], s));
    b->ring_seq[k] = seq;
    b->seq = seq;
    b->done = b->ring_ev[k];
    if (!b->fixed) b->pslot = sl ^ 1;
  }
  return SB_OK;
}

int batch_create(sb_context* ctx, hipStream_t stream, sb_batch** out) {
  sb_batch* b = new sb_batch();
  b->ctx = ctx;
  b->stream = stream;
  if (hipMalloc(reinterpret_cast<void**>(&b->counter), 256) != hipSuccess) {
    (void)hipGetLastError();
    delete b;
    return fail(SB_ERR_OOM, "launcher allocation failed");
  }
  for (hipEvent_t& e : b->staged) SB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  SB_HIP(hipEventCreateWithFlags(&b->done_own, hipEventDisableTiming));
  b->done = b->done_own;
  for (hipEvent_t& e : b->ring_ev) SB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  b->serial = ++ctx->next_serial;
  ctx->live.push_back(b);
  *out = b;
  return SB_OK;
}

void batch_free(sb_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->stream);
  auto& live = b->ctx->live;
  live.erase(std::remove(live.begin(), live.end(), b), live.end());
  for (hipEvent_t e : b->ring_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->g2_ev)
    if (e) (void)hipEventDestroy(e);
  if (b->aux) (void)hipStreamDestroy(b->aux);
  for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.second);
  for (auto* v : {&b->tev, &b->tfree})
    for (auto& e : *v) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  if (!b->fixed) {
    for (DevBuf* d : {&b->plan, &b->bt, &b->scratch, &b->part, &b->stage, &b->xaln, &b->xmask, &b->xout,
                      &b->xcols})
      if (d->p) (void)hipFree(d->p);
    if (b->counter) (void)hipFree(b->counter);
  }
  if (b->xev) (void)hipEventDestroy(b->xev);
  for (int k = 0; k < 2; ++k) {
    if (b->host[k]) (void)hipHostFree(b->host[k]);
    if (b->staged[k]) (void)hipEventDestroy(b->staged[k]);
  }
  if (b->done_own) (void)hipEventDestroy(b->done_own);
  delete b;
}

// The no-allocation launcher (set-up, from sb_create): its plan, tile
// counters, branch records and wave scratch slots are fixed parts of this
// GPU's static scratch array (whose device address the scratch source file
// looks up with hipGetSymbolAddress), and its pinned plan staging is
// allocated here, so a submission on it allocates nothing.
int noalloc_create(sb_context* ctx) {
  void* base = nullptr;
  int rc = sb_internal_noalloc_scratch(ctx->device, &base);
  if (rc) return rc;
  sb_batch* b = new sb_batch();
  b->ctx = ctx;
  b->fixed = true;
  char* p = static_cast<char*>(base);
  b->plan.p = p + SB_NA_PLAN_OFF;
  b->plan.cap = SB_NA_PLAN_BYTES;
  b->header = static_cast<SbLaunch*>(b->plan.p);
  b->counter = reinterpret_cast<int*>(p + SB_NA_COUNTER_OFF);
  b->bt.p = p + SB_NA_BT_OFF;
  b->bt.cap = SB_NA_BT_BYTES;
  b->scratch.p = p + SB_NA_WAVES_OFF;
  b->scratch.cap = SB_NA_WAVES_BYTES;
  if (hipHostMalloc(reinterpret_cast<void**>(&b->host[0]), SB_NA_PLAN_BYTES, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    delete b;
    return fail(SB_ERR_OOM, "pinned plan staging of the no-allocation launcher failed");
  }
  b->host_cap[0] = SB_NA_PLAN_BYTES;
  hipError_t e = hipEventCreateWithFlags(&b->staged[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done_own, hipEventDisableTiming);
  b->done = b->done_own;
  for (hipEvent_t& ev : b->ring_ev)
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  b->serial = ++ctx->next_serial;
  ctx->live.push_back(b);
  if (e != hipSuccess) {
    batch_free(b);
    return fail(SB_ERR_HIP, "no-allocation launcher events: %s", hipGetErrorString(e));
  }
  ctx->noalloc = b;
  return SB_OK;
}

// Single-block calls go through the context's own launcher; if the caller
// switches streams, the new stream first waits for the launcher's last work.
int single(sb_context* ctx, const BlockIn& bi, int dtype, void* stream) {
  if (!ctx) return fail(SB_ERR_ARG, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  SB_HIP(hipSetDevice(ctx->device));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!ctx->internal) {
    int rc = batch_create(ctx, s, &ctx->internal);
    if (rc) return rc;
  }
  sb_batch* b = ctx->internal;
  if (b->stream != s) {
    SB_HIP(hipStreamWaitEvent(s, b->done, 0));
    b->stream = s;
  }
  std::vector<BlockIn> v(1, bi);
  return submit(b, v, dtype, 0);
}

BlockIn make_block(int model_id, const int8_t* aln, const int32_t* parent, const double* dist, int R,
                   int C) {
  BlockIn bi;
  std::memset(&bi, 0, sizeof bi);
  bi.blk.model_id = model_id;
  bi.blk.R = R;
  bi.blk.C = C;
  bi.blk.d_alignment = aln;
  bi.blk.parent_index = parent;
  bi.blk.distances = dist;
  return bi;
}

// ---- calls with several models (model lists, per-branch models, grids, mixtures)

struct MultiOut {
  void* counts = nullptr;
  void* root = nullptr;
  double* ll = nullptr;
  void* bcounts = nullptr;
  void* npost = nullptr;   // posteriors: passes under the first column model's root distribution
  void* bpost = nullptr;
};

// Launcher of a context for a call on `stream` (the context lock is held).
int context_launcher(sb_context* ctx, void* stream, sb_batch** out) {
  SB_HIP(hipSetDevice(ctx->device));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!ctx->internal) {
    int rc = batch_create(ctx, s, &ctx->internal);
    if (rc) return rc;
  }
  sb_batch* b = ctx->internal;
  if (b->stream != s) {
    SB_HIP(hipStreamWaitEvent(s, b->done, 0));
    b->stream = s;
  }
  *out = b;
  return SB_OK;
}

// Host side of the column maps and per-group tables, uploaded to b->xcols.
int upload_cols(sb_batch* b, const std::vector<int>& ints, const std::vector<double>& dbl,
                size_t* dbl_off) {
  if (b->xpending) {
    SB_HIP(hipEventSynchronize(b->xev));
    b->xpending = false;
  }
  const size_t ib = align_up(ints.size() * sizeof(int), 256);
  const size_t bytes = ib + dbl.size() * sizeof(double);
  int rc = ensure(b, b->xcols, std::max<size_t>(bytes, 256));
  if (rc) return rc;
  b->xhost.resize(bytes);
  std::memcpy(b->xhost.data(), ints.data(), ints.size() * sizeof(int));
  if (!dbl.empty()) std::memcpy(b->xhost.data() + ib, dbl.data(), dbl.size() * sizeof(double));
  SB_HIP(hipMemcpyAsync(b->xcols.p, b->xhost.data(), bytes, hipMemcpyHostToDevice, b->stream));
  if (!b->xev) SB_HIP(hipEventCreateWithFlags(&b->xev, hipEventDisableTiming));
  SB_HIP(hipEventRecord(b->xev, b->stream));
  b->xpending = true;
  *dbl_off = ib;
  return SB_OK;
}

// Model table ids[n_cols][n_rows] (n_cols in {1, C}, n_rows in {1, R}):
//   (1, 1) one model; (C, 1) one model per column (model list);
//   (1, R) one model per branch (row 0: root distribution); (C, R) both.
// Columns with the same model row form one block; with more than one group the
// columns are gathered into staging blocks and the outputs scattered back.
int multi(sb_context* ctx, int n_cols, int n_rows, const int* ids, const int8_t* aln,
          const int32_t* parent, const double* dist, int R, int C, int dtype, int f81, int mask_mode,
          const uint8_t* d_mask, const MultiOut& o, void* stream) {
  if (!ctx || !ids || !parent || !dist) return fail(SB_ERR_ARG, "null argument");
  if (R < 1 || C < 0) return fail(SB_ERR_ARG, "bad sizes R=%d C=%d", R, C);
  if (!(n_cols == 1 || n_cols == C) || !(n_rows == 1 || n_rows == R))
    return fail(SB_ERR_ARG, "model table must be (1 or C) x (1 or R), got %d x %d", n_cols, n_rows);
  if (dtype != SB_F32 && dtype != SB_F64) return fail(SB_ERR_ARG, "dtype must be SB_F32 or SB_F64");
  std::lock_guard<std::mutex> lock(ctx->mu);
  std::vector<const Model*> tab(size_t(n_cols) * n_rows);
  for (size_t i = 0; i < tab.size(); ++i) {
    tab[i] = get_model(ctx, ids[i]);
    if (!tab[i]) return fail(SB_ERR_MODEL, "unknown model id %d (table entry %zu)", ids[i], i);
    if (tab[i]->A != tab[0]->A || tab[i]->irr != tab[0]->irr)
      return fail(SB_ERR_UNSUPPORTED, "all models of one call must share the alphabet and reversibility");
  }
  const int A = tab[0]->A;
  sb_batch* b = nullptr;
  int rc = context_launcher(ctx, stream, &b);
  if (rc) return rc;
  const int flags = (o.counts ? SB_F_COUNTS : 0) | (o.root ? SB_F_ROOT : 0) | (o.ll ? SB_F_LL : 0) |
                    (o.bcounts ? SB_F_BCOUNTS : 0) | (o.npost ? SB_F_NPOST : 0) | (o.bpost ? SB_F_BPOST : 0);
  if (!flags && C > 0) return fail(SB_ERR_ARG, "no output requested");
  const bool per_branch = n_rows == R && R > 1;
  const bool posts = o.npost || o.bpost;
  if (posts && (o.counts || o.root || o.bcounts))
    return fail(SB_ERR_ARG, "posteriors of model lists are requested apart from counts and RootProb");
  if (posts && n_rows != 1)
    return fail(SB_ERR_UNSUPPORTED, "posteriors take one model per column (model grids are not supported, "
                "as in the reference)");

  // ---- group columns by model row (first appearance order)
  std::vector<int> grp_of(C, 0), first_col;
  if (n_cols == 1 || C == 0) {
    first_col.push_back(0);
  } else {
    std::map<std::vector<int>, int> key;
    for (int c = 0; c < C; ++c) {
      std::vector<int> k(ids + size_t(c) * n_rows, ids + size_t(c + 1) * n_rows);
      auto it = key.find(k);
      if (it == key.end()) {
        it = key.emplace(k, int(first_col.size())).first;
        first_col.push_back(c);
      }
      grp_of[c] = it->second;
    }
  }
  const int G = int(first_col.size());
  auto row_of = [&](int g) { return &tab[size_t(n_cols == 1 ? 0 : first_col[g]) * n_rows]; };
  auto id_of = [&](int g) { return ids[size_t(n_cols == 1 ? 0 : first_col[g]) * n_rows]; };
  auto block_for = [&](int g) {
    BlockIn bi = make_block(id_of(g), aln, parent, dist, R, C);
    bi.flags = flags;
    bi.blk.f81_fast = f81 ? 1 : 0;
    bi.blk.mask_mode = mask_mode;
    bi.blk.d_mask = d_mask;
    bi.rows = per_branch ? row_of(g) : nullptr;
    // posteriors of a model list: passes under the first column model's root
    // distribution (the reference's InsideOutside over a list)
    bi.root_model = posts && row_of(g)[0] != tab[0] ? tab[0] : nullptr;
    return bi;
  };
  if (G == 1) {   // one model row for every column: straight on the caller's arrays
    BlockIn bi = block_for(0);
    bi.blk.d_counts = o.counts;
    bi.blk.d_root = o.root;
    bi.blk.d_loglike = o.ll;
    bi.bcounts = o.bcounts;
    bi.npost = o.npost;
    bi.bpost = o.bpost;
    std::vector<BlockIn> v(1, bi);
    return submit(b, v, dtype, 0);
  }

  // ---- model lists: LogLike / RootProb take the root distribution of the first
  // column's model (the reference stacks the models and passes stacked.pi[0])
  std::vector<int> swap(G, 0);
  std::vector<double> w(size_t(G) * A, 1.0);
  bool any_swap = false;
  if (!per_branch && !posts && (o.ll || o.root)) {
    const Model* m0 = tab[0];
    for (int g = 0; g < G; ++g) {
      const Model* mg = row_of(g)[0];
      for (int a = 0; a < A; ++a) {
        w[size_t(g) * A + a] = m0->pi[a] / mg->pi[a];
        if (m0->pi[a] != mg->pi[a]) swap[g] = 1;
      }
      any_swap = any_swap || swap[g];
    }
  }

  // ---- column maps: permuted order = groups one after another
  std::vector<int> gsize(G, 0), gstart(G, 0);
  for (int c = 0; c < C; ++c) ++gsize[grp_of[c]];
  for (int g = 1; g < G; ++g) gstart[g] = gstart[g - 1] + gsize[g - 1];
  std::vector<int> ints(size_t(2) * C + 3 * G);
  int* perm = ints.data();
  int* grp = perm + C;
  {
    std::vector<int> fill(gstart);
    for (int c = 0; c < C; ++c) {
      const int j = fill[grp_of[c]]++;
      perm[j] = c;
      grp[j] = grp_of[c];
    }
  }
  std::memcpy(grp + C, gstart.data(), G * sizeof(int));
  std::memcpy(grp + C + G, gsize.data(), G * sizeof(int));
  std::memcpy(grp + C + 2 * G, swap.data(), G * sizeof(int));
  size_t w_off = 0;
  rc = upload_cols(b, ints, any_swap ? w : std::vector<double>(), &w_off);
  if (rc) return rc;
  const int* dcols = static_cast<const int*>(b->xcols.p);
  const int* cols[4] = {dcols, dcols + C, dcols + 2 * C, dcols + 2 * C + G};
  const int* dswap = dcols + 2 * C + 2 * G;
  const double* dw = reinterpret_cast<const double*>(static_cast<char*>(b->xcols.p) + w_off);

  // ---- gather the alignment (and an explicit mask) into per-group blocks
  hipStream_t s = b->stream;
  rc = ensure(b, b->xaln, std::max<size_t>(size_t(R) * C, 256));
  if (rc) return rc;
  SB_HIP(sb_launch_gather_cols(aln, b->xaln.p, 1, R, C, cols, s));
  const bool emask = mask_mode == SB_MASK_EXPLICIT;
  if (emask) {
    if (!d_mask) return fail(SB_ERR_ARG, "explicit mask without mask buffer");
    rc = ensure(b, b->xmask, std::max<size_t>(size_t(R) * C, 256));
    if (rc) return rc;
    SB_HIP(sb_launch_gather_cols(d_mask, b->xmask.p, 1, R, C, cols, s));
  }
  // ---- staged outputs
  const size_t esz = dtype == SB_F32 ? 4 : 8;
  const bool want_root = o.root || any_swap;
  size_t off = 0;
  auto carve = [&](bool on, size_t bytes) {
    const size_t at = off;
    if (on) off = align_up(off + bytes, 256);
    return at;
  };
  const size_t x_cnt = carve(o.counts, size_t(A) * A * C * esz);
  const size_t x_root = carve(want_root, size_t(A) * C * esz);
  const size_t x_ll = carve(o.ll, size_t(C) * 8);
  const size_t x_bc = carve(o.bcounts, size_t(R) * A * A * C * esz);
  const size_t x_np = carve(o.npost, size_t(R) * A * C * esz);
  const size_t x_bp = carve(o.bpost, size_t(R) * A * A * C * esz);
  rc = ensure(b, b->xout, std::max<size_t>(off, 256));
  if (rc) return rc;
  char* X = static_cast<char*>(b->xout.p);
  std::vector<BlockIn> blocks(G);
  for (int g = 0; g < G; ++g) {
    BlockIn bi = block_for(g);
    const size_t st = size_t(gstart[g]);
    bi.flags |= any_swap && !o.root ? SB_F_ROOT : 0;
    bi.blk.C = gsize[g];
    bi.blk.d_alignment = static_cast<const int8_t*>(b->xaln.p) + size_t(R) * st;
    bi.blk.d_mask = emask ? static_cast<const uint8_t*>(b->xmask.p) + size_t(R) * st : nullptr;
    bi.blk.d_counts = o.counts ? X + x_cnt + size_t(A) * A * st * esz : nullptr;
    bi.blk.d_root = want_root ? X + x_root + size_t(A) * st * esz : nullptr;
    bi.blk.d_loglike = o.ll ? reinterpret_cast<double*>(X + x_ll) + st : nullptr;
    bi.bcounts = o.bcounts ? X + x_bc + size_t(R) * A * A * st * esz : nullptr;
    bi.npost = o.npost ? X + x_np + size_t(R) * A * st * esz : nullptr;
    bi.bpost = o.bpost ? X + x_bp + size_t(R) * A * A * st * esz : nullptr;
    blocks[g] = bi;
  }
  rc = submit(b, blocks, dtype, 0);
  if (rc) return rc;
  if (any_swap)
    SB_HIP(sb_launch_prior_swap(o.ll ? reinterpret_cast<double*>(X + x_ll) : nullptr, X + x_root, dtype, A, C,
                                cols, dw, dswap, s));
  if (o.counts) SB_HIP(sb_launch_scatter_cols(X + x_cnt, o.counts, int(esz), A * A, C, cols, s));
  if (o.root) SB_HIP(sb_launch_scatter_cols(X + x_root, o.root, int(esz), A, C, cols, s));
  if (o.ll) SB_HIP(sb_launch_scatter_cols(X + x_ll, o.ll, 8, 1, C, cols, s));
  if (o.bcounts) SB_HIP(sb_launch_scatter_cols(X + x_bc, o.bcounts, int(esz), R * A * A, C, cols, s));
  if (o.npost) SB_HIP(sb_launch_scatter_cols(X + x_np, o.npost, int(esz), R * A, C, cols, s));
  if (o.bpost) SB_HIP(sb_launch_scatter_cols(X + x_bp, o.bpost, int(esz), R * A * A, C, cols, s));
  SB_HIP(hipEventRecord(b->done, s));
  return SB_OK;
}

// A reversible model's device record and table entry, built on the host
// (validation and host math: no context state, so callers may run it on
// several threads); commit_model then places it under the context lock.
struct PreparedModel {
  Model m;
  std::vector<char> blob;
};

int prepare_model(int A, const double* eigenvalues, const double* eigenvectors, const double* pi,
                  PreparedModel* out) {
  if (!eigenvalues || !eigenvectors || !pi) return fail(SB_ERR_ARG, "null argument");
  const int AP = pad_alphabet(A);
  if (A < 2 || AP < 0) return fail(SB_ERR_ALPHABET, "alphabet size %d outside 2..64", A);
  double psum = 0.0;
  for (int a = 0; a < A; ++a) {
    if (!(pi[a] > 0.0) || !std::isfinite(pi[a])) return fail(SB_ERR_MODEL, "pi[%d]=%g must be > 0", a, pi[a]);
    if (!std::isfinite(eigenvalues[a])) return fail(SB_ERR_MODEL, "eigenvalue %d not finite", a);
    psum += pi[a];
  }
  for (int i = 0; i < A * A; ++i)
    if (!std::isfinite(eigenvectors[i])) return fail(SB_ERR_MODEL, "eigenvector entry %d not finite", i);
  if (!(psum < 1.0 + 1e-6 && psum > 1.0 - 1e-6)) return fail(SB_ERR_MODEL, "pi sums to %g, not 1", psum);
  // (host math below runs outside the context lock; only the arena and the
  // model table are shared)

  // FP64 master with inert pads: mu=0, V=identity block, pi=1 (pads only enter
  // M through sqrt(pi_j/pi_i) on the identity block).
  std::vector<double> m64(sb_m64_size(AP), 0.0);
  double f81 = 0.0;
  bool f81s = false;
  for (int a = 0; a < A; ++a) f81 += pi[a] * pi[a];
  f81 = 1.0 / (1.0 - f81);
  for (int k = 0; k < AP; ++k) m64[sb_m64_mu(AP) + k] = k < A ? eigenvalues[k] : 0.0;
  for (int a = 0; a < AP; ++a)
    for (int k = 0; k < AP; ++k)
      m64[sb_m64_V(AP) + a * AP + k] =
          (a < A && k < A) ? eigenvectors[a * A + k] : (a == k ? 1.0 : 0.0);
  for (int a = 0; a < AP; ++a) m64[sb_m64_pi(AP) + a] = a < A ? pi[a] : 1.0;
  m64[sb_m64_misc(AP)] = f81;
  m64[sb_m64_misc(AP) + 1] = A;

  std::vector<double> rec(sb_mr_size(AP), 0.0);
  const double* V = &m64[sb_m64_V(AP)];
  for (int a = 0; a < AP; ++a) {
    const double sp = a < A ? std::sqrt(pi[a]) : 1.0;
    for (int k = 0; k < AP; ++k) {
      rec[sb_mr_V(AP) + a * AP + k] = V[a * AP + k];
      rec[sb_mr_Vs(AP) + a * AP + k] = sp * V[a * AP + k];
      rec[sb_mr_Vi(AP) + a * AP + k] = V[a * AP + k] / sp;
    }
  }
  for (int i = 0; i < AP; ++i)   // S is symmetric; stored exactly so (kernels read S[j][i] for S[i][j])
    for (int j = i; j < AP; ++j) {
      double s = 0.0;
      for (int k = 0; k < AP; ++k) s += V[i * AP + k] * m64[sb_m64_mu(AP) + k] * V[j * AP + k];
      rec[sb_mr_S(AP) + i * AP + j] = rec[sb_mr_S(AP) + j * AP + i] = (i < A && j < A) ? s : 0.0;
    }
  for (int a = 0; a < AP; ++a) {
    rec[sb_mr_pi(AP) + a] = a < A ? pi[a] : 0.0;
    rec[sb_mr_ones(AP) + a] = a < A ? 1.0 : 0.0;
  }
  rec[sb_mr_misc(AP)] = f81;
  // F81 structure: one zero eigenvalue with eigenvector +-sqrt(pi), all others
  // equal to lambda.  Then M(t) = e^{lambda t} I + (1 - e^{lambda t}) 1 pi^T,
  // which the large-alphabet kernel applies in O(A).
  {
    int k0 = -1, ks = -1;
    double lam = 0.0, scale = 0.0;
    for (int k = 0; k < A; ++k) scale = std::max(scale, std::fabs(eigenvalues[k]));
    bool ok = scale > 0.0;
    for (int k = 0; k < A && ok; ++k) {
      if (std::fabs(eigenvalues[k]) <= 1e-12 * scale) {
        if (k0 >= 0) ok = false;
        k0 = k;
      } else if (ks < 0) {
        ks = k;
        lam = eigenvalues[k];
      } else if (std::fabs(eigenvalues[k] - lam) > 1e-12 * scale) {
        ok = false;
      }
    }
    if (ok && k0 >= 0 && ks >= 0) {
      double dot = 0.0;
      for (int a = 0; a < A; ++a) dot += eigenvectors[a * A + k0] * std::sqrt(pi[a]);
      ok = std::fabs(std::fabs(dot) - 1.0) < 1e-10;
    } else {
      ok = false;
    }
    rec[sb_mr_misc(AP) + 1] = ok ? 1.0 : 0.0;
    f81s = ok;
    rec[sb_mr_misc(AP) + 2] = ok ? double(ks) : 0.0;
  }
  std::vector<float> recf(rec.begin(), rec.end());

  Model& m = out->m;
  m.alive = true;
  m.A = A;
  m.AP = AP;
  m.f81s = f81s;
  m.mu.assign(eigenvalues, eigenvalues + A);
  m.V.assign(eigenvectors, eigenvectors + A * A);
  m.pi.assign(pi, pi + A);
  const size_t b64 = align_up(m64.size() * 8, 256), bf = align_up(recf.size() * 4, 256),
               bd = align_up(rec.size() * 8, 256);
  std::vector<char>& blob = out->blob;
  blob.assign(b64 + bf + bd, 0);
  std::memcpy(blob.data(), m64.data(), m64.size() * 8);
  std::memcpy(blob.data() + b64, recf.data(), recf.size() * 4);
  std::memcpy(blob.data() + b64 + bf, rec.data(), rec.size() * 8);

  return SB_OK;
}

// Place a prepared model: arena record, staged upload, table entry (the
// caller holds ctx->mu).
int commit_model(sb_context* ctx, PreparedModel& pm, int* model_id) {
  Model& m = pm.m;
  char* base = nullptr;
  int rc = arena_alloc(ctx, pm.blob.size(), &base);
  if (rc) return rc;
  rc = stage_model_bytes(ctx, base, pm.blob.data(), pm.blob.size());
  if (rc) return rc;
  const size_t b64 = align_up(sb_m64_size(m.AP) * 8, 256), bf = align_up(sb_mr_size(m.AP) * 4, 256);
  m.d64 = reinterpret_cast<double*>(base);
  m.rec[0] = base + b64;
  m.rec[1] = base + b64 + bf;
  m.base = base;
  m.bytes = pm.blob.size();
  int id;
  if (!ctx->free_ids.empty()) {
    id = ctx->free_ids.back();
    ctx->free_ids.pop_back();
    ctx->models[id] = std::move(m);
  } else {
    id = int(ctx->models.size());
    ctx->models.push_back(std::move(m));
  }
  *model_id = id;
  return SB_OK;
}

int drop_model(sb_context* ctx, int model_id) {
  Model* m = get_model(ctx, model_id);
  if (!m) return fail(SB_ERR_MODEL, "unknown model id %d", model_id);
  // the record is reused by a later model put once every launch that may
  // read it has finished
  if (m->base) {
    ctx->retired[m->bytes].push_back({m->base, m->bytes, std::move(m->uses)});
    ++ctx->n_retired;
  }
  m->uses.clear();
  m->base = nullptr;
  m->d64 = nullptr;
  m->rec[0] = m->rec[1] = nullptr;
  m->alive = false;
  ctx->free_ids.push_back(model_id);
  return SB_OK;
}

}  // namespace

extern "C" {

const char* sb_last_error(void) { return g_err.c_str(); }

// Internal hooks for the serving loop (sb_serve.cpp): set the calling
// thread's error message; the device a context was created on.
int sb_internal_error(int code, const char* msg) {
  g_err = msg;
  return code;
}
int sb_internal_device(sb_context* ctx) { return ctx ? ctx->device : -1; }
// Padded alphabet of an alphabet size (-1 outside 2..64): requests of one
// padded alphabet (and reversibility) share a launch.
int sb_internal_pad_alphabet(int A) { return A < 2 ? -1 : pad_alphabet(A); }
// 1 for an irreversible model, 0 for a reversible one, -1 for an unknown id.
int sb_internal_model_irr(sb_context* ctx, int model_id) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lock(ctx->mu);
  const Model* m = get_model(ctx, model_id);
  return m ? (m->irr ? 1 : 0) : -1;
}
int sb_version(void) { return 1; }

int sb_create(sb_context** out, int device) {
  if (!out) return fail(SB_ERR_ARG, "null output pointer");
  int n = 0;
  SB_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(SB_ERR_ARG, "device %d not present (%d devices)", device, n);
  SB_HIP(hipSetDevice(device));
  sb_context* ctx = new sb_context();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->num_cus = prop.multiProcessorCount;
  // model-upload events up front, so no call creates them later
  for (int k = 0; k < sb_context::kPend; ++k)
    SB_HIP(hipEventCreateWithFlags(&ctx->pend_ev[k], hipEventDisableTiming));
  SB_HIP(hipEventCreateWithFlags(&ctx->model_ev, hipEventDisableTiming));
  // the no-allocation entry point's launcher; if it cannot be set up, only
  // that entry point fails (with this message)
  if (noalloc_create(ctx) != SB_OK) ctx->noalloc_err = g_err;
  *out = ctx;
  return SB_OK;
}

int sb_destroy(sb_context* ctx) {
  if (!ctx) return SB_OK;
  (void)hipSetDevice(ctx->device);
  batch_free(ctx->internal);
  batch_free(ctx->noalloc);
  if (ctx->plain.p) (void)hipFree(ctx->plain.p);
  if (ctx->plain_bt.p) (void)hipFree(ctx->plain_bt.p);
  (void)hipDeviceSynchronize();
  for (void* c : ctx->arena_chunks) (void)hipFree(c);
  for (int k = 0; k < sb_context::kPend; ++k) {
    if (ctx->pend[k]) (void)hipHostFree(ctx->pend[k]);
    if (ctx->pend_ev[k]) (void)hipEventDestroy(ctx->pend_ev[k]);
    if (ctx->dstage[k]) (void)hipFree(ctx->dstage[k]);
  }
  if (ctx->model_ev) (void)hipEventDestroy(ctx->model_ev);
  delete ctx;
  return SB_OK;
}

int sb_model_put(sb_context* ctx, int A, const double* eigenvalues, const double* eigenvectors,
                 const double* pi, int* model_id) {
  if (!ctx || !model_id) return fail(SB_ERR_ARG, "null argument");
  PreparedModel pm;
  int rc = prepare_model(A, eigenvalues, eigenvectors, pi, &pm);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(ctx->mu);
  return commit_model(ctx, pm, model_id);
}

// Batched puts and drops for the serving loop (one context lock per batch):
// sb_internal_model_prepare builds a model off the lock (any thread) and
// returns it as an opaque handle; sb_internal_model_commit places n of them
// (consuming the handles, also on error); sb_internal_model_drop_many drops n ids.
void* sb_internal_model_prepare(int A, const double* eigenvalues, const double* eigenvectors,
                                const double* pi, int* rc) {
  PreparedModel* pm = new PreparedModel();
  *rc = prepare_model(A, eigenvalues, eigenvectors, pi, pm);
  if (*rc) {
    delete pm;
    return nullptr;
  }
  return pm;
}

int sb_internal_model_commit(sb_context* ctx, int n, void* const* handles, int* ids) {
  int rc = SB_OK;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (int i = 0; i < n; ++i) {
      ids[i] = -1;
      if (!rc && handles[i]) rc = commit_model(ctx, *static_cast<PreparedModel*>(handles[i]), &ids[i]);
    }
  }
  for (int i = 0; i < n; ++i) delete static_cast<PreparedModel*>(handles[i]);
  return rc;
}

int sb_internal_model_drop_many(sb_context* ctx, int n, const int* ids) {
  std::lock_guard<std::mutex> lock(ctx->mu);
  int rc = SB_OK;
  for (int i = 0; i < n; ++i) {
    const int e = drop_model(ctx, ids[i]);
    if (e && !rc) rc = e;
  }
  return rc;
}

int sb_model_put_irrev(sb_context* ctx, int A, const double* eigenvalues, const double* eigenvectors,
                       const double* eigenvectors_inv, const double* pi, int exact, int* model_id) {
  typedef std::complex<double> cd;
  if (!ctx || !eigenvalues || !eigenvectors || !eigenvectors_inv || !pi || !model_id)
    return fail(SB_ERR_ARG, "null argument");
  const int AP = pad_alphabet(A);
  if (A < 2 || AP < 0) return fail(SB_ERR_ALPHABET, "alphabet size %d outside 2..64", A);
  double psum = 0.0;
  for (int a = 0; a < A; ++a) {
    if (!(pi[a] > 0.0) || !std::isfinite(pi[a])) return fail(SB_ERR_MODEL, "pi[%d]=%g must be > 0", a, pi[a]);
    psum += pi[a];
  }
  if (!(psum < 1.0 + 1e-6 && psum > 1.0 - 1e-6)) return fail(SB_ERR_MODEL, "pi sums to %g, not 1", psum);
  for (int i = 0; i < 2 * A; ++i)
    if (!std::isfinite(eigenvalues[i])) return fail(SB_ERR_MODEL, "eigenvalue entry %d not finite", i);
  for (int i = 0; i < 2 * A * A; ++i)
    if (!std::isfinite(eigenvectors[i]) || !std::isfinite(eigenvectors_inv[i]))
      return fail(SB_ERR_MODEL, "eigenvector entry %d not finite", i);
  // padded complex matrices: V, V^{-1} extended by an identity block, mu by zeros
  std::vector<cd> mu(AP), V(AP * AP), Vi(AP * AP);
  for (int k = 0; k < AP; ++k) mu[k] = k < A ? cd(eigenvalues[2 * k], eigenvalues[2 * k + 1]) : cd(0.0);
  for (int a = 0; a < AP; ++a)
    for (int k = 0; k < AP; ++k) {
      const bool in = a < A && k < A;
      const int s = 2 * (a * A + k);
      V[a * AP + k] = in ? cd(eigenvectors[s], eigenvectors[s + 1]) : cd(a == k ? 1.0 : 0.0);
      Vi[a * AP + k] = in ? cd(eigenvectors_inv[s], eigenvectors_inv[s + 1]) : cd(a == k ? 1.0 : 0.0);
    }
  double err = 0.0;
  for (int i = 0; i < A; ++i)
    for (int j = 0; j < A; ++j) {
      cd s = 0.0;
      for (int k = 0; k < A; ++k) s += V[i * AP + k] * Vi[k * AP + j];
      err = std::max(err, std::abs(s - cd(i == j ? 1.0 : 0.0)));
    }
  if (!(err < 1e-6)) return fail(SB_ERR_MODEL, "eigenvectors_inv is not the inverse (|V Vinv - I| = %g)", err);

  std::vector<double> m64(sb_m64i_size(AP), 0.0);
  for (int k = 0; k < AP; ++k) {
    m64[sb_m64i_mu(AP) + k] = mu[k].real();
    m64[sb_m64i_mu(AP) + AP + k] = mu[k].imag();
  }
  for (int i = 0; i < AP * AP; ++i) {
    m64[sb_m64i_V(AP) + i] = V[i].real();
    m64[sb_m64i_V(AP) + AP * AP + i] = V[i].imag();
    m64[sb_m64i_Vi(AP) + i] = Vi[i].real();
    m64[sb_m64i_Vi(AP) + AP * AP + i] = Vi[i].imag();
  }
  for (int a = 0; a < AP; ++a) m64[sb_m64i_pi(AP) + a] = a < A ? pi[a] : 1.0;
  m64[sb_m64i_misc(AP) + 1] = A;

  // record: set 0 (model's convention), set 1 (exact convention), W, pi, ones
  std::vector<double> rec(sb_mri_size(AP), 0.0);
  auto put = [&](int off, int i, int j, cd v) {
    rec[off + i * AP + j] = v.real();
    rec[off + AP * AP + i * AP + j] = v.imag();
  };
  for (int s = 0; s < 2; ++s) {
    const bool ex = s == 1 || exact;
    const int base = sb_mri_set(AP, s);
    for (int i = 0; i < AP; ++i)
      for (int j = 0; j < AP; ++j) {
        const cd v = V[i * AP + j], vt = V[j * AP + i], w = Vi[i * AP + j], wt = Vi[j * AP + i];
        put(base + sb_mri_Pu(AP), i, j, ex ? wt : v);   // Pu = V^{-T} | V
        put(base + sb_mri_Pd(AP), i, j, ex ? v : wt);   // Pd = V | V^{-T}
        put(base + sb_mri_L(AP), i, j, ex ? wt : v);    // L  = V^{-T} | V
        put(base + sb_mri_Rt(AP), i, j, ex ? vt : w);   // Rt = V^T | V^{-1}
      }
  }
  for (int i = 0; i < AP; ++i)
    for (int j = 0; j < AP; ++j) {
      cd q = 0.0;
      if (i < A && j < A)
        for (int k = 0; k < A; ++k) q += V[i * AP + k] * mu[k] * Vi[k * AP + j];
      put(sb_mri_W(AP), i, j, i == j ? cd(1.0) : q);
    }
  for (int a = 0; a < AP; ++a) {
    rec[sb_mri_pi(AP) + a] = a < A ? pi[a] : 0.0;
    rec[sb_mri_ones(AP) + a] = a < A ? 1.0 : 0.0;
  }
  std::vector<float> recf(rec.begin(), rec.end());

  Model m;
  m.alive = true;
  m.A = A;
  m.AP = AP;
  m.irr = true;
  m.pi.assign(pi, pi + A);
  const size_t b64 = align_up(m64.size() * 8, 256), bf = align_up(recf.size() * 4, 256),
               bd = align_up(rec.size() * 8, 256);
  std::vector<char> blob(b64 + bf + bd, 0);
  std::memcpy(blob.data(), m64.data(), m64.size() * 8);
  std::memcpy(blob.data() + b64, recf.data(), recf.size() * 4);
  std::memcpy(blob.data() + b64 + bf, rec.data(), rec.size() * 8);

  std::lock_guard<std::mutex> lock(ctx->mu);
  SB_HIP(hipSetDevice(ctx->device));
  char* base = nullptr;
  int rc = arena_alloc(ctx, b64 + bf + bd, &base);
  if (rc) return rc;
  rc = stage_model_bytes(ctx, base, blob.data(), blob.size());
  if (rc) return rc;
  m.d64 = reinterpret_cast<double*>(base);
  m.rec[0] = base + b64;
  m.rec[1] = base + b64 + bf;
  m.base = base;
  m.bytes = b64 + bf + bd;
  int id;
  if (!ctx->free_ids.empty()) {
    id = ctx->free_ids.back();
    ctx->free_ids.pop_back();
    ctx->models[id] = m;
  } else {
    id = int(ctx->models.size());
    ctx->models.push_back(m);
  }
  *model_id = id;
  return SB_OK;
}

int sb_model_drop(sb_context* ctx, int model_id) {
  if (!ctx) return fail(SB_ERR_ARG, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  return drop_model(ctx, model_id);
}

int sb_arena_stats(sb_context* ctx, long long* reserved_bytes, long long* live_models,
                   int* retired_records) {
  if (!ctx) return fail(SB_ERR_ARG, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  long long n = 0;
  for (const Model& m : ctx->models) n += m.alive ? 1 : 0;
  if (reserved_bytes) *reserved_bytes = (long long)ctx->arena_reserved;
  if (live_models) *live_models = n;
  if (retired_records) *retired_records = int(ctx->n_retired);
  return SB_OK;
}

int sb_loglike(sb_context* ctx, int model_id, const int8_t* d_alignment, const int32_t* parent_index,
               const double* distances, int R, int C, int dtype, double* d_loglike, void* stream) {
  if (!d_loglike && C > 0) return fail(SB_ERR_ARG, "null loglike output");
  BlockIn bi = make_block(model_id, d_alignment, parent_index, distances, R, C);
  bi.flags = SB_F_LL;
  bi.blk.d_loglike = d_loglike;
  return single(ctx, bi, dtype, stream);
}

int sb_root_prob(sb_context* ctx, int model_id, const int8_t* d_alignment, const int32_t* parent_index,
                 const double* distances, int R, int C, int dtype, void* d_root, double* d_loglike,
                 void* stream) {
  if (!d_root && C > 0) return fail(SB_ERR_ARG, "null root output");
  BlockIn bi = make_block(model_id, d_alignment, parent_index, distances, R, C);
  bi.flags = SB_F_ROOT | (d_loglike ? SB_F_LL : 0);
  bi.blk.d_root = d_root;
  bi.blk.d_loglike = d_loglike;
  return single(ctx, bi, dtype, stream);
}

int sb_counts(sb_context* ctx, int model_id, const int8_t* d_alignment, const int32_t* parent_index,
              const double* distances, int R, int C, int dtype, int f81_fast, int mask_mode,
              const uint8_t* d_mask, void* d_counts, void* d_root, double* d_loglike, void* stream) {
  if (!d_counts && C > 0) return fail(SB_ERR_ARG, "null counts output");
  BlockIn bi = make_block(model_id, d_alignment, parent_index, distances, R, C);
  bi.flags = SB_F_COUNTS | (d_root ? SB_F_ROOT : 0) | (d_loglike ? SB_F_LL : 0);
  bi.blk.f81_fast = f81_fast ? 1 : 0;
  bi.blk.mask_mode = mask_mode;
  bi.blk.d_mask = d_mask;
  bi.blk.d_counts = d_counts;
  bi.blk.d_root = d_root;
  bi.blk.d_loglike = d_loglike;
  return single(ctx, bi, dtype, stream);
}

int sb_counts_noalloc(sb_context* ctx, int model_id, const int8_t* d_alignment, const int32_t* parent_index,
                      const double* distances, int R, int C, int dtype, int f81_fast, int mask_mode,
                      const uint8_t* d_mask, void* d_counts, void* d_root, double* d_loglike, void* stream) {
  if (!ctx) return fail(SB_ERR_ARG, "null context");
  if (!d_counts && C > 0) return fail(SB_ERR_ARG, "null counts output");
  BlockIn bi = make_block(model_id, d_alignment, parent_index, distances, R, C);
  bi.flags = SB_F_COUNTS | (d_root ? SB_F_ROOT : 0) | (d_loglike ? SB_F_LL : 0);
  bi.blk.f81_fast = f81_fast ? 1 : 0;
  bi.blk.mask_mode = mask_mode;
  bi.blk.d_mask = d_mask;
  bi.blk.d_counts = d_counts;
  bi.blk.d_root = d_root;
  bi.blk.d_loglike = d_loglike;
  std::lock_guard<std::mutex> lock(ctx->mu);
  SB_HIP(hipSetDevice(ctx->device));
  sb_batch* b = ctx->noalloc;
  if (!b) return fail(SB_ERR_UNSUPPORTED, "no-allocation scratch unavailable: %s", ctx->noalloc_err.c_str());
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (b->stream != s) {
    SB_HIP(hipStreamWaitEvent(s, b->done, 0));
    b->stream = s;
  }
  std::vector<BlockIn> v(1, bi);
  return submit(b, v, dtype, 0);
}

int sb_inside_outside(sb_context* ctx, int model_id, const int8_t* d_alignment,
                      const int32_t* parent_index, const double* distances, int R, int C, int dtype,
                      int f81_fast, int mask_mode, const uint8_t* d_mask, void* d_branch_counts,
                      void* d_node_post, void* d_branch_post, double* d_loglike, void* stream) {
  if (!d_branch_counts && !d_node_post && !d_branch_post && C > 0)
    return fail(SB_ERR_ARG, "no per-branch output requested");
  const Model* m = get_model(ctx, model_id);
  if (m && !is_small(m->AP) && d_branch_counts && (d_node_post || d_branch_post)) {
    // above A = 8 branch counts and posteriors run in different kernels
    int rc = sb_inside_outside(ctx, model_id, d_alignment, parent_index, distances, R, C, dtype, f81_fast,
                               mask_mode, d_mask, d_branch_counts, nullptr, nullptr, d_loglike, stream);
    if (rc) return rc;
    return sb_inside_outside(ctx, model_id, d_alignment, parent_index, distances, R, C, dtype, f81_fast,
                             mask_mode, d_mask, nullptr, d_node_post, d_branch_post, nullptr, stream);
  }
  BlockIn bi = make_block(model_id, d_alignment, parent_index, distances, R, C);
  bi.flags = (d_branch_counts ? SB_F_BCOUNTS : 0) | (d_node_post ? SB_F_NPOST : 0) |
             (d_branch_post ? SB_F_BPOST : 0) | (d_loglike ? SB_F_LL : 0);
  bi.blk.f81_fast = f81_fast ? 1 : 0;
  bi.blk.mask_mode = mask_mode;
  bi.blk.d_mask = d_mask;
  bi.blk.d_loglike = d_loglike;
  bi.bcounts = d_branch_counts;
  bi.npost = d_node_post;
  bi.bpost = d_branch_post;
  return single(ctx, bi, dtype, stream);
}

int sb_loglike_grad(sb_context* ctx, int model_id, const int8_t* d_alignment, const int32_t* parent_index,
                    const double* distances, int R, int C, int dtype, const double* d_g,
                    double* d_dist_grad, int n_rate, const double* rate_weights, double* d_rate_grad,
                    double* d_loglike, void* stream) {
  if (!d_dist_grad) return fail(SB_ERR_ARG, "null dist_grad output");
  if (n_rate < 0 || n_rate > SB_MAX_RATE) return fail(SB_ERR_ARG, "n_rate must be 0..%d", SB_MAX_RATE);
  if (n_rate > 0 && (!rate_weights || !d_rate_grad)) return fail(SB_ERR_ARG, "null rate arrays");
  BlockIn bi = make_block(model_id, d_alignment, parent_index, distances, R, C);
  bi.flags = SB_F_GRAD | (d_loglike ? SB_F_LL : 0);
  bi.blk.d_loglike = d_loglike;
  bi.g = d_g;
  bi.dist_grad = d_dist_grad;
  bi.rate_grad = d_rate_grad;
  bi.n_rate = n_rate;
  bi.rate_w_host = rate_weights;
  return single(ctx, bi, dtype, stream);
}

int sb_counts_multi(sb_context* ctx, int n_cols, int n_rows, const int* model_ids,
                    const int8_t* d_alignment, const int32_t* parent_index, const double* distances,
                    int R, int C, int dtype, int f81_fast, int mask_mode, const uint8_t* d_mask,
                    void* d_counts, void* d_root, double* d_loglike, void* stream) {
  MultiOut o;
  o.counts = d_counts;
  o.root = d_root;
  o.ll = d_loglike;
  return multi(ctx, n_cols, n_rows, model_ids, d_alignment, parent_index, distances, R, C, dtype,
               f81_fast, mask_mode, d_mask, o, stream);
}

int sb_branch_counts_multi(sb_context* ctx, int n_cols, int n_rows, const int* model_ids,
                           const int8_t* d_alignment, const int32_t* parent_index,
                           const double* distances, int R, int C, int dtype, int f81_fast,
                           int mask_mode, const uint8_t* d_mask, void* d_branch_counts,
                           double* d_loglike, void* stream) {
  if (!d_branch_counts && C > 0) return fail(SB_ERR_ARG, "null branch counts output");
  MultiOut o;
  o.bcounts = d_branch_counts;
  o.ll = d_loglike;
  return multi(ctx, n_cols, n_rows, model_ids, d_alignment, parent_index, distances, R, C, dtype,
               f81_fast, mask_mode, d_mask, o, stream);
}

int sb_inside_outside_multi(sb_context* ctx, int n_cols, const int* model_ids, const int8_t* d_alignment,
                            const int32_t* parent_index, const double* distances, int R, int C, int dtype,
                            void* d_node_post, void* d_branch_post, double* d_loglike, void* stream) {
  if (!d_node_post && !d_branch_post && C > 0) return fail(SB_ERR_ARG, "no posterior output requested");
  MultiOut o;
  o.npost = d_node_post;
  o.bpost = d_branch_post;
  o.ll = d_loglike;
  return multi(ctx, n_cols, 1, model_ids, d_alignment, parent_index, distances, R, C, dtype, 0,
               SB_MASK_NONE, nullptr, o, stream);
}

int sb_mixture_posterior(sb_context* ctx, int n_models, const int* model_ids,
                         const int8_t* d_alignment, const int32_t* parent_index,
                         const double* distances, int R, int C, int dtype, const double* log_weights,
                         double* d_post, double* d_loglikes, void* stream) {
  if (!ctx || !model_ids || !log_weights || !parent_index || !distances)
    return fail(SB_ERR_ARG, "null argument");
  if (n_models < 1 || n_models > SB_MAX_MIX)
    return fail(SB_ERR_ARG, "n_models must be 1..%d, got %d", SB_MAX_MIX, n_models);
  if (!d_post && C > 0) return fail(SB_ERR_ARG, "null posterior output");
  if (R < 1 || C < 0) return fail(SB_ERR_ARG, "bad sizes R=%d C=%d", R, C);
  for (int k = 0; k < n_models; ++k)
    if (!std::isfinite(log_weights[k]) && !(log_weights[k] < 0))
      return fail(SB_ERR_ARG, "log_weights[%d] is not a number or +inf", k);
  std::lock_guard<std::mutex> lock(ctx->mu);
  sb_batch* b = nullptr;
  int rc = context_launcher(ctx, stream, &b);
  if (rc) return rc;
  if (C == 0) return SB_OK;
  double* ll = d_loglikes;
  if (!ll) {
    rc = ensure(b, b->xout, size_t(n_models) * C * sizeof(double));
    if (rc) return rc;
    ll = static_cast<double*>(b->xout.p);
  }
  std::vector<BlockIn> blocks(n_models);
  for (int k = 0; k < n_models; ++k) {
    blocks[k] = make_block(model_ids[k], d_alignment, parent_index, distances, R, C);
    blocks[k].flags = SB_F_LL;
    blocks[k].blk.d_loglike = ll + size_t(k) * C;
  }
  for (int k = 1; k < n_models; ++k) {
    const Model* a = get_model(ctx, model_ids[0]);
    const Model* m = get_model(ctx, model_ids[k]);
    if (a && m && a->A != m->A) return fail(SB_ERR_UNSUPPORTED, "mixture components must share the alphabet");
  }
  rc = submit(b, blocks, dtype, 0);
  if (rc) return rc;
  SB_HIP(sb_launch_mixture_softmax(ll, d_post, n_models, C, log_weights, b->stream));
  SB_HIP(hipEventRecord(b->done, b->stream));
  return SB_OK;
}

int sb_batch_create(sb_context* ctx, void* stream, sb_batch** out) {
  if (!ctx || !out) return fail(SB_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lock(ctx->mu);
  SB_HIP(hipSetDevice(ctx->device));
  return batch_create(ctx, static_cast<hipStream_t>(stream), out);
}

int sb_batch_destroy(sb_batch* b) {
  if (!b) return SB_OK;
  std::lock_guard<std::mutex> lock(b->ctx->mu);
  batch_free(b);
  return SB_OK;
}

int sb_batch_set_tiling(sb_batch* b, int waves_per_group, int waves_per_cu) {
  if (!b) return fail(SB_ERR_ARG, "null launcher");
  if (waves_per_group) {
    if (waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4)
      return fail(SB_ERR_ARG, "waves_per_group must be 1, 2 or 4");
    b->waves_per_group = waves_per_group;
  }
  if (waves_per_cu >= 0) {
    if (waves_per_cu > 32) return fail(SB_ERR_ARG, "waves_per_cu must be 0..32 (0: from occupancy)");
    b->waves_per_cu = waves_per_cu;
  }
  return SB_OK;
}

int sb_batch_set_tune(sb_batch* b, int knob, int value) {
  if (!b) return fail(SB_ERR_ARG, "null launcher");
  switch (knob) {
    case SB_TUNE_WAVES_PER_GROUP: return sb_batch_set_tiling(b, value, -1);
    case SB_TUNE_WAVES_PER_CU: return sb_batch_set_tiling(b, 0, value);
    case SB_TUNE_GEN2_ROUND:
      if (value < 0) return fail(SB_ERR_ARG, "gen2 round must be >= 0 tiles");
      b->gen2_round = value;
      return SB_OK;
    case SB_TUNE_SPLIT:
      if (value < 0 || value > 1) return fail(SB_ERR_ARG, "split must be 0 or 1");
      b->split = value;
      return SB_OK;
    case SB_TUNE_SPLIT_MIN:
      if (value < 8) return fail(SB_ERR_ARG, "split_min must be >= 8 internal nodes");
      b->split_min = value;
      return SB_OK;
    default: return fail(SB_ERR_ARG, "unknown tuning knob %d", knob);
  }
}

int sb_context_set_tune(sb_context* ctx, int knob, int value) {
  if (!ctx) return fail(SB_ERR_ARG, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  SB_HIP(hipSetDevice(ctx->device));
  if (!ctx->internal) {
    // (its stream is set by the first single call, which then waits on this
    // recorded, already complete event)
    int rc = batch_create(ctx, nullptr, &ctx->internal);
    if (rc) return rc;
    SB_HIP(hipEventRecord(ctx->internal->done, nullptr));
  }
  return sb_batch_set_tune(ctx->internal, knob, value);
}

int sb_batch_get_tune(sb_batch* b, int knob, int* value) {
  if (!b || !value) return fail(SB_ERR_ARG, "null argument");
  switch (knob) {
    case SB_TUNE_WAVES_PER_GROUP: *value = b->waves_per_group; return SB_OK;
    case SB_TUNE_WAVES_PER_CU: *value = b->waves_per_cu; return SB_OK;
    case SB_TUNE_GEN2_ROUND: *value = b->gen2_round; return SB_OK;
    case SB_TUNE_SPLIT: *value = b->split; return SB_OK;
    case SB_TUNE_LAST_SPLIT: *value = b->last_split; return SB_OK;
    case SB_TUNE_SPLIT_MIN: *value = b->split_min; return SB_OK;
    default: return fail(SB_ERR_ARG, "unknown tuning knob %d", knob);
  }
}

int sb_batch_submit(sb_batch* b, int n_blocks, const sb_block* blocks, int dtype, int use_graph) {
  if (!b || (n_blocks > 0 && !blocks) || n_blocks < 0) return fail(SB_ERR_ARG, "bad batch arguments");
  std::lock_guard<std::mutex> lock(b->ctx->mu);
  SB_HIP(hipSetDevice(b->ctx->device));
  std::vector<BlockIn> v(n_blocks);
  for (int i = 0; i < n_blocks; ++i) {
    std::memset(&v[i], 0, sizeof v[i]);
    v[i].blk = blocks[i];
    int w = blocks[i].want;
    if (w & ~(SB_WANT_LOGLIKE | SB_WANT_ROOT | SB_WANT_COUNTS)) return fail(SB_ERR_ARG, "block %d: bad want bits", i);
    v[i].flags = ((w & SB_WANT_LOGLIKE) ? SB_F_LL : 0) | ((w & SB_WANT_ROOT) ? SB_F_ROOT : 0) |
                 ((w & SB_WANT_COUNTS) ? SB_F_COUNTS : 0);
  }
  return submit(b, v, dtype, use_graph);
}

int sb_counts_plain(sb_context* ctx, int model_id, const int8_t* d_alignment, const int32_t* parent_index,
                    const double* distances, int R, int C, int mask_mode, float* d_counts, float* d_root,
                    double* d_loglike, void* stream) {
  if (!ctx || !parent_index || !distances) return fail(SB_ERR_ARG, "null argument");
  if (!d_loglike) return fail(SB_ERR_ARG, "plain path needs a loglike buffer");
  if (mask_mode != SB_MASK_NONE && mask_mode != SB_MASK_AUTO)
    return fail(SB_ERR_UNSUPPORTED, "plain path supports mask_mode none or auto");
  std::lock_guard<std::mutex> lock(ctx->mu);
  SB_HIP(hipSetDevice(ctx->device));
  const Model* m = get_model(ctx, model_id);
  if (!m) return fail(SB_ERR_MODEL, "unknown model id %d", model_id);
  if (m->A != 4 || m->irr)
    return fail(SB_ERR_UNSUPPORTED, "plain path is implemented for reversible A = 4 models only");
  TreePlan tp;
  int rc = plan_tree(parent_index, R, tp);
  if (rc) return rc;
  if (C <= 0) return SB_OK;
  // M(t), J(t) in FP64 on the host (diagonalize.py:141-172, eigensub.py:7-50)
  const int A = 4, bts = sb_bt_stride(4);
  ctx->plain_bt_host.assign(size_t(R) * bts, 0.f);
  for (int n = 0; n < R; ++n) {
    const double t = distances[n];
    double E[4];
    for (int k = 0; k < A; ++k) E[k] = std::exp(m->mu[k] * t);
    float* o = &ctx->plain_bt_host[size_t(n) * bts];
    for (int i = 0; i < A; ++i)
      for (int j = 0; j < A; ++j) {
        double s = 0.0;
        for (int k = 0; k < A; ++k) s += m->V[i * A + k] * E[k] * m->V[j * A + k];
        o[sb_bt_M(4) + i * A + j] = float(s * std::sqrt(m->pi[j] / m->pi[i]));
        const double d = m->mu[i] - m->mu[j];
        o[sb_bt_J(4) + i * A + j] = float(std::fabs(d) < 1e-8 ? t * E[i] : E[j] * std::expm1(d * t) / d);
      }
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  rc = flush_models(ctx, s);
  if (rc) return rc;
  const size_t bytes = ctx->plain_bt_host.size() * sizeof(float);
  if (bytes > ctx->plain_bt.cap) {
    SB_HIP(hipStreamSynchronize(s));
    if (ctx->plain_bt.p) SB_HIP(hipFree(ctx->plain_bt.p));
    ctx->plain_bt.p = nullptr;
    ctx->plain_bt.cap = 0;
    if (hipMalloc(&ctx->plain_bt.p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SB_ERR_OOM, "plain branch table allocation failed");
    }
    ctx->plain_bt.cap = bytes;
  }
  SB_HIP(hipMemcpyAsync(ctx->plain_bt.p, ctx->plain_bt_host.data(), bytes, hipMemcpyHostToDevice, s));
  SB_HIP(sb_plain_run(d_alignment, parent_index, R, C, static_cast<const float*>(ctx->plain_bt.p),
                      static_cast<const float*>(m->rec[0]), mask_mode, d_counts, d_root, d_loglike,
                      &ctx->plain, s));
  return SB_OK;
}

int sb_batch_set_timing(sb_batch* b, int enable) {
  if (!b) return fail(SB_ERR_ARG, "null launcher");
  b->timing = enable != 0;
  return SB_OK;
}

int sb_batch_kernel_ms(sb_batch* b, double* total_ms, int* launches) {
  if (!b || !total_ms || !launches) return fail(SB_ERR_ARG, "null argument");
  double tot = 0.0;
  for (auto& e : b->tev) {
    SB_HIP(hipEventSynchronize(e.second));
    float ms = 0.f;
    SB_HIP(hipEventElapsedTime(&ms, e.first, e.second));
    tot += ms;
    b->tfree.push_back(e);
  }
  *launches = int(b->tev.size());
  b->tev.clear();
  *total_ms = tot;
  return SB_OK;
}

int sb_batch_sync(sb_batch* b) {
  if (!b) return fail(SB_ERR_ARG, "null launcher");
  SB_HIP(hipEventSynchronize(b->done));
  return SB_OK;
}

}  // extern "C"
