/*
 * subby_amd — C ABI of the MI355X (gfx950) per-column sufficient-statistics engine.
 *
 * Every entry point returns an int status (SB_OK == 0, negative on error) and
 * leaves a human-readable message retrievable with sb_last_error().  Buffers
 * named d_* are device pointers owned by the caller; small tree arrays
 * (parent_index, distances) are host pointers.  All work is stream-ordered on
 * the hipStream_t passed as `stream` (NULL = default stream) and returns
 * without synchronising.
 *
 * Layouts (row-major, the reference's own):
 *   alignment  int8  [R][C]   tokens 0..A-1 observed, A ungapped-unobserved,
 *                             A+1 or -1 gap (any other value reads as a gap)
 *   loglike    f64   [C]
 *   root       T     [A][C]   posterior root distribution
 *   counts     T     [A][A][C] diagonal = dwell times, off-diagonal = counts
 *   mask       u8    [R][C]   explicit branch mask (row 0 ignored)
 * T is float for SB_F32 and double for SB_F64.
 *
 * Trees are rooted binary trees in preorder: parent_index[0] == -1 and
 * 0 <= parent_index[n] < n; every node has 0 or 2 children.
 */
#ifndef SUBBY_AMD_H
#define SUBBY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SB_OK = 0,
  SB_ERR_ARG = -1,         /* bad argument (null pointer, size, unknown flag) */
  SB_ERR_TREE = -2,        /* parent_index not preorder or not binary */
  SB_ERR_ALPHABET = -3,    /* A outside 2..64 */
  SB_ERR_OOM = -4,         /* device or pinned allocation failed */
  SB_ERR_HIP = -5,         /* a HIP runtime call failed */
  SB_ERR_MODEL = -6,       /* unknown model id or invalid eigensystem */
  SB_ERR_UNSUPPORTED = -7  /* combination not supported (e.g. F81 on non-F81 model) */
};

enum { SB_F32 = 0, SB_F64 = 1 };
enum { SB_MASK_NONE = 0, SB_MASK_AUTO = 1, SB_MASK_EXPLICIT = 2 };

typedef struct sb_context sb_context;
typedef struct sb_batch sb_batch;

/* Message of the last failing call on this host thread. */
const char* sb_last_error(void);
int sb_version(void);

/* Context: one per device.  Holds the resident model table.
 * Replaces the module-level state of subby/jax/__init__.py (none: JAX is
 * functional); the context exists so eigensystems stay resident on device. */
int sb_create(sb_context** out, int device);
int sb_destroy(sb_context* ctx);

/* Upload an eigensystem (DiagModel, subby/jax/types.py:8-11): eigenvalues[A],
 * eigenvectors[A*A] row-major with V[a*A+k] = component a of eigenvector k,
 * pi[A].  The model stays resident until sb_model_drop / sb_destroy. */
int sb_model_put(sb_context* ctx, int A, const double* eigenvalues,
                 const double* eigenvectors, const double* pi, int* model_id);
/* Retire a model id.  Its arena record is reused by a later sb_model_put of
 * the same alphabet once every launch already submitted that reads it has
 * finished (each launcher's submissions are fenced by events). */
int sb_model_drop(sb_context* ctx, int model_id);
/* Model arena diagnostics: bytes of records handed out so far (live models
 * plus records waiting for reuse), live model count, retired records waiting
 * for reuse.  Any output may be NULL.  (No reference counterpart.) */
int sb_arena_stats(sb_context* ctx, long long* reserved_bytes, long long* live_models,
                   int* retired_records);

/* Upload an irreversible eigensystem (IrrevDiagModel, subby/jax/types.py:14-18;
 * built by diagonalize_irreversible, subby/jax/diagonalize.py:49-67):
 * Q = V diag(mu) V^{-1}.  Complex arrays are interleaved (re, im) pairs, as
 * C99 double complex / numpy complex128: eigenvalues[2A], eigenvectors[2*A*A]
 * with V[a][k] at 2*(a*A+k), eigenvectors_inv[2*A*A] = V^{-1}; pi[A] real.
 * exact == 0 reproduces the reference's counts (eigensub.py:271-434:
 * projections V / V^{-T}, back-transform V C V^{-1}, whose integrand is that
 * of Q^T); exact != 0 gives the expected counts under Q itself (projections
 * V^{-T} / V, back-transform V^{-T} C V^T).  LogLike, RootProb and the
 * gradient do not depend on it.  Every output, the gradient included, takes
 * alphabets up to 64.  The F81 fast path rejects them (SB_ERR_UNSUPPORTED). */
int sb_model_put_irrev(sb_context* ctx, int A, const double* eigenvalues,
                       const double* eigenvectors, const double* eigenvectors_inv,
                       const double* pi, int exact, int* model_id);

/* Single-block calls (drop-in for subby.jax LogLike / RootProb / Counts). */

/* subby/jax/__init__.py:207-291 (LogLike, single DiagModel branch :285-291). */
int sb_loglike(sb_context* ctx, int model_id, const int8_t* d_alignment,
               const int32_t* parent_index, const double* distances, int R, int C,
               int dtype, double* d_loglike, void* stream);

/* subby/jax/__init__.py:543-650 (RootProb, single model :635-650).
 * d_loglike may be NULL. */
int sb_root_prob(sb_context* ctx, int model_id, const int8_t* d_alignment,
                 const int32_t* parent_index, const double* distances, int R, int C,
                 int dtype, void* d_root, double* d_loglike, void* stream);

/* subby/jax/__init__.py:338-447 (Counts, single model :413-447).
 * f81_fast selects the closed-form F81/JC path (f81_fast.py:7-141);
 * mask_mode: SB_MASK_NONE, SB_MASK_AUTO (Steiner mask, oracle.py:1372-1435),
 * SB_MASK_EXPLICIT (d_mask).  d_root and d_loglike may be NULL. */
int sb_counts(sb_context* ctx, int model_id, const int8_t* d_alignment,
              const int32_t* parent_index, const double* distances, int R, int C,
              int dtype, int f81_fast, int mask_mode, const uint8_t* d_mask,
              void* d_counts, void* d_root, double* d_loglike, void* stream);

/* sb_counts without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  Same arguments, results
 * and errors as sb_counts (bitwise the same results), but the call's plan,
 * tile counters, branch records and per-wave scratch live in one static
 * device array of 256 MiB per GPU (reserved when the library loads on that
 * GPU, held until the process ends), and the call makes no hipMalloc,
 * hipFree or other allocating call.  Limits: launch plan <= 4 MiB (about
 * 500,000 column tiles), branch records <= 12 MiB - 64 KiB (R up to ~17,000
 * nodes at A <= 4 in FP32), at least one wave's scratch slot in 240 MiB;
 * beyond them it fails with SB_ERR_ARG and a message.  Fewer waves run when
 * the slots of a full grid do not fit (slower, same results); general models
 * above 8 states run the large-alphabet VALU kernel rather than the two-pass
 * path.  Every call on a GPU (from any context) shares that GPU's array and
 * leaves its contents behind: calls that overlap in time on one GPU race, so
 * the caller keeps them in one stream or otherwise ordered. */
int sb_counts_noalloc(sb_context* ctx, int model_id, const int8_t* d_alignment,
                      const int32_t* parent_index, const double* distances, int R, int C,
                      int dtype, int f81_fast, int mask_mode, const uint8_t* d_mask,
                      void* d_counts, void* d_root, double* d_loglike, void* stream);

/* Gradient of sum_c g_c logL_c (subby/jax/vjp.py:62-133, method 'eigen').
 * d_g: [C] f64 cotangent, or NULL for all ones.
 * d_dist_grad: [R] f64 output, entry 0 (root) is 0.
 * n_rate: number of rate parameters; rate_weights: host [n_rate][A*A] f64 with
 *   W_ij = dQ_ij/dtheta / Q_ij (i != j) and W_ii = dQ_ii/dtheta, so that
 *   dlogL/dtheta = sum_ij W_ij Counts_ij (unmasked) — HR2002 eqs. 3-4.
 *   (The reference has no rate-parameter gradient; this extends vjp.py.)
 * d_rate_grad: [n_rate] f64 output (may be NULL when n_rate == 0).
 * d_loglike may be NULL. */
int sb_loglike_grad(sb_context* ctx, int model_id, const int8_t* d_alignment,
                    const int32_t* parent_index, const double* distances, int R, int C,
                    int dtype, const double* d_g, double* d_dist_grad,
                    int n_rate, const double* rate_weights, double* d_rate_grad,
                    double* d_loglike, void* stream);

/* Per-branch statistics of the inside-outside tables: replaces
 * BranchCounts (subby/jax/__init__.py:450-540) and the InsideOutside queries
 * branch_counts / node_posterior / branch_posterior (:681-1040).  Any output
 * may be NULL (at least one of the three per-branch outputs must be given):
 *   d_branch_counts  T   [R][A][A][C]  per-branch counts (branch 0 zeros); mask_mode
 *                                       as in sb_counts (subby.jax applies none)
 *   d_node_post      T   [R][A][C]     node state posteriors (row 0: the root)
 *   d_branch_post    T   [R][A][A][C]  joint (parent = i, node = j) posteriors (branch 0 zeros)
 *   d_loglike        f64 [C]
 * Every output takes any alphabet.  Above 8 states the posteriors run in the
 * large-alphabet kernel's posterior variant and branch counts in its
 * per-branch counts variant (the back-transform after every branch into that
 * branch's slice), each one inside/outside pass; asking for both launches
 * the two in turn. */
int sb_inside_outside(sb_context* ctx, int model_id, const int8_t* d_alignment,
                      const int32_t* parent_index, const double* distances, int R, int C, int dtype,
                      int f81_fast, int mask_mode, const uint8_t* d_mask, void* d_branch_counts,
                      void* d_node_post, void* d_branch_post, double* d_loglike, void* stream);

/* Several models in one call: the reference's per-column model lists and
 * model grids (subby/jax/__init__.py:52-204; LogLike :230-283, Counts
 * :362-411, RootProb :562-632).  model_ids is a host table
 * [n_cols][n_rows] of resident model ids, n_cols = 1 or C, n_rows = 1 or R:
 *   (1, 1) one model (as sb_counts);
 *   (C, 1) model_ids[c] for column c (model list): counts use each column's
 *          own model; loglike and root use the root distribution of column
 *          0's model, as the reference's stacked models do (:264-274, :617-632);
 *   (1, R) model_ids[n] for branch n (model grid (1, R)); row 0 gives the root
 *          distribution; counts are back-transformed per branch in each
 *          branch's eigenbasis (:102-144);
 *   (C, R) both (model grid (C, R)).
 * All models share the alphabet and reversibility.  Per-branch models take
 * no F81 path (SB_ERR_UNSUPPORTED), and no gradient.  Outputs as sb_counts;
 * any may be NULL, at least one given. */
int sb_counts_multi(sb_context* ctx, int n_cols, int n_rows, const int* model_ids,
                    const int8_t* d_alignment, const int32_t* parent_index, const double* distances,
                    int R, int C, int dtype, int f81_fast, int mask_mode, const uint8_t* d_mask,
                    void* d_counts, void* d_root, double* d_loglike, void* stream);

/* BranchCounts with a model table as sb_counts_multi (subby/jax/__init__.py:474-500):
 * d_branch_counts T [R][A][A][C]; mask_mode as in sb_counts (the reference applies none). */
int sb_branch_counts_multi(sb_context* ctx, int n_cols, int n_rows, const int* model_ids,
                           const int8_t* d_alignment, const int32_t* parent_index,
                           const double* distances, int R, int C, int dtype, int f81_fast,
                           int mask_mode, const uint8_t* d_mask, void* d_branch_counts,
                           double* d_loglike, void* stream);

/* InsideOutside over a per-column model list (subby/jax/__init__.py:719-780,
 * node_posterior :870-960, branch_posterior :962-1040): model_ids is a host
 * table [n_cols] (n_cols = 1 or C) of resident models sharing the alphabet and
 * reversibility.  As in the reference, the inside and outside passes of
 * column c run under model c's matrices with the root distribution of
 * model_ids[0], and the root row of the node posteriors uses model c's own.
 *   d_node_post T [R][A][C], d_branch_post T [R][A][A][C] (branch 0 zeros),
 *   d_loglike f64 [C] (LogLike of the list); any may be NULL, not all. */
int sb_inside_outside_multi(sb_context* ctx, int n_cols, const int* model_ids, const int8_t* d_alignment,
                            const int32_t* parent_index, const double* distances, int R, int C, int dtype,
                            void* d_node_post, void* d_branch_post, double* d_loglike, void* stream);

/* MixturePosterior (subby/jax/__init__.py:653-678, mixture.py:6-30): the
 * log-likelihoods of n_models (1..64) resident models over one alignment in
 * one persistent launch, then d_post[k][c] = softmax_k(ll[k][c] + log_weights[k]).
 * log_weights: host [n_models]; d_post f64 [n_models][C]; d_loglikes f64
 * [n_models][C] or NULL. */
int sb_mixture_posterior(sb_context* ctx, int n_models, const int* model_ids,
                         const int8_t* d_alignment, const int32_t* parent_index,
                         const double* distances, int R, int C, int dtype, const double* log_weights,
                         double* d_post, double* d_loglikes, void* stream);

/* Batched calls: many (tree, alignment-block) pairs in one persistent launch.
 * Mirrors jax.vmap(Counts) of scripts/batch_eigensub_maf.py:528-532 without
 * requiring equal tree sizes. */
typedef struct sb_block {
  int model_id;
  int R, C;
  const int8_t* d_alignment;    /* [R][C] */
  const int32_t* parent_index;  /* host [R] */
  const double* distances;      /* host [R] */
  int want;                     /* SB_WANT_* bits */
  int f81_fast;
  int mask_mode;
  const uint8_t* d_mask;        /* [R][C] or NULL */
  void* d_counts;               /* [A][A][C] or NULL */
  void* d_root;                 /* [A][C] or NULL */
  double* d_loglike;            /* [C] or NULL */
} sb_block;

enum { SB_WANT_LOGLIKE = 1, SB_WANT_ROOT = 2, SB_WANT_COUNTS = 4 };

/* A batch launcher bound to one stream; owns its staging, plan and scratch
 * buffers so several launchers can run concurrently on different streams. */
int sb_batch_create(sb_context* ctx, void* stream, sb_batch** out);
int sb_batch_destroy(sb_batch* b);
/* Enqueue every block; all blocks must use one A and one dtype.
 * use_graph != 0 replays a captured hipGraph of the kernel sequence. */
int sb_batch_submit(sb_batch* b, int n_blocks, const sb_block* blocks, int dtype,
                    int use_graph);
/* Tile shape of the persistent grid: waves per workgroup (1, 2 or 4; 0 keeps
 * the current value) and waves per CU (1..32; 0 = as many as the kernel's
 * occupancy allows, the default; negative keeps the current value). */
int sb_batch_set_tiling(sb_batch* b, int waves_per_group, int waves_per_cu);
/* Launch-shape knobs of a launcher (the autotuner's search space; no
 * reference counterpart — the reference sizes its vmap batches by a memory
 * budget only, scripts/batch_eigensub_maf.py:296-304):
 *   SB_TUNE_WAVES_PER_GROUP  1, 2 or 4 (small-alphabet kernels)
 *   SB_TUNE_WAVES_PER_CU     0..32, 0 = from the kernel's occupancy
 *   SB_TUNE_GEN2_ROUND       two-pass path (general models above 8 states,
 *                            FP32): 64-column tiles staged per round, 0 = as
 *                            many as an 8 GiB stage holds (default); smaller
 *                            rounds keep the stage in cache
 *   SB_TUNE_SPLIT            small-alphabet single-model kinds at four states:
 *                            two waves walk each tile (a subtree each) -- 1
 *                            wherever a tree names a subtree (default; two
 *                            waves per workgroup, waves_per_group is then
 *                            ignored), 0 never
 *   SB_TUNE_LAST_SPLIT       read only: 1 when the launcher's last submission
 *                            walked its tiles with two waves
 *   SB_TUNE_SPLIT_MIN        internal nodes a tree needs to split (>= 8)
 * Results do not depend on any knob (bitwise), except that a split walk sums
 * the subtree's branch terms apart: its Counts (and rate gradients) may differ
 * from the one-wave walk's in the last bits; log-likelihoods, root
 * posteriors and branch-length gradients do not. */
enum { SB_TUNE_WAVES_PER_GROUP = 0, SB_TUNE_WAVES_PER_CU = 1, SB_TUNE_GEN2_ROUND = 3, SB_TUNE_SPLIT = 4,
       SB_TUNE_LAST_SPLIT = 5, SB_TUNE_SPLIT_MIN = 6 };
int sb_batch_set_tune(sb_batch* b, int knob, int value);
int sb_batch_get_tune(sb_batch* b, int knob, int* value);
/* The same knobs on the launcher behind the single-call entry points
 * (sb_loglike, sb_counts, sb_loglike_grad, ...), e.g. SB_TUNE_SPLIT for the
 * latency of small calls.  No reference counterpart. */
int sb_context_set_tune(sb_context* ctx, int knob, int value);
/* Block until the launcher's last submission has finished. */
int sb_batch_sync(sb_batch* b);
/* Kernel timing: when enabled, eager (non-graph) submissions bracket the
 * main kernel with HIP events on the launcher's stream; sb_batch_kernel_ms
 * synchronises, returns the summed main-kernel time and the number of
 * launches since the last call, and resets both.  (No reference counterpart:
 * measurement hook for the benchmark.) */
int sb_batch_set_timing(sb_batch* b, int enable);
int sb_batch_kernel_ms(sb_batch* b, double* total_ms, int* launches);

/* Continuous-batching server.  Replaces the batched driver's host loop
 * (scripts/batch_eigensub_maf.py:423-698: bucket blocks by size, pad, vmap
 * Counts per bucket, copy features back) with a native admission loop: a
 * worker thread admits queued requests first come first served into free
 * lanes (launcher + HIP stream + pinned and device buffers each) while the
 * bytes in flight stay within the HBM budget, uploads each batch with one
 * copy, runs it as one persistent launch, and copies results straight into
 * the callers' host buffers.  Results are bitwise those of the same request
 * run alone.  All host pointers of a request must stay valid until it
 * completes.  Outputs are FP32 counts / root and FP64 logLike.  Inline models
 * are kept resident by content (requests with the same eigensystem share one
 * record); beyond 4096 of them the least recently used idle one is dropped,
 * so a stream of distinct models runs in a bounded model arena. */
typedef struct sb_server sb_server;
typedef struct sb_request {
  int model_id;                 /* resident model, or -1: the inline model below */
  int A;                        /* inline model: alphabet size */
  const double* eigenvalues;    /* inline model [A] */
  const double* eigenvectors;   /* inline model [A*A], V[a*A + k] */
  const double* pi;             /* inline model [A] */
  int R, C;
  const int8_t* alignment;      /* host [R][C] */
  const int32_t* parent_index;  /* host [R] */
  const double* distances;      /* host [R] */
  int mask_mode;                /* SB_MASK_NONE or SB_MASK_AUTO */
  int want;                     /* SB_WANT_* bits */
  float* counts;                /* host [A][A][C] or NULL */
  float* root;                  /* host [A][C] or NULL */
  double* loglike;              /* host [C] or NULL */
  long long tag;                /* returned on completion */
} sb_request;
int sb_server_create(sb_context* ctx, int lanes, long long hbm_budget, int max_blocks,
                     long long max_cols, sb_server** out);
int sb_server_destroy(sb_server* s);   /* waits for everything submitted */
int sb_server_submit(sb_server* s, int n, const sb_request* requests);
/* Up to `max` completed tags with completion times (seconds since the
 * server was created); *n_done receives how many. */
int sb_server_poll(sb_server* s, int max, long long* tags, double* done_s, int* n_done);
int sb_server_drain(sb_server* s);     /* wait until every submitted request completed */
/* Load generator: submit request i at arrival_s[i] seconds after the call
 * (non-decreasing), wait for all, and write each request's completion time
 * (seconds after the call) to done_s[i].  Request tags are overwritten with
 * their index.  (No reference counterpart: measurement of the serving loop.) */
int sb_server_replay(sb_server* s, int n, sb_request* requests, const double* arrival_s,
                     double* done_s);
/* Where a request's latency went.  Times are seconds after the replay call,
 * like done_s; the *_ms fields are the request's batch's GPU spans (HIP
 * events on the lane's stream) and host planning time. */
typedef struct sb_request_timing {
  double admitted_s;     /* a lane took it off the queue */
  double enqueued_s;     /* its batch's upload, launch and download were enqueued */
  double gpu_done_s;     /* the lane saw its batch's download complete */
  double done_s;         /* its results were copied into the caller's arrays */
  double upload_ms;      /* batch alignments host -> device */
  double kernel_ms;      /* end of upload -> end of the batch's kernels (plan copy included) */
  double download_ms;    /* batch outputs device -> host */
  double plan_ms;        /* host time in the batch submission (planning, model table, launch) */
  double models_ms;      /* host time resolving the batch's inline models (sharing, evictions, arena puts) */
  long long batch_columns;
  int batch_requests;
  int lane;
} sb_request_timing;
/* sb_server_replay with the per-request breakdown above written to
 * timing[i].  (No reference counterpart: measurement of the serving loop.) */
int sb_server_replay_timed(sb_server* s, int n, sb_request* requests, const double* arrival_s,
                           sb_request_timing* timing);

/* Plain path: the same statistics through one launch per tree level per pass
 * with inside/outside partials materialised in HBM (the data flow of the
 * reference's WebGPU engine, subby/webgpu/phylo_gpu.js:611-724).  Kept as the
 * one-block-per-launch baseline. */
int sb_counts_plain(sb_context* ctx, int model_id, const int8_t* d_alignment,
                    const int32_t* parent_index, const double* distances, int R, int C,
                    int mask_mode, float* d_counts, float* d_root, double* d_loglike,
                    void* stream);

#ifdef __cplusplus
}
#endif
#endif
