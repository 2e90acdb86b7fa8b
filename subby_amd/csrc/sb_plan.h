// Plan records shared by the host planner and the gfx950 kernels.
//
// A launch processes a list of blocks (tree + alignment block + model).  The
// host planner (sb_host.cpp) validates each tree, numbers its internal nodes
// in increasing node index ("slots"), and writes one 8-int record per
// internal node:
//     {p, l, r, 0, slot(l) or -1, slot(r) or -1, 0, 0}
// where l is the first child of p in index order (the reference's left child,
// subby/oracle/oracle.py:19-49).  Walking the records backwards is a
// postorder (children have larger indices than parents); forwards a preorder.
#pragma once
#include <stdint.h>

#define SB_WAVE 64
#define SB_MAX_RATE 8
#define SB_MAX_MIX 64   // mixture components of one MixturePosterior call
// Tile counters: counter[0] for kernels with one tile queue; the
// large-alphabet kernel keeps one queue per XCD at counter[x * SB_XCD_GAP].
#define SB_NCOUNTER 64
#define SB_NXCD 8
#define SB_XCD_GAP 8
#define SB_LARGE_TC 4   // columns per wave tile on the large-alphabet path

// Per-branch parameter record, element type T (float or double), padded
// alphabet AP:  M[AP*AP] | J[AP*AP] | coef[4] = {alpha, beta, gamma, t} |
//               E[AP] = exp(mu_k t) | dE[AP] = mu_k exp(mu_k t) |
//               (AP > 8 only) Mt[AP*AP] = M transposed
__host__ __device__ inline int sb_bt_stride(int AP) {
  return (AP > 8 ? 3 : 2) * AP * AP + 4 + 2 * AP;
}
__host__ __device__ inline int sb_bt_Mt(int AP) { return 2 * AP * AP + 4 + 2 * AP; }
__host__ __device__ inline int sb_bt_M(int) { return 0; }
__host__ __device__ inline int sb_bt_J(int AP) { return AP * AP; }
__host__ __device__ inline int sb_bt_coef(int AP) { return 2 * AP * AP; }
__host__ __device__ inline int sb_bt_E(int AP) { return 2 * AP * AP + 4; }
__host__ __device__ inline int sb_bt_dE(int AP) { return 2 * AP * AP + 4 + AP; }

// Model record, element type T, padded alphabet AP (pads are inert states
// that stay exactly zero in every vector):
//   V[AP*AP]   eigenvectors, V[a*AP+k]
//   Vs[AP*AP]  sqrt(pi_a) V[a,k]      (inside projection,  eigensub.py:53-84)
//   Vi[AP*AP]  V[a,k] / sqrt(pi_a)    (outside projection)
//   S[AP*AP]   V diag(mu) V^T          (back-transform,     eigensub.py:138-172)
//   pi[AP]     (pads 0)
//   ones[AP]   1 for real states, 0 for pads (unobserved-token likelihood)
//   misc[4]    {f81_mu, F81-structured (0/1), index of a nonzero eigenvalue, 0}
__host__ __device__ inline int sb_mr_V(int) { return 0; }
__host__ __device__ inline int sb_mr_Vs(int AP) { return AP * AP; }
__host__ __device__ inline int sb_mr_Vi(int AP) { return 2 * AP * AP; }
__host__ __device__ inline int sb_mr_S(int AP) { return 3 * AP * AP; }
__host__ __device__ inline int sb_mr_pi(int AP) { return 4 * AP * AP; }
__host__ __device__ inline int sb_mr_ones(int AP) { return 4 * AP * AP + AP; }
__host__ __device__ inline int sb_mr_misc(int AP) { return 4 * AP * AP + 2 * AP; }
__host__ __device__ inline int sb_mr_size(int AP) { return 4 * AP * AP + 2 * AP + 4; }

// FP64 master copy of a model, used by the branch-parameter kernel:
//   mu[AP] | V[AP*AP] | pi[AP] | misc[4] = {f81_mu, A, 0, 0}
__host__ __device__ inline int sb_m64_mu(int) { return 0; }
__host__ __device__ inline int sb_m64_V(int AP) { return AP; }
__host__ __device__ inline int sb_m64_pi(int AP) { return AP + AP * AP; }
__host__ __device__ inline int sb_m64_misc(int AP) { return 2 * AP + AP * AP; }
__host__ __device__ inline int sb_m64_size(int AP) { return 2 * AP + AP * AP + 4; }

// ---- Irreversible models: Q = V diag(mu) V^{-1} with complex mu, V
// (subby/jax/types.py:14-18).  Complex matrices are stored as a real plane
// followed by an imaginary plane.
//
// Branch record (element T):  M[AP*AP] (real) | Jre[AP*AP] | Jim[AP*AP] |
//   coef[4] = {0, 0, 0, t} | dEre[AP] | dEim[AP]   (dE_k = mu_k e^{mu_k t}) |
//   (AP > 8 only) Mt[AP*AP] = M transposed
__host__ __device__ inline int sb_bti_stride(int AP) { return (AP > 8 ? 4 : 3) * AP * AP + 4 + 2 * AP; }
__host__ __device__ inline int sb_bti_Mt(int AP) { return 3 * AP * AP + 4 + 2 * AP; }
__host__ __device__ inline int sb_bti_Jre(int AP) { return AP * AP; }
__host__ __device__ inline int sb_bti_Jim(int AP) { return 2 * AP * AP; }
__host__ __device__ inline int sb_bti_coef(int AP) { return 3 * AP * AP; }
__host__ __device__ inline int sb_bti_dEre(int AP) { return 3 * AP * AP + 4; }
__host__ __device__ inline int sb_bti_dEim(int AP) { return 3 * AP * AP + 4 + AP; }
__host__ __device__ inline int sb_bt_stride_k(int AP, bool irr) {
  return irr ? sb_bti_stride(AP) : sb_bt_stride(AP);
}

// Model record (element T) of an irreversible model: two sets of complex
// matrices, set 0 for Counts in the model's convention and set 1 for the
// exact convention (gradients), each
//   Pu (AP x AP: U~_l = sum_b U_b Pu[b][l]) | Pd (D~_k = sum_a D_a Pd[a][k]) |
//   L, Rt (VCV = L C Rt)
// then W (counts_ij = Re(W_ij VCV_ij), W = Q off the diagonal, 1 on it),
// pi[AP], ones[AP], misc[4].
//   reference convention (eigensub.py:271-385): Pu = V, Pd = V^{-T}, L = V, Rt = V^{-1}
//   exact convention:                           Pu = V^{-T}, Pd = V, L = V^{-T}, Rt = V^T
__host__ __device__ inline int sb_mri_set(int AP, int s) { return s * 8 * AP * AP; }
__host__ __device__ inline int sb_mri_Pu(int) { return 0; }
__host__ __device__ inline int sb_mri_Pd(int AP) { return 2 * AP * AP; }
__host__ __device__ inline int sb_mri_L(int AP) { return 4 * AP * AP; }
__host__ __device__ inline int sb_mri_Rt(int AP) { return 6 * AP * AP; }
__host__ __device__ inline int sb_mri_W(int AP) { return 16 * AP * AP; }
__host__ __device__ inline int sb_mri_pi(int AP) { return 18 * AP * AP; }
__host__ __device__ inline int sb_mri_ones(int AP) { return 18 * AP * AP + AP; }
__host__ __device__ inline int sb_mri_misc(int AP) { return 18 * AP * AP + 2 * AP; }
__host__ __device__ inline int sb_mri_size(int AP) { return 18 * AP * AP + 2 * AP + 4; }
__host__ __device__ inline int sb_mr_pi_k(int AP, bool irr) { return irr ? sb_mri_pi(AP) : sb_mr_pi(AP); }
__host__ __device__ inline int sb_mr_ones_k(int AP, bool irr) { return irr ? sb_mri_ones(AP) : sb_mr_ones(AP); }

// FP64 master of an irreversible model:
//   mu_re[AP] | mu_im[AP] | V_re, V_im [AP*AP] | Vinv_re, Vinv_im [AP*AP] | pi[AP] | misc[4]
__host__ __device__ inline int sb_m64i_mu(int) { return 0; }
__host__ __device__ inline int sb_m64i_V(int AP) { return 2 * AP; }
__host__ __device__ inline int sb_m64i_Vi(int AP) { return 2 * AP + 2 * AP * AP; }
__host__ __device__ inline int sb_m64i_pi(int AP) { return 2 * AP + 4 * AP * AP; }
__host__ __device__ inline int sb_m64i_misc(int AP) { return 3 * AP + 4 * AP * AP; }
__host__ __device__ inline int sb_m64i_size(int AP) { return 3 * AP + 4 * AP * AP + 4; }
__host__ __device__ inline int sb_m64_pi_k(int AP, bool irr) { return irr ? sb_m64i_pi(AP) : sb_m64_pi(AP); }

enum {
  SB_F_F81 = 1,
  SB_F_MASK_SHIFT = 1,     // 2 bits: 0 none, 1 auto, 2 explicit
  SB_F_COUNTS = 8,
  SB_F_ROOT = 16,
  SB_F_LL = 32,
  SB_F_GRAD = 64,
  SB_F_BCOUNTS = 128,      // per-branch counts   [R][A][A][C]
  SB_F_NPOST = 256,        // node posteriors     [R][A][C]
  SB_F_BPOST = 512,        // branch joint posteriors [R][A][A][C]
  SB_F_IRR = 1024,         // irreversible model records (set by the host)
};

struct SbBlockDesc {
  const int8_t* aln;       // [R][C]
  const uint8_t* mask;     // [R][C] or null
  void* counts;            // [A][A][C] of T or null
  void* root;              // [A][C] of T or null
  double* loglike;         // [C] or null
  const int* sched;        // nint records of 8 ints
  const void* bt;          // R branch records of T
  const void* model;       // model record of T
  const double* model64;   // FP64 master (pi used for the log-likelihood)
  const double* g;         // [C] cotangent of logLike (gradient) or null = ones
  const double* rate_w;    // [n_rate][AP*AP] Fisher weights (gradient)
  double* dist_grad;       // [R] output (gradient)
  double* rate_grad;       // [n_rate] output (gradient)
  void* bcounts;           // [R][A][A][C] of T or null (per-branch outputs)
  void* npost;             // [R][A][C] of T or null
  void* bpost;             // [R][A][A][C] of T or null
  const void* const* row_models;  // [R] model records (T) of per-branch models, or null
  // Root distribution of the inside/outside passes when it is not the
  // model's own (model lists: that of the first column's model): its model
  // record (T) and FP64 master, or null.  The root row of the node
  // posteriors keeps the model's own distribution.
  const void* root_model;
  const double* root_model64;
  int R, C, nint, flags;
  int A, tile_base, ntiles, n_rate;
};

// One run of a model-record upload: `len` bytes from offset `src` of the
// device staging copy to device address `dst`.
struct SbScatterRun {
  unsigned long long dst, src, len;
};

// Device memory of the no-allocation Counts entry point: one static array of
// SB_NOALLOC_BYTES per GPU, split into the launch plan (header first), the
// tile counters, the branch records and the per-wave scratch slots.
#define SB_NOALLOC_BYTES (256ull << 20)
#define SB_NOALLOC_MAX_DEVICES 64
#define SB_NA_PLAN_OFF 0ull
#define SB_NA_PLAN_BYTES (4ull << 20)
#define SB_NA_COUNTER_OFF (4ull << 20)
#define SB_NA_BT_OFF ((4ull << 20) + (64ull << 10))
#define SB_NA_BT_BYTES ((12ull << 20) - (64ull << 10))
#define SB_NA_WAVES_OFF (16ull << 20)
#define SB_NA_WAVES_BYTES (SB_NOALLOC_BYTES - SB_NA_WAVES_OFF)

struct SbBranchJob {       // one per (block, node): M(t), J(t) ... in FP64
  const double* model64;
  double t;
};

// Launch header: lives at a fixed device address per launcher, so the kernel
// sequence of one variant can be captured in a hipGraph once and replayed for
// any batch (the header and the plan are re-uploaded before each replay).
struct SbLaunch {
  const SbBlockDesc* blocks;
  const int2* tiles;       // {block, tile-in-block}, longest trees first
  const SbBranchJob* jobs; // one per (block, node)
  void* bt;                // branch records, filled by the params kernel
  int* counter;            // SB_NCOUNTER dynamic tile counters (zeroed by the params / reset kernels)
  char* scratch;           // nwaves * wave_stride bytes
  double* grad_part;       // [tile][grad_stride]: per-tile branch-length partials
  double* rate_part;       // [tile][SB_MAX_RATE]
  void* stage;             // two-pass general path: per-tile node vectors
  long long wave_stride;
  long long stage_stride;  // bytes of stage per tile (two-pass general path)
  int nblocks, ntiles, njobs, nint_max;
  int grad_stride, AP, tile_cols, rmax;   // rmax: largest tree (nodes) in the launch
  // small-alphabet kernels: rows of a wave's LDS token stage (a multiple of
  // 32, >= rmax; the tile's R x 64 tokens are copied there at tile start),
  // 0 = read tokens from the alignment in HBM
};
