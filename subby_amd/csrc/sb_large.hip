// Large-alphabet path (padded alphabet AP = 16, 20, 32 or 64), gfx950.
//
// One wave64 owns a tile of TC columns of one block; lane a holds component a
// of every state vector of those columns, so the inside pass
// (pruning.py:94-118), the outside pass (outside.py:94-126), the eigenbasis
// projection (eigensub.py:53-84) and the accumulation (eigensub.py:87-135 or
// f81_fast.py:7-141) run lane-parallel over states and loop over the tree.
// Products with an AP x AP matrix stage the matrix in LDS (one coalesced row
// per step, loads issued in batches) and broadcast the vector being contracted
// from LDS.  Lane k keeps row k of the per-column A x A accumulator in
// registers.  Per-column rescale maxima and dot products are DPP wavefront
// reductions.  The next tree node's child / parent data is fetched one
// iteration ahead, and a vector computed for the next node is forwarded in
// registers.
// Rescaling, Steiner flags and scales follow the small-alphabet kernel:
// exact power-of-two rescaling with integer exponents, and
// scale = ldexp(1/s0, eD + eU - e0).
// For F81-structured models M(t) = e I + (1 - e) 1 pi^T is applied in O(A)
// (the same matrix, evaluated in closed form).
#include <hip/hip_runtime.h>

#include "../../include/subby_amd.h"
#include "sb_plan.h"

#define SB_CONST __attribute__((address_space(4)))
#define SB_GLOBAL __attribute__((address_space(1)))

namespace {

template <typename T>
__device__ inline const SB_CONST T* cst(const void* p) { return (const SB_CONST T*)(p); }
template <typename T>
__device__ inline SB_GLOBAL T* glb(void* p) { return (SB_GLOBAL T*)(p); }
template <typename T>
__device__ inline const SB_GLOBAL T* glb(const void* p) { return (const SB_GLOBAL T*)(p); }

__device__ inline int frexp_exp(float x) { return __builtin_amdgcn_frexp_expf(x); }
__device__ inline int frexp_exp(double x) { return __builtin_amdgcn_frexp_exp(x); }
__device__ inline float ldexp_(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
__device__ inline double ldexp_(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }

// Wavefront reductions with DPP: quad swaps, half-row and row mirrors give every
// lane its 16-lane row total; the four row totals are combined via readlane.
template <int CTRL>
__device__ inline int dpp_i(int x) { return __builtin_amdgcn_mov_dpp(x, CTRL, 0xf, 0xf, true); }
template <int CTRL>
__device__ inline float dpp(float x) { return __int_as_float(dpp_i<CTRL>(__float_as_int(x))); }
template <int CTRL>
__device__ inline double dpp(double x) {
  const int lo = dpp_i<CTRL>(__double2loint(x)), hi = dpp_i<CTRL>(__double2hiint(x));
  return __hiloint2double(hi, lo);
}
__device__ inline float rl(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }
__device__ inline double rl(double x, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                          __builtin_amdgcn_readlane(__double2loint(x), l));
}

template <typename T>
__device__ inline T wsum(T v) {
  v += dpp<0xb1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4e>(v);   // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);  // row_half_mirror
  v += dpp<0x140>(v);  // row_mirror
  return (rl(v, 0) + rl(v, 16)) + (rl(v, 32) + rl(v, 48));
}
template <typename T>
__device__ inline T wmax(T v) {
  v = fmax(v, dpp<0xb1>(v));
  v = fmax(v, dpp<0x4e>(v));
  v = fmax(v, dpp<0x141>(v));
  v = fmax(v, dpp<0x140>(v));
  return fmax(fmax(rl(v, 0), rl(v, 16)), fmax(rl(v, 32), rl(v, 48)));
}

// Output variants of the kernel (see sb_large_kernel).
enum { LOUT_COUNTS = 0, LOUT_POST = 1, LOUT_BRANCH = 2, LOUT_IRR = 3 };

// Alphabet tile width for a padded alphabet: 20, 32 or 64 states.
inline constexpr int large_aw(int AP) { return AP <= 20 ? 20 : AP <= 32 ? 32 : 64; }

// Columns per wave tile: 4, except where the A x A accumulator is widest
// (FP64, or the general 64-state FP32 model) where 2 keeps it in registers;
// FP32 per-branch counts take 2 (three waves per SIMD instead of one); the
// irreversible counts hold complex temporaries beside the accumulator (1
// column where 2 would not fit).  Host and kernel share this rule.
__host__ __device__ inline constexpr int large_tc_of(bool dbl, int AW, bool f81, int out) {
  return (out == LOUT_IRR && (dbl || AW == 64)) ? 1
         : (out == LOUT_BRANCH && !f81 && !dbl && AW <= 32) ? 2
         : (dbl || (AW == 64 && !f81)) ? 2
         : 4;
}


// Per-lane view of TC column vectors: v[c] = component `lane` of column c.
template <typename T, int TC>
struct Vec {
  T v[TC];
};

// Child data as loaded: internal node -> vector + word per column, leaf -> token per column.
template <typename T, int TC>
struct Raw {
  T u[TC];
  int w[TC];
};

template <int TC>
struct Flags {
  int e[TC];
  bool has[TC], branch[TC];
  bool leaf;
};

// Stage P (AP x AP, read as P[j*AP + lane]) in LDS as sm[j*64 + lane].
template <typename T>
__device__ inline void stage(const SB_GLOBAL T* P, int AP, T* sm, int lane) {
  const bool on = lane < AP;
  for (int j0 = 0; j0 < AP; j0 += 16) {
    T r[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) r[q] = (on && j0 + q < AP) ? P[(j0 + q) * AP + lane] : T(0);
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (j0 + q < AP) sm[(j0 + q) * SB_WAVE + lane] = r[q];
  }
}

// w_c[lane] = sum_j P[j*AP + lane] * u_c[j]; lanes >= AP give 0.
template <typename T, int TC>
__device__ inline void contract(const SB_GLOBAL T* P, int AP, const Vec<T, TC>& u, Vec<T, TC>& w, T* bc, T* sm,
                                int lane) {
  __syncthreads();
  stage<T>(P, AP, sm, lane);
#pragma unroll
  for (int c = 0; c < TC; ++c) bc[c * SB_WAVE + lane] = u.v[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < TC; ++c) w.v[c] = T(0);
  for (int j = 0; j < AP; j += 4) {
    const T m0 = sm[(j + 0) * SB_WAVE + lane];
    const T m1 = sm[(j + 1) * SB_WAVE + lane];
    const T m2 = sm[(j + 2) * SB_WAVE + lane];
    const T m3 = sm[(j + 3) * SB_WAVE + lane];
#pragma unroll
    for (int c = 0; c < TC; ++c) {
      const T* b = bc + c * SB_WAVE + j;
      w.v[c] = fma(m0, b[0], w.v[c]);
      w.v[c] = fma(m1, b[1], w.v[c]);
      w.v[c] = fma(m2, b[2], w.v[c]);
      w.v[c] = fma(m3, b[3], w.v[c]);
    }
  }
}

// w = M u for an F81-structured M: w_a = e u_a + (1 - e) (pi . u) on real states.
template <typename T, int TC>
__device__ inline void f81_matvec(T e, T pia, T onea, const Vec<T, TC>& u, Vec<T, TC>& w) {
  T pu[TC];
#pragma unroll
  for (int c = 0; c < TC; ++c) pu[c] = wsum(pia * u.v[c]);
#pragma unroll
  for (int c = 0; c < TC; ++c) w.v[c] = onea * fma(T(1) - e, pu[c], e * u.v[c]);
}

// x = d^T M for an F81-structured M: x_a = e d_a + (1 - e) pi_a sum_i d_i.
template <typename T, int TC>
__device__ inline void f81_vecmat(T e, T pia, const Vec<T, TC>& d, Vec<T, TC>& x) {
  T sd[TC];
#pragma unroll
  for (int c = 0; c < TC; ++c) sd[c] = wsum(d.v[c]);
#pragma unroll
  for (int c = 0; c < TC; ++c) x.v[c] = fma(T(1) - e, pia * sd[c], e * d.v[c]);
}

template <typename T, int TC>
__device__ inline void rescale(Vec<T, TC>& w, int* e) {
#pragma unroll
  for (int c = 0; c < TC; ++c) {
    const int x = frexp_exp(wmax(w.v[c]));
    e[c] = x;
    w.v[c] = ldexp_(w.v[c], -x);
  }
}

// Next column tile of a wave.  The tile list is cut into SB_NXCD contiguous
// ranges with a queue each; a wave starts on the range of its own XCD
// (workgroups are dispatched to the XCDs round-robin by id), so the waves
// that run together on one XCD take neighbouring tiles and the few columns
// each writes into an output row [..][C] fill whole cache lines in that XCD's
// L2 before write-back.  An exhausted range sends the wave on to the next.
// `k` (initially 0) is the wave's number of exhausted ranges; -1 = no tiles.
__device__ inline int claim_tile(int* counter, int ntiles, int lane, int& k) {
  while (k < SB_NXCD) {
    const int r = (int(blockIdx.x) + k) % SB_NXCD;
    const int lo = int((long long)ntiles * r / SB_NXCD), hi = int((long long)ntiles * (r + 1) / SB_NXCD);
    int t = 0;
    if (lane == 0) t = atomicAdd(counter + r * SB_XCD_GAP, 1);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, SB_WAVE));
    if (lo + t < hi) return lo + t;
    ++k;
  }
  return -1;
}

}  // namespace

// OUT selects the outputs beside logLike and root:
//  LOUT_COUNTS  counts summed over branches (and the gradient when GRAD);
//  LOUT_POST    node and branch posteriors (InsideOutside queries,
//               subby/jax/__init__.py:800-1040): per branch n, with outside
//               vector d, inside vector u and M = M(t_n), node posterior
//               (d^T M) * u / z and branch joint posterior d_i M_ij u_j / z,
//               z = d . (M u) (the scales cancel);
//  LOUT_BRANCH  per-branch counts (BranchCounts, subby/jax/__init__.py:450-540):
//               the epilogue runs after every branch into that branch's slice;
//  LOUT_IRR     counts or per-branch counts of irreversible models (complex
//               eigensystem, eigensub.py:234-434): each branch's
//               Re(W o (L C_n Rt)) is added to natural-basis accumulators
//               (lane = column j of the counts, acc[c][i] = row i).
template <typename T, int AW, bool F81, bool GRAD, int OUT = LOUT_COUNTS>
__global__ void __launch_bounds__(64) sb_large_kernel(const SbLaunch* __restrict__ Lp) {
  constexpr bool POST = OUT == LOUT_POST, BRC = OUT == LOUT_BRANCH, IRC = OUT == LOUT_IRR;
  static_assert(!(GRAD && OUT != LOUT_COUNTS && OUT != LOUT_IRR) && !(F81 && (POST || IRC)), "unsupported variant");
  // irreversible counts hold their accumulators beside complex temporaries:
  // one column per wave where 2 would not fit the registers
  constexpr int TC = large_tc_of(sizeof(T) == 8, AW, F81, OUT);
  typedef Vec<T, TC> V_;
  typedef Raw<T, TC> R_;
  typedef Flags<TC> F_;
  const SbLaunch& L = *Lp;
  __shared__ T bc[TC * SB_WAVE];
  // rows j < AP <= AW of 64 floats: the staged AP x AP matrix, and the epilogue's
  // transposition buffer (rows of lanes < AW)
  __shared__ T sm[AW * 64];
  __shared__ T sv[(F81 || POST) ? 1 : AW * 64];   // V^T for the back-transform
  __shared__ T px[POST ? TC * SB_WAVE : 1];        // posteriors: d_i of each column
  __shared__ T pnode[POST ? TC * SB_WAVE : 1];     // posteriors: the node row
  __shared__ T dim_[IRC ? TC * SB_WAVE : 1];       // irreversible: Im of the projected outside vectors
  __shared__ T yrow[IRC ? 2 * SB_WAVE : 1];         // irreversible: one row of L C_n (re | im)
  T* accb = sm;
  const int lane = threadIdx.x;
  const int gw = blockIdx.x;
  const int AP = L.AP;
  const int NI = L.nint_max;
  const long long vbytes = (long long)NI * TC * SB_WAVE * sizeof(T);
  SB_GLOBAL char* wscr = glb<char>(L.scratch) + (long long)gw * L.wave_stride;
  SB_GLOBAL T* Ureg = (SB_GLOBAL T*)wscr;
  SB_GLOBAL T* Dreg = (SB_GLOBAL T*)(wscr + vbytes);
  SB_GLOBAL int* Uw = (SB_GLOBAL int*)(wscr + 2 * vbytes);
  SB_GLOBAL int* Dw = Uw + NI * TC;

  int kx = 0;
  for (;;) {
    const int t = claim_tile(L.counter, L.ntiles, lane, kx);
    if (t < 0) break;
    const SB_CONST int* tbp = cst<int>(L.tiles) + 2 * t;
    const int bidx = tbp[0], tidx = tbp[1];
    const SB_CONST SbBlockDesc* B = cst<SbBlockDesc>(L.blocks) + bidx;
    const long long C = B->C;
    const int A = B->A;
    const int flags = B->flags;
    const int nint = B->nint;
    long long cc[TC];
    bool valid[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c) {
      const long long col = (long long)tidx * TC + c;
      valid[c] = col < C;
      cc[c] = valid[c] ? col : C - 1;
    }
    const SB_GLOBAL int8_t* aln = glb<int8_t>(B->aln);
    const SB_GLOBAL T* bt = glb<T>(B->bt);
    const SB_GLOBAL T* md = glb<T>(B->model);
    const SB_CONST T* mdc = cst<T>(B->model);
    const SB_CONST int* sched = cst<int>(B->sched);
    const bool on = lane < AP;
    // irreversible models (log-likelihood, root and posteriors only): their
    // record layouts; M and its transpose are real
    const bool irr = (flags & SB_F_IRR) != 0;
    const int BTS = sb_bt_stride_k(AP, irr);
    const int o_Mt = irr ? sb_bti_Mt(AP) : sb_bt_Mt(AP);
    const int o_pi = sb_mr_pi_k(AP, irr), o_misc = irr ? sb_mri_misc(AP) : sb_mr_misc(AP);
    const T onea = on ? md[sb_mr_ones_k(AP, irr) + lane] : T(0);
    const T pia = on ? md[o_pi + lane] : T(0);
    const T pra = (on && B->root_model) ? glb<T>(B->root_model)[o_pi + lane] : pia;   // root prior
    // F81-structured M (never for irreversible or per-branch models)
    const bool f81m = mdc[o_misc + 1] != T(0) && B->row_models == nullptr;
    const int kstar = int(mdc[o_misc + 2]);

    auto apply_M = [&](int n, const V_& u, V_& w) {
      const SB_GLOBAL T* bn = bt + (long long)n * BTS;
      if (f81m) f81_matvec<T, TC>(bn[sb_bt_E(AP) + kstar], pia, onea, u, w);
      else contract<T, TC>(bn + o_Mt, AP, u, w, bc, sm, lane);
    };
    auto tokens = [&](int n, int* tok) {
#pragma unroll
      for (int c = 0; c < TC; ++c) tok[c] = aln[(long long)n * C + cc[c]];
    };
    auto token_vec = [&](const int* tok, V_& u) {
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        const bool obs = tok[c] >= 0 && tok[c] < A;
        u.v[c] = obs ? (lane == tok[c] ? T(1) : T(0)) : onea;
      }
    };
    auto load_raw = [&](R_& r, int n, int slot, const SB_GLOBAL T* vreg, const SB_GLOBAL int* wreg) {
      if (slot < 0) {
        tokens(n, r.w);
      } else {
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          r.u[c] = vreg[((long long)slot * TC + c) * SB_WAVE + lane];
          r.w[c] = wreg[slot * TC + c];
        }
      }
    };
    auto decode = [&](const R_& r, int slot, V_& u, F_& f) {
      if (slot < 0) {
        token_vec(r.w, u);
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          f.e[c] = 0;
          f.has[c] = r.w[c] >= 0 && r.w[c] <= A;
          f.branch[c] = false;
        }
        f.leaf = true;
      } else {
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          u.v[c] = r.u[c];
          f.e[c] = r.w[c] >> 2;
          f.has[c] = (r.w[c] & 1) != 0;
          f.branch[c] = (r.w[c] & 2) != 0;
        }
        f.leaf = false;
      }
    };

    // ---------------- inside pass (postorder, one node ahead) ----------------
    V_ u0;
    int e0[TC];
    bool rbr[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c) { e0[c] = 0; rbr[c] = false; }
    if (nint == 0) {
      int tok[TC];
      tokens(0, tok);
      token_vec(tok, u0);
    }
    R_ nl, nr;
    int np_ = 0, nl_ = 0, nr_ = 0, nsl = -1, nsr = -1;
    if (nint > 0) {
      const SB_CONST int* rec = sched + (nint - 1) * 8;
      np_ = rec[0]; nl_ = rec[1]; nr_ = rec[2]; nsl = rec[4]; nsr = rec[5];
      load_raw(nl, nl_, nsl, Ureg, Uw);
      load_raw(nr, nr_, nsr, Ureg, Uw);
    }
    for (int i = nint - 1; i >= 0; --i) {
      const int p = np_, l = nl_, r = nr_, sl = nsl, sr = nsr;
      const R_ rl_ = nl, rr_ = nr;
      int tokp[TC];
      tokens(p, tokp);
      bool fwd_l = false, fwd_r = false;
      if (i > 0) {
        const SB_CONST int* rec = sched + (i - 1) * 8;
        np_ = rec[0]; nl_ = rec[1]; nr_ = rec[2]; nsl = rec[4]; nsr = rec[5];
        fwd_l = nsl == i;
        fwd_r = nsr == i;
        if (!fwd_l) load_raw(nl, nl_, nsl, Ureg, Uw);
        if (!fwd_r) load_raw(nr, nr_, nsr, Ureg, Uw);
      }
      V_ u, wl, up;
      F_ fl, fr;
      decode(rl_, sl, u, fl);
      apply_M(l, u, wl);
      decode(rr_, sr, u, fr);
      apply_M(r, u, up);
      token_vec(tokp, u);
#pragma unroll
      for (int c = 0; c < TC; ++c) up.v[c] = u.v[c] * wl.v[c] * up.v[c];
      int e[TC];
      rescale<T, TC>(up, e);
      int word[TC];
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        const int ep = fl.e[c] + fr.e[c] + e[c];
        word[c] = (ep << 2) | ((fl.has[c] && fr.has[c]) ? 2 : 0) | ((fl.has[c] || fr.has[c]) ? 1 : 0);
      }
      if (i > 0) {
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          Ureg[((long long)i * TC + c) * SB_WAVE + lane] = up.v[c];
          if (lane == 0) Uw[i * TC + c] = word[c];
        }
        if (fwd_l) {
#pragma unroll
          for (int c = 0; c < TC; ++c) { nl.u[c] = up.v[c]; nl.w[c] = word[c]; }
        }
        if (fwd_r) {
#pragma unroll
          for (int c = 0; c < TC; ++c) { nr.u[c] = up.v[c]; nr.w[c] = word[c]; }
        }
      } else {
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          u0.v[c] = up.v[c];
          e0[c] = fl.e[c] + fr.e[c] + e[c];
          rbr[c] = fl.has[c] && fr.has[c];
        }
      }
    }
    // log-likelihood, root posterior
    const SB_CONST double* pi64 =
        B->root_model64 ? cst<double>(B->root_model64) + sb_m64_pi_k(AP, irr)
                        : cst<double>(B->model64) + sb_m64_pi_k(AP, irr);
    T inv_s0[TC];
    double s0s[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c) s0s[c] = wsum((lane < A) ? pi64[lane < A ? lane : 0] * double(u0.v[c]) : 0.0);
#pragma unroll
    for (int c = 0; c < TC; ++c) {
      const double ll = double(e0[c]) * 0.69314718055994530942 + log(s0s[c]);
      inv_s0[c] = T(1.0 / s0s[c]);
      if (valid[c] && lane == 0 && (flags & SB_F_LL) && B->loglike) glb<double>(B->loglike)[cc[c]] = ll;
      if (valid[c] && lane < A && (flags & SB_F_ROOT) && B->root)
        glb<T>(B->root)[(long long)lane * C + cc[c]] = pia * u0.v[c] * inv_s0[c];
    }
    SB_GLOBAL T* npost = (POST && (flags & SB_F_NPOST)) ? glb<T>(B->npost) : nullptr;
    SB_GLOBAL T* bpost = (POST && (flags & SB_F_BPOST)) ? glb<T>(B->bpost) : nullptr;
    // POST outputs of node / branch n from LDS: pnode[i*TC + c] (node row n),
    // px[i*TC + c] = d_i, bc[j*TC + c] = u_j / z and mij(i, j) (branch n > 0).
    // Lanes run over (row, column) so that each store writes TC consecutive
    // columns of 64 / TC rows [..][C].
    const long long col0 = (long long)tidx * TC;
    const int ncol = int(C - col0 < TC ? C - col0 : TC);
    const float invA = 1.0f / float(A);
    auto post_store = [&](int n, auto mij) {
      if (npost)
        for (int f = lane; f < A * TC; f += SB_WAVE) {
          const int i = f / TC, c = f % TC;
          if (c < ncol) npost[((long long)n * A + i) * C + col0 + c] = pnode[f];
        }
      if (bpost)
        for (int f = lane; f < A * A * TC; f += SB_WAVE) {
          const int r = f / TC, c = f % TC;
          const int i = int((float(r) + 0.5f) * invA), j = r - i * A;   // exact: r < 4096, A <= 64
          if (c < ncol)
            bpost[((long long)n * A * A + r) * C + col0 + c] = n == 0 ? T(0) : px[i * TC + c] * mij(i, j) * bc[j * TC + c];
        }
    };
    // BRC: per-branch slices of B->bcounts, or (per-branch models, Counts) the
    // sum over branches accumulated in B->counts
    const bool rows = B->row_models != nullptr;
    const bool bsum = BRC && rows && (flags & SB_F_COUNTS) && B->counts;
    SB_GLOBAL T* bcnt = !BRC ? nullptr
                        : bsum ? glb<T>(B->counts)
                        : ((flags & SB_F_BCOUNTS) && B->bcounts) ? glb<T>(B->bcounts) : nullptr;
    // IRC: per-branch slices (BranchCounts) or the sum (Counts)
    SB_GLOBAL T* icnt = !IRC ? nullptr
                        : ((flags & SB_F_BCOUNTS) && B->bcounts) ? glb<T>(B->bcounts)
                        : ((flags & SB_F_COUNTS) && B->counts) ? glb<T>(B->counts) : nullptr;
    const bool islice = IRC && (flags & SB_F_BCOUNTS);
    if constexpr (POST) {
      // row 0: the root posterior (under the model's own root distribution);
      // branch 0 (above the root) is zero
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        const T inv_own = B->root_model ? T(1) / wsum(pia * u0.v[c]) : inv_s0[c];
        pnode[lane * TC + c] = pia * u0.v[c] * inv_own;
      }
      __syncthreads();
      post_store(0, [&](int, int) { return T(0); });
      if (!(npost || bpost)) continue;
    } else if constexpr (IRC) {
      if (!icnt && !GRAD) continue;
      if (islice)
#pragma unroll
        for (int c = 0; c < TC; ++c) {    // branch 0 (above the root) holds no counts
          if (!valid[c] || lane >= A) continue;
          for (int jj = 0; jj < A; ++jj) icnt[((long long)lane * A + jj) * C + cc[c]] = T(0);
        }
    } else if constexpr (BRC) {
      if (!bcnt) continue;
#pragma unroll
      for (int c = 0; c < TC; ++c) {      // branch 0 (above the root) holds no counts; a sum starts at 0
        if (!valid[c] || lane >= A) continue;
        for (int jj = 0; jj < A; ++jj) bcnt[((long long)lane * A + jj) * C + cc[c]] = T(0);
      }
    } else {
      if (!(flags & (SB_F_COUNTS | SB_F_GRAD))) continue;
    }

    // ---------------- outside pass + accumulation (preorder, one node ahead) ----------------
    const int mask_mode = (flags >> SB_F_MASK_SHIFT) & 3;
    const SB_GLOBAL uint8_t* emask = B->mask ? glb<uint8_t>(B->mask) : nullptr;
    double gcol[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c)
      gcol[c] = GRAD ? (valid[c] ? (B->g ? glb<double>(B->g)[cc[c]] : 1.0) : 0.0) : 0.0;
    SB_GLOBAL double* gpart =
        GRAD ? glb<double>(L.grad_part) + (long long)(B->tile_base + tidx) * L.grad_stride : nullptr;

    T acc[TC][AW];
#pragma unroll
    for (int c = 0; c < TC; ++c)
#pragma unroll
      for (int q = 0; q < AW; ++q) acc[c][q] = T(0);
    T t2a[TC], t2b[TC], t3[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c) { t2a[c] = T(0); t2b[c] = T(0); t3[c] = T(0); }
    // acc (lane = column j, acc[c][i] = row i) into out[i][j][C]
    auto irc_store = [&](SB_GLOBAL T* out) {
#pragma unroll
      for (int c = 0; c < TC; ++c)
#pragma unroll
        for (int i = 0; i < AW; ++i)
          if (i < A && valid[c] && lane < A) out[((long long)i * A + lane) * C + cc[c]] = acc[c][i];
    };
    // model record of branch n (per-branch models: its own)
    auto model_of = [&](int n) -> const SB_GLOBAL T* { return rows ? glb<T>(B->row_models[n]) : md; };

    // ---- epilogue: natural-basis counts per column from acc (and t2a/t2b/t3)
    // into out[A][A][C] (null: none); GRAD adds the rate-parameter partials
    const T fmu = mdc[sb_mr_misc(AP)];
    // sv[k*64 + lane] = V[lane][k] of model record mv
    auto stage_V = [&](const SB_GLOBAL T* mv) {
      __syncthreads();
      const SB_GLOBAL T* Vg = mv + sb_mr_V(AP);
      for (int k = 0; k < AP; ++k) sv[k * SB_WAVE + lane] = on ? Vg[lane * AP + k] : T(0);
    };
    if constexpr (!F81 && !POST && !IRC) stage_V(md);
    double rsum[SB_MAX_RATE];
#pragma unroll
    for (int q = 0; q < SB_MAX_RATE; ++q) rsum[q] = 0.0;
    // out: written, or added to when `add` (the wave owns these columns)
    auto epilogue = [&](SB_GLOBAL T* out, const void* me, bool add) {
      const SB_CONST T* Vc = cst<T>(me) + sb_mr_V(AP);   // uniform reads: scalar loads
      const SB_GLOBAL T* S = glb<T>(me) + sb_mr_S(AP);
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        // natural-basis counts of column c into accb[j*64 + i] (i = lane = row)
        if constexpr (F81) {
          __syncthreads();
          bc[lane] = t2b[c];
          __syncthreads();
#pragma unroll
          for (int j = 0; j < AW; ++j) {
            const T I = acc[c][j] + t2a[c] + pia * (bc[j] + t3[c]);
            const T pj = md[sb_mr_pi(AP) + (j < AP ? j : 0)];
            accb[j * SB_WAVE + lane] = (j == lane) ? I : fmu * pj * I;
          }
        } else {
          __syncthreads();
#pragma unroll
          for (int q = 0; q < AW; ++q)
            if (lane < AW) accb[lane * 64 + q] = acc[c][q];   // lanes >= AW hold zeros
          __syncthreads();
          T tmp[AW];
#pragma unroll
          for (int q = 0; q < AW; ++q) tmp[q] = T(0);
          for (int k = 0; k < AP; ++k) {
            const T vik = sv[k * SB_WAVE + lane];
#pragma unroll
            for (int q = 0; q < AW; ++q) tmp[q] = fma(vik, accb[k * 64 + q], tmp[q]);
          }
          __syncthreads();
          // tmp[q] = 0 for q >= AP (acc holds zeros there), so the row of V is
          // read unguarded up to AW (into the record's next matrix, finite
          // values): contiguous uniform reads the compiler merges into wide
          // scalar loads
          for (int jj = 0; jj < AP; ++jj) {
            T sacc = T(0);
#pragma unroll
            for (int q = 0; q < AW; ++q) sacc = fma(tmp[q], Vc[jj * AP + q], sacc);
            accb[jj * SB_WAVE + lane] = (jj == lane) ? sacc : (on ? S[jj * AP + lane] : T(0)) * sacc;   // S symmetric
          }
        }
        if (valid[c] && out && lane < A)
          for (int jj = 0; jj < A; ++jj) {
            SB_GLOBAL T* o = out + ((long long)lane * A + jj) * C + cc[c];
            *o = add ? *o + accb[jj * SB_WAVE + lane] : accb[jj * SB_WAVE + lane];
          }
        if constexpr (GRAD) {
          for (int q = 0; q < B->n_rate; ++q) {
            const SB_CONST double* W = cst<double>(B->rate_w) + (long long)q * AP * AP;
            double sq = 0.0;
            if (on)
              for (int jj = 0; jj < AP; ++jj) sq = fma(W[lane * AP + jj], double(accb[jj * SB_WAVE + lane]), sq);
            const double v = gcol[c] * wsum(sq);
#pragma unroll
            for (int qq = 0; qq < SB_MAX_RATE; ++qq)
              if (qq == q) rsum[qq] += v;
          }
        }
      }
    };

    R_ ql, qr, qd;
    int qp = 0, qls = 0, qrs = 0, qsl = -1, qsr = -1;
    if (nint > 0) {
      qp = sched[0]; qls = sched[1]; qrs = sched[2]; qsl = sched[4]; qsr = sched[5];
      load_raw(ql, qls, qsl, Ureg, Uw);
      load_raw(qr, qrs, qsr, Ureg, Uw);
    }
    for (int i = 0; i < nint; ++i) {
      const int p = qp, l = qls, r = qrs, sl = qsl, sr = qsr;
      const R_ rl_ = ql, rr_ = qr, rd_ = qd;
      int tokp[TC];
      tokens(p, tokp);
      bool fwd0 = false, fwd1 = false;
      if (i + 1 < nint) {
        const SB_CONST int* rec = sched + (i + 1) * 8;
        qp = rec[0]; qls = rec[1]; qrs = rec[2]; qsl = rec[4]; qsr = rec[5];
        fwd0 = qp == l;
        fwd1 = qp == r;
        if (!fwd0 && !fwd1) load_raw(qd, 0, i + 1, Dreg, Dw);
        load_raw(ql, qls, qsl, Ureg, Uw);
        load_raw(qr, qrs, qsr, Ureg, Uw);
      }
      V_ x;
      int ex[TC];
      bool stp[TC];
      {
        V_ o;
        token_vec(tokp, o);
        if (i == 0) {
#pragma unroll
          for (int c = 0; c < TC; ++c) { x.v[c] = o.v[c] * pra; ex[c] = 0; stp[c] = rbr[c]; }
        } else {
          V_ dp;
#pragma unroll
          for (int c = 0; c < TC; ++c) {
            dp.v[c] = rd_.u[c];
            ex[c] = rd_.w[c] >> 1;
            stp[c] = (rd_.w[c] & 1) != 0;
          }
          const SB_GLOBAL T* bp = bt + (long long)p * BTS;
          if (f81m) f81_vecmat<T, TC>(bp[sb_bt_E(AP) + kstar], pia, dp, x);
          else contract<T, TC>(bp + sb_bt_M(AP), AP, dp, x, bc, sm, lane);
#pragma unroll
          for (int c = 0; c < TC; ++c) x.v[c] *= o.v[c];
        }
      }
#pragma unroll 1
      for (int s = 0; s < 2; ++s) {
        const int n = s == 0 ? l : r, sn = s == 0 ? sl : sr;
        const int sib = s == 0 ? r : l, ssib = s == 0 ? sr : sl;
        V_ u, d;
        F_ fs, fn;
        decode(s == 0 ? rr_ : rl_, ssib, u, fs);
        apply_M(sib, u, d);
        int e[TC], ed[TC];
#pragma unroll
        for (int c = 0; c < TC; ++c) d.v[c] *= x.v[c];
        rescale<T, TC>(d, e);
        decode(s == 0 ? rl_ : rr_, sn, u, fn);
        bool act[TC];
        T sc[TC], scr[TC];
        const bool fwd = s == 0 ? fwd0 : fwd1;
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          ed[c] = fs.e[c] + ex[c] + e[c];
          const bool st = (fn.leaf ? fn.has[c] : fn.branch[c]) || (stp[c] && fn.has[c]);
          const int dword = (ed[c] << 1) | (st ? 1 : 0);
          if (sn >= 0 && !fwd) {     // a forwarded outside vector is never read back
            Dreg[((long long)sn * TC + c) * SB_WAVE + lane] = d.v[c];
            if (lane == 0) Dw[sn * TC + c] = dword;
          }
          if (fwd) {
            qd.u[c] = d.v[c];
            qd.w[c] = dword;
          }
          if (GRAD || mask_mode == SB_MASK_NONE) act[c] = true;
          else if (mask_mode == SB_MASK_AUTO) act[c] = st && stp[c];
          else act[c] = emask[(long long)n * C + cc[c]] != 0;
          scr[c] = ldexp_(inv_s0[c], ed[c] + fn.e[c] - e0[c]);
          sc[c] = act[c] ? scr[c] : T(0);
        }
        const SB_GLOBAL T* bn = bt + (long long)n * BTS;
        if constexpr (POST) {
          // y = d^T M_n; the contraction leaves M_n staged as sm[i*64 + j] = M[i][j].
          // z = d . (M_n u) = y . u
          V_ y;
          if (f81m) f81_vecmat<T, TC>(bn[sb_bt_E(AP) + kstar], pia, d, y);
          else contract<T, TC>(bn + sb_bt_M(AP), AP, d, y, bc, sm, lane);
          T iz[TC];
#pragma unroll
          for (int c = 0; c < TC; ++c) iz[c] = T(1) / wsum(y.v[c] * u.v[c]);
          __syncthreads();
#pragma unroll
          for (int c = 0; c < TC; ++c) {
            px[lane * TC + c] = d.v[c];
            bc[lane * TC + c] = u.v[c] * iz[c];
            pnode[lane * TC + c] = y.v[c] * u.v[c] * iz[c];
          }
          __syncthreads();
          const T ef = f81m ? bn[sb_bt_E(AP) + kstar] : T(0);
          auto mij = [&](int i, int j) -> T {
            return f81m ? ((i == j ? ef : T(0)) + (T(1) - ef) * md[o_pi + j]) : sm[i * SB_WAVE + j];
          };
          post_store(n, mij);
        } else if constexpr (F81) {
          const T al = bn[sb_bt_coef(AP) + 0], be = bn[sb_bt_coef(AP) + 1], ga = bn[sb_bt_coef(AP) + 2];
          __syncthreads();
#pragma unroll
          for (int c = 0; c < TC; ++c) bc[c * SB_WAVE + lane] = u.v[c];
          __syncthreads();
          T Ds[TC], piU[TC], Dsum[TC];
#pragma unroll
          for (int c = 0; c < TC; ++c) {
            Ds[c] = d.v[c] * sc[c];
            piU[c] = wsum(pia * u.v[c]);
            Dsum[c] = wsum(Ds[c]);
          }
#pragma unroll
          for (int c = 0; c < TC; ++c) {
            const T ad = al * Ds[c];
#pragma unroll
            for (int q = 0; q < AW; ++q)
              if (q < AP) acc[c][q] = fma(ad, bc[c * SB_WAVE + q], acc[c][q]);
            t2a[c] = fma(be * Ds[c], piU[c], t2a[c]);
            t2b[c] = fma(be * Dsum[c], u.v[c], t2b[c]);
            t3[c] = fma(ga * Dsum[c], piU[c], t3[c]);
          }
        } else if constexpr (IRC) {
          // U~ = u Pu, D~ = d Pd (complex; the model's counts convention, set 0;
          // gradients: the exact set 1, vjp.py:80-131)
          const SB_GLOBAL T* mI = model_of(n);
          const SB_GLOBAL T* P0 = mI + sb_mri_set(AP, GRAD ? 1 : 0);
          V_ Ur, Ui, Dr, Di;
          contract<T, TC>(P0 + sb_mri_Pu(AP), AP, u, Ur, bc, sm, lane);
          contract<T, TC>(P0 + sb_mri_Pu(AP) + AP * AP, AP, u, Ui, bc, sm, lane);
          contract<T, TC>(P0 + sb_mri_Pd(AP), AP, d, Dr, bc, sm, lane);
          contract<T, TC>(P0 + sb_mri_Pd(AP) + AP * AP, AP, d, Di, bc, sm, lane);
          if constexpr (GRAD) {
            // d logL / d t_n = scale * Re sum_k D~_k mu_k e^{mu_k t} U~_k  (vjp.py:92-125)
            const T er = on ? bn[sb_bti_dEre(AP) + lane] : T(0), ei = on ? bn[sb_bti_dEim(AP) + lane] : T(0);
            double v = 0.0;
#pragma unroll
            for (int c = 0; c < TC; ++c) {
              const T yr = fma(Dr.v[c], er, -Di.v[c] * ei), yi = fma(Dr.v[c], ei, Di.v[c] * er);
              v += gcol[c] * double(scr[c]) * double(wsum(fma(yr, Ur.v[c], -yi * Ui.v[c])));
            }
            if (lane == 0) gpart[n] = v;
            if (B->n_rate == 0) continue;   // no counts needed
          }
          __syncthreads();
          stage<T>(bn + sb_bti_Jre(AP), AP, sm, lane);   // sm[k*64 + q] = Re J_kq
          stage<T>(bn + sb_bti_Jim(AP), AP, sv, lane);   // sv[k*64 + q] = Im J_kq
#pragma unroll
          for (int c = 0; c < TC; ++c) {
            bc[c * SB_WAVE + lane] = Dr.v[c] * sc[c];
            dim_[c * SB_WAVE + lane] = Di.v[c] * sc[c];
          }
          __syncthreads();
          const SB_CONST T* Lr = cst<T>(rows ? B->row_models[n] : B->model) + sb_mri_set(AP, GRAD ? 1 : 0) + sb_mri_L(AP);
          const SB_CONST T* Li = Lr + AP * AP;
          const SB_GLOBAL T* Rr = P0 + sb_mri_Rt(AP);
          const SB_GLOBAL T* Ri = Rr + AP * AP;
          const SB_GLOBAL T* Wr = mI + sb_mri_W(AP);
          const SB_GLOBAL T* Wi = Wr + AP * AP;
#pragma unroll
          for (int c = 0; c < TC; ++c) {
#pragma unroll
            for (int i = 0; i < AW; ++i) {
              if (i >= AP) continue;
              // (L C_n)[i][q = lane] = U~_q sum_k (L_ik D~_k sc) J_kq
              T sr = T(0), si = T(0);
              for (int k = 0; k < AP; ++k) {
                const T lr = Lr[i * AP + k], li = Li[i * AP + k];
                const T dr = bc[c * SB_WAVE + k], di = dim_[c * SB_WAVE + k];
                const T zr = fma(lr, dr, -li * di), zi = fma(lr, di, li * dr);
                const T jr = sm[k * SB_WAVE + lane], ji = sv[k * SB_WAVE + lane];
                sr = fma(zr, jr, fma(-zi, ji, sr));
                si = fma(zr, ji, fma(zi, jr, si));
              }
              __syncthreads();
              yrow[lane] = fma(sr, Ur.v[c], -si * Ui.v[c]);
              yrow[SB_WAVE + lane] = fma(sr, Ui.v[c], si * Ur.v[c]);
              __syncthreads();
              // (L C_n Rt)[i][j = lane], then counts_ij += Re(W_ij (L C_n Rt)_ij)
              T vr = T(0), vi = T(0);
              for (int q = 0; q < AP; ++q) {
                const T rr = on ? Rr[q * AP + lane] : T(0), ri = on ? Ri[q * AP + lane] : T(0);
                const T yr = yrow[q], yi = yrow[SB_WAVE + q];
                vr = fma(yr, rr, fma(-yi, ri, vr));
                vi = fma(yr, ri, fma(yi, rr, vi));
              }
              const T wr = on ? Wr[i * AP + lane] : T(0), wi = on ? Wi[i * AP + lane] : T(0);
              acc[c][i] += fma(wr, vr, -wi * vi);
            }
          }
          if (islice) {
            irc_store(icnt + (long long)n * A * A * C);
#pragma unroll
            for (int c = 0; c < TC; ++c)
#pragma unroll
              for (int q = 0; q < AW; ++q) acc[c][q] = T(0);
          }
        } else {
          V_ Ut, Dt;
          const SB_GLOBAL T* mn = model_of(n);
          contract<T, TC>(mn + sb_mr_Vs(AP), AP, u, Ut, bc, sm, lane);
          contract<T, TC>(mn + sb_mr_Vi(AP), AP, d, Dt, bc, sm, lane);
          __syncthreads();
          stage<T>(bn + sb_bt_J(AP), AP, sm, lane);   // J is symmetric: row q across lanes
#pragma unroll
          for (int c = 0; c < TC; ++c) bc[c * SB_WAVE + lane] = Ut.v[c];
          __syncthreads();
          T dk[TC];
#pragma unroll
          for (int c = 0; c < TC; ++c) dk[c] = Dt.v[c] * sc[c];
#pragma unroll
          for (int q = 0; q < AW; ++q) {
            if (q < AP) {
              const T jq = sm[q * SB_WAVE + lane];
#pragma unroll
              for (int c = 0; c < TC; ++c) acc[c][q] = fma(dk[c] * jq, bc[c * SB_WAVE + q], acc[c][q]);
            }
          }
          if constexpr (GRAD) {
            const T dE = on ? bn[sb_bt_dE(AP) + lane] : T(0);
            double v = 0.0;
#pragma unroll
            for (int c = 0; c < TC; ++c)
              v += gcol[c] * double(scr[c]) * double(wsum(Dt.v[c] * dE * Ut.v[c]));
            if (lane == 0) gpart[n] = v;
          }
        }
        if constexpr (BRC) {
          if (rows) stage_V(model_of(n));
          epilogue(bsum ? bcnt : bcnt + (long long)n * A * A * C, rows ? B->row_models[n] : B->model, bsum);
#pragma unroll
          for (int c = 0; c < TC; ++c) {
#pragma unroll
            for (int q = 0; q < AW; ++q) acc[c][q] = T(0);
            t2a[c] = T(0); t2b[c] = T(0); t3[c] = T(0);
          }
        }
      }
    }

    if constexpr (IRC) {
      if (!islice && icnt) irc_store(icnt);
      if constexpr (GRAD) {
        // rate parameters: sum_ij W_q[i][j] counts_ij (exact counts), lane = j
        SB_GLOBAL double* rpart = glb<double>(L.rate_part) + (long long)(B->tile_base + tidx) * SB_MAX_RATE;
        for (int q = 0; q < B->n_rate; ++q) {
          const SB_CONST double* W = cst<double>(B->rate_w) + (long long)q * AP * AP;
          double v = 0.0;
#pragma unroll
          for (int c = 0; c < TC; ++c) {
            double sq = 0.0;
#pragma unroll
            for (int i = 0; i < AW; ++i)
              if (i < AP && on) sq = fma(W[i * AP + lane], double(acc[c][i]), sq);
            v += gcol[c] * wsum(sq);
          }
          if (lane == 0) rpart[q] = v;
        }
      }
      continue;
    }
    if constexpr (POST || BRC) continue;
    epilogue(((flags & SB_F_COUNTS) && B->counts) ? glb<T>(B->counts) : nullptr, B->model, false);
    if constexpr (GRAD) {
      SB_GLOBAL double* rpart = glb<double>(L.rate_part) + (long long)(B->tile_base + tidx) * SB_MAX_RATE;
#pragma unroll
      for (int q = 0; q < SB_MAX_RATE; ++q)
        if (lane == 0 && q < B->n_rate) rpart[q] = rsum[q];
    }
  }
}

#define SB_LINST(T, AW)                                                                      \
  template __global__ void sb_large_kernel<T, AW, false, false>(const SbLaunch*);               \
  template __global__ void sb_large_kernel<T, AW, true, false>(const SbLaunch*);                \
  template __global__ void sb_large_kernel<T, AW, false, true>(const SbLaunch*);                \
  template __global__ void sb_large_kernel<T, AW, false, false, LOUT_POST>(const SbLaunch*);    \
  template __global__ void sb_large_kernel<T, AW, false, false, LOUT_BRANCH>(const SbLaunch*);  \
  template __global__ void sb_large_kernel<T, AW, true, false, LOUT_BRANCH>(const SbLaunch*);    \
  template __global__ void sb_large_kernel<T, AW, false, false, LOUT_IRR>(const SbLaunch*);      \
  template __global__ void sb_large_kernel<T, AW, false, true, LOUT_IRR>(const SbLaunch*);
SB_LINST(float, 32)
SB_LINST(float, 64)
SB_LINST(double, 32)
SB_LINST(double, 64)
SB_LINST(float, 20)
SB_LINST(double, 20)

static int large_tc(int dtype, int AP, bool f81, int out) {
  return large_tc_of(dtype != 0, large_aw(AP), f81, out);
}

extern "C" int sb_large_tile_cols(int dtype, int AP, bool f81, int out) { return large_tc(dtype, AP, f81, out); }

extern "C" long long sb_large_wave_stride(int dtype, int AP, int nint_max, bool f81, int out) {
  const long long esz = dtype == 0 ? 4 : 8;
  const int TC = large_tc(dtype, AP, f81, out);
  return (long long)nint_max * (2 * TC * SB_WAVE * esz + 2 * TC * 4);
}

// out: 0 counts (and gradient), 1 posteriors, 2 per-branch counts, 3 irreversible counts
static void (*pick_large(int dtype, int AP, bool f81, bool grad, int out))(const SbLaunch*) {
  void (*k)(const SbLaunch*) = nullptr;
#define SB_LPICK(T, AWV)                                                          \
  {                                                                               \
    if (out == LOUT_POST) k = sb_large_kernel<T, AWV, false, false, LOUT_POST>;  \
    else if (out == LOUT_IRR)                                                     \
      k = grad ? sb_large_kernel<T, AWV, false, true, LOUT_IRR>                   \
               : sb_large_kernel<T, AWV, false, false, LOUT_IRR>;                 \
    else if (out == LOUT_BRANCH)                                                  \
      k = f81 ? sb_large_kernel<T, AWV, true, false, LOUT_BRANCH>                 \
              : sb_large_kernel<T, AWV, false, false, LOUT_BRANCH>;               \
    else if (grad) k = sb_large_kernel<T, AWV, false, true>;                      \
    else if (f81) k = sb_large_kernel<T, AWV, true, false>;                       \
    else k = sb_large_kernel<T, AWV, false, false>;                               \
  }
  if (dtype == 0) {
    if (AP <= 20) SB_LPICK(float, 20) else if (AP <= 32) SB_LPICK(float, 32) else SB_LPICK(float, 64)
  } else {
    if (AP <= 20) SB_LPICK(double, 20) else if (AP <= 32) SB_LPICK(double, 32) else SB_LPICK(double, 64)
  }
#undef SB_LPICK
  return k;
}

extern "C" int sb_large_occupancy(int dtype, int AP, bool f81, bool grad, int out) {
  void (*k)(const SbLaunch*) = pick_large(dtype, AP, f81, grad, out);
  int n = 0;
  if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, 64, 0) != hipSuccess) return 0;
  return n;
}

extern "C" hipError_t sb_launch_large(int dtype, int AP, bool f81, bool grad, int out, int nwaves,
                                      const SbLaunch* L, hipStream_t s) {
  void (*k)(const SbLaunch*) = pick_large(dtype, AP, f81, grad, out);
  if (!k || AP > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3(nwaves), dim3(64), 0, s, L);
  return hipGetLastError();
}
