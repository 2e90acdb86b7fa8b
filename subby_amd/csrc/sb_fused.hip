// Fused inside / outside / count-accumulation kernel for small alphabets
// (padded alphabet AP = 4 or 8), gfx950.
//
// One thread owns one alignment column; one wave64 owns a tile of 64
// consecutive columns of one block and walks the whole tree for it:
//
//   inside pass  (subby/jax/pruning.py:94-118, oracle.py:784-838)
//     internal nodes in postorder; U[p] = obs(p) * (M_l U_l) * (M_r U_r)
//   outside pass (subby/jax/outside.py:94-126, oracle.py:845-924) fused with
//     count accumulation (eigensub.py:87-135 or f81_fast.py:7-141):
//     internal nodes p in preorder; x = obs(p) * (pi or D[p] M_p);
//     D[l] = (M_r U_r) * x, D[r] = (M_l U_l) * x, and both branches are
//     accumulated right away, so leaf vectors never touch memory.
//   epilogue: back-transform V C V^T (eigensub.py:138-172) or the F81
//     combination, root posterior (subby/jax/__init__.py:635-650) and the
//     log-likelihood.
//
// Only internal-node vectors are staged, in a per-wave scratch slot that the
// same wave reads back (stays in L2 / MALL).  All per-branch matrices and the
// tree records are wave-uniform and arrive through scalar loads.
//
// Rescaling: every node vector is scaled by an exact power of two 2^-e with
// e = frexp-exponent of its max (max lands in [0.5, 1)), and the per-node
// log-normaliser is kept as the integer e summed over the subtree.  This is
// the reference's per-node max-rescaling (clamp 1e-300) with an exact
// log-scale: ln-normaliser = e * ln 2.  The log-likelihood is formed in FP64
// and scale(n,c) = exp(lnD + lnU - ll) = ldexp(1/s0, eD + eU - e0) exactly.
#include <hip/hip_runtime.h>
#include "../../include/subby_amd.h"
#include "sb_plan.h"

// Read-only, wave-uniform data (tree records, branch matrices, model) is read
// through the constant address space so it arrives by scalar loads; per-lane
// data goes through the global address space (no flat instructions).
#define SB_CONST __attribute__((address_space(4)))
#define SB_GLOBAL __attribute__((address_space(1)))

namespace {

template <typename T>
__device__ inline const SB_CONST T* cst(const void* p) { return (const SB_CONST T*)(p); }
template <typename T>
__device__ inline SB_GLOBAL T* glb(void* p) { return (SB_GLOBAL T*)(p); }
template <typename T>
__device__ inline const SB_GLOBAL T* glb(const void* p) { return (const SB_GLOBAL T*)(p); }

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename T> struct Chunk;
template <> struct Chunk<float> { typedef f32x4 type; static constexpr int n = 4; };
template <> struct Chunk<double> { typedef f64x2 type; static constexpr int n = 2; };

// A lane offset the compiler cannot hoist out of the walk: a loop-invariant
// per-lane 64-bit address (base + offset) would cost a 64-bit add per access,
// while a uniform base plus this 32-bit offset is the load's own addressing.
__device__ inline unsigned opq(unsigned x) {
  asm volatile("" : "+v"(x));
  return x;
}

__device__ inline int frexp_exp(float x) { return __builtin_amdgcn_frexp_expf(x); }
__device__ inline int frexp_exp(double x) { return __builtin_amdgcn_frexp_exp(x); }
__device__ inline float ldexp_(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
__device__ inline double ldexp_(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }

template <typename T, int AP>
struct Tile {
  static constexpr int EPC = Chunk<T>::n;        // elements per 16-byte chunk
  static constexpr int NCH = AP / EPC;           // chunks per vector
  typedef typename Chunk<T>::type CT;

  // Scratch slot of one wave: U vectors, D vectors, U words, D words.
  static __device__ inline long long vec_bytes(int nint) {
    return (long long)nint * NCH * SB_WAVE * 16;
  }
  // (the slot's base is wave-uniform: a scalar base plus a 32-bit lane offset)
  // (a uniform slot address plus an unsigned 32-bit lane offset: no 64-bit
  // address arithmetic per lane)
  static __device__ inline void store_vec(SB_GLOBAL char* region, int slot, int lane, const T* v) {
    SB_GLOBAL char* p = region + (long long)slot * (NCH * SB_WAVE * 16);
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      CT c;
      if constexpr (EPC == 4) { c.x = v[q * 4]; c.y = v[q * 4 + 1]; c.z = v[q * 4 + 2]; c.w = v[q * 4 + 3]; }
      else { c.x = v[q * 2]; c.y = v[q * 2 + 1]; }
      *(SB_GLOBAL CT*)(p + q * SB_WAVE * 16 + opq(unsigned(lane) * 16u)) = c;
    }
  }
  static __device__ inline void load_vec(const SB_GLOBAL char* region, int slot, int lane, T* v) {
    const SB_GLOBAL char* p = region + (long long)slot * (NCH * SB_WAVE * 16);
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      CT c = *(const SB_GLOBAL CT*)(p + q * SB_WAVE * 16 + opq(unsigned(lane) * 16u));
      if constexpr (EPC == 4) { v[q * 4] = c.x; v[q * 4 + 1] = c.y; v[q * 4 + 2] = c.z; v[q * 4 + 3] = c.w; }
      else { v[q * 2] = c.x; v[q * 2 + 1] = c.y; }
    }
  }
};

// Likelihood vectors of tokens by byte masks: word w of a token's mask holds
// one byte per state 4w..4w+3 (1 if the state is possible), so each element
// is one unsigned-byte-to-float conversion.  Observed tokens 0..A-1 give one
// byte; every other token gives the real states (padded states stay zero).
template <typename T, int AP>
struct TokDec {
  static constexpr int NW = (AP + 3) / 4;
  unsigned unobs[NW];
  int A;
  __device__ inline void init(int A_) {
    A = A_;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      unsigned m = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) m |= (4 * w + b < A_ ? 1u : 0u) << (8 * b);
      unobs[w] = m;
    }
  }
  __device__ inline void decode(int tok, T* u) const {
    const bool obs = unsigned(tok) < unsigned(A);
    const unsigned one = 1u << (8 * (tok & 3));
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      // (with one word, an observed token is always in it)
      const unsigned m = obs ? ((NW == 1 || (tok >> 2) == w) ? one : 0u) : unobs[w];
#pragma unroll
      for (int b = 0; b < 4; ++b) u[4 * w + b] = T(float((m >> (8 * b)) & 0xffu));
    }
  }
};

// Token of node n for this lane's column, as an unsigned byte: a gap token
// -1 reads as 255, which every test below treats as a gap (not observed,
// not ungapped).  Reading bytes unsigned leaves nothing to do between the
// load and the use, so a token fetched a record ahead is not waited for at
// the end of the record that fetched it.
struct TokSrc {
  const SB_GLOBAL uint8_t* aln;     // the tile's first column of the alignment (row stride C)
  unsigned lc;                      // this lane's column in the tile (clamped to the block)
  long long C;
};
__device__ inline int tok_at(const TokSrc& ts, int n) { return int((ts.aln + (long long)n * ts.C)[opq(ts.lc)]); }

// w = M u  (w_a = sum_j M[a][j] u_j)
template <typename T, int AP>
__device__ inline void matvec(const SB_CONST T* M, const T* u, T* w) {
#pragma unroll
  for (int a = 0; a < AP; ++a) {
    T s = M[a * AP] * u[0];
#pragma unroll
    for (int j = 1; j < AP; ++j) s = fma(M[a * AP + j], u[j], s);
    w[a] = s;
  }
}

template <typename T, int AP>
__device__ inline void matvec(const T* M, const T* u, T* w) {
#pragma unroll
  for (int a = 0; a < AP; ++a) {
    T s = M[a * AP] * u[0];
#pragma unroll
    for (int j = 1; j < AP; ++j) s = fma(M[a * AP + j], u[j], s);
    w[a] = s;
  }
}

// x = d^T M  (x_a = sum_i d_i M[i][a])
template <typename T, int AP>
__device__ inline void vecmat(const T* d, const SB_CONST T* M, T* x) {
#pragma unroll
  for (int a = 0; a < AP; ++a) {
    T s = d[0] * M[a];
#pragma unroll
    for (int i = 1; i < AP; ++i) s = fma(d[i], M[i * AP + a], s);
    x[a] = s;
  }
}

// Scale v by 2^-e, e = frexp exponent of max(v); returns e (0 for a zero vector).
template <typename T, int AP>
__device__ inline int rescale(T* v) {
  T m = v[0];
#pragma unroll
  for (int a = 1; a < AP; ++a) m = fmax(m, v[a]);
  const int e = frexp_exp(m);
#pragma unroll
  for (int a = 0; a < AP; ++a) v[a] = ldexp_(v[a], -e);
  return e;
}

template <int CTRL>
__device__ inline double dppd(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
// Sum over the wave, the same bits in every lane, in registers only: a
// butterfly inside each 16-lane row (quad swaps, half-row mirror, row mirror),
// then the row pairs and the wave halves traded by v_permlane16_swap and
// v_permlane32_swap, the two lanes of each trade adding the same two values
// in the same order.  (A shuffle butterfly would be six dependent LDS round
// trips per sum.)
__device__ inline double wave_sum(double v) {
  v += dppd<0xb1>(v);    // quad_perm [1,0,3,2]
  v += dppd<0x4e>(v);    // quad_perm [2,3,0,1]
  v += dppd<0x141>(v);   // row_half_mirror
  v += dppd<0x140>(v);   // row_mirror
  {
    const unsigned lo = __double2loint(v), hi = __double2hiint(v);
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    v = __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
  }
  {
    const unsigned lo = __double2loint(v), hi = __double2hiint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    v = __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
  }
  return v;
}

// Child of an internal node: its inside vector, cumulative exponent and flags.
template <typename T, int AP>
struct Child {
  T u[AP];
  int e;
  bool has;     // subtree contains an ungapped leaf (token 0..A)
  bool branch;  // internal node with an ungapped leaf under both children
  bool leaf;
};

// One tree record as the walk reads it (scalar data).
struct InRec {
  int p, l, r, sl, sr;
};
__device__ inline InRec in_rec(const SB_CONST int* sched, int i) {
  const SB_CONST int* q = sched + i * 8;
  return InRec{q[0], q[1], q[2], q[4], q[5]};
}

// What the walk reads of a record's children and parent, fetched a record
// ahead: both children's staged vectors and words and the tokens of both
// children and the parent.  A fetch whose value will not be used (a leaf, or
// a child forwarded in registers: slot == fwd_slot) reads slot 0 instead,
// which stays in cache, so the instruction stream has no branch and the
// wasted reads cost no memory round trip.
template <typename T, int AP>
struct Fetch {
  T ul[AP], ur[AP];
  int wl, wr;
  // tokens of the left child, right child and parent, each in its own
  // register: combining them at fetch time would wait for the loads there
  int tkl, tkr, tkp;
  __device__ inline int tl() const { return tkl; }
  __device__ inline int tr() const { return tkr; }
  __device__ inline int tp() const { return tkp; }
};

template <typename T, int AP>
__device__ inline void fetch_vec(T* v, int& w, int slot, const SB_GLOBAL char* reg, const SB_GLOBAL int* words,
                                 int lane) {
  Tile<T, AP>::load_vec(reg, slot, lane, v);
  w = *(const SB_GLOBAL int*)((const SB_GLOBAL char*)(words + slot * SB_WAVE) + opq(unsigned(lane) * 4u));
}

template <typename T, int AP, bool LEFT = true>
__device__ inline void fetch_children(Fetch<T, AP>& f, const InRec& q, int fwd_slot, const TokSrc& ts,
                                      const SB_GLOBAL char* Ureg, const SB_GLOBAL int* Uw, int lane) {
  if constexpr (LEFT) fetch_vec<T, AP>(f.ul, f.wl, (q.sl < 0 || q.sl == fwd_slot) ? 0 : q.sl, Ureg, Uw, lane);
  fetch_vec<T, AP>(f.ur, f.wr, (q.sr < 0 || q.sr == fwd_slot) ? 0 : q.sr, Ureg, Uw, lane);
  f.tkl = tok_at(ts, q.l);
  f.tkr = tok_at(ts, q.r);
  f.tkp = tok_at(ts, q.p);
}

// Where a record's child vector comes from.  The kind is uniform per record,
// so the walk runs one specialised copy of the record per combination of
// kinds and neither decodes a token nor selects a vector it does not use.
template <int K>
struct Kind {
  static constexpr int v = K;
};
constexpr int kLeaf = 0;    // decoded from the leaf's token
constexpr int kFwd = 1;     // forwarded in registers from the record computed just before (fu / fw)
constexpr int kStaged = 2;  // staged in the wave's scratch (su / sw)
constexpr int kAny = 3;     // chosen per record by selects (slot < 0: leaf; fwd: forwarded; else staged)
constexpr int kLeafOrStaged = 4;  // leaf or staged, by a uniform branch

template <int K, typename T, int AP>
__device__ inline void take_child(int slot, bool fwd, int tok, const T* su, int sw, const T* fu, int fw,
                                  const TokDec<T, AP>& td, Child<T, AP>& ch) {
  if constexpr (K == kAny) {
    const bool leaf = slot < 0;
    const int A = td.A;
    T tv[AP];
    td.decode(tok, tv);
#pragma unroll
    for (int a = 0; a < AP; ++a) ch.u[a] = leaf ? tv[a] : (fwd ? fu[a] : su[a]);
    const int w = fwd ? fw : sw;
    ch.e = leaf ? 0 : (w >> 2);
    ch.has = leaf ? (unsigned(tok) <= unsigned(A)) : ((w & 1) != 0);
    ch.branch = !leaf && ((w & 2) != 0);
    ch.leaf = leaf;
  } else if constexpr (K == kLeafOrStaged) {
    if (slot < 0) {
      td.decode(tok, ch.u);
      ch.e = 0;
      ch.has = unsigned(tok) <= unsigned(td.A);
      ch.branch = false;
    } else {
#pragma unroll
      for (int a = 0; a < AP; ++a) ch.u[a] = su[a];
      ch.e = sw >> 2;
      ch.has = (sw & 1) != 0;
      ch.branch = (sw & 2) != 0;
    }
    ch.leaf = slot < 0;
  } else if constexpr (K == kLeaf) {
    td.decode(tok, ch.u);
    ch.e = 0;
    ch.has = unsigned(tok) <= unsigned(td.A);
    ch.branch = false;
    ch.leaf = true;
  } else {
    const T* v = K == kFwd ? fu : su;
    const int w = K == kFwd ? fw : sw;
#pragma unroll
    for (int a = 0; a < AP; ++a) ch.u[a] = v[a];
    ch.e = w >> 2;
    ch.has = (w & 1) != 0;
    ch.branch = (w & 2) != 0;
    ch.leaf = false;
  }
}

// v *= the likelihood vector of an internal node's token.  Internal nodes'
// rows are usually all unobserved (token A, or a gap), whose vector is 1 on
// every real state (padded states are zero in every vector anyway), so the
// product is skipped unless a lane of the wave observes the node.  Token
// vectors hold 0 or 1, so the product is exact and the result does not
// depend on the skip.
template <typename T, int AP>
__device__ inline void times_token(int tok, const TokDec<T, AP>& td, T* v) {
  if (__any(unsigned(tok) < unsigned(td.A))) {
    T tv[AP];
    td.decode(tok, tv);
#pragma unroll
    for (int a = 0; a < AP; ++a) v[a] = tv[a] * v[a];
  }
}

template <typename T, int AP, bool F81, bool IRR>
struct Acc {
  // eigen: c[k][l];  irreversible: c = Re C[k][l], then Im C[k][l];
  // F81: c = term1[i][j], t2a[i], t2b[j], t3
  static constexpr int NC = (IRR ? 2 : 1) * AP * AP;
  T c[NC];
  T t2a[F81 ? AP : 1], t2b[F81 ? AP : 1], t3;
  __device__ inline void zero() {
#pragma unroll
    for (int i = 0; i < NC; ++i) c[i] = T(0);
    if constexpr (F81) {
#pragma unroll
      for (int i = 0; i < AP; ++i) { t2a[i] = T(0); t2b[i] = T(0); }
    }
    t3 = T(0);
  }
  // The accumulators as one list of NV values (c, then t2a, t2b, t3 for F81),
  // kept in LDS one value per lane: v[x * 64 + lane].
  static constexpr int NV = NC + (F81 ? 2 * AP + 1 : 0);
  __device__ inline T& at(int x) {
    if constexpr (F81) {
      if (x >= NC + 2 * AP) return t3;
      if (x >= NC + AP) return t2b[x - NC - AP];
      if (x >= NC) return t2a[x - NC];
    }
    return c[x];
  }
  __device__ inline void add_lds(const T* v, int lane) {
#pragma unroll
    for (int x = 0; x < NV; ++x) at(x) += v[x * SB_WAVE + lane];
  }
  __device__ inline void store_lds(T* v, int lane) {
#pragma unroll
    for (int x = 0; x < NV; ++x) v[x * SB_WAVE + lane] = at(x);
  }
};

}  // namespace

// Branch parameters in FP64 (subby/jax/diagonalize.py:141-172, eigensub.py:7-50,
// f81_fast.py:56-66), one workgroup of 256 threads per (block, node) record.
template <typename T>
__global__ void __launch_bounds__(256) sb_branch_params_kernel(const SbLaunch* __restrict__ L) {
  // One workgroup per branch job: V^T, mu, E = exp(mu t) and sqrt(pi) in LDS;
  // wave w computes rows i = w, w + 4, ... four at a time, lane j one column,
  // so each coalesced read of V^T[q][j] feeds four FMAs.
  const int AP = L->AP;
  const int njobs = L->njobs;
  SB_GLOBAL T* out = glb<T>(L->bt);
  __shared__ double vt[64 * 64];          // vt[q * AP + j] = V[j][q]
  __shared__ double mu[64], E[64], sq[64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (blockIdx.x == 0 && tid < SB_NCOUNTER) L->counter[tid] = 0;  // tile counters of the launch that follows
  for (int jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
    const SB_CONST SbBranchJob* job = cst<SbBranchJob>(L->jobs) + jb;
    const SB_CONST double* m = cst<double>(job->model64);
    const double t = job->t;
    const SB_CONST double* V = m + sb_m64_V(AP);
    __syncthreads();
    if (tid < AP) {
      mu[tid] = m[sb_m64_mu(AP) + tid];
      E[tid] = exp(mu[tid] * t);
      sq[tid] = sqrt(m[sb_m64_pi(AP) + tid]);
    }
    for (int idx = tid; idx < AP * AP; idx += 256) {
      const int j = idx / AP, q = idx - j * AP;
      vt[q * AP + j] = V[idx];
    }
    __syncthreads();
    SB_GLOBAL T* o = out + (long long)jb * sb_bt_stride(AP);
    const int j = lane;
    if (j < AP) {
      for (int i0 = wv; i0 < AP; i0 += 16) {
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < AP; ++q) {
          const double vjq = vt[q * AP + j];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = i0 + 4 * r;
            if (i < AP) s[r] = fma(vt[q * AP + i] * E[q], vjq, s[r]);
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = i0 + 4 * r;
          if (i >= AP) continue;
          const T mij = T(s[r] * sq[j] / sq[i]);
          o[sb_bt_M(AP) + i * AP + j] = mij;
          if (AP > 8) o[sb_bt_Mt(AP) + j * AP + i] = mij;
          const double d = mu[i] - mu[j];
          const double J = fabs(d) < 1e-8 ? t * E[i] : E[j] * expm1(d * t) / d;
          o[sb_bt_J(AP) + i * AP + j] = T(J);
        }
      }
    }
    if (tid < AP) {
      o[sb_bt_E(AP) + tid] = T(E[tid]);
      o[sb_bt_dE(AP) + tid] = T(mu[tid] * E[tid]);
    }
    if (tid == 0) {
      const double fm = m[sb_m64_misc(AP)];
      const double e = exp(-fm * t);
      const double p = -expm1(-fm * t);
      o[sb_bt_coef(AP) + 0] = T(t * e);
      o[sb_bt_coef(AP) + 1] = T(p / fm - t * e);
      o[sb_bt_coef(AP) + 2] = T(t * (1.0 + e) - 2.0 * p / fm);
      o[sb_bt_coef(AP) + 3] = T(t);
    }
  }
}

// Branch parameters of an irreversible model in FP64 (oracle.py:474-535,
// eigensub.py:234-268): E_q = e^{mu_q t} (complex), M_ij = Re sum_q V_iq E_q
// Vinv_qj, J_kl = E_l expm1((mu_k - mu_l) t) / (mu_k - mu_l) (t E_k when
// |mu_k - mu_l| < 1e-8), dE_k = mu_k E_k.  One 256-thread workgroup per job,
// one thread per matrix entry.
template <typename T>
__global__ void __launch_bounds__(256) sb_branch_params_irrev_kernel(const SbLaunch* __restrict__ L) {
  const int AP = L->AP;
  const int njobs = L->njobs;
  SB_GLOBAL T* out = glb<T>(L->bt);
  __shared__ double er[64], ei[64], mr[64], mi[64];
  const int tid = threadIdx.x;
  if (blockIdx.x == 0 && tid < SB_NCOUNTER) L->counter[tid] = 0;  // tile counters of the launch that follows
  for (int jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
    const SB_CONST SbBranchJob* job = cst<SbBranchJob>(L->jobs) + jb;
    const SB_CONST double* m = cst<double>(job->model64);
    const double t = job->t;
    __syncthreads();
    if (tid < AP) {
      const double a = m[sb_m64i_mu(AP) + tid], b = m[sb_m64i_mu(AP) + AP + tid];
      const double mag = exp(a * t);
      double sn, cs;
      sincos(b * t, &sn, &cs);
      mr[tid] = a;
      mi[tid] = b;
      er[tid] = mag * cs;
      ei[tid] = mag * sn;
    }
    __syncthreads();
    const SB_CONST double* Vr = m + sb_m64i_V(AP);
    const SB_CONST double* Vi = Vr + AP * AP;
    const SB_CONST double* Wr = m + sb_m64i_Vi(AP);
    const SB_CONST double* Wi = Wr + AP * AP;
    SB_GLOBAL T* o = out + (long long)jb * sb_bti_stride(AP);
    for (int idx = tid; idx < AP * AP; idx += 256) {
      const int i = idx / AP, j = idx - i * AP;
      double s = 0.0;
      for (int q = 0; q < AP; ++q) {
        // Re(V_iq E_q Vinv_qj)
        const double xr = Vr[i * AP + q] * er[q] - Vi[i * AP + q] * ei[q];
        const double xi = Vr[i * AP + q] * ei[q] + Vi[i * AP + q] * er[q];
        s += xr * Wr[q * AP + j] - xi * Wi[q * AP + j];
      }
      o[i * AP + j] = T(s);
      if (AP > 8) o[sb_bti_Mt(AP) + j * AP + i] = T(s);   // the large-alphabet kernels' transposed copy
      // J_ij (k = i, l = j)
      const double dr = mr[i] - mr[j], di = mi[i] - mi[j];
      double jr, ji;
      if (sqrt(dr * dr + di * di) < 1e-8) {
        jr = t * er[i];
        ji = t * ei[i];
      } else {
        // expm1(z), z = (dr + i di) t: Re = expm1(x) cos y - 2 sin^2(y/2), Im = e^x sin y
        const double x = dr * t, y = di * t;
        double sy, cy, sh, ch;
        sincos(y, &sy, &cy);
        sincos(0.5 * y, &sh, &ch);
        const double zr = expm1(x) * cy - 2.0 * sh * sh, zi = exp(x) * sy;
        const double den = dr * dr + di * di;            // z / (mu_i - mu_j)
        const double qr = (zr * dr + zi * di) / den, qi = (zi * dr - zr * di) / den;
        jr = er[j] * qr - ei[j] * qi;                    // times E_j
        ji = er[j] * qi + ei[j] * qr;
      }
      o[sb_bti_Jre(AP) + idx] = T(jr);
      o[sb_bti_Jim(AP) + idx] = T(ji);
    }
    if (tid < AP) {
      o[sb_bti_dEre(AP) + tid] = T(mr[tid] * er[tid] - mi[tid] * ei[tid]);
      o[sb_bti_dEim(AP) + tid] = T(mr[tid] * ei[tid] + mi[tid] * er[tid]);
    }
    if (tid < 4) o[sb_bti_coef(AP) + tid] = T(tid == 3 ? t : 0.0);
  }
}

__global__ void sb_reset_kernel(const SbLaunch* __restrict__ L) {
  if (threadIdx.x < SB_NCOUNTER && blockIdx.x == 0) L->counter[threadIdx.x] = 0;
}

// Register budget: the FP32 four-state reversible variants with one model
// keep four waves per SIMD (at most 128 registers), so a full grid of tiles
// is resident at once.
template <typename T, int AP, bool IRR, bool ROWS, bool SPLIT>
struct SmallWavesPerEU {
  // (split walks too: the F81 variant then spills 28 bytes per lane, and at
  // 3,125 tiles runs 0.232 ms against 0.251 ms at three waves per SIMD)
  static constexpr int v = (sizeof(T) == 4 && AP == 4 && !IRR && !ROWS) ? 4 : 1;
};

// SPLIT: a workgroup of two waves walks each tile together.  The planner
// names a subtree of internal records [s0, s1) (records in depth-first
// preorder, so a subtree is contiguous) that the second wave walks while the
// first walks the rest: inside pass, first wave [nint-1 .. s1], then after a
// barrier [s0-1 .. 0]; second wave [s1-1 .. s0]; outside pass, first wave
// [0 .. s0-1], then after a barrier [s1 .. nint-1]; second wave, after it,
// [s0 .. s1-1].  Nothing is forwarded in registers across a range boundary
// (a range's first record reads its children staged, its last record stores
// what it computes), so the waves exchange only through the tile's scratch
// slot (the first wave's) and, for the log-likelihood scale and the
// accumulators, through LDS.  For launches with spare wave slots: it cuts the
// latency of a tile's walk (s0 = s1 = 0: the first wave walks all).  The
// subtree's branch terms are summed apart and added last, so counts may differ
// from the one-wave walk's in the last bits.
template <typename T, int AP, bool F81, bool GRAD, bool IRR, bool ROWS, bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SmallWavesPerEU<T, AP, IRR, ROWS, SPLIT>::v)))
sb_fused_small_kernel(const SbLaunch* __restrict__ Lp) {
  typedef Tile<T, AP> TL;
  static_assert(!SPLIT || (!IRR && !ROWS), "split walks are built for the single-model kinds only");
  const SbLaunch& args = *Lp;
  const int lane = threadIdx.x & (SB_WAVE - 1);
  const int wpb = blockDim.x / SB_WAVE;
  const int wig = __builtin_amdgcn_readfirstlane(threadIdx.x / SB_WAVE);   // wave in the workgroup
  const int role = SPLIT ? wig : 0;            // 0: first wave, 1: second (split walks)
  // (split walks share the first wave's scratch slot)
  const int gw = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (SPLIT ? 0 : wig));
  constexpr int NCC = AP * AP + (F81 ? 2 * AP + 1 : 0);   // accumulator values per lane
  __shared__ T split_acc[SPLIT ? NCC * SB_WAVE : 1];
  __shared__ T split_inv[SPLIT ? SB_WAVE : 1];
  __shared__ int split_e0[SPLIT ? SB_WAVE : 1];
  __shared__ int split_tile;
  SB_GLOBAL char* wscr = glb<char>(args.scratch) + (long long)gw * args.wave_stride;
  const int NI = args.nint_max;
  SB_GLOBAL char* Ureg = wscr;
  SB_GLOBAL char* Dreg = Ureg + TL::vec_bytes(NI);
  SB_GLOBAL int* Uw = (SB_GLOBAL int*)(Dreg + TL::vec_bytes(NI));
  SB_GLOBAL int* Dw = Uw + NI * SB_WAVE;

  for (;;) {
    int t = 0;
    if constexpr (SPLIT) {
      // (every later LDS read of the previous tile came before this tile's barriers)
      if (threadIdx.x == 0) split_tile = atomicAdd(args.counter, 1);
      __syncthreads();
      t = __builtin_amdgcn_readfirstlane(split_tile);
    } else {
      if (lane == 0) t = atomicAdd(args.counter, 1);
      t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, SB_WAVE));
    }
    if (t >= args.ntiles) break;
    const SB_CONST int* tbp = cst<int>(args.tiles) + 2 * t;
    const int2 tb = make_int2(tbp[0], tbp[1]);
    const SB_CONST SbBlockDesc* B = cst<SbBlockDesc>(args.blocks) + tb.x;
    const long long C = B->C;
    const int A = B->A;
    const int flags = B->flags;
    const int nint = B->nint;
    const int c = tb.y * SB_WAVE + lane;
    const bool valid = c < C;
    const long long cc = valid ? c : C - 1;
    const SB_GLOBAL int8_t* aln = glb<int8_t>(B->aln) + cc;
    const long long tcol0 = (long long)tb.y * SB_WAVE;   // the tile's first column
    const SB_CONST T* bt = cst<T>(B->bt);
    const SB_CONST T* md = cst<T>(B->model);
    const SB_CONST int* sched = cst<int>(B->sched);
    T pi[AP];
#pragma unroll
    for (int a = 0; a < AP; ++a) pi[a] = md[sb_mr_pi_k(AP, IRR) + a];
    const int BTS = sb_bt_stride_k(AP, IRR);
    TokSrc ts;
    ts.aln = glb<uint8_t>(B->aln) + tcol0;
    ts.lc = unsigned(cc - tcol0);
    ts.C = C;

    const bool outside = (flags & (SB_F_COUNTS | SB_F_GRAD)) != 0;

    // ---------------- inside pass ----------------
    // Internal nodes in postorder (records walked backwards), software-
    // pipelined: while record i is computed, everything record i - 1 reads
    // (its right child's staged U and word, both children's tokens and the
    // parent's token) is already in flight.  The fetches are unconditional (a leaf's slot is
    // clamped to 0, a forwarded child's slot is read but not used), so the
    // wave waits only on fetches issued a whole record earlier.  A child whose
    // slot is i + 1 (an internal left child, or an internal right child next
    // to a leaf) is the record computed just before: forwarded in registers.
    // Where each child's vector comes from is uniform per record: the
    // single-model reversible and F81 variants (SPEC) branch to a copy of the
    // record for its kinds; the others choose by selects.
    constexpr bool SPEC = !IRR && !ROWS;
    T u0[AP];
    int e0 = 0;
    if constexpr (SPLIT) {
#pragma unroll
      for (int a = 0; a < AP; ++a) u0[a] = T(1);   // (the second wave never reaches the root)
    }
    bool root_branch = false;
    TokDec<T, AP> td;
    td.init(A);
    if (nint == 0) td.decode(tok_at(ts, 0), u0);  // single-node tree
    // (a tree of one node has no records: nothing is fetched)
    // split walks: the second wave's subtree of records [s0, s1) (from the
    // root record; s0 = s1 = 0: none, the first wave walks the whole tree)
    const int s0 = SPLIT && nint > 0 ? sched[6] : 0, s1 = SPLIT && nint > 0 ? sched[7] : 0;
    InRec q{0, 0, 0, -1, -1};
    Fetch<T, AP> fa, fb;
    if (!SPLIT && nint > 0) {
      q = in_rec(sched, nint - 1);
      fetch_children<T, AP, false>(fa, q, nint, ts, Ureg, Uw, lane);
    }
    T fu[AP];                      // U and word of the record computed last
    int fw = 0;
    int lo_in = 0;                 // the last record of the range walked
#pragma unroll
    for (int a = 0; a < AP; ++a) fu[a] = T(0);
    // (an internal left child is the next node in preorder: slot i + 1, forwarded)
    auto istep = [&](int i, const Fetch<T, AP>& cur, Fetch<T, AP>& nxt, auto LK, auto RK)
                     __attribute__((always_inline)) {
      // the record's two matrices and the next record are requested together,
      // so one scalar round trip covers them
      T ml[AP * AP], mr[AP * AP];
      {
        const SB_CONST T* Ml = bt + (long long)q.l * BTS + sb_bt_M(AP);
        const SB_CONST T* Mr = bt + (long long)q.r * BTS + sb_bt_M(AP);
#pragma unroll
        for (int k = 0; k < AP * AP; ++k) {
          ml[k] = Ml[k];
          mr[k] = Mr[k];
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      const InRec qn = in_rec(sched, i > 0 ? i - 1 : 0);
      __builtin_amdgcn_sched_barrier(0);
      fetch_children<T, AP, false>(nxt, qn, i, ts, Ureg, Uw, lane);
      Child<T, AP> cl, cr;
      take_child<decltype(LK)::v, T, AP>(q.sl, true, cur.tl(), cur.ul, cur.wl, fu, fw, td, cl);
      take_child<decltype(RK)::v, T, AP>(q.sr, q.sr == i + 1, cur.tr(), cur.ur, cur.wr, fu, fw, td, cr);
      T wl[AP], wr[AP], up[AP];
      matvec<T, AP>(ml, cl.u, wl);
      matvec<T, AP>(mr, cr.u, wr);
      if constexpr (SPEC) {
#pragma unroll
        for (int a = 0; a < AP; ++a) up[a] = wl[a] * wr[a];
        times_token<T, AP>(cur.tp(), td, up);
      } else {
        td.decode(cur.tp(), up);
#pragma unroll
        for (int a = 0; a < AP; ++a) up[a] = up[a] * wl[a] * wr[a];
      }
      const int e = rescale<T, AP>(up);
      const int ep = cl.e + cr.e + e;
      const bool has = cl.has || cr.has;
      const bool br = cl.has && cr.has;
      const int word = (ep << 2) | (br ? 2 : 0) | (has ? 1 : 0);
      if (i > 0) {
        // the outside pass reads every internal U; without it, a U forwarded
        // to the next record is never read back (a range's last record has no
        // next record in this wave)
        if (outside || i == lo_in || !(qn.sl == i || qn.sr == i)) {
          TL::store_vec(Ureg, i, lane, up);
          *(SB_GLOBAL int*)((SB_GLOBAL char*)(Uw + i * SB_WAVE) + opq(unsigned(lane) * 4u)) = word;
        }
      } else {
#pragma unroll
        for (int a = 0; a < AP; ++a) u0[a] = up[a];
        e0 = ep;
        root_branch = br;
      }
#pragma unroll
      for (int a = 0; a < AP; ++a) fu[a] = up[a];
      fw = word;
      q = qn;
    };
    auto irec = [&](int i, const Fetch<T, AP>& cur, Fetch<T, AP>& nxt) __attribute__((always_inline)) {
      if constexpr (SPEC) {
        const int rk = q.sr < 0 ? kLeaf : (q.sr == i + 1 ? kFwd : kStaged);
        if (q.sl < 0) {
          if (rk == kLeaf) istep(i, cur, nxt, Kind<kLeaf>{}, Kind<kLeaf>{});
          else if (rk == kFwd) istep(i, cur, nxt, Kind<kLeaf>{}, Kind<kFwd>{});
          else istep(i, cur, nxt, Kind<kLeaf>{}, Kind<kStaged>{});
        } else {
          if (rk == kLeaf) istep(i, cur, nxt, Kind<kFwd>{}, Kind<kLeaf>{});
          else if (rk == kFwd) istep(i, cur, nxt, Kind<kFwd>{}, Kind<kFwd>{});
          else istep(i, cur, nxt, Kind<kFwd>{}, Kind<kStaged>{});
        }
      } else {
        istep(i, cur, nxt, Kind<kAny>{}, Kind<kAny>{});
      }
    };
    // unrolled by two over ping-pong fetch buffers, so no loaded value is copied
    if constexpr (!SPLIT) {
      for (int i = nint - 1; i >= 0; i -= 2) {
        irec(i, fa, fb);
        if (i == 0) break;
        irec(i - 1, fb, fa);
      }
    } else {
      // records hi down to lo; the first one takes both children from their
      // tokens or staged (nothing computed in this range before it)
      auto walk_in = [&](int hi, int lo) __attribute__((always_inline)) {
        if (hi < lo) return;
        lo_in = lo;
        q = in_rec(sched, hi);
        fetch_children<T, AP, true>(fa, q, -3, ts, Ureg, Uw, lane);
        istep(hi, fa, fb, Kind<kLeafOrStaged>{}, Kind<kLeafOrStaged>{});
        for (int i = hi - 1; i >= lo; i -= 2) {
          irec(i, fb, fa);
          if (i == lo) break;
          irec(i - 1, fa, fb);
        }
      };
      if (role == 0) walk_in(nint - 1, s1);
      else walk_in(s1 - 1, s0);
      __syncthreads();               // the second wave's inside vectors are stored
      if (role == 0) walk_in(s0 - 1, 0);
    }
    // log-likelihood and root posterior (pruning.py:84-87; __init__.py:642-650)
    const SB_CONST double* pi64 = cst<double>(B->model64) + sb_m64_pi_k(AP, IRR);
    double s0d = 0.0;
#pragma unroll
    for (int a = 0; a < AP; ++a)
      if (a < A) s0d = fma(pi64[a], double(u0[a]), s0d);
    const double ll = double(e0) * 0.69314718055994530942 + log(s0d);
    T inv_s0 = T(1.0 / s0d);
    if (valid && role == 0 && (flags & SB_F_LL) && B->loglike) glb<double>(B->loglike)[c] = ll;
    if (valid && role == 0 && (flags & SB_F_ROOT) && B->root) {
      SB_GLOBAL T* root = glb<T>(B->root);
#pragma unroll
      for (int a = 0; a < AP; ++a)
        if (a < A) root[(long long)a * C + c] = pi[a] * u0[a] * inv_s0;
    }
    if (!outside) continue;
    if constexpr (SPLIT) {
      // the second wave scales its branches by the root's likelihood too
      if (role == 0) {
        split_e0[lane] = e0;
        split_inv[lane] = inv_s0;
      }
    }

    // ---------------- outside pass + accumulation ----------------
    // Reversible models with one model per block carry each outside vector
    // divided by pi (HAT): detailed balance, pi_i M_ia = pi_a M_ai, gives
    // (d^T M)_a / pi_a = (M (d / pi))_a, so the parent's outside product is a
    // matvec, the root's outside vector is its token alone, and the outside
    // projection Vi^T d (Vi = V / sqrt(pi)) equals Vs^T (d / pi): one
    // projection matrix instead of two.  The stored and forwarded outside
    // vectors of the walk are these; nothing else reads them.
    constexpr bool HAT = !F81 && !IRR && !ROWS;
    const int mask_mode = (flags >> SB_F_MASK_SHIFT) & 3;
    const SB_GLOBAL uint8_t* emask = B->mask ? glb<uint8_t>(B->mask) + cc : nullptr;
    const SB_CONST T* Vs = md + sb_mr_Vs(AP);
    const double gc = GRAD ? ((B->g && valid) ? glb<double>(B->g)[cc] : (valid ? 1.0 : 0.0)) : 0.0;
    SB_GLOBAL double* gpart =
        GRAD ? glb<double>(args.grad_part) + (long long)(B->tile_base + tb.y) * args.grad_stride : nullptr;

    Acc<T, AP, F81, IRR && !ROWS> acc;   // IRR + ROWS: natural-basis real counts
    acc.zero();

    // Internal nodes in preorder, pipelined as the inside pass: record i + 1's
    // own staged outside vector (slot i + 1), its children's staged U and words
    // and its tokens are fetched while record i is computed.  An outside vector
    // computed for a child of record i that is record i + 1's node is
    // forwarded in registers and never stored (an internal left child always,
    // a right child when the left one is a leaf).
    InRec oq{0, 0, 0, -1, -1};
    Fetch<T, AP> ocur, onxt;
    T cd[AP], nd[AP];              // staged outside vector of the record's node and its word
    int cdw = 0, ndw = 0;
    int hi_out = nint - 1;         // the last record of the range walked
    if (!SPLIT && nint > 0) {
      oq = in_rec(sched, 0);
      fetch_children<T, AP>(ocur, oq, -2, ts, Ureg, Uw, lane);
      fetch_vec<T, AP>(cd, cdw, 0, Dreg, Dw, lane);
    }
    T fd[AP];                      // the outside vector (and word) forwarded to this record
    int fdw = 0;
    bool pf = false;               // this record's node is a child of the last record
#pragma unroll
    for (int a = 0; a < AP; ++a) fd[a] = T(0);
    // The loop is unrolled by two over ping-pong buffers (ocur / onxt, cd / nd): the record
    // computed reads one buffer while the next record's fetches land in the
    // other, so no loaded value is copied (a copy would wait for the load).
    auto ostep = [&](int i, const Fetch<T, AP>& cur, const T* cdv, int cdwv, Fetch<T, AP>& nxt, T* ndv,
                     int& ndwv) __attribute__((always_inline)) {
      // HAT: the record's three transition matrices and the next record are
      // requested together, so one scalar round trip covers them
      T mp[HAT ? AP * AP : 1], mlr[HAT ? 2 : 1][HAT ? AP * AP : 1];
      if constexpr (HAT) {
        const SB_CONST T* Mp = bt + (long long)oq.p * BTS + sb_bt_M(AP);
        const SB_CONST T* Ml = bt + (long long)oq.l * BTS + sb_bt_M(AP);
        const SB_CONST T* Mr = bt + (long long)oq.r * BTS + sb_bt_M(AP);
#pragma unroll
        for (int k = 0; k < AP * AP; ++k) {
          mp[k] = Mp[k];
          mlr[0][k] = Ml[k];
          mlr[1][k] = Mr[k];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      const int inext = i + 1 <= hi_out ? i + 1 : i;
      const InRec oqn = in_rec(sched, inext);
      if constexpr (HAT) __builtin_amdgcn_sched_barrier(0);
      const bool last = i >= hi_out;   // (no next record in this wave: store what is computed)
      const bool fwd0 = !last && oqn.p == oq.l;
      const bool fwd1 = !last && oqn.p == oq.r;
      fetch_children<T, AP>(nxt, oqn, -2, ts, Ureg, Uw, lane);
      fetch_vec<T, AP>(ndv, ndwv, (fwd0 || fwd1) ? 0 : inext, Dreg, Dw, lane);
      const int p = oq.p, l = oq.l, r = oq.r, sl = oq.sl, sr = oq.sr;
      // the node's own outside vector: staged, or forwarded by its parent's record
      T rd[AP];
      int rdw;
      {
#pragma unroll
        for (int a = 0; a < AP; ++a) rd[a] = pf ? fd[a] : cdv[a];
        rdw = pf ? fdw : cdwv;
      }
      T x[AP];
      int ex;
      bool stp;
      if (i == 0) {
        td.decode(cur.tp(), x);
        if constexpr (!HAT) {
#pragma unroll
          for (int a = 0; a < AP; ++a) x[a] *= pi[a];
        }
        ex = 0;
        stp = root_branch;
      } else {
        T xm[AP];
        ex = rdw >> 1;
        stp = (rdw & 1) != 0;
        if constexpr (HAT) matvec<T, AP>(mp, rd, xm);
        else vecmat<T, AP>(rd, bt + (long long)p * BTS + sb_bt_M(AP), xm);
        if constexpr (SPEC) {
#pragma unroll
          for (int a = 0; a < AP; ++a) x[a] = xm[a];
          times_token<T, AP>(cur.tp(), td, x);
        } else {
          td.decode(cur.tp(), x);
#pragma unroll
          for (int a = 0; a < AP; ++a) x[a] *= xm[a];
        }
      }
      Child<T, AP> ch[2];
      take_child<SPEC ? kLeafOrStaged : kAny, T, AP>(sl, false, cur.tl(), cur.ul, cur.wl, cur.ul, 0, td, ch[0]);
      take_child<SPEC ? kLeafOrStaged : kAny, T, AP>(sr, false, cur.tr(), cur.ur, cur.wr, cur.ur, 0, td, ch[1]);
      T w[2][AP];
      if constexpr (HAT) {
        matvec<T, AP>(mlr[0], ch[0].u, w[0]);
        matvec<T, AP>(mlr[1], ch[1].u, w[1]);
      } else {
        matvec<T, AP>(bt + (long long)l * BTS + sb_bt_M(AP), ch[0].u, w[0]);
        matvec<T, AP>(bt + (long long)r * BTS + sb_bt_M(AP), ch[1].u, w[1]);
      }
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int n = s == 0 ? l : r;
        const int sn = s == 0 ? sl : sr;
        T d[AP];
#pragma unroll
        for (int a = 0; a < AP; ++a) d[a] = w[1 - s][a] * x[a];
        const int ed = ch[1 - s].e + ex + rescale<T, AP>(d);
        // Steiner membership (oracle.py:1398-1431)
        const bool st = (ch[s].leaf ? ch[s].has : ch[s].branch) || (stp && ch[s].has);
        const int dword = (ed << 1) | (st ? 1 : 0);
        // an internal child's outside vector is read once, by its own record:
        // when that record is the next one it is forwarded in registers and
        // never stored
        const bool fwc = (s == 0) ? fwd0 : fwd1;
        if (!ch[s].leaf && !fwc) {
          TL::store_vec(Dreg, sn, lane, d);
          *(SB_GLOBAL int*)((SB_GLOBAL char*)(Dw + sn * SB_WAVE) + opq(unsigned(lane) * 4u)) = dword;
        }
        if (fwc) {   // uniform
#pragma unroll
          for (int a = 0; a < AP; ++a) fd[a] = d[a];
          fdw = dword;
        }
        bool active;
        if (GRAD || mask_mode == SB_MASK_NONE) active = true;
        else if (mask_mode == SB_MASK_AUTO) active = st && stp;
        else active = emask[(long long)n * C] != 0;
        const T sc_raw = ldexp_(inv_s0, ed + ch[s].e - e0);
        const T sc = active ? sc_raw : T(0);
        const SB_CONST T* bn = bt + (long long)n * BTS;
        const T* u = ch[s].u;
        if constexpr (F81) {
          const T al = bn[sb_bt_coef(AP) + 0], be = bn[sb_bt_coef(AP) + 1], ga = bn[sb_bt_coef(AP) + 2];
          T Ds[AP];
          T piU = T(0), Dsum = T(0);
#pragma unroll
          for (int a = 0; a < AP; ++a) {
            Ds[a] = d[a] * sc;
            piU = fma(pi[a], u[a], piU);
            Dsum += Ds[a];
          }
#pragma unroll
          for (int a = 0; a < AP; ++a) {
            const T ad = al * Ds[a];
#pragma unroll
            for (int b = 0; b < AP; ++b) acc.c[a * AP + b] = fma(ad, u[b], acc.c[a * AP + b]);
            acc.t2a[a] = fma(be * Ds[a], piU, acc.t2a[a]);
            acc.t2b[a] = fma(be * Dsum, u[a], acc.t2b[a]);
          }
          acc.t3 = fma(ga * Dsum, piU, acc.t3);
        } else if constexpr (IRR) {
          // Complex eigenbasis (eigensub.py:271-385): U~ = u Pu, D~ = d Pd,
          // C[k][q] += D~_k sc J_kq U~_q.  Gradients use the exact set, whose
          // projections are those of vjp.py:80-88.  Per-branch models (ROWS):
          // branch n in its own model's eigenbasis, back-transformed at once.
          const SB_CONST T* mI = (ROWS && B->row_models)
              ? cst<T>((const void*)(cst<unsigned long long>(B->row_models)[n])) : md;
          const SB_CONST T* P = mI + sb_mri_set(AP, GRAD ? 1 : 0);
          const SB_CONST T* Pu = P + sb_mri_Pu(AP);
          const SB_CONST T* Pd = P + sb_mri_Pd(AP);
          T Ur[AP], Ui[AP], Dr[AP], Di[AP];
#pragma unroll
          for (int k = 0; k < AP; ++k) {
            T ur = T(0), ui = T(0), dr = T(0), di = T(0);
#pragma unroll
            for (int b = 0; b < AP; ++b) {
              ur = fma(u[b], Pu[b * AP + k], ur);
              ui = fma(u[b], Pu[AP * AP + b * AP + k], ui);
              dr = fma(d[b], Pd[b * AP + k], dr);
              di = fma(d[b], Pd[AP * AP + b * AP + k], di);
            }
            Ur[k] = ur;
            Ui[k] = ui;
            Dr[k] = dr;
            Di[k] = di;
          }
          const SB_CONST T* Jr = bn + sb_bti_Jre(AP);
          const SB_CONST T* Ji = bn + sb_bti_Jim(AP);
          if constexpr (ROWS) {
            // counts_ij += Re(W_ij (L C Rt)_ij) with C_kq = D~_k sc J_kq U~_q:
            // row i of L C is t_q = U~_q sum_k (L_ik D~_k sc) J_kq
            const SB_CONST T* Lr = P + sb_mri_L(AP);
            const SB_CONST T* Li = Lr + AP * AP;
            const SB_CONST T* Rr = P + sb_mri_Rt(AP);
            const SB_CONST T* Ri = Rr + AP * AP;
            const SB_CONST T* Wr = mI + sb_mri_W(AP);
            const SB_CONST T* Wi = Wr + AP * AP;
#pragma unroll
            for (int i = 0; i < AP; ++i) {
              T yr[AP], yi[AP];
#pragma unroll
              for (int k = 0; k < AP; ++k) {
                const T lr = Lr[i * AP + k], li = Li[i * AP + k], dr = Dr[k] * sc, di = Di[k] * sc;
                yr[k] = fma(lr, dr, -li * di);
                yi[k] = fma(lr, di, li * dr);
              }
              T tr[AP], ti[AP];
#pragma unroll
              for (int q = 0; q < AP; ++q) {
                T sr = T(0), si = T(0);
#pragma unroll
                for (int k = 0; k < AP; ++k) {
                  const T jr = Jr[k * AP + q], ji = Ji[k * AP + q];
                  sr = fma(yr[k], jr, fma(-yi[k], ji, sr));
                  si = fma(yr[k], ji, fma(yi[k], jr, si));
                }
                tr[q] = fma(sr, Ur[q], -si * Ui[q]);
                ti[q] = fma(sr, Ui[q], si * Ur[q]);
              }
#pragma unroll
              for (int j = 0; j < AP; ++j) {
                T vr = T(0), vi = T(0);
#pragma unroll
                for (int q = 0; q < AP; ++q) {
                  const T rr = Rr[q * AP + j], ri = Ri[q * AP + j];
                  vr = fma(tr[q], rr, fma(-ti[q], ri, vr));
                  vi = fma(tr[q], ri, fma(ti[q], rr, vi));
                }
                acc.c[i * AP + j] += fma(Wr[i * AP + j], vr, -Wi[i * AP + j] * vi);
              }
            }
          } else {
#pragma unroll
          for (int k = 0; k < AP; ++k) {
            const T dr = Dr[k] * sc, di = Di[k] * sc;
#pragma unroll
            for (int q = 0; q < AP; ++q) {
              const T jr = Jr[k * AP + q], ji = Ji[k * AP + q];
              const T xr = fma(dr, jr, -di * ji), xi = fma(dr, ji, di * jr);
              T& cr = acc.c[k * AP + q];
              T& ci = acc.c[AP * AP + k * AP + q];
              cr = fma(xr, Ur[q], fma(-xi, Ui[q], cr));
              ci = fma(xr, Ui[q], fma(xi, Ur[q], ci));
            }
          }
          }
          if constexpr (GRAD) {
            // d logL / d t_n = scale * Re sum_k D~_k mu_k e^{mu_k t} U~_k  (vjp.py:92-125)
            const SB_CONST T* Er = bn + sb_bti_dEre(AP);
            const SB_CONST T* Ei = bn + sb_bti_dEim(AP);
            T s = T(0);
#pragma unroll
            for (int k = 0; k < AP; ++k) {
              const T yr = fma(Dr[k], Er[k], -Di[k] * Ei[k]), yi = fma(Dr[k], Ei[k], Di[k] * Er[k]);
              s = fma(yr, Ur[k], fma(-yi, Ui[k], s));
            }
            const double v = wave_sum(gc * double(s) * double(sc_raw));
            if (lane == 0) gpart[n] = v;
          }
        } else if constexpr (ROWS) {
          // Per-branch models (_branch_counts_per_row, subby/jax/__init__.py:102-144):
          // branch n is projected, accumulated and back-transformed in the
          // eigenbasis of its own model, straight into natural-basis counts.
          const SB_CONST T* mn = B->row_models
              ? cst<T>((const void*)(cst<unsigned long long>(B->row_models)[n])) : md;
          const SB_CONST T* Vsn = mn + sb_mr_Vs(AP);
          const SB_CONST T* Vin = mn + sb_mr_Vi(AP);
          const SB_CONST T* Vn = mn + sb_mr_V(AP);
          const SB_CONST T* Sn = mn + sb_mr_S(AP);
          T Ut[AP], Dt[AP];
#pragma unroll
          for (int k = 0; k < AP; ++k) {
            T su = T(0), sd = T(0);
#pragma unroll
            for (int b = 0; b < AP; ++b) {
              su = fma(u[b], Vsn[b * AP + k], su);
              sd = fma(d[b], Vin[b * AP + k], sd);
            }
            Ut[k] = su;
            Dt[k] = sd * sc;
          }
          const SB_CONST T* J = bn + sb_bt_J(AP);
          T Ce[AP * AP];
#pragma unroll
          for (int k = 0; k < AP; ++k)
#pragma unroll
            for (int q = 0; q < AP; ++q) Ce[k * AP + q] = Dt[k] * J[k * AP + q] * Ut[q];
#pragma unroll
          for (int i = 0; i < AP; ++i) {
            T tmp[AP];
#pragma unroll
            for (int q = 0; q < AP; ++q) {
              T v = T(0);
#pragma unroll
              for (int k = 0; k < AP; ++k) v = fma(Vn[i * AP + k], Ce[k * AP + q], v);
              tmp[q] = v;
            }
#pragma unroll
            for (int j = 0; j < AP; ++j) {
              T v = T(0);
#pragma unroll
              for (int q = 0; q < AP; ++q) v = fma(tmp[q], Vn[j * AP + q], v);
              acc.c[i * AP + j] = (i == j) ? acc.c[i * AP + j] + v : fma(Sn[i * AP + j], v, acc.c[i * AP + j]);
            }
          }
        } else {
          T Ut[AP], Dt[AP];
#pragma unroll
          for (int k = 0; k < AP; ++k) {
            T su = T(0), sd = T(0);
#pragma unroll
            for (int b = 0; b < AP; ++b) {
              // (d here is D / pi, so the outside projection uses Vs too)
              su = fma(u[b], Vs[b * AP + k], su);
              sd = fma(d[b], Vs[b * AP + k], sd);
            }
            Ut[k] = su;
            Dt[k] = sd;
          }
          const SB_CONST T* J = bn + sb_bt_J(AP);
#pragma unroll
          for (int k = 0; k < AP; ++k) {
            const T dk = Dt[k] * sc;
#pragma unroll
            for (int q = 0; q < AP; ++q) acc.c[k * AP + q] = fma(dk * J[k * AP + q], Ut[q], acc.c[k * AP + q]);
          }
          if constexpr (GRAD) {
            // d logL / d t_n = scale * sum_k D~_k mu_k e^{mu_k t} U~_k  (vjp.py:92-117)
            const SB_CONST T* dE = bn + sb_bt_dE(AP);
            T s = T(0);
#pragma unroll
            for (int k = 0; k < AP; ++k) s = fma(Dt[k] * dE[k], Ut[k], s);
            const double v = wave_sum(gc * double(s) * double(sc_raw));
            if (lane == 0) gpart[n] = v;
          }
        }
      }
      pf = fwd0 || fwd1;
      oq = oqn;
    };
    if constexpr (!SPLIT) {
      for (int i = 0; i < nint; i += 2) {
        ostep(i, ocur, cd, cdw, onxt, nd, ndw);
        if (i + 1 >= nint) break;
        ostep(i + 1, onxt, nd, ndw, ocur, cd, cdw);
      }
    } else {
      // records lo up to hi; the first one reads its own outside vector staged
      auto walk_out = [&](int lo, int hi) __attribute__((always_inline)) {
        if (hi < lo) return;
        hi_out = hi;
        oq = in_rec(sched, lo);
        fetch_children<T, AP>(ocur, oq, -2, ts, Ureg, Uw, lane);
        fetch_vec<T, AP>(cd, cdw, lo, Dreg, Dw, lane);
        pf = false;
        for (int i = lo; i <= hi; i += 2) {
          ostep(i, ocur, cd, cdw, onxt, nd, ndw);
          if (i + 1 > hi) break;
          ostep(i + 1, onxt, nd, ndw, ocur, cd, cdw);
        }
      };
      if (role == 0) walk_out(0, s0 - 1);
      __syncthreads();               // the subtree's outside vector and the root's scale are stored
      if (role == 1) {
        e0 = split_e0[lane];
        inv_s0 = split_inv[lane];
        walk_out(s0, s1 - 1);
      } else {
        walk_out(s1, nint - 1);
      }
      // the second wave's accumulators join the first wave's
      if (role == 1) acc.store_lds(split_acc, lane);
      __syncthreads();
      if (role == 1) continue;
      if (s1 > s0) acc.add_lds(split_acc, lane);
    }

    // ---------------- epilogue: natural-basis counts ----------------
    T cnt[AP * AP];
    if constexpr (ROWS) {
#pragma unroll
      for (int i = 0; i < AP * AP; ++i) cnt[i] = acc.c[i];
    } else if constexpr (F81) {
      const T fmu = md[sb_mr_misc(AP)];
#pragma unroll
      for (int i = 0; i < AP; ++i)
#pragma unroll
        for (int j = 0; j < AP; ++j) {
          const T I = acc.c[i * AP + j] + acc.t2a[i] + pi[i] * (acc.t2b[j] + acc.t3);
          cnt[i * AP + j] = (i == j) ? I : fmu * pi[j] * I;
        }
    } else if constexpr (IRR) {
      // VCV = L C Rt (complex), one row of L C at a time; counts_ii = Re VCV_ii,
      // counts_ij = Re(Q_ij VCV_ij)  (eigensub.py:388-434)
      const SB_CONST T* P = md + sb_mri_set(AP, GRAD ? 1 : 0);
      const SB_CONST T* Lr = P + sb_mri_L(AP);
      const SB_CONST T* Li = Lr + AP * AP;
      const SB_CONST T* Rr = P + sb_mri_Rt(AP);
      const SB_CONST T* Ri = Rr + AP * AP;
      const SB_CONST T* Wr = md + sb_mri_W(AP);
      const SB_CONST T* Wi = Wr + AP * AP;
#pragma unroll
      for (int i = 0; i < AP; ++i) {
        T tr[AP], ti[AP];
#pragma unroll
        for (int q = 0; q < AP; ++q) {
          T sr = T(0), si = T(0);
#pragma unroll
          for (int k = 0; k < AP; ++k) {
            const T lr = Lr[i * AP + k], li = Li[i * AP + k];
            const T cr = acc.c[k * AP + q], ci = acc.c[AP * AP + k * AP + q];
            sr = fma(lr, cr, fma(-li, ci, sr));
            si = fma(lr, ci, fma(li, cr, si));
          }
          tr[q] = sr;
          ti[q] = si;
        }
#pragma unroll
        for (int j = 0; j < AP; ++j) {
          T vr = T(0), vi = T(0);
#pragma unroll
          for (int q = 0; q < AP; ++q) {
            const T rr = Rr[q * AP + j], ri = Ri[q * AP + j];
            vr = fma(tr[q], rr, fma(-ti[q], ri, vr));
            vi = fma(tr[q], ri, fma(ti[q], rr, vi));
          }
          cnt[i * AP + j] = (i == j) ? vr : fma(Wr[i * AP + j], vr, -Wi[i * AP + j] * vi);
        }
      }
    } else {
      const SB_CONST T* V = md + sb_mr_V(AP);
      const SB_CONST T* S = md + sb_mr_S(AP);
      T tmp[AP * AP];
#pragma unroll
      for (int i = 0; i < AP; ++i)
#pragma unroll
        for (int q = 0; q < AP; ++q) {
          T s = T(0);
#pragma unroll
          for (int k = 0; k < AP; ++k) s = fma(V[i * AP + k], acc.c[k * AP + q], s);
          tmp[i * AP + q] = s;
        }
#pragma unroll
      for (int i = 0; i < AP; ++i)
#pragma unroll
        for (int j = 0; j < AP; ++j) {
          T s = T(0);
#pragma unroll
          for (int q = 0; q < AP; ++q) s = fma(tmp[i * AP + q], V[j * AP + q], s);
          cnt[i * AP + j] = (i == j) ? s : S[i * AP + j] * s;
        }
    }
    if (valid && (flags & SB_F_COUNTS) && B->counts) {
      SB_GLOBAL T* out = glb<T>(B->counts);
#pragma unroll
      for (int i = 0; i < AP; ++i)
#pragma unroll
        for (int j = 0; j < AP; ++j)
          if (i < A && j < A) out[((long long)i * A + j) * C + c] = cnt[i * AP + j];
    }
    if constexpr (GRAD) {
      // Rate-parameter gradients by the Fisher identity: sum_ij W_ij N_ij.
      SB_GLOBAL double* rpart = glb<double>(args.rate_part) + (long long)(B->tile_base + tb.y) * SB_MAX_RATE;
      for (int q = 0; q < B->n_rate; ++q) {
        const SB_CONST double* W = cst<double>(B->rate_w) + (long long)q * AP * AP;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < AP * AP; ++i) s = fma(W[i], double(cnt[i]), s);
        const double v = wave_sum(gc * s);
        if (lane == 0) rpart[q] = v;
      }
    }
  }
}

// Sum of the per-tile gradient partials.  A workgroup takes 16 nodes (or the
// rate parameters) of one block; 16 threads per node add fixed tile strides
// and a fixed tree adds their sums, so the result does not depend on timing.
__global__ void __launch_bounds__(256) sb_grad_reduce_kernel(const SbLaunch* __restrict__ L) {
  __shared__ double red[256];
  const SB_GLOBAL double* grad_part = glb<double>(L->grad_part);
  const SB_GLOBAL double* rate_part = glb<double>(L->rate_part);
  const long long grad_stride = L->grad_stride;
  const int tid = threadIdx.x, nl = tid & 15, part = tid >> 4;
  long long g = 0;
  for (int b = 0; b < L->nblocks; ++b) {
    const SbBlockDesc* B = L->blocks + b;
    if (!(B->flags & SB_F_GRAD)) continue;
    const int ngrp = (B->R + 15) / 16 + 1;     // node groups, then the rate parameters
    for (int gr = 0; gr < ngrp; ++gr, ++g) {
      if (g % gridDim.x != blockIdx.x) continue;
      const bool rates = gr == ngrp - 1;
      const int n = rates ? nl : gr * 16 + nl;
      const int lim = rates ? B->n_rate : B->R;
      const long long stride = rates ? SB_MAX_RATE : grad_stride;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      if (n < lim && (rates || n > 0)) {
        const SB_GLOBAL double* base = (rates ? rate_part : grad_part) + (long long)B->tile_base * stride + n;
        int t = part;
        for (; t + 48 < B->ntiles; t += 64) {
          s0 += base[(long long)t * stride];
          s1 += base[(long long)(t + 16) * stride];
          s2 += base[(long long)(t + 32) * stride];
          s3 += base[(long long)(t + 48) * stride];
        }
        for (; t < B->ntiles; t += 16) s0 += base[(long long)t * stride];
      }
      red[tid] = (s0 + s1) + (s2 + s3);
      __syncthreads();
      for (int w = 8; w >= 1; w >>= 1) {
        if (part < w) red[tid] += red[tid + 16 * w];
        __syncthreads();
      }
      if (part == 0 && n < lim) {
        if (rates) {
          if (B->rate_grad) B->rate_grad[n] = red[tid];
        } else if (B->dist_grad) {
          B->dist_grad[n] = n > 0 ? red[tid] : 0.0;
        }
      }
      __syncthreads();
    }
  }
}

// Explicit instantiations and host-visible launch helpers.
#define SB_INST(T, AP, F81, GRAD, IRR, ROWS, SPLIT) \
  template __global__ void sb_fused_small_kernel<T, AP, F81, GRAD, IRR, ROWS, SPLIT>(const SbLaunch*);
#define SB_INST_T(T, AP)                              \
  SB_INST(T, AP, false, false, false, false, false)   \
  SB_INST(T, AP, true, false, false, false, false)    \
  SB_INST(T, AP, false, true, false, false, false)    \
  SB_INST(T, AP, false, false, true, false, false)    \
  SB_INST(T, AP, false, true, true, false, false)     \
  SB_INST(T, AP, false, false, false, true, false)    \
  SB_INST(T, AP, false, false, true, true, false)
// split walks: the single-model reversible and F81 kinds at four states
#define SB_INST_SPLIT(T)                              \
  SB_INST(T, 4, false, false, false, false, true)     \
  SB_INST(T, 4, true, false, false, false, true)      \
  SB_INST(T, 4, false, true, false, false, true)
SB_INST_T(float, 4)
SB_INST_T(float, 8)
SB_INST_T(double, 4)
SB_INST_T(double, 8)
SB_INST_SPLIT(float)
SB_INST_SPLIT(double)
template __global__ void sb_branch_params_kernel<float>(const SbLaunch*);
template __global__ void sb_branch_params_kernel<double>(const SbLaunch*);
template __global__ void sb_branch_params_irrev_kernel<float>(const SbLaunch*);
template __global__ void sb_branch_params_irrev_kernel<double>(const SbLaunch*);

static void (*pick_small(int dtype, int AP, bool f81, bool grad, bool irr,
                         bool rows, bool split))(const SbLaunch*) {
  void (*k)(const SbLaunch*) = nullptr;
  if ((irr && f81) || (rows && (f81 || grad)) || (f81 && grad)) return nullptr;
  if (split) {
    // two waves per tile: the single-model reversible and F81 kinds, four states
    if (AP != 4 || irr || rows) return nullptr;
    if (dtype == 0)
      k = grad ? sb_fused_small_kernel<float, 4, false, true, false, false, true>
                : f81 ? sb_fused_small_kernel<float, 4, true, false, false, false, true>
                      : sb_fused_small_kernel<float, 4, false, false, false, false, true>;
    else
      k = grad ? sb_fused_small_kernel<double, 4, false, true, false, false, true>
                : f81 ? sb_fused_small_kernel<double, 4, true, false, false, false, true>
                      : sb_fused_small_kernel<double, 4, false, false, false, false, true>;
    return k;
  }
#define SB_PICK(T, APV)                                                                          \
  if (AP == APV) {                                                                               \
    if (rows) k = irr ? sb_fused_small_kernel<T, APV, false, false, true, true, false>           \
                      : sb_fused_small_kernel<T, APV, false, false, false, true, false>;         \
    else if (irr) k = grad ? sb_fused_small_kernel<T, APV, false, true, true, false, false>      \
                           : sb_fused_small_kernel<T, APV, false, false, true, false, false>;    \
    else if (grad) k = sb_fused_small_kernel<T, APV, false, true, false, false, false>;          \
    else if (f81) k = sb_fused_small_kernel<T, APV, true, false, false, false, false>;           \
    else k = sb_fused_small_kernel<T, APV, false, false, false, false, false>;                   \
  }
  if (dtype == 0) { SB_PICK(float, 4) SB_PICK(float, 8) }
  else { SB_PICK(double, 4) SB_PICK(double, 8) }
#undef SB_PICK
  return k;
}

// Resident workgroups per CU for a block of `threads` threads with `lds`
// bytes of token stage (0 on error).
extern "C" int sb_fused_small_occupancy(int dtype, int AP, bool f81, bool grad, bool irr,
                                        bool rows, bool split, int threads, int lds) {
  void (*k)(const SbLaunch*) = pick_small(dtype, AP, f81, grad, irr, rows, split);
  if (!k) return 0;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, threads, size_t(lds)) != hipSuccess) return 0;
  return n;
}

extern "C" hipError_t sb_launch_fused_small(int dtype, int AP, bool f81, bool grad, bool irr,
                                            bool rows, bool split, dim3 grid, dim3 block, int lds,
                                            const SbLaunch* L, hipStream_t s) {
  void (*k)(const SbLaunch*) = pick_small(dtype, AP, f81, grad, irr, rows, split);
  if (!k) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, grid, block, size_t(lds), s, L);
  return hipGetLastError();
}

extern "C" hipError_t sb_launch_branch_params(int dtype, bool irr, int nwg, const SbLaunch* L,
                                              hipStream_t s) {
  if (irr) {
    if (dtype == 0)
      hipLaunchKernelGGL(sb_branch_params_irrev_kernel<float>, dim3(nwg), dim3(256), 0, s, L);
    else
      hipLaunchKernelGGL(sb_branch_params_irrev_kernel<double>, dim3(nwg), dim3(256), 0, s, L);
    return hipGetLastError();
  }
  if (dtype == 0)
    hipLaunchKernelGGL(sb_branch_params_kernel<float>, dim3(nwg), dim3(256), 0, s, L);
  else
    hipLaunchKernelGGL(sb_branch_params_kernel<double>, dim3(nwg), dim3(256), 0, s, L);
  return hipGetLastError();
}

extern "C" hipError_t sb_launch_reset(const SbLaunch* L, hipStream_t s) {
  hipLaunchKernelGGL(sb_reset_kernel, dim3(1), dim3(64), 0, s, L);
  return hipGetLastError();
}

extern "C" hipError_t sb_launch_grad_reduce(int nwg, const SbLaunch* L, hipStream_t s) {
  hipLaunchKernelGGL(sb_grad_reduce_kernel, dim3(nwg), dim3(256), 0, s, L);
  return hipGetLastError();
}
