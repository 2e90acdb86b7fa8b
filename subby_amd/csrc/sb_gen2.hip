// Two-pass path for general reversible models, padded alphabets 16..64, FP32, gfx950.
//
// The per-column A x A accumulator limits a wave to a few columns, while
// matrix products want many.  So the work is split at the accumulation:
//
// Pass 1 (sb_gwalk_kernel): a workgroup of 4 waves owns a tile of 64 columns
// of one block (16 per wave) and walks the tree: the inside pass
// (pruning.py:94-118), the outside pass (outside.py:94-126) and the eigenbasis
// projections (eigensub.py:53-84).  Every product with an AW x AW matrix --
// M(t) U, M(t)^T D, Vs^T U, Vi^T D -- is a GEMM over the wave's 16 columns on
// v_mfma_f32_16x16x4f32, with the matrix staged once per workgroup in LDS.
// Children that are leaves use a column of the staged matrix instead.
// Rescaling, Steiner flags and scales are those of the fused kernels (exact
// power-of-two rescaling, integer exponents, scale = ldexp(1/s0, eD+eU-e0)).
// Per branch n and column it leaves in the tile's stage area
//     Ut_n = Vs^T U_n   and   Dt_n = (Vi^T D_n) * scale * mask,
// the operands of accumulate_C (eigensub.py:87-135), and writes logLike and
// the root posterior.
//
// Pass 2 (sb_gacc_kernel): lane = eigen-state k, two columns per wave, eight
// waves share J_n(t) staged in LDS:  C[k][q] = sum_n Dt_n[k] J_n[k][q] Ut_n[q],
// then counts = (V C V^T) o S with the diagonal unscaled (eigensub.py:138-172).
//
// Vector layout in pass 1: lane l holds column l % 16 of the wave and, in
// register v, state 16 (v / 4) + 4 (l / 16) + v % 4 -- the output layout of
// v_mfma_f32_16x16x4f32, which also serves as its B operand when the K index
// of the A operand is permuted to match.
#include <hip/hip_runtime.h>

#include "../../include/subby_amd.h"
#include "sb_plan.h"

#define SB_CONST __attribute__((address_space(4)))
#define SB_GLOBAL __attribute__((address_space(1)))

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWC = 16;            // columns per wave, pass 1
constexpr int kWPG = 4;            // waves per workgroup, pass 1
constexpr int kTW = kWC * kWPG;    // columns per tile
constexpr int kTC2 = 2;            // columns per wave, pass 2
constexpr int kWPG2 = 8;           // waves per workgroup, pass 2
constexpr int kQ2 = kTW / (kTC2 * kWPG2);   // pass-2 workgroups per tile

template <typename T>
__device__ inline const SB_CONST T* cst(const void* p) { return (const SB_CONST T*)(p); }
template <typename T>
__device__ inline SB_GLOBAL T* glb(void* p) { return (SB_GLOBAL T*)(p); }
template <typename T>
__device__ inline const SB_GLOBAL T* glb(const void* p) { return (const SB_GLOBAL T*)(p); }

// Workgroup barrier that orders LDS only: loads and stores to global memory
// stay in flight across it (a plain __syncthreads() waits for all of them).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Reductions over the four quarters of a wave (lanes j, j+16, j+32, j+48 hold
// one column): row swap then half swap; every lane gets identical bits.
__device__ inline float qmax(float x) {
  const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  x = fmaxf(__uint_as_float(a[0]), __uint_as_float(a[1]));
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return fmaxf(__uint_as_float(b[0]), __uint_as_float(b[1]));
}
__device__ inline double qsum(double x) {
  unsigned lo = __double2loint(x), hi = __double2hiint(x);
  auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  x = __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
  lo = __double2loint(x);
  hi = __double2hiint(x);
  a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
}

// Stage an AP x AP row-major matrix into LDS as AW x AW with row stride
// AW + 4 and zero pads, transposed when TR.
template <int AW, bool TR>
__device__ inline void stage(float* dst, const SB_GLOBAL float* src, int AP, int tid, int nthr) {
  constexpr int PS = AW + 4;
#pragma unroll 4
  for (int idx = tid; idx < AW * AW; idx += nthr) {
    const int r = idx / AW, c = idx % AW;
    const float v = (r < AP && c < AP) ? src[r * AP + c] : 0.f;
    if (TR) dst[c * PS + r] = v;
    else dst[r * PS + c] = v;
  }
}

// w = P u for the wave's 16 columns (P: AW x AW in LDS, row stride AW + 4).
// Output tile t, K step (t2, r2): the A operand of lane (i, h) is
// P[16t + i][16 t2 + 4h + r2], the B operand its own register 4 t2 + r2.
template <int AW>
__device__ inline void gemm(const float* P, const float* u, float* w, int lane) {
  constexpr int NT = AW / 16, PS = AW + 4;
  const int i = lane & 15, h = lane >> 4;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t2 = 0; t2 < NT; ++t2) {
    f32x4 a[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) a[t] = *(const f32x4*)(P + (16 * t + i) * PS + 16 * t2 + 4 * h);
#pragma unroll
    for (int r2 = 0; r2 < 4; ++r2)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][r2], u[4 * t2 + r2], acc[t], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) w[4 * t + r] = acc[t][r];
}

struct Kid {
  int n, tok, e;
  bool leaf, has, branch;
};

}  // namespace

// ------------------------------------------------------------------ pass 1
// Stage area of a tile (floats unless noted), R = rmax nodes, 64 columns:
//   UU[R][64][AW]  U of internal nodes, then Ut of every branch
//   DD[R][64][AW]  D of internal nodes
//   DT[R][64][AW]  Dt * scale * mask of every branch
//   UW[R][64], DW[R][64]  ints: U word {exp << 2 | both | any}, D word {exp << 1 | steiner}
template <int AW>
__global__ void __launch_bounds__(256) sb_gwalk_kernel(const SbLaunch* __restrict__ Lp, int tile0) {
  constexpr int NV = AW / 4;
  constexpr int PS = AW + 4;
  __shared__ __attribute__((aligned(16))) float sVs[AW * PS];   // Vs^T
  __shared__ __attribute__((aligned(16))) float sVi[AW * PS];   // Vi^T
  __shared__ __attribute__((aligned(16))) float sP[2][AW * PS];
  __shared__ float sv1[AW];                                     // Vs^T (all ones on real states)
  const SbLaunch& L = *Lp;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int jc = lane & 15, h = lane >> 4;
  const int tl = blockIdx.x;
  const SB_CONST int* tbp = cst<int>(L.tiles) + 2 * (tile0 + tl);
  const int bidx = tbp[0], tidx = tbp[1];
  const SB_CONST SbBlockDesc* B = cst<SbBlockDesc>(L.blocks) + bidx;
  const long long C = B->C;
  const int A = B->A;
  const int flags = B->flags;
  const int nint = B->nint;
  const int AP = L.AP, RM = L.rmax;
  const int wcol = wv * kWC + jc;
  const long long col = (long long)tidx * kTW + wcol;
  const bool valid = col < C;
  const long long cc = valid ? col : C - 1;
  SB_GLOBAL char* st = glb<char>(L.stage) + (long long)tl * L.stage_stride;
  const long long vplane = (long long)RM * kTW * AW;
  SB_GLOBAL float* UU = (SB_GLOBAL float*)st;
  SB_GLOBAL float* DD = UU + vplane;
  SB_GLOBAL float* DT = DD + vplane;
  SB_GLOBAL int* UW = (SB_GLOBAL int*)(DT + vplane);
  SB_GLOBAL int* DWd = UW + (long long)RM * kTW;
  const SB_GLOBAL int8_t* aln = glb<int8_t>(B->aln);
  const SB_GLOBAL float* bt = glb<float>(B->bt);
  const SB_GLOBAL float* md = glb<float>(B->model);
  const SB_CONST int* sched = cst<int>(B->sched);
  const int BTS = sb_bt_stride(AP);

  auto state = [&](int v) { return 16 * (v >> 2) + 4 * h + (v & 3); };
  float pis[NV], ones[NV];   // pi and 1 on this lane's real states
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int s = state(v);
    pis[v] = s < A ? md[sb_mr_pi(AP) + s] : 0.f;
    ones[v] = s < A ? 1.f : 0.f;
  }
  stage<AW, true>(sVs, md + sb_mr_Vs(AP), AP, tid, 256);
  stage<AW, true>(sVi, md + sb_mr_Vi(AP), AP, tid, 256);
  lds_barrier();
  if (tid < AW) {
    float s = 0.f;
    for (int a = 0; a < A; ++a) s += sVs[tid * PS + a];
    sv1[tid] = s;
  }
  lds_barrier();
  float vs1r[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) vs1r[v] = sv1[state(v)];

  auto vptr = [&](SB_GLOBAL float* base, int n) { return base + ((long long)n * kTW + wcol) * AW + 4 * h; };
  auto load_vec = [&](SB_GLOBAL float* base, int n, float* u) {
    const SB_GLOBAL float* p = vptr(base, n);
#pragma unroll
    for (int t = 0; t < NV / 4; ++t) {
      const f32x4 x = *(const SB_GLOBAL f32x4*)(p + 16 * t);
#pragma unroll
      for (int r = 0; r < 4; ++r) u[4 * t + r] = x[r];
    }
  };
  auto store_vec = [&](SB_GLOBAL float* base, int n, const float* u) {
    SB_GLOBAL float* p = vptr(base, n);
#pragma unroll
    for (int t = 0; t < NV / 4; ++t)
      *(SB_GLOBAL f32x4*)(p + 16 * t) = f32x4{u[4 * t], u[4 * t + 1], u[4 * t + 2], u[4 * t + 3]};
  };
  auto token = [&](int n) -> int { return aln[(long long)n * C + cc]; };
  auto token_vec = [&](int tok, float* u) {
    const bool obs = tok >= 0 && tok < A;
#pragma unroll
    for (int v = 0; v < NV; ++v) u[v] = obs ? (state(v) == tok ? 1.f : 0.f) : ones[v];
  };
  // w = P e_tok (a column of the staged matrix), or the all-ones image
  // (rows of M(t) sum to one) for an unobserved token
  auto gather = [&](const float* P, int tok, const float* all, float* w) {
    const bool obs = tok >= 0 && tok < A;
#pragma unroll
    for (int v = 0; v < NV; ++v) w[v] = obs ? P[state(v) * PS + tok] : all[v];
  };
  auto rescale = [&](float* v) -> int {
    float m = v[0];
#pragma unroll
    for (int i = 1; i < NV; ++i) m = fmaxf(m, v[i]);
    const int x = __builtin_amdgcn_frexp_expf(qmax(m));
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = __builtin_amdgcn_ldexpf(v[i], -x);
    return x;
  };
  auto kid = [&](int n, int slot) {
    Kid k;
    k.n = n;
    k.leaf = slot < 0;
    if (k.leaf) {
      k.tok = token(n);
      k.e = 0;
      k.has = k.tok >= 0 && k.tok <= A;
      k.branch = false;
    } else {
      const int w = UW[(long long)n * kTW + wcol];
      k.tok = 0;
      k.e = w >> 2;
      k.has = (w & 1) != 0;
      k.branch = (w & 2) != 0;
    }
    return k;
  };
  // u (the child's inside vector) and w = M u, with M staged in P
  auto child_vecs = [&](const Kid& k, const float* P, float* u, float* w) {
    if (k.leaf) {
      token_vec(k.tok, u);
      gather(P, k.tok, ones, w);
    } else {
      load_vec(UU, k.n, u);
      gemm<AW>(P, u, w, lane);
    }
  };

  // ---------------- inside pass (postorder) ----------------
  float u0[NV];
  int e0 = 0;
  bool rbr = false;
  if (nint == 0) token_vec(token(0), u0);
  for (int i = nint - 1; i >= 0; --i) {
    const SB_CONST int* rec = sched + i * 8;
    const int p = rec[0], l = rec[1], r = rec[2], sl = rec[4], sr = rec[5];
    const Kid kl = kid(l, sl), kr = kid(r, sr);
    const int tokp = token(p);
    lds_barrier();   // previous readers of sP are done
    stage<AW, false>(sP[0], bt + (long long)l * BTS + sb_bt_M(AP), AP, tid, 256);
    stage<AW, false>(sP[1], bt + (long long)r * BTS + sb_bt_M(AP), AP, tid, 256);
    lds_barrier();
    float u[NV], wl[NV], wr[NV];
    child_vecs(kl, sP[0], u, wl);
    child_vecs(kr, sP[1], u, wr);
    float up[NV];
    token_vec(tokp, up);
#pragma unroll
    for (int v = 0; v < NV; ++v) up[v] *= wl[v] * wr[v];
    const int e = rescale(up);
    const int ep = kl.e + kr.e + e;
    if (i > 0) {
      const int word = (ep << 2) | ((kl.has && kr.has) ? 2 : 0) | ((kl.has || kr.has) ? 1 : 0);
      store_vec(UU, p, up);
      if (h == 0) UW[(long long)p * kTW + wcol] = word;
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v) u0[v] = up[v];
      e0 = ep;
      rbr = kl.has && kr.has;
    }
  }
  // log-likelihood (FP64 pi . u), root posterior
  const SB_CONST double* pi64 = cst<double>(B->model64) + sb_m64_pi(AP);
  double s0 = 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int s = state(v);
    if (s < A) s0 = fma(pi64[s], double(u0[v]), s0);
  }
  s0 = qsum(s0);
  const float inv_s0 = float(1.0 / s0);
  if (valid && h == 0 && (flags & SB_F_LL) && B->loglike)
    glb<double>(B->loglike)[cc] = double(e0) * 0.69314718055994530942 + log(s0);
  if (valid && (flags & SB_F_ROOT) && B->root) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int s = state(v);
      if (s < A) glb<float>(B->root)[(long long)s * C + cc] = pis[v] * u0[v] * inv_s0;
    }
  }
  if (!(flags & SB_F_COUNTS)) return;

  // ---------------- outside pass + projections (preorder) ----------------
  const int mask_mode = (flags >> SB_F_MASK_SHIFT) & 3;
  const SB_GLOBAL uint8_t* emask = B->mask ? glb<uint8_t>(B->mask) : nullptr;
  for (int i = 0; i < nint; ++i) {
    const SB_CONST int* rec = sched + i * 8;
    const int p = rec[0], l = rec[1], r = rec[2], sl = rec[4], sr = rec[5];
    const Kid kl = kid(l, sl), kr = kid(r, sr);
    // x = obs(p) * (pi at the root, else M_p^T d_p)
    float x[NV];
    int ex;
    bool stp;
    token_vec(token(p), x);
    if (i == 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) x[v] *= pis[v];
      ex = 0;
      stp = rbr;
    } else {
      float dp[NV], y[NV];
      load_vec(DD, p, dp);
      const int dw = DWd[(long long)p * kTW + wcol];
      ex = dw >> 1;
      stp = (dw & 1) != 0;
      lds_barrier();
      stage<AW, false>(sP[0], bt + (long long)p * BTS + sb_bt_Mt(AP), AP, tid, 256);
      lds_barrier();
      gemm<AW>(sP[0], dp, y, lane);
#pragma unroll
      for (int v = 0; v < NV; ++v) x[v] *= y[v];
    }
    lds_barrier();
    stage<AW, false>(sP[0], bt + (long long)l * BTS + sb_bt_M(AP), AP, tid, 256);
    stage<AW, false>(sP[1], bt + (long long)r * BTS + sb_bt_M(AP), AP, tid, 256);
    lds_barrier();
    float ul[NV], ur[NV], wl[NV], wr[NV];
    child_vecs(kl, sP[0], ul, wl);
    child_vecs(kr, sP[1], ur, wr);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const Kid& kn = s == 0 ? kl : kr;
      const Kid& ks = s == 0 ? kr : kl;
      const int sn = s == 0 ? sl : sr;
      float d[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) d[v] = x[v] * (s == 0 ? wr[v] : wl[v]);
      const int e = rescale(d);
      const int ed = ks.e + ex + e;
      const bool stn = (kn.leaf ? kn.has : kn.branch) || (stp && kn.has);
      if (sn >= 0) {
        store_vec(DD, kn.n, d);
        if (h == 0) DWd[(long long)kn.n * kTW + wcol] = (ed << 1) | (stn ? 1 : 0);
      }
      bool act;
      if (mask_mode == SB_MASK_NONE) act = true;
      else if (mask_mode == SB_MASK_AUTO) act = stn && stp;
      else act = emask[(long long)kn.n * C + cc] != 0;
      const float sc = act ? __builtin_amdgcn_ldexpf(inv_s0, ed + kn.e - e0) : 0.f;
      // projections: Ut = Vs^T u, Dt = Vi^T d * scale
      float t[NV];
      if (kn.leaf) gather(sVs, kn.tok, vs1r, t);
      else gemm<AW>(sVs, s == 0 ? ul : ur, t, lane);
      store_vec(UU, kn.n, t);
      gemm<AW>(sVi, d, t, lane);
#pragma unroll
      for (int v = 0; v < NV; ++v) t[v] *= sc;
      store_vec(DT, kn.n, t);
    }
  }
}

// ------------------------------------------------------------------ pass 2
template <int AW>
__global__ void __launch_bounds__(512) sb_gacc_kernel(const SbLaunch* __restrict__ Lp, int tile0) {
  constexpr int TC = kTC2;
  constexpr int JPT = AW * AW / (64 * kWPG2);   // J elements staged per thread
  // LDS: J double buffer, reused after the branch loop as one AW x AW
  // transpose buffer per wave; V^T for the back-transform; U~ broadcast rows
  constexpr int kAcc = kWPG2 * AW * AW;
  __shared__ __attribute__((aligned(16))) float sBig[kAcc];
  __shared__ __attribute__((aligned(16))) float sV[AW * AW];
  __shared__ __attribute__((aligned(16))) float sb[kWPG2][TC][AW];
  float* sJ0 = sBig;
  float* sJ1 = sBig + AW * AW;
  const SbLaunch& L = *Lp;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tl = blockIdx.x / kQ2, qq = blockIdx.x % kQ2;
  const SB_CONST int* tbp = cst<int>(L.tiles) + 2 * (tile0 + tl);
  const int bidx = tbp[0], tidx = tbp[1];
  const SB_CONST SbBlockDesc* B = cst<SbBlockDesc>(L.blocks) + bidx;
  const int flags = B->flags;
  if (!(flags & SB_F_COUNTS) || !B->counts) return;
  const long long C = B->C;
  const int A = B->A, R = B->R;
  const int AP = L.AP, RM = L.rmax;
  const int wcol0 = qq * (TC * kWPG2) + wv * TC;
  SB_GLOBAL char* st = glb<char>(L.stage) + (long long)tl * L.stage_stride;
  const long long vplane = (long long)RM * kTW * AW;
  const SB_GLOBAL float* UU = (const SB_GLOBAL float*)st;
  const SB_GLOBAL float* DT = UU + 2 * vplane;
  const SB_GLOBAL float* bt = glb<float>(B->bt);
  const SB_GLOBAL float* md = glb<float>(B->model);
  const SB_CONST float* mdc = cst<float>(B->model);
  const int BTS = sb_bt_stride(AP);
  const int k = lane;
  const bool on = k < AW;

  // V^T: sV[kk * AW + i] = V[i][kk]
  for (int idx = tid; idx < AW * AW; idx += 64 * kWPG2) {
    const int i = idx / AW, kk = idx % AW;
    sV[kk * AW + i] = (i < AP && kk < AP) ? md[sb_mr_V(AP) + i * AP + kk] : 0.f;
  }
  // J_n staged as sJ[q * AW + kk] = J_n[q][kk] (J is symmetric), pads zero
  float jr[JPT];
  auto load_J = [&](int n) {
    const SB_GLOBAL float* J = bt + (long long)n * BTS + sb_bt_J(AP);
#pragma unroll
    for (int m = 0; m < JPT; ++m) {
      const int idx = tid + m * 64 * kWPG2;
      const int q = idx / AW, kk = idx % AW;
      jr[m] = (q < AP && kk < AP) ? J[q * AP + kk] : 0.f;
    }
  };
  auto put_J = [&](float* dst) {
#pragma unroll
    for (int m = 0; m < JPT; ++m) dst[tid + m * 64 * kWPG2] = jr[m];
  };

  float acc[TC][AW];
#pragma unroll
  for (int c = 0; c < TC; ++c)
#pragma unroll
    for (int q = 0; q < AW; ++q) acc[c][q] = 0.f;
  if (R > 1) {
    load_J(1);
    put_J(sJ1);
    if (R > 2) load_J(2);
  }
  for (int n = 1; n < R; ++n) {
    lds_barrier();                       // J_n is in buffer n & 1; buffer (n+1) & 1 is free
    if (n + 1 < R) {
      put_J((n + 1) & 1 ? sJ1 : sJ0);
      if (n + 2 < R) load_J(n + 2);
    }
    const float* J = (n & 1) ? sJ1 : sJ0;
    float dk[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c) {
      const long long base = ((long long)n * kTW + wcol0 + c) * AW;
      dk[c] = on ? DT[base + k] : 0.f;
      if (on) sb[wv][c][k] = UU[base + k];
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < AW; q += 4) {
      float jq[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) jq[m] = J[(q + m) * AW + (on ? k : 0)];
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        const f32x4 u4 = *(const f32x4*)&sb[wv][c][q];
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[c][q + m] = fmaf(dk[c] * jq[m], u4[m], acc[c][q + m]);
      }
    }
    wave_lds_sync();
  }

  // ---------------- back-transform: counts = (V C V^T) o S ----------------
  lds_barrier();                         // every wave is done with the J buffers
  float* accb = sBig + wv * AW * AW;     // this wave's transpose buffer
  SB_GLOBAL float* out = glb<float>(B->counts);
  const SB_GLOBAL float* S = md + sb_mr_S(AP);
  const SB_CONST float* Vc = mdc + sb_mr_V(AP);
  const int i = lane;
#pragma unroll
  for (int c = 0; c < TC; ++c) {
    if (on) {
#pragma unroll
      for (int q = 0; q < AW; q += 4)
        *(f32x4*)&accb[k * AW + q] = f32x4{acc[c][q], acc[c][q + 1], acc[c][q + 2], acc[c][q + 3]};
    }
    wave_lds_sync();
    // tmp[q] = sum_kk V[i][kk] C[kk][q]
    float tmp[AW];
#pragma unroll
    for (int q = 0; q < AW; ++q) tmp[q] = 0.f;
    for (int kk = 0; kk < AP; ++kk) {
      const float vik = sV[kk * AW + (on ? i : 0)];
#pragma unroll
      for (int q = 0; q < AW; q += 4) {
        const f32x4 a4 = *(const f32x4*)&accb[kk * AW + q];
#pragma unroll
        for (int m = 0; m < 4; ++m) tmp[q + m] = fmaf(vik, a4[m], tmp[q + m]);
      }
    }
    wave_lds_sync();
    const long long col = (long long)tidx * kTW + wcol0 + c;
    if (col < C && i < A) {
      for (int jj = 0; jj < A; ++jj) {
        float sacc = 0.f;
        // tmp[q] = 0 for q >= A (pads of U~ are zero), so reading past row jj is harmless
#pragma unroll
        for (int q = 0; q < AW; ++q) sacc = fmaf(tmp[q], Vc[jj * AP + q], sacc);
        const float v = jj == i ? sacc : S[jj * AP + i] * sacc;   // S is symmetric
        out[((long long)i * A + jj) * C + col] = v;
      }
    }
  }
}

template __global__ void sb_gwalk_kernel<32>(const SbLaunch*, int);
template __global__ void sb_gwalk_kernel<64>(const SbLaunch*, int);
template __global__ void sb_gacc_kernel<32>(const SbLaunch*, int);
template __global__ void sb_gacc_kernel<64>(const SbLaunch*, int);

extern "C" int sb_gen2_tile_cols() { return kTW; }

extern "C" long long sb_gen2_tile_bytes(int AP, int rmax) {
  const long long AW = AP <= 32 ? 32 : 64;
  const long long b = 3LL * rmax * kTW * AW * 4 + 2LL * rmax * kTW * 4;
  return (b + 255) / 256 * 256;
}

extern "C" hipError_t sb_launch_gen2(int AP, int tile0, int ntiles, const SbLaunch* L, hipStream_t s) {
  if (AP > 64 || ntiles <= 0) return hipErrorInvalidValue;
  if (AP <= 32) {
    hipLaunchKernelGGL(sb_gwalk_kernel<32>, dim3(ntiles), dim3(64 * kWPG), 0, s, L, tile0);
    hipLaunchKernelGGL(sb_gacc_kernel<32>, dim3(ntiles * kQ2), dim3(64 * kWPG2), 0, s, L, tile0);
  } else {
    hipLaunchKernelGGL(sb_gwalk_kernel<64>, dim3(ntiles), dim3(64 * kWPG), 0, s, L, tile0);
    hipLaunchKernelGGL(sb_gacc_kernel<64>, dim3(ntiles * kQ2), dim3(64 * kWPG2), 0, s, L, tile0);
  }
  return hipGetLastError();
}
