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
// Pass 2 (sb_gacc_kernel): two columns per wave, eight waves share J_n(t)
// staged in LDS:  C[k][q] = sum_n Dt_n[k] J_n[k][q] Ut_n[q], the accumulator
// held in the matrix cores' output layout; then counts = (V C V^T) o S with
// the diagonal unscaled (eigensub.py:138-172) as two GEMMs on the matrix
// cores.  With branch counts, pass 2 is sb_gbranch_kernel instead.
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
constexpr int kWPG2 = 4;           // waves per workgroup, pass 2
// columns per wave, pass 2 (two waves per SIMD: at AW = 64 two columns' 128
// accumulators are half of the registers a wave can address)
template <int AW> struct Gacc { static constexpr int TC = AW == 64 ? 2 : 4; };

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

// Dense swizzled LDS layout of an AW x AW matrix: the 16-byte chunk j of row r
// sits at chunk r * (AW / 4) + (j ^ (r % (AW / 4))).  A wave's A-operand read
// (rows 16 t + i, chunk 4 t2 + h) puts its sixteen lanes of a quarter on
// sixteen different chunk columns, i.e. distinct banks.
template <int AW>
__device__ inline int swz(int r, int c) {
  constexpr int CPR = AW / 4;
  return r * AW + 4 * (((c >> 2) ^ r) & (CPR - 1)) + (c & 3);
}

// Stage an AP x AP row-major matrix into LDS as AW x AW in the swizzled
// layout with zero pads, transposed when TR.
template <int AW, bool TR>
__device__ inline void stage_sw(float* dst, const SB_GLOBAL float* src, int AP, int tid, int nthr) {
#pragma unroll 4
  for (int idx = tid; idx < AW * AW; idx += nthr) {
    const int r = idx / AW, c = idx % AW;
    const float v = (r < AP && c < AP) ? src[r * AP + c] : 0.f;
    dst[TR ? swz<AW>(c, r) : swz<AW>(r, c)] = v;
  }
}

// w = P u as gemm() below, P in the swizzled layout.  The A operands are read
// one K block ahead into two register sets, and scheduling barriers keep the
// compiler from hoisting more.
template <int AW>
__device__ inline void gemm_sw(const float* P, const float* u, float* w, int lane) {
  constexpr int NT = AW / 16;
  const int i = lane & 15, h = lane >> 4;
  f32x4 acc[NT], a[2][NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < NT; ++t) a[0][t] = *(const f32x4*)(P + swz<AW>(16 * t + i, 4 * h));
#pragma unroll
  for (int t2 = 0; t2 < NT; ++t2) {
    if (t2 + 1 < NT) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
        a[(t2 + 1) & 1][t] = *(const f32x4*)(P + swz<AW>(16 * t + i, 16 * (t2 + 1) + 4 * h));
    }
#pragma unroll
    for (int r2 = 0; r2 < 4; ++r2)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t2 & 1][t][r2], u[4 * t2 + r2], acc[t], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
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
//   UU[R][64][AW]  U_n of internal nodes (inside pass); W_n = M_n U_n of every
//                  branch replaces it at the parent's record (U_n is dead then)
//   UT[R][64][AW]  Ut_n = Vs^T U_n of every branch
//   DD[R][64][AW]  D_n of internal nodes
//   DT[R][64][AW]  Dt_n = Vi^T D_n * scale * mask of every branch
//   UW[R][64], DW[R][64]  ints: U word {exp << 2 | both | any}, D word {exp << 1 | steiner}
//
// Records are in depth-first preorder (the planner's order), so the first
// internal child of a record's node is the node of the record computed just
// before it in the inside pass (postorder, records walked backwards) and just
// after it in the outside pass: that vector is forwarded in registers.  Per
// record, the matrices arrive by direct-to-LDS loads one record ahead into
// double buffers (A: M_l or M_p^T, B: M_r), and the staged vectors it reads
// (the other internal child's U, both children's W, the node's D) by register
// loads one record ahead, so a record waits only on loads issued a whole record
// earlier.  One LDS matrix holds Vs^T during the inside pass and Vi^T during
// the outside pass: at AW = 64 the workgroup uses 80 KiB, two per CU.
// Rows that are loaded but not needed (a leaf's, a forwarded child's) are
// discarded by selects.
template <int AW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AW == 64 ? 1 : 2)))
sb_gwalk_kernel(const SbLaunch* __restrict__ Lp, int tile0, int slot0) {
  constexpr int NV = AW / 4;
  constexpr int CPR = AW / 4;
  constexpr int NTH = 64 * kWPG;
  constexpr int MPW = AW * AW / 4 / NTH;   // direct-to-LDS loads per thread per matrix
  static_assert(AW * AW / 4 % NTH == 0, "matrix chunks must split evenly over the workgroup");
  __shared__ __attribute__((aligned(16))) float sV[AW * AW];     // Vs^T, then Vi^T
  __shared__ __attribute__((aligned(16))) float sA0[AW * AW];
  __shared__ __attribute__((aligned(16))) float sA1[AW * AW];
  __shared__ __attribute__((aligned(16))) float sB0[AW * AW];
  __shared__ __attribute__((aligned(16))) float sB1[AW * AW];
  const SbLaunch& L = *Lp;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
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
  SB_GLOBAL char* st = glb<char>(L.stage) + (long long)(slot0 + tl) * L.stage_stride;
  const long long vplane = (long long)RM * kTW * AW;
  SB_GLOBAL float* UU = (SB_GLOBAL float*)st;
  SB_GLOBAL float* UT = UU + vplane;
  SB_GLOBAL float* DD = UT + vplane;
  SB_GLOBAL float* DT = DD + vplane;
  SB_GLOBAL int* UW = (SB_GLOBAL int*)(DT + vplane);
  SB_GLOBAL int* DWd = UW + (long long)RM * kTW;
  // per-lane data is addressed as a uniform base plus a 32-bit lane offset;
  // tokens and mask bytes are read unsigned (a gap token -1 reads as 255,
  // which every test below treats as a gap), so a prefetched byte needs no
  // conversion that would wait for its load
  const SB_GLOBAL uint8_t* aln = glb<uint8_t>(B->aln) + (long long)tidx * kTW;
  const unsigned lcol = unsigned(cc - (long long)tidx * kTW);     // this lane's column in the tile
  const SB_GLOBAL float* bt = glb<float>(B->bt);
  const SB_GLOBAL float* md = glb<float>(B->model);
  const SB_CONST int* sched = cst<int>(B->sched);
  const int BTS = sb_bt_stride(AP);
  const int mask_mode = (flags >> SB_F_MASK_SHIFT) & 3;
  // rows of the explicit mask (the alignment stands in when there is none; never used then)
  const SB_GLOBAL uint8_t* mrow = (B->mask ? glb<uint8_t>(B->mask) : glb<uint8_t>(B->aln)) + (long long)tidx * kTW;

  auto state = [&](int v) { return 16 * (v >> 2) + 4 * h + (v & 3); };
  auto real = [&](int v) { return state(v) < A ? 1.f : 0.f; };

  // this thread's chunks of a staged matrix (position p -> row r, chunk j);
  // a pad chunk (AP < AW) reads the matrix's first chunk: finite values that
  // only ever meet pad states, which are exactly zero in every vector or are
  // zeroed by the observation vector before use
  // (byte offsets, unsigned: a uniform base plus a 32-bit lane offset)
  unsigned moff[MPW];
#pragma unroll
  for (int m = 0; m < MPW; ++m) {
    const int p = (wv + kWPG * m) * 64 + lane;
    const int r = p / CPR, j = (p % CPR) ^ (r % CPR);
    moff[m] = (r < AP && 4 * j < AP) ? unsigned(4 * (r * AP + 4 * j)) : 0u;
  }
  auto dma = [&](float* dst, int n, int off) __attribute__((always_inline)) {
    const SB_GLOBAL char* src = (const SB_GLOBAL char*)(bt + (long long)n * BTS + off);
#pragma unroll
    for (int m = 0; m < MPW; ++m)
      __builtin_amdgcn_global_load_lds((const SB_GLOBAL void*)(src + moff[m]),
                                       (__attribute__((address_space(3))) void*)(dst + 4 * 64 * (wv + kWPG * m)),
                                       16, 0, 0);
  };

  const unsigned lvec = unsigned((wcol * AW + 4 * h) * 4);   // lane byte offset in a node's row of a vector plane
  auto vptr = [&](SB_GLOBAL float* base, int n) {
    return (SB_GLOBAL float*)((SB_GLOBAL char*)(base + (long long)n * kTW * AW) + lvec);
  };
  auto word_at = [&](SB_GLOBAL int* base, int n) -> SB_GLOBAL int& { return (base + (long long)n * kTW)[wcol]; };
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
  auto token = [&](int n) -> uint8_t { return (aln + (long long)n * C)[lcol]; };
  auto token_vec = [&](int tok, float* u) {
    const bool obs = tok < A;
#pragma unroll
    for (int v = 0; v < NV; ++v) u[v] = obs ? (state(v) == tok ? 1.f : 0.f) : real(v);
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

  // One tree record's indices, tokens and words (scalar record, per-lane loads)
  struct Rec {
    int p, l, r, sl, sr;
    uint8_t tp, tl, tr, mkl, mkr;     // (bytes as loaded: widened where used)
    int wl, wr;
  };
  auto fetch = [&](Rec& x, int k, bool outside) {
    const SB_CONST int* rec = sched + k * 8;
    x.p = rec[0]; x.l = rec[1]; x.r = rec[2]; x.sl = rec[4]; x.sr = rec[5];
    x.tp = token(x.p);
    x.tl = token(x.l);
    x.tr = token(x.r);
    x.wr = word_at(UW, x.r);
    x.wl = outside ? word_at(UW, x.l) : 0;
    x.mkl = outside ? (mrow + (long long)x.l * C)[lcol] : uint8_t(0);
    x.mkr = outside ? (mrow + (long long)x.r * C)[lcol] : uint8_t(0);
  };
  auto kid = [&](int n, int slot, int tok, int w) {
    Kid k;
    k.n = n;
    k.leaf = slot < 0;
    k.tok = tok;
    k.e = k.leaf ? 0 : w >> 2;
    k.has = k.leaf ? tok <= A : (w & 1) != 0;
    k.branch = !k.leaf && (w & 2) != 0;
    return k;
  };

  stage_sw<AW, true>(sV, md + sb_mr_Vs(AP), AP, tid, NTH);
  lds_barrier();
  // Vs^T u for an unobserved leaf: the sums of the rows of Vs^T over real states
  float sv1[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    float s = 0.f;
    for (int a = 0; a < A; ++a) s += sV[swz<AW>(state(v), a)];
    sv1[v] = s;
  }

  // ---------------- inside pass (postorder: records walked backwards) ----------------
  float u0[NV];
  int e0 = 0;
  bool rbr = false;
  if (nint == 0) token_vec(token(0), u0);
  if (nint > 0) {
    // ping-pong records: record i reads one while record i - 1's token and word
    // loads land in the other, so those are never copied (a copy of a loaded
    // value waits for its load); the prefetched vector is copied at the end of
    // the record, after its compute
    Rec ra, rb;
    float ur[NV], nur[NV];           // the right child's staged U (both children internal)
    float fu[NV];                    // U of the record computed last (its node is the first internal child)
    int fw = 0;
#pragma unroll
    for (int v = 0; v < NV; ++v) fu[v] = 0.f;
    fetch(ra, nint - 1, false);
    if (ra.sl >= 0 && ra.sr >= 0) load_vec(UU, ra.r, ur);
    dma((nint - 1) & 1 ? sA1 : sA0, ra.l, sb_bt_M(AP));
    dma((nint - 1) & 1 ? sB1 : sB0, ra.r, sb_bt_M(AP));
    auto istep = [&](int i, const Rec& cur, Rec& nxt) __attribute__((always_inline)) {
      __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0): this wave's loads for record i have landed
      lds_barrier();                        // ... and every wave's; the other buffers are free
      if (i > 0) {
        fetch(nxt, i - 1, false);
        if (nxt.sl >= 0 && nxt.sr >= 0) load_vec(UU, nxt.r, nur);   // (uniform)
        dma((i - 1) & 1 ? sA1 : sA0, nxt.l, sb_bt_M(AP));
        dma((i - 1) & 1 ? sB1 : sB0, nxt.r, sb_bt_M(AP));
      }
      const float* MA = i & 1 ? sA1 : sA0;
      const float* MB = i & 1 ? sB1 : sB0;
      // first internal child: the record computed last; the right child, when the
      // left one is internal too, from the stage area
      const bool lint = cur.sl >= 0;
      const Kid kl = kid(cur.l, cur.sl, cur.tl, fw);
      const Kid kr = kid(cur.r, cur.sr, cur.tr, lint ? cur.wr : fw);
      float wl[NV], wr[NV];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const Kid& k = s == 0 ? kl : kr;
        float* w = s == 0 ? wl : wr;
        const float* P = s == 0 ? MA : MB;
        if (k.leaf) {
          const bool obs = k.tok < A;
          float t[NV];
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            w[v] = obs ? P[swz<AW>(state(v), k.tok)] : real(v);
            t[v] = obs ? sV[swz<AW>(state(v), k.tok)] : sv1[v];
          }
          store_vec(UT, k.n, t);
        } else {
          float u[NV];
#pragma unroll
          for (int v = 0; v < NV; ++v) u[v] = (s == 1 && lint) ? ur[v] : fu[v];
          gemm_sw<AW>(P, u, w, lane);
        }
        store_vec(UU, k.n, w);       // W_n replaces U_n
      }
      float up[NV];
      token_vec(cur.tp, up);
#pragma unroll
      for (int v = 0; v < NV; ++v) up[v] *= wl[v] * wr[v];
      const int e = rescale(up);
      const int ep = kl.e + kr.e + e;
      const int word = (ep << 2) | ((kl.has && kr.has) ? 2 : 0) | ((kl.has || kr.has) ? 1 : 0);
      if (i > 0) {
        store_vec(UU, cur.p, up);
        if (h == 0) word_at(UW, cur.p) = word;
        float t[NV];
        gemm_sw<AW>(sV, up, t, lane);
        store_vec(UT, cur.p, t);
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) u0[v] = up[v];
        e0 = ep;
        rbr = kl.has && kr.has;
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        fu[v] = up[v];
        ur[v] = nur[v];
      }
      fw = word;
    };
    for (int i = nint - 1; i >= 0; i -= 2) {
      istep(i, ra, rb);
      if (i == 0) break;
      istep(i - 1, rb, ra);
    }
  }
  // log-likelihood (FP64 pi . u), root posterior
  const SB_CONST double* pi64 = cst<double>(B->model64) + sb_m64_pi(AP);
  const SB_GLOBAL float* pif = md + sb_mr_pi(AP);
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
      if (s < A) glb<float>(B->root)[(long long)s * C + cc] = pif[s] * u0[v] * inv_s0;
    }
  }
  if (!(flags & (SB_F_COUNTS | SB_F_BCOUNTS))) return;

  // ---------------- outside pass + Dt (preorder) ----------------
  __builtin_amdgcn_s_waitcnt(0x0f70);
  lds_barrier();                       // every wave is done with Vs^T
  stage_sw<AW, true>(sV, md + sb_mr_Vi(AP), AP, tid, NTH);
  {
    // ping-pong records as in the inside pass; the prefetched vectors (both
    // children's W, the node's D: loaded, or forwarded from the parent's record)
    // are copied at the end of the record
    Rec ra, rb;
    float wl[NV], wr[NV], dp[NV], nwl[NV], nwr[NV], ndp[NV];
    int dw = 0, ndw = 0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      dp[v] = 0.f;
      ndp[v] = 0.f;
    }
    fetch(ra, 0, true);
    load_vec(UU, ra.l, wl);
    load_vec(UU, ra.r, wr);
    auto ostep = [&](int i, const Rec& cur, Rec& nxt) __attribute__((always_inline)) {
      __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0): this wave's loads for record i have landed
      lds_barrier();                        // ... and every wave's (and Vi^T is staged)
      // the next record's node is a child of this one (forwarded) or was reached earlier
      const SB_CONST int* rn = sched + (i + 1 < nint ? i + 1 : i) * 8;
      const int np = rn[0];
      const bool nfwd = np == cur.l || np == cur.r;
      if (i + 1 < nint) {
        fetch(nxt, i + 1, true);
        load_vec(UU, nxt.l, nwl);
        load_vec(UU, nxt.r, nwr);
        if (!nfwd) {                   // (uniform)
          load_vec(DD, np, ndp);
          ndw = word_at(DWd, np);
        }
        dma((i + 1) & 1 ? sA1 : sA0, np, sb_bt_Mt(AP));
      }
      float x[NV];
      int ex;
      bool stp;
      if (i == 0) {
        token_vec(cur.tp, x);
#pragma unroll
        for (int v = 0; v < NV; ++v) x[v] *= state(v) < A ? pif[state(v)] : 0.f;
        ex = 0;
        stp = rbr;
      } else {
        // x = obs(p) * M_p^T d_p
        float o[NV];
        ex = dw >> 1;
        stp = (dw & 1) != 0;
        gemm_sw<AW>(i & 1 ? sA1 : sA0, dp, x, lane);
        token_vec(cur.tp, o);
#pragma unroll
        for (int v = 0; v < NV; ++v) x[v] *= o[v];
      }
      const Kid kl = kid(cur.l, cur.sl, cur.tl, cur.wl), kr = kid(cur.r, cur.sr, cur.tr, cur.wr);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const Kid& kn_ = s == 0 ? kl : kr;
        const Kid& ks = s == 0 ? kr : kl;
        const int sn = s == 0 ? cur.sl : cur.sr;
        float d[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) d[v] = x[v] * (s == 0 ? wr[v] : wl[v]);
        const int e = rescale(d);
        const int ed = ks.e + ex + e;
        const bool stn = (kn_.leaf ? kn_.has : kn_.branch) || (stp && kn_.has);
        const int dword = (ed << 1) | (stn ? 1 : 0);
        if (sn >= 0) {
          if (i + 1 < nint && np == kn_.n) {   // the next record's node: forwarded
#pragma unroll
            for (int v = 0; v < NV; ++v) ndp[v] = d[v];
            ndw = dword;
          } else {
            store_vec(DD, kn_.n, d);
            if (h == 0) word_at(DWd, kn_.n) = dword;
          }
        }
        bool act;
        if (mask_mode == SB_MASK_NONE) act = true;
        else if (mask_mode == SB_MASK_AUTO) act = stn && stp;
        else act = (s == 0 ? cur.mkl : cur.mkr) != 0;
        const float sc = act ? __builtin_amdgcn_ldexpf(inv_s0, ed + kn_.e - e0) : 0.f;
        float t[NV];
        gemm_sw<AW>(sV, d, t, lane);
#pragma unroll
        for (int v = 0; v < NV; ++v) t[v] *= sc;
        store_vec(DT, kn_.n, t);
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        wl[v] = nwl[v];
        wr[v] = nwr[v];
        dp[v] = ndp[v];
      }
      dw = ndw;
    };
    for (int i = 0; i < nint; i += 2) {
      ostep(i, ra, rb);
      if (i + 1 >= nint) break;
      ostep(i + 1, rb, ra);
    }
  }
}

// ------------------------------------------------------------------ pass 2
// C[k][q] = sum_n Dt_n[k] J_n[k][q] Ut_n[q] for TC = 2 columns per wave and
// WPG waves per workgroup.  The accumulator is held in the output layout of
// v_mfma_f32_16x16x4f32: lane (c = l % 16, h = l / 16) holds
// C[16 mt + 4 h + r][16 nq + c] for every tile (mt, nq) and r < 4.
//
// Operands arrive by direct-to-LDS loads (global_load_lds_dwordx4), one branch
// ahead into double buffers, so no register holds a prefetched operand and the
// registers left over keep LDS reads in flight:
//   * J_n (shared by the workgroup): AW x AW, 16-byte chunks stored densely
//     with chunk j of row r at r * CPR + (j ^ (r % CPR)), CPR = AW / 4 chunks
//     per row; a lane (c, h) reads J[16 nq + c][16 mt + 4 h .. + 3] (J is
//     symmetric: that is J_n[k][q] for its four k), and the sixteen lanes of a
//     quarter hit sixteen different chunk columns, i.e. distinct banks.  Pad
//     rows and columns (AP < AW) are never loaded and stay zero.
//   * Dt_n and Ut_n of the wave's two columns (per wave): 4 AW floats.
//
// Back-transform (eigensub.py:138-172) on the matrix cores without any
// transpose: as an A operand, a value in the output layout is the transpose of
// what it holds as a B operand.  So with the accumulator as A (it reads as
// C^T) and V^T from LDS as B, D1 = C^T V^T = (V C)^T; D1 as A reads as V C,
// and D2 = (V C) V^T = V C V^T; counts = D2 o S with the diagonal unscaled.
template <int AW, int WPG, int TC>
__global__ void __launch_bounds__(64 * WPG) __attribute__((amdgpu_waves_per_eu(2)))
sb_gacc_kernel(const SbLaunch* __restrict__ Lp, int tile0, int slot0) {
  constexpr int MT = AW / 16;
  constexpr int NA = MT * MT * 4;                 // accumulator registers per column
  constexpr int PS = AW + 4;
  constexpr int NTH = 64 * WPG;
  constexpr int CPR = AW / 4;                     // 16-byte chunks per row of J
  constexpr int JCH = AW * AW / 4;                // chunks of J
  constexpr int VF = 2 * TC * AW;                 // floats of a wave's Dt / Ut block
  constexpr int Q2 = kTW / (TC * WPG);            // workgroups per tile
  // (separate objects per buffer, selected statically, so the compiler can
  // tell the buffer being loaded from the one being read)
  __shared__ __attribute__((aligned(16))) float sJ0[AW * AW];
  __shared__ __attribute__((aligned(16))) float sJ1[AW * AW];
  __shared__ __attribute__((aligned(16))) float sW0[WPG][VF];
  __shared__ __attribute__((aligned(16))) float sW1[WPG][VF];
  __shared__ __attribute__((aligned(16))) float sV[AW * PS];   // V[i][k]
  const SbLaunch& L = *Lp;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int jc = lane & 15, h = lane >> 4;
  const int tl = blockIdx.x / Q2, qq = blockIdx.x % Q2;
  const SB_CONST int* tbp = cst<int>(L.tiles) + 2 * (tile0 + tl);
  const int bidx = tbp[0], tidx = tbp[1];
  const SB_CONST SbBlockDesc* B = cst<SbBlockDesc>(L.blocks) + bidx;
  const int flags = B->flags;
  if (!(flags & SB_F_COUNTS) || !B->counts) return;
  const long long C = B->C;
  const int A = B->A, R = B->R;
  const int AP = L.AP, RM = L.rmax;
  const int wcol0 = qq * (TC * WPG) + wv * TC;
  const SB_GLOBAL char* st = glb<char>(L.stage) + (long long)(slot0 + tl) * L.stage_stride;
  const long long vplane = (long long)RM * kTW * AW;
  const SB_GLOBAL float* UT = (const SB_GLOBAL float*)st + vplane;
  const SB_GLOBAL float* DT = UT + 2 * vplane;
  const SB_GLOBAL float* bt = glb<float>(B->bt);
  const SB_GLOBAL float* md = glb<float>(B->model);
  const int BTS = sb_bt_stride(AP);

  stage<AW, false>(sV, md + sb_mr_V(AP), AP, tid, NTH);

  // this lane's J chunks: position p = (wv + WPG m) * 64 + lane -> row r, chunk j.
  // Every lane loads (no branches in the loop): a pad chunk (AP < AW) reads
  // J_n[0][0..3] instead, finite values that only ever multiply the pad
  // states of Dt and Ut, which are exactly zero.
  constexpr int JPW = JCH / NTH;                 // load instructions per wave
  static_assert(JCH % NTH == 0 && (JPW > 0), "J chunks must split evenly over the workgroup");
  unsigned joff[JPW];                            // byte offset of the chunk in J_n
#pragma unroll
  for (int m = 0; m < JPW; ++m) {
    const int p = (wv + WPG * m) * 64 + lane;
    const int r = p / CPR, j = (p % CPR) ^ (r % CPR);
    joff[m] = (r < AP && 4 * j < AP) ? unsigned(4 * (r * AP + 4 * j)) : 0u;
  }
  // the wave's Dt / Ut block: chunks p = 64 m + lane, Dt first
  constexpr int VPW = VF / 4 / 64;               // load instructions per wave
  static_assert(VF % 256 == 0, "the Dt / Ut block must be whole wave loads");
  auto issue = [&](int n, float* sJb, float (*sWb)[VF]) __attribute__((always_inline)) {
    const SB_GLOBAL char* J = (const SB_GLOBAL char*)(bt + (long long)n * BTS + sb_bt_J(AP));
#pragma unroll
    for (int m = 0; m < JPW; ++m)
      __builtin_amdgcn_global_load_lds((const SB_GLOBAL void*)(J + joff[m]),
                                       (__attribute__((address_space(3))) void*)(sJb + 4 * 64 * (wv + WPG * m)),
                                       16, 0, 0);
    const long long base = ((long long)n * kTW + wcol0) * AW;
#pragma unroll
    for (int m = 0; m < VPW; ++m) {
      const int p = 64 * m + lane;
      __builtin_amdgcn_global_load_lds(
          (const SB_GLOBAL void*)(p < VF / 8 ? DT + base + 4 * p : UT + base + 4 * (p - VF / 8)),
          (__attribute__((address_space(3))) void*)(&sWb[wv][0] + 4 * 64 * m), 16, 0, 0);
    }
  };

  float acc[TC][NA];
#pragma unroll
  for (int c = 0; c < TC; ++c)
#pragma unroll
    for (int x = 0; x < NA; ++x) acc[c][x] = 0.f;

  auto step = [&](int n, const float* J, const float (*Wb)[VF], float* sJn, float (*sWn)[VF])
      __attribute__((always_inline)) {
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's loads of branch n have landed
    lds_barrier();                       // ... and every wave's; the other buffers are free
    issue(n + 1 < R ? n + 1 : n, sJn, sWn);   // (the last step reloads its own branch)
    const float* W = &Wb[wv][0];
    float uk[TC][MT];
#pragma unroll
    for (int c = 0; c < TC; ++c)
#pragma unroll
      for (int nq = 0; nq < MT; ++nq) uk[c][nq] = W[TC * AW + c * AW + 16 * nq + jc];
    // one K block mt of the state k at a time: its J rows and Dt chunks are
    // read, then used; a scheduling barrier between blocks keeps the
    // compiler from hoisting every read (and spilling accumulators): the
    // other wave on the SIMD covers the LDS latency
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      f32x4 jv[MT], dk[TC];
#pragma unroll
      for (int c = 0; c < TC; ++c) dk[c] = *(const f32x4*)(W + c * AW + 16 * mt + 4 * h);
#pragma unroll
      for (int nq = 0; nq < MT; ++nq) {
        const int r = 16 * nq + jc;
        jv[nq] = *(const f32x4*)(J + 4 * (r * CPR + ((4 * mt + h) ^ (r % CPR))));
      }
#pragma unroll
      for (int nq = 0; nq < MT; ++nq)
#pragma unroll
        for (int c = 0; c < TC; ++c)
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            float& a = acc[c][(mt * MT + nq) * 4 + x];
            a = fmaf(dk[c][x] * jv[nq][x], uk[c][nq], a);
          }
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  if (R > 1) issue(1, sJ1, sW1);
  int n = 1;
  for (; n + 1 < R; n += 2) {
    step(n, sJ1, sW1, sJ0, sW0);
    step(n + 1, sJ0, sW0, sJ1, sW1);
  }
  if (n < R) step(n, sJ1, sW1, sJ0, sW0);
  __builtin_amdgcn_s_waitcnt(0x0f70);    // no load into LDS outlives the loop

  // ---------------- back-transform on the matrix cores ----------------
  SB_GLOBAL float* out = glb<float>(B->counts);
  const SB_GLOBAL float* S = md + sb_mr_S(AP);
#pragma unroll
  for (int c = 0; c < TC; ++c) {
    // D1 = C^T V^T: tile (nq, ni), lane (jc, h) holds D1[16 nq + 4 h + r][16 ni + jc]
    f32x4 d1[MT][MT];
#pragma unroll
    for (int ni = 0; ni < MT; ++ni) {
      f32x4 b[MT];
#pragma unroll
      for (int t2 = 0; t2 < MT; ++t2) b[t2] = *(const f32x4*)(sV + (16 * ni + jc) * PS + 16 * t2 + 4 * h);
#pragma unroll
      for (int nq = 0; nq < MT; ++nq) {
        f32x4 t = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t2 = 0; t2 < MT; ++t2)
#pragma unroll
          for (int r2 = 0; r2 < 4; ++r2)
            t = __builtin_amdgcn_mfma_f32_16x16x4f32(acc[c][(t2 * MT + nq) * 4 + r2], b[t2][r2], t, 0, 0, 0);
        d1[nq][ni] = t;
      }
    }
    // D2 = (V C) V^T: tile (mi, nj), lane (jc, h) holds VCV[16 mi + 4 h + r][16 nj + jc]
    const long long col = (long long)tidx * kTW + wcol0 + c;
#pragma unroll
    for (int nj = 0; nj < MT; ++nj) {
      if (16 * nj >= A) break;
      f32x4 b[MT];
#pragma unroll
      for (int t2 = 0; t2 < MT; ++t2) b[t2] = *(const f32x4*)(sV + (16 * nj + jc) * PS + 16 * t2 + 4 * h);
      const int j = 16 * nj + jc;
#pragma unroll
      for (int mi = 0; mi < MT; ++mi) {
        if (16 * mi >= A) break;
        f32x4 o = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t2 = 0; t2 < MT; ++t2)
#pragma unroll
          for (int r2 = 0; r2 < 4; ++r2) o = __builtin_amdgcn_mfma_f32_16x16x4f32(d1[t2][mi][r2], b[t2][r2], o, 0, 0, 0);
        if (col < C && j < A) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * mi + 4 * h + r;
            if (i < A) out[((long long)i * A + j) * C + col] = i == j ? o[r] : S[i * AP + j] * o[r];
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------ pass 2, per-branch counts
// BranchCounts (subby/jax/__init__.py:450-540): for branch n >= 1 and column c,
//     counts_n,c = (V diag(Dt_c) J_n diag(Ut_c) V^T) o S     (diagonal unscaled)
// with Ut_c = Vs^T U_n, Dt_c = Vi^T D_n * scale * mask from pass 1.  Over a
// wave's 16 columns both products are GEMMs with a shared left operand:
//     W[k][(j,c)]   = sum_q J_n[k][q] (Ut_c[q] V[j][q])
//     out[i][(j,c)] = sum_k V[i][k] (Dt_c[k] W[k][(j,c)])
// on v_mfma_f32_16x16x4f32, N index (j, c) with c fastest: one N tile is one j
// for the wave's 16 columns, so every (i, j) row of the output is a 64-byte
// run of 16 consecutive columns.  K steps use the permuted index of pass 1
// (lane l, step (t2, r2): k = 16 t2 + 4 (l / 16) + r2), under which the output
// layout of the first GEMM (lane l: rows 16 mt + 4 (l / 16) + r) is already the
// B operand of the second, and Ut / Dt in pass 1's register layout are the
// factors each lane needs.  One workgroup of 4 waves per (tile, branch);
// branch 0 writes zeros.
template <int AW>
__global__ void __launch_bounds__(256) sb_gbranch_kernel(const SbLaunch* __restrict__ Lp, int tile0, int slot0) {
  constexpr int NV = AW / 4;
  constexpr int MT = AW / 16;
  constexpr int PS = AW + 4;
  constexpr int NTH = 64 * kWPG;
  __shared__ __attribute__((aligned(16))) float sJ[AW * PS];   // J_n[k][q]
  __shared__ __attribute__((aligned(16))) float sV[AW * PS];   // V[i][k]
  __shared__ __attribute__((aligned(16))) float sS[AW * PS];   // S[j][i]
  const SbLaunch& L = *Lp;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int jc = lane & 15, h = lane >> 4;
  const int RM = L.rmax;
  const int tl = blockIdx.x / RM, n = blockIdx.x % RM;
  const SB_CONST int* tbp = cst<int>(L.tiles) + 2 * (tile0 + tl);
  const int bidx = tbp[0], tidx = tbp[1];
  const SB_CONST SbBlockDesc* B = cst<SbBlockDesc>(L.blocks) + bidx;
  if (!(B->flags & SB_F_BCOUNTS) || !B->bcounts || n >= B->R) return;
  const long long C = B->C;
  const int A = B->A, AP = L.AP;
  const int wcol = wv * kWC + jc;
  const long long col = (long long)tidx * kTW + wcol;
  const bool valid = col < C;
  SB_GLOBAL float* out = glb<float>(B->bcounts) + (long long)n * A * A * C;
  if (n == 0) {   // the root has no branch: zeros
    if (valid)
      for (int e = h; e < A * A; e += 4) out[(long long)e * C + col] = 0.f;
    return;
  }
  const SB_GLOBAL float* md = glb<float>(B->model);
  const SB_GLOBAL float* bt = glb<float>(B->bt);
  stage<AW, false>(sJ, bt + (long long)n * sb_bt_stride(AP) + sb_bt_J(AP), AP, tid, NTH);
  stage<AW, false>(sV, md + sb_mr_V(AP), AP, tid, NTH);
  stage<AW, false>(sS, md + sb_mr_S(AP), AP, tid, NTH);   // S symmetric: S[j][i] = S[i][j]
  // this lane's factors: Ut_c and Dt_c at states 16 (v / 4) + 4 h + v % 4
  const SB_GLOBAL char* st = glb<char>(L.stage) + (long long)(slot0 + tl) * L.stage_stride;
  const long long vplane = (long long)RM * kTW * AW;
  const SB_GLOBAL float* UT = (const SB_GLOBAL float*)st + vplane;
  const SB_GLOBAL float* DT = UT + 2 * vplane;
  float ut[NV], dt[NV];
  {
    const SB_GLOBAL float* pu = UT + ((long long)n * kTW + wcol) * AW + 4 * h;
    const SB_GLOBAL float* pd = DT + ((long long)n * kTW + wcol) * AW + 4 * h;
#pragma unroll
    for (int t = 0; t < NV / 4; ++t) {
      const f32x4 x = *(const SB_GLOBAL f32x4*)(pu + 16 * t);
      const f32x4 y = *(const SB_GLOBAL f32x4*)(pd + 16 * t);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        ut[4 * t + r] = x[r];
        dt[4 * t + r] = y[r];
      }
    }
  }
  lds_barrier();
  for (int j = 0; j < A; ++j) {
    // B operand of the first GEMM: Ut_c[k] V[j][k] at this lane's K indices
    float b1[NV];
#pragma unroll
    for (int t2 = 0; t2 < MT; ++t2) {
      const f32x4 vj = *(const f32x4*)(sV + j * PS + 16 * t2 + 4 * h);
#pragma unroll
      for (int r2 = 0; r2 < 4; ++r2) b1[4 * t2 + r2] = ut[4 * t2 + r2] * vj[r2];
    }
    // W = J_n (Ut o V[j]) for the wave's 16 columns, rows scaled by Dt_c
    float w[NV];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t2 = 0; t2 < MT; ++t2) {
        const f32x4 a = *(const f32x4*)(sJ + (16 * mt + jc) * PS + 16 * t2 + 4 * h);
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r2], b1[4 * t2 + r2], acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) w[4 * mt + r] = acc[r] * dt[4 * mt + r];
    }
    // out = V W; lane l holds rows i = 16 mi + 4 h + r of column c = l % 16
#pragma unroll
    for (int mi = 0; mi < MT; ++mi) {
      if (16 * mi >= A) break;
      f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t2 = 0; t2 < MT; ++t2) {
        const f32x4 a = *(const f32x4*)(sV + (16 * mi + jc) * PS + 16 * t2 + 4 * h);
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r2], w[4 * t2 + r2], acc, 0, 0, 0);
      }
      const f32x4 sj = *(const f32x4*)(sS + j * PS + 16 * mi + 4 * h);
      if (valid) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * mi + 4 * h + r;
          if (i < A) out[((long long)i * A + j) * C + col] = i == j ? acc[r] : sj[r] * acc[r];
        }
      }
    }
  }
}

template __global__ void sb_gwalk_kernel<32>(const SbLaunch*, int, int);
template __global__ void sb_gwalk_kernel<64>(const SbLaunch*, int, int);
template __global__ void sb_gacc_kernel<32, kWPG2, Gacc<32>::TC>(const SbLaunch*, int, int);
template __global__ void sb_gacc_kernel<64, kWPG2, Gacc<64>::TC>(const SbLaunch*, int, int);
template __global__ void sb_gbranch_kernel<32>(const SbLaunch*, int, int);
template __global__ void sb_gbranch_kernel<64>(const SbLaunch*, int, int);

extern "C" int sb_gen2_tile_cols() { return kTW; }

extern "C" long long sb_gen2_tile_bytes(int AP, int rmax) {
  const long long AW = AP <= 32 ? 32 : 64;
  const long long b = 4LL * rmax * kTW * AW * 4 + 2LL * rmax * kTW * 4;
  return (b + 255) / 256 * 256;
}

// One pass of tiles tile0 .. tile0 + ntiles - 1 of the launch, whose stage
// slots are slot0 .. slot0 + ntiles - 1: pass 1 (the walk) or pass 2 (branch
// != 0: per-branch counts, rmax workgroups per tile; else summed counts).
extern "C" hipError_t sb_launch_gen2_pass(int AP, int tile0, int slot0, int ntiles, int rmax, int branch, int pass,
                                          const SbLaunch* L, hipStream_t s) {
  if (AP > 64 || ntiles <= 0 || rmax <= 0 || (pass != 1 && pass != 2)) return hipErrorInvalidValue;
  if (AP <= 32) {
    if (pass == 1)
      hipLaunchKernelGGL(sb_gwalk_kernel<32>, dim3(ntiles), dim3(64 * kWPG), 0, s, L, tile0, slot0);
    else if (branch)
      hipLaunchKernelGGL(sb_gbranch_kernel<32>, dim3(ntiles * rmax), dim3(64 * kWPG), 0, s, L, tile0, slot0);
    else
      hipLaunchKernelGGL((sb_gacc_kernel<32, kWPG2, Gacc<32>::TC>), dim3(ntiles * (kTW / (Gacc<32>::TC * kWPG2))),
                         dim3(64 * kWPG2), 0, s, L, tile0, slot0);
  } else {
    if (pass == 1)
      hipLaunchKernelGGL(sb_gwalk_kernel<64>, dim3(ntiles), dim3(64 * kWPG), 0, s, L, tile0, slot0);
    else if (branch)
      hipLaunchKernelGGL(sb_gbranch_kernel<64>, dim3(ntiles * rmax), dim3(64 * kWPG), 0, s, L, tile0, slot0);
    else
      hipLaunchKernelGGL((sb_gacc_kernel<64, kWPG2, Gacc<64>::TC>), dim3(ntiles * (kTW / (Gacc<64>::TC * kWPG2))),
                         dim3(64 * kWPG2), 0, s, L, tile0, slot0);
  }
  return hipGetLastError();
}

// Pass-1 workgroups one CU holds at once (its register and LDS budget).
extern "C" int sb_gen2_pass1_per_cu(int AP) { return AP <= 32 ? 2 : 1; }

// Both passes of one round of tiles (stage slots from 0).
extern "C" hipError_t sb_launch_gen2(int AP, int tile0, int ntiles, int rmax, int branch, const SbLaunch* L,
                                     hipStream_t s) {
  hipError_t e = sb_launch_gen2_pass(AP, tile0, 0, ntiles, rmax, branch, 1, L, s);
  if (e == hipSuccess) e = sb_launch_gen2_pass(AP, tile0, 0, ntiles, rmax, branch, 2, L, s);
  return e;
}
