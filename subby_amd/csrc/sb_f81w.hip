// F81 fast path for padded alphabets 16..64 in FP32 on gfx950 (MFMA accumulation).
//
// Applies to F81-structured models (one zero eigenvalue with eigenvector
// sqrt(pi), all other eigenvalues equal), for which
//     M(t) = e I + (1 - e) 1 pi^T,   e = exp(-mu t),
// so the inside pass (pruning.py:94-118) and the outside pass
// (outside.py:94-126) cost O(A) per column and branch.  The counts follow
// f81_fast.py:7-141:
//     I_ij = sum_n alpha_n Ds_i U_j + beta_n (Ds_i piU + pi_i sumDs U_j)
//            + gamma_n pi_i sumDs piU,
//     counts_ii = I_ii,  counts_ij = mu pi_j I_ij,
// with Ds = D * scale * mask.  The first term is, per column, a GEMM over the
// branches (K = 2 per internal node: its two child branches); it runs on the
// matrix cores (v_mfma_f32_32x32x2f32) with the AW x AW accumulator of each
// column in 16 * (AW/32)^2 accumulation registers.
//
// Layout: one wave owns a tile of 2 columns.  Lane l holds, for column l/32,
// states (l%32) + 32 j, j < AW/32, of every state vector, so per-column sums and
// maxima are four DPP steps inside a 16-lane row plus one row-pair swap, and
// every lane of the column ends with the identical total.  Two columns'
// accumulators fill half of the accumulation registers, so two waves share a
// SIMD and one's walk latency hides behind the other's work.  Scaled child
// vectors go through LDS into the MFMA operand layout (lane = state within a
// 32-state half x branch of the pair).
//
// Rescaling, Steiner flags and scales are those of the other kernels: exact
// power-of-two rescaling with integer exponents, scale = ldexp(1/s0, eD+eU-e0).
#include <hip/hip_runtime.h>

#include "../../include/subby_amd.h"
#include "sb_plan.h"

#define SB_CONST __attribute__((address_space(4)))
#define SB_GLOBAL __attribute__((address_space(1)))

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x2 __attribute__((ext_vector_type(2)));

constexpr int kLPC = 32;             // lanes per column
constexpr int kTC = SB_WAVE / kLPC;  // columns per wave
constexpr int kWG = 8;               // waves per workgroup: a group tile is kTC * kWG = 16 consecutive columns
constexpr int kXS = 20;  // row stride (floats) of the count exchange buffer: 16 columns + pad

template <typename T>
__device__ inline const SB_CONST T* cst(const void* p) { return (const SB_CONST T*)(p); }
template <typename T>
__device__ inline SB_GLOBAL T* glb(void* p) { return (SB_GLOBAL T*)(p); }
template <typename T>
__device__ inline const SB_GLOBAL T* glb(const void* p) { return (const SB_GLOBAL T*)(p); }

template <int CTRL>
__device__ inline float dppf(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ inline double dppd(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
// Butterfly over the 16 lanes of a row (quad swaps, half-row mirror, row
// mirror): every lane ends with the same total, bit for bit.
__device__ inline float row_sum(float v) {
  v += dppf<0xb1>(v);
  v += dppf<0x4e>(v);
  v += dppf<0x141>(v);
  v += dppf<0x140>(v);
  return v;
}
__device__ inline double row_sum(double v) {
  v += dppd<0xb1>(v);
  v += dppd<0x4e>(v);
  v += dppd<0x141>(v);
  v += dppd<0x140>(v);
  return v;
}
// Orders this wave's LDS writes before its later LDS reads (and reads before
// later writes).  LDS executes a wave's instructions in order, so no wait on
// global memory is needed, unlike __syncthreads(), whose fence drains every
// load and store in flight.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}
// Workgroup barrier ordering LDS only (global loads and stores stay in flight).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// Over the 32 lanes of a column (two rows): the row result, then the pair of
// rows traded by v_permlane16_swap; both rows see (row 0, row 1) in the same
// order, so every lane ends with the same bits.  (Vectors are rescaled by the
// power of two of their sum, which the walk needs anyway, so no maximum is
// reduced.)
__device__ inline float col_sum(float v) {
  v = row_sum(v);
  const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(a[0]) + __uint_as_float(a[1]);
}
__device__ inline double col_sum(double v) {
  v = row_sum(v);
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
}
// A lane offset the compiler cannot hoist out of the walk: a loop-invariant
// per-lane 64-bit address (base + offset) would cost a 64-bit add per access,
// while a uniform base plus this 32-bit offset is the load's own addressing.
__device__ inline unsigned opq(unsigned x) {
  asm volatile("" : "+v"(x));
  return x;
}
// the element at a uniform base plus a byte offset
template <typename T, typename P>
__device__ inline T& at(P* base, unsigned off) {
  return *(T*)((SB_GLOBAL char*)base + off);
}

template <int NS>
struct Child {
  float u[NS];
  float pu;   // pi . u
  int e;      // log2 scale
  bool has, branch, leaf;
};

}  // namespace

// A workgroup of eight waves takes a group tile of 16 consecutive columns of
// one block; wave w owns columns 2w, 2w + 1 and walks the tree on its own
// (wave-scope LDS exchanges only).  At the end the eight waves trade counts
// through LDS so that each count plane gets 64-byte stores of 16 columns.
// Scratch per wave: U and D vectors of internal nodes ([slot][NS][64] floats)
// followed by per-column words ([slot][kTC] int pairs: {word, pi.u or sum d}).
template <int AW>
__global__ void __launch_bounds__(64 * kWG) __attribute__((amdgpu_waves_per_eu(2)))
sb_f81w_kernel(const SbLaunch* __restrict__ Lp) {
  constexpr int NS = AW / kLPC;   // states per lane
  constexpr int NT = AW / 32;   // MFMA tiles per dimension
  typedef Child<NS> C_;
  const SbLaunch& L = *Lp;
  __shared__ float opnd_all[kWG][2][2][kTC][AW];   // per wave: [a|u][branch of the pair][column][state]
  __shared__ __attribute__((aligned(16))) float xb[256 * kXS];   // count exchange: [plane][column]
  __shared__ float piT[AW + 1];
  __shared__ int gtile;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // (wave-uniform)
  auto& opnd = opnd_all[wv];
  const int cl = lane / kLPC;     // column of this lane
  const int rl = lane % kLPC;
  const int AP = L.AP;
  const int NI = L.nint_max;
  const long long vbytes = (long long)NI * NS * SB_WAVE * 4;
  SB_GLOBAL char* wscr = glb<char>(L.scratch) + ((long long)blockIdx.x * kWG + wv) * L.wave_stride;
  SB_GLOBAL float* Ureg = (SB_GLOBAL float*)wscr;
  SB_GLOBAL float* Dreg = (SB_GLOBAL float*)(wscr + vbytes);
  SB_GLOBAL i32x2* Uw = (SB_GLOBAL i32x2*)(wscr + 2 * vbytes);
  SB_GLOBAL i32x2* Dw = Uw + (long long)NI * kTC;
  const int BTS = sb_bt_stride(AP);

  for (;;) {
    if (threadIdx.x == 0) gtile = atomicAdd(L.counter, 1);
    lds_barrier();
    const int t = __builtin_amdgcn_readfirstlane(gtile);
    lds_barrier();                 // everyone has read gtile before it is rewritten
    if (t >= L.ntiles) break;
    const SB_CONST int* tbp = cst<int>(L.tiles) + 2 * t;
    const int bidx = tbp[0], tidx = tbp[1];
    const SB_CONST SbBlockDesc* B = cst<SbBlockDesc>(L.blocks) + bidx;
    const long long C = B->C;
    const int A = B->A;
    const int flags = B->flags;
    const int nint = B->nint;
    const long long gcol0 = (long long)tidx * (kTC * kWG);   // first column of the group tile
    const long long col0 = gcol0 + kTC * wv;
    const long long mycol = col0 + cl;
    const bool valid = mycol < C;
    const long long cc = valid ? mycol : C - 1;
    const SB_GLOBAL int8_t* aln = glb<int8_t>(B->aln);
    // per-lane data addressed as a uniform base plus an unsigned 32-bit lane offset
    // (the group tile's first column is < C, so the clamped column cc of a lane
    // past the block's end is still at or after it: lcl is in 0 .. 15)
    const SB_GLOBAL int8_t* alnt = aln + gcol0;         // the group tile's first column
    const unsigned lcl = unsigned(cc - gcol0);          // this lane's column in the tile (clamped)
    const unsigned ul = unsigned(4 * lane);        // byte offsets
    const unsigned ucl = unsigned(8 * cl);
    const SB_CONST float* bt = cst<float>(B->bt);
    const SB_CONST float* mdc = cst<float>(B->model);
    const SB_CONST int* sched = cst<int>(B->sched);
    const int kstar = int(mdc[sb_mr_misc(AP) + 2]);
    float pis[NS], ones[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int s = rl + kLPC * j;
      pis[j] = s < A ? mdc[sb_mr_pi(AP) + s] : 0.f;
      ones[j] = s < A ? 1.f : 0.f;
    }
    if (wv == 0) {                 // (the previous tile's readers passed the barrier above)
      if (lane < AW) piT[lane] = lane < A ? mdc[sb_mr_pi(AP) + (lane < A ? lane : 0)] : 0.f;
      if (lane == 0) piT[AW] = 1.f;   // pi . (all ones)
    }
    lds_barrier();

    auto token = [&](int n) -> int { return (alnt + (long long)n * C)[opq(lcl)]; };
    auto emul = [&](int n) -> float { return bt[(long long)n * BTS + sb_bt_E(AP) + kstar]; };
    const int mask_mode = (flags >> SB_F_MASK_SHIFT) & 3;
    // rows of the explicit mask (the alignment stands in when there is none; never used then)
    const SB_GLOBAL uint8_t* mrow = (B->mask ? glb<uint8_t>(B->mask) : (const SB_GLOBAL uint8_t*)aln) + gcol0;

    // Everything one tree record needs is fetched one record ahead (its node
    // indices two ahead) with unconditional loads: slots of leaf children are
    // clamped to 0 and the values discarded, so no load or store sits in a
    // branch and the loops are unrolled by two over ping-pong buffers (no
    // loop-carried copies of loaded registers).
    struct Idx {
      int p, l, r, sl, sr;
    };
    auto fetch_idx = [&](Idx& x, int k) {
      const SB_CONST int* rec = sched + k * 8;
      x.p = rec[0]; x.l = rec[1]; x.r = rec[2]; x.sl = rec[4]; x.sr = rec[5];
    };
    struct Rec {
      Idx ix;
      int tp, tl, tr;
      float ul[NS], ur[NS];
      i32x2 wl, wr;
      float el, er;
    };
    auto fetch = [&](Rec& x, const Idx& ix) {
      x.ix = ix;
      x.tp = token(ix.p);
      x.tl = token(ix.l);
      x.tr = token(ix.r);
      const long long a = ix.sl < 0 ? 0 : ix.sl, b = ix.sr < 0 ? 0 : ix.sr;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        x.ul[j] = at<SB_GLOBAL float>(Ureg + (a * NS + j) * SB_WAVE, opq(ul));
        x.ur[j] = at<SB_GLOBAL float>(Ureg + (b * NS + j) * SB_WAVE, opq(ul));
      }
      x.wl = at<SB_GLOBAL i32x2>(Uw + a * kTC, opq(ucl));
      x.wr = at<SB_GLOBAL i32x2>(Uw + b * kTC, opq(ucl));
      x.el = emul(ix.l);
      x.er = emul(ix.r);
    };
    // one-hot / all-ones vector of a token, and pi . u of it
    auto token_vec = [&](int tok, float* u) -> float {
      const bool obs = tok >= 0 && tok < A;
#pragma unroll
      for (int j = 0; j < NS; ++j) u[j] = obs ? ((rl + kLPC * j) == tok ? 1.f : 0.f) : ones[j];
      return piT[obs ? tok : AW];
    };
    // child of a record: leaf (token), internal computed by the previous
    // record (forwarded in registers), or internal loaded from scratch
    auto child = [&](int slot, int tok, const float* ul, i32x2 wl, int fslot, const float* fu, int fw,
                     float fpu, C_& ch) {
      if (slot < 0) {
        ch.pu = token_vec(tok, ch.u);
        ch.e = 0;
        ch.has = tok >= 0 && tok <= A;
        ch.branch = false;
        ch.leaf = true;
      } else {
        const bool f = slot == fslot;
        const int w = f ? fw : wl.x;
#pragma unroll
        for (int j = 0; j < NS; ++j) ch.u[j] = f ? fu[j] : ul[j];
        ch.pu = f ? fpu : __int_as_float(wl.y);
        ch.e = w >> 2;
        ch.has = (w & 1) != 0;
        ch.branch = (w & 2) != 0;
        ch.leaf = false;
      }
    };
    // w = M u on real states
    auto apply_M = [&](float e, const C_& ch, float* w) {
      const float b = (1.f - e) * ch.pu;
#pragma unroll
      for (int j = 0; j < NS; ++j) w[j] = ones[j] * fmaf(e, ch.u[j], b);
    };
    auto dot_pi = [&](const float* v) -> float {
      float s = 0.f;
#pragma unroll
      for (int j = 0; j < NS; ++j) s = fmaf(pis[j], v[j], s);
      return col_sum(s);
    };

    // ---------------- inside pass (postorder) ----------------
    float u0[NS];
    int e0 = 0;
    bool rbr = false;
    if (nint == 0) (void)token_vec(token(0), u0);
    if (nint > 0) {
      Rec RA, RB;
      Idx i2;
      fetch_idx(i2, nint - 1);
      fetch(RA, i2);
      fetch_idx(i2, nint > 1 ? nint - 2 : 0);
      float pv[NS];       // result of the previous record (slot i + 1)
      int pw = 0;
      float ppu = 0.f;
#pragma unroll
      for (int j = 0; j < NS; ++j) pv[j] = 0.f;
      auto step = [&](int i, const Rec& cur, Rec& nxt) {
        fetch(nxt, i2);                          // record i - 1 (clamped)
        fetch_idx(i2, i > 1 ? i - 2 : 0);        // record i - 2 (clamped)
        C_ fl, fr;
        child(cur.ix.sl, cur.tl, cur.ul, cur.wl, i + 1, pv, pw, ppu, fl);
        child(cur.ix.sr, cur.tr, cur.ur, cur.wr, i + 1, pv, pw, ppu, fr);
        float wl[NS], up[NS], o[NS];
        apply_M(cur.el, fl, wl);
        apply_M(cur.er, fr, up);
        (void)token_vec(cur.tp, o);
#pragma unroll
        for (int j = 0; j < NS; ++j) up[j] = o[j] * wl[j] * up[j];
        // scaled by the power of two that puts pi . u in [0.5, 1): the sum is
        // needed anyway, and an exact power-of-two scale changes no rounding
        const float pu0 = dot_pi(up);
        const int e = __builtin_amdgcn_frexp_expf(pu0);
#pragma unroll
        for (int j = 0; j < NS; ++j) up[j] = __builtin_amdgcn_ldexpf(up[j], -e);
        const float pu = __builtin_amdgcn_ldexpf(pu0, -e);
        const int ep = fl.e + fr.e + e;
        const int word = (ep << 2) | ((fl.has && fr.has) ? 2 : 0) | ((fl.has || fr.has) ? 1 : 0);
        // (the root's own slot 0 only ever serves as the clamped dummy)
#pragma unroll
        for (int j = 0; j < NS; ++j) at<SB_GLOBAL float>(Ureg + ((long long)i * NS + j) * SB_WAVE, opq(ul)) = up[j];
        at<SB_GLOBAL i32x2>(Uw + i * kTC, opq(ucl)) = i32x2{word, __float_as_int(pu)};   // same value from the whole column
#pragma unroll
        for (int j = 0; j < NS; ++j) pv[j] = up[j];
        pw = word;
        ppu = pu;
        if (i == 0) {
#pragma unroll
          for (int j = 0; j < NS; ++j) u0[j] = up[j];
          e0 = ep;
          rbr = fl.has && fr.has;
        }
      };
      int i = nint - 1;
      for (; i >= 1; i -= 2) {
        step(i, RA, RB);
        step(i - 1, RB, RA);
      }
      if (i == 0) step(0, RA, RB);
    }
    // log-likelihood (FP64 pi . u), root posterior
    const SB_CONST double* pi64 = cst<double>(B->model64) + sb_m64_pi(AP);
    double s0 = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int s = rl + kLPC * j;
      if (s < A) s0 = fma(pi64[s], double(u0[j]), s0);
    }
    s0 = col_sum(s0);
    const float inv_s0 = float(1.0 / s0);
    if (valid && rl == 0 && (flags & SB_F_LL) && B->loglike)
      glb<double>(B->loglike)[cc] = double(e0) * 0.69314718055994530942 + log(s0);
    if (valid && (flags & SB_F_ROOT) && B->root) {
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const int s = rl + kLPC * j;
        if (s < A) glb<float>(B->root)[(long long)s * C + cc] = pis[j] * u0[j] * inv_s0;
      }
    }
    if (!(flags & SB_F_COUNTS)) continue;   // (uniform over the workgroup: one block)

    // ---------------- outside pass + accumulation (preorder) ----------------
    f32x16 acc[kTC][NT][NT];
#pragma unroll
    for (int c = 0; c < kTC; ++c)
#pragma unroll
      for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) acc[c][a][b] = f32x16{};
    float t2a[NS], t2b[NS], t3 = 0.f;
#pragma unroll
    for (int j = 0; j < NS; ++j) { t2a[j] = 0.f; t2b[j] = 0.f; }

    // outside-pass extras of a record: D of p (slot k), e of p's branch,
    // F81 coefficients of the child branches, explicit-mask bytes
    struct Out {
      float dp[NS];
      i32x2 wp;
      float ep, al[2], be[2], ga[2];
      int mk[2];
    };
    auto fetch_out = [&](Out& y, const Idx& ix, int k) {
#pragma unroll
      for (int j = 0; j < NS; ++j) y.dp[j] = at<SB_GLOBAL float>(Dreg + ((long long)k * NS + j) * SB_WAVE, opq(ul));
      y.wp = at<SB_GLOBAL i32x2>(Dw + k * kTC, opq(ucl));
      y.ep = emul(ix.p);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int n = s == 0 ? ix.l : ix.r;
        const SB_CONST float* cf = bt + (long long)n * BTS + sb_bt_coef(AP);
        y.al[s] = cf[0];
        y.be[s] = cf[1];
        y.ga[s] = cf[2];
        y.mk[s] = (mrow + (long long)n * C)[opq(lcl)];
      }
    };
    if (nint > 0) {
      Rec RA, RB;
      Out OA, OB;
      Idx i2;
      fetch_idx(i2, 0);
      fetch(RA, i2);
      fetch_out(OA, i2, 0);
      int k2 = nint > 1 ? 1 : 0;   // record held in i2
      fetch_idx(i2, k2);
      // D vectors computed by the previous record for its two children
      float pd[2][NS];
      int pdw[2] = {0, 0}, pn[2] = {-1, -1};
      float psd[2] = {0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < NS; ++j) pd[s][j] = 0.f;
      auto step = [&](int i, const Rec& cur, const Out& oc, Rec& nxt, Out& on) {
        fetch(nxt, i2);                 // record i + 1 (clamped)
        fetch_out(on, i2, k2);
        k2 = i + 2 < nint ? i + 2 : nint - 1;
        fetch_idx(i2, k2);              // record i + 2 (clamped)
        // x = obs(p) * (pi at the root, else M_p^T d_p)
        float x[NS];
        int ex;
        bool stp;
        {
          float o[NS];
          (void)token_vec(cur.tp, o);
          if (i == 0) {
#pragma unroll
            for (int j = 0; j < NS; ++j) x[j] = o[j] * pis[j];
            ex = 0;
            stp = rbr;
          } else {
            const bool f0 = pn[0] == cur.ix.p, f1 = pn[1] == cur.ix.p;
            const int w = f0 ? pdw[0] : f1 ? pdw[1] : oc.wp.x;
            const float sdp = f0 ? psd[0] : f1 ? psd[1] : __int_as_float(oc.wp.y);
            ex = w >> 1;
            stp = (w & 1) != 0;
            const float e = oc.ep;
            const float b = (1.f - e) * sdp;
#pragma unroll
            for (int j = 0; j < NS; ++j) {
              const float dpj = f0 ? pd[0][j] : f1 ? pd[1][j] : oc.dp[j];
              x[j] = o[j] * fmaf(e, dpj, b * pis[j]);
            }
          }
        }
        C_ ch[2];
        child(cur.ix.sl, cur.tl, cur.ul, cur.wl, -2, cur.ul, 0, 0.f, ch[0]);   // all U stored by now
        child(cur.ix.sr, cur.tr, cur.ur, cur.wr, -2, cur.ur, 0, 0.f, ch[1]);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int n = s == 0 ? cur.ix.l : cur.ix.r, sn = s == 0 ? cur.ix.sl : cur.ix.sr;
          const C_& fs = ch[1 - s];
          const C_& fn = ch[s];
          float d[NS];
          apply_M(s == 0 ? cur.er : cur.el, fs, d);
#pragma unroll
          for (int j = 0; j < NS; ++j) d[j] *= x[j];
          // scaled by the power of two that puts sum(d) in [0.5, 1) (the sum
          // is needed anyway; an exact power-of-two scale changes no rounding)
          float sd = 0.f;
#pragma unroll
          for (int j = 0; j < NS; ++j) sd += d[j];
          sd = col_sum(sd);
          const int e = __builtin_amdgcn_frexp_expf(sd);
#pragma unroll
          for (int j = 0; j < NS; ++j) d[j] = __builtin_amdgcn_ldexpf(d[j], -e);
          sd = __builtin_amdgcn_ldexpf(sd, -e);
          const int ed = fs.e + ex + e;
          const bool st = (fn.leaf ? fn.has : fn.branch) || (stp && fn.has);
          const int dword = (ed << 1) | (st ? 1 : 0);
          // leaf children write the never-read slot 0 (the root has no D)
          const long long ds_slot = sn < 0 ? 0 : sn;
#pragma unroll
          for (int j = 0; j < NS; ++j) at<SB_GLOBAL float>(Dreg + (ds_slot * NS + j) * SB_WAVE, opq(ul)) = d[j];
          at<SB_GLOBAL i32x2>(Dw + ds_slot * kTC, opq(ucl)) = i32x2{dword, __float_as_int(sd)};
#pragma unroll
          for (int j = 0; j < NS; ++j) pd[s][j] = d[j];
          pdw[s] = dword;
          psd[s] = sd;
          pn[s] = n;
          bool act;
          if (mask_mode == SB_MASK_NONE) act = true;
          else if (mask_mode == SB_MASK_AUTO) act = st && stp;
          else act = oc.mk[s] != 0;
          const float sc = act ? __builtin_amdgcn_ldexpf(inv_s0, ed + fn.e - e0) : 0.f;
          const float dsum = sd * sc;
          const float al = oc.al[s], be = oc.be[s];
#pragma unroll
          for (int j = 0; j < NS; ++j) {
            const float ds = d[j] * sc;
            t2a[j] = fmaf(be * ds, fn.pu, t2a[j]);
            t2b[j] = fmaf(be * dsum, fn.u[j], t2b[j]);
            opnd[0][s][cl][rl + kLPC * j] = al * ds;
            opnd[1][s][cl][rl + kLPC * j] = fn.u[j];
          }
          t3 = fmaf(oc.ga[s] * dsum, fn.pu, t3);
        }
        wave_sync();
        // acc_c += [a_l a_r] [u_l u_r]^T  (K = 2), per column and 32 x 32 tile
        const int kb = lane >> 5, ii = lane & 31;
        float av[kTC][NT], bv[kTC][NT];
#pragma unroll
        for (int c = 0; c < kTC; ++c)
#pragma unroll
          for (int a = 0; a < NT; ++a) {
            av[c][a] = opnd[0][kb][c][32 * a + ii];
            bv[c][a] = opnd[1][kb][c][32 * a + ii];
          }
#pragma unroll
        for (int c = 0; c < kTC; ++c)
#pragma unroll
          for (int a = 0; a < NT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b)
              acc[c][a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[c][a], bv[c][b], acc[c][a][b], 0, 0, 0);
        wave_sync();
      };
      int i = 0;
      for (; i + 1 < nint; i += 2) {
        step(i, RA, OA, RB, OB);
        step(i + 1, RB, OB, RA, OA);
      }
      if (i < nint) step(i, RA, OA, RB, OB);
    }

    // ---------------- epilogue: natural-basis counts ----------------
    // The rank-two correction t2a_i + pi_i (t2b_j + t3) is one more K = 2 step:
    // [t2a pi] [1 (t2b + t3)]^T.
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      opnd[0][0][cl][rl + kLPC * j] = t2a[j];
      opnd[0][1][cl][rl + kLPC * j] = pis[j];
      opnd[1][0][cl][rl + kLPC * j] = 1.f;
      opnd[1][1][cl][rl + kLPC * j] = t2b[j] + t3;
    }
    wave_sync();
    const int kb = lane >> 5, ii = lane & 31;
#pragma unroll
    for (int c = 0; c < kTC; ++c) {
      float av[NT], bv[NT];
#pragma unroll
      for (int a = 0; a < NT; ++a) {
        av[a] = opnd[0][kb][c][32 * a + ii];
        bv[a] = opnd[1][kb][c][32 * a + ii];
      }
#pragma unroll
      for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
          acc[c][a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[b], acc[c][a][b], 0, 0, 0);
    }
    // counts_ii = I_ii, counts_ij = mu pi_j I_ij; accumulator element q of tile
    // (a, b) in this lane is I[32a + 8(q/4) + 4(lane/32) + q%4][32b + lane%32].
    // Chunks of 256 planes (tile (a, b), q / 4 fixed) go through LDS: wave w
    // writes its kTC columns of each plane, then the workgroup stores 16
    // columns per plane.
    const float fmu = mdc[sb_mr_misc(AP)];
    SB_GLOBAL float* out = glb<float>(B->counts);
    const bool full = (C % 4 == 0) && (gcol0 + kTC * kWG <= C);
    if (out) {
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        const int j = 32 * b + ii;
        const float fj = fmu * piT[j];
#pragma unroll
        for (int a = 0; a < NT; ++a) {
#pragma unroll
          for (int qg = 0; qg < 4; ++qg) {
            lds_barrier();           // previous chunk's readers are done
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int q = 4 * qg + r;
              const int i = 32 * a + 8 * qg + 4 * kb + r;
              const float f = i == j ? 1.f : fj;
#pragma unroll
              for (int c = 0; c < kTC; ++c) xb[(r * 64 + lane) * kXS + kTC * wv + c] = f * acc[c][a][b][q];
            }
            lds_barrier();
#pragma unroll
            for (int m = 0; m < 256 * 4 / (64 * kWG); ++m) {
              const int e = threadIdx.x + 64 * kWG * m;   // float4 index: plane e / 4, part e % 4
              const int pl = e >> 2, part = e & 3;
              const int r = pl >> 6, Lp_ = pl & 63;
              const int i = 32 * a + 8 * qg + 4 * (Lp_ >> 5) + r;
              const int jj = 32 * b + (Lp_ & 31);
              if (i < A && jj < A) {
                const f32x4 v = *(const f32x4*)&xb[pl * kXS + 4 * part];
                const long long c0 = gcol0 + 4 * part;
                SB_GLOBAL float* dst = out + ((long long)i * A + jj) * C + c0;
                if (full) {
                  __builtin_nontemporal_store(v, (SB_GLOBAL f32x4*)dst);
                } else {
#pragma unroll
                  for (int z = 0; z < 4; ++z)
                    if (c0 + z < C) dst[z] = v[z];
                }
              }
            }
          }
        }
      }
    }
  }
}

template __global__ void sb_f81w_kernel<32>(const SbLaunch*);
template __global__ void sb_f81w_kernel<64>(const SbLaunch*);

static void (*pick_f81w(int AP))(const SbLaunch*) {
  if (AP <= 32) return sb_f81w_kernel<32>;
  if (AP <= 64) return sb_f81w_kernel<64>;
  return nullptr;
}

extern "C" int sb_f81w_tile_cols() { return kTC * kWG; }
extern "C" int sb_f81w_waves_per_group() { return kWG; }

extern "C" long long sb_f81w_wave_stride(int AP, int nint_max) {
  const int AW = AP <= 32 ? 32 : 64;
  const long long NS = AW / kLPC;
  return (long long)nint_max * (2 * NS * SB_WAVE * 4 + 2 * kTC * 8);
}

extern "C" int sb_f81w_occupancy(int AP) {
  void (*k)(const SbLaunch*) = pick_f81w(AP);
  int n = 0;
  if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, 64 * kWG, 0) != hipSuccess) return 0;
  return n * kWG;   // resident waves per CU
}

// nwaves: a multiple of kWG (one workgroup per kWG waves)
extern "C" hipError_t sb_launch_f81w(int AP, int nwaves, const SbLaunch* L, hipStream_t s) {
  void (*k)(const SbLaunch*) = pick_f81w(AP);
  if (!k) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3((nwaves + kWG - 1) / kWG), dim3(64 * kWG), 0, s, L);
  return hipGetLastError();
}
