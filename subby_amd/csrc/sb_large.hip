// Large-alphabet path (padded alphabet AP = 16, 20, 32 or 64), gfx950.
//
// One wave64 owns a tile of SB_LARGE_TC columns of one block; lane a holds
// component a of every state vector of those columns, so the inside pass
// (pruning.py:94-118), the outside pass (outside.py:94-126), the eigenbasis
// projection (eigensub.py:53-84) and the accumulation (eigensub.py:87-135 or
// f81_fast.py:7-141) run lane-parallel over states and loop over the tree.
// Products with an AP x AP matrix read one matrix row per lane-step from
// L1/L2 (coalesced) and broadcast the vector being contracted from LDS.
// Lane k keeps row k of the per-column A x A accumulator in registers.
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

// Columns per wave tile: 4 in FP32, 2 in FP64 (the accumulator rows are 2x wider).
template <typename T> struct TCof { static constexpr int v = SB_LARGE_TC; };
template <> struct TCof<double> { static constexpr int v = SB_LARGE_TC / 2; };

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

// Per-lane view of TC column vectors: v[c] = component `lane` of column c.
template <typename T>
struct Vec {
  static constexpr int TC = TCof<T>::v;
  T v[TC];
};

// w_c[lane] = sum_j P[j*AP + lane] * u_c[j]; lanes >= AP give 0.  P is first
// staged in LDS (sm[j*64 + lane], one coalesced row per step, all loads in
// flight together), u is published to LDS and read back as broadcasts.
template <typename T>
__device__ inline void contract(const SB_GLOBAL T* P, int AP, const Vec<T>& u, Vec<T>& w, T* bc, T* sm,
                                int lane) {
  constexpr int TC = TCof<T>::v;
  __syncthreads();
  const bool on = lane < AP;
#pragma unroll 8
  for (int j = 0; j < AP; ++j) sm[j * SB_WAVE + lane] = on ? P[j * AP + lane] : T(0);
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
template <typename T>
__device__ inline void f81_matvec(T e, T pia, T onea, const Vec<T>& u, Vec<T>& w) {
  constexpr int TC = TCof<T>::v;
#pragma unroll
  for (int c = 0; c < TC; ++c) {
    const T pu = wsum(pia * u.v[c]);
    w.v[c] = onea * fma(T(1) - e, pu, e * u.v[c]);
  }
}

// x = d^T M for an F81-structured M: x_a = e d_a + (1 - e) pi_a sum_i d_i.
template <typename T>
__device__ inline void f81_vecmat(T e, T pia, const Vec<T>& d, Vec<T>& x) {
  constexpr int TC = TCof<T>::v;
#pragma unroll
  for (int c = 0; c < TC; ++c) {
    const T sd = wsum(d.v[c]);
    x.v[c] = fma(T(1) - e, pia * sd, e * d.v[c]);
  }
}

template <typename T>
__device__ inline void token_vec(const SB_GLOBAL int8_t* aln, long long row_off, const long long* cc, int A,
                                 T onea, int lane, Vec<T>& u, int* tok) {
  constexpr int TC = TCof<T>::v;
#pragma unroll
  for (int c = 0; c < TC; ++c) {
    const int t = aln[row_off + cc[c]];
    tok[c] = t;
    const bool obs = t >= 0 && t < A;
    u.v[c] = obs ? (lane == t ? T(1) : T(0)) : onea;
  }
}

template <typename T>
__device__ inline void rescale(Vec<T>& w, int* e) {
  constexpr int TC = TCof<T>::v;
#pragma unroll
  for (int c = 0; c < TC; ++c) {
    const int x = frexp_exp(wmax(w.v[c]));
    e[c] = x;
    w.v[c] = ldexp_(w.v[c], -x);
  }
}

template <int TC>
struct Flags {
  int e[TC];
  bool has[TC], branch[TC];
  bool leaf;
};

}  // namespace

template <typename T, int AW, bool F81, bool GRAD>
__global__ void __launch_bounds__(64) sb_large_kernel(const SbLaunch* __restrict__ Lp) {
  constexpr int TC = TCof<T>::v;
  typedef Flags<TC> Flags;
  const SbLaunch& L = *Lp;
  __shared__ T bc[TC * SB_WAVE];
  __shared__ T sm[64 * 64];      // staged AP x AP matrix; also the epilogue accumulator
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
  const int BTS = sb_bt_stride(AP);

  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(L.counter, 1);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, SB_WAVE));
    if (t >= L.ntiles) break;
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
    const T onea = on ? md[sb_mr_ones(AP) + lane] : T(0);
    const T pia = on ? md[sb_mr_pi(AP) + lane] : T(0);
    const bool f81m = mdc[sb_mr_misc(AP) + 1] != T(0);    // F81-structured M
    const int kstar = int(mdc[sb_mr_misc(AP) + 2]);

    auto apply_M = [&](int n, const Vec<T>& u, Vec<T>& w) {
      const SB_GLOBAL T* bn = bt + (long long)n * BTS;
      if (f81m) f81_matvec<T>(bn[sb_bt_E(AP) + kstar], pia, onea, u, w);
      else contract<T>(bn + sb_bt_Mt(AP), AP, u, w, bc, sm, lane);
    };
    auto load_child = [&](int n, int slot, Vec<T>& u, Flags& f) {
      if (slot < 0) {
        int tok[TC];
        token_vec<T>(aln, (long long)n * C, cc, A, onea, lane, u, tok);
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          f.e[c] = 0;
          f.has[c] = tok[c] >= 0 && tok[c] <= A;
          f.branch[c] = false;
        }
        f.leaf = true;
      } else {
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          u.v[c] = Ureg[((long long)slot * TC + c) * SB_WAVE + lane];
          const int w = Uw[slot * TC + c];
          f.e[c] = w >> 2;
          f.has[c] = (w & 1) != 0;
          f.branch[c] = (w & 2) != 0;
        }
        f.leaf = false;
      }
    };

    // ---------------- inside pass ----------------
    Vec<T> u0;
    int e0[TC];
    bool rbr[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c) { e0[c] = 0; rbr[c] = false; }
    if (nint == 0) {
      int tok[TC];
      token_vec<T>(aln, 0, cc, A, onea, lane, u0, tok);
    }
    for (int i = nint - 1; i >= 0; --i) {
      const SB_CONST int* rec = sched + i * 8;
      const int p = rec[0], l = rec[1], r = rec[2], sl = rec[4], sr = rec[5];
      Vec<T> u, wl, up;
      Flags fl, fr;
      load_child(l, sl, u, fl);
      apply_M(l, u, wl);
      load_child(r, sr, u, fr);
      apply_M(r, u, up);
      int tok[TC];
      token_vec<T>(aln, (long long)p * C, cc, A, onea, lane, u, tok);
#pragma unroll
      for (int c = 0; c < TC; ++c) up.v[c] = u.v[c] * wl.v[c] * up.v[c];
      int e[TC];
      rescale<T>(up, e);
      if (i > 0) {
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          Ureg[((long long)i * TC + c) * SB_WAVE + lane] = up.v[c];
          const int ep = fl.e[c] + fr.e[c] + e[c];
          const int w = (ep << 2) | ((fl.has[c] && fr.has[c]) ? 2 : 0) | ((fl.has[c] || fr.has[c]) ? 1 : 0);
          if (lane == 0) Uw[i * TC + c] = w;
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
    const SB_CONST double* pi64 = cst<double>(B->model64) + sb_m64_pi(AP);
    T inv_s0[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c) {
      const double s0 = wsum((lane < A) ? pi64[lane < A ? lane : 0] * double(u0.v[c]) : 0.0);
      const double ll = double(e0[c]) * 0.69314718055994530942 + log(s0);
      inv_s0[c] = T(1.0 / s0);
      if (valid[c] && lane == 0 && (flags & SB_F_LL) && B->loglike) glb<double>(B->loglike)[cc[c]] = ll;
      if (valid[c] && lane < A && (flags & SB_F_ROOT) && B->root)
        glb<T>(B->root)[(long long)lane * C + cc[c]] = pia * u0.v[c] * inv_s0[c];
    }
    if (!(flags & (SB_F_COUNTS | SB_F_GRAD))) continue;

    // ---------------- outside pass + accumulation ----------------
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
    const SB_GLOBAL T* Vs = md + sb_mr_Vs(AP);
    const SB_GLOBAL T* Vi = md + sb_mr_Vi(AP);

    for (int i = 0; i < nint; ++i) {
      const SB_CONST int* rec = sched + i * 8;
      const int p = rec[0], l = rec[1], r = rec[2], sl = rec[4], sr = rec[5];
      Vec<T> x;
      int ex[TC];
      bool stp[TC];
      {
        Vec<T> o;
        int tok[TC];
        token_vec<T>(aln, (long long)p * C, cc, A, onea, lane, o, tok);
        if (i == 0) {
#pragma unroll
          for (int c = 0; c < TC; ++c) { x.v[c] = o.v[c] * pia; ex[c] = 0; stp[c] = rbr[c]; }
        } else {
          Vec<T> dp;
#pragma unroll
          for (int c = 0; c < TC; ++c) {
            dp.v[c] = Dreg[((long long)i * TC + c) * SB_WAVE + lane];
            const int w = Dw[i * TC + c];
            ex[c] = w >> 1;
            stp[c] = (w & 1) != 0;
          }
          const SB_GLOBAL T* bp = bt + (long long)p * BTS;
          if (f81m) f81_vecmat<T>(bp[sb_bt_E(AP) + kstar], pia, dp, x);
          else contract<T>(bp + sb_bt_M(AP), AP, dp, x, bc, sm, lane);
#pragma unroll
          for (int c = 0; c < TC; ++c) x.v[c] *= o.v[c];
        }
      }
#pragma unroll 1
      for (int s = 0; s < 2; ++s) {
        const int n = s == 0 ? l : r, sn = s == 0 ? sl : sr;
        const int sib = s == 0 ? r : l, ssib = s == 0 ? sr : sl;
        Vec<T> u, d;
        Flags fs, fn;
        load_child(sib, ssib, u, fs);
        apply_M(sib, u, d);
        int e[TC], ed[TC];
#pragma unroll
        for (int c = 0; c < TC; ++c) d.v[c] *= x.v[c];
        rescale<T>(d, e);
        load_child(n, sn, u, fn);
        bool act[TC];
        T sc[TC], scr[TC];
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          ed[c] = fs.e[c] + ex[c] + e[c];
          const bool st = (fn.leaf ? fn.has[c] : fn.branch[c]) || (stp[c] && fn.has[c]);
          if (sn >= 0) {
            Dreg[((long long)sn * TC + c) * SB_WAVE + lane] = d.v[c];
            if (lane == 0) Dw[sn * TC + c] = (ed[c] << 1) | (st ? 1 : 0);
          }
          if (GRAD || mask_mode == SB_MASK_NONE) act[c] = true;
          else if (mask_mode == SB_MASK_AUTO) act[c] = st && stp[c];
          else act[c] = emask[(long long)n * C + cc[c]] != 0;
          scr[c] = ldexp_(inv_s0[c], ed[c] + fn.e[c] - e0[c]);
          sc[c] = act[c] ? scr[c] : T(0);
        }
        const SB_GLOBAL T* bn = bt + (long long)n * BTS;
        if constexpr (F81) {
          const T al = bn[sb_bt_coef(AP) + 0], be = bn[sb_bt_coef(AP) + 1], ga = bn[sb_bt_coef(AP) + 2];
          __syncthreads();
#pragma unroll
          for (int c = 0; c < TC; ++c) bc[c * SB_WAVE + lane] = u.v[c];
          __syncthreads();
#pragma unroll
          for (int c = 0; c < TC; ++c) {
            const T Ds = d.v[c] * sc[c];
            const T piU = wsum(pia * u.v[c]);
            const T Dsum = wsum(Ds);
            const T ad = al * Ds;
#pragma unroll
            for (int q = 0; q < AW; ++q)
              if (q < AP) acc[c][q] = fma(ad, bc[c * SB_WAVE + q], acc[c][q]);
            t2a[c] = fma(be * Ds, piU, t2a[c]);
            t2b[c] = fma(be * Dsum, u.v[c], t2b[c]);
            t3[c] = fma(ga * Dsum, piU, t3[c]);
          }
        } else {
          Vec<T> Ut, Dt;
          contract<T>(Vs, AP, u, Ut, bc, sm, lane);
          contract<T>(Vi, AP, d, Dt, bc, sm, lane);
          __syncthreads();
#pragma unroll
          for (int c = 0; c < TC; ++c) bc[c * SB_WAVE + lane] = Ut.v[c];
          __syncthreads();
          const SB_GLOBAL T* J = bn + sb_bt_J(AP);   // symmetric: row q read across lanes
          T dk[TC];
#pragma unroll
          for (int c = 0; c < TC; ++c) dk[c] = Dt.v[c] * sc[c];
#pragma unroll
          for (int q = 0; q < AW; ++q) {
            if (q < AP) {
              const T jq = on ? J[q * AP + lane] : T(0);
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
      }
    }

    // ---------------- epilogue: natural-basis counts per column ----------------
    const SB_GLOBAL T* V = md + sb_mr_V(AP);
    const SB_CONST T* Vc = mdc + sb_mr_V(AP);
    const SB_GLOBAL T* S = md + sb_mr_S(AP);
    const T fmu = mdc[sb_mr_misc(AP)];
    double rsum[SB_MAX_RATE];
#pragma unroll
    for (int q = 0; q < SB_MAX_RATE; ++q) rsum[q] = 0.0;
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
        for (int q = 0; q < AW; ++q) accb[lane * 64 + q] = acc[c][q];
        __syncthreads();
        T tmp[AW];
#pragma unroll
        for (int q = 0; q < AW; ++q) tmp[q] = T(0);
        for (int k = 0; k < AP; ++k) {
          const T vik = on ? V[lane * AP + k] : T(0);
#pragma unroll
          for (int q = 0; q < AW; ++q) tmp[q] = fma(vik, accb[k * 64 + q], tmp[q]);
        }
        __syncthreads();
        for (int jj = 0; jj < AP; ++jj) {
          T sacc = T(0);
#pragma unroll
          for (int q = 0; q < AW; ++q)
            if (q < AP) sacc = fma(tmp[q], Vc[jj * AP + q], sacc);
          accb[jj * SB_WAVE + lane] = (jj == lane) ? sacc : (on ? S[lane * AP + jj] : T(0)) * sacc;
        }
      }
      if (valid[c] && (flags & SB_F_COUNTS) && B->counts && lane < A) {
        SB_GLOBAL T* out = glb<T>(B->counts);
        for (int jj = 0; jj < A; ++jj) out[((long long)lane * A + jj) * C + cc[c]] = accb[jj * SB_WAVE + lane];
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
    if constexpr (GRAD) {
      SB_GLOBAL double* rpart = glb<double>(L.rate_part) + (long long)(B->tile_base + tidx) * SB_MAX_RATE;
#pragma unroll
      for (int q = 0; q < SB_MAX_RATE; ++q)
        if (lane == 0 && q < B->n_rate) rpart[q] = rsum[q];
    }
  }
}

#define SB_LINST(T, AW)                                                        \
  template __global__ void sb_large_kernel<T, AW, false, false>(const SbLaunch*); \
  template __global__ void sb_large_kernel<T, AW, true, false>(const SbLaunch*);  \
  template __global__ void sb_large_kernel<T, AW, false, true>(const SbLaunch*);
SB_LINST(float, 32)
SB_LINST(float, 64)
SB_LINST(double, 32)
SB_LINST(double, 64)

extern "C" int sb_large_tile_cols(int dtype) { return dtype == 0 ? TCof<float>::v : TCof<double>::v; }

extern "C" long long sb_large_wave_stride(int dtype, int AP, int nint_max) {
  const long long esz = dtype == 0 ? 4 : 8;
  const int TC = sb_large_tile_cols(dtype);
  (void)AP;
  return (long long)nint_max * (2 * TC * SB_WAVE * esz + 2 * TC * 4);
}

static void (*pick_large(int dtype, int AP, bool f81, bool grad))(const SbLaunch*) {
  void (*k)(const SbLaunch*) = nullptr;
#define SB_LPICK(T, AWV)                                             \
  {                                                                  \
    if (grad) k = sb_large_kernel<T, AWV, false, true>;              \
    else if (f81) k = sb_large_kernel<T, AWV, true, false>;          \
    else k = sb_large_kernel<T, AWV, false, false>;                  \
  }
  if (dtype == 0) {
    if (AP <= 32) SB_LPICK(float, 32) else SB_LPICK(float, 64)
  } else {
    if (AP <= 32) SB_LPICK(double, 32) else SB_LPICK(double, 64)
  }
#undef SB_LPICK
  return k;
}

extern "C" int sb_large_occupancy(int dtype, int AP, bool f81, bool grad) {
  void (*k)(const SbLaunch*) = pick_large(dtype, AP, f81, grad);
  int n = 0;
  if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, 64, 0) != hipSuccess) return 0;
  return n;
}

extern "C" hipError_t sb_launch_large(int dtype, int AP, bool f81, bool grad, int, int nwaves,
                                      const SbLaunch* L, hipStream_t s) {
  void (*k)(const SbLaunch*) = pick_large(dtype, AP, f81, grad);
  if (!k || AP > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3(nwaves), dim3(64), 0, s, L);
  return hipGetLastError();
}

