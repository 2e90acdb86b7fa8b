// Per-branch counts and posteriors of the inside-outside tables, small
// alphabets (padded to 4 or 8), gfx950.
//
// Outputs (any subset), per column:
//   branch counts  [R][A][A][C]  per-branch dwell / substitution counts
//                  (accumulate_C_per_branch + back_transform_per_branch,
//                  eigensub.py:175-231; f81_counts_per_branch, f81_fast.py:144-);
//                  branch 0 zeros
//   node posterior [R][A][C]     P(x_n = j | column) (InsideOutside.node_posterior,
//                  subby/jax/__init__.py:864-946): root pi_a U_0a, others
//                  (sum_a D_na M_n[a][j]) U_nj, normalised
//   branch joint   [R][A][A][C]  P(x_parent = i, x_n = j | column)
//                  (InsideOutside.branch_posterior, :948-1040): D_ni M_n[i][j] U_nj,
//                  normalised; branch 0 zeros
// Both posteriors share the normaliser Z = D_n . (M_n U_n), so the scales cancel.
//
// One thread = one column, one wave64 = 64 columns of one block, persistent
// over tiles; the inside pass stores internal U (as the fused kernel does),
// the outside pass forms D for both children of each internal node and writes
// the requested outputs of both branches, all stores coalesced over columns.
// Rescaling, Steiner flags and scales follow the fused kernel.
#include <hip/hip_runtime.h>

#include "../../include/subby_amd.h"
#include "sb_plan.h"

#define SB_CONST __attribute__((address_space(4)))
#define SB_GLOBAL __attribute__((address_space(1)))
#define SB_INLINE __attribute__((always_inline))

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

template <typename T, int AP>
__device__ inline void matvec(const SB_CONST T* M, const T* u, T* w) {   // w_a = sum_j M[a][j] u_j
#pragma unroll
  for (int a = 0; a < AP; ++a) {
    T s = M[a * AP] * u[0];
#pragma unroll
    for (int j = 1; j < AP; ++j) s = fma(M[a * AP + j], u[j], s);
    w[a] = s;
  }
}
template <typename T, int AP>
__device__ inline void vecmat(const T* d, const SB_CONST T* M, T* x) {   // x_a = sum_i d_i M[i][a]
#pragma unroll
  for (int a = 0; a < AP; ++a) {
    T s = d[0] * M[a];
#pragma unroll
    for (int i = 1; i < AP; ++i) s = fma(d[i], M[i * AP + a], s);
    x[a] = s;
  }
}
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

template <typename T, int AP>
struct Kid {
  T u[AP];
  int e, n, tok;
  bool has, branch, leaf;
};

template <typename T>
__device__ inline void out_store(SB_GLOBAL T* p, T v) { __builtin_nontemporal_store(v, p); }

}  // namespace

template <typename T, int AP, bool F81, bool IRR>
__global__ void __launch_bounds__(64, (sizeof(T) == 4 && AP == 4) ? 4 : 1) sb_post_kernel(const SbLaunch* __restrict__ Lp) {
  const SbLaunch& L = *Lp;
  const int lane = threadIdx.x;
  const int NI = L.nint_max;
  SB_GLOBAL char* wscr = glb<char>(L.scratch) + (long long)blockIdx.x * L.wave_stride;
  SB_GLOBAL T* Ureg = (SB_GLOBAL T*)wscr;                    // [slot][AP][64]
  SB_GLOBAL T* Dreg = Ureg + (long long)NI * AP * SB_WAVE;   // [slot][AP][64]
  SB_GLOBAL int* Uw = (SB_GLOBAL int*)(Dreg + (long long)NI * AP * SB_WAVE);
  SB_GLOBAL int* Dw = Uw + (long long)NI * SB_WAVE;
  const int BTS = sb_bt_stride_k(AP, IRR);

  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(L.counter, 1);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, SB_WAVE));
    if (t >= L.ntiles) break;
    const SB_CONST int* tbp = cst<int>(L.tiles) + 2 * t;
    const SB_CONST SbBlockDesc* B = cst<SbBlockDesc>(L.blocks) + tbp[0];
    const long long C = B->C;
    const int A = B->A;
    const int flags = B->flags;
    const int nint = B->nint;
    const long long col = (long long)tbp[1] * SB_WAVE + lane;
    const bool valid = col < C;
    const long long cc = valid ? col : C - 1;
    const SB_GLOBAL int8_t* aln = glb<int8_t>(B->aln) + cc;
    const SB_CONST T* bt = cst<T>(B->bt);
    const SB_CONST T* md = cst<T>(B->model);
    const SB_CONST int* sched = cst<int>(B->sched);
    const SB_CONST T* pi = md + sb_mr_pi_k(AP, IRR);
    const SB_CONST T* ones = md + sb_mr_ones_k(AP, IRR);
    const SB_CONST T* pr = B->root_model ? cst<T>(B->root_model) + sb_mr_pi_k(AP, IRR) : pi;
    auto Mat = [&](int n) SB_INLINE { return bt + (long long)n * BTS + sb_bt_M(AP); };
    auto vstore = [&](SB_GLOBAL T* base, int slot, const T* v) SB_INLINE {
#pragma unroll
      for (int a = 0; a < AP; ++a) base[((long long)slot * AP + a) * SB_WAVE + lane] = v[a];
    };
    auto vload = [&](const SB_GLOBAL T* base, int slot, T* v) SB_INLINE {
#pragma unroll
      for (int a = 0; a < AP; ++a) v[a] = base[((long long)slot * AP + a) * SB_WAVE + lane];
    };
    auto token_vec = [&](int tok, T* u) SB_INLINE {
      const bool obs = tok >= 0 && tok < A;
#pragma unroll
      for (int a = 0; a < AP; ++a) u[a] = obs ? (a == tok ? T(1) : T(0)) : ones[a];
    };
    // Loads of one child, issued a record ahead and decoded when used (so a
    // wait on them never waits for the previous record's stores).  slot < 0:
    // a leaf (token of row n); else the inside vector and word of that slot.
    // The branch is uniform over the wave.
    auto fetch = [&](int n, int slot, T* u, int& w, int& tok) SB_INLINE {
      if (slot < 0) {
        tok = aln[(long long)n * C];
      } else {
        vload(Ureg, slot, u);
        w = Uw[slot * SB_WAVE + lane];
      }
    };
    auto decode = [&](Kid<T, AP>& k, int n, int slot, const T* u, int w, int tok) SB_INLINE {
      k.n = n;
      k.leaf = slot < 0;
      if (k.leaf) {
        k.tok = tok;
        token_vec(tok, k.u);
        k.e = 0;
        k.has = tok >= 0 && tok <= A;
        k.branch = false;
      } else {
#pragma unroll
        for (int a = 0; a < AP; ++a) k.u[a] = u[a];
        k.tok = 0;
        k.e = w >> 2;
        k.has = (w & 1) != 0;
        k.branch = (w & 2) != 0;
      }
    };

    // ---------------- inside pass (postorder) ----------------
    T u0[AP];
    int e0 = 0;
    bool rbr = false;
    if (nint == 0) token_vec(aln[0], u0);
    T lu[AP], ru[AP];
    int lw = 0, rw = 0, lt = 0, rt = 0, pt = 0;
    if (nint > 0) {
      const SB_CONST int* rec = sched + (nint - 1) * 8;
      fetch(rec[1], rec[4], lu, lw, lt);
      fetch(rec[2], rec[5], ru, rw, rt);
      pt = aln[(long long)rec[0] * C];
    }
    for (int i = nint - 1; i >= 0; --i) {
      const SB_CONST int* rec = sched + i * 8;
      const SB_CONST int* rn = sched + (i > 0 ? i - 1 : 0) * 8;
      T nlu[AP], nru[AP];
      int nlw = 0, nrw = 0, nlt = 0, nrt = 0, npt = 0;
      if (i > 0) {
        fetch(rn[1], rn[4], nlu, nlw, nlt);
        fetch(rn[2], rn[5], nru, nrw, nrt);
        npt = aln[(long long)rn[0] * C];
      }
      Kid<T, AP> kl, kr;
      decode(kl, rec[1], rec[4], lu, lw, lt);
      decode(kr, rec[2], rec[5], ru, rw, rt);
      T wl[AP], up[AP];
      matvec<T, AP>(Mat(kl.n), kl.u, wl);
      matvec<T, AP>(Mat(kr.n), kr.u, up);
      T o[AP];
      token_vec(pt, o);
#pragma unroll
      for (int a = 0; a < AP; ++a) up[a] *= o[a] * wl[a];
      const int ep = kl.e + kr.e + rescale<T, AP>(up);
      if (i > 0) {
        const int word = (ep << 2) | ((kl.has && kr.has) ? 2 : 0) | ((kl.has || kr.has) ? 1 : 0);
        vstore(Ureg, i, up);
        Uw[i * SB_WAVE + lane] = word;
        // the next record's loads of slot i went out before this store
        if (rn[4] == i) {
#pragma unroll
          for (int a = 0; a < AP; ++a) nlu[a] = up[a];
          nlw = word;
        }
        if (rn[5] == i) {
#pragma unroll
          for (int a = 0; a < AP; ++a) nru[a] = up[a];
          nrw = word;
        }
      } else {
#pragma unroll
        for (int a = 0; a < AP; ++a) u0[a] = up[a];
        e0 = ep;
        rbr = kl.has && kr.has;
      }
#pragma unroll
      for (int a = 0; a < AP; ++a) {
        lu[a] = nlu[a];
        ru[a] = nru[a];
      }
      lw = nlw;
      rw = nrw;
      lt = nlt;
      rt = nrt;
      pt = npt;
    }
    // root distribution of the passes: the model's own or (model lists) the
    // first column model's; the root row of the node posteriors keeps the own
    const SB_CONST double* pi64 = cst<double>(B->model64) + sb_m64_pi_k(AP, IRR);
    const SB_CONST double* pr64 = B->root_model64 ? cst<double>(B->root_model64) + sb_m64_pi_k(AP, IRR) : pi64;
    double s0 = 0.0, s0own = 0.0;
#pragma unroll
    for (int a = 0; a < AP; ++a)
      if (a < A) {
        s0 = fma(pr64[a], double(u0[a]), s0);
        s0own = fma(pi64[a], double(u0[a]), s0own);
      }
    const T inv_s0 = T(1.0 / s0);
    const T inv_own = T(1.0 / s0own);
    SB_GLOBAL T* bcnt = (flags & SB_F_BCOUNTS) ? glb<T>(B->bcounts) : nullptr;
    SB_GLOBAL T* npost = (flags & SB_F_NPOST) ? glb<T>(B->npost) : nullptr;
    SB_GLOBAL T* bpost = (flags & SB_F_BPOST) ? glb<T>(B->bpost) : nullptr;
    if (valid) {
      if ((flags & SB_F_LL) && B->loglike)
        glb<double>(B->loglike)[cc] = double(e0) * 0.69314718055994530942 + log(s0);
      if (npost) {
#pragma unroll
        for (int a = 0; a < AP; ++a)
          if (a < A) out_store(npost + (long long)a * C + cc, pi[a] * u0[a] * inv_own);
      }
      for (int ij = 0; ij < A * A; ++ij) {     // branch 0 (above the root) is zero
        if (bcnt) out_store(bcnt + (long long)ij * C + cc, T(0));
        if (bpost) out_store(bpost + (long long)ij * C + cc, T(0));
      }
    }
    if (!(bcnt || npost || bpost)) continue;

    // ---------------- outside pass (preorder) ----------------
    const int mask_mode = (flags >> SB_F_MASK_SHIFT) & 3;
    const SB_GLOBAL uint8_t* emask = B->mask ? glb<uint8_t>(B->mask) + cc : nullptr;

    // Branch above child kn (sibling ks) of a parent with outside product x:
    // its outside vector d (stored for an internal child), then the outputs.
    auto branch = [&](const Kid<T, AP>& kn, const T* wn, const Kid<T, AP>& ks, const T* ws, int sn,
                      const T* x, int ex, bool stp, T* d, int& word) SB_INLINE {
#pragma unroll
      for (int a = 0; a < AP; ++a) d[a] = x[a] * ws[a];
      const int ed = ks.e + ex + rescale<T, AP>(d);
      const bool st = (kn.leaf ? kn.has : kn.branch) || (stp && kn.has);
      word = (ed << 1) | (st ? 1 : 0);
      if (sn >= 0) {
        vstore(Dreg, sn, d);
        Dw[sn * SB_WAVE + lane] = word;
      }
      if (!valid) return;
      const int n = kn.n;
      const SB_CONST T* Mn = Mat(n);
      if (npost || bpost) {
        T z = T(0);
#pragma unroll
        for (int a = 0; a < AP; ++a) z = fma(d[a], wn[a], z);
        const T iz = T(1) / z;
        if (npost) {
          T y[AP];
          vecmat<T, AP>(d, Mn, y);
#pragma unroll
          for (int j = 0; j < AP; ++j)
            if (j < A) out_store(npost + ((long long)n * A + j) * C + cc, y[j] * kn.u[j] * iz);
        }
        if (bpost) {
#pragma unroll
          for (int a = 0; a < AP; ++a) {
            if (a >= A) continue;
            const T da = d[a] * iz;
#pragma unroll
            for (int j = 0; j < AP; ++j)
              if (j < A) out_store(bpost + (((long long)n * A + a) * A + j) * C + cc, da * Mn[a * AP + j] * kn.u[j]);
          }
        }
      }
      if (bcnt) {
        bool act;
        if (mask_mode == SB_MASK_NONE) act = true;
        else if (mask_mode == SB_MASK_AUTO) act = st && stp;
        else act = emask[(long long)n * C] != 0;
        const T sc = act ? ldexp_(inv_s0, ed + kn.e - e0) : T(0);
        SB_GLOBAL T* out = bcnt + (long long)n * A * A * C + cc;
        if constexpr (F81) {
          const SB_CONST T* cf = bt + (long long)n * BTS + sb_bt_coef(AP);
          const T al = cf[0], be = cf[1], ga = cf[2];
          const T fmu = md[sb_mr_misc(AP)];
          T Ds[AP], piU = T(0), Dsum = T(0);
#pragma unroll
          for (int a = 0; a < AP; ++a) {
            Ds[a] = d[a] * sc;
            piU = fma(pi[a], kn.u[a], piU);
            Dsum += Ds[a];
          }
#pragma unroll
          for (int a = 0; a < AP; ++a) {
            if (a >= A) continue;
#pragma unroll
            for (int j = 0; j < AP; ++j) {
              if (j >= A) continue;
              const T I = al * Ds[a] * kn.u[j] + be * (Ds[a] * piU + pi[a] * Dsum * kn.u[j]) +
                          ga * pi[a] * Dsum * piU;
              out_store(out + (long long)(a * A + j) * C, a == j ? I : fmu * pi[j] * I);
            }
          }
        } else if constexpr (IRR) {
          // complex eigenbasis of an irreversible model, per branch
          // (accumulate_C_complex_per_branch / back_transform_irrev_per_branch,
          // subby/jax/eigensub.py:175-231, 388-434): counts convention set 0;
          // per-branch models: branch n's own
          const SB_CONST T* mI = B->row_models
              ? cst<T>((const void*)(cst<unsigned long long>(B->row_models)[n])) : md;
          const SB_CONST T* Pu = mI + sb_mri_set(AP, 0) + sb_mri_Pu(AP);
          const SB_CONST T* Pd = mI + sb_mri_set(AP, 0) + sb_mri_Pd(AP);
          const SB_CONST T* Lr = mI + sb_mri_set(AP, 0) + sb_mri_L(AP);
          const SB_CONST T* Rr = mI + sb_mri_set(AP, 0) + sb_mri_Rt(AP);
          const SB_CONST T* Wr = mI + sb_mri_W(AP);
          const SB_CONST T* Jr = bt + (long long)n * BTS + sb_bti_Jre(AP);
          const SB_CONST T* Ji = bt + (long long)n * BTS + sb_bti_Jim(AP);
          constexpr int Q2 = AP * AP;
          T Ur[AP], Ui[AP], Dr[AP], Di[AP];
#pragma unroll
          for (int k = 0; k < AP; ++k) {
            T ur = T(0), ui = T(0), dr = T(0), di = T(0);
#pragma unroll
            for (int b = 0; b < AP; ++b) {
              ur = fma(kn.u[b], Pu[b * AP + k], ur);
              ui = fma(kn.u[b], Pu[Q2 + b * AP + k], ui);
              dr = fma(d[b], Pd[b * AP + k], dr);
              di = fma(d[b], Pd[Q2 + b * AP + k], di);
            }
            Ur[k] = ur;
            Ui[k] = ui;
            Dr[k] = dr * sc;
            Di[k] = di * sc;
          }
          T Cr[AP * AP], Ci[AP * AP];
#pragma unroll
          for (int k = 0; k < AP; ++k)
#pragma unroll
            for (int q = 0; q < AP; ++q) {
              const T jr = Jr[k * AP + q], ji = Ji[k * AP + q];
              const T xr = fma(Dr[k], jr, -Di[k] * ji), xi = fma(Dr[k], ji, Di[k] * jr);
              Cr[k * AP + q] = fma(xr, Ur[q], -xi * Ui[q]);
              Ci[k * AP + q] = fma(xr, Ui[q], xi * Ur[q]);
            }
#pragma unroll
          for (int a = 0; a < AP; ++a) {
            if (a >= A) continue;
            T tr[AP], ti[AP];
#pragma unroll
            for (int q = 0; q < AP; ++q) {
              T sr = T(0), si = T(0);
#pragma unroll
              for (int k = 0; k < AP; ++k) {
                const T lr = Lr[a * AP + k], li = Lr[Q2 + a * AP + k];
                sr = fma(lr, Cr[k * AP + q], fma(-li, Ci[k * AP + q], sr));
                si = fma(lr, Ci[k * AP + q], fma(li, Cr[k * AP + q], si));
              }
              tr[q] = sr;
              ti[q] = si;
            }
#pragma unroll
            for (int j = 0; j < AP; ++j) {
              if (j >= A) continue;
              T vr = T(0), vi = T(0);
#pragma unroll
              for (int q = 0; q < AP; ++q) {
                const T rr = Rr[q * AP + j], ri = Rr[Q2 + q * AP + j];
                vr = fma(tr[q], rr, fma(-ti[q], ri, vr));
                vi = fma(tr[q], ri, fma(ti[q], rr, vi));
              }
              const T wv = Wr[a * AP + j], wi = Wr[Q2 + a * AP + j];
              out_store(out + (long long)(a * A + j) * C, a == j ? vr : T(fma(wv, vr, -wi * vi)));
            }
          }
        } else {
          // per-branch models: branch n in the eigenbasis of its own model
          // (_branch_counts_per_row, subby/jax/__init__.py:102-144)
          const SB_CONST T* mn = B->row_models
              ? cst<T>((const void*)(cst<unsigned long long>(B->row_models)[n])) : md;
          const SB_CONST T* Vs = mn + sb_mr_Vs(AP);
          const SB_CONST T* Vi = mn + sb_mr_Vi(AP);
          const SB_CONST T* V = mn + sb_mr_V(AP);
          const SB_CONST T* S = mn + sb_mr_S(AP);
          const SB_CONST T* J = bt + (long long)n * BTS + sb_bt_J(AP);
          T Ut[AP], Dt[AP];
          vecmat<T, AP>(kn.u, Vs, Ut);   // Ut_q = sum_a Vs[a][q] u_a
          vecmat<T, AP>(d, Vi, Dt);      // Dt_k = sum_a Vi[a][k] d_a
          T Ce[AP * AP];
#pragma unroll
          for (int k = 0; k < AP; ++k)
#pragma unroll
            for (int q = 0; q < AP; ++q) Ce[k * AP + q] = Dt[k] * sc * J[k * AP + q] * Ut[q];
#pragma unroll
          for (int a = 0; a < AP; ++a) {
            if (a >= A) continue;
            T tmp[AP];
#pragma unroll
            for (int q = 0; q < AP; ++q) {
              T v = T(0);
#pragma unroll
              for (int k = 0; k < AP; ++k) v = fma(V[a * AP + k], Ce[k * AP + q], v);
              tmp[q] = v;
            }
#pragma unroll
            for (int j = 0; j < AP; ++j) {
              if (j >= A) continue;
              T v = T(0);
#pragma unroll
              for (int q = 0; q < AP; ++q) v = fma(tmp[q], V[j * AP + q], v);
              out_store(out + (long long)(a * A + j) * C, a == j ? v : S[a * AP + j] * v);
            }
          }
        }
      }
    };

    // loads of record i: parent token, its outside vector (slot i, i > 0), both children
    T xd[AP], au[AP], bu[AP];
    int xdw = 0, xt = 0, aw = 0, bw = 0, at = 0, btk = 0;
    auto ofetch = [&](int i, T* dp, int& dw, int& tp, T* u0_, int& w0, int& t0, T* u1_, int& w1,
                      int& t1) SB_INLINE {
      const SB_CONST int* rq = sched + i * 8;
      tp = aln[(long long)rq[0] * C];
      if (i > 0) {
        vload(Dreg, i, dp);
        dw = Dw[i * SB_WAVE + lane];
      }
      fetch(rq[1], rq[4], u0_, w0, t0);
      fetch(rq[2], rq[5], u1_, w1, t1);
    };
    if (nint > 0) ofetch(0, xd, xdw, xt, au, aw, at, bu, bw, btk);
    for (int i = 0; i < nint; ++i) {
      const SB_CONST int* rec = sched + i * 8;
      const int p = rec[0];
      T nxd[AP], nau[AP], nbu[AP];
      int nxdw = 0, nxt = 0, naw = 0, nbw = 0, nat = 0, nbt = 0;
      if (i + 1 < nint) ofetch(i + 1, nxd, nxdw, nxt, nau, naw, nat, nbu, nbw, nbt);
      T x[AP];
      int ex;
      bool stp;
      {
        T o[AP];
        token_vec(xt, o);
        if (i == 0) {
#pragma unroll
          for (int a = 0; a < AP; ++a) x[a] = o[a] * pr[a];
          ex = 0;
          stp = rbr;
        } else {
          ex = xdw >> 1;
          stp = (xdw & 1) != 0;
          vecmat<T, AP>(xd, Mat(p), x);
#pragma unroll
          for (int a = 0; a < AP; ++a) x[a] *= o[a];
        }
      }
      Kid<T, AP> k0, k1;
      decode(k0, rec[1], rec[4], au, aw, at);
      decode(k1, rec[2], rec[5], bu, bw, btk);
      T w0[AP], w1[AP];
      matvec<T, AP>(Mat(k0.n), k0.u, w0);
      matvec<T, AP>(Mat(k1.n), k1.u, w1);
      T d0[AP], d1[AP];
      int word0, word1;
      // a child whose record is the next one gets its outside vector forwarded
      // below, so it is not stored
      branch(k0, w0, k1, w1, rec[4] == i + 1 ? -1 : rec[4], x, ex, stp, d0, word0);
      branch(k1, w1, k0, w0, rec[5] == i + 1 ? -1 : rec[5], x, ex, stp, d1, word1);
      // the next record's load of slot i + 1 went out before this record stored it
      if (rec[4] == i + 1) {
#pragma unroll
        for (int a = 0; a < AP; ++a) nxd[a] = d0[a];
        nxdw = word0;
      }
      if (rec[5] == i + 1) {
#pragma unroll
        for (int a = 0; a < AP; ++a) nxd[a] = d1[a];
        nxdw = word1;
      }
#pragma unroll
      for (int a = 0; a < AP; ++a) {
        xd[a] = nxd[a];
        au[a] = nau[a];
        bu[a] = nbu[a];
      }
      xdw = nxdw;
      xt = nxt;
      aw = naw;
      bw = nbw;
      at = nat;
      btk = nbt;
    }
  }
}

#define SB_PINST(T, AP)                                                          \
  template __global__ void sb_post_kernel<T, AP, false, false>(const SbLaunch*); \
  template __global__ void sb_post_kernel<T, AP, true, false>(const SbLaunch*);  \
  template __global__ void sb_post_kernel<T, AP, false, true>(const SbLaunch*);
SB_PINST(float, 4)
SB_PINST(float, 8)
SB_PINST(double, 4)
SB_PINST(double, 8)

static void (*pick_post(int dtype, int AP, bool f81, bool irr))(const SbLaunch*) {
  if (f81 && irr) return nullptr;
#define SB_PP(T, APV)                                                                  \
  if (AP == APV) return irr ? sb_post_kernel<T, APV, false, true>                      \
                            : f81 ? sb_post_kernel<T, APV, true, false> : sb_post_kernel<T, APV, false, false>;
  if (dtype == 0) { SB_PP(float, 4) SB_PP(float, 8) }
  else { SB_PP(double, 4) SB_PP(double, 8) }
#undef SB_PP
  return nullptr;
}

extern "C" int sb_post_occupancy(int dtype, int AP, bool f81, bool irr) {
  void (*k)(const SbLaunch*) = pick_post(dtype, AP, f81, irr);
  int n = 0;
  if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, 64, 0) != hipSuccess) return 0;
  return n;
}

extern "C" hipError_t sb_launch_post(int dtype, int AP, bool f81, bool irr, int nwaves,
                                     const SbLaunch* L, hipStream_t s) {
  void (*k)(const SbLaunch*) = pick_post(dtype, AP, f81, irr);
  if (!k) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3(nwaves), dim3(64), 0, s, L);
  return hipGetLastError();
}
