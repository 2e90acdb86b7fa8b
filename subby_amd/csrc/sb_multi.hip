// Column regrouping and small epilogues for calls with several models, gfx950.
//
// A call whose columns use different models (the reference's per-column
// model lists and (C, R) model grids, subby/jax/__init__.py:52-204) runs one
// block per distinct model (or per-branch model row): the columns of each
// group are gathered into a contiguous [R][C_g] staging block, the persistent
// kernels run all groups in one launch, and the outputs are scattered back to
// the caller's column positions.  Columns are visited in "permuted" order:
// the groups one after another, so group g owns permuted positions
// [gstart, gstart + gsize) and its staging block of `rows` output rows starts
// at rows * gstart.
//
//   perm[j]   original column of permuted position j
//   grp[j]    group of permuted position j
//   gstart[g], gsize[g]
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sb_plan.h"

namespace {

struct Cols {
  const int* perm;
  const int* grp;
  const int* gstart;
  const int* gsize;
};

// dst (staging, per group [rows][C_g]) <- src [rows][C] at the permuted columns
template <typename T>
__global__ void __launch_bounds__(256) sb_gather_cols_kernel(const T* __restrict__ src, T* __restrict__ dst,
                                                             int rows, int C, Cols cm) {
  const long long total = (long long)rows * C;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int r = int(i / C), j = int(i - (long long)r * C);
    const int g = cm.grp[j], s = cm.gstart[g], n = cm.gsize[g];
    dst[(long long)rows * s + (long long)r * n + (j - s)] = src[(long long)r * C + cm.perm[j]];
  }
}

// dst [rows][C] at the permuted columns <- src (staging, per group [rows][C_g])
template <typename T>
__global__ void __launch_bounds__(256) sb_scatter_cols_kernel(const T* __restrict__ src, T* __restrict__ dst,
                                                              int rows, int C, Cols cm) {
  const long long total = (long long)rows * C;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int r = int(i / C), j = int(i - (long long)r * C);
    const int g = cm.grp[j], s = cm.gstart[g], n = cm.gsize[g];
    dst[(long long)r * C + cm.perm[j]] = src[(long long)rows * s + (long long)r * n + (j - s)];
  }
}

// Root-distribution swap in staging: a column computed with its own model's
// pi_g at the root gets the log-likelihood and root posterior of the same
// column with the root distribution pi_0 (the stacked-model convention of the
// reference's model lists, subby/jax/__init__.py:264-274, 617-632):
//   s = sum_a w_a q_a with w = pi_0 / pi_g and q the own-model root posterior,
//   ll += log s,  q /= s   (so q = pi_g U0 / sum_a pi_0a U0a).
// w: [ngroups][A] FP64; groups with swap[g] == 0 are left alone.
template <typename T>
__global__ void __launch_bounds__(256) sb_prior_swap_kernel(double* __restrict__ ll, T* __restrict__ root,
                                                            int A, int C, Cols cm, const double* __restrict__ w,
                                                            const int* __restrict__ swap) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < C; j += gridDim.x * 256) {
    const int g = cm.grp[j];
    if (!swap[g]) continue;
    const int s0 = cm.gstart[g], n = cm.gsize[g];
    T* q = root + (long long)A * s0 + (j - s0);
    double s = 0.0;
    for (int a = 0; a < A; ++a) s += w[(long long)g * A + a] * double(q[(long long)a * n]);
    if (ll) ll[j] += log(s);           // ll staging is [C_g] per group: offset s0 + (j - s0) = j
    const double inv = 1.0 / s;
    for (int a = 0; a < A; ++a) q[(long long)a * n] = T(double(q[(long long)a * n]) * inv);
  }
}

// Mixture posterior: post[k][c] = softmax_k(ll[k][c] + logw[k])
// (subby/jax/mixture.py:6-30).  One thread per column.
struct MixW {
  double logw[SB_MAX_MIX];
};

__global__ void __launch_bounds__(256) sb_mixture_softmax_kernel(const double* __restrict__ ll,
                                                                 double* __restrict__ post, int K, int C,
                                                                 MixW w) {
  for (int c = blockIdx.x * 256 + threadIdx.x; c < C; c += gridDim.x * 256) {
    double mx = -INFINITY;
    for (int k = 0; k < K; ++k) mx = fmax(mx, ll[(long long)k * C + c] + w.logw[k]);
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += exp(ll[(long long)k * C + c] + w.logw[k] - mx);
    const double inv = 1.0 / s;
    for (int k = 0; k < K; ++k) post[(long long)k * C + c] = exp(ll[(long long)k * C + c] + w.logw[k] - mx) * inv;
  }
}

int grid_for(long long n) {
  long long g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return int(g);
}

Cols make_cols(const int* const* p) { return Cols{p[0], p[1], p[2], p[3]}; }

}  // namespace

// cols: device pointers {perm, grp, gstart, gsize}.  esz: element size (1, 4 or 8).
extern "C" hipError_t sb_launch_gather_cols(const void* src, void* dst, int esz, int rows, int C,
                                            const int* const* cols, hipStream_t s) {
  const Cols cm = make_cols(cols);
  const int g = grid_for((long long)rows * C);
  if (esz == 1)
    hipLaunchKernelGGL(sb_gather_cols_kernel<uint8_t>, dim3(g), dim3(256), 0, s,
                       static_cast<const uint8_t*>(src), static_cast<uint8_t*>(dst), rows, C, cm);
  else if (esz == 4)
    hipLaunchKernelGGL(sb_gather_cols_kernel<float>, dim3(g), dim3(256), 0, s,
                       static_cast<const float*>(src), static_cast<float*>(dst), rows, C, cm);
  else if (esz == 8)
    hipLaunchKernelGGL(sb_gather_cols_kernel<double>, dim3(g), dim3(256), 0, s,
                       static_cast<const double*>(src), static_cast<double*>(dst), rows, C, cm);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

extern "C" hipError_t sb_launch_scatter_cols(const void* src, void* dst, int esz, int rows, int C,
                                             const int* const* cols, hipStream_t s) {
  const Cols cm = make_cols(cols);
  const int g = grid_for((long long)rows * C);
  if (esz == 4)
    hipLaunchKernelGGL(sb_scatter_cols_kernel<float>, dim3(g), dim3(256), 0, s,
                       static_cast<const float*>(src), static_cast<float*>(dst), rows, C, cm);
  else if (esz == 8)
    hipLaunchKernelGGL(sb_scatter_cols_kernel<double>, dim3(g), dim3(256), 0, s,
                       static_cast<const double*>(src), static_cast<double*>(dst), rows, C, cm);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

extern "C" hipError_t sb_launch_prior_swap(double* ll, void* root, int dtype, int A, int C,
                                           const int* const* cols, const double* w, const int* swap,
                                           hipStream_t s) {
  const Cols cm = make_cols(cols);
  const int g = grid_for(C);
  if (dtype == 0)
    hipLaunchKernelGGL(sb_prior_swap_kernel<float>, dim3(g), dim3(256), 0, s, ll,
                       static_cast<float*>(root), A, C, cm, w, swap);
  else
    hipLaunchKernelGGL(sb_prior_swap_kernel<double>, dim3(g), dim3(256), 0, s, ll,
                       static_cast<double*>(root), A, C, cm, w, swap);
  return hipGetLastError();
}

extern "C" hipError_t sb_launch_mixture_softmax(const double* ll, double* post, int K, int C,
                                                const double* logw, hipStream_t s) {
  if (K < 1 || K > SB_MAX_MIX) return hipErrorInvalidValue;
  MixW w;
  for (int k = 0; k < SB_MAX_MIX; ++k) w.logw[k] = k < K ? logw[k] : 0.0;
  hipLaunchKernelGGL(sb_mixture_softmax_kernel, dim3(grid_for(C)), dim3(256), 0, s, ll, post, K, C, w);
  return hipGetLastError();
}
