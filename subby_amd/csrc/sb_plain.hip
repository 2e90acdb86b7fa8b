// Plain one-block path (A = 4, FP32 partials, FP64 log-normalisers).
//
// The data flow of the reference's WebGPU engine (subby/webgpu/phylo_gpu.js:
// 611-724, shaders upward_step.wgsl / downward_step.wgsl / accumulate_C.wgsl /
// back_transform.wgsl), organised by tree level instead of by branch: every
// inside and outside partial U, D (R x C x 4 floats) and its log-normaliser
// (R x C doubles) is written to HBM once and read back, one launch per
// postorder level and per preorder level.  It is the one-block-per-launch
// baseline the fused persistent kernel is measured against, and a second,
// independent GPU restatement of the same statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/subby_amd.h"
#include "sb_plan.h"

namespace {

struct PlainNode {  // per node, uploaded once per call
  int parent, sibling, left, right;
};

__device__ inline void tokvec(int tok, float* u) {
  const bool obs = tok >= 0 && tok < 4;
#pragma unroll
  for (int a = 0; a < 4; ++a) u[a] = obs ? (a == tok ? 1.f : 0.f) : 1.f;
}

__device__ inline void mv(const float* __restrict__ M, const float* u, float* w) {
#pragma unroll
  for (int a = 0; a < 4; ++a) w[a] = M[a * 4] * u[0] + M[a * 4 + 1] * u[1] + M[a * 4 + 2] * u[2] + M[a * 4 + 3] * u[3];
}

__global__ void plain_init(const int8_t* __restrict__ aln, int R, int C, float4* __restrict__ U,
                           double* __restrict__ lnU) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)R * C) return;
  float u[4];
  tokvec(aln[i], u);
  U[i] = make_float4(u[0], u[1], u[2], u[3]);
  lnU[i] = 0.0;
}

// One postorder level: U[p] = obs(p) * (M_l U_l) * (M_r U_r), max-rescaled.
__global__ void plain_up(const int* __restrict__ nodes, int count, const PlainNode* __restrict__ tn,
                         const float* __restrict__ bt, int bts, int C, float4* __restrict__ U,
                         double* __restrict__ lnU) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (c >= C || k >= count) return;
  const int p = nodes[k];
  const int l = tn[p].left, r = tn[p].right;
  const long long ip = (long long)p * C + c, il = (long long)l * C + c, ir = (long long)r * C + c;
  float4 up = U[ip], ul = U[il], ur = U[ir];
  float a[4] = {ul.x, ul.y, ul.z, ul.w}, b[4] = {ur.x, ur.y, ur.z, ur.w}, wl[4], wr[4];
  mv(bt + (long long)l * bts, a, wl);
  mv(bt + (long long)r * bts, b, wr);
  float v[4] = {up.x * wl[0] * wr[0], up.y * wl[1] * wr[1], up.z * wl[2] * wr[2], up.w * wl[3] * wr[3]};
  float m = fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3]));
  m = fmaxf(m, 1e-30f);
  const float im = 1.f / m;
  U[ip] = make_float4(v[0] * im, v[1] * im, v[2] * im, v[3] * im);
  lnU[ip] = lnU[il] + lnU[ir] + log((double)m);
}

__global__ void plain_root(const float* __restrict__ pi, int C, const float4* __restrict__ U,
                           const double* __restrict__ lnU, double* __restrict__ ll,
                           float* __restrict__ root) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const float4 u = U[c];
  const double s = (double)pi[0] * u.x + (double)pi[1] * u.y + (double)pi[2] * u.z + (double)pi[3] * u.w;
  const double l = lnU[c] + log(s);
  ll[c] = l;
  if (root) {
    const float is = float(1.0 / s);
    root[c] = pi[0] * u.x * is;
    root[(long long)C + c] = pi[1] * u.y * is;
    root[2LL * C + c] = pi[2] * u.z * is;
    root[3LL * C + c] = pi[3] * u.w * is;
  }
}

// One preorder level: D[n] = (M_s U_s) * obs(p) * (pi or D[p] M_p), max-rescaled.
__global__ void plain_down(const int* __restrict__ nodes, int count, const PlainNode* __restrict__ tn,
                           const float* __restrict__ bt, int bts, const float* __restrict__ pi,
                           const int8_t* __restrict__ aln, int C, const float4* __restrict__ U,
                           const double* __restrict__ lnU, float4* __restrict__ D,
                           double* __restrict__ lnD) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (c >= C || k >= count) return;
  const int n = nodes[k];
  const int p = tn[n].parent, s = tn[n].sibling;
  const long long is = (long long)s * C + c;
  const float4 us = U[is];
  float a[4] = {us.x, us.y, us.z, us.w}, ws[4], x[4], o[4];
  mv(bt + (long long)s * bts, a, ws);
  double prior = 0.0;
  if (p == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = pi[q];
  } else {
    const long long ipp = (long long)p * C + c;
    const float4 dp = D[ipp];
    const float d[4] = {dp.x, dp.y, dp.z, dp.w};
    const float* M = bt + (long long)p * bts;
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = d[0] * M[q] + d[1] * M[4 + q] + d[2] * M[8 + q] + d[3] * M[12 + q];
    prior = lnD[ipp];
  }
  tokvec(aln[(long long)p * C + c], o);
  float v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = ws[q] * x[q] * o[q];
  float m = fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3]));
  m = fmaxf(m, 1e-30f);
  const float im = 1.f / m;
  const long long in = (long long)n * C + c;
  D[in] = make_float4(v[0] * im, v[1] * im, v[2] * im, v[3] * im);
  lnD[in] = lnU[is] + prior + log((double)m);
}

// Steiner mask per column (oracle.py:1372-1435) into an (R, C) byte array.
__global__ void plain_mask(const int8_t* __restrict__ aln, const PlainNode* __restrict__ tn, int R, int C,
                           uint8_t* __restrict__ has, uint8_t* __restrict__ st, uint8_t* __restrict__ mask) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  for (int n = R - 1; n >= 0; --n) {
    const long long i = (long long)n * C + c;
    if (tn[n].left < 0) {
      const int t = aln[i];
      const uint8_t h = (t >= 0 && t <= 4) ? 1 : 0;
      has[i] = h;
      st[i] = h;
    } else {
      const uint8_t hl = has[(long long)tn[n].left * C + c], hr = has[(long long)tn[n].right * C + c];
      has[i] = hl | hr;
      st[i] = hl & hr;
    }
  }
  mask[c] = 0;
  for (int n = 1; n < R; ++n) {
    const long long i = (long long)n * C + c;
    const uint8_t sp = st[(long long)tn[n].parent * C + c];
    const uint8_t sn = st[i] | (sp & has[i]);
    st[i] = sn;
    mask[i] = sn & sp;
  }
}

// accumulate_C + back_transform (eigensub.py:87-172) per column.
__global__ void plain_accum(int R, int C, const float* __restrict__ bt, int bts, const float* __restrict__ md,
                            const float4* __restrict__ U, const double* __restrict__ lnU,
                            const float4* __restrict__ D, const double* __restrict__ lnD,
                            const double* __restrict__ ll, const uint8_t* __restrict__ mask,
                            float* __restrict__ counts) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const float* Vs = md + sb_mr_Vs(4);
  const float* Vi = md + sb_mr_Vi(4);
  float acc[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  const double l0 = ll[c];
  for (int n = 1; n < R; ++n) {
    const long long i = (long long)n * C + c;
    if (mask && !mask[i]) continue;
    const float sc = float(exp(lnD[i] + lnU[i] - l0));
    const float4 uu = U[i], dd = D[i];
    const float u[4] = {uu.x, uu.y, uu.z, uu.w}, d[4] = {dd.x, dd.y, dd.z, dd.w};
    float ut[4], dt[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ut[k] = u[0] * Vs[k] + u[1] * Vs[4 + k] + u[2] * Vs[8 + k] + u[3] * Vs[12 + k];
      dt[k] = (d[0] * Vi[k] + d[1] * Vi[4 + k] + d[2] * Vi[8 + k] + d[3] * Vi[12 + k]) * sc;
    }
    const float* J = bt + (long long)n * bts + sb_bt_J(4);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[k * 4 + q] += dt[k] * J[k * 4 + q] * ut[q];
  }
  const float* V = md + sb_mr_V(4);
  const float* S = md + sb_mr_S(4);
  float tmp[16];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      tmp[i * 4 + q] = V[i * 4] * acc[q] + V[i * 4 + 1] * acc[4 + q] + V[i * 4 + 2] * acc[8 + q] + V[i * 4 + 3] * acc[12 + q];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float s = tmp[i * 4] * V[j * 4] + tmp[i * 4 + 1] * V[j * 4 + 1] + tmp[i * 4 + 2] * V[j * 4 + 2] +
                      tmp[i * 4 + 3] * V[j * 4 + 3];
      counts[(long long)(i * 4 + j) * C + c] = (i == j) ? s : S[i * 4 + j] * s;
    }
}

}  // namespace

struct SbPlainBuffers {
  void* p = nullptr;
  size_t cap = 0;
};

extern "C" hipError_t sb_plain_run(const int8_t* aln, const int32_t* parent, int R, int C,
                                   const float* d_bt, const float* d_model, int mask_mode,
                                   float* counts, float* root, double* ll, SbPlainBuffers* buf,
                                   hipStream_t s) {
  // levels
  std::vector<PlainNode> tn(R);
  for (int n = 0; n < R; ++n) tn[n] = {n == 0 ? -1 : parent[n], -1, -1, -1};
  for (int n = 1; n < R; ++n) {
    PlainNode& P = tn[parent[n]];
    if (P.left < 0) P.left = n;
    else P.right = n;
  }
  for (int n = 1; n < R; ++n) {
    const PlainNode& P = tn[parent[n]];
    tn[n].sibling = P.left == n ? P.right : P.left;
  }
  std::vector<int> height(R, 0), depth(R, 0);
  for (int n = R - 1; n >= 1; --n) height[parent[n]] = std::max(height[parent[n]], height[n] + 1);
  for (int n = 1; n < R; ++n) depth[n] = depth[parent[n]] + 1;
  const int H = height[0];
  int Dm = 0;
  for (int n = 0; n < R; ++n) Dm = std::max(Dm, depth[n]);
  std::vector<int> up_nodes, up_off(H + 2, 0), dn_nodes, dn_off(Dm + 2, 0);
  for (int h = 1; h <= H; ++h) {
    up_off[h] = int(up_nodes.size());
    for (int n = 0; n < R; ++n)
      if (tn[n].left >= 0 && height[n] == h) up_nodes.push_back(n);
  }
  up_off[H + 1] = int(up_nodes.size());
  for (int d = 1; d <= Dm; ++d) {
    dn_off[d] = int(dn_nodes.size());
    for (int n = 1; n < R; ++n)
      if (depth[n] == d) dn_nodes.push_back(n);
  }
  dn_off[Dm + 1] = int(dn_nodes.size());

  const size_t RC = size_t(R) * C;
  const size_t meta = sizeof(PlainNode) * R + sizeof(int) * (up_nodes.size() + dn_nodes.size() + 2);
  const size_t need = 2 * RC * 16 + 2 * RC * 8 + 3 * RC + (meta + 255) / 256 * 256 + 256;
  if (need > buf->cap) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (buf->p) (void)hipFree(buf->p);
    buf->p = nullptr;
    buf->cap = 0;
    e = hipMalloc(&buf->p, need);
    if (e != hipSuccess) return e;
    buf->cap = need;
  }
  char* base = static_cast<char*>(buf->p);
  float4* U = reinterpret_cast<float4*>(base);
  float4* D = U + RC;
  double* lnU = reinterpret_cast<double*>(D + RC);
  double* lnD = lnU + RC;
  char* meta_p = reinterpret_cast<char*>(lnD + RC);
  PlainNode* d_tn = reinterpret_cast<PlainNode*>(meta_p);
  int* d_up = reinterpret_cast<int*>(d_tn + R);
  int* d_dn = d_up + up_nodes.size() + 1;
  uint8_t* has = reinterpret_cast<uint8_t*>(meta_p + (meta + 255) / 256 * 256);
  uint8_t* st = has + RC;
  uint8_t* mask = st + RC;

  std::vector<char> hm(meta);
  std::memcpy(hm.data(), tn.data(), sizeof(PlainNode) * R);
  std::memcpy(hm.data() + sizeof(PlainNode) * R, up_nodes.data(), sizeof(int) * up_nodes.size());
  std::memcpy(hm.data() + sizeof(PlainNode) * R + sizeof(int) * (up_nodes.size() + 1), dn_nodes.data(),
              sizeof(int) * dn_nodes.size());
  hipError_t e = hipMemcpyAsync(meta_p, hm.data(), meta, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;

  const int bts = sb_bt_stride(4);
  const float* pi = d_model + sb_mr_pi(4);
  const int TB = 256;
  hipLaunchKernelGGL(plain_init, dim3(unsigned((RC + TB - 1) / TB)), dim3(TB), 0, s, aln, R, C, U, lnU);
  for (int h = 1; h <= H; ++h) {
    const int cnt = up_off[h + 1] - up_off[h];
    hipLaunchKernelGGL(plain_up, dim3((C + TB - 1) / TB, cnt), dim3(TB), 0, s, d_up + up_off[h], cnt, d_tn,
                       d_bt, bts, C, U, lnU);
  }
  hipLaunchKernelGGL(plain_root, dim3((C + TB - 1) / TB), dim3(TB), 0, s, pi, C, U, lnU, ll, root);
  if (counts) {
    for (int d = 1; d <= Dm; ++d) {
      const int cnt = dn_off[d + 1] - dn_off[d];
      hipLaunchKernelGGL(plain_down, dim3((C + TB - 1) / TB, cnt), dim3(TB), 0, s, d_dn + dn_off[d], cnt, d_tn,
                         d_bt, bts, pi, aln, C, U, lnU, D, lnD);
    }
    const uint8_t* mk = nullptr;
    if (mask_mode == SB_MASK_AUTO) {
      hipLaunchKernelGGL(plain_mask, dim3((C + TB - 1) / TB), dim3(TB), 0, s, aln, d_tn, R, C, has, st, mask);
      mk = mask;
    }
    hipLaunchKernelGGL(plain_accum, dim3((C + TB - 1) / TB), dim3(TB), 0, s, R, C, d_bt, bts, d_model, U, lnU,
                       D, lnD, ll, mk, counts);
  }
  // hm (pageable) must outlive its copy; the plain path also returns only
  // when the block's results are complete, like the reference's readbacks.
  return hipStreamSynchronize(s);
}
