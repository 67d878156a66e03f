// Learned warm start: a 4-layer 3x3 convolutional network on BF16 MFMA
// (gfx950), FP32 accumulation, with hand-written forward and backward passes
// and an Adam step, so the network is trained on the GPU by these kernels.
//
// Network (channels 16 -> 32 -> 32 -> 32 -> 16; input channels 0, 1 are the
// features log|A^H y|^2 and log gamma, the rest zero padding; output channel 0
// is the per-cell log-variance correction):
//     h_l = relu(conv3x3(h_{l-1}, W_l) + b_l),  l = 1..3;   out = conv3x3(h_3, W_4) + b_4
// Layouts: activations [B, H, W, C] bf16 (NHWC); master weights fp32
// W[tap][ci][co] with tap = 3 kh + kw reading input cell (h + kh - 1,
// w + kw - 1); biases fp32.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16_bf16: a workgroup
// (4 wavefronts) computes 64 consecutive pixels of one image row for all
// output channels; the input patch (3 rows x 66 columns) and the weights are
// staged in LDS.  dX uses the same kernel with the taps reversed and the
// channel roles swapped; dW reduces over pixels (K = pixels) per workgroup
// into deterministic FP32 partials.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fbme_hip.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int BLK = 256;
constexpr int TILE_W = 64;   // pixels per workgroup along the row
constexpr int NL = 4;
constexpr int CH[NL + 1] = {16, 32, 32, 32, 16};

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;   // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);                          // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// EPI 0: + bias, relu, bf16 out.  EPI 1: + bias, fp32 channel 0 out.
// EPI 2: no bias, zeroed where mask <= 0 (relu derivative), bf16 out.
template <int CIN, int COUT, int EPI>
__global__ void __launch_bounds__(BLK)
k_conv3x3(const uint16_t* __restrict__ x, const uint16_t* __restrict__ wt, const float* __restrict__ bias,
          const uint16_t* __restrict__ mask, uint16_t* __restrict__ y, float* __restrict__ yf, int H, int W) {
  constexpr int CP = CIN + 4;      // LDS channel stride (keeps 8-byte alignment)
  constexpr int CQ = CIN / 4;      // 8-byte chunks per pixel
  __shared__ uint16_t patch[3][TILE_W + 2][CP];
  __shared__ uint16_t wl[9][COUT][CP];
  const int b = blockIdx.z, h = blockIdx.y, w0 = blockIdx.x * TILE_W;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int e = tid; e < 3 * (TILE_W + 2) * CQ; e += BLK) {
    const int r = e / ((TILE_W + 2) * CQ);
    const int rem = e - r * (TILE_W + 2) * CQ;
    const int c = rem / CQ, q = rem - c * CQ;
    const int hh = h + r - 1, ww = w0 + c - 1;
    uint2 v = make_uint2(0u, 0u);
    if (hh >= 0 && hh < H && ww >= 0 && ww < W)
      v = *(const uint2*)(x + (((size_t)b * H + hh) * W + ww) * CIN + 4 * q);
    *(uint2*)(&patch[r][c][4 * q]) = v;
  }
  for (int e = tid; e < 9 * COUT * CQ; e += BLK) {
    const int tap = e / (COUT * CQ);
    const int rem = e - tap * COUT * CQ;
    const int co = rem / CQ, q = rem - co * CQ;
    *(uint2*)(&wl[tap][co][4 * q]) = *(const uint2*)(wt + ((size_t)tap * COUT + co) * CIN + 4 * q);
  }
  __syncthreads();
  constexpr int NT = COUT / 16;
  f4 acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
  const int i = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const int kh = tap / 3, kw = tap - 3 * kh;
#pragma unroll
    for (int s = 0; s < CIN / 16; ++s) {
      const s4 a = *(const s4*)(&patch[kh][16 * wave + i + kw][16 * s + 4 * kq]);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const s4 bb = *(const s4*)(&wl[tap][16 * nt + i][16 * s + 4 * kq]);
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[nt], 0, 0, 0);
      }
    }
  }
  // lane holds D[4 kq + t][i]: pixel 16 wave + 4 kq + t, output channel 16 nt + i
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int px = w0 + 16 * wave + 4 * kq + t;
      if (px >= W) continue;
      const int co = 16 * nt + i;
      const size_t o = ((size_t)b * H + h) * W + px;
      float v = acc[nt][t];
      if (EPI == 0) {
        v = fmaxf(v + bias[co], 0.f);
        y[o * COUT + co] = f2bf(v);
      } else if (EPI == 1) {
        if (co == 0) yf[o] = v + bias[0];
      } else {
        if (!(bf2f(mask[o * COUT + co]) > 0.f)) v = 0.f;
        y[o * COUT + co] = f2bf(v);
      }
    }
  }
}

// Partial weight gradients of one workgroup: dW[tap][ci][co] = sum over its
// pixels p of x[p + tap][ci] dz[p][co], for `rows` image rows starting at
// row0 and TILE_W columns starting at w0.  Wave w handles taps w, w+4, w+8.
template <int CIN, int COUT>
__global__ void __launch_bounds__(BLK)
k_conv3x3_dw(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dz, float* __restrict__ part, int H,
             int W, int rows) {
  constexpr int XP = TILE_W + 2 + 6;   // column stride of the transposed patch
  constexpr int DP = TILE_W + 8;
  __shared__ uint16_t xt[3][CIN][XP];
  __shared__ uint16_t dt[COUT][DP];
  const int b = blockIdx.z, row0 = blockIdx.y * rows, w0 = blockIdx.x * TILE_W;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int i = lane & 15, kq = lane >> 4;
  constexpr int CT = CIN / 16, NT = COUT / 16;
  f4 acc[3][CT][NT];
#pragma unroll
  for (int u = 0; u < 3; ++u)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[u][ct][nt] = f4{0.f, 0.f, 0.f, 0.f};
  for (int r = 0; r < rows; ++r) {
    const int h = row0 + r;
    if (h >= H) break;
    __syncthreads();
    for (int e = tid; e < 3 * (TILE_W + 2) * CIN; e += BLK) {
      const int rr = e / ((TILE_W + 2) * CIN);
      const int rem = e - rr * (TILE_W + 2) * CIN;
      const int c = rem / CIN, ci = rem - c * CIN;
      const int hh = h + rr - 1, ww = w0 + c - 1;
      uint16_t v = 0;
      if (hh >= 0 && hh < H && ww >= 0 && ww < W) v = x[(((size_t)b * H + hh) * W + ww) * CIN + ci];
      xt[rr][ci][c] = v;
    }
    for (int e = tid; e < TILE_W * COUT; e += BLK) {
      const int c = e / COUT, co = e - c * COUT;
      const int ww = w0 + c;
      dt[co][c] = (ww < W) ? dz[(((size_t)b * H + h) * W + ww) * COUT + co] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int tap = wave + 4 * u;
      if (tap >= 9) break;
      const int kh = tap / 3, kw = tap - 3 * kh;
#pragma unroll
      for (int ks = 0; ks < TILE_W / 16; ++ks) {
        const int p0 = 16 * ks + 4 * kq;
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const uint16_t* xr = &xt[kh][16 * ct + i][p0 + kw];
          const s4 a = s4{(short)xr[0], (short)xr[1], (short)xr[2], (short)xr[3]};
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) {
            const s4 bb = *(const s4*)(&dt[16 * nt + i][p0]);
            acc[u][ct][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[u][ct][nt], 0, 0, 0);
          }
        }
      }
    }
  }
  const size_t blk = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  float* pb = part + blk * 9 * CIN * COUT;
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int tap = wave + 4 * u;
    if (tap >= 9) break;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int ci = 16 * ct + 4 * kq + t, co = 16 * nt + i;
          pb[((size_t)tap * CIN + ci) * COUT + co] = acc[u][ct][nt][t];
        }
  }
}

// out[e] = sum over nblk partials (fixed order), e < n.
__global__ void k_reduce_parts(const float* __restrict__ part, int nblk, int n, float* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int k = 0; k < nblk; ++k) s += (double)part[(size_t)k * n + e];
  out[e] = (float)s;
}

// Bias gradient partials: one workgroup per image row, part[row][co].
template <int COUT>
__global__ void __launch_bounds__(BLK)
k_bias_grad(const uint16_t* __restrict__ dz, int W, float* __restrict__ part) {
  __shared__ float red[BLK / COUT][COUT];
  const size_t row = blockIdx.x;
  const int co = threadIdx.x % COUT, g = threadIdx.x / COUT;
  float s = 0.f;
  for (int w = g; w < W; w += BLK / COUT) s += bf2f(dz[(row * W + w) * COUT + co]);
  red[g][co] = s;
  __syncthreads();
  if (threadIdx.x < COUT) {
    float t = 0.f;
    for (int k = 0; k < BLK / COUT; ++k) t += red[k][threadIdx.x];
    part[row * COUT + threadIdx.x] = t;
  }
}

// Forward weights wf[tap][co][ci] = W[tap][ci][co]; backward weights for
// dX, wb[tap][ci][co] = W[8 - tap][ci][co] (read as [tap][co'][ci'] with
// co' = ci, ci' = co by the convolution kernel).
__global__ void k_prep(const float* __restrict__ wm, int cin, int cout, uint16_t* __restrict__ wf,
                       uint16_t* __restrict__ wb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 9 * cin * cout) return;
  const int tap = e / (cin * cout), rem = e - tap * cin * cout;
  const int ci = rem / cout, co = rem - ci * cout;
  const uint16_t v = f2bf(wm[e]);
  wf[((size_t)tap * cout + co) * cin + ci] = v;
  wb[((size_t)(8 - tap) * cin + ci) * cout + co] = v;
}

// Input features (channels 0, 1; 2..15 zero).
__global__ void k_features(const float2* __restrict__ bp, const float* __restrict__ gamma,
                           const double* __restrict__ bpe, const double* __restrict__ gsum, int n,
                           uint16_t* __restrict__ x) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = bp[(size_t)b * n + i];
  const float me = (float)(bpe[b] / n), mg = (float)(gsum[b] / n);
  float f0 = logf((v.x * v.x + v.y * v.y) / fmaxf(me, 1e-30f) + 1e-6f);
  float f1 = logf(fmaxf(gamma[(size_t)b * n + i], 1e-37f) / fmaxf(mg, 1e-30f));
  f0 = fminf(fmaxf(f0, -30.f), 30.f);
  f1 = fminf(fmaxf(f1, -30.f), 30.f);
  uint16_t* o = x + ((size_t)b * n + i) * 16;
  o[0] = f2bf(f0);
  o[1] = f2bf(f1);
#pragma unroll
  for (int c = 2; c < 16; ++c) o[c] = 0;
}

// Inputs of the preconditioner network: channels 0 and 1 as k_features,
// channel 2 = log(1 + N gamma / noise) / 8, the cell's loading of Q
// relative to the noise floor (the quantity the grid-diagonal background
// w = 1 / (N (noise + N gamma)) is built from).
__global__ void k_features_precond(const float2* __restrict__ bp, const float* __restrict__ gamma,
                                   const double* __restrict__ bpe, const double* __restrict__ gsum,
                                   const float* __restrict__ noise, int n, uint16_t* __restrict__ x) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = bp[(size_t)b * n + i];
  const float g = gamma[(size_t)b * n + i];
  const float me = (float)(bpe[b] / n), mg = (float)(gsum[b] / n);
  float f0 = logf((v.x * v.x + v.y * v.y) / fmaxf(me, 1e-30f) + 1e-6f);
  float f1 = logf(fmaxf(g, 1e-37f) / fmaxf(mg, 1e-30f));
  float f2 = log1pf((float)n * fmaxf(g, 0.f) / fmaxf(noise[b], 1e-37f)) * 0.125f;
  f0 = fminf(fmaxf(f0, -30.f), 30.f);
  f1 = fminf(fmaxf(f1, -30.f), 30.f);
  f2 = fminf(f2, 30.f);
  uint16_t* o = x + ((size_t)b * n + i) * 16;
  o[0] = f2bf(f0);
  o[1] = f2bf(f1);
  o[2] = f2bf(f2);
#pragma unroll
  for (int c = 3; c < 16; ++c) o[c] = 0;
}

// dz4 channel 0 = g (a loss gradient at the output formed elsewhere), other
// channels 0.
__global__ void k_output_grad(const float* __restrict__ g, long long count, uint16_t* __restrict__ dz) {
  const long long e = (long long)blockIdx.x * BLK + threadIdx.x;
  if (e >= count) return;
  uint16_t* o = dz + e * 16;
  o[0] = f2bf(g[e]);
#pragma unroll
  for (int c = 1; c < 16; ++c) o[c] = 0;
}

// dz4 channel 0 = 2 (out - target) / count, other channels 0; FP64 loss partials.
__global__ void __launch_bounds__(BLK)
k_loss(const float* __restrict__ out, const float* __restrict__ target, long long count, uint16_t* __restrict__ dz,
       double* __restrict__ part) {
  __shared__ double red[BLK / 64];
  const long long e = (long long)blockIdx.x * BLK + threadIdx.x;
  double l = 0.0;
  if (e < count) {
    const float d = out[e] - target[e];
    l = (double)d * d;
    uint16_t* o = dz + e * 16;
    o[0] = f2bf(2.0f * d / (float)count);
#pragma unroll
    for (int c = 1; c < 16; ++c) o[c] = 0;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < BLK / 64; ++k) s += red[k];
    part[blockIdx.x] = s;
  }
}

__global__ void k_sum_parts(const double* __restrict__ part, int n, double scale, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += part[i];
  out[0] = s * scale;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// gamma <- max(gamma exp(clamp(out)), floor).
__global__ void k_warm(float* __restrict__ gamma, const float* __restrict__ out, const float* __restrict__ floor_,
                       int n, float clampv) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t k = (size_t)b * n + i;
  const float d = fminf(fmaxf(out[k], -clampv), clampv);
  gamma[k] = fmaxf(gamma[k] * expf(d), floor_[b]);
}

inline int nblk(long long n, int bs = BLK) { return (int)((n + bs - 1) / bs); }

// parameter layout: for l in 0..3: W_l [9][CH[l]][CH[l+1]] then b_l [CH[l+1]]
inline size_t w_off(int l) {
  size_t o = 0;
  for (int k = 0; k < l; ++k) o += (size_t)9 * CH[k] * CH[k + 1] + CH[k + 1];
  return o;
}
inline size_t b_off(int l) { return w_off(l) + (size_t)9 * CH[l] * CH[l + 1]; }
inline size_t wbf_off(int l) {
  size_t o = 0;
  for (int k = 0; k < l; ++k) o += (size_t)9 * CH[k] * CH[k + 1];
  return o;
}

}  // namespace

extern "C" {

long long fbme_cnn_param_count(void) { return (long long)w_off(NL); }
long long fbme_cnn_weight_count(void) { return (long long)wbf_off(NL); }

int fbme_cnn_prepare(void* stream, const float* params, void* wf, void* wb) {
  hipStream_t st = (hipStream_t)stream;
  for (int l = 0; l < NL; ++l) {
    const int n = 9 * CH[l] * CH[l + 1];
    hipLaunchKernelGGL(k_prep, dim3(nblk(n)), dim3(BLK), 0, st, params + w_off(l), CH[l], CH[l + 1],
                       (uint16_t*)wf + wbf_off(l), (uint16_t*)wb + wbf_off(l));
  }
  return (int)hipGetLastError();
}

int fbme_cnn_features(void* stream, int batch, int n, const void* bp, const float* gamma, const double* bpe,
                      const double* gsum, void* x) {
  if (batch < 1 || n < 1) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_features, dim3(nblk(n), batch), dim3(BLK), 0, (hipStream_t)stream, (const float2*)bp, gamma,
                     bpe, gsum, n, (uint16_t*)x);
  return (int)hipGetLastError();
}

int fbme_cnn_features_precond(void* stream, int batch, int n, const void* bp, const float* gamma,
                              const double* bpe, const double* gsum, const float* noise, void* x) {
  if (batch < 1 || n < 1 || !noise) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_features_precond, dim3(nblk(n), batch), dim3(BLK), 0, (hipStream_t)stream,
                     (const float2*)bp, gamma, bpe, gsum, noise, n, (uint16_t*)x);
  return (int)hipGetLastError();
}

int fbme_cnn_output_grad(void* stream, long long count, const float* g, void* dz4) {
  if (count < 1 || !g || !dz4) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_output_grad, dim3(nblk(count)), dim3(BLK), 0, (hipStream_t)stream, g, count, (uint16_t*)dz4);
  return (int)hipGetLastError();
}

// act: 3 hidden activations, each batch * h * w * 32 bf16; out [B, H, W] fp32.
int fbme_cnn_forward(void* stream, int batch, int h, int w, const void* x, const void* wf, const float* params,
                     void* act, float* out) {
  if (batch < 1 || h < 1 || w < 1) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const size_t plane = (size_t)batch * h * w * 32;
  uint16_t* a = (uint16_t*)act;
  const uint16_t* wt = (const uint16_t*)wf;
  dim3 grid((w + TILE_W - 1) / TILE_W, h, batch);
  hipLaunchKernelGGL((k_conv3x3<16, 32, 0>), grid, dim3(BLK), 0, st, (const uint16_t*)x, wt + wbf_off(0),
                     params + b_off(0), nullptr, a, nullptr, h, w);
  hipLaunchKernelGGL((k_conv3x3<32, 32, 0>), grid, dim3(BLK), 0, st, a, wt + wbf_off(1), params + b_off(1), nullptr,
                     a + plane, nullptr, h, w);
  hipLaunchKernelGGL((k_conv3x3<32, 32, 0>), grid, dim3(BLK), 0, st, a + plane, wt + wbf_off(2), params + b_off(2),
                     nullptr, a + 2 * plane, nullptr, h, w);
  hipLaunchKernelGGL((k_conv3x3<32, 16, 1>), grid, dim3(BLK), 0, st, a + 2 * plane, wt + wbf_off(3),
                     params + b_off(3), nullptr, nullptr, out, h, w);
  return (int)hipGetLastError();
}

// Workspace sizes for fbme_cnn_backward.
long long fbme_cnn_grad_rows(int h) { return h < 16 ? h : 16; }
long long fbme_cnn_part_floats(int batch, int h, int w) {
  const long long rows = fbme_cnn_grad_rows(h);
  const long long blocks = (long long)batch * ((h + rows - 1) / rows) * ((w + TILE_W - 1) / TILE_W);
  const long long wparts = blocks * 9 * 32 * 32;
  const long long bparts = (long long)batch * h * 32;
  return wparts > bparts ? wparts : bparts;
}

int fbme_cnn_loss(void* stream, long long count, const float* out, const float* target, void* dz4, double* part,
                  double* loss) {
  hipStream_t st = (hipStream_t)stream;
  const int nb = nblk(count);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(BLK), 0, st, out, target, count, (uint16_t*)dz4, part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, st, part, nb, 1.0 / (double)count, loss);
  return (int)hipGetLastError();
}

// grads: fp32 in the parameter layout.  dz: 2 buffers of batch*h*w*32 bf16;
// dz4: batch*h*w*16 bf16 (written by fbme_cnn_loss); part: fbme_cnn_part_floats.
int fbme_cnn_backward(void* stream, int batch, int h, int w, const void* x, const void* act, const void* wb,
                      const void* dz4, void* dz, float* part, float* grads) {
  if (batch < 1 || h < 1 || w < 1) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const size_t plane = (size_t)batch * h * w * 32;
  const uint16_t* a = (const uint16_t*)act;
  const uint16_t* wt = (const uint16_t*)wb;
  uint16_t* d0 = (uint16_t*)dz;
  uint16_t* d1 = d0 + plane;
  const int rows = (int)fbme_cnn_grad_rows(h);
  dim3 grid((w + TILE_W - 1) / TILE_W, h, batch);
  dim3 ggrid((w + TILE_W - 1) / TILE_W, (h + rows - 1) / rows, batch);
  const int gblocks = ggrid.x * ggrid.y * ggrid.z;
  const int rows_all = batch * h;
  auto wgrad = [&](int l, const uint16_t* xin, const uint16_t* dzl) {
    const int cin = CH[l], cout = CH[l + 1];
    if (cin == 16 && cout == 32)
      hipLaunchKernelGGL((k_conv3x3_dw<16, 32>), ggrid, dim3(BLK), 0, st, xin, dzl, part, h, w, rows);
    else if (cin == 32 && cout == 32)
      hipLaunchKernelGGL((k_conv3x3_dw<32, 32>), ggrid, dim3(BLK), 0, st, xin, dzl, part, h, w, rows);
    else
      hipLaunchKernelGGL((k_conv3x3_dw<32, 16>), ggrid, dim3(BLK), 0, st, xin, dzl, part, h, w, rows);
    const int n = 9 * cin * cout;
    hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(n)), dim3(BLK), 0, st, part, gblocks, n, grads + w_off(l));
    if (cout == 32)
      hipLaunchKernelGGL((k_bias_grad<32>), dim3(rows_all), dim3(BLK), 0, st, dzl, w, part);
    else
      hipLaunchKernelGGL((k_bias_grad<16>), dim3(rows_all), dim3(BLK), 0, st, dzl, w, part);
    hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(cout)), dim3(BLK), 0, st, part, rows_all, cout, grads + b_off(l));
  };
  // layer 4 (32 -> 16): dW from h3 and dz4; dz3 = conv(dz4, W4 reversed) masked by h3
  wgrad(3, a + 2 * plane, (const uint16_t*)dz4);
  hipLaunchKernelGGL((k_conv3x3<16, 32, 2>), grid, dim3(BLK), 0, st, (const uint16_t*)dz4, wt + wbf_off(3), nullptr,
                     a + 2 * plane, d0, nullptr, h, w);
  // layer 3
  wgrad(2, a + plane, d0);
  hipLaunchKernelGGL((k_conv3x3<32, 32, 2>), grid, dim3(BLK), 0, st, d0, wt + wbf_off(2), nullptr, a + plane, d1,
                     nullptr, h, w);
  // layer 2
  wgrad(1, a, d1);
  hipLaunchKernelGGL((k_conv3x3<32, 32, 2>), grid, dim3(BLK), 0, st, d1, wt + wbf_off(1), nullptr, a, d0, nullptr, h,
                     w);
  // layer 1 (input gradient not needed)
  wgrad(0, (const uint16_t*)x, d0);
  return (int)hipGetLastError();
}

int fbme_cnn_adam(void* stream, long long n, float* params, const float* grads, float* m, float* v, float lr,
                  int step) {
  if (n < 1 || step < 1) return FBME_ERR_ARG;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(BLK), 0, (hipStream_t)stream, params, grads, m, v, (int)n, lr, b1,
                     b2, 1e-8f, c1, c2);
  return (int)hipGetLastError();
}

int fbme_cnn_apply_warm_start(void* stream, int batch, int n, float* gamma, const float* out, const float* floor_,
                              float clampv) {
  if (batch < 1 || n < 1) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_warm, dim3(nblk(n), batch), dim3(BLK), 0, (hipStream_t)stream, gamma, out, floor_, n, clampv);
  return (int)hipGetLastError();
}

}  // extern "C"
