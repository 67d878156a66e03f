// Standalone HIP operators: the prefix DFT operator A / A^H on a uniform
// batch and the elementwise UAMP kernels with the contracts of the
// reference's Triton wrappers (torch_sbl/triton_kernels.py:476-904).
//
// These are the kernel-level drop-ins used by the eager Python facade, the
// PCG path and the parity tests; the locked plans use the fused passes of
// fbme_uamp.hip instead.
#include <algorithm>

#include <cfloat>

#include "fbme_device.h"

using namespace fbme;

namespace {

constexpr int OP_TILE = FBME_TILE_MAX;

// Column pass on a [batch, rows_in, cols] array (row-major) producing
// [batch, rows_out, cols]: FFT of length n <= TWN along the rows axis,
// reading the first `rows_in` entries (zero-padded) and keeping the first
// `rows_out`.
template <bool INV, int TWN>
__global__ void __launch_bounds__(FBME_BLOCK)
k_op_col(const float2* __restrict__ in, float2* __restrict__ out, int n, int rows_in, int rows_out, int cols,
         int tile_cols, unsigned code) {
  __shared__ float2 buf[FBME_LDS_SLOTS(OP_TILE)];
  __shared__ float2 tw[TWN];
  const int b = blockIdx.y;
  const int c0 = blockIdx.x * tile_cols;
  const int cc = min(tile_cols, cols - c0);
  const int stride = n + 1;
  build_twiddles(tw, n);
  const float2* src = in + (long long)b * rows_in * cols;
  for (int e = threadIdx.x; e < n * cc; e += FBME_BLOCK) {
    const int i = e / cc, c = e - i * cc;
    buf[lds_idx(c * stride + i)] = (i < rows_in) ? src[(long long)i * cols + c0 + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lines<16, INV>(buf, n, cc, stride, code, tw);
  float2* dst = out + (long long)b * rows_out * cols;
  for (int e = threadIdx.x; e < rows_out * cc; e += FBME_BLOCK) {
    const int i = e / cc, c = e - i * cc;
    dst[(long long)i * cols + c0 + c] = buf[lds_idx(c * stride + i)];
  }
}

// Row pass on [batch, rows, len_in] -> [batch, rows, len_out], FFT length
// n <= TWN.
template <bool INV, int TWN>
__global__ void __launch_bounds__(FBME_BLOCK)
k_op_row(const float2* __restrict__ in, float2* __restrict__ out, int n, int len_in, int len_out, int rows,
         int tile_rows, unsigned code) {
  __shared__ float2 buf[FBME_LDS_SLOTS(OP_TILE)];
  __shared__ float2 tw[TWN];
  const int b = blockIdx.y;
  const int r0 = blockIdx.x * tile_rows;
  const int rr = min(tile_rows, rows - r0);
  const int stride = n + 1;
  build_twiddles(tw, n);
  const float2* src = in + ((long long)b * rows + r0) * len_in;
  for (int e = threadIdx.x; e < rr * n; e += FBME_BLOCK) {
    const int r = e / n, j = e - r * n;
    buf[lds_idx(r * stride + j)] = (j < len_in) ? src[(long long)r * len_in + j] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lines<16, INV>(buf, n, rr, stride, code, tw);
  float2* dst = out + ((long long)b * rows + r0) * len_out;
  for (int e = threadIdx.x; e < rr * len_out; e += FBME_BLOCK) {
    const int r = e / len_out, j = e - r * len_out;
    dst[(long long)r * len_out + j] = buf[lds_idx(r * stride + j)];
  }
}

unsigned radix_code_host(int len) {
  unsigned code = 0;
  int shift = 0, rem = len;
  auto push = [&](int r) {
    code |= (unsigned)r << shift;
    shift += 4;
  };
  while (rem % 5 == 0) { push(5); rem /= 5; }
  while (rem % 3 == 0) { push(3); rem /= 3; }
  while (rem % 8 == 0) { push(8); rem /= 8; }
  while (rem % 4 == 0) { push(4); rem /= 4; }
  while (rem % 2 == 0) { push(2); rem /= 2; }
  if (rem != 1 || shift > 28) return 0xFFFFFFFFu;
  return code;
}

// ---------------------------------------------------------- UAMP elementwise
__global__ void k_message(const float2* __restrict__ y, const float2* __restrict__ pred,
                          const float2* __restrict__ old, const float* __restrict__ dv_,
                          const float* __restrict__ noise_, float2* __restrict__ out, float* __restrict__ tau,
                          int m, float damping) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const float dv = dv_[b], noise = noise_[b];
  const float precision = 1.0f / (dv + noise);
  if (i < m) {
    const long long gi = (long long)b * m + i;
    const float2 yy = y[gi], po = pred[gi], mo = old[gi];
    const float2 pseudo = make_float2(po.x - dv * mo.x, po.y - dv * mo.y);
    const float2 cand = make_float2(precision * (yy.x - pseudo.x), precision * (yy.y - pseudo.y));
    const float keep = damping, update = 1.0f - damping;
    out[gi] = make_float2(keep * mo.x + update * cand.x, keep * mo.y + update * cand.y);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) tau[b] = (dv + noise) / (float)m;
}

__global__ void __launch_bounds__(FBME_BLOCK)
k_denoise(const float* __restrict__ prior_, const float* __restrict__ tau_, const float2* __restrict__ lifted,
          const float2* __restrict__ old_mean, const float* __restrict__ old_var, float2* __restrict__ mean,
          float* __restrict__ var, const float* __restrict__ floor_, float* __restrict__ out_prior, int n,
          float damping, float tiny, float hyper_shape, float hyper_rate, float prior_damping, float prior_max,
          double* __restrict__ partial) {
  __shared__ double red[8];
  const int b = blockIdx.y;
  const int i = blockIdx.x * FBME_BLOCK + threadIdx.x;
  const float tau = tau_[b];
  double rel_acc = 0.0, var_acc = 0.0;
  if (i < n) {
    const long long gi = (long long)b * n + i;
    const float prior = prior_[gi];
    const float2 om = old_mean[gi], lm = lifted[gi];
    const float ov = old_var[gi];
    const float denom = fmaxf(prior + tau, tiny);
    const float gain = prior / denom;
    const float2 pm = make_float2(om.x + tau * lm.x, om.y + tau * lm.y);
    const float2 cm = make_float2(gain * pm.x, gain * pm.y);
    const float cvar = prior * tau / denom;
    const float keep = damping, update = 1.0f - damping;
    const float2 mu = make_float2(keep * om.x + update * cm.x, keep * om.y + update * cm.y);
    const float v = fmaxf(keep * ov + update * cvar, tiny);
    float rel = 1.0f - v / fmaxf(prior, tiny);
    rel = fmaxf(0.0f, fminf(1.0f, rel));
    mean[gi] = mu;
    var[gi] = v;
    rel_acc = rel;
    var_acc = v;
    if (floor_) {
      const float fl = floor_[b];
      const float second = mu.x * mu.x + mu.y * mu.y + v;
      float cp = (second + hyper_rate) / (1.0f + hyper_shape);
      cp = fmaxf(fl, fminf(prior_max, cp));
      const float nxt = prior_damping * prior + (1.0f - prior_damping) * cp;
      out_prior[gi] = fmaxf(nxt, fl);
    }
  }
  block_sum2(rel_acc, var_acc, red);
  if (threadIdx.x == 0) {
    const long long slot = (long long)b * gridDim.x + blockIdx.x;
    partial[2 * slot] = rel_acc;
    partial[2 * slot + 1] = var_acc;
  }
}

__global__ void k_finish_pairs(const double* __restrict__ partial, int nblocks, float* __restrict__ out_a,
                               float* __restrict__ out_b) {
  const int b = blockIdx.x;
  double a = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += 64) {
    a += partial[2 * ((long long)b * nblocks + i)];
    c += partial[2 * ((long long)b * nblocks + i) + 1];
  }
  a = wave_sum(a);
  c = wave_sum(c);
  if (threadIdx.x == 0) {
    out_a[b] = (float)a;
    out_b[b] = (float)c;
  }
}

__global__ void k_prior(const float2* __restrict__ mean, const float* __restrict__ pvar, const float* __restrict__ old,
                        const float* __restrict__ floor_, float* __restrict__ out, int n, float hyper_shape,
                        float hyper_rate, float damping, float maximum) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long gi = (long long)b * n + i;
  const float fl = floor_[b];
  const float2 mu = mean[gi];
  const float second = mu.x * mu.x + mu.y * mu.y + pvar[gi];
  float cand = (second + hyper_rate) / (1.0f + hyper_shape);
  cand = fmaxf(fl, fminf(maximum, cand));
  const float upd = damping * old[gi] + (1.0f - damping) * cand;
  out[gi] = fmaxf(upd, fl);
}

__global__ void __launch_bounds__(FBME_BLOCK)
k_residual_partial(const float2* __restrict__ data, const float2* __restrict__ pred, int m,
                   double* __restrict__ partial, const unsigned char* __restrict__ mask) {
  __shared__ double red[8];
  const int b = blockIdx.y;
  const int i = blockIdx.x * FBME_BLOCK + threadIdx.x;
  double acc = 0.0;
  if (i < m && (!mask || mask[i])) {
    const long long gi = (long long)b * m + i;
    const float rx = data[gi].x - pred[gi].x, ry = data[gi].y - pred[gi].y;
    acc = (double)rx * rx + (double)ry * ry;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) partial[(long long)b * gridDim.x + blockIdx.x] = acc;
}

__global__ void k_residual_finish(const double* __restrict__ partial, int nblocks, float* __restrict__ out,
                                  const float* __restrict__ old_noise, const float* __restrict__ dof_,
                                  const float* __restrict__ floor_, int m, float noise_shape, float noise_rate,
                                  float damping, int evidence, float epsilon, int noise_mode) {
  const int b = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += 64) a += partial[(long long)b * nblocks + i];
  a = wave_sum(a);
  if (threadIdx.x != 0) return;
  if (!noise_mode) {
    out[b] = (float)a;
    return;
  }
  const float energy = (float)a;
  const float old = old_noise[b], dof = dof_[b], fl = floor_[b];
  float cand;
  if (evidence) {
    cand = (energy + noise_rate) / fmaxf((float)m - dof + noise_shape, epsilon);
  } else {
    cand = (energy + old * dof + noise_rate) / ((float)m + noise_shape);
  }
  cand = fmaxf(cand, fl);
  out[b] = damping * old + (1.0f - damping) * cand;
}

// Shared noise of joint sparsity: each snapshot's update (EM or evidence
// form, floored at its own floor), averaged over the snapshots, floored at
// the largest floor, then damped (solver.py:1737-1774).  One wavefront.
__global__ void __launch_bounds__(64)
k_joint_noise_finish(const double* __restrict__ partial, int nblocks, int batch, float* __restrict__ out,
                     const float* __restrict__ old_noise, const float* __restrict__ dof_,
                     const float* __restrict__ floor_, int m, float noise_shape, float noise_rate, float damping,
                     int evidence, float epsilon) {
  float acc = 0.f, fmax_ = 0.f;
  for (int b = 0; b < batch; ++b) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += 64) a += partial[(long long)b * nblocks + i];
    a = wave_sum_all(a);
    const float energy = (float)a;
    const float old = old_noise[b], dof = dof_[b], fl = floor_[b];
    float cand;
    if (evidence) {
      cand = (energy + noise_rate) / fmaxf((float)m - dof + noise_shape, epsilon);
    } else {
      cand = (energy + old * dof + noise_rate) / ((float)m + noise_shape);
    }
    acc += fmaxf(cand, fl);
    fmax_ = fmaxf(fmax_, fl);
  }
  const float shared = fmaxf(acc / (float)batch, fmax_);
  if (threadIdx.x == 0)
    for (int b = 0; b < batch; ++b) out[b] = damping * old_noise[b] + (1.0f - damping) * shared;
}

// A fixed noise variance: the value itself, or max(value, floor) in place
// of an update (shared: the batch's largest floor; the mean of the clamped
// values, clamped again at the largest floor, is that).
__global__ void k_fixed_noise(int batch, float value, const float* __restrict__ floor_, float* __restrict__ noise,
                              int shared) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  float f = 0.f;
  if (floor_) {
    if (shared)
      for (int i = 0; i < batch; ++i) f = fmaxf(f, floor_[i]);
    else
      f = floor_[b];
  }
  noise[b] = floor_ ? fmaxf(value, f) : value;
}

inline int nblk(int n) { return (n + FBME_BLOCK - 1) / FBME_BLOCK; }

// Plain HBM streams (the measured ceiling the kernels are compared with):
// one pass over the buffer, U independent 16-byte accesses in flight per
// thread, coalesced across the workgroup.
template <int U>
__global__ void __launch_bounds__(256) k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     long long n) {
  const long long base = (long long)blockIdx.x * 256 * U + threadIdx.x;
  float4 v[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const long long i = base + 256 * k;
    if (i < n) v[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const long long i = base + 256 * k;
    if (i < n) dst[i] = v[k];
  }
}

template <int U>
__global__ void __launch_bounds__(256) k_stream_fill(float4* __restrict__ dst, long long n, float value) {
  const long long base = (long long)blockIdx.x * 256 * U + threadIdx.x;
  const float4 f = make_float4(value, value, value, value);
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const long long i = base + 256 * k;
    if (i < n) dst[i] = f;
  }
}

}  // namespace

extern "C" {

int fbme_stream_copy(void* stream, const void* src, void* dst, long long bytes, int unroll) {
  // bytes: a multiple of 16; src, dst 16-byte aligned and disjoint; unroll 1, 2, 4 or 8.
  if (!src || !dst || bytes <= 0 || (bytes & 15) || ((uintptr_t)src & 15) || ((uintptr_t)dst & 15))
    return FBME_ERR_ARG;
  const long long n = bytes / 16;
  hipStream_t st = (hipStream_t)stream;
#define FBME_COPY(U)                                                                                         \
  hipLaunchKernelGGL(k_stream_copy<U>, dim3((unsigned)((n + 256 * U - 1) / (256 * U))), dim3(256), 0, st,   \
                     (const float4*)src, (float4*)dst, n)
  switch (unroll) {
    case 1: FBME_COPY(1); break;
    case 2: FBME_COPY(2); break;
    case 4: FBME_COPY(4); break;
    case 8: FBME_COPY(8); break;
    default: return FBME_ERR_ARG;
  }
#undef FBME_COPY
  return (int)hipGetLastError();
}

int fbme_stream_fill(void* stream, void* dst, long long bytes, float value, int unroll) {
  if (!dst || bytes <= 0 || (bytes & 15) || ((uintptr_t)dst & 15)) return FBME_ERR_ARG;
  const long long n = bytes / 16;
  hipStream_t st = (hipStream_t)stream;
#define FBME_FILL(U)                                                                                         \
  hipLaunchKernelGGL(k_stream_fill<U>, dim3((unsigned)((n + 256 * U - 1) / (256 * U))), dim3(256), 0, st,   \
                     (float4*)dst, n, value)
  switch (unroll) {
    case 1: FBME_FILL(1); break;
    case 2: FBME_FILL(2); break;
    case 4: FBME_FILL(4); break;
    case 8: FBME_FILL(8); break;
    default: return FBME_ERR_ARG;
  }
#undef FBME_FILL
  return (int)hipGetLastError();
}

unsigned fbme_radix_code(int len) {
  if (len < 1 || len > FBME_FFT_MAX) return 0xFFFFFFFFu;
  return radix_code_host(len);
}

int fbme_column_fft(void* stream, int batch, int n, int rows_in, int rows_out, int cols, const void* in, void* out,
                    int inverse) {
  if (batch < 1 || n < 1 || n > FBME_FFT_MAX || rows_in > n || rows_out > n || cols < 1) return FBME_ERR_ARG;
  const unsigned code = radix_code_host(n);
  if (code == 0xFFFFFFFFu) return FBME_ERR_ARG;
  const int tc = std::max(1, std::min(cols, std::min(16, OP_TILE / n)));
  dim3 grid((cols + tc - 1) / tc, batch);
#define FBME_OPCOL(INV, TWN)                                                                                   \
  hipLaunchKernelGGL((k_op_col<INV, TWN>), grid, dim3(FBME_BLOCK), 0, (hipStream_t)stream, (const float2*)in, \
                     (float2*)out, n, rows_in, rows_out, cols, tc, code)
  if (n > 1024) {
    if (inverse) FBME_OPCOL(true, FBME_FFT_MAX);
    else FBME_OPCOL(false, FBME_FFT_MAX);
  } else {
    if (inverse) FBME_OPCOL(true, 1024);
    else FBME_OPCOL(false, 1024);
  }
#undef FBME_OPCOL
  return (int)hipGetLastError();
}

int fbme_prefix_forward(void* stream, int batch, int n1, int n2, int m1, int m2, const void* x, void* y,
                        void* mid) {
  if (batch < 1 || m1 < 1 || m2 < 1 || m1 > n1 || m2 > n2 || n1 > 1024 || n2 > 1024) return FBME_ERR_ARG;
  const unsigned c1 = radix_code_host(n1), c2 = radix_code_host(n2);
  if (c1 == 0xFFFFFFFFu || c2 == 0xFFFFFFFFu) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int tr = std::max(1, std::min(n1, std::min(8, OP_TILE / n2)));
  hipLaunchKernelGGL((k_op_row<false, 1024>), dim3((n1 + tr - 1) / tr, batch), dim3(FBME_BLOCK), 0, st,
                     (const float2*)x, (float2*)mid, n2, n2, m2, n1, tr, c2);
  const int tc = std::max(1, std::min(m2, std::min(16, OP_TILE / n1)));
  hipLaunchKernelGGL((k_op_col<false, 1024>), dim3((m2 + tc - 1) / tc, batch), dim3(FBME_BLOCK), 0, st,
                     (const float2*)mid, (float2*)y, n1, n1, m1, m2, tc, c1);
  return (int)hipGetLastError();
}

int fbme_prefix_adjoint(void* stream, int batch, int n1, int n2, int m1, int m2, const void* v, void* x,
                        void* mid) {
  if (batch < 1 || m1 < 1 || m2 < 1 || m1 > n1 || m2 > n2 || n1 > 1024 || n2 > 1024) return FBME_ERR_ARG;
  const unsigned c1 = radix_code_host(n1), c2 = radix_code_host(n2);
  if (c1 == 0xFFFFFFFFu || c2 == 0xFFFFFFFFu) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int tc = std::max(1, std::min(m2, std::min(16, OP_TILE / n1)));
  hipLaunchKernelGGL((k_op_col<true, 1024>), dim3((m2 + tc - 1) / tc, batch), dim3(FBME_BLOCK), 0, st,
                     (const float2*)v, (float2*)mid, n1, m1, n1, m2, tc, c1);
  const int tr = std::max(1, std::min(n1, std::min(8, OP_TILE / n2)));
  hipLaunchKernelGGL((k_op_row<true, 1024>), dim3((n1 + tr - 1) / tr, batch), dim3(FBME_BLOCK), 0, st,
                     (const float2*)mid, (float2*)x, n2, m2, n2, n1, tr, c2);
  return (int)hipGetLastError();
}

int fbme_uamp_message_update(void* stream, int batch, int m, const void* y, const void* prediction,
                             const void* old_message, const float* data_variance, const float* noise_variance,
                             void* message, float* pseudo_variance, float damping) {
  if (batch < 1 || m < 1) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_message, dim3(nblk(m), batch), dim3(FBME_BLOCK), 0, (hipStream_t)stream,
                     (const float2*)y, (const float2*)prediction, (const float2*)old_message, data_variance,
                     noise_variance, (float2*)message, pseudo_variance, m, damping);
  return (int)hipGetLastError();
}

int fbme_uamp_denoise(void* stream, int batch, int n, const float* prior_variance, const float* pseudo_variance,
                      const void* lifted_message, const void* old_mean, const float* old_variance, void* mean,
                      float* variance, float* effective_dof, float* variance_sum, const float* prior_floor,
                      float* output_prior, float damping, float tiny, float hyper_shape, float hyper_rate,
                      float prior_damping, float prior_maximum, double* partial) {
  if (batch < 1 || n < 1 || !partial) return FBME_ERR_ARG;
  if (prior_floor && !output_prior) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int nb = nblk(n);
  hipLaunchKernelGGL(k_denoise, dim3(nb, batch), dim3(FBME_BLOCK), 0, st, prior_variance, pseudo_variance,
                     (const float2*)lifted_message, (const float2*)old_mean, old_variance, (float2*)mean, variance,
                     prior_floor, output_prior, n, damping, tiny, hyper_shape, hyper_rate, prior_damping,
                     prior_maximum, partial);
  hipLaunchKernelGGL(k_finish_pairs, dim3(batch), dim3(64), 0, st, (const double*)partial, nb, effective_dof,
                     variance_sum);
  return (int)hipGetLastError();
}

int fbme_uamp_prior_update(void* stream, int batch, int n, const void* mean, const float* posterior_variance,
                           const float* old_prior, const float* prior_floor, float* output, float hyper_shape,
                           float hyper_rate, float damping, float maximum) {
  if (batch < 1 || n < 1) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_prior, dim3(nblk(n), batch), dim3(FBME_BLOCK), 0, (hipStream_t)stream,
                     (const float2*)mean, posterior_variance, old_prior, prior_floor, output, n, hyper_shape,
                     hyper_rate, damping, maximum);
  return (int)hipGetLastError();
}

int fbme_complex_residual_energy(void* stream, int batch, int m, const void* data, const void* prediction,
                                 float* output, double* partial) {
  if (batch < 1 || m < 1 || !partial) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int nb = nblk(m);
  hipLaunchKernelGGL(k_residual_partial, dim3(nb, batch), dim3(FBME_BLOCK), 0, st, (const float2*)data,
                     (const float2*)prediction, m, partial, (const unsigned char*)nullptr);
  hipLaunchKernelGGL(k_residual_finish, dim3(batch), dim3(64), 0, st, (const double*)partial, nb, output,
                     (const float*)nullptr, (const float*)nullptr, (const float*)nullptr, m, 0.f, 0.f, 0.f, 0,
                     0.f, 0);
  return (int)hipGetLastError();
}

int fbme_uamp_noise_update(void* stream, int batch, int m, const void* data, const void* prediction,
                           const float* old_noise, const float* effective_dof, const float* variance_floor,
                           float* output, float noise_shape, float noise_rate, float damping, int evidence_update,
                           float epsilon, double* partial) {
  return fbme_uamp_noise_update_masked(stream, batch, m, nullptr, m, data, prediction, old_noise, effective_dof,
                                       variance_floor, output, noise_shape, noise_rate, damping, evidence_update,
                                       epsilon, partial);
}

int fbme_fixed_noise(void* stream, int batch, float value, const float* variance_floor, float* noise, int shared) {
  if (batch < 1 || !noise || !(value > 0.f) || !(value <= FLT_MAX)) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_fixed_noise, dim3((batch + 63) / 64), dim3(64), 0, (hipStream_t)stream, batch, value,
                     variance_floor, noise, shared);
  return (int)hipGetLastError();
}

int fbme_joint_noise_update(void* stream, int batch, int m, const unsigned char* mask, int observed,
                            const void* data, const void* prediction, const float* old_noise,
                            const float* effective_dof, const float* variance_floor, float* output, float noise_shape,
                            float noise_rate, float damping, int evidence_update, float epsilon, double* partial) {
  if (batch < 1 || m < 1 || observed < 1 || observed > m || !partial) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int nb = nblk(m);
  hipLaunchKernelGGL(k_residual_partial, dim3(nb, batch), dim3(FBME_BLOCK), 0, st, (const float2*)data,
                     (const float2*)prediction, m, partial, mask);
  hipLaunchKernelGGL(k_joint_noise_finish, dim3(1), dim3(64), 0, st, (const double*)partial, nb, batch, output,
                     old_noise, effective_dof, variance_floor, observed, noise_shape, noise_rate, damping,
                     evidence_update, epsilon);
  return (int)hipGetLastError();
}

int fbme_uamp_noise_update_masked(void* stream, int batch, int m, const unsigned char* mask, int observed,
                                  const void* data, const void* prediction, const float* old_noise,
                                  const float* effective_dof, const float* variance_floor, float* output,
                                  float noise_shape, float noise_rate, float damping, int evidence_update,
                                  float epsilon, double* partial) {
  if (batch < 1 || m < 1 || observed < 1 || observed > m || !partial) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int nb = nblk(m);
  hipLaunchKernelGGL(k_residual_partial, dim3(nb, batch), dim3(FBME_BLOCK), 0, st, (const float2*)data,
                     (const float2*)prediction, m, partial, mask);
  hipLaunchKernelGGL(k_residual_finish, dim3(batch), dim3(64), 0, st, (const double*)partial, nb, output, old_noise,
                     effective_dof, variance_floor, observed, noise_shape, noise_rate, damping, evidence_update,
                     epsilon, 1);
  return (int)hipGetLastError();
}

}  // extern "C"
