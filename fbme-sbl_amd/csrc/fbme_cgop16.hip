// Covariance matvec of the general solver on the 256 x 256 grid (gfx950).
//
// Q d = noise d + A Gamma A^H d (the operator torch_sbl/operators.py:328-355
// applies through the BTTB circulant embedding) evaluated through the grid
// instead: three passes on the register-resident 256-point FFT of
// fbme_fft16.h, with the probes batched in the lane dimension like the BTTB
// path.  Per lane (d and out [M1, M2], intermediate [256, M2]):
//   1. column pass of A^H: mid[n1, c] = sum_{m1 < M1} e^{+2 pi i m1 n1 / 256} d[m1, c];
//   2. row pass: the row's unnormalised inverse transform (M2 -> 256, zero
//      padded), times gamma[n1, :], forward transform, first M2 outputs back
//      into mid (in place);
//   3. column pass of A: out[k1, c] = noise d[k1, c] + sum_{n1} e^{-2 pi i n1 k1 / 256} mid[n1, c], k1 < M1.
// HBM traffic per lane: 16 M1 M2 + 32 * 256 M2 bytes (+ gamma rows, shared
// by the lanes of a scene through L2) against 16 M + 32 M1 E2 (+ the
// embedded spectrum) for the BTTB passes on the E1 x E2 = 320 x 320
// embedding, and no per-iteration spectrum to build.
#include "fbme_device.h"
#include "fbme_fft16.h"

using namespace fbme;

namespace {

constexpr int GC = 16;             // columns per column-pass workgroup
constexpr int XS = 16 * 17 + 8;    // LDS slots of one 256-point transpose

__global__ void __launch_bounds__(16 * GC)
k_gop_col_adjoint(const float2* __restrict__ d, float2* __restrict__ mid, int m1, int m2,
                  const float2* __restrict__ twtab) {
  __shared__ float2 xs[16 * 17 * GC];   // column c's element i at xs[i * GC + c]
  __shared__ float2 tw[256];
  const int c = threadIdx.x % GC, t = threadIdx.x / GC;
  const int lane = blockIdx.y;
  const int col = blockIdx.x * GC + c;
  const bool act = col < m2;
  load_tw256(tw, twtab);
  const float2* src = d + (long long)lane * m1 * m2;
  float2 v[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int i = t + 16 * j;
    v[j] = (act && i < m1) ? src[(long long)i * m2 + col] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft256<true, GC>(v, tw, xs + c, t);
  if (act) {
    float2* dst = mid + (long long)lane * 256 * m2 + col;
#pragma unroll
    for (int n = 0; n < 16; ++n) dst[(long long)(t + 16 * n) * m2] = v[n];
  }
}

__global__ void __launch_bounds__(64)
k_gop_row(float2* __restrict__ mid, const float* __restrict__ gamma, int m2, int rhs,
          const float2* __restrict__ twtab) {
  __shared__ float2 xs[4][XS];
  __shared__ float2 tw[256];
  const int g = threadIdx.x >> 4, t = threadIdx.x & 15;
  const int lane = blockIdx.y;
  const int b = lane / rhs;
  const int row = blockIdx.x * 4 + g;
  load_tw256(tw, twtab);
  float2* line = mid + ((long long)lane * 256 + row) * m2;
  const float* gr = gamma + ((long long)b * 256 + row) * 256;
  float2 v[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int c = t + 16 * j;
    v[j] = (c < m2) ? line[c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft256<true>(v, tw, xs[g], t);
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const float gv = gr[t + 16 * n];
    v[n] = make_float2(v[n].x * gv, v[n].y * gv);
  }
  fft256<false>(v, tw, xs[g], t);
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int k = t + 16 * n;
    if (k < m2) line[k] = v[n];
  }
}

// DOT: also part[lane, block] = sum Re(conj(d) out) over the block's
// outputs (FP64, fixed order), the PCG curvature <d, Q d>.
template <bool DOT>
__global__ void __launch_bounds__(16 * GC)
k_gop_col_forward(const float2* __restrict__ mid, const float2* __restrict__ d, const float* __restrict__ noise,
                  float2* __restrict__ out, int m1, int m2, int rhs, const float2* __restrict__ twtab,
                  double* __restrict__ part) {
  __shared__ float2 xs[16 * 17 * GC];
  __shared__ float2 tw[256];
  __shared__ double red[8];
  const int c = threadIdx.x % GC, t = threadIdx.x / GC;
  const int lane = blockIdx.y;
  const int col = blockIdx.x * GC + c;
  const bool act = col < m2;
  load_tw256(tw, twtab);
  const float2* src = mid + (long long)lane * 256 * m2 + col;
  float2 v[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = act ? src[(long long)(t + 16 * j) * m2] : make_float2(0.f, 0.f);
  __syncthreads();
  fft256<false, GC>(v, tw, xs + c, t);
  const float nv = noise[lane / rhs];
  const long long base = (long long)lane * m1 * m2 + col;
  double acc = 0.0;
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int k = t + 16 * n;
    if (act && k < m1) {
      const float2 dv = d[base + (long long)k * m2];
      const float2 o = make_float2(v[n].x + nv * dv.x, v[n].y + nv * dv.y);
      out[base + (long long)k * m2] = o;
      if (DOT) acc += (double)dv.x * o.x + (double)dv.y * o.y;
    }
  }
  if (DOT) {
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[(long long)lane * gridDim.x + blockIdx.x] = acc;
  }
}

// out[l] = sum_i part[l, i], one wavefront per lane (fixed shuffle tree).
__global__ void __launch_bounds__(64) k_gop_finish(const double* __restrict__ part, int blocks, int lanes,
                                                   double* __restrict__ out) {
  const int l = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < blocks; i += 64) s += part[(long long)l * blocks + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[l] = s;
}

}  // namespace

extern "C" {

int fbme_grid_matvec_dot(void* stream, int lanes, int rhs, int m1, int m2, const float* gamma, const float* noise,
                         const void* d, void* out, void* mid, const void* tw256, double* part, double* dot) {
  if (lanes < 1 || rhs < 1 || lanes % rhs != 0 || m1 < 1 || m2 < 1 || m1 > 256 || m2 > 256 || d == out)
    return FBME_ERR_ARG;
  if ((part == nullptr) != (dot == nullptr)) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const dim3 cg((m2 + GC - 1) / GC, lanes);
  hipLaunchKernelGGL(k_gop_col_adjoint, cg, dim3(16 * GC), 0, st, (const float2*)d, (float2*)mid, m1, m2,
                     (const float2*)tw256);
  hipLaunchKernelGGL(k_gop_row, dim3(64, lanes), dim3(64), 0, st, (float2*)mid, gamma, m2, rhs,
                     (const float2*)tw256);
  if (dot) {
    hipLaunchKernelGGL(k_gop_col_forward<true>, cg, dim3(16 * GC), 0, st, (const float2*)mid, (const float2*)d, noise,
                       (float2*)out, m1, m2, rhs, (const float2*)tw256, part);
    hipLaunchKernelGGL(k_gop_finish, dim3(lanes), dim3(64), 0, st, (const double*)part, (int)cg.x, lanes,
                       dot);
  } else {
    hipLaunchKernelGGL(k_gop_col_forward<false>, cg, dim3(16 * GC), 0, st, (const float2*)mid, (const float2*)d,
                       noise, (float2*)out, m1, m2, rhs, (const float2*)tw256, (double*)nullptr);
  }
  return (int)hipGetLastError();
}

int fbme_grid_matvec(void* stream, int lanes, int rhs, int m1, int m2, const float* gamma, const float* noise,
                     const void* d, void* out, void* mid, const void* tw256) {
  return fbme_grid_matvec_dot(stream, lanes, rhs, m1, m2, gamma, noise, d, out, mid, tw256, nullptr, nullptr);
}

}  // extern "C"
