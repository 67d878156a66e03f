// Register-resident 256-point FFT passes of the FBME plans (gfx950).
//
// Same three HBM passes per UAMP iteration as fbme_uamp.hip (column pass of
// A^H fused with the message update, row pass fused with denoise/ARD, column
// pass of A fused with the residual), specialised for FFT length 256 =
// 16 x 16 along the transformed axis:
//
//   * a 256-point line is owned by 16 lanes; lane t holds the 16 elements
//     t + 16 j (j = 0..15) in registers, so global loads and stores are
//     16-lane contiguous segments and the input of the first radix-16 stage
//     comes straight from HBM;
//   * radix-16 DFTs run in registers (4 x 4 split), one twiddle multiply
//     exp(-+2 pi i t n / 256), one transpose through LDS (two barriers), a
//     second radix-16 DFT; lane t then holds outputs t + 16 n;
//   * the row pass uses one wavefront per workgroup (4 rows), so its
//     barriers are single-wave; the column passes use COLS columns x 16 lanes.
//
// Versus the LDS Stockham kernels (three radix stages, two barriers each, per
// FFT) this removes two thirds of the LDS traffic and most barriers, which is
// what bounds those kernels below the HBM roofline.  Numerics are the same
// unnormalised transforms (forward exp(-2 pi i jk/L), inverse exp(+...)), the
// same fused element formulas (triton_kernels.py:36-192, 321-354) and the same
// deterministic FP64 partial sums, so results agree with the generic kernels
// to float32 rounding.
#include "fbme_device.h"

using namespace fbme;

namespace {

constexpr float F32_EPS = 1.1920928955078125e-07f;
constexpr float F32_TINY = 1.17549435082228750797e-38f;
constexpr int XS = 16 * 17 + 8;   // LDS slots of one 256-point transpose (padded)

struct UampArgs16 {
  float damping, prior_damping, max_precision, hyper_shape, hyper_rate, blend;
};

// v * (c + i s)
__device__ __forceinline__ float2 cmulc(float2 v, float c, float s) {
  return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
}

// W16^e = exp(-+2 pi i e / 16) for the exponents e = n1 k1 of the 4 x 4 split.
template <bool INV, int E>
__device__ __forceinline__ float2 tw16(float2 v) {
  constexpr float c8 = 0.92387953251128675613f, s8 = 0.38268343236508977173f;
  constexpr float h = 0.70710678118654752440f;
  constexpr float sg = INV ? 1.0f : -1.0f;   // sign of the imaginary part
  if constexpr (E == 0) return v;
  else if constexpr (E == 1) return cmulc(v, c8, sg * s8);
  else if constexpr (E == 2) return cmulc(v, h, sg * h);
  else if constexpr (E == 3) return cmulc(v, s8, sg * c8);
  else if constexpr (E == 4) return rot90<INV>(v);
  else if constexpr (E == 6) return cmulc(v, -h, sg * h);
  else return cmulc(v, -c8, -sg * s8);       // E == 9
}

// In-place natural-order 16-point DFT of v (4 x 4 Cooley-Tukey).
template <bool INV>
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    float2 q[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    bfly4<INV>(q);
    v[n1] = q[0];
    v[n1 + 4] = q[1];
    v[n1 + 8] = q[2];
    v[n1 + 12] = q[3];
  }
  // v[n1 + 4 k1] *= W16^(n1 k1)
  v[5] = tw16<INV, 1>(v[5]);
  v[9] = tw16<INV, 2>(v[9]);
  v[13] = tw16<INV, 3>(v[13]);
  v[6] = tw16<INV, 2>(v[6]);
  v[10] = tw16<INV, 4>(v[10]);
  v[14] = tw16<INV, 6>(v[14]);
  v[7] = tw16<INV, 3>(v[7]);
  v[11] = tw16<INV, 6>(v[11]);
  v[15] = tw16<INV, 9>(v[15]);
  float2 out[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 q[4] = {v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]};
    bfly4<INV>(q);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) out[k1 + 4 * k2] = q[k2];
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = out[i];
}

// 256-point transform of the line owned by the 16 lanes of a group.  On entry
// v[j] = a[t + 16 j]; on exit v[n] = b[t + 16 n].  w[n] = exp(-2 pi i t n / 256).
// x: this group's LDS transpose buffer; every lane of the workgroup must call
// (the barriers are workgroup-wide).
template <bool INV>
__device__ __forceinline__ void fft256(float2 (&v)[16], const float2 (&w)[16], float2* x, int t) {
  dft16<INV>(v);
#pragma unroll
  for (int n = 1; n < 16; ++n) {
    float2 ww = w[n];
    if (INV) ww.y = -ww.y;
    v[n] = cmul(v[n], ww);
  }
#pragma unroll
  for (int n = 0; n < 16; ++n) x[n * 17 + t] = v[n];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = x[t * 17 + j];
  __syncthreads();
  dft16<INV>(v);
}

__device__ __forceinline__ void lane_twiddles(float2 (&w)[16], const float2* __restrict__ table, int t) {
#pragma unroll
  for (int n = 0; n < 16; ++n) w[n] = table[t * n];
}

// ------------------------------------------------------------------- row pass
// Modes as k_row in fbme_uamp.hip: 0 backprojection rows, 1 first UAMP step
// (no inverse FFT), 2 UAMP step, 3 final E-step with top-k candidates.
template <int MODE>
__global__ void __launch_bounds__(64)
k_row16(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
        float2* __restrict__ mid, float2* __restrict__ bp, float* __restrict__ gamma,
        float2* __restrict__ mean, float* __restrict__ var, const double* __restrict__ sc,
        double* __restrict__ part, int pstride, UampArgs16 a, uint32_t* __restrict__ cand_key,
        int* __restrict__ cand_idx, int store_final, const float2* __restrict__ twtab) {
  __shared__ float2 xs[4][XS];
  __shared__ double red[8];
  __shared__ unsigned long long top[MODE == 3 ? FBME_KMAX : 1];
  const int lane = threadIdx.x, g = lane >> 4, t = lane & 15;
  float2* x = xs[g];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const FbmeWork wk = work[wi];
    const FbmeScene s = scenes[wk.scene];
    const bool act = g < wk.count;
    const int row = wk.first + g;
    const long long gb = s.grid_off + (long long)row * 256;
    const long long mb = s.mid_off + (long long)row * s.m2;
    const long long n_grid = (long long)s.n1 * 256;
    const int n_samp = s.m1 * s.m2;
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float noise = (float)scs[SC_NOISE];
    const float floor_v = (float)scs[SC_FLOOR];
    const float power = (float)scs[SC_POWER];
    float2 w[16];
    lane_twiddles(w, twtab + s.tw2_off, t);
    float gv[16], vv[16];
    float2 mv[16], v[16];
    if (MODE >= 1) {
#pragma unroll
      for (int n = 0; n < 16; ++n) {
        const long long k = gb + t + 16 * n;
        if (MODE == 1) {
          mv[n] = act ? bp[k] : make_float2(0.f, 0.f);
        } else {
          gv[n] = act ? gamma[k] : 1.0f;
          mv[n] = act ? mean[k] : make_float2(0.f, 0.f);
          vv[n] = act ? var[k] : 1.0f;
        }
      }
    }
    if (MODE != 1) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int c = t + 16 * j;
        v[j] = (act && c < s.m2) ? mid[mb + c] : make_float2(0.f, 0.f);
      }
      fft256<true>(v, w, x, t);
    }
    if (MODE == 0) {
      double acc = 0.0;
#pragma unroll
      for (int n = 0; n < 16; ++n) {
        if (act) {
          bp[gb + t + 16 * n] = v[n];
          acc += (double)v[n].x * v[n].x + (double)v[n].y * v[n].y;
        }
      }
      acc = block_sum(acc, red);
      if (threadIdx.x == 0) part[(long long)PT_BPE * pstride + s.row_part + wk.slot] = acc;
      continue;
    }
    const float tau = fmaxf((dv + noise) / (float)n_samp, F32_TINY);
    const float keep = a.damping, update = 1.0f - a.damping;
    const float first_scale = update / (dv + noise);
    const float bpe_clamped = fmaxf((float)scs[SC_BPE], F32_EPS);
    const float uniform = power / (float)n_grid;
    double dof_acc = 0.0, vs_acc = 0.0;
    unsigned long long pk[16];
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const long long k = gb + t + 16 * n;
      float2 lifted, mean_old;
      float prior, v_old;
      if (MODE == 1) {
        const float2 b = mv[n];
        lifted = cscale(b, first_scale);
        const float energy = b.x * b.x + b.y * b.y;
        const float structured = energy / bpe_clamped * power;
        prior = fmaxf(a.blend * structured + (1.0f - a.blend) * uniform, floor_v);
        mean_old = make_float2(0.f, 0.f);
        v_old = fmaxf(prior, F32_TINY);
      } else {
        lifted = v[n];
        prior = gv[n];
        mean_old = mv[n];
        v_old = fmaxf(vv[n], F32_TINY);
      }
      // Denoise (triton_kernels.py:131-158).
      const float denom = fmaxf(prior + tau, F32_TINY);
      const float gain = prior / denom;
      const float2 pm = make_float2(mean_old.x + tau * lifted.x, mean_old.y + tau * lifted.y);
      const float2 cm = make_float2(gain * pm.x, gain * pm.y);
      const float cvar = prior * tau / denom;
      const float2 mu = make_float2(keep * mean_old.x + update * cm.x, keep * mean_old.y + update * cm.y);
      const float vn = fmaxf(keep * v_old + update * cvar, F32_TINY);
      float rel = 1.0f - vn / fmaxf(prior, F32_TINY);
      rel = fmaxf(0.0f, fminf(1.0f, rel));
      if (act) {
        dof_acc += (double)rel;
        vs_acc += (double)vn;
      }
      if (MODE == 3) {
        pk[n] = act ? pack_key(score_key(mu.x * mu.x + mu.y * mu.y + vn), g * 256 + t + 16 * n) : 0ull;
        if (store_final && act) {
          mean[k] = mu;
          var[k] = vn;
        }
      } else {
        // Fused ARD update (triton_kernels.py:159-180).
        const float second = mu.x * mu.x + mu.y * mu.y + vn;
        float cp = (second + a.hyper_rate) / (1.0f + a.hyper_shape);
        cp = fmaxf(floor_v, fminf(a.max_precision, cp));
        const float nxt = a.prior_damping * prior + (1.0f - a.prior_damping) * cp;
        if (act) {
          gamma[k] = fmaxf(nxt, floor_v);
          mean[k] = mu;
          var[k] = vn;
        }
        v[n] = act ? mu : make_float2(0.f, 0.f);
      }
    }
    block_sum2(dof_acc, vs_acc, red);
    if (threadIdx.x == 0) {
      part[(long long)PT_DOF * pstride + s.row_part + wk.slot] = dof_acc;
      part[(long long)PT_VS * pstride + s.row_part + wk.slot] = vs_acc;
    }
    if (MODE == 3) {
      // Top-k candidates of the tile for the polish selection
      // (solver.py:1926-1933): score |mu|^2 + v, ties to the lower index.
      wave_topk<16>(pk, s.k, top);
      __syncthreads();
      const long long cbase = (long long)s.cand_off + (long long)wk.slot * s.k;
      for (int i = threadIdx.x; i < s.k; i += 64) {
        const unsigned long long p = top[i];
        cand_key[cbase + i] = (uint32_t)(p >> 32);
        cand_idx[cbase + i] = p ? wk.first * 256 + packed_pos(p) : -1;
      }
      __syncthreads();
      continue;
    }
    // Row pass of A mu keeping the M2 leading outputs.
    fft256<false>(v, w, x, t);
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int c = t + 16 * n;
      if (act && c < s.m2) mid[mb + c] = v[n];
    }
  }
}

// ------------------------------------------------------------- column passes
// COLS columns per workgroup, 16 lanes per column: lane = c + COLS * t.
template <int COLS, int MODE>
__global__ void __launch_bounds__(16 * COLS)
k_col16_adjoint(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
                const float2* __restrict__ y, const float2* __restrict__ pred, float2* __restrict__ msg,
                float2* __restrict__ mid, const double* __restrict__ sc, double* __restrict__ part,
                int pstride, int store_msg, float damping, const float2* __restrict__ twtab) {
  __shared__ float2 xs[COLS][XS];
  __shared__ double red[8];
  const int c = threadIdx.x % COLS, t = threadIdx.x / COLS;
  float2* x = xs[c];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const FbmeWork wk = work[wi];
    const FbmeScene s = scenes[wk.scene];
    const bool act = c < wk.count;
    const int col = wk.first + c;
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float precision = 1.0f / (dv + (float)scs[SC_NOISE]);
    const float keep = damping, update = 1.0f - damping;
    float2 w[16];
    lane_twiddles(w, twtab + s.tw1_off, t);
    float2 v[16];
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int i = t + 16 * j;
      v[j] = make_float2(0.f, 0.f);
      if (act && i < s.m1) {
        const long long gi = s.samp_off + (long long)i * s.m2 + col;
        const float2 yv = y[gi];
        if (MODE == 0) {
          v[j] = yv;
          acc += (double)yv.x * yv.x + (double)yv.y * yv.y;
        } else {
          const float2 pv = pred[gi], mv = msg[gi];
          const float2 pseudo = make_float2(pv.x - dv * mv.x, pv.y - dv * mv.y);
          const float2 cand = make_float2(precision * (yv.x - pseudo.x), precision * (yv.y - pseudo.y));
          v[j] = make_float2(keep * mv.x + update * cand.x, keep * mv.y + update * cand.y);
          if (store_msg) msg[gi] = v[j];
        }
      }
    }
    if (MODE == 0) {
      acc = block_sum(acc, red);
      if (threadIdx.x == 0) part[(long long)PT_YE * pstride + s.col_part + wk.slot] = acc;
    }
    fft256<true>(v, w, x, t);
    if (act) {
#pragma unroll
      for (int n = 0; n < 16; ++n) mid[s.mid_off + (long long)(t + 16 * n) * s.m2 + col] = v[n];
    }
  }
}

template <int COLS>
__global__ void __launch_bounds__(16 * COLS)
k_col16_forward(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
                const float2* __restrict__ mid, const float2* __restrict__ y, float2* __restrict__ pred,
                float2* __restrict__ msg, const double* __restrict__ sc, double* __restrict__ part,
                int pstride, int first_message, float damping, const float2* __restrict__ twtab) {
  __shared__ float2 xs[COLS][XS];
  __shared__ double red[8];
  const int c = threadIdx.x % COLS, t = threadIdx.x / COLS;
  float2* x = xs[c];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const FbmeWork wk = work[wi];
    const FbmeScene s = scenes[wk.scene];
    const bool act = c < wk.count;
    const int col = wk.first + c;
    float2 w[16];
    lane_twiddles(w, twtab + s.tw1_off, t);
    float2 v[16], yv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int i = t + 16 * j;
      v[j] = act ? mid[s.mid_off + (long long)i * s.m2 + col] : make_float2(0.f, 0.f);
      yv[j] = (act && i < s.m1) ? y[s.samp_off + (long long)i * s.m2 + col] : make_float2(0.f, 0.f);
    }
    fft256<false>(v, w, x, t);
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float precision = 1.0f / (dv + (float)scs[SC_NOISE]);
    double acc = 0.0;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int i = t + 16 * n;
      if (act && i < s.m1) {
        const long long gi = s.samp_off + (long long)i * s.m2 + col;
        pred[gi] = v[n];
        const float rx = yv[n].x - v[n].x, ry = yv[n].y - v[n].y;
        acc += (double)rx * rx + (double)ry * ry;
        if (first_message) {
          // First message from zero state (triton_kernels.py:66-77 with
          // prediction = message = 0).
          msg[gi] = make_float2((1.0f - damping) * (precision * yv[n].x),
                                (1.0f - damping) * (precision * yv[n].y));
        }
      }
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[(long long)PT_RES * pstride + s.col_part + wk.slot] = acc;
  }
}

}  // namespace

extern "C" {

int fbme_launch_row16(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, void* mid, void* bp,
                      float* gamma, void* mean, float* var, const double* sc, double* part, int pstride, int mode,
                      float damping, float prior_damping, float max_precision, float hyper_shape, float hyper_rate,
                      float blend, uint32_t* cand_key, int* cand_idx, int store_final, const void* twtab,
                      hipStream_t stream) {
  if (nwork <= 0) return 0;
  UampArgs16 a{damping, prior_damping, max_precision, hyper_shape, hyper_rate, blend};
#define FBME_ROW16(M)                                                                                        \
  hipLaunchKernelGGL((k_row16<M>), dim3(grid), dim3(64), 0, stream, scenes, work, nwork, (float2*)mid,       \
                     (float2*)bp, gamma, (float2*)mean, var, sc, part, pstride, a, cand_key, cand_idx,        \
                     store_final, (const float2*)twtab)
  switch (mode) {
    case 0: FBME_ROW16(0); break;
    case 1: FBME_ROW16(1); break;
    case 2: FBME_ROW16(2); break;
    default: FBME_ROW16(3); break;
  }
#undef FBME_ROW16
  return (int)hipGetLastError();
}

int fbme_launch_col16_adjoint(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, const void* y,
                              const void* pred, void* msg, void* mid, const double* sc, double* part, int pstride,
                              int mode, int store_msg, float damping, const void* twtab, hipStream_t stream) {
  if (nwork <= 0) return 0;
  constexpr int C = FBME_R16_COLS;
  if (mode == 0)
    hipLaunchKernelGGL((k_col16_adjoint<C, 0>), dim3(grid), dim3(16 * C), 0, stream, scenes, work, nwork,
                       (const float2*)y, (const float2*)pred, (float2*)msg, (float2*)mid, sc, part, pstride,
                       store_msg, damping, (const float2*)twtab);
  else
    hipLaunchKernelGGL((k_col16_adjoint<C, 1>), dim3(grid), dim3(16 * C), 0, stream, scenes, work, nwork,
                       (const float2*)y, (const float2*)pred, (float2*)msg, (float2*)mid, sc, part, pstride,
                       store_msg, damping, (const float2*)twtab);
  return (int)hipGetLastError();
}

int fbme_launch_col16_forward(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, const void* mid,
                              const void* y, void* pred, void* msg, const double* sc, double* part, int pstride,
                              int first_message, float damping, const void* twtab, hipStream_t stream) {
  if (nwork <= 0) return 0;
  constexpr int C = FBME_R16_COLS;
  hipLaunchKernelGGL((k_col16_forward<C>), dim3(grid), dim3(16 * C), 0, stream, scenes, work, nwork,
                     (const float2*)mid, (const float2*)y, (float2*)pred, (float2*)msg, sc, part, pstride,
                     first_message, damping, (const float2*)twtab);
  return (int)hipGetLastError();
}

}  // extern "C"
