// Fused FFT-UAMP passes of the FBME plans (gfx950).
//
// One UAMP iteration of the reference (torch_sbl/solver.py:2726-2870 with the
// Triton kernels triton_kernels.py:36-192, the noise update :266-417 and the
// loop body solver.py:1390-1826) is executed as three HBM passes plus a tiny
// per-scene scalar kernel:
//
//   col_adjoint : message update (triton_kernels.py:36-89) fused into the
//                 load of the column pass of A^H (operators.py:255-280);
//                 writes the new message and the [N1, M2] intermediate.
//   row         : row pass of A^H, denoise + damping + fused ARD update
//                 (triton_kernels.py:92-192) and the row pass of A mu
//                 (operators.py:237-253), keeping only M2 outputs.
//   col_forward : column pass of A mu keeping M1 outputs (the prediction) and
//                 the residual-energy partial sums (triton_kernels.py:321-354).
//   scalars     : noise EM update (triton_kernels.py:375-417), dof clamp
//                 (solver.py:1470-1474), data variance sum(v).
//
// All reductions are two-level and deterministic (double partials in fixed
// order); the reference's float atomics (triton_kernels.py:185-192) are not
// reproduced.
#include "fbme_device.h"

using namespace fbme;

namespace {

constexpr float F32_EPS = 1.1920928955078125e-07f;
constexpr float F32_TINY = 1.17549435082228750797e-38f;

struct UampArgs {
  float damping;        // UAMP damping (0.05)
  float prior_damping;  // ARD damping (0.15)
  float max_precision;  // 1e12
  float hyper_shape, hyper_rate;
  float blend;          // backprojection blend (0.1)
};

__device__ __forceinline__ double part_get(const double* part, int pstride, int q, int slot) {
  return part[(long long)q * pstride + slot];
}

// ------------------------------------------------------------- column adjoint
// mode 0: input is the data y (backprojection); partial sum |y|^2.
// mode 1: input is the UAMP message update of (y, pred, msg).
template <int EPT>
__global__ void __launch_bounds__(FBME_BLOCK)
k_col_adjoint(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
              const float2* __restrict__ y, const float2* __restrict__ pred, float2* __restrict__ msg,
              float2* __restrict__ mid, const double* __restrict__ sc, double* __restrict__ part,
              int pstride, int mode, int store_msg, float damping) {
  __shared__ float2 buf[FBME_TILE_MAX + 64];
  __shared__ float2 tw[1024];
  __shared__ double red[8];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const FbmeWork wk = work[w];
    const FbmeScene s = scenes[wk.scene];
    const int len = s.n1, cols = wk.count, stride = len + 1;
    build_twiddles(tw, len);
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float noise = (float)scs[SC_NOISE];
    const float precision = 1.0f / (dv + noise);
    const float keep = damping, update = 1.0f - damping;
    double acc = 0.0;
    for (int e = threadIdx.x; e < len * cols; e += FBME_BLOCK) {
      const int i = e / cols, c = e - i * cols;
      float2 v = make_float2(0.f, 0.f);
      if (i < s.m1) {
        const long long gi = s.samp_off + (long long)i * s.m2 + wk.first + c;
        const float2 yy = y[gi];
        if (mode == 0) {
          v = yy;
          acc += (double)yy.x * yy.x + (double)yy.y * yy.y;
        } else {
          const float2 po = pred[gi], mo = msg[gi];
          const float2 pseudo = make_float2(po.x - dv * mo.x, po.y - dv * mo.y);
          const float2 cand = make_float2(precision * (yy.x - pseudo.x), precision * (yy.y - pseudo.y));
          v = make_float2(keep * mo.x + update * cand.x, keep * mo.y + update * cand.y);
          if (store_msg) msg[gi] = v;
        }
      }
      buf[c * stride + i] = v;
    }
    if (mode == 0) {
      acc = block_sum(acc, red);
      if (threadIdx.x == 0) part[(long long)PT_YE * pstride + s.col_part + wk.slot] = acc;
    }
    __syncthreads();
    fft_lines<EPT, true>(buf, len, cols, stride, s.rad1, tw);
    for (int e = threadIdx.x; e < len * cols; e += FBME_BLOCK) {
      const int i = e / cols, c = e - i * cols;
      mid[s.mid_off + (long long)i * s.m2 + wk.first + c] = buf[c * stride + i];
    }
    __syncthreads();
  }
}

// -------------------------------------------------------------------- row pass
// mode 0: backprojection rows (A^H y), partial sum |A^H y|^2.
// mode 1: first UAMP step: lifted = bp * (1-d)/(dv+sigma^2), gamma0 built from
//         the backprojection (solver.py:1063-1099), mean_old = 0, v_old = gamma0.
// mode 2: UAMP step: lifted = row pass of A^H of the column output.
// mode 3: final E-step: no ARD update, no prediction; writes the polish score
//         |mu|^2 + v as the top-k candidates of this work item.
template <int EPT>
__global__ void __launch_bounds__(FBME_BLOCK)
k_row(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
      float2* __restrict__ mid, float2* __restrict__ bp, float* __restrict__ gamma,
      float2* __restrict__ mean, float* __restrict__ var, const double* __restrict__ sc,
      double* __restrict__ part, int pstride, int mode, UampArgs a, uint32_t* __restrict__ cand_key,
      int* __restrict__ cand_idx, int store_final) {
  __shared__ float2 buf[FBME_TILE_MAX + 64];
  __shared__ float2 tw[1024];
  __shared__ double red[8];
  __shared__ uint32_t keys[FBME_TILE_MAX];
  __shared__ uint32_t hist[256];
  __shared__ int pos[FBME_KMAX];
  __shared__ int misc[4];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const FbmeWork wk = work[w];
    const FbmeScene s = scenes[wk.scene];
    const int len = s.n2, rows = wk.count, stride = len + 1;
    const long long n_grid = (long long)s.n1 * s.n2;
    const int n_samp = s.m1 * s.m2;
    if (mode != 1) build_twiddles(tw, len);
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float noise = (float)scs[SC_NOISE];
    const float floor_v = (float)scs[SC_FLOOR];
    const float power = (float)scs[SC_POWER];
    // Row pass of A^H: zero-padded inverse FFT of the M2 leading entries.
    if (mode != 1) {
      for (int e = threadIdx.x; e < rows * len; e += FBME_BLOCK) {
        const int r = e / len, j = e - r * len;
        float2 v = make_float2(0.f, 0.f);
        if (j < s.m2) v = mid[s.mid_off + (long long)(wk.first + r) * s.m2 + j];
        buf[r * stride + j] = v;
      }
      __syncthreads();
      fft_lines<EPT, true>(buf, len, rows, stride, s.rad2, tw);
    }
    if (mode == 0) {
      double acc = 0.0;
      for (int e = threadIdx.x; e < rows * len; e += FBME_BLOCK) {
        const int r = e / len, j = e - r * len;
        const float2 v = buf[r * stride + j];
        bp[s.grid_off + (long long)(wk.first + r) * len + j] = v;
        acc += (double)v.x * v.x + (double)v.y * v.y;
      }
      acc = block_sum(acc, red);
      if (threadIdx.x == 0) part[(long long)PT_BPE * pstride + s.row_part + wk.slot] = acc;
      __syncthreads();
      continue;
    }
    const float tau = fmaxf((dv + noise) / (float)n_samp, F32_TINY);
    const float keep = a.damping, update = 1.0f - a.damping;
    // First step quantities (solver.py:1063-1099, 2775-2782).
    const float first_scale = update / (dv + noise);
    const float bpe_clamped = fmaxf((float)scs[SC_BPE], F32_EPS);
    const float uniform = power / (float)n_grid;
    double dof_acc = 0.0, vs_acc = 0.0;
    for (int e = threadIdx.x; e < rows * len; e += FBME_BLOCK) {
      const int r = e / len, j = e - r * len;
      const long long gi = s.grid_off + (long long)(wk.first + r) * len + j;
      float2 lifted, mean_old;
      float prior, v_old;
      if (mode == 1) {
        const float2 b = bp[gi];
        lifted = cscale(b, first_scale);
        const float energy = b.x * b.x + b.y * b.y;
        const float structured = energy / bpe_clamped * power;
        prior = fmaxf(a.blend * structured + (1.0f - a.blend) * uniform, floor_v);
        mean_old = make_float2(0.f, 0.f);
        v_old = fmaxf(prior, F32_TINY);
      } else {
        lifted = buf[r * stride + j];
        prior = gamma[gi];
        mean_old = mean[gi];
        v_old = fmaxf(var[gi], F32_TINY);
      }
      // Denoise (triton_kernels.py:131-158).
      const float denom = fmaxf(prior + tau, F32_TINY);
      const float gain = prior / denom;
      const float2 pm = make_float2(mean_old.x + tau * lifted.x, mean_old.y + tau * lifted.y);
      const float2 cm = make_float2(gain * pm.x, gain * pm.y);
      const float cvar = prior * tau / denom;
      const float2 mu = make_float2(keep * mean_old.x + update * cm.x, keep * mean_old.y + update * cm.y);
      const float v = fmaxf(keep * v_old + update * cvar, F32_TINY);
      float rel = 1.0f - v / fmaxf(prior, F32_TINY);
      rel = fmaxf(0.0f, fminf(1.0f, rel));
      dof_acc += (double)rel;
      vs_acc += (double)v;
      if (mode == 3) {
        const float score = mu.x * mu.x + mu.y * mu.y + v;
        keys[e] = score_key(score);
        if (store_final) {
          mean[gi] = mu;
          var[gi] = v;
        }
        continue;
      }
      // Fused ARD update (triton_kernels.py:159-180).
      const float second = mu.x * mu.x + mu.y * mu.y + v;
      float cp = (second + a.hyper_rate) / (1.0f + a.hyper_shape);
      cp = fmaxf(floor_v, fminf(a.max_precision, cp));
      const float nxt = a.prior_damping * prior + (1.0f - a.prior_damping) * cp;
      gamma[gi] = fmaxf(nxt, floor_v);
      mean[gi] = mu;
      var[gi] = v;
      buf[r * stride + j] = mu;
    }
    block_sum2(dof_acc, vs_acc, red);
    if (threadIdx.x == 0) {
      part[(long long)PT_DOF * pstride + s.row_part + wk.slot] = dof_acc;
      part[(long long)PT_VS * pstride + s.row_part + wk.slot] = vs_acc;
    }
    if (mode == 3) {
      // Top-k candidates of this tile for the polish selection (solver.py:1926-1933).
      __syncthreads();
      const int n_el = rows * len;
      const int got = block_select_topk(keys, n_el, s.k, pos, hist, misc);
      const long long cbase = (long long)s.cand_off + (long long)wk.slot * s.k;
      for (int i = threadIdx.x; i < s.k; i += FBME_BLOCK) {
        if (i < got) {
          cand_key[cbase + i] = keys[pos[i]];
          cand_idx[cbase + i] = wk.first * len + pos[i];
        } else {
          cand_key[cbase + i] = 0u;
          cand_idx[cbase + i] = -1;
        }
      }
      __syncthreads();
      continue;
    }
    if (mode == 1) {
      // The first step skipped the inverse row FFT, so the twiddles are built here.
      build_twiddles(tw, len);
    }
    __syncthreads();
    // Row pass of A mu keeping the M2 leading outputs.
    fft_lines<EPT, false>(buf, len, rows, stride, s.rad2, tw);
    for (int e = threadIdx.x; e < rows * s.m2; e += FBME_BLOCK) {
      const int r = e / s.m2, j = e - r * s.m2;
      mid[s.mid_off + (long long)(wk.first + r) * s.m2 + j] = buf[r * stride + j];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- column forward
template <int EPT>
__global__ void __launch_bounds__(FBME_BLOCK)
k_col_forward(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
              const float2* __restrict__ mid, const float2* __restrict__ y, float2* __restrict__ pred,
              float2* __restrict__ msg, const double* __restrict__ sc, double* __restrict__ part,
              int pstride, int first_message, float damping) {
  __shared__ float2 buf[FBME_TILE_MAX + 64];
  __shared__ float2 tw[1024];
  __shared__ double red[8];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const FbmeWork wk = work[w];
    const FbmeScene s = scenes[wk.scene];
    const int len = s.n1, cols = wk.count, stride = len + 1;
    build_twiddles(tw, len);
    for (int e = threadIdx.x; e < len * cols; e += FBME_BLOCK) {
      const int i = e / cols, c = e - i * cols;
      buf[c * stride + i] = mid[s.mid_off + (long long)i * s.m2 + wk.first + c];
    }
    __syncthreads();
    fft_lines<EPT, false>(buf, len, cols, stride, s.rad1, tw);
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float precision = 1.0f / (dv + (float)scs[SC_NOISE]);
    double acc = 0.0;
    for (int e = threadIdx.x; e < s.m1 * cols; e += FBME_BLOCK) {
      const int i = e / cols, c = e - i * cols;
      const long long gi = s.samp_off + (long long)i * s.m2 + wk.first + c;
      const float2 p = buf[c * stride + i];
      const float2 yy = y[gi];
      pred[gi] = p;
      const float rx = yy.x - p.x, ry = yy.y - p.y;
      acc += (double)rx * rx + (double)ry * ry;
      if (first_message) {
        // First message from zero state (triton_kernels.py:66-77 with
        // prediction = message = 0).
        msg[gi] = make_float2((1.0f - damping) * (precision * yy.x),
                              (1.0f - damping) * (precision * yy.y));
      }
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[(long long)PT_RES * pstride + s.col_part + wk.slot] = acc;
    __syncthreads();
  }
}

// --------------------------------------------------------------- scalar update
struct ScalarArgs {
  float min_variance, noise_fraction, blend, damping;
  float noise_shape, noise_rate;
  int evidence_noise;
};

// mode 0: initialisation (solver.py:1063-1189): power, floor, sigma^2_0,
//         sum |A^H y|^2 and sum gamma0 (closed form, the floor clamp is
//         inactive because gamma0 >= 0.9 P / N > max(1e-12 P, tiny)).
// mode 1: end of a UAMP iteration: dof clamp, noise EM/evidence update with
//         damping (triton_kernels.py:375-417, solver.py:1764-1798), data
//         variance = sum v for the next message update.
__global__ void __launch_bounds__(64)
k_scalars(const FbmeScene* __restrict__ scenes, int nscenes, double* __restrict__ sc,
          const double* __restrict__ part, int pstride, int mode, const unsigned char* __restrict__ active,
          ScalarArgs a) {
  const int si = blockIdx.x;
  if (si >= nscenes) return;
  if (active && !active[si]) return;
  const FbmeScene s = scenes[si];
  double* scs = sc + (long long)si * SC_COUNT;
  const int n_samp = s.m1 * s.m2;
  const double n_grid = (double)s.n1 * s.n2;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;
  if (mode == 0) {
    for (int i = threadIdx.x; i < s.col_items; i += 64) q0 += part_get(part, pstride, PT_YE, s.col_part + i);
    for (int i = threadIdx.x; i < s.row_items; i += 64) q1 += part_get(part, pstride, PT_BPE, s.row_part + i);
  } else {
    for (int i = threadIdx.x; i < s.row_items; i += 64) {
      q0 += part_get(part, pstride, PT_DOF, s.row_part + i);
      q1 += part_get(part, pstride, PT_VS, s.row_part + i);
    }
    for (int i = threadIdx.x; i < s.col_items; i += 64) q2 += part_get(part, pstride, PT_RES, s.col_part + i);
  }
  q0 = wave_sum(q0);
  q1 = wave_sum(q1);
  q2 = wave_sum(q2);
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    const float power = fmaxf((float)(q0 / n_samp), F32_EPS);
    const float floor_v = fmaxf(power * a.min_variance, F32_TINY);
    const float bpe = (float)q1;
    const double ratio = (double)bpe / (double)fmaxf(bpe, F32_EPS);
    const float dv0 = (float)(a.blend * (double)power * ratio + (1.0 - a.blend) * (double)power);
    scs[SC_YE] = q0;
    scs[SC_POWER] = power;
    scs[SC_FLOOR] = floor_v;
    scs[SC_BPE] = bpe;
    scs[SC_NOISE] = fmaxf(a.noise_fraction * power, floor_v);
    scs[SC_DV] = dv0;
    scs[SC_DOF] = 0.0;
    (void)n_grid;
  } else {
    const double max_dof = (double)n_samp * (1.0 - 10.0 * (double)F32_EPS);
    const double dof = fmin(q0, max_dof);
    const double energy = q2;
    const double noise = scs[SC_NOISE];
    const double floor_v = scs[SC_FLOOR];
    double cand;
    if (a.evidence_noise) {
      cand = (energy + a.noise_rate) / fmax((double)n_samp - dof + a.noise_shape, (double)F32_EPS);
    } else {
      cand = (energy + noise * dof + a.noise_rate) / ((double)n_samp + a.noise_shape);
    }
    cand = fmax(cand, floor_v);
    const double updated = fmax(a.damping * noise + (1.0 - a.damping) * cand, floor_v);
    scs[SC_NOISE] = (float)updated;
    scs[SC_DOF] = (float)dof;
    scs[SC_DV] = (float)q1;
  }
}

}  // namespace

// ----------------------------------------------------------- host launchers
extern "C" {

int fbme_launch_col_adjoint(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid,
                            const void* y, const void* pred, void* msg, void* mid, const double* sc,
                            double* part, int pstride, int mode, int store_msg, float damping,
                            hipStream_t stream) {
  if (nwork <= 0) return 0;
  hipLaunchKernelGGL(k_col_adjoint<16>, dim3(grid), dim3(FBME_BLOCK), 0, stream, scenes, work, nwork,
                     (const float2*)y, (const float2*)pred, (float2*)msg, (float2*)mid, sc, part, pstride,
                     mode, store_msg, damping);
  return (int)hipGetLastError();
}

int fbme_launch_row(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, void* mid, void* bp,
                    float* gamma, void* mean, float* var, const double* sc, double* part, int pstride,
                    int mode, float damping, float prior_damping, float max_precision, float hyper_shape,
                    float hyper_rate, float blend, uint32_t* cand_key, int* cand_idx, int store_final,
                    hipStream_t stream) {
  if (nwork <= 0) return 0;
  UampArgs a{damping, prior_damping, max_precision, hyper_shape, hyper_rate, blend};
  hipLaunchKernelGGL(k_row<16>, dim3(grid), dim3(FBME_BLOCK), 0, stream, scenes, work, nwork, (float2*)mid,
                     (float2*)bp, gamma, (float2*)mean, var, sc, part, pstride, mode, a, cand_key, cand_idx,
                     store_final);
  return (int)hipGetLastError();
}

int fbme_launch_col_forward(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid,
                            const void* mid, const void* y, void* pred, void* msg, const double* sc,
                            double* part, int pstride, int first_message, float damping,
                            hipStream_t stream) {
  if (nwork <= 0) return 0;
  hipLaunchKernelGGL(k_col_forward<16>, dim3(grid), dim3(FBME_BLOCK), 0, stream, scenes, work, nwork,
                     (const float2*)mid, (const float2*)y, (float2*)pred, (float2*)msg, sc, part, pstride,
                     first_message, damping);
  return (int)hipGetLastError();
}

int fbme_launch_scalars(const FbmeScene* scenes, int nscenes, double* sc, const double* part, int pstride,
                        int mode, const unsigned char* active, float min_variance, float noise_fraction,
                        float blend, float damping, float noise_shape, float noise_rate, int evidence_noise,
                        hipStream_t stream) {
  if (nscenes <= 0) return 0;
  ScalarArgs a{min_variance, noise_fraction, blend, damping, noise_shape, noise_rate, evidence_noise};
  hipLaunchKernelGGL(k_scalars, dim3(nscenes), dim3(64), 0, stream, scenes, nscenes, sc, part, pstride, mode,
                     active, a);
  return (int)hipGetLastError();
}

}  // extern "C"
