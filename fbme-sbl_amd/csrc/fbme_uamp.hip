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
// MODE 0: input is the data y (backprojection); partial sum |y|^2.
// MODE 1: input is the UAMP message update of (r, msg), r = y - A mu being
// the residual the forward column pass stored.
// With an observation mask, unobserved samples contribute zero data and a
// zero message (the restricted operator, operators.py:57-75).
// All global loads of the tile are issued before the first LDS store.
template <int EPT, int MODE>
__global__ void __launch_bounds__(FBME_BLOCK)
k_col_adjoint(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
              const float2* __restrict__ y, const unsigned char* __restrict__ mask,
              const float2* __restrict__ resid, float2* __restrict__ msg,
              float2* __restrict__ mid, const double* __restrict__ sc, double* __restrict__ part,
              int pstride, int store_msg, float damping, const float2* __restrict__ twtab) {
  __shared__ float2 buf[FBME_LDS_SLOTS(EPT * FBME_BLOCK)];
  __shared__ float2 tw[1024];
  __shared__ double red[8];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const FbmeWork wk = work[w];
    const FbmeScene s = scenes[wk.scene];
    const int len = s.n1, cols = wk.count, stride = len + 1;
    const int nin = s.m1 * cols;     // nonzero inputs of the tile
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float noise = (float)scs[SC_NOISE];
    const float precision = 1.0f / (dv + noise);
    const float keep = damping, update = 1.0f - damping;
    float2 yv[EPT], pv[EPT], mv[EPT];
    bool ob[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + q * FBME_BLOCK;
      if (e < nin) {
        const int i = e / cols, c = e - i * cols;
        const long long gi = s.samp_off + (long long)i * s.m2 + wk.first + c;
        ob[q] = !mask || mask[gi];
        if (MODE == 0) yv[q] = y[gi];
        if (MODE == 1) {
          pv[q] = resid[gi];
          mv[q] = msg[gi];
        }
      }
    }
    load_twiddles(tw, twtab + s.tw1_off, len);
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + q * FBME_BLOCK;
      if (e < nin) {
        const int i = e / cols, c = e - i * cols;
        float2 v;
        if (MODE == 0) {
          v = ob[q] ? yv[q] : make_float2(0.f, 0.f);
          acc += (double)v.x * v.x + (double)v.y * v.y;
        } else {
          // y - (pred - dv msg) from the stored residual y - pred
          const float2 cand = make_float2(precision * (pv[q].x + dv * mv[q].x), precision * (pv[q].y + dv * mv[q].y));
          v = ob[q] ? make_float2(keep * mv[q].x + update * cand.x, keep * mv[q].y + update * cand.y)
                    : make_float2(0.f, 0.f);
          if (store_msg) msg[s.samp_off + (long long)i * s.m2 + wk.first + c] = v;
        }
        buf[lds_idx(c * stride + i)] = v;
      }
    }
    // zero padding rows M1..N1-1
    for (int e = nin + threadIdx.x; e < len * cols; e += FBME_BLOCK) {
      const int i = e / cols, c = e - i * cols;
      buf[lds_idx(c * stride + i)] = make_float2(0.f, 0.f);
    }
    if (MODE == 0) {
      acc = block_sum(acc, red);
      if (threadIdx.x == 0) part[(long long)PT_YE * pstride + s.col_part + wk.slot] = acc;
    }
    __syncthreads();
    fft_lines<EPT, true>(buf, len, cols, stride, s.rad1, tw);
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + q * FBME_BLOCK;
      if (e < len * cols) {
        const int i = e / cols, c = e - i * cols;
        mid[s.mid_off + (long long)i * s.m2 + wk.first + c] = buf[lds_idx(c * stride + i)];
      }
    }
    __syncthreads();
  }
}

// -------------------------------------------------------------------- row pass
// MODE 0: backprojection rows (A^H y).
// MODE 1: first UAMP step: lifted = bp * (1-d)/(dv+sigma^2), gamma0 built from
//         the backprojection (solver.py:1063-1099), mean_old = 0, v_old = gamma0.
// MODE 2: UAMP step: lifted = row pass of A^H of the column output.
// MODE 3: final E-step: no ARD update, no prediction; writes the polish score
//         |mu|^2 + v as the top-k candidates of this work item.
// TRACK (modes 1, 2): also the partial sums of |mu_t - mu_t-1|^2 and |mu_t|^2.
// MODE 3 without cand_key: the final E-step of a plan without a polish (no
// candidate selection).
template <int EPT, int MODE, bool TRACK>
__global__ void __launch_bounds__(FBME_BLOCK)
k_row(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
      float2* __restrict__ mid, float2* __restrict__ bp, float* __restrict__ gamma,
      float2* __restrict__ mean, float* __restrict__ var, const double* __restrict__ sc,
      double* __restrict__ part, int pstride, UampArgs a, uint32_t* __restrict__ cand_key,
      int* __restrict__ cand_idx, int store_final, const float2* __restrict__ twtab) {
  constexpr int TILE = EPT * FBME_BLOCK;
  __shared__ float2 buf[FBME_LDS_SLOTS(TILE)];
  __shared__ float2 tw[1024];
  __shared__ double red[8];
  __shared__ uint32_t keys[MODE == 3 ? TILE : 1];
  __shared__ uint32_t hist[MODE == 3 ? 256 : 1];
  __shared__ int pos[FBME_KMAX];
  __shared__ int misc[4];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const FbmeWork wk = work[w];
    const FbmeScene s = scenes[wk.scene];
    const int len = s.n2, rows = wk.count, stride = len + 1;
    const int nel = rows * len;
    const long long n_grid = (long long)s.n1 * s.n2;
    const int n_samp = s.m_obs;
    const long long gbase = s.grid_off + (long long)wk.first * len;
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float noise = (float)scs[SC_NOISE];
    const float floor_v = (float)scs[SC_FLOOR];
    const float power = (float)scs[SC_POWER];
    // Prefetch the per-element state of this tile.
    float gv[EPT], vv[EPT];
    float2 mv[EPT];
    if (MODE >= 1) {
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * FBME_BLOCK;
        if (e < nel) {
          if (MODE == 1) {
            mv[q] = bp[gbase + e];
          } else {
            gv[q] = gamma[gbase + e];
            mv[q] = mean[gbase + e];
            vv[q] = var[gbase + e];
          }
        }
      }
    }
    if (MODE != 1) {
      // Row pass of A^H: zero-padded inverse FFT of the M2 leading entries.
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * FBME_BLOCK;
        if (e < nel) {
          const int r = e / len, j = e - r * len;
          float2 v = make_float2(0.f, 0.f);
          if (j < s.m2) v = mid[s.mid_off + (long long)(wk.first + r) * s.m2 + j];
          buf[lds_idx(r * stride + j)] = v;
        }
      }
    }
    load_twiddles(tw, twtab + s.tw2_off, len);
    __syncthreads();
    if (MODE != 1) fft_lines<EPT, true>(buf, len, rows, stride, s.rad2, tw);
    if (MODE == 0) {
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * FBME_BLOCK;
        if (e < nel) {
          const int r = e / len, j = e - r * len;
          bp[gbase + e] = buf[lds_idx(r * stride + j)];
        }
      }
      __syncthreads();
      continue;
    }
    const float tau = fmaxf((dv + noise) / (float)n_samp, F32_TINY);
    const float keep = a.damping, update = 1.0f - a.damping;
    const float first_scale = update / (dv + noise);
    const float bpe_clamped = fmaxf((float)scs[SC_BPE], F32_EPS);
    const float uniform = power / (float)n_grid;
    double dof_acc = 0.0, vs_acc = 0.0, dmu_acc = 0.0, mu_acc = 0.0;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + q * FBME_BLOCK;
      if (e < nel) {
        const int r = e / len, j = e - r * len;
        float2 lifted, mean_old;
        float prior, v_old;
        if (MODE == 1) {
          const float2 b = mv[q];
          lifted = cscale(b, first_scale);
          const float energy = b.x * b.x + b.y * b.y;
          const float structured = energy / bpe_clamped * power;
          prior = fmaxf(a.blend * structured + (1.0f - a.blend) * uniform, floor_v);
          mean_old = make_float2(0.f, 0.f);
          v_old = fmaxf(prior, F32_TINY);
        } else {
          lifted = buf[lds_idx(r * stride + j)];
          prior = gv[q];
          mean_old = mv[q];
          v_old = fmaxf(vv[q], F32_TINY);
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
        if (TRACK) {
          const float dx = mu.x - mean_old.x, dy = mu.y - mean_old.y;
          dmu_acc += (double)(dx * dx + dy * dy);
          mu_acc += (double)(mu.x * mu.x + mu.y * mu.y);
        }
        if (MODE == 3) {
          keys[e] = score_key(mu.x * mu.x + mu.y * mu.y + v);
          if (store_final) {
            mean[gbase + e] = mu;
            var[gbase + e] = v;
          }
        } else {
          // Fused ARD update (triton_kernels.py:159-180).
          const float second = mu.x * mu.x + mu.y * mu.y + v;
          float cp = (second + a.hyper_rate) / (1.0f + a.hyper_shape);
          cp = fmaxf(floor_v, fminf(a.max_precision, cp));
          const float nxt = a.prior_damping * prior + (1.0f - a.prior_damping) * cp;
          gamma[gbase + e] = fmaxf(nxt, floor_v);
          mean[gbase + e] = mu;
          var[gbase + e] = v;
          buf[lds_idx(r * stride + j)] = mu;
        }
      }
    }
    block_sum2(dof_acc, vs_acc, red);
    if (threadIdx.x == 0) {
      part[(long long)PT_DOF * pstride + s.row_part + wk.slot] = dof_acc;
      part[(long long)PT_VS * pstride + s.row_part + wk.slot] = vs_acc;
    }
    if (TRACK) {
      block_sum2(dmu_acc, mu_acc, red);
      if (threadIdx.x == 0) {
        part[(long long)PT_DMU * pstride + s.row_part + wk.slot] = dmu_acc;
        part[(long long)PT_MU * pstride + s.row_part + wk.slot] = mu_acc;
      }
    }
    if (MODE == 3) {
      if (!cand_key) {   // uniform: a kernel argument
        __syncthreads();
        continue;
      }
      // Top-k candidates of this tile for the polish selection (solver.py:1926-1933).
      __syncthreads();
      const int got = block_select_topk(keys, nel, s.k, pos, hist, misc);
      write_sorted_candidates(keys, pos, got, s.k, wk.first * len,
                              cand_key + (long long)s.cand_off + (long long)wk.slot * s.k,
                              cand_idx + (long long)s.cand_off + (long long)wk.slot * s.k);
      __syncthreads();
      continue;
    }
    // Row pass of A mu keeping the M2 leading outputs.
    fft_lines<EPT, false>(buf, len, rows, stride, s.rad2, tw);
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + q * FBME_BLOCK;
      if (e < rows * s.m2) {
        const int r = e / s.m2, j = e - r * s.m2;
        mid[s.mid_off + (long long)(wk.first + r) * s.m2 + j] = buf[lds_idx(r * stride + j)];
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- column forward
template <int EPT>
__global__ void __launch_bounds__(FBME_BLOCK)
k_col_forward(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
              const float2* __restrict__ mid, const float2* __restrict__ y,
              const unsigned char* __restrict__ mask, float2* __restrict__ resid,
              float2* __restrict__ msg, const double* __restrict__ sc, double* __restrict__ part,
              int pstride, int first_message, float damping, const float2* __restrict__ twtab) {
  __shared__ float2 buf[FBME_LDS_SLOTS(EPT * FBME_BLOCK)];
  __shared__ float2 tw[1024];
  __shared__ double red[8];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const FbmeWork wk = work[w];
    const FbmeScene s = scenes[wk.scene];
    const int len = s.n1, cols = wk.count, stride = len + 1;
    const int nout = s.m1 * cols;
    float2 yv[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + q * FBME_BLOCK;
      if (e < len * cols) {
        const int i = e / cols, c = e - i * cols;
        buf[lds_idx(c * stride + i)] = mid[s.mid_off + (long long)i * s.m2 + wk.first + c];
      }
      if (e < nout) {
        const int i = e / cols, c = e - i * cols;
        yv[q] = y[s.samp_off + (long long)i * s.m2 + wk.first + c];
      }
    }
    load_twiddles(tw, twtab + s.tw1_off, len);
    __syncthreads();
    fft_lines<EPT, false>(buf, len, cols, stride, s.rad1, tw);
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float precision = 1.0f / (dv + (float)scs[SC_NOISE]);
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + q * FBME_BLOCK;
      if (e < nout) {
        const int i = e / cols, c = e - i * cols;
        const long long gi = s.samp_off + (long long)i * s.m2 + wk.first + c;
        const float2 p = buf[lds_idx(c * stride + i)];
        const bool ob = !mask || mask[gi];
        const float rx = yv[q].x - p.x, ry = yv[q].y - p.y;
        resid[gi] = make_float2(rx, ry);
        if (ob) acc += (double)rx * rx + (double)ry * ry;
        if (first_message) {
          // First message from zero state (triton_kernels.py:66-77 with
          // prediction = message = 0).
          msg[gi] = ob ? make_float2((1.0f - damping) * (precision * yv[q].x),
                                     (1.0f - damping) * (precision * yv[q].y))
                       : make_float2(0.f, 0.f);
        }
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
  float fixed_noise;   // > 0: fixed_noise_variance
};

// mode 0: initialisation (solver.py:1063-1189): power, floor, sigma^2_0,
//         sum |A^H y|^2 and sum gamma0 (closed form, the floor clamp is
//         inactive because gamma0 >= 0.9 P / N > max(1e-12 P, tiny)).
//         A^H y is the unnormalised 2-D inverse DFT of the zero-filled
//         samples, so by Parseval sum |A^H y|^2 = N1 N2 sum |y|^2 (masked
//         samples are zero); no pass over the backprojection is needed.
// mode 1: end of a UAMP iteration: dof clamp, noise EM/evidence update with
//         damping (triton_kernels.py:375-417, solver.py:1764-1798), data
//         variance = sum v for the next message update.  A fixed noise
//         variance starts unclamped and is max(fixed, floor) after every
//         update, undamped (solver.py:1151-1156, 1747-1750, 1788-1798).
// mode 2: mode 1 and the relative change ||mu_t - mu_t-1|| / max(||mu_t||,
//         eps) of the step (solver.py:1815-1825) from the tracked row pass.
// mode 3: the final E-step's clamped dof only (the UAMP posterior without a
//         polish, solver.py:1907-1925, 2081-2087).
__global__ void __launch_bounds__(64)
k_scalars(const FbmeScene* __restrict__ scenes, int nscenes, double* __restrict__ sc,
          const double* __restrict__ part, int pstride, int mode, const unsigned char* __restrict__ active,
          ScalarArgs a) {
  const int si = blockIdx.x;
  if (si >= nscenes) return;
  if (active && !active[si]) return;
  const FbmeScene s = scenes[si];
  double* scs = sc + (long long)si * SC_COUNT;
  const int n_samp = s.m_obs;
  const double n_grid = (double)s.n1 * s.n2;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, q4 = 0.0;
  if (mode == 0) {
    for (int i = threadIdx.x; i < s.col_items; i += 64) q0 += part_get(part, pstride, PT_YE, s.col_part + i);
  } else {
    for (int i = threadIdx.x; i < s.row_items; i += 64) {
      q0 += part_get(part, pstride, PT_DOF, s.row_part + i);
      q1 += part_get(part, pstride, PT_VS, s.row_part + i);
      if (mode == 2) {
        q3 += part_get(part, pstride, PT_DMU, s.row_part + i);
        q4 += part_get(part, pstride, PT_MU, s.row_part + i);
      }
    }
    if (mode != 3)
      for (int i = threadIdx.x; i < s.col_items; i += 64) q2 += part_get(part, pstride, PT_RES, s.col_part + i);
  }
  q0 = wave_sum(q0);
  q1 = wave_sum(q1);
  q2 = wave_sum(q2);
  q3 = wave_sum(q3);
  q4 = wave_sum(q4);
  if (threadIdx.x != 0) return;
  if (mode == 3) {
    scs[SC_DOF] = (float)fmin(q0, (double)n_samp * (1.0 - 10.0 * (double)F32_EPS));
    return;
  }
  if (mode == 2) scs[SC_CHANGE] = sqrt(q3) / fmax(sqrt(q4), (double)F32_EPS);
  if (mode == 0) {
    const float power = fmaxf((float)(q0 / n_samp), F32_EPS);
    const float floor_v = fmaxf(power * a.min_variance, F32_TINY);
    const float bpe = (float)(n_grid * q0);
    const double ratio = (double)bpe / (double)fmaxf(bpe, F32_EPS);
    const float dv0 = (float)(a.blend * (double)power * ratio + (1.0 - a.blend) * (double)power);
    scs[SC_YE] = q0;
    scs[SC_POWER] = power;
    scs[SC_FLOOR] = floor_v;
    scs[SC_BPE] = bpe;
    scs[SC_NOISE] = a.fixed_noise > 0.f ? a.fixed_noise : fmaxf(a.noise_fraction * power, floor_v);
    scs[SC_DV] = dv0;
    scs[SC_DOF] = 0.0;
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
    scs[SC_NOISE] = a.fixed_noise > 0.f ? fmaxf(a.fixed_noise, (float)floor_v) : (float)updated;
    scs[SC_DOF] = (float)dof;
    scs[SC_DV] = (float)q1;
  }
}

}  // namespace

// ----------------------------------------------------------- host launchers
#define FBME_EPT_DISPATCH(ept, ...)            \
  switch (ept) {                                \
    case 4: { constexpr int E = 4; __VA_ARGS__; break; }   \
    case 8: { constexpr int E = 8; __VA_ARGS__; break; }   \
    default: { constexpr int E = 16; __VA_ARGS__; break; } \
  }

extern "C" {

int fbme_launch_row16(const FbmeScene*, const FbmeWork*, int, int, void*, void*, float*, void*, float*, const double*,
                      double*, int, int, float, float, float, float, float, float, uint32_t*, int*, int, uint32_t*,
                      const void*, hipStream_t);
int fbme_launch_col16_adjoint(const FbmeScene*, const FbmeWork*, int, int, const void*, const uint16_t*,
                              const void*, void*, void*, const double*, double*, int, int, int, float, const void*,
                              hipStream_t);
int fbme_launch_col16_forward(const FbmeScene*, const FbmeWork*, int, int, const void*, const void*,
                              const uint16_t*, void*, void*, const double*, double*, int, int, float,
                              const void*, hipStream_t);

int fbme_launch_col_adjoint(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, int ept,
                            const void* y, const unsigned char* mask, const void* mbits, const void* pred, void* msg,
                            void* mid, const double* sc, double* part, int pstride, int mode, int store_msg,
                            float damping, const void* twtab, hipStream_t stream) {
  if (nwork <= 0) return 0;
  if (ept == FBME_EPT_R16)
    return fbme_launch_col16_adjoint(scenes, work, nwork, grid, y, (const uint16_t*)mbits, pred, msg, mid, sc, part,
                                     pstride, mode, store_msg, damping, twtab, stream);
  FBME_EPT_DISPATCH(ept, {
    if (mode == 0)
      hipLaunchKernelGGL((k_col_adjoint<E, 0>), dim3(grid), dim3(FBME_BLOCK), 0, stream, scenes, work, nwork,
                         (const float2*)y, mask, (const float2*)pred, (float2*)msg, (float2*)mid, sc, part, pstride,
                         store_msg, damping, (const float2*)twtab);
    else
      hipLaunchKernelGGL((k_col_adjoint<E, 1>), dim3(grid), dim3(FBME_BLOCK), 0, stream, scenes, work, nwork,
                         (const float2*)y, mask, (const float2*)pred, (float2*)msg, (float2*)mid, sc, part, pstride,
                         store_msg, damping, (const float2*)twtab);
  })
  return (int)hipGetLastError();
}

int fbme_launch_row(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, int ept, void* mid, void* bp,
                    float* gamma, void* mean, float* var, const double* sc, double* part, int pstride,
                    int mode, float damping, float prior_damping, float max_precision, float hyper_shape,
                    float hyper_rate, float blend, uint32_t* cand_key, int* cand_idx, int store_final,
                    uint32_t* score, const void* twtab, hipStream_t stream) {
  if (nwork <= 0) return 0;
  if (ept == FBME_EPT_R16)
    return fbme_launch_row16(scenes, work, nwork, grid, mid, bp, gamma, mean, var, sc, part, pstride, mode, damping,
                             prior_damping, max_precision, hyper_shape, hyper_rate, blend, cand_key, cand_idx,
                             store_final, score, twtab, stream);
  UampArgs a{damping, prior_damping, max_precision, hyper_shape, hyper_rate, blend};
#define FBME_ROW(M, T)                                                                                   \
  hipLaunchKernelGGL((k_row<E, M, T>), dim3(grid), dim3(FBME_BLOCK), 0, stream, scenes, work, nwork,    \
                     (float2*)mid, (float2*)bp, gamma, (float2*)mean, var, sc, part, pstride, a, cand_key, \
                     cand_idx, store_final, (const float2*)twtab)
  // modes 4 and 5: modes 1 and 2 with the change partials
  FBME_EPT_DISPATCH(ept, {
    switch (mode) {
      case 0: FBME_ROW(0, false); break;
      case 1: FBME_ROW(1, false); break;
      case 2: FBME_ROW(2, false); break;
      case 4: FBME_ROW(1, true); break;
      case 5: FBME_ROW(2, true); break;
      default: FBME_ROW(3, false); break;
    }
  })
#undef FBME_ROW
  return (int)hipGetLastError();
}

int fbme_launch_col_forward(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, int ept,
                            const void* mid, const void* y, const unsigned char* mask, const void* mbits, void* pred,
                            void* msg, const double* sc, double* part, int pstride, int first_message, float damping,
                            const void* twtab, hipStream_t stream) {
  if (nwork <= 0) return 0;
  if (ept == FBME_EPT_R16)
    return fbme_launch_col16_forward(scenes, work, nwork, grid, mid, y, (const uint16_t*)mbits, pred, msg, sc, part,
                                     pstride, first_message, damping, twtab, stream);
  FBME_EPT_DISPATCH(ept, {
    hipLaunchKernelGGL((k_col_forward<E>), dim3(grid), dim3(FBME_BLOCK), 0, stream, scenes, work, nwork,
                       (const float2*)mid, (const float2*)y, mask, (float2*)pred, (float2*)msg, sc, part, pstride,
                       first_message, damping, (const float2*)twtab);
  })
  return (int)hipGetLastError();
}

int fbme_launch_scalars(const FbmeScene* scenes, int nscenes, double* sc, const double* part, int pstride,
                        int mode, const unsigned char* active, float min_variance, float noise_fraction,
                        float blend, float damping, float noise_shape, float noise_rate, int evidence_noise,
                        float fixed_noise, hipStream_t stream) {
  if (nscenes <= 0) return 0;
  ScalarArgs a{min_variance, noise_fraction, blend, damping, noise_shape, noise_rate, evidence_noise, fixed_noise};
  hipLaunchKernelGGL(k_scalars, dim3(nscenes), dim3(64), 0, stream, scenes, nscenes, sc, part, pstride, mode,
                     active, a);
  return (int)hipGetLastError();
}

}  // extern "C"
