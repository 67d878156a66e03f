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
#include "fbme_fft16.h"

using namespace fbme;

namespace {

constexpr float F32_EPS = 1.1920928955078125e-07f;
constexpr float F32_TINY = 1.17549435082228750797e-38f;
constexpr int XS = 16 * 17 + 8;   // LDS slots of one 256-point transpose (padded)

struct UampArgs16 {
  float damping, prior_damping, max_precision, hyper_shape, hyper_rate, blend;
};

// Top-k (k <= 32) of the keys a wavefront holds 16 per lane (lane (g, t)
// holds tile position g * 256 + 16 n + t; key 0 marks an empty slot),
// written sorted by (key desc, position asc) to out_key / out_idx with
// `base` added to positions.
//  1. The k largest lane maxima are k distinct keys >= L, the k-th largest
//     lane maximum, so every top-k key is >= L.
//  2. While more than 64 keys survive, bisection raises the threshold lo
//     keeping count(key >= lo) >= k.
//  3. At most `cap` survivors are ranked directly.  Otherwise the bisection
//     ended on the exact k-th largest key T with massive ties: keys > T and
//     the lowest positions among keys == T are taken.
// Called by every lane of a single-wavefront workgroup.
__device__ void wave_tile_topk(const uint32_t (&kv)[16], int g, int t, int k, int base, uint32_t* sk, int* sp,
                               int cap, uint32_t* lmax, uint32_t* out_key, int* out_idx) {
  const int lane = threadIdx.x & 63;
  uint32_t v = 0u;
#pragma unroll
  for (int n = 0; n < 16; ++n) v = max(v, kv[n]);
  lmax[lane] = v;
  __syncthreads();
  int rank = 0;
#pragma unroll 16
  for (int j = 0; j < 64; ++j) {
    const uint32_t u = lmax[j];
    rank += (u > v) || (u == v && j < lane);
  }
  __syncthreads();
  if (rank == k - 1) lmax[0] = v;
  __syncthreads();
  uint32_t lo = lmax[0];
  __syncthreads();
  auto count_ge = [&](uint32_t th) {
    int cnt = 0;
#pragma unroll
    for (int n = 0; n < 16; ++n) cnt += __popcll(__ballot(kv[n] != 0u && kv[n] >= th));
    return cnt;
  };
  unsigned long long hi = (unsigned long long)wave_max_u64((unsigned long long)v) + 1ull;  // count(>= hi) = 0
  int cnt = count_ge(lo);
  while (cnt > 64 && hi > (unsigned long long)lo + 1ull) {
    const uint32_t mid = (uint32_t)((unsigned long long)lo + (hi - (unsigned long long)lo) / 2ull);
    const int c2 = count_ge(mid);
    if (c2 >= k) {
      lo = mid;
      cnt = c2;
    } else {
      hi = mid;
    }
  }
  const unsigned long long below = __lanemask_lt();
  int total = 0;
  if (cnt <= cap) {
    // all survivors (keys >= lo), compacted in (n, lane) order
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const bool keep = kv[n] != 0u && kv[n] >= lo;
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const int o = total + __popcll(m & below);
        sk[o] = kv[n];
        sp[o] = g * 256 + 16 * n + t;
      }
      total += __popcll(m);
    }
  } else {
    // lo is the k-th largest key T: take keys > T, then the lowest
    // positions among keys == T (position order is (g, n, t)).
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const bool keep = kv[n] > lo;
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const int o = total + __popcll(m & below);
        sk[o] = kv[n];
        sp[o] = g * 256 + 16 * n + t;
      }
      total += __popcll(m);
    }
    for (int gg = 0; gg < 4 && total < k; ++gg) {
#pragma unroll
      for (int n = 0; n < 16; ++n) {
        if (total >= k) break;
        const bool tie = g == gg && kv[n] == lo;
        const unsigned long long m = __ballot(tie);
        if (tie) {
          const int o = total + __popcll(m & below);
          if (o < k) {
            sk[o] = kv[n];
            sp[o] = g * 256 + 16 * n + t;
          }
        }
        total = min(k, total + __popcll(m));
      }
    }
  }
  __syncthreads();
  for (int i = lane; i < total; i += 64) {
    const uint32_t ki = sk[i];
    const int pi = sp[i];
    int r = 0;
    for (int j = 0; j < total; ++j) {
      const uint32_t kj = sk[j];
      r += (kj > ki) || (kj == ki && sp[j] < pi);
    }
    if (r < k) {
      out_key[r] = ki;
      out_idx[r] = base + pi;
    }
  }
  for (int i = total + lane; i < k; i += 64) {
    out_key[i] = 0u;
    out_idx[i] = -1;
  }
}

// ------------------------------------------------------------------- row pass
// Modes as k_row in fbme_uamp.hip: 0 backprojection rows, 1 first UAMP step,
// 2 UAMP step, 3 final E-step with top-k candidates.  Mode 1 transforms the
// column output itself and stores the backprojection A^H y (read by the
// polish) on the way, so the plans launch no separate mode-0 pass for
// register-FFT rows.
// TRACK (modes 1, 2): also the partial sums of |mu_t - mu_t-1|^2 and |mu_t|^2
// for the relative change of the step (eager convergence checks).
template <int MODE, bool TRACK>
__global__ void __launch_bounds__(64)
k_row16(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
        float2* __restrict__ mid, float2* __restrict__ bp, float* __restrict__ gamma,
        float2* __restrict__ mean, float* __restrict__ var, const double* __restrict__ sc,
        double* __restrict__ part, int pstride, UampArgs16 a, int store_final, uint32_t* __restrict__ score,
        const float2* __restrict__ twtab) {
  __shared__ float2 xs[4][XS];
  __shared__ float2 tw[256];
  __shared__ double red[8];
  const int lane = threadIdx.x, g = lane >> 4, t = lane & 15;
  float2* x = xs[g];
  int tw_loaded = -1;   // twiddle table in LDS (all 256-point axes share one)
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const FbmeWork wk = work[wi];
    const FbmeScene s = scenes[wk.scene];
    const bool act = g < wk.count;
    const int row = wk.first + g;
    const long long gb = s.grid_off + (long long)row * 256;
    const long long mb = s.mid_off + (long long)row * s.m2;
    const long long n_grid = (long long)s.n1 * 256;
    const int n_samp = s.m_obs;
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float noise = (float)scs[SC_NOISE];
    const float floor_v = (float)scs[SC_FLOOR];
    const float power = (float)scs[SC_POWER];
    if (s.tw2_off != tw_loaded) {   // uniform: every lane reads the same scene
      load_tw256(tw, twtab + s.tw2_off);
      tw_loaded = s.tw2_off;
    }
    float2 v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int c = t + 16 * j;
      v[j] = (act && c < s.m2) ? mid[mb + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft256<true>(v, tw, x, t);
    if (MODE <= 1) {
#pragma unroll
      for (int n = 0; n < 16; ++n)
        if (act) bp[gb + t + 16 * n] = v[n];
    }
    if (MODE == 0) {
      __syncthreads();
      continue;
    }
    const float tau = fmaxf((dv + noise) / (float)n_samp, F32_TINY);
    const float keep = a.damping, update = 1.0f - a.damping;
    const float first_scale = update / (dv + noise);
    const float bpe_clamped = fmaxf((float)scs[SC_BPE], F32_EPS);
    const float uniform = power / (float)n_grid;
    const float struct_scale = power / bpe_clamped;
    const float ard_scale = 1.0f / (1.0f + a.hyper_shape);
    double dof_acc = 0.0, vs_acc = 0.0, dmu_acc = 0.0, mu_acc = 0.0;
    // The per-element state is loaded after the inverse transform, in two
    // halves, so that little of it is live at once (fewer registers, more
    // resident waves to cover the HBM latency).
#pragma unroll
    for (int h = 0; h < 16; h += 8) {
      float gv[8], vv[8];
      float2 mv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const long long k = gb + t + 16 * (h + q);
        if (MODE != 1) {
          gv[q] = act ? gamma[k] : 1.0f;
          mv[q] = act ? mean[k] : make_float2(0.f, 0.f);
          vv[q] = act ? var[k] : 1.0f;
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int n = h + q;
        const long long k = gb + t + 16 * n;
        float2 lifted, mean_old;
        float prior, v_old;
        if (MODE == 1) {
          const float2 b = v[n];
          lifted = cscale(b, first_scale);
          const float energy = b.x * b.x + b.y * b.y;
          const float structured = energy * struct_scale;
          prior = fmaxf(a.blend * structured + (1.0f - a.blend) * uniform, floor_v);
          mean_old = make_float2(0.f, 0.f);
          v_old = fmaxf(prior, F32_TINY);
        } else {
          lifted = v[n];
          prior = gv[q];
          mean_old = mv[q];
          v_old = fmaxf(vv[q], F32_TINY);
        }
        // Denoise (triton_kernels.py:131-158).  Quotients use the hardware
        // reciprocal (1 ulp), the precision class of the reference's Triton
        // float division.
        const float denom = fmaxf(prior + tau, F32_TINY);
        const float rden = __builtin_amdgcn_rcpf(denom);
        const float gain = prior * rden;
        const float2 pm = make_float2(mean_old.x + tau * lifted.x, mean_old.y + tau * lifted.y);
        const float2 cm = make_float2(gain * pm.x, gain * pm.y);
        const float cvar = prior * tau * rden;
        const float2 mu = make_float2(keep * mean_old.x + update * cm.x, keep * mean_old.y + update * cm.y);
        const float vn = fmaxf(keep * v_old + update * cvar, F32_TINY);
        float rel = 1.0f - vn * __builtin_amdgcn_rcpf(fmaxf(prior, F32_TINY));
        rel = fmaxf(0.0f, fminf(1.0f, rel));
        if (act) {
          dof_acc += (double)rel;
          vs_acc += (double)vn;
          if (TRACK) {
            const float dx = mu.x - mean_old.x, dy = mu.y - mean_old.y;
            dmu_acc += (double)(dx * dx + dy * dy);
            mu_acc += (double)(mu.x * mu.x + mu.y * mu.y);
          }
        }
        if (MODE == 3) {
          // polish score |mu|^2 + v as an order-preserving key; the tile's
          // top-k runs in k_tile_topk16 (a latency-bound select would hold
          // this bandwidth-bound pass's registers and waves).  No score
          // buffer: the plan has no polish.
          if (score && act) score[k] = score_key(mu.x * mu.x + mu.y * mu.y + vn);
          if (store_final && act) {
            mean[k] = mu;
            var[k] = vn;
          }
        } else {
          // Fused ARD update (triton_kernels.py:159-180).
          const float second = mu.x * mu.x + mu.y * mu.y + vn;
          float cp = (second + a.hyper_rate) * ard_scale;
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
      __builtin_amdgcn_sched_barrier(0);
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
      __syncthreads();
      continue;
    }
    // Row pass of A mu keeping the M2 leading outputs.
    fft256<false>(v, tw, x, t);
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int c = t + 16 * n;
      if (act && c < s.m2) mid[mb + c] = v[n];
    }
    __syncthreads();
  }
}

// Top-k candidates of each four-row tile for the polish selection
// (solver.py:1926-1933): score |mu|^2 + v written by k_row16<3>, sorted by
// (score desc, grid index asc), one wavefront per tile.
__global__ void __launch_bounds__(64)
k_tile_topk16(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
              const uint32_t* __restrict__ score, uint32_t* __restrict__ cand_key, int* __restrict__ cand_idx) {
  __shared__ uint32_t surv_key[64];
  __shared__ int surv_pos[64];
  __shared__ uint32_t lane_max[64];
  const int lane = threadIdx.x, g = lane >> 4, t = lane & 15;
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const FbmeWork wk = work[wi];
    const FbmeScene s = scenes[wk.scene];
    const bool act = g < wk.count;
    const long long gb = s.grid_off + (long long)(wk.first + g) * 256;
    uint32_t kv[16];
#pragma unroll
    for (int n = 0; n < 16; ++n) kv[n] = act ? score[gb + t + 16 * n] : 0u;
    uint32_t* ck = cand_key + (long long)s.cand_off + (long long)wk.slot * s.k;
    int* ci = cand_idx + (long long)s.cand_off + (long long)wk.slot * s.k;
    wave_tile_topk(kv, g, t, s.k, wk.first * 256, surv_key, surv_pos, 64, lane_max, ck, ci);
    __syncthreads();
  }
}

// ------------------------------------------------------------- column passes
// COLS columns per workgroup, 16 lanes per column: lane = c + COLS * t.
// With an observation mask, unobserved samples contribute zero data, a zero
// message and no residual.  The mask reaches these kernels as 16-bit words
// (bit j of word [col][t] = sample (t + 16 j, col) observed), one load per
// lane and work item.
template <int COLS, int MODE>
__global__ void __launch_bounds__(16 * COLS)
k_col16_adjoint(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
                const float2* __restrict__ y, const uint16_t* __restrict__ mbits,
                const float2* __restrict__ resid, float2* __restrict__ msg,
                float2* __restrict__ mid, const double* __restrict__ sc, double* __restrict__ part,
                int pstride, int store_msg, float damping, const float2* __restrict__ twtab) {
  __shared__ float2 xs[16 * 17 * COLS];   // column c's element i at xs[i * COLS + c]
  __shared__ float2 tw[256];
  __shared__ double red[8];
  const int c = threadIdx.x % COLS, t = threadIdx.x / COLS;
  float2* x = xs + c;
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const FbmeWork wk = work[wi];
    const FbmeScene s = scenes[wk.scene];
    const bool act = c < wk.count;
    const int col = wk.first + c;
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float precision = 1.0f / (dv + (float)scs[SC_NOISE]);
    const float keep = damping, update = 1.0f - damping;
    load_tw256(tw, twtab + s.tw1_off);
    const unsigned bits = (mbits && s.mbits_off >= 0 && act) ? mbits[s.mbits_off + col * 16 + t] : 0xFFFFu;
    float2 v[16];
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int i = t + 16 * j;
      v[j] = make_float2(0.f, 0.f);
      if (act && i < s.m1) {
        const long long gi = s.samp_off + (long long)i * s.m2 + col;
        const bool ob = (bits >> j) & 1u;
        if (MODE == 0) {
          if (ob) v[j] = y[gi];
          acc += (double)v[j].x * v[j].x + (double)v[j].y * v[j].y;
        } else {
          // y - (pred - dv msg) with the residual y - pred stored by the
          // forward pass (one sample array less to read)
          const float2 rv = resid[gi], mv = msg[gi];
          const float2 cand = make_float2(precision * (rv.x + dv * mv.x), precision * (rv.y + dv * mv.y));
          if (ob) v[j] = make_float2(keep * mv.x + update * cand.x, keep * mv.y + update * cand.y);
          if (store_msg) msg[gi] = v[j];
        }
      }
    }
    if (MODE == 0) {
      acc = block_sum(acc, red);
      if (threadIdx.x == 0) part[(long long)PT_YE * pstride + s.col_part + wk.slot] = acc;
    }
    __syncthreads();
    fft256<true, COLS>(v, tw, x, t);
    if (act) {
#pragma unroll
      for (int n = 0; n < 16; ++n) mid[s.mid_off + (long long)(t + 16 * n) * s.m2 + col] = v[n];
    }
    __syncthreads();
  }
}

template <int COLS>
__global__ void __launch_bounds__(16 * COLS)
k_col16_forward(const FbmeScene* __restrict__ scenes, const FbmeWork* __restrict__ work, int nwork,
                const float2* __restrict__ mid, const float2* __restrict__ y,
                const uint16_t* __restrict__ mbits, float2* __restrict__ resid,
                float2* __restrict__ msg, const double* __restrict__ sc, double* __restrict__ part,
                int pstride, int first_message, float damping, const float2* __restrict__ twtab) {
  __shared__ float2 xs[16 * 17 * COLS];   // column c's element i at xs[i * COLS + c]
  __shared__ float2 tw[256];
  __shared__ double red[8];
  const int c = threadIdx.x % COLS, t = threadIdx.x / COLS;
  float2* x = xs + c;
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const FbmeWork wk = work[wi];
    const FbmeScene s = scenes[wk.scene];
    const bool act = c < wk.count;
    const int col = wk.first + c;
    load_tw256(tw, twtab + s.tw1_off);
    float2 v[16], yv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int i = t + 16 * j;
      v[j] = act ? mid[s.mid_off + (long long)i * s.m2 + col] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft256<false, COLS>(v, tw, x, t);
    // data loaded after the transform (not live across it)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int i = t + 16 * j;
      yv[j] = (act && i < s.m1) ? y[s.samp_off + (long long)i * s.m2 + col] : make_float2(0.f, 0.f);
    }
    const unsigned bits = (mbits && s.mbits_off >= 0 && act) ? mbits[s.mbits_off + col * 16 + t] : 0xFFFFu;
    const double* scs = sc + (long long)wk.scene * SC_COUNT;
    const float dv = fmaxf((float)scs[SC_DV], F32_TINY);
    const float precision = 1.0f / (dv + (float)scs[SC_NOISE]);
    double acc = 0.0;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int i = t + 16 * n;
      if (act && i < s.m1) {
        const long long gi = s.samp_off + (long long)i * s.m2 + col;
        const bool ob = (bits >> n) & 1u;
        const float rx = yv[n].x - v[n].x, ry = yv[n].y - v[n].y;
        resid[gi] = make_float2(rx, ry);
        if (ob) acc += (double)rx * rx + (double)ry * ry;
        if (first_message) {
          // First message from zero state (triton_kernels.py:66-77 with
          // prediction = message = 0).
          msg[gi] = ob ? make_float2((1.0f - damping) * (precision * yv[n].x),
                                     (1.0f - damping) * (precision * yv[n].y))
                       : make_float2(0.f, 0.f);
        }
      }
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[(long long)PT_RES * pstride + s.col_part + wk.slot] = acc;
    __syncthreads();
  }
}

// Mask words of the register column passes: word [col][t] bit j = sample
// (t + 16 j, col) observed (rows t + 16 j >= M1 give 0).  One thread per word.
__global__ void k_mask_bits16(const unsigned char* __restrict__ mask, int m1, int m2, uint16_t* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m2 * 16) return;
  const int col = e >> 4, t = e & 15;
  unsigned bits = 0u;
  for (int j = 0; j < 16; ++j) {
    const int i = t + 16 * j;
    if (i < m1 && mask[(long long)i * m2 + col]) bits |= 1u << j;
  }
  out[e] = (uint16_t)bits;
}

}  // namespace

extern "C" {

int fbme_launch_mask_bits16(const unsigned char* mask, int m1, int m2, void* out, hipStream_t stream) {
  if (m1 < 1 || m1 > 256 || m2 < 1) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_mask_bits16, dim3((m2 * 16 + 255) / 256), dim3(256), 0, stream, mask, m1, m2, (uint16_t*)out);
  return (int)hipGetLastError();
}


int fbme_launch_row16(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, void* mid, void* bp,
                      float* gamma, void* mean, float* var, const double* sc, double* part, int pstride, int mode,
                      float damping, float prior_damping, float max_precision, float hyper_shape, float hyper_rate,
                      float blend, uint32_t* cand_key, int* cand_idx, int store_final, uint32_t* score,
                      const void* twtab, hipStream_t stream) {
  if (nwork <= 0) return 0;
  // mode 3 without a candidate buffer: the final E-step of a plan without a
  // polish (no scores, no selection)
  if (mode == 3 && cand_key && !score) return FBME_ERR_ARG;
  if (mode == 3 && !cand_key) score = nullptr;
  UampArgs16 a{damping, prior_damping, max_precision, hyper_shape, hyper_rate, blend};
#define FBME_ROW16(M, T)                                                                                     \
  hipLaunchKernelGGL((k_row16<M, T>), dim3(grid), dim3(64), 0, stream, scenes, work, nwork, (float2*)mid,    \
                     (float2*)bp, gamma, (float2*)mean, var, sc, part, pstride, a, store_final, score,       \
                     (const float2*)twtab)
  // modes 4 and 5: modes 1 and 2 with the change partials
  switch (mode) {
    case 0: FBME_ROW16(0, false); break;
    case 1: FBME_ROW16(1, false); break;
    case 2: FBME_ROW16(2, false); break;
    case 4: FBME_ROW16(1, true); break;
    case 5: FBME_ROW16(2, true); break;
    default: FBME_ROW16(3, false); break;
  }
#undef FBME_ROW16
  if (mode == 3 && cand_key) {
    const int e = (int)hipGetLastError();
    if (e) return e;
    hipLaunchKernelGGL(k_tile_topk16, dim3(grid), dim3(64), 0, stream, scenes, work, nwork, score, cand_key,
                       cand_idx);
  }
  return (int)hipGetLastError();
}

int fbme_launch_col16_adjoint(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, const void* y,
                              const uint16_t* mbits, const void* pred, void* msg, void* mid, const double* sc,
                              double* part, int pstride, int mode, int store_msg, float damping, const void* twtab,
                              hipStream_t stream) {
  if (nwork <= 0) return 0;
  constexpr int C = FBME_R16_COLS;
  if (mode == 0)
    hipLaunchKernelGGL((k_col16_adjoint<C, 0>), dim3(grid), dim3(16 * C), 0, stream, scenes, work, nwork,
                       (const float2*)y, mbits, (const float2*)pred, (float2*)msg, (float2*)mid, sc, part, pstride,
                       store_msg, damping, (const float2*)twtab);
  else
    hipLaunchKernelGGL((k_col16_adjoint<C, 1>), dim3(grid), dim3(16 * C), 0, stream, scenes, work, nwork,
                       (const float2*)y, mbits, (const float2*)pred, (float2*)msg, (float2*)mid, sc, part, pstride,
                       store_msg, damping, (const float2*)twtab);
  return (int)hipGetLastError();
}

int fbme_launch_col16_forward(const FbmeScene* scenes, const FbmeWork* work, int nwork, int grid, const void* mid,
                              const void* y, const uint16_t* mbits, void* pred, void* msg, const double* sc,
                              double* part, int pstride, int first_message, float damping, const void* twtab,
                              hipStream_t stream) {
  if (nwork <= 0) return 0;
  constexpr int C = FBME_R16_COLS;
  hipLaunchKernelGGL((k_col16_forward<C>), dim3(grid), dim3(16 * C), 0, stream, scenes, work, nwork,
                     (const float2*)mid, (const float2*)y, mbits, (float2*)pred, (float2*)msg, sc, part, pstride,
                     first_message, damping, (const float2*)twtab);
  return (int)hipGetLastError();
}

}  // extern "C"
