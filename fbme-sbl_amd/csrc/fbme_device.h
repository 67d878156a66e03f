// Device-side building blocks shared by the FBME HIP kernels (gfx950).
//
//  * complex64 helpers on float2,
//  * a mixed-radix (2,3,4,5,8) Stockham FFT over a tile of lines held in LDS,
//    with each stage staged through registers so one LDS buffer suffices,
//  * deterministic block reductions in double (fixed tree order, no atomics),
//  * a block-wide radix select of the K largest 32-bit keys.
//
// Every FFT here is unnormalised: forward uses exp(-2 pi i jk/L), inverse uses
// exp(+2 pi i jk/L), matching torch.fft.fft / (L * torch.fft.ifft) which the
// reference operator uses (torch_sbl/operators.py:237-280).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fbme_types.h"

#define FBME_BLOCK 256
// Longest line of the LDS Stockham passes (twiddle tables up to this length;
// the BTTB embedding of a 1024-sample axis is 2048 long).
#define FBME_FFT_MAX 2048

namespace fbme {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }
// Multiply by -i (forward) or +i (inverse).
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------- butterflies
template <bool INV>
__device__ __forceinline__ void bfly2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void bfly4(float2* v) {
  float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  float2 s13 = cadd(v[1], v[3]), d13 = rot90<INV>(csub(v[1], v[3]));
  v[0] = cadd(s02, s13);
  v[2] = csub(s02, s13);
  v[1] = cadd(d02, d13);
  v[3] = csub(d02, d13);
}

template <bool INV>
__device__ __forceinline__ void bfly8(float2* v) {
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<INV>(e);
  bfly4<INV>(o);
  const float h = 0.70710678118654752440f;
  // W8^k for k = 1, 2, 3 (forward: exp(-i pi k / 4)).
  float2 o1 = INV ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                  : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
  float2 o2 = rot90<INV>(o[2]);
  float2 o3 = INV ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                  : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <bool INV>
__device__ __forceinline__ void bfly3(float2* v) {
  const float s = 0.86602540378443864676f;
  float2 t1 = cadd(v[1], v[2]);
  float2 t2 = csub(v[1], v[2]);
  float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  float2 r = rot90<INV>(cscale(t2, s));
  v[0] = cadd(v[0], t1);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <bool INV>
__device__ __forceinline__ void bfly5(float2* v) {
  const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  float2 b1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  float2 b2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  float2 r1 = rot90<INV>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  float2 r2 = rot90<INV>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(b1, r1);
  v[4] = csub(b1, r1);
  v[2] = cadd(b2, r2);
  v[3] = csub(b2, r2);
}

template <int R, bool INV>
__device__ __forceinline__ void butterfly(float2* v) {
  if constexpr (R == 2) bfly2<INV>(v);
  else if constexpr (R == 3) bfly3<INV>(v);
  else if constexpr (R == 4) bfly4<INV>(v);
  else if constexpr (R == 5) bfly5<INV>(v);
  else bfly8<INV>(v);
}

// Twiddle table of one line length, tw[i] = exp(-2 pi i * i / L), built in
// double and rounded once so the FFT error stays at the float32 level.
__device__ __forceinline__ void build_twiddles(float2* tw, int len) {
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    double s, c;
    sincospi(2.0 * (double)i / (double)len, &s, &c);
    tw[i] = make_float2((float)c, (float)-s);
  }
}

// LDS index with one pad slot every 16 complex values: stride-R butterfly
// writes of the first stages then spread over distinct banks.
__device__ __forceinline__ int lds_idx(int i) { return i + (i >> 4); }
// LDS complex slots needed for `elems` logical elements.
#define FBME_LDS_SLOTS(elems) ((elems) + ((elems) >> 4) + 80)

// Copy a precomputed twiddle table (rounded once from double on the host)
// into LDS.
__device__ __forceinline__ void load_twiddles(float2* tw, const float2* __restrict__ table, int len) {
  for (int i = threadIdx.x; i < len; i += blockDim.x) tw[i] = table[i];
}

// One Stockham stage of radix R over `nlines` lines of length `len` stored at
// buf[lds_idx(line * stride + j)].  Ns is the product of the radices already
// applied.  All inputs of the stage are loaded to registers before the
// barrier and all outputs written after it, so the stage runs in place.
// EPT bounds the elements per thread: nlines * len <= EPT * blockDim.x.
template <int R, int EPT, bool INV>
__device__ __forceinline__ void fft_stage(float2* buf, int len, int nlines, int stride, int ns,
                                          const float2* tw) {
  constexpr int NB = (EPT + R - 1) / R;
  const int per_line = len / R;
  const int total = per_line * nlines;
  float2 v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int g = threadIdx.x + b * FBME_BLOCK;
    if (g < total) {
      const int line = g / per_line;
      const int j = g - line * per_line;
      const int base = line * stride + j;
#pragma unroll
      for (int t = 0; t < R; ++t) v[b][t] = buf[lds_idx(base + t * per_line)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int g = threadIdx.x + b * FBME_BLOCK;
    if (g < total) {
      const int line = g / per_line;
      const int j = g - line * per_line;
      const int k = j % ns;
      if (ns > 1) {
        const int step = len / (ns * R);
#pragma unroll
        for (int t = 1; t < R; ++t) {
          float2 w = tw[t * k * step];
          if (INV) w.y = -w.y;
          v[b][t] = cmul(v[b][t], w);
        }
      }
      butterfly<R, INV>(v[b]);
      const int base = line * stride + (j - k) * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) buf[lds_idx(base + t * ns)] = v[b][t];
    }
  }
  __syncthreads();
}

// Full line FFT; `code` packs the radix sequence 4 bits per stage (first
// stage in the low nibble, 0 terminates).  The caller has synchronised after
// filling `buf` and the twiddle table.
template <int EPT, bool INV>
__device__ __forceinline__ void fft_lines(float2* buf, int len, int nlines, int stride, unsigned code,
                                          const float2* tw) {
  int ns = 1;
  while (code) {
    const int r = code & 15u;
    code >>= 4;
    switch (r) {
      case 2: fft_stage<2, EPT, INV>(buf, len, nlines, stride, ns, tw); break;
      case 3: fft_stage<3, EPT, INV>(buf, len, nlines, stride, ns, tw); break;
      case 4: fft_stage<4, EPT, INV>(buf, len, nlines, stride, ns, tw); break;
      case 5: fft_stage<5, EPT, INV>(buf, len, nlines, stride, ns, tw); break;
      default: fft_stage<8, EPT, INV>(buf, len, nlines, stride, ns, tw); break;
    }
    ns *= r;
  }
}

// ----------------------------------------------------------------- reductions
// Deterministic block sum of one double per thread (fixed shuffle/LDS tree).
// Sum over the 64 lanes; the result is valid in lane 0 only.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over the 64 lanes, broadcast to every lane (same value as wave_sum).
__device__ __forceinline__ double wave_sum_all(double v) { return __shfl(wave_sum(v), 0, 64); }

// DPP move of a double between the lanes of a row of 16.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 64 lanes, in every lane, by row-local DPP steps (quad
// permutes [1,0,3,2] and [2,3,0,1], half-row mirror, row mirror: each step
// adds two equal-shaped partial sums, so all lanes of a row hold the same
// bits) and the four row sums added in lane order.  Deterministic; call with
// all 64 lanes active.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  auto lane_f64 = [&](int l) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  };
  return ((lane_f64(0) + lane_f64(16)) + lane_f64(32)) + lane_f64(48);
}

__device__ __forceinline__ double block_sum(double v, double* scratch /* >= 4 */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double r = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) r += scratch[i];
  __syncthreads();
  return r;
}

// Two sums at once (same tree for both).
__device__ __forceinline__ void block_sum2(double& a, double& b, double* scratch /* >= 8 */) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    scratch[2 * wid] = a;
    scratch[2 * wid + 1] = b;
  }
  __syncthreads();
  double ra = 0.0, rb = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) {
    ra += scratch[2 * i];
    rb += scratch[2 * i + 1];
  }
  __syncthreads();
  a = ra;
  b = rb;
}

// Sortable key of a score: larger score -> larger key.  Scores here are
// non-negative; -inf (excluded entries) maps to 0 and NaN to the top, which
// is where torch.topk places NaN.
__device__ __forceinline__ uint32_t score_key(float s) {
  if (s != s) return 0xFFFFFFFFu;
  if (!(s > -INFINITY)) return 0u;
  uint32_t u = __float_as_uint(s);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}


// Packed selection key: score key in the high word, the complement of the
// position in the low word, so the larger score wins and equal scores go to
// the lower position (a deterministic tie rule).
__device__ __forceinline__ unsigned long long pack_key(uint32_t key, int pos) {
  return ((unsigned long long)key << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)pos);
}
__device__ __forceinline__ int packed_pos(unsigned long long p) { return (int)(0xFFFFFFFFu - (uint32_t)p); }

// DPP move of a 64-bit value between the lanes of a row of 16 (the same
// dpp_ctrl on both halves).
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}

// Maximum over the 64 lanes, in every lane.  Four row-local DPP steps (quad
// permutes [1,0,3,2] and [2,3,0,1], half-row mirror, row mirror) leave every
// row of 16 lanes holding its maximum; the row maxima are read from lanes 0,
// 16, 32 and 48.  Call with all 64 lanes active.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  auto mx = [](unsigned long long a, unsigned long long b) { return a > b ? a : b; };
  v = mx(v, dpp_u64<0xB1>(v));
  v = mx(v, dpp_u64<0x4E>(v));
  v = mx(v, dpp_u64<0x141>(v));
  v = mx(v, dpp_u64<0x140>(v));
  return mx(mx(readlane_u64(v, 0), readlane_u64(v, 16)), mx(readlane_u64(v, 32), readlane_u64(v, 48)));
}

// Top-k of the NPL packed keys held by each lane of one wavefront, by k
// rounds of wave-wide arg-max (no barriers, no atomics).  Lane 0 of the
// wave writes the winners in decreasing order to out[0..k).  Entries equal
// to 0 are padding and never win against a real key.
template <int NPL>
__device__ __forceinline__ void wave_topk(unsigned long long (&v)[NPL], int k, unsigned long long* out) {
  const int lane = threadIdx.x & 63;
  unsigned long long best = 0ull;
#pragma unroll
  for (int i = 0; i < NPL; ++i) best = v[i] > best ? v[i] : best;
  for (int r = 0; r < k; ++r) {
    const unsigned long long w = wave_max_u64(best);
    if (lane == 0) out[r] = w;
    if (w != 0ull && best == w) {
      // the unique owner (positions are distinct) retires it
      best = 0ull;
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        if (v[i] == w) v[i] = 0ull;
        best = v[i] > best ? v[i] : best;
      }
    }
  }
}

// Block-wide selection of the k largest keys among keys[0..n) (in LDS).
// Writes the chosen positions to out_pos[0..k) in unspecified order (ties on
// the threshold key are broken towards the lower position) and returns the
// number written (min(k, n)).  `hist` needs 256 uint32 of LDS, `misc` 4.
__device__ int block_select_topk(const uint32_t* keys, int n, int k, int* out_pos, uint32_t* hist,
                                 int* misc) {
  if (k > n) k = n;
  uint32_t prefix = 0u, mask = 0u;
  int remaining = k;
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0u;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const uint32_t key = keys[i];
      if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      // Wave-parallel search from the top digit: lane l owns digits
      // 255-4l .. 252-4l; an inclusive scan of the lane sums finds the lane
      // where the count of larger-or-equal keys first reaches `remaining`.
      const int lane = threadIdx.x;
      const int b0 = 255 - 4 * lane;
      const int c0 = (int)hist[b0], c1 = (int)hist[b0 - 1], c2 = (int)hist[b0 - 2], c3 = (int)hist[b0 - 3];
      const int local = c0 + c1 + c2 + c3;
      int incl = local;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
      }
      const int excl = incl - local;
      if (excl < remaining && remaining <= incl) {
        int cum = excl, d;
        if (cum + c0 >= remaining) {
          d = b0;
        } else if ((cum += c0) + c1 >= remaining) {
          d = b0 - 1;
        } else if ((cum += c1) + c2 >= remaining) {
          d = b0 - 2;
        } else {
          cum += c2;
          d = b0 - 3;
        }
        misc[0] = d;
        misc[1] = remaining - cum;
      }
    }
    __syncthreads();
    prefix |= ((uint32_t)misc[0]) << shift;
    mask |= 255u << shift;
    remaining = misc[1];
    __syncthreads();
  }
  // Keys > prefix are all taken; `remaining` of the keys == prefix are taken
  // in position order.
  if (threadIdx.x == 0) {
    misc[2] = 0;
    misc[3] = 0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t key = keys[i];
    if (key > prefix) {
      const int slot = atomicAdd(&misc[2], 1);
      out_pos[slot] = i;
    } else if (key == prefix) {
      atomicAdd(&misc[3], 1);
    }
  }
  __syncthreads();
  if (misc[3] == remaining) {
    // Usual case: every key equal to the threshold is taken.
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      if (keys[i] == prefix) {
        const int slot = atomicAdd(&misc[2], 1);
        out_pos[slot] = i;
      }
    }
  } else if (threadIdx.x == 0) {
    // Exact ties across the threshold: take the lowest positions.
    int slot = misc[2];
    for (int i = 0; i < n && remaining > 0; ++i) {
      if (keys[i] == prefix) {
        out_pos[slot++] = i;
        --remaining;
      }
    }
    misc[2] = slot;
  }
  __syncthreads();
  return k;
}

// Write the `got` selected tile positions (pos[0..got)) as the tile's K
// candidates sorted by (key desc, position asc), padded with key 0 / index
// -1; `base` converts a tile position to a grid index.  Threads < K work.
__device__ __forceinline__ void write_sorted_candidates(const uint32_t* keys, const int* pos, int got, int k,
                                                        int base, uint32_t* out_key, int* out_idx) {
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    if (i < got) {
      const uint32_t ki = keys[pos[i]];
      const int pi = pos[i];
      int rank = 0;
      for (int j = 0; j < got; ++j) {
        const uint32_t kj = keys[pos[j]];
        rank += (kj > ki) || (kj == ki && pos[j] < pi);
      }
      out_key[rank] = ki;
      out_idx[rank] = base + pi;
    } else {
      out_key[i] = 0u;
      out_idx[i] = -1;
    }
  }
}

}  // namespace fbme
