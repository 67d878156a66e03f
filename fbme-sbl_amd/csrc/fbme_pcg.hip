// Covariance-free PCG + Hutchinson path of the general solver (gfx950).
//
// Restates, on HIP kernels, the mode="cg" pieces of torch_sbl/solver.py and
// torch_sbl/operators.py:
//   * exact BTTB covariance spectrum (operators.py:299-326) and matvec
//     Q d = crop(IFFT(FFT(pad d) * spectrum)) (operators.py:328-355), with
//     the 1 + P right-hand sides (data + Rademacher probes) batched into the
//     leading FFT dimension; the column FFT, the spectrum product and the
//     inverse column FFT are fused into one pass;
//   * low-rank Woodbury preconditioner (solver.py:5786-5848): strongest
//     atoms, exact-phase Fourier factors, Dirichlet-kernel Gram, FP64
//     Cholesky and inverse of the 32 x 32 system;
//   * PCG lane arithmetic with masks (solver.py:5582-5723), inner products in
//     FP64;
//   * Hutchinson posterior diagonal and moments (solver.py:3016-3036),
//     relevance projection (3156-3181), MacKay update (1669-1677) and
//     damping (1776-1798).
// Vectors of lane (b, r) are contiguous: X[(b * R + r) * M + m].
#include <algorithm>
#include <mutex>

#include "fbme_device.h"

using namespace fbme;

namespace {

constexpr double F32_EPS_D = 1.1920928955078125e-07;
constexpr double F32_TINY_D = 1.17549435082228750797e-38;
constexpr int RANK = 32;

// ------------------------------------------------------------ BTTB spectrum
__global__ void k_real_to_complex(const float* __restrict__ in, float2* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2(in[i], 0.f);
}

// kern[b, e1, e2] = lags[b, lag1 mod N1, lag2 mod N2] at the signed-lag
// positions of the circulant embedding, plus the noise at the origin.
__global__ void k_embed(const float2* __restrict__ lags, const float* __restrict__ noise,
                        float2* __restrict__ kern, int n1, int n2, int m1, int m2, int e1, int e2) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= e1 * e2) return;
  const int a1 = i / e2, a2 = i - a1 * e2;
  int l1 = -1, l2 = -1;
  if (a1 < m1) l1 = a1;
  else if (a1 > e1 - m1) l1 = a1 - e1 + n1;       // negative lag, mod n1
  if (a2 < m2) l2 = a2;
  else if (a2 > e2 - m2) l2 = a2 - e2 + n2;
  float2 v = make_float2(0.f, 0.f);
  if (l1 >= 0 && l2 >= 0) v = lags[((long long)b * n1 + l1) * n2 + l2];
  if (i == 0 && noise) v.x += noise[b];
  kern[(long long)b * e1 * e2 + i] = v;
}

// Fused column pass of the Toeplitz matvec on [BR, M1, E2]: for each column,
// zero-pad M1 -> E1, forward FFT, multiply by spectrum[b][:, c] / (E1 E2),
// inverse FFT, keep the first M1 rows.
// TWN bounds E1 (the twiddle table's LDS length).
template <int EPT, int TWN>
__global__ void __launch_bounds__(FBME_BLOCK)
k_toeplitz_col(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ spec,
               int rhs, int m1, int e1, int e2, int tile_cols, unsigned code, const float2* __restrict__ twtab) {
  __shared__ float2 buf[FBME_LDS_SLOTS(EPT * FBME_BLOCK)];
  __shared__ float2 tw[TWN];
  const int lanev = blockIdx.y;          // b * rhs + r
  const int b = lanev / rhs;
  const int c0 = blockIdx.x * tile_cols;
  const int cc = min(tile_cols, e2 - c0);
  const int stride = e1 + 1;
  load_twiddles(tw, twtab, e1);
  const float2* src = in + (long long)lanev * m1 * e2;
  const float2* sp = spec + (long long)b * e1 * e2;
  const float scale = 1.0f / ((float)e1 * (float)e2);
  for (int e = threadIdx.x; e < e1 * cc; e += FBME_BLOCK) {
    const int i = e / cc, c = e - i * cc;
    buf[lds_idx(c * stride + i)] = (i < m1) ? src[(long long)i * e2 + c0 + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lines<EPT, false>(buf, e1, cc, stride, code, tw);
  for (int e = threadIdx.x; e < e1 * cc; e += FBME_BLOCK) {
    const int i = e / cc, c = e - i * cc;
    const int k = lds_idx(c * stride + i);
    buf[k] = cscale(cmul(buf[k], sp[(long long)i * e2 + c0 + c]), scale);
  }
  __syncthreads();
  fft_lines<EPT, true>(buf, e1, cc, stride, code, tw);
  float2* dst = out + (long long)lanev * m1 * e2;
  for (int e = threadIdx.x; e < m1 * cc; e += FBME_BLOCK) {
    const int i = e / cc, c = e - i * cc;
    dst[(long long)i * e2 + c0 + c] = buf[lds_idx(c * stride + i)];
  }
}

// Row pass on [rows, len_in] -> [rows, len_out] with an FFT of length n <= TWN,
// twiddles from a table (plan-owned).
template <int EPT, bool INV, int TWN>
__global__ void __launch_bounds__(FBME_BLOCK)
k_rowpass(const float2* __restrict__ in, float2* __restrict__ out, int n, int len_in, int len_out,
          long long rows, int tile_rows, unsigned code, const float2* __restrict__ twtab) {
  __shared__ float2 buf[FBME_LDS_SLOTS(EPT * FBME_BLOCK)];
  __shared__ float2 tw[TWN];
  const long long r0 = (long long)blockIdx.x * tile_rows;
  const int rr = (int)min((long long)tile_rows, rows - r0);
  const int stride = n + 1;
  load_twiddles(tw, twtab, n);
  const float2* src = in + r0 * len_in;
  for (int e = threadIdx.x; e < rr * n; e += FBME_BLOCK) {
    const int r = e / n, j = e - r * n;
    buf[lds_idx(r * stride + j)] = (j < len_in) ? src[(long long)r * len_in + j] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lines<EPT, INV>(buf, n, rr, stride, code, tw);
  float2* dst = out + r0 * len_out;
  for (int e = threadIdx.x; e < rr * len_out; e += FBME_BLOCK) {
    const int r = e / len_out, j = e - r * len_out;
    dst[(long long)r * len_out + j] = buf[lds_idx(r * stride + j)];
  }
}

// ------------------------------------------- Strang circulant preconditioner
// (operators.py:373-416): the circulant column of Q on the M1 x M2 sample
// grid takes lag (k <= M / 2 ? k : k - M) on each axis from fft2(gamma);
// its eigenvalues fft2(column) + noise are floored at max(eps_f32 max|eig|,
// noise), and M^-1 v = ifft2(fft2(v) / eig).
__global__ void k_circ_gather(const float2* __restrict__ lags, int n1, int n2, int m1, int m2,
                              float2* __restrict__ col) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m1 * m2) return;
  const int k1 = i / m2, k2 = i - k1 * m2;
  const int l1 = k1 <= m1 / 2 ? k1 : n1 - (m1 - k1);
  const int l2 = k2 <= m2 / 2 ? k2 : n2 - (m2 - k2);
  col[(long long)b * m1 * m2 + i] = lags[((long long)b * n1 + l1) * n2 + l2];
}

__global__ void __launch_bounds__(FBME_BLOCK)
k_circ_eig(const float2* __restrict__ spec, const float* __restrict__ noise, int m, float* __restrict__ eig) {
  __shared__ float red[FBME_BLOCK / 64];
  const int b = blockIdx.x;
  const float nz = noise[b];
  const float2* sp = spec + (long long)b * m;
  float mx = 0.f;
  for (int i = threadIdx.x; i < m; i += FBME_BLOCK) mx = fmaxf(mx, fabsf(sp[i].x + nz));
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  float amax = 0.f;
  for (int w = 0; w < FBME_BLOCK / 64; ++w) amax = fmaxf(amax, red[w]);
  const float fl = fmaxf(1.1920928955078125e-07f * amax, nz);
  for (int i = threadIdx.x; i < m; i += FBME_BLOCK) eig[(long long)b * m + i] = fmaxf(sp[i].x + nz, fl);
}

// data[lane] /= eig[scene] * M (the unnormalised inverse transform follows).
__global__ void k_circ_scale(float2* __restrict__ data, const float* __restrict__ eig, int rhs, int m) {
  const int l = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const float s = 1.0f / (eig[(long long)(l / rhs) * m + i] * (float)m);
  float2* d = data + (long long)l * m + i;
  *d = make_float2(d->x * s, d->y * s);
}

// -------------------------------------------------------------- tile top-k
// Top RANK entries of each 1024-value tile of x[b, :] (by value, then lower
// index), as packed keys.
__global__ void __launch_bounds__(FBME_BLOCK)
k_topk_tile(const float* __restrict__ x, long long n, int tiles, unsigned long long* __restrict__ cand) {
  __shared__ unsigned long long wtop[4][RANK];
  const int b = blockIdx.y;
  const int t = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned long long v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long long e = (long long)t * 1024 + wave * 256 + lane + 64 * i;
    v[i] = (e < n) ? pack_key(score_key(x[(long long)b * n + e]), (int)e) : 0ull;
  }
  wave_topk<4>(v, RANK, wtop[wave]);
  __syncthreads();
  if (wave == 0) {
    unsigned long long u[2];
    u[0] = wtop[lane >> 5][lane & 31];
    u[1] = wtop[2 + (lane >> 5)][lane & 31];
    wave_topk<2>(u, RANK, cand + ((long long)b * tiles + t) * RANK);
  }
}

// One level of the candidate merge for large grids: every group of up to
// 2048 candidate keys of a scene (one wavefront) keeps its top RANK.  Keys
// pack (value, position), so each level is exact and keeps the tie order.
__global__ void __launch_bounds__(64)
k_topk_merge_keys(const unsigned long long* __restrict__ cand_in, int count, int groups,
                  unsigned long long* __restrict__ cand_out) {
  const int g = blockIdx.x, b = blockIdx.y;
  unsigned long long v[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int slot = g * 2048 + threadIdx.x + 64 * i;
    v[i] = (slot < count) ? cand_in[(long long)b * count + slot] : 0ull;
  }
  wave_topk<32>(v, RANK, cand_out + ((long long)b * groups + g) * RANK);
}

// Top k <= 256 (the active-set selection, solver.py:3184-3216): every
// 1024-value tile keeps its top k, then groups of 2048 candidates keep
// theirs, level by level, until one wavefront holds them all.  Keys pack
// (value, position), ties go to the lower position.
__global__ void __launch_bounds__(64)
k_topkk_tile(const float* __restrict__ x, long long n, int tiles, int k, unsigned long long* __restrict__ cand) {
  const int t = blockIdx.x, b = blockIdx.y;
  unsigned long long v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const long long e = (long long)t * 1024 + threadIdx.x + 64 * i;
    v[i] = (e < n) ? pack_key(score_key(x[(long long)b * n + e]), (int)e) : 0ull;
  }
  wave_topk<16>(v, k, cand + ((long long)b * tiles + t) * k);
}

__global__ void __launch_bounds__(64)
k_topkk_merge(const unsigned long long* __restrict__ cand_in, int count, int groups, int k,
              unsigned long long* __restrict__ cand_out) {
  const int g = blockIdx.x, b = blockIdx.y;
  unsigned long long v[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int slot = g * 2048 + threadIdx.x + 64 * i;
    v[i] = (slot < count) ? cand_in[(long long)b * count + slot] : 0ull;
  }
  wave_topk<32>(v, k, cand_out + ((long long)b * groups + g) * k);
}

__global__ void __launch_bounds__(64)
k_topkk_final(const unsigned long long* __restrict__ cand, int count, int k, const float* __restrict__ x, long long n,
              int* __restrict__ idx_out, float* __restrict__ val_out) {
  __shared__ unsigned long long top[256];
  const int b = blockIdx.x;
  unsigned long long v[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int slot = threadIdx.x + 64 * i;
    v[i] = (slot < count) ? cand[(long long)b * count + slot] : 0ull;
  }
  wave_topk<32>(v, k, top);
  __syncthreads();
  for (int r = threadIdx.x; r < k; r += 64) {
    const int e = packed_pos(top[r]);
    idx_out[(long long)b * k + r] = e;
    val_out[(long long)b * k + r] = x[(long long)b * n + e];
  }
}

// Merge the tile candidates of each scene into the sorted top RANK
// (values and indices), one wavefront per scene (tiles * RANK <= 2048).
__global__ void __launch_bounds__(64)
k_topk_merge(const unsigned long long* __restrict__ cand, int tiles, const float* __restrict__ x, long long n,
             int* __restrict__ idx_out, float* __restrict__ val_out) {
  __shared__ unsigned long long top[RANK];
  const int b = blockIdx.x;
  unsigned long long v[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int slot = threadIdx.x + 64 * i;
    v[i] = (slot < tiles * RANK) ? cand[(long long)b * tiles * RANK + slot] : 0ull;
  }
  wave_topk<32>(v, RANK, top);
  __syncthreads();
  if (threadIdx.x < RANK) {
    const int e = packed_pos(top[threadIdx.x]);
    idx_out[b * RANK + threadIdx.x] = e;
    val_out[b * RANK + threadIdx.x] = x[(long long)b * n + e];
  }
}

// ------------------------------------------------------ low-rank factors
// The factor matrix F[m, k] = exp(-2 pi i (m1 k1 / N1 + m2 k2 / N2))
// sqrt(gamma_k) is separable: F[m1 M2 + m2, k] = U[m1, k] V[m2, k] with
// U[m1, k] = exp(-2 pi i (m1 k1 mod N1) / N1) and
// V[m2, k] = exp(-2 pi i (m2 k2 mod N2) / N2) sqrt(gamma_k) (phases reduced
// modulo N in integers).  tab[b] = U [M1, RANK] followed by V [M2, RANK];
// the M x RANK matrix itself is never formed.
__global__ void k_lr_tables(const int* __restrict__ idx, const float* __restrict__ val, float2* __restrict__ tab,
                            int n1, int n2, int m1, int m2) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (m1 + m2) * RANK) return;
  const int row = i / RANK, k = i - row * RANK;
  const int g = idx[b * RANK + k];
  const int k1 = g / n2, k2 = g - k1 * n2;
  double ph;
  float amp = 1.f;
  if (row < m1) {
    ph = (double)(((long long)row * k1) % n1) / n1;
  } else {
    ph = (double)(((long long)(row - m1) * k2) % n2) / n2;
    amp = sqrtf(fmaxf(val[b * RANK + k], 0.f));
  }
  double s, c;
  sincospi(-2.0 * ph, &s, &c);
  tab[(long long)b * (m1 + m2) * RANK + i] = make_float2((float)c * amp, (float)s * amp);
}

// Inverse of the Hermitian RANK x RANK system held in a (one wavefront,
// lane < RANK owns row `lane`): symmetrised, loaded with 10 eps mean|diag|
// (solver.py:5827-5833), Cholesky-factored in registers, S^-1 = L^-H L^-1
// (solver.py:5848) written to sinv (row-major complex64).  a is
// overwritten by the factor; returns nonzero on a non-positive pivot.
__device__ int lr_chol_inverse(double2 (*a)[RANK + 1], double2 (*xs)[RANK + 1], int lane,
                               float2* __restrict__ sinv) {
  double2 row[RANK];
#pragma unroll
  for (int q = 0; q < RANK; ++q) {
    double2 v = make_double2(lane == q ? 1.0 : 0.0, 0.0);
    if (lane < RANK) {
      const double2 u = a[lane][q], l = a[q][lane];
      v = make_double2(0.5 * (u.x + l.x), 0.5 * (u.y - l.y));
    }
    row[q] = v;
  }
  // mean |diag| of the symmetrised system (read from LDS: row[lane] would
  // need a dynamic register index)
  double md = (lane < RANK) ? fabs(a[lane][lane].x) : 0.0;
  md = wave_sum_all(md) / RANK;
  const double jit = 10.0 * F32_EPS_D * fmax(md, F32_TINY_D);
  __syncthreads();
  int fail = 0;
#pragma unroll
  for (int j = 0; j < RANK; ++j) {
    if (lane == j) row[j].x += jit;
    const double d = __shfl(row[j].x, j, 64);
    fail |= !(d > 0.0);
    const double piv = sqrt(fmax(d, 0.0));
    if (lane > j) {
      row[j].x /= piv;
      row[j].y /= piv;
    } else if (lane == j) {
      row[j] = make_double2(piv, 0.0);
    }
#pragma unroll
    for (int q = j + 1; q < RANK; ++q) {
      const double lx = __shfl(row[j].x, q, 64), ly = __shfl(row[j].y, q, 64);
      if (lane >= q) {
        row[q].x -= row[j].x * lx + row[j].y * ly;
        row[q].y -= row[j].y * lx - row[j].x * ly;
      }
    }
  }
  if (lane < RANK) {
#pragma unroll
    for (int q = 0; q < RANK; ++q) a[lane][q] = row[q];
  }
  __syncthreads();
  double2 x[RANK];
#pragma unroll
  for (int r = 0; r < RANK; ++r) {
    double2 acc = make_double2(r == lane ? 1.0 : 0.0, 0.0);
#pragma unroll
    for (int p = 0; p < r; ++p) {
      const double2 l = a[r][p];
      acc.x -= l.x * x[p].x - l.y * x[p].y;
      acc.y -= l.x * x[p].y + l.y * x[p].x;
    }
    const double dg = a[r][r].x;
    x[r] = make_double2(acc.x / dg, acc.y / dg);
  }
  if (lane < RANK) {
#pragma unroll
    for (int r = 0; r < RANK; ++r) xs[r][lane] = x[r];
  }
  __syncthreads();
  // S^-1[i][j] = sum_r conj(X[r][i]) X[r][j]; lane j computes column j.
  if (lane < RANK) {
    for (int i = 0; i < RANK; ++i) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int r = 0; r < RANK; ++r) {
        const double2 xi = xs[r][i];
        acc.x += xi.x * x[r].x + xi.y * x[r].y;
        acc.y += xi.x * x[r].y - xi.y * x[r].x;
      }
      sinv[(long long)i * RANK + lane] = make_float2((float)acc.x, (float)acc.y);
    }
  }
  return fail;
}

// One wavefront per scene: small = I + F^H F / bg with F^H F from the
// Dirichlet Gram (or, under an observation mask, from the mask spectrum mh:
// the Gram of the factor columns on the observed rows), GPU Cholesky loading
// (10 eps mean|diag|, solver.py:5827-5833), inverse S^-1 = L^-H L^-1
// (solver.py:5848).
__global__ void __launch_bounds__(64)
k_lr_setup(const int* __restrict__ idx, const float* __restrict__ val, const double* __restrict__ sums,
           const float* __restrict__ noise, const double2* __restrict__ d1, const double2* __restrict__ d2,
           const double2* __restrict__ mh, int n1, int n2, float2* __restrict__ sinv, float* __restrict__ bg_out,
           int* __restrict__ status) {
  __shared__ double2 a[RANK][RANK + 1];
  __shared__ double2 xs[RANK][RANK + 1];
  __shared__ double sq[RANK];
  __shared__ int gi[RANK];
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  double vsum = 0.0;
  if (lane < RANK) {
    gi[lane] = idx[b * RANK + lane];
    sq[lane] = sqrt(fmax((double)val[b * RANK + lane], 0.0));
    vsum = (double)val[b * RANK + lane];
  }
  vsum = wave_sum_all(vsum);
  const double total = sums[b];
  const double omitted = fmax(total - vsum, 0.0);
  const double bg = fmax((double)noise[b] + omitted, F32_EPS_D * total);
  __syncthreads();
  for (int e = lane; e < RANK * RANK; e += 64) {
    const int r = e / RANK, q = e - r * RANK;
    int dd1 = gi[r] / n2 - gi[q] / n2;
    int dd2 = gi[r] % n2 - gi[q] % n2;
    dd1 += dd1 < 0 ? n1 : 0;
    dd2 += dd2 < 0 ? n2 : 0;
    double2 gv;
    if (mh) {
      gv = mh[(long long)dd1 * n2 + dd2];
    } else {
      const double2 u = d1[dd1], w = d2[dd2];
      gv = make_double2(u.x * w.x - u.y * w.y, u.x * w.y + u.y * w.x);
    }
    const double s = sq[r] * sq[q] / bg;
    a[r][q] = make_double2((r == q ? 1.0 : 0.0) + gv.x * s, gv.y * s);
  }
  __syncthreads();
  const int fail = lr_chol_inverse(a, xs, lane, sinv + (long long)b * RANK * RANK);
  if (lane == 0) {
    bg_out[b] = (float)bg;
    if (status && fail) status[b] = 1;
  }
}

// Projection partials proj[(b, r), k] = sum_m conj(F[m, k]) v[(b, r), m] over
// the rows [row0, row0 + rb) of scene b, for the LG lanes r0 .. r0 + LG - 1
// of one lane group: F is generated from the U / V tables once per sample
// and applied to every lane of the group.  Thread (g, k) = (tid / 32,
// tid % 32) takes the samples m2 = g (mod 8) of each row; the V values of
// its samples stay in registers across the rows of the block.
constexpr int LR_TILE = 128;

template <int LG>
__global__ void __launch_bounds__(FBME_BLOCK)
k_lr_project(const float2* __restrict__ tab, const float2* __restrict__ v, int m1, int m2, int rhs, int rb,
             double2* __restrict__ part) {
  constexpr int LDS = (LG * LR_TILE > 8 * LG * RANK) ? LG * LR_TILE : 8 * LG * RANK;
  __shared__ float2 buf[LDS];
  const int b = blockIdx.z;
  const int r0 = blockIdx.y * LG;
  const int nl = min(LG, rhs - r0);
  const int k = threadIdx.x & (RANK - 1), g = threadIdx.x >> 5;
  const int row0 = blockIdx.x * rb, row1 = min(m1, row0 + rb);
  const long long m = (long long)m1 * m2;
  const float2* U = tab + (long long)b * (m1 + m2) * RANK;
  const float2* V = U + (long long)m1 * RANK;
  const float2* vb = v + ((long long)b * rhs + r0) * m;
  float2 acc[LG];
#pragma unroll
  for (int j = 0; j < LG; ++j) acc[j] = make_float2(0.f, 0.f);
  for (int t0 = 0; t0 < m2; t0 += LR_TILE) {
    const int tl = min(LR_TILE, m2 - t0);
    float2 cv[LR_TILE / 8];   // conj(V) of this thread's samples in the tile
#pragma unroll
    for (int i = 0; i < LR_TILE / 8; ++i) {
      const int c = g + 8 * i;
      const float2 w = (c < tl) ? V[(long long)(t0 + c) * RANK + k] : make_float2(0.f, 0.f);
      cv[i] = make_float2(w.x, -w.y);
    }
    for (int row = row0; row < row1; ++row) {
      __syncthreads();
      for (int e = threadIdx.x; e < LG * tl; e += FBME_BLOCK) {
        const int j = e / tl, c = e - j * tl;
        buf[j * LR_TILE + c] = (j < nl) ? vb[(long long)j * m + (long long)row * m2 + t0 + c] : make_float2(0.f, 0.f);
      }
      __syncthreads();
      const float2 u = U[(long long)row * RANK + k];
      const float2 cu = make_float2(u.x, -u.y);
#pragma unroll
      for (int i = 0; i < LR_TILE / 8; ++i) {
        const int c = g + 8 * i;
        if (c < tl) {
          const float2 f = cmul(cu, cv[i]);
#pragma unroll
          for (int j = 0; j < LG; ++j) {
            const float2 w = buf[j * LR_TILE + c];
            acc[j].x += f.x * w.x - f.y * w.y;
            acc[j].y += f.x * w.y + f.y * w.x;
          }
        }
      }
    }
  }
  // reduce the 8 sample groups through LDS, in FP64
  __syncthreads();
#pragma unroll
  for (int j = 0; j < LG; ++j) buf[(g * LG + j) * RANK + k] = acc[j];
  __syncthreads();
  for (int e = threadIdx.x; e < nl * RANK; e += FBME_BLOCK) {
    const int j = e / RANK, kk = e - j * RANK;
    double2 sum = make_double2(0.0, 0.0);
    for (int q = 0; q < 8; ++q) {
      const float2 w = buf[(q * LG + j) * RANK + kk];
      sum.x += w.x;
      sum.y += w.y;
    }
    part[(((long long)b * rhs + r0 + j) * gridDim.x + blockIdx.x) * RANK + kk] = sum;
  }
}

// Projection partials proj[(b, r), k] = sum_m conj(U[b, k, m]) v[(b, r), m]
// onto RANK stored columns (the Nystrom eigenvectors) over the samples
// [chunk * ch, (chunk + 1) * ch) of scene b, for the LG lanes r0 .. r0 + LG - 1
// of one lane group.  Sample tiles of U and of the lanes are staged in LDS;
// thread (g, k) = (tid / 32, tid % 32) takes the samples g (mod 8) of a
// tile.  The eight partial sums are added in FP64 in a fixed order; the
// layout is k_lr_project's, for k_lr_corr.
constexpr int SP_TILE = 64;
constexpr int SP_CHUNK = 1024;

template <int LG>
__global__ void __launch_bounds__(FBME_BLOCK)
k_stored_project(const float2* __restrict__ ucols, const float2* __restrict__ v, long long m, int rhs,
                 double2* __restrict__ part) {
  __shared__ float2 us[RANK][SP_TILE + 1];
  __shared__ float2 vs[LG][SP_TILE];
  __shared__ float2 buf[8][LG][RANK];
  const int b = blockIdx.z;
  const int r0 = blockIdx.y * LG;
  const int nl = min(LG, rhs - r0);
  const int k = threadIdx.x & (RANK - 1), g = threadIdx.x >> 5;
  const float2* U = ucols + (long long)b * RANK * m;
  const float2* vb = v + ((long long)b * rhs + r0) * m;
  const long long s0 = (long long)blockIdx.x * SP_CHUNK, s1 = min(m, s0 + SP_CHUNK);
  float2 acc[LG];
#pragma unroll
  for (int j = 0; j < LG; ++j) acc[j] = make_float2(0.f, 0.f);
  for (long long base = s0; base < s1; base += SP_TILE) {
    const int tl = (int)min((long long)SP_TILE, s1 - base);
    __syncthreads();
    for (int e = threadIdx.x; e < RANK * SP_TILE; e += FBME_BLOCK) {
      const int q = e / SP_TILE, c = e - q * SP_TILE;
      us[q][c] = c < tl ? U[(long long)q * m + base + c] : make_float2(0.f, 0.f);
    }
    for (int e = threadIdx.x; e < LG * SP_TILE; e += FBME_BLOCK) {
      const int j = e / SP_TILE, c = e - j * SP_TILE;
      vs[j][c] = (j < nl && c < tl) ? vb[(long long)j * m + base + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int c = g; c < tl; c += 8) {
      const float2 u = us[k][c];
      const float2 f = make_float2(u.x, -u.y);
#pragma unroll
      for (int j = 0; j < LG; ++j) {
        const float2 w = vs[j][c];
        acc[j].x += f.x * w.x - f.y * w.y;
        acc[j].y += f.x * w.y + f.y * w.x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < LG; ++j) buf[g][j][k] = acc[j];
  __syncthreads();
  for (int e = threadIdx.x; e < nl * RANK; e += FBME_BLOCK) {
    const int j = e / RANK, kk = e - j * RANK;
    double2 sum = make_double2(0.0, 0.0);
    for (int q = 0; q < 8; ++q) {
      sum.x += buf[q][j][kk].x;
      sum.y += buf[q][j][kk].y;
    }
    part[(((long long)b * rhs + r0 + j) * gridDim.x + blockIdx.x) * RANK + kk] = sum;
  }
}

// corr[(b, r), :] = S^-1 (proj / bg) / bg, from the row-chunk partials:
// thread (q, k) = (tid / 32, tid % 32) sums chunks q, q + 8, ... of atom k,
// the eight partial sums are added in a fixed order, then thread k < 32
// forms row k of S^-1 proj.
// With rr_part: also nrho[lane] = <r, z> of z = r / bg - F corr without z,
// ||r||^2 / bg - Re(proj^H corr) (||r||^2 from the update pass's block
// partials), both sums in FP64 and fixed order.
__global__ void __launch_bounds__(FBME_BLOCK)
k_lr_corr(const double2* __restrict__ part, int chunks, const float2* __restrict__ sinv,
          const float* __restrict__ bg, int rhs, float2* __restrict__ corr, const double* __restrict__ rr_part,
          int rr_blocks, double* __restrict__ nrho) {
  __shared__ double2 acc8[8][RANK];
  __shared__ double2 proj[RANK];
  __shared__ double pcs[RANK];
  const int lanev = blockIdx.x;
  const int b = lanev / rhs;
  const int k = threadIdx.x & (RANK - 1), q = threadIdx.x >> 5;
  double2 s = make_double2(0.0, 0.0);
  for (int c = q; c < chunks; c += 8) {
    const double2 v = part[((long long)lanev * chunks + c) * RANK + k];
    s.x += v.x;
    s.y += v.y;
  }
  acc8[q][k] = s;
  __syncthreads();
  if (threadIdx.x < RANK) {
    double2 t = acc8[0][k];
    for (int w = 1; w < 8; ++w) {
      t.x += acc8[w][k].x;
      t.y += acc8[w][k].y;
    }
    proj[k] = t;
  }
  __syncthreads();
  if (threadIdx.x < RANK) {
    const double ib = 1.0 / (double)bg[b];
    double2 a = make_double2(0.0, 0.0);
    for (int jj = 0; jj < RANK; ++jj) {
      const float2 w = sinv[((long long)b * RANK + k) * RANK + jj];
      a.x += (double)w.x * proj[jj].x - (double)w.y * proj[jj].y;
      a.y += (double)w.x * proj[jj].y + (double)w.y * proj[jj].x;
    }
    const double2 c = make_double2(a.x * ib * ib, a.y * ib * ib);
    const float2 cf = make_float2((float)c.x, (float)c.y);
    corr[(long long)lanev * RANK + k] = cf;
    // the synthesis uses the rounded corr, so the identity does too
    pcs[k] = proj[k].x * (double)cf.x + proj[k].y * (double)cf.y;
  }
  if (rr_part) {   // uniform: a kernel argument
    __syncthreads();
    if (threadIdx.x < 64) {
      double rr = 0.0;
      for (int i = threadIdx.x; i < rr_blocks; i += 64) rr += rr_part[(long long)lanev * rr_blocks + i];
      rr = wave_sum(rr);
      double pc = threadIdx.x < RANK ? pcs[threadIdx.x] : 0.0;
      pc = wave_sum(pc);
      if (threadIdx.x == 0) nrho[lanev] = rr / (double)bg[b] - pc;
    }
  }
}

// out = v / bg - F corr (corr already carries the 1 / bg^2 factor): one
// sample per thread, the 32 factor entries of the sample generated once from
// U / V and applied to every lane of the group (corr staged in LDS).  DOT:
// also part[lane, block] = sum Re(conj(v) out)
// over the block's samples (FP64, fixed order), the PCG inner product
// <r, M^-1 r>.
template <int LG, bool DOT>
__global__ void __launch_bounds__(FBME_BLOCK)
k_lr_apply(const float2* __restrict__ tab, const float2* __restrict__ v, const float2* __restrict__ corr,
           const float* __restrict__ bg, int m1, int m2, int rhs, float2* __restrict__ out,
           double* __restrict__ part) {
  __shared__ double red[FBME_BLOCK / 64][LG];
  __shared__ float2 cs[LG][RANK];
  const int b = blockIdx.z;
  const int r0 = blockIdx.y * LG;
  const int nl = min(LG, rhs - r0);
  for (int e = threadIdx.x; e < nl * RANK; e += FBME_BLOCK)
    cs[e / RANK][e % RANK] = corr[((long long)b * rhs + r0) * RANK + e];
  __syncthreads();
  const long long m = (long long)m1 * m2;
  const long long i = (long long)blockIdx.x * FBME_BLOCK + threadIdx.x;
  const bool live = i < m;
  const int a1 = live ? (int)(i / m2) : 0;
  const int a2 = live ? (int)(i - (long long)a1 * m2) : 0;
  const float2* U = tab + (long long)b * (m1 + m2) * RANK;
  const float2* V = U + (long long)m1 * RANK;
  float2 f[RANK];
  const float4* u4 = reinterpret_cast<const float4*>(U + (long long)a1 * RANK);
  const float4* v4 = reinterpret_cast<const float4*>(V + (long long)a2 * RANK);
#pragma unroll
  for (int q = 0; q < RANK / 2; ++q) {
    const float4 uu = u4[q], vv = v4[q];
    f[2 * q] = cmul(make_float2(uu.x, uu.y), make_float2(vv.x, vv.y));
    f[2 * q + 1] = cmul(make_float2(uu.z, uu.w), make_float2(vv.z, vv.w));
  }
  const float ib = 1.0f / bg[b];
  const long long lane0 = (long long)b * rhs + r0;
  double dacc[LG];
#pragma unroll
  for (int j = 0; j < LG; ++j) {
    dacc[j] = 0.0;
    if (j < nl && live) {
      float ax = 0.f, ay = 0.f;
#pragma unroll
      for (int q = 0; q < RANK; ++q) {
        const float2 c = cs[j][q];
        ax += f[q].x * c.x - f[q].y * c.y;
        ay += f[q].x * c.y + f[q].y * c.x;
      }
      const float2 vv = v[(lane0 + j) * m + i];
      const float2 o = make_float2(vv.x * ib - ax, vv.y * ib - ay);
      out[(lane0 + j) * m + i] = o;
      if (DOT) dacc[j] = (double)vv.x * o.x + (double)vv.y * o.y;
    }
  }
  if (DOT) {
    const int wave = threadIdx.x >> 6, wl = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < LG; ++j) {
      const double w = wave_sum(dacc[j]);
      if (wl == 0) red[wave][j] = w;
    }
    __syncthreads();
    if (threadIdx.x < nl) {
      double sum = 0.0;
      for (int w = 0; w < FBME_BLOCK / 64; ++w) sum += red[w][threadIdx.x];
      part[(lane0 + threadIdx.x) * gridDim.x + blockIdx.x] = sum;
    }
  }
}

// ------------------------------------------- grid-diagonal background
// Background of the low-rank preconditioner taken on the grid instead of as a
// scalar: B^-1 = A diag(w) A^H with w > 0, so B^-1 is Hermitian positive
// definite for any w (A A^H = N I).  For a uniform variance g,
// Q = (noise + N g) I exactly, so w = 1 / (N (noise + N g)) makes B = Q; the
// closed form w_n = 1 / (N (noise + N gamma_T,n)) takes gamma_T = gamma
// without the RANK exact atoms of the Woodbury part.  A per-cell log-scale o
// (the learned correction, clamped to +-cl) multiplies w by exp(o).
__global__ void k_gdiag_weights(const float* __restrict__ gamma, const float* __restrict__ noise,
                                const float* __restrict__ logscale, float cl, long long n,
                                float* __restrict__ w) {
  const int b = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long e = (long long)b * n + i;
  const double nn = (double)n;
  double v = 1.0 / (nn * ((double)noise[b] + nn * (double)fmaxf(gamma[e], 0.f)));
  if (logscale) v *= exp((double)fminf(fmaxf(logscale[e], -cl), cl));
  w[e] = (float)v;
}

// The RANK exact atoms carry no background variance: w = 1 / (N noise) there.
__global__ void k_gdiag_atoms(const int* __restrict__ idx, const float* __restrict__ noise,
                              const float* __restrict__ logscale, float cl, long long n, float* __restrict__ w) {
  const int b = blockIdx.x, k = threadIdx.x;
  if (k >= RANK) return;
  const long long e = (long long)b * n + idx[b * RANK + k];
  double v = 1.0 / ((double)n * (double)noise[b]);
  if (logscale) v *= exp((double)fminf(fmaxf(logscale[e], -cl), cl));
  w[e] = (float)v;
}

// Factor columns as lanes: out[(b, k), m] = U[m1, k] V[m2, k] (the phase
// tables of k_lr_tables), so that B^-1 F can go through the matvec passes.
__global__ void k_lr_columns(const float2* __restrict__ tab, int m1, int m2, float2* __restrict__ out) {
  const int b = blockIdx.z, k = blockIdx.y;
  const long long m = (long long)m1 * m2;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int a1 = (int)(i / m2), a2 = (int)(i - (long long)a1 * m2);
  const float2* U = tab + (long long)b * (m1 + m2) * RANK;
  const float2* V = U + (long long)m1 * RANK;
  out[((long long)b * RANK + k) * m + i] = cmul(U[(long long)a1 * RANK + k], V[(long long)a2 * RANK + k]);
}

// One wavefront per scene: C = I + F^H B^-1 F from the projection partials of
// the RANK lanes B^-1 F_l onto the atoms (part[(b RANK + l) chunks + c][k] =
// partial of (F^H B^-1 F)[k][l], summed over the chunks in order), then
// C^-1 by the Cholesky inverse of the low-rank setup.  C >= I, so the
// factorisation cannot fail in exact arithmetic; a failure is reported in
// status[b].
__global__ void __launch_bounds__(64)
k_gdiag_system(const double2* __restrict__ part, int chunks, float2* __restrict__ cinv, int* __restrict__ status) {
  __shared__ double2 a[RANK][RANK + 1];
  __shared__ double2 xs[RANK][RANK + 1];
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  for (int e = lane; e < RANK * RANK; e += 64) {
    const int l = e / RANK, k = e - l * RANK;
    const double2* pp = part + ((long long)(b * RANK + l) * chunks) * RANK + k;
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < chunks; ++c) {
      const double2 v = pp[(long long)c * RANK];
      s.x += v.x;
      s.y += v.y;
    }
    a[k][l] = make_double2((k == l ? 1.0 : 0.0) + s.x, s.y);
  }
  __syncthreads();
  const int fail = lr_chol_inverse(a, xs, lane, cinv + (long long)b * RANK * RANK);
  if (lane == 0 && status && fail) status[b] = 1;
}

// out = b - U corr with U = B^-1 F stored as RANK lanes per scene (the
// grid-diagonal preconditioner (B + F F^H)^-1 v = b - U C^-1 U^H v,
// b = B^-1 v, U^H v = F^H b): one sample per thread, the RANK column values
// of the sample loaded once and applied to every lane of the group.  DOT:
// part[lane, block] = sum Re(conj(vd) out) (FP64, fixed order), the PCG inner
// product <r, M^-1 r> with vd = r.
template <int LG, bool DOT>
__global__ void __launch_bounds__(FBME_BLOCK)
k_lr_apply_stored(const float2* __restrict__ ucols, const float2* __restrict__ b, const float2* __restrict__ vd,
                  const float2* __restrict__ corr, long long m, int rhs, float2* __restrict__ out,
                  double* __restrict__ part) {
  __shared__ double red[FBME_BLOCK / 64][LG];
  __shared__ float2 cs[LG][RANK];
  const int sc = blockIdx.z;
  const int r0 = blockIdx.y * LG;
  const int nl = min(LG, rhs - r0);
  for (int e = threadIdx.x; e < nl * RANK; e += FBME_BLOCK)
    cs[e / RANK][e % RANK] = corr[((long long)sc * rhs + r0) * RANK + e];
  __syncthreads();
  const long long i = (long long)blockIdx.x * FBME_BLOCK + threadIdx.x;
  const bool live = i < m;
  float2 f[RANK];
  const float2* U = ucols + (long long)sc * RANK * m;
#pragma unroll
  for (int q = 0; q < RANK; ++q) f[q] = live ? U[(long long)q * m + i] : make_float2(0.f, 0.f);
  const long long lane0 = (long long)sc * rhs + r0;
  double dacc[LG];
#pragma unroll
  for (int j = 0; j < LG; ++j) {
    dacc[j] = 0.0;
    if (j < nl && live) {
      float ax = 0.f, ay = 0.f;
#pragma unroll
      for (int q = 0; q < RANK; ++q) {
        const float2 c = cs[j][q];
        ax += f[q].x * c.x - f[q].y * c.y;
        ay += f[q].x * c.y + f[q].y * c.x;
      }
      const float2 bb = b[(lane0 + j) * m + i];
      const float2 o = make_float2(bb.x - ax, bb.y - ay);
      out[(lane0 + j) * m + i] = o;
      if (DOT) {
        const float2 vv = vd[(lane0 + j) * m + i];
        dacc[j] = (double)vv.x * o.x + (double)vv.y * o.y;
      }
    }
  }
  if (DOT) {
    const int wave = threadIdx.x >> 6, wl = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < LG; ++j) {
      const double w = wave_sum(dacc[j]);
      if (wl == 0) red[wave][j] = w;
    }
    __syncthreads();
    if (threadIdx.x < nl) {
      double sum = 0.0;
      for (int w = 0; w < FBME_BLOCK / 64; ++w) sum += red[w][threadIdx.x];
      part[(lane0 + threadIdx.x) * gridDim.x + blockIdx.x] = sum;
    }
  }
}

// ------------------------------------------------------------- lane algebra
// part[lane, blk] = sum Re(conj(a) b) over one block of samples (FP64).
__global__ void __launch_bounds__(FBME_BLOCK)
k_lane_dot(const float2* __restrict__ a, const float2* __restrict__ b, int m, double* __restrict__ part) {
  __shared__ double red[8];
  const int lanev = blockIdx.y;
  double acc = 0.0;
  for (int i = blockIdx.x * FBME_BLOCK + threadIdx.x; i < m; i += gridDim.x * FBME_BLOCK) {
    const float2 u = a[(long long)lanev * m + i], w = b[(long long)lanev * m + i];
    acc += (double)u.x * w.x + (double)u.y * w.y;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[(long long)lanev * gridDim.x + blockIdx.x] = acc;
}

// out[l] = sum_i part[l, i]: one wavefront per lane, lane t of the wave
// summing i = t, t + 64, ... and a fixed shuffle tree (deterministic).
__global__ void __launch_bounds__(64) k_lane_finish(const double* __restrict__ part, int blocks, int lanes,
                                                    double* __restrict__ out) {
  const int l = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < blocks; i += 64) s += part[(long long)l * blocks + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[l] = s;
}

// ---------------------------------------------------------- observation masks
// Vectors keep all M1 M2 samples with zeros on the unobserved ones; the
// operator restricted to the observed rows (operators.py:57-75, 473-496) is
// the full one followed by the mask.

// out[lane, i] *= mask[i]; with part, also the block partials of
// Re <a, out> for k_lane_finish.
__global__ void __launch_bounds__(FBME_BLOCK)
k_mask_lanes(const unsigned char* __restrict__ mask, const float2* __restrict__ a, float2* __restrict__ out, int m,
             double* __restrict__ part) {
  __shared__ double red[8];
  const int lanev = blockIdx.y;
  double acc = 0.0;
  for (int i = blockIdx.x * FBME_BLOCK + threadIdx.x; i < m; i += gridDim.x * FBME_BLOCK) {
    const long long g = (long long)lanev * m + i;
    if (!mask[i]) {
      out[g] = make_float2(0.f, 0.f);
    } else if (part) {
      const float2 u = a[g], w = out[g];
      acc += (double)u.x * w.x + (double)u.y * w.y;
    }
  }
  if (!part) return;
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[(long long)lanev * gridDim.x + blockIdx.x] = acc;
}

// t[m1, d2] = sum over observed m2 of exp(+2 pi i m2 d2 / N2), FP64, the
// phase reduced modulo N2 exactly in integers.  One thread per (m1, d2).
__global__ void k_mask_rows(const unsigned char* __restrict__ mask, int m1, int m2, int n2, double2* __restrict__ t) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m1 * n2) return;
  const int r = (int)(e / n2), d2 = (int)(e - (long long)r * n2);
  double sx = 0.0, sy = 0.0;
  for (int j = 0; j < m2; ++j) {
    if (!mask[(long long)r * m2 + j]) continue;
    const long long ph = ((long long)j * d2) % n2;
    double sn, cs;
    sincospi(2.0 * (double)ph / (double)n2, &sn, &cs);
    sx += cs;
    sy += sn;
  }
  t[e] = make_double2(sx, sy);
}

// s[d1, d2] = sum_m1 exp(+2 pi i m1 d1 / N1) t[m1, d2].  One thread per
// (d1, d2); consecutive threads read consecutive d2.
__global__ void k_mask_cols(const double2* __restrict__ t, int m1, int n1, int n2, double2* __restrict__ s) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n1 * n2) return;
  const int d1 = (int)(e / n2), d2 = (int)(e - (long long)d1 * n2);
  double sx = 0.0, sy = 0.0;
  for (int r = 0; r < m1; ++r) {
    const long long ph = ((long long)r * d1) % n1;
    double sn, cs;
    sincospi(2.0 * (double)ph / (double)n1, &sn, &cs);
    const double2 v = t[(long long)r * n2 + d2];
    sx += cs * v.x - sn * v.y;
    sy += cs * v.y + sn * v.x;
  }
  s[e] = make_double2(sx, sy);
}

// x += alpha d, r -= alpha q with alpha from (rho, curv, active)
// (solver.py:5656-5662).
__global__ void k_pcg_update(float2* __restrict__ x, float2* __restrict__ r, const float2* __restrict__ d,
                             const float2* __restrict__ q, const double* __restrict__ rho,
                             const double* __restrict__ curv, int* __restrict__ active, int m) {
  const int lanev = blockIdx.y;
  const double cv = curv[lanev], rh = rho[lanev];
  const bool valid = active[lanev] && isfinite(cv) && isfinite(rh) && cv > F32_TINY_D && rh > F32_TINY_D;
  const float alpha = valid ? (float)(rh / fmax(cv, F32_TINY_D)) : 0.f;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && alpha != 0.f) {
    const long long g = (long long)lanev * m + i;
    const float2 dv = d[g], qv = q[g];
    float2 xv = x[g], rv = r[g];
    xv.x += alpha * dv.x;
    xv.y += alpha * dv.y;
    rv.x -= alpha * qv.x;
    rv.y -= alpha * qv.y;
    x[g] = xv;
    r[g] = rv;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    // the flag is read above by every block; a separate pass commits it
  }
}

__global__ void k_pcg_commit_valid(const double* __restrict__ rho, const double* __restrict__ curv,
                                   int* __restrict__ active, int lanes) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= lanes) return;
  const double cv = curv[l], rh = rho[l];
  active[l] = active[l] && isfinite(cv) && isfinite(rh) && cv > F32_TINY_D && rh > F32_TINY_D;
}

// d = z + beta d (masked), rho and active updated (solver.py:5709-5719).
__global__ void k_pcg_direction(float2* __restrict__ d, const float2* __restrict__ z,
                                const double* __restrict__ rho, const double* __restrict__ nrho,
                                const int* __restrict__ active, int m, int restart) {
  const int lanev = blockIdx.y;
  const double nr = nrho[lanev], rh = rho[lanev];
  const bool vu = active[lanev] && isfinite(nr) && nr > F32_TINY_D;
  const float beta = (vu && !restart) ? (float)(nr / fmax(rh, F32_TINY_D)) : 0.f;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const long long g = (long long)lanev * m + i;
  if (vu) {
    const float2 zv = z[g];
    float2 dv = restart ? make_float2(0.f, 0.f) : d[g];
    d[g] = make_float2(zv.x + beta * dv.x, zv.y + beta * dv.y);
  } else {
    d[g] = make_float2(0.f, 0.f);
  }
}

__global__ void k_pcg_commit_rho(double* __restrict__ rho, const double* __restrict__ nrho,
                                 int* __restrict__ active, int lanes) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= lanes) return;
  const double nr = nrho[l];
  const bool vu = active[l] && isfinite(nr) && nr > F32_TINY_D;
  if (vu) rho[l] = nr;
  active[l] = vu;
}

// Variants of the update and direction passes that take their lane inner
// product as per-block partials of the producing pass (summed here in a fixed
// order, so no separate reduction launch) and fold the validity commit of
// k_pcg_commit_valid into the rho commit: the lane logic is the same as
// update -> commit_valid -> direction -> commit_rho.
__device__ __forceinline__ double lane_parts(const double* __restrict__ part, int nparts, int lanev) {
  // up to 16 partials (M2 <= 256) loaded together, then added in index order
  const double* p = part + (long long)lanev * nparts;
  double v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = i < nparts ? p[i] : 0.0;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) s += v[i];
  for (int i = 16; i < nparts; ++i) s += p[i];
  return s;
}

__device__ __forceinline__ bool pcg_valid(double cv, double rh) {
  return isfinite(cv) && isfinite(rh) && cv > F32_TINY_D && rh > F32_TINY_D;
}

// x += alpha d, r -= alpha q with curv = sum of the matvec's partials; block
// (0, lane) stores curv[lane].  RR: also rr_part[lane, block] = ||r||^2 over
// the block's samples after the update (FP64, fixed tree order).
template <bool RR>
__global__ void k_pcg_update_parts(float2* __restrict__ x, float2* __restrict__ r, const float2* __restrict__ d,
                                   const float2* __restrict__ q, const double* __restrict__ rho,
                                   const double* __restrict__ part, int nparts, double* __restrict__ curv,
                                   const int* __restrict__ active, int m, double* __restrict__ rr_part) {
  const int lanev = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const long long g = (long long)lanev * m + i;
  // vector loads issued before the lane scalars are reduced (lanes already
  // converged load nothing)
  const bool act = active[lanev] != 0;
  float2 dv = make_float2(0.f, 0.f), qv = dv, xv = dv, rv = dv;
  if (i < m && act) {
    dv = d[g];
    qv = q[g];
    xv = x[g];
    rv = r[g];
  }
  const double cv = lane_parts(part, nparts, lanev), rh = rho[lanev];
  if (blockIdx.x == 0 && threadIdx.x == 0) curv[lanev] = cv;
  const bool valid = act && pcg_valid(cv, rh);
  const float alpha = valid ? (float)(rh / fmax(cv, F32_TINY_D)) : 0.f;
  if (i < m && alpha != 0.f) {
    xv.x += alpha * dv.x;
    xv.y += alpha * dv.y;
    rv.x -= alpha * qv.x;
    rv.y -= alpha * qv.y;
    x[g] = xv;
    r[g] = rv;
  }
  if (RR) {
    __shared__ double red[4];
    const double rr = block_sum((double)rv.x * rv.x + (double)rv.y * rv.y, red);
    if (threadIdx.x == 0) rr_part[(long long)lanev * gridDim.x + blockIdx.x] = rr;
  }
}

// d = z + beta d with nrho = sum of the preconditioner's partials and the
// lane active iff it was, its update was valid and nrho is; block (0, lane)
// stores nrho[lane].
__global__ void k_pcg_direction_parts(float2* __restrict__ d, const float2* __restrict__ z,
                                      const double* __restrict__ rho, const double* __restrict__ curv,
                                      const double* __restrict__ part, int nparts, double* __restrict__ nrho,
                                      const int* __restrict__ active, int m) {
  const int lanev = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const long long g = (long long)lanev * m + i;
  // vector loads issued before the lane scalars are reduced (lanes already
  // converged load nothing)
  const bool act = active[lanev] != 0;
  float2 zv = make_float2(0.f, 0.f), dv = zv;
  if (i < m && act) {
    zv = z[g];
    dv = d[g];
  }
  const double nr = lane_parts(part, nparts, lanev), rh = rho[lanev];
  if (blockIdx.x == 0 && threadIdx.x == 0) nrho[lanev] = nr;
  const bool vu = act && pcg_valid(curv[lanev], rh) && isfinite(nr) && nr > F32_TINY_D;
  const float beta = vu ? (float)(nr / fmax(rh, F32_TINY_D)) : 0.f;
  if (i >= m) return;
  d[g] = vu ? make_float2(zv.x + beta * dv.x, zv.y + beta * dv.y) : make_float2(0.f, 0.f);
}

__global__ void k_pcg_commit_parts(double* __restrict__ rho, const double* __restrict__ curv,
                                   const double* __restrict__ nrho, int* __restrict__ active, int lanes) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= lanes) return;
  const double nr = nrho[l];
  const bool vu = active[l] && pcg_valid(curv[l], rho[l]) && isfinite(nr) && nr > F32_TINY_D;
  if (vu) rho[l] = nr;
  active[l] = vu;
}

// r = rhs - Qx (true residual at a check iteration).
__global__ void k_sub(const float2* __restrict__ a, const float2* __restrict__ b, float2* __restrict__ out,
                      long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2(a[i].x - b[i].x, a[i].y - b[i].y);
}

// Adaptive/fixed residual test at a check iteration (solver.py:5668-5690):
// rel = ||r|| / init; active = [active &] finite(rel) & rel > tol.
__global__ void k_pcg_check(const double* __restrict__ rr, const double* __restrict__ init, int* __restrict__ active,
                            double* __restrict__ best_rel, int* __restrict__ improved, float tol, int fixed,
                            int lanes) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= lanes) return;
  const double rel = sqrt(fmax(rr[l], 0.0)) / fmax(sqrt(fmax(init[l], 0.0)), F32_EPS_D);
  if (fixed) {
    const bool better = rel < best_rel[l];
    improved[l] = better;
    if (better) best_rel[l] = rel;
  }
  const bool ra = isfinite(rel) && rel > (double)tol;
  active[l] = fixed ? (active[l] && ra) : ra;
}

// ------------------------------------------------- Hutchinson and EM updates
// dd = max(mean_p Re(L_p conj(LP_p)), 0); mu = gamma L_0;
// v = max(gamma - gamma^2 dd, 0); partial sums of clamp(gamma dd, 0, 1).
__global__ void __launch_bounds__(FBME_BLOCK)
k_hutch(const float2* __restrict__ lifted, const float2* __restrict__ lprobes, const float* __restrict__ gamma,
        int n, int probes, int used, float2* __restrict__ mean, float* __restrict__ var, float* __restrict__ dd_out,
        double* __restrict__ part) {
  __shared__ double red[8];
  const int b = blockIdx.y;
  const int i = blockIdx.x * FBME_BLOCK + threadIdx.x;
  double rel = 0.0;
  if (i < n) {
    const int rhs = probes + 1;
    double acc = 0.0;
    for (int p = 0; p < used; ++p) {
      const float2 u = lifted[((long long)b * rhs + 1 + p) * n + i];
      const float2 w = lprobes[((long long)b * probes + p) * n + i];
      acc += (double)u.x * w.x + (double)u.y * w.y;
    }
    const float dd = fmaxf((float)(acc / used), 0.f);
    const float g = gamma[(long long)b * n + i];
    const float2 l0 = lifted[((long long)b * rhs) * n + i];
    mean[(long long)b * n + i] = make_float2(g * l0.x, g * l0.y);
    var[(long long)b * n + i] = fmaxf(g - g * g * dd, 0.f);
    dd_out[(long long)b * n + i] = dd;
    rel = fminf(fmaxf(g * dd, 0.f), 1.f);
  }
  rel = block_sum(rel, red);
  if (threadIdx.x == 0) part[(long long)b * gridDim.x + blockIdx.x] = rel;
}

// Relevance projection, MacKay update with damping and floor; optional
// active-set mask keeps gamma off the set (solver.py:1669-1677, 1776-1798,
// 1700-1712).  Partial sums of the projected relevance (dof).
__global__ void __launch_bounds__(FBME_BLOCK)
k_cg_em(float* __restrict__ gamma, const float2* __restrict__ mean, const float* __restrict__ dd,
        const double* __restrict__ rel_sum, const float* __restrict__ floor_, const unsigned char* __restrict__ mask,
        int n, int m, float hyper_shape, float hyper_rate, float max_prec, float keep, int update,
        double* __restrict__ part) {
  __shared__ double red[8];
  const int b = blockIdx.y;
  const int i = blockIdx.x * FBME_BLOCK + threadIdx.x;
  const double raw = rel_sum[b];
  const double cap = (double)m * (1.0 - 10.0 * F32_EPS_D);
  const float scale = (float)fmin(1.0, cap / fmax(raw, F32_EPS_D));
  double dof = 0.0;
  if (i < n) {
    const long long g = (long long)b * n + i;
    const float gm = gamma[g];
    const float rel = fminf(fmaxf(gm * dd[g], 0.f), 1.f) * scale;
    dof = rel;
    const float2 mu = mean[g];
    float prec = (rel + hyper_shape) / (mu.x * mu.x + mu.y * mu.y + hyper_rate);
    prec = fminf(fmaxf(prec, 1.0f / max_prec), max_prec);
    float ng = 1.0f / prec;
    if (mask && !mask[g]) ng = gm;
    if (update) gamma[g] = fmaxf(keep * gm + (1.0f - keep) * ng, floor_[b]);
  }
  dof = block_sum(dof, red);
  if (threadIdx.x == 0) part[(long long)b * gridDim.x + blockIdx.x] = dof;
}

// ------------------------------------------------- pattern-coupled prior
// Fixed coupling beta (solver.py:1100-1150, 1489-1590, 2396-2720): the
// precision of cell i is eta_i = alpha_i + beta sum_{j in N(i)} alpha_j over
// the four axial neighbours (circular: wrapped; else zero outside), added in
// the order (i-1, j), (i+1, j), (i, j-1), (i, j+1).
__device__ __forceinline__ float nsum4(const float* __restrict__ f, int n1, int n2, int i, int j, int circular) {
  float acc = 0.f;
  const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int a = i + di[q], c = j + dj[q];
    if (circular) {
      a = (a + n1) % n1;
      c = (c + n2) % n2;
    } else if (a < 0 || a >= n1 || c < 0 || c >= n2) {
      continue;
    }
    acc += f[(long long)a * n2 + c];
  }
  return acc;
}

// pattern = (1 / gamma) / (1 + beta degree) (solver.py:2674-2687)
__global__ void k_spatial_init(const float* __restrict__ gamma, int n1, int n2, float beta, int circular,
                               float* __restrict__ pattern) {
  const int b = blockIdx.y;
  const long long n = (long long)n1 * n2;
  const long long i = (long long)blockIdx.x * FBME_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int r = (int)(i / n2), c = (int)(i - (long long)r * n2);
  int deg = 0;
  if (circular) {
    deg = 4;
  } else {
    deg = (r > 0) + (r < n1 - 1) + (c > 0) + (c < n2 - 1);
  }
  pattern[b * n + i] = (1.0f / gamma[b * n + i]) / (1.0f + beta * (float)deg);
}

// gamma = max(1 / clamp(pattern + beta N(pattern)), floor) (solver.py:2662-2672, 1590-1610)
__global__ void k_spatial_gamma(const float* __restrict__ pattern, int n1, int n2, float beta, int circular,
                                float max_prec, const float* __restrict__ floor_, float* __restrict__ gamma) {
  const int b = blockIdx.y;
  const long long n = (long long)n1 * n2;
  const long long i = (long long)blockIdx.x * FBME_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int r = (int)(i / n2), c = (int)(i - (long long)r * n2);
  const float* pb = pattern + b * n;
  const float eta = pb[i] + beta * nsum4(pb, n1, n2, r, c, circular);
  const float prec = fminf(fmaxf(eta, 1.0f / max_prec), max_prec);
  gamma[b * n + i] = fmaxf(1.0f / prec, floor_[b]);
}

// second = |mu|^2 + v
__global__ void k_spatial_second(const float2* __restrict__ mean, const float* __restrict__ var, long long n,
                                 float* __restrict__ second) {
  const int b = blockIdx.y;
  const long long i = (long long)blockIdx.x * FBME_BLOCK + threadIdx.x;
  if (i >= n) return;
  const float2 mu = mean[b * n + i];
  second[b * n + i] = mu.x * mu.x + mu.y * mu.y + var[b * n + i];
}

// pattern <- keep pattern + (1 - keep) clamp((1 + a) / (second + beta N(second) + b))
// (_pattern_precision_update, solver.py:2690-2720, damped at 1575-1580)
__global__ void k_spatial_pattern(const float* __restrict__ second, int n1, int n2, float beta, int circular,
                                  float hyper_shape, float hyper_rate, float max_prec, float keep,
                                  float* __restrict__ pattern) {
  const int b = blockIdx.y;
  const long long n = (long long)n1 * n2;
  const long long i = (long long)blockIdx.x * FBME_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int r = (int)(i / n2), c = (int)(i - (long long)r * n2);
  const float* sb = second + b * n;
  const float coupled = sb[i] + beta * nsum4(sb, n1, n2, r, c, circular);
  float base = (1.0f + hyper_shape) / (coupled + hyper_rate);
  base = fminf(fmaxf(base, 1.0f / max_prec), max_prec);
  pattern[b * n + i] = keep * pattern[b * n + i] + (1.0f - keep) * base;
}

// ------------------------------------------------------------- joint sparsity
// Snapshots b = 0..B-1 share one variance per cell (M-SBL).  gamma is kept
// as B identical rows so the per-snapshot kernels read it unchanged.

// gamma[b, i] = max(mean_l(blend |bp_l|^2 / ||bp_l||^2 P_l + (1 - blend) P_l / N),
// max_l floor_l) for every b (solver.py:1071-1099).
__global__ void k_joint_gamma0(const float2* __restrict__ bp, const double* __restrict__ bpe,
                               const float* __restrict__ power, const float* __restrict__ floor_, float blend,
                               int batch, int n, float* __restrict__ gamma) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc = 0.f, fmax_ = 0.f;
  for (int b = 0; b < batch; ++b) {
    const float2 v = bp[(long long)b * n + i];
    const float p = power[b];
    const float structured = (v.x * v.x + v.y * v.y) / fmaxf((float)bpe[b], (float)F32_EPS_D) * p;
    acc += blend * structured + (1.0f - blend) * (p / (float)n);
    fmax_ = fmaxf(fmax_, floor_[b]);
  }
  const float g = fmaxf(acc / (float)batch, fmax_);
  for (int b = 0; b < batch; ++b) gamma[(long long)b * n + i] = g;
}

// noise[b] = mean_l noise_l for every b (solver.py:1157-1159).
__global__ void k_joint_noise0(float* __restrict__ noise, int batch) {
  if (threadIdx.x != 0) return;
  float s = 0.f;
  for (int b = 0; b < batch; ++b) s += noise[b];
  s /= (float)batch;
  for (int b = 0; b < batch; ++b) noise[b] = s;
}

// Shared-noise posterior: lanes 0..B-1 of `lifted` are A^H Q^-1 y_b, lanes
// B..B+P-1 the probes' A^H Q^-1 p (one system, solver.py:3039-3153).
// mean[b] = gamma lifted[b]; the Hutchinson diagonal, variance and relevance
// partials are the same for every snapshot and written for each.
__global__ void __launch_bounds__(FBME_BLOCK)
k_hutch_shared(const float2* __restrict__ lifted, const float2* __restrict__ lprobes,
               const float* __restrict__ gamma, int batch, int n, int probes, float2* __restrict__ mean,
               float* __restrict__ var, float* __restrict__ dd_out, double* __restrict__ part) {
  __shared__ double red[8];
  const int i = blockIdx.x * FBME_BLOCK + threadIdx.x;
  double rel = 0.0;
  if (i < n) {
    double acc = 0.0;
    for (int p = 0; p < probes; ++p) {
      const float2 u = lifted[((long long)batch + p) * n + i];
      const float2 w = lprobes[(long long)p * n + i];
      acc += (double)u.x * w.x + (double)u.y * w.y;
    }
    const float dd = fmaxf((float)(acc / probes), 0.f);
    const float g = gamma[i];
    const float v = fmaxf(g - g * g * dd, 0.f);
    for (int b = 0; b < batch; ++b) {
      const long long gi = (long long)b * n + i;
      const float2 l = lifted[gi];
      mean[gi] = make_float2(g * l.x, g * l.y);
      var[gi] = v;
      dd_out[gi] = dd;
    }
    rel = fminf(fmaxf(g * dd, 0.f), 1.f);
  }
  rel = block_sum(rel, red);
  if (threadIdx.x == 0)
    for (int b = 0; b < batch; ++b) part[(long long)b * gridDim.x + blockIdx.x] = rel;
}

// Projected relevance (dof partials per snapshot) and the complex M-SBL EM
// update gamma <- (mean_b(|mu_b|^2 + v_b) + rate) / (1 + shape), clamped to
// [max_b floor_b, max_prec], damped and floored (solver.py:1653-1666,
// 1776-1798).
__global__ void __launch_bounds__(FBME_BLOCK)
k_cg_em_joint(float* __restrict__ gamma, const float2* __restrict__ mean, const float* __restrict__ var,
              const float* __restrict__ dd, const double* __restrict__ rel_sum, const float* __restrict__ floor_,
              int batch, int n, int m, float hyper_shape, float hyper_rate, float max_prec, float keep, int update,
              double* __restrict__ part) {
  __shared__ double red[8];
  const int i = blockIdx.x * FBME_BLOCK + threadIdx.x;
  const double cap = (double)m * (1.0 - 10.0 * F32_EPS_D);
  float fmax_ = 0.f;
  for (int b = 0; b < batch; ++b) fmax_ = fmaxf(fmax_, floor_[b]);
  float gm = 0.f, second = 0.f;
  if (i < n) {
    gm = gamma[i];
    for (int b = 0; b < batch; ++b) {
      const float2 mu = mean[(long long)b * n + i];
      second += mu.x * mu.x + mu.y * mu.y + var[(long long)b * n + i];
    }
    second /= (float)batch;
  }
  for (int b = 0; b < batch; ++b) {
    double rel = 0.0;
    if (i < n) {
      const float scale = (float)fmin(1.0, cap / fmax(rel_sum[b], F32_EPS_D));
      rel = fminf(fmaxf(gamma[(long long)b * n + i] * dd[(long long)b * n + i], 0.f), 1.f) * scale;
    }
    rel = block_sum(rel, red);
    if (threadIdx.x == 0) part[(long long)b * gridDim.x + blockIdx.x] = rel;
  }
  if (update && i < n) {
    const float ng = fminf(fmaxf((second + hyper_rate) / (1.0f + hyper_shape), fmax_), max_prec);
    const float g = fmaxf(keep * gm + (1.0f - keep) * ng, fmax_);
    for (int b = 0; b < batch; ++b) gamma[(long long)b * n + i] = g;
  }
}

__global__ void k_seg_sum(const double* __restrict__ part, int blocks, double* __restrict__ out,
                          float* __restrict__ out_f) {
  const int b = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < blocks; i += 64) s += part[(long long)b * blocks + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) {
    out[b] = s;
    if (out_f) out_f[b] = (float)s;
  }
}

__global__ void k_masked_copy(float2* __restrict__ dst, const float2* __restrict__ src, const int* __restrict__ mask,
                              int m) {
  const int l = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && mask[l]) dst[(long long)l * m + i] = src[(long long)l * m + i];
}

__global__ void k_twiddle(float2* __restrict__ tw, int len) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double s, c;
  sincospi(2.0 * (double)i / (double)len, &s, &c);
  tw[i] = make_float2((float)c, (float)-s);
}

// Partial sums of gamma (for the preconditioner background) and of
// ||mu - mu_prev||^2, ||mu||^2 (relative change, solver.py:1813-1826).
__global__ void __launch_bounds__(FBME_BLOCK)
k_sums(const float* __restrict__ gamma, const float2* __restrict__ mean, const float2* __restrict__ prev, int n,
       double* __restrict__ part) {
  __shared__ double red[8];
  const int b = blockIdx.y;
  const int i = blockIdx.x * FBME_BLOCK + threadIdx.x;
  double sg = 0.0, dm = 0.0, nm = 0.0;
  if (i < n) {
    const long long g = (long long)b * n + i;
    sg = gamma[g];
    if (mean) {
      const float2 u = mean[g], w = prev[g];
      dm = (double)(u.x - w.x) * (u.x - w.x) + (double)(u.y - w.y) * (u.y - w.y);
      nm = (double)u.x * u.x + (double)u.y * u.y;
    }
  }
  block_sum2(sg, dm, red);
  nm = block_sum(nm, red);
  if (threadIdx.x == 0) {
    const long long slot = ((long long)b * gridDim.x + blockIdx.x) * 3;
    part[slot] = sg;
    part[slot + 1] = dm;
    part[slot + 2] = nm;
  }
}

__global__ void k_sums_finish(const double* __restrict__ part, int blocks, double* __restrict__ out) {
  const int b = blockIdx.x;
  double a = 0.0, c = 0.0, d = 0.0;
  for (int i = threadIdx.x; i < blocks; i += 64) {
    a += part[((long long)b * blocks + i) * 3];
    c += part[((long long)b * blocks + i) * 3 + 1];
    d += part[((long long)b * blocks + i) * 3 + 2];
  }
  a = wave_sum(a);
  c = wave_sum(c);
  d = wave_sum(d);
  if (threadIdx.x == 0) {
    out[b * 3] = a;
    out[b * 3 + 1] = c;
    out[b * 3 + 2] = d;
  }
}


// Initial scalars (solver.py:1063-1070, 1151-1156): power = max(||y||^2/M,
// eps), floor = max(1e-12 power, tiny), noise = max(frac power, floor).
__global__ void k_cg_scalars(const double* __restrict__ ye, int m, float min_variance, float noise_fraction,
                             float* __restrict__ power, float* __restrict__ floor_, float* __restrict__ noise,
                             int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const float p = fmaxf((float)(ye[b] / m), (float)F32_EPS_D);
  const float f = fmaxf(p * min_variance, (float)F32_TINY_D);
  power[b] = p;
  floor_[b] = f;
  noise[b] = fmaxf(noise_fraction * p, f);
}

// Backprojection initialisation of gamma (solver.py:1071-1089).
__global__ void k_cg_gamma0(const float2* __restrict__ bp, const double* __restrict__ bpe,
                            const float* __restrict__ power, const float* __restrict__ floor_, float blend, int n,
                            float* __restrict__ gamma) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = bp[(long long)b * n + i];
  const float p = power[b];
  const float structured = (v.x * v.x + v.y * v.y) / fmaxf((float)bpe[b], (float)F32_EPS_D) * p;
  const float uniform = p / (float)n;
  gamma[(long long)b * n + i] = fmaxf(blend * structured + (1.0f - blend) * uniform, floor_[b]);
}

__global__ void __launch_bounds__(FBME_BLOCK)
k_rel_sum(const float* __restrict__ gamma, const float* __restrict__ dd, int n, double* __restrict__ part) {
  __shared__ double red[8];
  const int b = blockIdx.y;
  const int i = blockIdx.x * FBME_BLOCK + threadIdx.x;
  double r = 0.0;
  if (i < n) r = fminf(fmaxf(gamma[(long long)b * n + i] * dd[(long long)b * n + i], 0.f), 1.f);
  r = block_sum(r, red);
  if (threadIdx.x == 0) part[(long long)b * gridDim.x + blockIdx.x] = r;
}

__global__ void k_reciprocal(const float* __restrict__ in, float* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = 1.0f / in[i];
}

inline int nblk(long long n, int bs = FBME_BLOCK) { return (int)((n + bs - 1) / bs); }

// Row passes: `want` rows per workgroup (0: as many as a 4-element-per-thread
// LDS tile holds), at most 64 and at most what a 16-element tile holds; the
// kernel instance is the smallest tile that fits the rows.
int rowpass_cap(int n) { return std::max(1, std::min(16 * FBME_BLOCK / (n + 1), 64)); }

template <bool INV>
void launch_rowpass(const float2* in, float2* out, int n, int len_in, int len_out, long long rows, unsigned code,
                    const float2* tw, hipStream_t st, int want_rows = 0) {
  const int def = std::max(1, std::min(4 * FBME_BLOCK / (n + 1), 64));
  const int tr = want_rows > 0 ? std::min(want_rows, rowpass_cap(n)) : def;
  const int grid = (int)((rows + tr - 1) / tr);
  const int need = tr * (n + 1);
#define FBME_ROWPASS(EPT, TWN)                                                                                \
  hipLaunchKernelGGL((k_rowpass<EPT, INV, TWN>), dim3(grid), dim3(FBME_BLOCK), 0, st, in, out, n, len_in, len_out, \
                     rows, tr, code, tw)
  if (n > 1024)
    FBME_ROWPASS(16, FBME_FFT_MAX);
  else if (need <= 4 * FBME_BLOCK)
    FBME_ROWPASS(4, 1024);
  else if (need <= 8 * FBME_BLOCK)
    FBME_ROWPASS(8, 1024);
  else
    FBME_ROWPASS(16, 1024);
#undef FBME_ROWPASS
}

// Columns per workgroup of the fused Toeplitz column pass: at most 16 and at
// most what a 24-element-per-thread LDS tile holds.
int toeplitz_col_cap(int e1, int e2) { return std::max(1, std::min({e2, 16, 24 * FBME_BLOCK / (e1 + 1)})); }

}  // namespace

// Scratch of fbme_topk32_noalloc: a fixed device array instead of a
// workspace argument, for callers that cannot allocate.  Each GPU holds its
// own copy at its own address; host code uses the address that
// hipGetSymbolAddress returns for the current device, never the array's name.
static __device__ __align__(256) unsigned long long scratch[FBME_TOPK32_NOALLOC_SCRATCH_BYTES / 8];

namespace {

constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mutex;
void* scratch_ptr[SCRATCH_DEVICES] = {};

// Device address of `scratch` on the current device, resolved once per device.
int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_DEVICES) return FBME_ERR_HIP;
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FBME_ERR_HIP;
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return 0;
}

}  // namespace

// Launcher of k_lr_corr shared with the FFT-based preconditioner passes.
int fbme_lr_corr_launch(int lanes, int rhs, int chunks, const void* part, const void* sinv, const float* bg,
                        void* corr, hipStream_t st, const double* rr_part, int rr_blocks, double* nrho) {
  hipLaunchKernelGGL(k_lr_corr, dim3(lanes), dim3(FBME_BLOCK), 0, st, (const double2*)part, chunks,
                     (const float2*)sinv, bg, rhs, (float2*)corr, rr_part, rr_blocks, nrho);
  return (int)hipGetLastError();
}

// Launcher of k_pcg_commit_parts shared with the fused preconditioner pass.
int fbme_pcg_commit_launch(int lanes, double* rho, const double* curv, const double* nrho, int* active,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_pcg_commit_parts, dim3(nblk(lanes, 64)), dim3(64), 0, st, rho, curv, nrho, active, lanes);
  return (int)hipGetLastError();
}

extern "C" {

int fbme_column_fft(void* stream, int batch, int n, int rows_in, int rows_out, int cols, const void* in, void* out,
                    int inverse);

int fbme_gamma_spectrum(void* stream, int batch, int n1, int n2, int m1, int m2, int e1, int e2, unsigned code1,
                        unsigned code2, unsigned codee1, unsigned codee2, const float* gamma, const float* noise,
                        void* work_a, void* work_b, void* spectrum, const void* tw_n1, const void* tw_n2,
                        const void* tw_e1, const void* tw_e2) {
  // work_a, work_b: batch * max(N, E1 E2) complex each.
  if (batch < 1 || n1 < 1 || n2 < 1 || n1 > 1024 || n2 > 1024 || m1 > n1 || m2 > n2 || e1 < 2 * m1 - 1 ||
      e2 < 2 * m2 - 1 || e1 > FBME_FFT_MAX || e2 > FBME_FFT_MAX)
    return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const long long n = (long long)n1 * n2;
  float2* a = (float2*)work_a;
  float2* bwork = (float2*)work_b;
  hipLaunchKernelGGL(k_real_to_complex, dim3(nblk(batch * n)), dim3(FBME_BLOCK), 0, st, gamma, a, batch * n);
  // lags = fft2(gamma): rows then columns (full lengths)
  launch_rowpass<false>(a, bwork, n2, n2, n2, (long long)batch * n1, code2, (const float2*)tw_n2, st);
  int rc = fbme_column_fft(stream, batch, n1, n1, n1, n2, bwork, a, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_embed, dim3(nblk((long long)e1 * e2), batch), dim3(FBME_BLOCK), 0, st, a, noise, bwork, n1, n2,
                     m1, m2, e1, e2);
  launch_rowpass<false>(bwork, a, e2, e2, e2, (long long)batch * e1, codee2, (const float2*)tw_e2, st);
  rc = fbme_column_fft(stream, batch, e1, e1, e1, e2, a, spectrum, 0);
  (void)code1;
  (void)codee1;
  (void)tw_n1;
  (void)tw_e1;
  if (rc) return rc;
  return (int)hipGetLastError();
}

int fbme_circulant_spectrum(void* stream, int batch, int n1, int n2, int m1, int m2, unsigned code_n2,
                            unsigned code_m2, const float* gamma, const float* noise, void* work_a, void* work_b,
                            float* eig, const void* tw_n2, const void* tw_m2) {
  // work_a, work_b: batch * max(N1 N2, M1 M2) complex64 each; eig: batch * M1 M2.
  if (batch < 1 || m1 < 1 || m2 < 1 || m1 > n1 || m2 > n2 || n1 > 1024 || n2 > 1024) return FBME_ERR_ARG;
  if (code_n2 == 0xFFFFFFFFu || code_m2 == 0xFFFFFFFFu || !gamma || !noise || !eig) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const long long n = (long long)n1 * n2, m = (long long)m1 * m2;
  float2* a = (float2*)work_a;
  float2* b = (float2*)work_b;
  hipLaunchKernelGGL(k_real_to_complex, dim3(nblk(batch * n)), dim3(FBME_BLOCK), 0, st, gamma, a, batch * n);
  launch_rowpass<false>(a, b, n2, n2, n2, (long long)batch * n1, code_n2, (const float2*)tw_n2, st);
  int rc = fbme_column_fft(stream, batch, n1, n1, n1, n2, b, a, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_circ_gather, dim3(nblk(m), batch), dim3(FBME_BLOCK), 0, st, a, n1, n2, m1, m2, b);
  launch_rowpass<false>(b, a, m2, m2, m2, (long long)batch * m1, code_m2, (const float2*)tw_m2, st);
  rc = fbme_column_fft(stream, batch, m1, m1, m1, m2, a, b, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_circ_eig, dim3(batch), dim3(FBME_BLOCK), 0, st, b, noise, (int)m, eig);
  return (int)hipGetLastError();
}

int fbme_circulant_apply(void* stream, int lanes, int rhs, int m1, int m2, unsigned code_m2, const float* eig,
                         const void* v, void* out, void* work, const void* tw_m2) {
  // work: lanes * M1 M2 complex64.
  if (lanes < 1 || rhs < 1 || lanes % rhs != 0 || m1 < 1 || m2 < 1 || m1 > 1024 || m2 > 1024) return FBME_ERR_ARG;
  if (code_m2 == 0xFFFFFFFFu || !eig || !v || !out || !work || v == out) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int m = m1 * m2;
  launch_rowpass<false>((const float2*)v, (float2*)work, m2, m2, m2, (long long)lanes * m1, code_m2,
                        (const float2*)tw_m2, st);
  int rc = fbme_column_fft(stream, lanes, m1, m1, m1, m2, work, out, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_circ_scale, dim3(nblk(m), lanes), dim3(FBME_BLOCK), 0, st, (float2*)out, eig, rhs, m);
  rc = fbme_column_fft(stream, lanes, m1, m1, m1, m2, out, work, 1);
  if (rc) return rc;
  launch_rowpass<true>((const float2*)work, (float2*)out, m2, m2, m2, (long long)lanes * m1, code_m2,
                       (const float2*)tw_m2, st);
  return (int)hipGetLastError();
}

int fbme_toeplitz_matvec_tiled(void* stream, int lanes, int rhs, int m1, int m2, int e1, int e2, unsigned code_e1,
                               unsigned code_e2, const void* spectrum, const void* d, void* out, void* work_a,
                               void* work_b, const void* tw_e1, const void* tw_e2, int row_tile, int col_tile) {
  // work_a, work_b: lanes * M1 * E2 complex each.
  if (row_tile < 0 || col_tile < 0 || e1 < 1 || e2 < 1 || e1 > FBME_FFT_MAX || e2 > FBME_FFT_MAX)
    return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  launch_rowpass<false>((const float2*)d, (float2*)work_a, e2, m2, e2, (long long)lanes * m1, code_e2,
                        (const float2*)tw_e2, st, row_tile);
  const int cap = toeplitz_col_cap(e1, e2);
  const int tc = col_tile > 0 ? std::min(col_tile, cap) : cap;
  const int need = tc * (e1 + 1);
  dim3 grid((e2 + tc - 1) / tc, lanes);
#define FBME_TCOL(EPT, TWN)                                                                                     \
  hipLaunchKernelGGL((k_toeplitz_col<EPT, TWN>), grid, dim3(FBME_BLOCK), 0, st, (const float2*)work_a,         \
                     (float2*)work_b, (const float2*)spectrum, rhs, m1, e1, e2, tc, code_e1, (const float2*)tw_e1)
  if (e1 > 1024) {
    if (need <= 16 * FBME_BLOCK)
      FBME_TCOL(16, FBME_FFT_MAX);
    else
      FBME_TCOL(24, FBME_FFT_MAX);
  } else if (need <= 8 * FBME_BLOCK) {
    FBME_TCOL(8, 1024);
  } else if (need <= 16 * FBME_BLOCK) {
    FBME_TCOL(16, 1024);
  } else {
    FBME_TCOL(24, 1024);
  }
#undef FBME_TCOL
  launch_rowpass<true>((const float2*)work_b, (float2*)out, e2, e2, m2, (long long)lanes * m1, code_e2,
                       (const float2*)tw_e2, st, row_tile);
  return (int)hipGetLastError();
}

int fbme_toeplitz_matvec(void* stream, int lanes, int rhs, int m1, int m2, int e1, int e2, unsigned code_e1,
                         unsigned code_e2, const void* spectrum, const void* d, void* out, void* work_a,
                         void* work_b, const void* tw_e1, const void* tw_e2) {
  return fbme_toeplitz_matvec_tiled(stream, lanes, rhs, m1, m2, e1, e2, code_e1, code_e2, spectrum, d, out, work_a,
                                    work_b, tw_e1, tw_e2, 0, 0);
}

int fbme_toeplitz_tiles(int e1, int e2, int* row_tile_max, int* col_tile_max) {
  if (e1 < 1 || e2 < 1 || e1 > FBME_FFT_MAX || e2 > FBME_FFT_MAX) return FBME_ERR_ARG;
  *row_tile_max = rowpass_cap(e2);
  *col_tile_max = toeplitz_col_cap(e1, e2);
  return 0;
}

size_t fbme_topk32_scratch_bytes(int batch, long long n) {
  if (batch < 1 || n < RANK || n > (1ll << 26)) return 0;
  long long count = (n + 1023) / 1024 * RANK, total = count;
  while (count > 2048) {
    count = (count + 2047) / 2048 * RANK;
    total += count;
  }
  return (size_t)batch * (size_t)total * sizeof(unsigned long long);
}

int fbme_topk32(void* stream, int batch, long long n, const float* x, void* cand, int* idx, float* val) {
  // cand: fbme_topk32_scratch_bytes(batch, n) bytes.  Per-tile top 32, then
  // merge levels of 2048 candidates until one wavefront holds them all.
  if (fbme_topk32_scratch_bytes(batch, n) == 0 || !cand) return FBME_ERR_ARG;
  const int tiles = (int)((n + 1023) / 1024);
  hipStream_t st = (hipStream_t)stream;
  unsigned long long* cur = (unsigned long long*)cand;
  hipLaunchKernelGGL(k_topk_tile, dim3(tiles, batch), dim3(FBME_BLOCK), 0, st, x, n, tiles, cur);
  int count = tiles * RANK;
  while (count > 2048) {
    const int groups = (count + 2047) / 2048;
    unsigned long long* next = cur + (long long)batch * count;
    hipLaunchKernelGGL(k_topk_merge_keys, dim3(groups, batch), dim3(64), 0, st, cur, count, groups, next);
    cur = next;
    count = groups * RANK;
  }
  hipLaunchKernelGGL(k_topk_merge, dim3(batch), dim3(64), 0, st, (const unsigned long long*)cur, count / RANK, x, n,
                     idx, val);
  return (int)hipGetLastError();
}

size_t fbme_topk_scratch_bytes(int batch, long long n, int k) {
  if (batch < 1 || batch > 65535 || k < 1 || k > 256 || n < k || n > (1ll << 26)) return 0;
  long long count = (n + 1023) / 1024 * k, total = count;
  while (count > 2048) {
    count = (count + 2047) / 2048 * k;
    total += count;
  }
  return (size_t)batch * (size_t)total * sizeof(unsigned long long);
}

int fbme_topk(void* stream, int batch, long long n, int k, const float* x, void* cand, int* idx, float* val) {
  if (fbme_topk_scratch_bytes(batch, n, k) == 0 || !cand || !x || !idx || !val) return FBME_ERR_ARG;
  const int tiles = (int)((n + 1023) / 1024);
  hipStream_t st = (hipStream_t)stream;
  unsigned long long* cur = (unsigned long long*)cand;
  hipLaunchKernelGGL(k_topkk_tile, dim3(tiles, batch), dim3(64), 0, st, x, n, tiles, k, cur);
  int count = tiles * k;
  while (count > 2048) {
    const int groups = (count + 2047) / 2048;
    unsigned long long* next = cur + (long long)batch * count;
    hipLaunchKernelGGL(k_topkk_merge, dim3(groups, batch), dim3(64), 0, st, cur, count, groups, k, next);
    cur = next;
    count = groups * k;
  }
  hipLaunchKernelGGL(k_topkk_final, dim3(batch), dim3(64), 0, st, (const unsigned long long*)cur, count, k, x, n,
                     idx, val);
  return (int)hipGetLastError();
}

int fbme_topk32_noalloc(void* stream, int batch, long long n, const float* x, int* idx, float* val) {
  // Every candidate slot is written before it is read, so the shared scratch
  // needs no clearing between calls.
  const size_t need = fbme_topk32_scratch_bytes(batch, n);
  if (need == 0 || need > FBME_TOPK32_NOALLOC_SCRATCH_BYTES) return FBME_ERR_ARG;
  void* ptr = nullptr;
  const int rc = scratch_address(&ptr);
  if (rc) return rc;
  return fbme_topk32(stream, batch, n, x, ptr, idx, val);
}

int fbme_lowrank_setup(void* stream, int batch, int n1, int n2, int m1, int m2, const int* idx, const float* val,
                       const double* gamma_sum, const float* noise, const void* dir1, const void* dir2,
                       void* factors, void* sinv, float* bg, int* status) {
  return fbme_lowrank_setup_masked(stream, batch, n1, n2, m1, m2, idx, val, gamma_sum, noise, dir1, dir2, nullptr,
                                   factors, sinv, bg, status);
}

int fbme_lowrank_setup_masked(void* stream, int batch, int n1, int n2, int m1, int m2, const int* idx,
                              const float* val, const double* gamma_sum, const float* noise, const void* dir1,
                              const void* dir2, const void* mask_spectrum, void* factors, void* sinv, float* bg,
                              int* status) {
  if (batch < 1 || n1 < 1 || n2 < 1 || m1 < 1 || m2 < 1) return FBME_ERR_ARG;
  if (!mask_spectrum && (!dir1 || !dir2)) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_lr_tables, dim3(nblk((long long)(m1 + m2) * RANK), batch), dim3(FBME_BLOCK), 0, st, idx, val,
                     (float2*)factors, n1, n2, m1, m2);
  hipLaunchKernelGGL(k_lr_setup, dim3(batch), dim3(64), 0, st, idx, val, gamma_sum, noise, (const double2*)dir1,
                     (const double2*)dir2, (const double2*)mask_spectrum, n1, n2, (float2*)sinv, bg, status);
  return (int)hipGetLastError();
}

namespace {

// out = vbase / bg - F S^-1 F^H vproj / bg^2 (vproj = vbase: the low-rank
// preconditioner); with dot, dot[lane] = Re <vbase, out>.
int lowrank_apply_general(hipStream_t st, int lanes, int rhs, int m1, int m2, const void* factors,
                          const void* sinv, const float* bg, const void* vproj, const void* v, void* out, void* part,
                          void* corr, double* dot_part, double* dot) {
  // part: lanes * 64 * 32 double2; corr: lanes * 32 complex; dot_part: lanes * ceil(M / 256) doubles.
  if (rhs < 1 || lanes % rhs != 0 || m1 < 1 || m2 < 1) return FBME_ERR_ARG;
  if ((dot_part == nullptr) != (dot == nullptr)) return FBME_ERR_ARG;
  const int batch = lanes / rhs;
  const int rb = (m1 + 63) / 64;
  const int chunks = (m1 + rb - 1) / rb;
  const int groups = (rhs + 23) / 24;
  const int per = (rhs + groups - 1) / groups;
  const long long m = (long long)m1 * m2;
  const dim3 pg(chunks, groups, batch), ag((unsigned)nblk(m), groups, batch);
  const float2* tab = (const float2*)factors;
#define FBME_LR_LAUNCH(LG)                                                                                      \
  hipLaunchKernelGGL((k_lr_project<LG>), pg, dim3(FBME_BLOCK), 0, st, tab, (const float2*)vproj, m1, m2, rhs,  \
                     rb,                                                                                        \
                     (double2*)part);                                                                           \
  hipLaunchKernelGGL(k_lr_corr, dim3(lanes), dim3(FBME_BLOCK), 0, st, (const double2*)part, chunks,             \
                     (const float2*)sinv, bg, rhs, (float2*)corr, (const double*)nullptr, 0,                    \
                     (double*)nullptr);                                                                         \
  if (dot)                                                                                                      \
    hipLaunchKernelGGL((k_lr_apply<LG, true>), ag, dim3(FBME_BLOCK), 0, st, tab, (const float2*)v,              \
                       (const float2*)corr, bg, m1, m2, rhs, (float2*)out, dot_part);                           \
  else                                                                                                          \
    hipLaunchKernelGGL((k_lr_apply<LG, false>), ag, dim3(FBME_BLOCK), 0, st, tab, (const float2*)v,             \
                       (const float2*)corr, bg, m1, m2, rhs, (float2*)out, (double*)nullptr)
  if (per <= 8) {
    FBME_LR_LAUNCH(8);
  } else if (per <= 12) {
    FBME_LR_LAUNCH(12);
  } else if (per <= 17) {
    FBME_LR_LAUNCH(17);
  } else {
    FBME_LR_LAUNCH(24);
  }
#undef FBME_LR_LAUNCH
  if (dot)
    hipLaunchKernelGGL(k_lane_finish, dim3(lanes), dim3(64), 0, st, (const double*)dot_part, (int)ag.x,
                       lanes, dot);
  return (int)hipGetLastError();
}

}  // namespace

int fbme_lowrank_apply_dot(void* stream, int lanes, int rhs, int m1, int m2, const void* factors, const void* sinv,
                           const float* bg, const void* v, void* out, void* part, void* corr, double* dot_part,
                           double* dot) {
  return lowrank_apply_general((hipStream_t)stream, lanes, rhs, m1, m2, factors, sinv, bg, v, v, out, part, corr,
                               dot_part, dot);
}

int fbme_lowrank_apply_split(void* stream, int lanes, int rhs, int m1, int m2, const void* factors,
                             const void* sinv, const float* bg, const void* vproj, const void* vbase, void* out,
                             void* part, void* corr) {
  if (!vproj || !vbase || vbase == out) return FBME_ERR_ARG;
  return lowrank_apply_general((hipStream_t)stream, lanes, rhs, m1, m2, factors, sinv, bg, vproj, vbase, out, part,
                               corr, nullptr, nullptr);
}

int fbme_gdiag_weights(void* stream, int batch, long long n, const float* gamma, const float* noise, const int* idx,
                       const float* logscale, float clamp, float* w) {
  if (batch < 1 || n < RANK || !gamma || !noise || !idx || !w || !(clamp >= 0.f)) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gdiag_weights, dim3((unsigned)nblk(n), batch), dim3(FBME_BLOCK), 0, st, gamma, noise,
                     logscale, clamp, n, w);
  hipLaunchKernelGGL(k_gdiag_atoms, dim3(batch), dim3(64), 0, st, idx, noise, logscale, clamp, n, w);
  return (int)hipGetLastError();
}

int fbme_lowrank_columns(void* stream, int batch, int m1, int m2, const void* factors, void* out) {
  if (batch < 1 || m1 < 1 || m2 < 1 || !factors || !out) return FBME_ERR_ARG;
  const long long m = (long long)m1 * m2;
  hipLaunchKernelGGL(k_lr_columns, dim3((unsigned)nblk(m), RANK, batch), dim3(FBME_BLOCK), 0, (hipStream_t)stream,
                     (const float2*)factors, m1, m2, (float2*)out);
  return (int)hipGetLastError();
}

int fbme_gdiag_system(void* stream, int batch, int m1, int m2, const void* factors, const void* bf, void* part,
                      void* cinv, int* status) {
  // part: batch * RANK * 64 * RANK double2 (the projection partials of the RANK lanes of bf)
  if (batch < 1 || m1 < 1 || m2 < 1 || !factors || !bf || !part || !cinv) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int rb = (m1 + 63) / 64;
  const int chunks = (m1 + rb - 1) / rb;
  const dim3 pg(chunks, 2, batch);   // two groups of 16 lanes per scene
  hipLaunchKernelGGL((k_lr_project<17>), pg, dim3(FBME_BLOCK), 0, st, (const float2*)factors, (const float2*)bf, m1,
                     m2, RANK, rb, (double2*)part);
  hipLaunchKernelGGL(k_gdiag_system, dim3(batch), dim3(64), 0, st, (const double2*)part, chunks, (float2*)cinv,
                     status);
  return (int)hipGetLastError();
}

int fbme_stored_apply(void* stream, int lanes, int rhs, long long m, const void* ucols, const void* cinv,
                      const float* ones, const void* v, void* out, void* part, void* corr, double* dot_part,
                      double* dot) {
  // part: lanes * ceil(m / 1024) * 32 double2; corr: lanes * 32 complex; dot_part: lanes * ceil(m / 256) doubles.
  if (rhs < 1 || lanes % rhs != 0 || m < 1 || !ucols || !cinv || !ones || !v || !out || !part || !corr || v == out)
    return FBME_ERR_ARG;
  if ((dot_part == nullptr) != (dot == nullptr)) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int batch = lanes / rhs;
  const int chunks = (int)((m + SP_CHUNK - 1) / SP_CHUNK);
  const int groups = (rhs + 23) / 24;
  const int per = (rhs + groups - 1) / groups;
  const dim3 pg(chunks, groups, batch), ag((unsigned)nblk(m), groups, batch);
  const float2* uc = (const float2*)ucols;
#define FBME_SA_LAUNCH(LG)                                                                                       \
  hipLaunchKernelGGL((k_stored_project<LG>), pg, dim3(FBME_BLOCK), 0, st, uc, (const float2*)v, m, rhs,          \
                     (double2*)part);                                                                            \
  hipLaunchKernelGGL(k_lr_corr, dim3(lanes), dim3(FBME_BLOCK), 0, st, (const double2*)part, chunks,              \
                     (const float2*)cinv, ones, rhs, (float2*)corr, (const double*)nullptr, 0, (double*)nullptr); \
  if (dot)                                                                                                       \
    hipLaunchKernelGGL((k_lr_apply_stored<LG, true>), ag, dim3(FBME_BLOCK), 0, st, uc, (const float2*)v,          \
                       (const float2*)v, (const float2*)corr, m, rhs, (float2*)out, dot_part);                    \
  else                                                                                                           \
    hipLaunchKernelGGL((k_lr_apply_stored<LG, false>), ag, dim3(FBME_BLOCK), 0, st, uc, (const float2*)v,         \
                       (const float2*)v, (const float2*)corr, m, rhs, (float2*)out, (double*)nullptr)
  if (per <= 8) {
    FBME_SA_LAUNCH(8);
  } else if (per <= 12) {
    FBME_SA_LAUNCH(12);
  } else if (per <= 17) {
    FBME_SA_LAUNCH(17);
  } else {
    FBME_SA_LAUNCH(24);
  }
#undef FBME_SA_LAUNCH
  if (dot)
    hipLaunchKernelGGL(k_lane_finish, dim3(lanes), dim3(64), 0, st, (const double*)dot_part, (int)ag.x, lanes, dot);
  return (int)hipGetLastError();
}

int fbme_gdiag_apply(void* stream, int lanes, int rhs, int m1, int m2, const void* factors, const void* cinv,
                     const float* ones, const void* bf, const void* b, const void* v, void* out, void* part,
                     void* corr, double* dot_part, double* dot) {
  // part: lanes * 64 * 32 double2; corr: lanes * 32 complex; dot_part: lanes * ceil(M / 256) doubles.
  if (rhs < 1 || lanes % rhs != 0 || m1 < 1 || m2 < 1 || !factors || !cinv || !ones || !bf || !b || !out)
    return FBME_ERR_ARG;
  if ((dot_part == nullptr) != (dot == nullptr) || (dot && !v) || b == out) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int batch = lanes / rhs;
  const int rb = (m1 + 63) / 64;
  const int chunks = (m1 + rb - 1) / rb;
  const int groups = (rhs + 23) / 24;
  const int per = (rhs + groups - 1) / groups;
  const long long m = (long long)m1 * m2;
  const dim3 pg(chunks, groups, batch), ag((unsigned)nblk(m), groups, batch);
  const float2* tab = (const float2*)factors;
#define FBME_GD_LAUNCH(LG)                                                                                       \
  hipLaunchKernelGGL((k_lr_project<LG>), pg, dim3(FBME_BLOCK), 0, st, tab, (const float2*)b, m1, m2, rhs, rb,    \
                     (double2*)part);                                                                            \
  hipLaunchKernelGGL(k_lr_corr, dim3(lanes), dim3(FBME_BLOCK), 0, st, (const double2*)part, chunks,              \
                     (const float2*)cinv, ones, rhs, (float2*)corr, (const double*)nullptr, 0, (double*)nullptr); \
  if (dot)                                                                                                       \
    hipLaunchKernelGGL((k_lr_apply_stored<LG, true>), ag, dim3(FBME_BLOCK), 0, st, (const float2*)bf,             \
                       (const float2*)b, (const float2*)v, (const float2*)corr, m, rhs, (float2*)out, dot_part); \
  else                                                                                                           \
    hipLaunchKernelGGL((k_lr_apply_stored<LG, false>), ag, dim3(FBME_BLOCK), 0, st, (const float2*)bf,            \
                       (const float2*)b, (const float2*)v, (const float2*)corr, m, rhs, (float2*)out,            \
                       (double*)nullptr)
  if (per <= 8) {
    FBME_GD_LAUNCH(8);
  } else if (per <= 12) {
    FBME_GD_LAUNCH(12);
  } else if (per <= 17) {
    FBME_GD_LAUNCH(17);
  } else {
    FBME_GD_LAUNCH(24);
  }
#undef FBME_GD_LAUNCH
  if (dot)
    hipLaunchKernelGGL(k_lane_finish, dim3(lanes), dim3(64), 0, st, (const double*)dot_part, (int)ag.x, lanes, dot);
  return (int)hipGetLastError();
}

int fbme_lowrank_apply(void* stream, int lanes, int rhs, int m1, int m2, const void* factors, const void* sinv,
                       const float* bg, const void* v, void* out, void* part, void* corr) {
  return fbme_lowrank_apply_dot(stream, lanes, rhs, m1, m2, factors, sinv, bg, v, out, part, corr, nullptr,
                                nullptr);
}

int fbme_mask_spectrum(void* stream, int n1, int n2, int m1, int m2, const unsigned char* mask, void* spectrum,
                       void* work) {
  if (n1 < 1 || n2 < 1 || m1 < 1 || m2 < 1 || m1 > n1 || m2 > n2 || !mask || !spectrum || !work) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mask_rows, dim3(nblk((long long)m1 * n2)), dim3(FBME_BLOCK), 0, st, mask, m1, m2, n2,
                     (double2*)work);
  hipLaunchKernelGGL(k_mask_cols, dim3(nblk((long long)n1 * n2)), dim3(FBME_BLOCK), 0, st, (const double2*)work, m1,
                     n1, n2, (double2*)spectrum);
  return (int)hipGetLastError();
}

int fbme_mask_lanes_dot(void* stream, int lanes, int m, const unsigned char* mask, const void* a, void* out,
                        double* part, double* dot) {
  if (lanes < 1 || m < 1 || !mask || !out) return FBME_ERR_ARG;
  if (dot && (!part || !a)) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int blocks = std::min(64, nblk(m));
  hipLaunchKernelGGL(k_mask_lanes, dim3(blocks, lanes), dim3(FBME_BLOCK), 0, st, mask, (const float2*)a,
                     (float2*)out, m, dot ? part : (double*)nullptr);
  if (dot)
    hipLaunchKernelGGL(k_lane_finish, dim3(lanes), dim3(64), 0, st, (const double*)part, blocks, lanes, dot);
  return (int)hipGetLastError();
}

int fbme_lane_dot(void* stream, int lanes, int m, const void* a, const void* b, double* part, double* out) {
  // part: lanes * 64 doubles.
  hipStream_t st = (hipStream_t)stream;
  const int blocks = std::min(64, nblk(m));
  hipLaunchKernelGGL(k_lane_dot, dim3(blocks, lanes), dim3(FBME_BLOCK), 0, st, (const float2*)a, (const float2*)b, m,
                     part);
  hipLaunchKernelGGL(k_lane_finish, dim3(lanes), dim3(64), 0, st, (const double*)part, blocks, lanes, out);
  return (int)hipGetLastError();
}

int fbme_pcg_update(void* stream, int lanes, int m, void* x, void* r, const void* d, const void* q,
                    const double* rho, const double* curv, int* active) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pcg_update, dim3(nblk(m), lanes), dim3(FBME_BLOCK), 0, st, (float2*)x, (float2*)r,
                     (const float2*)d, (const float2*)q, rho, curv, active, m);
  hipLaunchKernelGGL(k_pcg_commit_valid, dim3(nblk(lanes, 64)), dim3(64), 0, st, rho, curv, active, lanes);
  return (int)hipGetLastError();
}

int fbme_pcg_update_parts(void* stream, int lanes, int m, void* x, void* r, const void* d, const void* q,
                          const double* rho, const double* part, int nparts, double* curv, const int* active) {
  if (lanes < 1 || m < 1 || nparts < 1) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_pcg_update_parts<false>, dim3(nblk(m), lanes), dim3(FBME_BLOCK), 0, (hipStream_t)stream,
                     (float2*)x, (float2*)r, (const float2*)d, (const float2*)q, rho, part, nparts, curv, active, m,
                     (double*)nullptr);
  return (int)hipGetLastError();
}

int fbme_pcg_update_parts_rr(void* stream, int lanes, int m, void* x, void* r, const void* d, const void* q,
                             const double* rho, const double* part, int nparts, double* curv, const int* active,
                             double* rr_part) {
  if (lanes < 1 || m < 1 || nparts < 1 || !rr_part) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_pcg_update_parts<true>, dim3(nblk(m), lanes), dim3(FBME_BLOCK), 0, (hipStream_t)stream,
                     (float2*)x, (float2*)r, (const float2*)d, (const float2*)q, rho, part, nparts, curv, active, m,
                     rr_part);
  return (int)hipGetLastError();
}

int fbme_pcg_direction_parts(void* stream, int lanes, int m, void* d, const void* z, double* rho,
                             const double* curv, const double* part, int nparts, double* nrho, int* active) {
  if (lanes < 1 || m < 1 || nparts < 1) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pcg_direction_parts, dim3(nblk(m), lanes), dim3(FBME_BLOCK), 0, st, (float2*)d,
                     (const float2*)z, rho, curv, part, nparts, nrho, active, m);
  hipLaunchKernelGGL(k_pcg_commit_parts, dim3(nblk(lanes, 64)), dim3(64), 0, st, rho, curv, nrho, active, lanes);
  return (int)hipGetLastError();
}

int fbme_pcg_commit_valid(void* stream, int lanes, const double* rho, const double* curv, int* active) {
  hipLaunchKernelGGL(k_pcg_commit_valid, dim3(nblk(lanes, 64)), dim3(64), 0, (hipStream_t)stream, rho, curv, active,
                     lanes);
  return (int)hipGetLastError();
}

int fbme_pcg_direction(void* stream, int lanes, int m, void* d, const void* z, double* rho, const double* nrho,
                       int* active, int restart) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pcg_direction, dim3(nblk(m), lanes), dim3(FBME_BLOCK), 0, st, (float2*)d, (const float2*)z,
                     rho, nrho, active, m, restart);
  hipLaunchKernelGGL(k_pcg_commit_rho, dim3(nblk(lanes, 64)), dim3(64), 0, st, rho, nrho, active, lanes);
  return (int)hipGetLastError();
}

int fbme_sub(void* stream, long long n, const void* a, const void* b, void* out) {
  hipLaunchKernelGGL(k_sub, dim3(nblk(n)), dim3(FBME_BLOCK), 0, (hipStream_t)stream, (const float2*)a,
                     (const float2*)b, (float2*)out, n);
  return (int)hipGetLastError();
}

int fbme_pcg_check(void* stream, int lanes, const double* rr, const double* init, int* active, double* best_rel,
                   int* improved, float tol, int fixed) {
  hipLaunchKernelGGL(k_pcg_check, dim3(nblk(lanes, 64)), dim3(64), 0, (hipStream_t)stream, rr, init, active,
                     best_rel, improved, tol, fixed, lanes);
  return (int)hipGetLastError();
}

int fbme_spatial_init(void* stream, int batch, int n1, int n2, float coupling, int circular, float max_prec,
                      const float* floor_, float* gamma, float* pattern) {
  if (batch < 1 || n1 < 1 || n2 < 1 || !(coupling >= 0.f) || !floor_ || !gamma || !pattern) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const long long n = (long long)n1 * n2;
  const dim3 g((unsigned)nblk(n), batch);
  hipLaunchKernelGGL(k_spatial_init, g, dim3(FBME_BLOCK), 0, st, (const float*)gamma, n1, n2, coupling, circular,
                     pattern);
  hipLaunchKernelGGL(k_spatial_gamma, g, dim3(FBME_BLOCK), 0, st, (const float*)pattern, n1, n2, coupling, circular,
                     max_prec, floor_, gamma);
  return (int)hipGetLastError();
}

int fbme_spatial_update(void* stream, int batch, int n1, int n2, float coupling, int circular, float hyper_shape,
                        float hyper_rate, float max_prec, float keep, const void* mean, const float* var,
                        const float* floor_, float* second, float* pattern, float* gamma) {
  if (batch < 1 || n1 < 1 || n2 < 1 || !(coupling >= 0.f) || !mean || !var || !floor_ || !second || !pattern ||
      !gamma)
    return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const long long n = (long long)n1 * n2;
  const dim3 g((unsigned)nblk(n), batch);
  hipLaunchKernelGGL(k_spatial_second, g, dim3(FBME_BLOCK), 0, st, (const float2*)mean, var, n, second);
  hipLaunchKernelGGL(k_spatial_pattern, g, dim3(FBME_BLOCK), 0, st, (const float*)second, n1, n2, coupling, circular,
                     hyper_shape, hyper_rate, max_prec, keep, pattern);
  hipLaunchKernelGGL(k_spatial_gamma, g, dim3(FBME_BLOCK), 0, st, (const float*)pattern, n1, n2, coupling, circular,
                     max_prec, floor_, gamma);
  return (int)hipGetLastError();
}

int fbme_hutchinson_used(void* stream, int batch, int n, int probes, int used, const void* lifted,
                         const void* lprobes, const float* gamma, void* mean, float* var, float* dd, double* part,
                         double* rel_sum) {
  if (batch < 1 || n < 1 || used < 1 || used > probes) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int blocks = nblk(n);
  hipLaunchKernelGGL(k_hutch, dim3(blocks, batch), dim3(FBME_BLOCK), 0, st, (const float2*)lifted,
                     (const float2*)lprobes, gamma, n, probes, used, (float2*)mean, var, dd, part);
  hipLaunchKernelGGL(k_seg_sum, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, rel_sum,
                     (float*)nullptr);
  return (int)hipGetLastError();
}

int fbme_hutchinson(void* stream, int batch, int n, int probes, const void* lifted, const void* lprobes,
                    const float* gamma, void* mean, float* var, float* dd, double* part, double* rel_sum) {
  hipStream_t st = (hipStream_t)stream;
  const int blocks = nblk(n);
  hipLaunchKernelGGL(k_hutch, dim3(blocks, batch), dim3(FBME_BLOCK), 0, st, (const float2*)lifted,
                     (const float2*)lprobes, gamma, n, probes, probes, (float2*)mean, var, dd, part);
  hipLaunchKernelGGL(k_seg_sum, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, rel_sum,
                     (float*)nullptr);
  return (int)hipGetLastError();
}

int fbme_joint_init(void* stream, int batch, int n, const void* bp, const double* bpe, const float* power,
                    const float* floor_, float blend, int shared_noise, float* gamma, float* noise) {
  if (batch < 1 || n < 1) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_joint_gamma0, dim3(nblk(n)), dim3(FBME_BLOCK), 0, st, (const float2*)bp, bpe, power, floor_,
                     blend, batch, n, gamma);
  if (shared_noise) hipLaunchKernelGGL(k_joint_noise0, dim3(1), dim3(64), 0, st, noise, batch);
  return (int)hipGetLastError();
}

int fbme_hutchinson_shared(void* stream, int batch, int n, int probes, const void* lifted, const void* lprobes,
                           const float* gamma, void* mean, float* var, float* dd, double* part, double* rel_sum) {
  if (batch < 1 || n < 1 || probes < 1) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int blocks = nblk(n);
  hipLaunchKernelGGL(k_hutch_shared, dim3(blocks), dim3(FBME_BLOCK), 0, st, (const float2*)lifted,
                     (const float2*)lprobes, gamma, batch, n, probes, (float2*)mean, var, dd, part);
  hipLaunchKernelGGL(k_seg_sum, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, rel_sum,
                     (float*)nullptr);
  return (int)hipGetLastError();
}

int fbme_cg_em_joint(void* stream, int batch, int n, int m, float* gamma, const void* mean, const float* var,
                     const float* dd, const double* rel_sum, const float* floor_, float hyper_shape, float hyper_rate,
                     float max_prec, float keep, int update, double* part, double* dof, float* dof_f) {
  if (batch < 1 || n < 1 || m < 1) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int blocks = nblk(n);
  hipLaunchKernelGGL(k_cg_em_joint, dim3(blocks), dim3(FBME_BLOCK), 0, st, gamma, (const float2*)mean, var, dd,
                     rel_sum, floor_, batch, n, m, hyper_shape, hyper_rate, max_prec, keep, update, part);
  hipLaunchKernelGGL(k_seg_sum, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, dof, dof_f);
  return (int)hipGetLastError();
}

int fbme_cg_em(void* stream, int batch, int n, int m, float* gamma, const void* mean, const float* dd,
               const double* rel_sum, const float* floor_, const unsigned char* mask, float hyper_shape,
               float hyper_rate, float max_prec, float keep, int update, double* part, double* dof,
               float* dof_f) {
  hipStream_t st = (hipStream_t)stream;
  const int blocks = nblk(n);
  hipLaunchKernelGGL(k_cg_em, dim3(blocks, batch), dim3(FBME_BLOCK), 0, st, gamma, (const float2*)mean, dd, rel_sum,
                     floor_, mask, n, m, hyper_shape, hyper_rate, max_prec, keep, update, part);
  hipLaunchKernelGGL(k_seg_sum, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, dof, dof_f);
  return (int)hipGetLastError();
}

int fbme_cg_init(void* stream, int batch, int n, int m, const void* bp, const double* ye, const double* bpe,
                 float min_variance, float noise_fraction, float blend, float* power, float* floor_, float* noise,
                 float* gamma) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cg_scalars, dim3(nblk(batch, 64)), dim3(64), 0, st, ye, m, min_variance, noise_fraction, power,
                     floor_, noise, batch);
  hipLaunchKernelGGL(k_cg_gamma0, dim3(nblk(n), batch), dim3(FBME_BLOCK), 0, st, (const float2*)bp, bpe, power, floor_,
                     blend, n, gamma);
  return (int)hipGetLastError();
}

int fbme_relevance_sum(void* stream, int batch, int n, const float* gamma, const float* dd, double* part,
                       double* out) {
  hipStream_t st = (hipStream_t)stream;
  const int blocks = nblk(n);
  hipLaunchKernelGGL(k_rel_sum, dim3(blocks, batch), dim3(FBME_BLOCK), 0, st, gamma, dd, n, part);
  hipLaunchKernelGGL(k_seg_sum, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, out, (float*)nullptr);
  return (int)hipGetLastError();
}

int fbme_reciprocal(void* stream, long long n, const float* in, float* out) {
  hipLaunchKernelGGL(k_reciprocal, dim3(nblk(n)), dim3(FBME_BLOCK), 0, (hipStream_t)stream, in, out, n);
  return (int)hipGetLastError();
}

int fbme_masked_copy(void* stream, int lanes, int m, void* dst, const void* src, const int* mask) {
  hipLaunchKernelGGL(k_masked_copy, dim3(nblk(m), lanes), dim3(FBME_BLOCK), 0, (hipStream_t)stream, (float2*)dst,
                     (const float2*)src, mask, m);
  return (int)hipGetLastError();
}

int fbme_twiddle_table(void* stream, int len, void* out) {
  if (len < 1 || len > FBME_FFT_MAX) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_twiddle, dim3(nblk(len)), dim3(FBME_BLOCK), 0, (hipStream_t)stream, (float2*)out, len);
  return (int)hipGetLastError();
}

int fbme_grid_sums(void* stream, int batch, int n, const float* gamma, const void* mean, const void* prev,
                   double* part, double* out) {
  hipStream_t st = (hipStream_t)stream;
  const int blocks = nblk(n);
  hipLaunchKernelGGL(k_sums, dim3(blocks, batch), dim3(FBME_BLOCK), 0, st, gamma, (const float2*)mean,
                     (const float2*)prev, n, part);
  hipLaunchKernelGGL(k_sums_finish, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, out);
  return (int)hipGetLastError();
}

}  // extern "C"
