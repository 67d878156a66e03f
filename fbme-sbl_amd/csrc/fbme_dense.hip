// Exact dense posterior of the general solver, mode="dense" (gfx950):
// _posterior's dense branch (torch_sbl/solver.py:2962-2990).
//
//   * Q = noise I + A Gamma A^H over the observed samples, from the Fourier
//     lags of gamma (operators.py:282-297): c[l] = sum_n gamma_n
//     exp(-2 pi i l.n / N) at the (2 M1 - 1)(2 M2 - 1) sample-index
//     differences, a separable FP64 DFT;
//   * Cholesky with the GPU's one deterministic loading jitter * mean|diag|
//     (solver.py:2898-2920), then the triangular inverse W = L^-1, FP64, in
//     32 x 32 tiles staged in LDS: per block column a diagonal-tile kernel
//     (factor and inverse), a panel kernel and a trailing-update kernel over
//     all tile pairs, then one launch per block row of W; the matrix is padded
//     to a multiple of 32 with an identity block;
//   * Q^-1 y = W^H (W y): one wavefront per row for W y, then column sums
//     of W over row chunks (coalesced, fixed-order partials);
//   * the data diagonal diag(A^H Q^-1 A)_n = sum_k |(A^H conj(w_k))_n|^2 over
//     the rows w_k of W (a_n^H Q^-1 a_n = ||W a_n||^2; the reference
//     accumulates Q^-1 along its Fourier lags, operators.py:418-442, equally
//     exact), and A^H Q^-1 y, from separable FP64 prefix transforms of those
//     rows: the first axis reads W in place through the sample -> observed
//     index map, the second runs over chunks of rows in parallel with
//     fixed-order partial sums (twiddles from per-block LDS tables, phases
//     reduced mod N; grid axes up to 1024).
#include "fbme_device.h"

using namespace fbme;

namespace {

constexpr int DN_THREADS = 256;
constexpr int DN_MMAX = 2048;

__device__ __forceinline__ double2 cis(double turns) {
  double s, c;
  sincospi(2.0 * turns, &s, &c);
  return make_double2(c, s);
}

// t[b, n1, j] = sum_n2 gamma[b, n1, n2] exp(-2 pi i l2 n2 / N2), l2 = j - (M2 - 1).
__global__ void k_lag_rows(const float* __restrict__ gamma, int n1, int n2, int m2, double2* __restrict__ t) {
  const int l2c = 2 * m2 - 1;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= (long long)n1 * l2c) return;
  const int r = (int)(e / l2c), j = (int)(e - (long long)r * l2c);
  const long long l2 = j - (m2 - 1);
  const float* g = gamma + ((long long)b * n1 + r) * n2;
  double sx = 0.0, sy = 0.0;
  for (int c = 0; c < n2; ++c) {
    long long ph = (l2 * c) % n2;
    if (ph < 0) ph += n2;
    const double2 w = cis(-(double)ph / n2);
    sx += g[c] * w.x;
    sy += g[c] * w.y;
  }
  t[(long long)b * n1 * l2c + e] = make_double2(sx, sy);
}

// lags[b, i, j] = sum_n1 t[b, n1, j] exp(-2 pi i l1 n1 / N1), l1 = i - (M1 - 1).
__global__ void k_lag_cols(const double2* __restrict__ t, int n1, int m1, int m2, double2* __restrict__ lags) {
  const int l1c = 2 * m1 - 1, l2c = 2 * m2 - 1;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= (long long)l1c * l2c) return;
  const int i = (int)(e / l2c), j = (int)(e - (long long)i * l2c);
  const long long l1 = i - (m1 - 1);
  const double2* tb = t + (long long)b * n1 * l2c;
  double sx = 0.0, sy = 0.0;
  for (int r = 0; r < n1; ++r) {
    long long ph = (l1 * r) % n1;
    if (ph < 0) ph += n1;
    const double2 w = cis(-(double)ph / n1);
    const double2 v = tb[(long long)r * l2c + j];
    sx += v.x * w.x - v.y * w.y;
    sy += v.x * w.y + v.y * w.x;
  }
  lags[(long long)b * l1c * l2c + e] = make_double2(sx, sy);
}

// Q[b, r, c] = lags[s_r - s_c] + noise delta_rc over the observed samples
// (obs[r] = sample index m1 * M2 + m2 of row r; NULL: all samples), plus
// the loading jitter * mean|diag| on the diagonal (every diagonal entry is
// c[0] + noise); rows and columns m..mp-1 are an identity block.
// Sample index of observed sample k, or -1 when the caller's index lies
// outside [0, M1 M2) (such a sample is left out and status flags it).
__device__ __forceinline__ int obs_at(const int* __restrict__ obs, int k, int mfull) {
  if (!obs) return k;
  const int v = obs[k];
  return (v >= 0 && v < mfull) ? v : -1;
}

__global__ void k_dense_build(const double2* __restrict__ lags, const int* __restrict__ obs, int m, int mp, int m1,
                              int m2, const float* __restrict__ noise, double jitter, double tiny,
                              double2* __restrict__ q, int* __restrict__ status) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= (long long)mp * mp) return;
  const int r = (int)(e / mp), c = (int)(e - (long long)r * mp);
  double2 v = make_double2(r == c ? 1.0 : 0.0, 0.0);
  const int sr = r < m ? obs_at(obs, r, m1 * m2) : 0, sc = c < m ? obs_at(obs, c, m1 * m2) : 0;
  if (r < m && r == c && sr < 0 && status) status[b] = 2;
  if (r < m && c < m && sr >= 0 && sc >= 0) {
    const long long base = (long long)b * (2 * m1 - 1) * (2 * m2 - 1);
    const int d1 = sr / m2 - sc / m2, d2 = sr % m2 - sc % m2;
    v = lags[base + (long long)(d1 + m1 - 1) * (2 * m2 - 1) + (d2 + m2 - 1)];
    if (r == c) {
      const double diag = lags[base + (long long)(m1 - 1) * (2 * m2 - 1) + (m2 - 1)].x + (double)noise[b];
      v.x += (double)noise[b] + jitter * fmax(fabs(diag), tiny);
    }
  }
  q[(long long)b * mp * mp + e] = v;
}

__device__ double block_sum_dn(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = DN_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// ------------------------------------------------ tiled FP64 factorisation
constexpr int TS = 32;
static_assert(DN_THREADS * 4 == TS * TS, "tile kernels give every thread four elements");

__device__ __forceinline__ double2 zm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zcm(double2 a, double2 b) {   // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Diagonal tile jb: in-LDS Cholesky L_jj and its inverse, L_jj written back
// to Q and the inverse to linv[b, jb].
__global__ void __launch_bounds__(DN_THREADS)
k_tile_diag(double2* __restrict__ qs, int mp, int jb, double tiny, double2* __restrict__ linv,
            int* __restrict__ status) {
  __shared__ double2 a[TS][TS + 1];
  __shared__ double2 x[TS][TS + 1];
  __shared__ double piv[TS];
  const int t = threadIdx.x, b = blockIdx.x, nb = mp / TS;
  double2* q = qs + (long long)b * mp * mp;
  const long long o = (long long)jb * TS * mp + (long long)jb * TS;
  for (int e = t; e < TS * TS; e += DN_THREADS) a[e / TS][e % TS] = q[o + (long long)(e / TS) * mp + e % TS];
  __syncthreads();
  // Right-looking factorisation, one barrier per column: column j stays
  // unscaled (L[i][j] = a[i][j] / piv_j is applied at the end), so step j
  // reads only column j and writes only columns > j.
  for (int j = 0; j < TS; ++j) {
    const double d = a[j][j].x;
    const double dc = fmax(d, tiny);
    if (t == 0) {
      if ((!(d > 0.0) || !isfinite(d)) && status) status[b] = 1;
      piv[j] = sqrt(dc);
    }
    const double id = 1.0 / dc;
    // thread t owns row t / 8, columns t % 8 + 8u (a fixed map: no
    // division in the step loop)
    const int i = t >> 3;
    if (i > j) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = (t & 7) + 8 * u;
        if (k > j && k <= i) {
          const double2 v = zcm(a[i][j], a[k][j]);
          a[i][k] = make_double2(a[i][k].x - v.x * id, a[i][k].y - v.y * id);
        }
      }
    }
    __syncthreads();
  }
  for (int e = t; e < TS * TS; e += DN_THREADS) {
    const int i = e / TS, j = e % TS;
    if (i > j) {
      const double sj = 1.0 / piv[j];
      a[i][j] = make_double2(a[i][j].x * sj, a[i][j].y * sj);
    } else if (i == j) {
      a[i][i] = make_double2(piv[i], 0.0);
    }
    x[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  // X = L^-1 for all 32 columns at once, one barrier per row step: row k of
  // X stays unscaled (divided by L[k][k] at the end), so step k reads row k
  // and writes rows > k.
  for (int k = 0; k < TS - 1; ++k) {
    const double il = 1.0 / a[k][k].x;
    const int i = t >> 3;
    if (i > k) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = (t & 7) + 8 * u;
        const double2 v = zm(a[i][k], x[k][c]);
        x[i][c] = make_double2(x[i][c].x - v.x * il, x[i][c].y - v.y * il);
      }
    }
    __syncthreads();
  }
  for (int e = t; e < TS * TS; e += DN_THREADS) {
    const int i = e / TS, c = e % TS;
    const double il = 1.0 / a[i][i].x;
    x[i][c] = make_double2(x[i][c].x * il, x[i][c].y * il);
  }
  __syncthreads();
  double2* li = linv + ((long long)b * nb + jb) * TS * TS;
  for (int e = t; e < TS * TS; e += DN_THREADS) {
    const int i = e / TS, k = e % TS;
    q[o + (long long)i * mp + k] = k <= i ? a[i][k] : make_double2(0.0, 0.0);
    li[e] = x[i][k];
  }
}

// Panel: L_ib,jb = A_ib,jb L_jj^-H for ib = jb + 1 + blockIdx.x.
__global__ void __launch_bounds__(DN_THREADS)
k_tile_panel(double2* __restrict__ qs, int mp, int jb, const double2* __restrict__ linv) {
  __shared__ double2 a[TS][TS + 1];
  __shared__ double2 x[TS][TS + 1];
  const int t = threadIdx.x, b = blockIdx.y, nb = mp / TS;
  const int ib = jb + 1 + blockIdx.x;
  double2* q = qs + (long long)b * mp * mp;
  const long long o = (long long)ib * TS * mp + (long long)jb * TS;
  const double2* li = linv + ((long long)b * nb + jb) * TS * TS;
  for (int e = t; e < TS * TS; e += DN_THREADS) {
    a[e / TS][e % TS] = q[o + (long long)(e / TS) * mp + e % TS];
    x[e / TS][e % TS] = li[e];
  }
  __syncthreads();
  const int r = t >> 3, c0 = t & 7;
  double2 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int k = 0; k < TS; ++k) {
    const double2 v = a[r][k];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double2 p = zcm(v, x[c0 + 8 * u][k]);   // A[r][k] conj(Linv[c][k])
      acc[u] = make_double2(acc[u].x + p.x, acc[u].y + p.y);
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) q[o + (long long)r * mp + c0 + 8 * u] = acc[u];
}

// Trailing update: A_ib,kb -= L_ib,jb L_kb,jb^H for jb < kb <= ib, one
// workgroup per tile pair (blockIdx.x enumerates the lower-triangular pairs).
__global__ void __launch_bounds__(DN_THREADS)
k_tile_update(double2* __restrict__ qs, int mp, int jb) {
  __shared__ double2 l1[TS][TS + 1];
  __shared__ double2 l2[TS][TS + 1];
  const int t = threadIdx.x, b = blockIdx.y;
  // pair index p -> (i, k) with 0 <= k <= i over the remaining tiles
  int p = blockIdx.x, i = 0;
  while (p > i) {
    p -= i + 1;
    ++i;
  }
  const int ib = jb + 1 + i, kb = jb + 1 + p;
  double2* q = qs + (long long)b * mp * mp;
  const long long oi = (long long)ib * TS * mp + (long long)jb * TS, ok = (long long)kb * TS * mp + (long long)jb * TS;
  for (int e = t; e < TS * TS; e += DN_THREADS) {
    l1[e / TS][e % TS] = q[oi + (long long)(e / TS) * mp + e % TS];
    l2[e / TS][e % TS] = q[ok + (long long)(e / TS) * mp + e % TS];
  }
  __syncthreads();
  const int r = t >> 3, c0 = t & 7;
  double2 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int k = 0; k < TS; ++k) {
    const double2 v = l1[r][k];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double2 w = zcm(v, l2[c0 + 8 * u][k]);
      acc[u] = make_double2(acc[u].x + w.x, acc[u].y + w.y);
    }
  }
  const long long o = (long long)ib * TS * mp + (long long)kb * TS;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    double2* d = q + o + (long long)r * mp + c0 + 8 * u;
    *d = make_double2(d->x - acc[u].x, d->y - acc[u].y);
  }
}

constexpr int DN_KS = 8;   // interleaved partial sums per tile of W

// Partial sums of block row ib of W = L^-1: for j = blockIdx.x < ib and
// p = blockIdx.y, spart[j, p] = sum over k = j + p, j + p + KS, .. < ib of
// L_ik W_kj (the row's serial chain split KS ways).
__global__ void __launch_bounds__(DN_THREADS)
k_tile_inverse_part(const double2* __restrict__ qs, int mp, int ib, const double2* __restrict__ ws,
                    double2* __restrict__ spart) {
  __shared__ double2 l[TS][TS + 1];
  __shared__ double2 x[TS][TS + 1];
  const int t = threadIdx.x, jb = blockIdx.x, pk = blockIdx.y, b = blockIdx.z, nb = mp / TS;
  const double2* q = qs + (long long)b * mp * mp;
  const double2* w = ws + (long long)b * mp * mp;
  const int r = t >> 3, c0 = t & 7;
  double2 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int kb = jb + pk; kb < ib; kb += DN_KS) {
    for (int e = t; e < TS * TS; e += DN_THREADS) {
      l[e / TS][e % TS] = q[(long long)(ib * TS + e / TS) * mp + kb * TS + e % TS];
      x[e / TS][e % TS] = w[(long long)(kb * TS + e / TS) * mp + jb * TS + e % TS];
    }
    __syncthreads();
    for (int k = 0; k < TS; ++k) {
      const double2 v = l[r][k];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double2 p = zm(v, x[k][c0 + 8 * u]);
        acc[u] = make_double2(acc[u].x + p.x, acc[u].y + p.y);
      }
    }
    __syncthreads();
  }
  double2* sp = spart + (((long long)b * nb + jb) * DN_KS + pk) * TS * TS;
#pragma unroll
  for (int u = 0; u < 4; ++u) sp[r * TS + c0 + 8 * u] = acc[u];
}

// Block row i of W = L^-1 (into w, separate from L): W_ii = L_ii^-1 and, for
// j = blockIdx.x < i, W_ij = -L_ii^-1 sum_{k=j}^{i-1} L_ik W_kj from the KS
// partial sums (added in a fixed order).
__global__ void __launch_bounds__(DN_THREADS)
k_tile_inverse_row(int mp, int ib, const double2* __restrict__ linv, const double2* __restrict__ spart,
                   double2* __restrict__ ws) {
  __shared__ double2 l[TS][TS + 1];
  __shared__ double2 s[TS][TS + 1];
  const int t = threadIdx.x, b = blockIdx.y, nb = mp / TS, jb = blockIdx.x;
  double2* w = ws + (long long)b * mp * mp;
  const double2* lii = linv + ((long long)b * nb + ib) * TS * TS;
  const int r = t >> 3, c0 = t & 7;
  if (jb == ib) {
    for (int e = t; e < TS * TS; e += DN_THREADS)
      w[(long long)(ib * TS + e / TS) * mp + ib * TS + e % TS] = lii[e];
    return;
  }
  const double2* sp = spart + ((long long)b * nb + jb) * DN_KS * TS * TS;
  for (int e = t; e < TS * TS; e += DN_THREADS) {
    double sx = 0.0, sy = 0.0;
    for (int p = 0; p < DN_KS; ++p) {
      const double2 v = sp[(long long)p * TS * TS + e];
      sx += v.x;
      sy += v.y;
    }
    s[e / TS][e % TS] = make_double2(sx, sy);
    l[e / TS][e % TS] = lii[e];
  }
  __syncthreads();
  double2 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int k = 0; k < TS; ++k) {
    const double2 v = l[r][k];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double2 p = zm(v, s[k][c0 + 8 * u]);
      acc[u] = make_double2(acc[u].x + p.x, acc[u].y + p.y);
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
    w[(long long)(ib * TS + r) * mp + jb * TS + c0 + 8 * u] = make_double2(-acc[u].x, -acc[u].y);
}

// Rows of the transforms, FP64, in the full sample layout (zeros on the
// unobserved samples): rows 0..m-1 = conj(W[k, :]), row m = z = W^H (W y).
constexpr int DN_VCH = 128;   // rows of W per partial sum of W^H u
constexpr int DN_RCH = 64;    // rows of conj(W) per partial data diagonal
constexpr int DN_AXMAX = 1024;   // longest grid axis (LDS twiddle tables)

// pos[j] = observed index of sample j, -1 where sample j is not observed:
// identity without a list, else filled with -1 and scattered from the list.
__global__ void k_dense_pos_fill(int listed, int mfull, int* __restrict__ pos) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < mfull) pos[j] = listed ? -1 : j;
}

__global__ void k_dense_pos_scatter(const int* __restrict__ obs, int m, int mfull, int* __restrict__ pos) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int s = obs_at(obs, k, mfull);
  if (s >= 0) pos[s] = k;
}

// u[b, i] = sum_{k <= i} W[i][k] y[obs k]: one wavefront per row of W.
__global__ void __launch_bounds__(DN_THREADS)
k_dense_u(const double2* __restrict__ ws, const float2* __restrict__ y, const int* __restrict__ obs, int m, int mp,
          int mfull, double2* __restrict__ u) {
  const int lane = threadIdx.x & 63, b = blockIdx.y;
  const int i = blockIdx.x * (DN_THREADS / 64) + (threadIdx.x >> 6);
  if (i >= m) return;   // uniform per wavefront
  const double2* w = ws + (long long)b * mp * mp + (long long)i * mp;
  const float2* yb = y + (long long)b * mfull;
  double ax = 0.0, ay = 0.0;
  for (int k = lane; k <= i; k += 64) {
    const int sk = obs_at(obs, k, mfull);
    if (sk < 0) continue;
    const double2 a = w[k];
    const float2 v = yb[sk];
    ax += a.x * v.x - a.y * v.y;
    ay += a.x * v.y + a.y * v.x;
  }
  ax = wave_sum(ax);
  ay = wave_sum(ay);
  if (lane == 0) u[(long long)b * m + i] = make_double2(ax, ay);
}

// vpart[b, c, k] = sum over rows i of chunk c with i >= k of conj(W[i][k]) u_i:
// lane l owns column k (row-contiguous loads), the four wavefronts split the
// chunk's rows, fixed-order sums.
__global__ void __launch_bounds__(DN_THREADS)
k_dense_wh(const double2* __restrict__ ws, const double2* __restrict__ u, int m, int mp, int nch,
           double2* __restrict__ vpart) {
  __shared__ double2 red[DN_THREADS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, c = blockIdx.y, b = blockIdx.z;
  const int k = blockIdx.x * 64 + lane;
  const double2* w = ws + (long long)b * mp * mp;
  const double2* ub = u + (long long)b * m;
  const int i1 = min(m, (c + 1) * DN_VCH);
  double ax = 0.0, ay = 0.0;
  if (k < m) {
    for (int i = c * DN_VCH + wave; i < i1; i += DN_THREADS / 64) {
      if (i < k) continue;   // W is lower triangular
      const double2 a = w[(long long)i * mp + k], v = ub[i];
      ax += a.x * v.x + a.y * v.y;
      ay += a.x * v.y - a.y * v.x;
    }
  }
  red[threadIdx.x] = make_double2(ax, ay);
  __syncthreads();
  if (wave == 0 && k < m) {
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int q = 0; q < DN_THREADS / 64; ++q) {
      sx += red[q * 64 + lane].x;
      sy += red[q * 64 + lane].y;
    }
    vpart[((long long)b * nch + c) * m + k] = make_double2(sx, sy);
  }
}

// v[b, k] = sum_c vpart[b, c, k] (W^H W y over the observed samples).
__global__ void k_dense_vsum(const double2* __restrict__ vpart, int m, int nch, double2* __restrict__ v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (k >= m) return;
  double ax = 0.0, ay = 0.0;
  for (int c = 0; c < nch; ++c) {
    const double2 p = vpart[((long long)b * nch + c) * m + k];
    ax += p.x;
    ay += p.y;
  }
  v[(long long)b * m + k] = make_double2(ax, ay);
}

// t1[b, r, a, c] = sum_m2 R[r, a M2 + m2] exp(+2 pi i m2 c / N2), where R in
// the full sample layout holds the rows of conj(W) (r < m: R[r, j] =
// conj(W[r][pos j]) for 0 <= pos j <= r) and W^H W y (r = m: v[pos j]),
// zero elsewhere.  W is read in place through pos.
__global__ void __launch_bounds__(DN_THREADS)
k_dense_t1(const double2* __restrict__ ws, const double2* __restrict__ v, const int* __restrict__ pos, int m,
           int mp, int m1, int m2, int n2, double2* __restrict__ t1) {
  __shared__ double2 tw[DN_AXMAX];
  for (int i = threadIdx.x; i < n2; i += DN_THREADS) {
    double sn, cs;
    sincospi(2.0 * (double)i / n2, &sn, &cs);
    tw[i] = make_double2(cs, sn);
  }
  __syncthreads();
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * DN_THREADS + threadIdx.x;
  if (e >= (long long)(m + 1) * m1 * n2) return;
  const int c = (int)(e % n2);
  const long long rr = e / n2;   // r * m1 + a
  const int r = (int)(rr / m1), a = (int)(rr - (long long)r * m1);
  const double2* wr = ws + (long long)b * mp * mp + (long long)r * mp;
  const double2* vb = v + (long long)b * m;
  const int* pa = pos + a * m2;
  double ax = 0.0, ay = 0.0;
  int ph = 0;
  for (int k = 0; k < m2; ++k) {
    const int pk = pa[k];
    if (pk >= 0 && (r == m || pk <= r)) {
      double2 val;
      if (r == m) {
        val = vb[pk];
      } else {
        const double2 wv = wr[pk];
        val = make_double2(wv.x, -wv.y);
      }
      const double2 w = tw[ph];
      ax += val.x * w.x - val.y * w.y;
      ay += val.x * w.y + val.y * w.x;
    }
    ph += c;
    if (ph >= n2) ph -= n2;
  }
  t1[(long long)b * (m + 1) * m1 * n2 + e] = make_double2(ax, ay);
}

// Row transform of t1 row r at grid point (n1, n2):
// sum_m1 t1[r, m1, n2] exp(+2 pi i m1 n1 / N1).
__device__ __forceinline__ double2 dense_row_at(const double2* __restrict__ tb, int r, int m1, int n1, int n2, int r1,
                                                int c2, const double2* tw) {
  double ux = 0.0, uy = 0.0;
  int ph = 0;
  for (int k = 0; k < m1; ++k) {
    const double2 v = tb[((long long)r * m1 + k) * n2 + c2], w = tw[ph];
    ux += v.x * w.x - v.y * w.y;
    uy += v.x * w.y + v.y * w.x;
    ph += r1;
    if (ph >= n1) ph -= n1;
  }
  return make_double2(ux, uy);
}

// ddpart[b, ch, n] = sum over rows r < m of chunk ch of |row transform|^2.
__global__ void __launch_bounds__(DN_THREADS)
k_dense_ddpart(const double2* __restrict__ t1, int m, int m1, int n1, int n2, int nrch, double* __restrict__ ddpart) {
  __shared__ double2 tw[DN_AXMAX];
  for (int i = threadIdx.x; i < n1; i += DN_THREADS) {
    double sn, cs;
    sincospi(2.0 * (double)i / n1, &sn, &cs);
    tw[i] = make_double2(cs, sn);
  }
  __syncthreads();
  const int ch = blockIdx.y, b = blockIdx.z;
  const int n = n1 * n2;
  const int i = blockIdx.x * DN_THREADS + threadIdx.x;
  if (i >= n) return;
  const int r1 = i / n2, c2 = i - r1 * n2;
  const double2* tb = t1 + (long long)b * (m + 1) * m1 * n2;
  const int r_end = min(m, (ch + 1) * DN_RCH);
  double acc = 0.0;
  for (int r = ch * DN_RCH; r < r_end; ++r) {
    const double2 u = dense_row_at(tb, r, m1, n1, n2, r1, c2, tw);
    acc += u.x * u.x + u.y * u.y;
  }
  ddpart[((long long)b * nrch + ch) * n + i] = acc;
}

// Per grid point: dd = sum of the chunk partials, lifted = row transform of
// t1 row m; then mean = gamma lifted, variance, data diagonal and relevance
// partials (solver.py:2985-2990, 3016-3036).
__global__ void __launch_bounds__(DN_THREADS)
k_dense_moments(const double2* __restrict__ t1, const double* __restrict__ ddpart, int m, int m1, int n1, int n2,
                int nrch, const float* __restrict__ gamma, float2* __restrict__ mean, float* __restrict__ var,
                float* __restrict__ dd_out, double* __restrict__ part) {
  __shared__ double2 tw[DN_AXMAX];
  __shared__ double red[DN_THREADS];
  for (int i = threadIdx.x; i < n1; i += DN_THREADS) {
    double sn, cs;
    sincospi(2.0 * (double)i / n1, &sn, &cs);
    tw[i] = make_double2(cs, sn);
  }
  __syncthreads();
  const int b = blockIdx.y;
  const int n = n1 * n2;
  const int i = blockIdx.x * DN_THREADS + threadIdx.x;
  double rel = 0.0;
  if (i < n) {
    const int r1 = i / n2, c2 = i - r1 * n2;
    const double2* tb = t1 + (long long)b * (m + 1) * m1 * n2;
    double acc = 0.0;
    for (int ch = 0; ch < nrch; ++ch) acc += ddpart[((long long)b * nrch + ch) * n + i];
    const double2 l = dense_row_at(tb, m, m1, n1, n2, r1, c2, tw);
    const long long g = (long long)b * n + i;
    const double gm = gamma[g];
    mean[g] = make_float2((float)(gm * l.x), (float)(gm * l.y));
    var[g] = (float)fmax(gm - gm * gm * acc, 0.0);
    dd_out[g] = (float)acc;
    rel = fmin(fmax(gm * acc, 0.0), 1.0);
  }
  rel = block_sum_dn(rel, red);
  if (threadIdx.x == 0) part[(long long)b * gridDim.x + blockIdx.x] = rel;
}

__global__ void k_dense_seg(const double* __restrict__ part, int blocks, double* __restrict__ out) {
  const int b = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < blocks; i += 64) s += part[(long long)b * blocks + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[b] = s;
}

struct DenseLayout {
  size_t t, lags, q, w, linv, spart, pos, u, vpart, v, t1, ddpart, total;
};

inline size_t align256(size_t v) { return (v + 255) / 256 * 256; }

DenseLayout dense_layout(int batch, int n1, int n2, int m1, int m2, int m) {
  const size_t b = (size_t)batch, n = (size_t)n1 * n2, mf = (size_t)m1 * m2;
  DenseLayout l;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align256(off + bytes);
    return at;
  };
  l.t = take(16 * b * n1 * (2 * m2 - 1));
  l.lags = take(16 * b * (2 * m1 - 1) * (2 * m2 - 1));
  const size_t mp = (size_t)((m + 31) / 32) * 32;
  l.q = take(16 * b * mp * mp);
  l.w = take(16 * b * mp * mp);
  l.linv = take(16 * b * mp * 32);
  l.spart = take(16 * b * (mp / 32) * DN_KS * 32 * 32);
  const size_t nch = (size_t)((m + DN_VCH - 1) / DN_VCH), nrch = (size_t)((m + DN_RCH - 1) / DN_RCH);
  l.pos = take(4 * mf);
  l.u = take(16 * b * m);
  l.vpart = take(16 * b * nch * m);
  l.v = take(16 * b * m);
  l.t1 = take(16 * b * (m + 1) * (size_t)m1 * n2);
  l.ddpart = take(8 * b * nrch * n);
  l.total = off;
  return l;
}

// ------------------------------------ reduced posterior for K up to 256
// The general solver's active-set posterior (_active_posterior with the
// diffuse tail loading nu = noise + max(sum gamma - sum gamma_S, 0),
// solver.py:3591-3779, 3897-3924) for supports larger than one wavefront's
// 32 atoms: S = diag(1 / gamma_S) + G / nu (Gram G = C_S^H C_S from the
// Dirichlet tables or an observation mask's spectrum), equilibrated by
// D = sqrt(diag S), loaded with jitter * eps_f32 * mean|diag| (the GPU
// branch of _stable_cholesky, solver.py:2898-2920), factored and inverted on
// the FP64 32 x 32 tile kernels above (padded to a multiple of 32 with an
// identity block).
constexpr int RP_KMAX = 256;
constexpr double F32_EPS_DN = 1.1920928955078125e-07;
constexpr double F32_TINY_DN = 1.17549435082228750797e-38;

__device__ __forceinline__ double2 rp_lag(const double2* d1, const double2* d2, const double2* mh, int n1, int n2,
                                          int a, int b) {
  int l1 = a / n2 - b / n2, l2 = a % n2 - b % n2;
  l1 += (l1 < 0) ? n1 : 0;
  l2 += (l2 < 0) ? n2 : 0;
  if (mh) return mh[(long long)l1 * n2 + l2];
  return zm(d1[l1], d2[l2]);
}

// Per scene: nu, the equilibration D, the equilibrated right-hand side
// (C^H y / nu) / D with C^H y = (A^H y)[S], and the loading.
__global__ void __launch_bounds__(RP_KMAX)
k_rp_prep(int k, int n1, int n2, const int* __restrict__ support, const float* __restrict__ gamma,
          const double* __restrict__ gamma_sum, const float* __restrict__ noise, const float2* __restrict__ bp,
          const double2* __restrict__ d1, const double2* __restrict__ d2, const double2* __restrict__ mh,
          double jitter_multiplier, double* __restrict__ dsc, double2* __restrict__ rhs, double* __restrict__ rsc) {
  __shared__ double red[RP_KMAX];
  const int b = blockIdx.x, r = threadIdx.x;
  const long long n = (long long)n1 * n2;
  const int id = r < k ? support[(long long)b * k + r] : 0;
  const double g = r < k ? (double)gamma[b * n + id] : 0.0;
  red[r] = g;
  __syncthreads();
  for (int w = RP_KMAX / 2; w > 0; w >>= 1) {
    if (r < w) red[r] += red[r + w];
    __syncthreads();
  }
  const double nu = (double)noise[b] + fmax(gamma_sum[b] - red[0], 0.0);
  __syncthreads();
  double diag = 0.0;
  if (r < k) {
    const double srr = rp_lag(d1, d2, mh, n1, n2, id, id).x / nu + 1.0 / g;
    const double d = sqrt(fmax(srr, F32_TINY_DN));
    dsc[(long long)b * RP_KMAX + r] = d;
    diag = fabs(srr / (d * d));
    const float2 v = bp[b * n + id];
    rhs[(long long)b * RP_KMAX + r] = make_double2((double)v.x / nu / d, (double)v.y / nu / d);
  }
  red[r] = diag;
  __syncthreads();
  for (int w = RP_KMAX / 2; w > 0; w >>= 1) {
    if (r < w) red[r] += red[r + w];
    __syncthreads();
  }
  if (r == 0) {
    rsc[2 * b] = nu;
    rsc[2 * b + 1] = jitter_multiplier * F32_EPS_DN * fmax(red[0] / k, F32_TINY_DN);
  }
}

// Equilibrated, Hermitian-symmetrised and loaded system, padded to mp.
__global__ void k_rp_build(int k, int mp, int n1, int n2, const int* __restrict__ support,
                           const float* __restrict__ gamma, const double2* __restrict__ d1,
                           const double2* __restrict__ d2, const double2* __restrict__ mh,
                           const double* __restrict__ dsc, const double* __restrict__ rsc, double2* __restrict__ qs) {
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)mp * mp) return;
  const int r = (int)(e / mp), q = (int)(e % mp);
  double2 v = make_double2(r == q ? 1.0 : 0.0, 0.0);
  if (r < k && q < k) {
    const long long n = (long long)n1 * n2;
    const int ir = support[(long long)b * k + r], iq = support[(long long)b * k + q];
    const double inu = 1.0 / rsc[2 * b];
    const double2 grq = rp_lag(d1, d2, mh, n1, n2, ir, iq), gqr = rp_lag(d1, d2, mh, n1, n2, iq, ir);
    // 0.5 (S + S^H), S = G / nu + diag(1 / gamma_S)
    v = make_double2(0.5 * (grq.x + gqr.x) * inu, 0.5 * (grq.y - gqr.y) * inu);
    const double sc = 1.0 / (dsc[(long long)b * RP_KMAX + r] * dsc[(long long)b * RP_KMAX + q]);
    if (r == q) v = make_double2(v.x + 1.0 / (double)gamma[b * n + ir], 0.0);
    v = make_double2(v.x * sc, v.y * sc);
    if (r == q) v.x += rsc[2 * b + 1];
  }
  qs[(long long)b * mp * mp + e] = v;
}

// mu = W^H (W rhs) / D and Sigma_jj = ||W[:, j]||^2 / D_j^2 from W = L^-1,
// data diagonal (gamma - Sigma) / gamma^2, scattered to the support.
__global__ void __launch_bounds__(RP_KMAX)
k_rp_finish(int k, int mp, long long n, const int* __restrict__ support, const float* __restrict__ gamma,
            const double* __restrict__ dsc, const double2* __restrict__ rhs, const double2* __restrict__ ws,
            float2* __restrict__ mean, float* __restrict__ var, float* __restrict__ dd) {
  __shared__ double2 z[RP_KMAX];
  const int b = blockIdx.x, r = threadIdx.x;
  const double2* w = ws + (long long)b * mp * mp;
  const double2* rh = rhs + (long long)b * RP_KMAX;
  if (r < k) {
    double sx = 0.0, sy = 0.0;
    for (int p = 0; p <= r; ++p) {
      const double2 t = zm(w[(long long)r * mp + p], rh[p]);
      sx += t.x;
      sy += t.y;
    }
    z[r] = make_double2(sx, sy);
  }
  __syncthreads();
  if (r < k) {
    double mx = 0.0, my = 0.0, ev = 0.0;
    for (int i = r; i < k; ++i) {
      const double2 a = w[(long long)i * mp + r];
      const double2 zi = z[i];
      mx += a.x * zi.x + a.y * zi.y;   // conj(W[i][r]) z[i]
      my += a.x * zi.y - a.y * zi.x;
      ev += a.x * a.x + a.y * a.y;
    }
    const double d = dsc[(long long)b * RP_KMAX + r];
    const int id = support[(long long)b * k + r];
    const double g = (double)gamma[b * n + id];
    const double sg = fmax(ev / (d * d), 0.0);
    mean[b * n + id] = make_float2((float)(mx / d), (float)(my / d));
    var[b * n + id] = (float)sg;
    dd[b * n + id] = (float)fmax((g - sg) / (g * g), 0.0);
  }
}

__global__ void k_rp_defaults(const float* __restrict__ gamma, float2* __restrict__ mean, float* __restrict__ var,
                              float* __restrict__ dd, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  mean[i] = make_float2(0.f, 0.f);
  var[i] = gamma[i];
  dd[i] = 0.f;
}

struct RpLayout {
  size_t qs, ws, linv, spart, dsc, rhs, rsc, total;
};

RpLayout rp_layout(int batch, int k) {
  const long long mp = (k + TS - 1) / TS * TS, nb = mp / TS;
  RpLayout l{};
  size_t off = 0;
  auto reg = [&](size_t& at, size_t bytes) {
    off = (off + 255) / 256 * 256;
    at = off;
    off += bytes;
  };
  reg(l.qs, 16 * (size_t)(batch * mp * mp));
  reg(l.ws, 16 * (size_t)(batch * mp * mp));
  reg(l.linv, 16 * (size_t)(batch * nb * TS * TS));
  reg(l.spart, 16 * (size_t)(batch * nb * DN_KS * TS * TS));
  reg(l.dsc, 8 * (size_t)(batch * RP_KMAX));
  reg(l.rhs, 16 * (size_t)(batch * RP_KMAX));
  reg(l.rsc, 8 * 2 * (size_t)batch);
  l.total = (off + 255) / 256 * 256;
  return l;
}

}  // namespace

extern "C" {

static bool dense_args_ok(int batch, int n1, int n2, int m1, int m2, int observed) {
  // scenes run on grid.y (at most 65535); grid indices stay in int range
  return batch >= 1 && batch <= 65535 && m1 >= 1 && m2 >= 1 && m1 <= n1 && m2 <= n2 && n1 <= DN_AXMAX &&
         n2 <= DN_AXMAX && observed >= 1 && observed <= m1 * m2 && observed <= DN_MMAX;
}

size_t fbme_dense_workspace_bytes(int batch, int n1, int n2, int m1, int m2, int observed) {
  if (!dense_args_ok(batch, n1, n2, m1, m2, observed)) return 0;
  return dense_layout(batch, n1, n2, m1, m2, observed).total;
}

int fbme_dense_posterior(void* stream, int batch, int n1, int n2, int m1, int m2, int observed, const int* obs_idx,
                         const void* y, const float* gamma, const float* noise, double jitter, double tiny,
                         void* work, size_t work_bytes, void* mean, float* var, float* dd, double* part,
                         double* rel_sum, int* status) {
  if (!dense_args_ok(batch, n1, n2, m1, m2, observed) || !work || !y || !gamma || !noise || !mean || !var || !dd ||
      !part || !rel_sum)
    return FBME_ERR_ARG;
  if (!obs_idx && observed != m1 * m2) return FBME_ERR_ARG;
  if (work_bytes < dense_layout(batch, n1, n2, m1, m2, observed).total) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int m = observed, n = n1 * n2, mf = m1 * m2;
  const DenseLayout l = dense_layout(batch, n1, n2, m1, m2, m);
  char* w = (char*)work;
  double2* t = (double2*)(w + l.t);
  double2* lags = (double2*)(w + l.lags);
  double2* q = (double2*)(w + l.q);
  double2* wmat = (double2*)(w + l.w);
  double2* linv = (double2*)(w + l.linv);
  const int mp = (m + TS - 1) / TS * TS, nb = mp / TS;
  int* pos = (int*)(w + l.pos);
  double2* u = (double2*)(w + l.u);
  double2* vpart = (double2*)(w + l.vpart);
  double2* v = (double2*)(w + l.v);
  double2* t1 = (double2*)(w + l.t1);
  double* ddpart = (double*)(w + l.ddpart);
  const int nch = (m + DN_VCH - 1) / DN_VCH, nrch = (m + DN_RCH - 1) / DN_RCH;
  const long long rows_e = (long long)n1 * (2 * m2 - 1), lag_e = (long long)(2 * m1 - 1) * (2 * m2 - 1);
  hipLaunchKernelGGL(k_lag_rows, dim3((unsigned)((rows_e + 255) / 256), batch), dim3(256), 0, st, gamma, n1, n2, m2,
                     t);
  hipLaunchKernelGGL(k_lag_cols, dim3((unsigned)((lag_e + 255) / 256), batch), dim3(256), 0, st, t, n1, m1, m2,
                     lags);
  hipLaunchKernelGGL(k_dense_build, dim3((unsigned)(((long long)mp * mp + 255) / 256), batch), dim3(256), 0, st,
                     lags, obs_idx, m, mp, m1, m2, noise, jitter, tiny, q, status);
  for (int jb = 0; jb < nb; ++jb) {
    hipLaunchKernelGGL(k_tile_diag, dim3(batch), dim3(DN_THREADS), 0, st, q, mp, jb, tiny, linv, status);
    const int rest = nb - jb - 1;
    if (rest > 0) {
      hipLaunchKernelGGL(k_tile_panel, dim3(rest, batch), dim3(DN_THREADS), 0, st, q, mp, jb, (const double2*)linv);
      hipLaunchKernelGGL(k_tile_update, dim3(rest * (rest + 1) / 2, batch), dim3(DN_THREADS), 0, st, q, mp, jb);
    }
  }
  double2* spart = (double2*)(w + l.spart);
  for (int ib = 0; ib < nb; ++ib) {
    if (ib > 0)
      hipLaunchKernelGGL(k_tile_inverse_part, dim3(ib, DN_KS, batch), dim3(DN_THREADS), 0, st, (const double2*)q, mp,
                         ib, (const double2*)wmat, spart);
    hipLaunchKernelGGL(k_tile_inverse_row, dim3(ib + 1, batch), dim3(DN_THREADS), 0, st, mp, ib,
                       (const double2*)linv, (const double2*)spart, wmat);
  }
  // Q^-1 y = W^H W y and the rows of conj(W), both in the full sample layout
  hipLaunchKernelGGL(k_dense_pos_fill, dim3((mf + 255) / 256), dim3(256), 0, st, obs_idx ? 1 : 0, mf, pos);
  if (obs_idx) hipLaunchKernelGGL(k_dense_pos_scatter, dim3((m + 255) / 256), dim3(256), 0, st, obs_idx, m, mf, pos);
  hipLaunchKernelGGL(k_dense_u, dim3((m + DN_THREADS / 64 - 1) / (DN_THREADS / 64), batch), dim3(DN_THREADS), 0, st,
                     (const double2*)wmat, (const float2*)y, obs_idx, m, mp, mf, u);
  hipLaunchKernelGGL(k_dense_wh, dim3((m + 63) / 64, nch, batch), dim3(DN_THREADS), 0, st, (const double2*)wmat,
                     (const double2*)u, m, mp, nch, vpart);
  hipLaunchKernelGGL(k_dense_vsum, dim3((m + 255) / 256, batch), dim3(256), 0, st, (const double2*)vpart, m, nch, v);
  const long long t1_e = (long long)(m + 1) * m1 * n2;
  hipLaunchKernelGGL(k_dense_t1, dim3((unsigned)((t1_e + DN_THREADS - 1) / DN_THREADS), batch), dim3(DN_THREADS), 0,
                     st, (const double2*)wmat, (const double2*)v, (const int*)pos, m, mp, m1, m2, n2, t1);
  const int blocks = (n + DN_THREADS - 1) / DN_THREADS;
  hipLaunchKernelGGL(k_dense_ddpart, dim3(blocks, nrch, batch), dim3(DN_THREADS), 0, st, (const double2*)t1, m, m1,
                     n1, n2, nrch, ddpart);
  hipLaunchKernelGGL(k_dense_moments, dim3(blocks, batch), dim3(DN_THREADS), 0, st, (const double2*)t1,
                     (const double*)ddpart, m, m1, n1, n2, nrch, gamma, (float2*)mean, var, dd, part);
  hipLaunchKernelGGL(k_dense_seg, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, rel_sum);
  return (int)hipGetLastError();
}


size_t fbme_reduced_posterior_large_bytes(int batch, int k) {
  if (batch < 1 || batch > 65535 || k < 1 || k > RP_KMAX) return 0;
  return rp_layout(batch, k).total;
}

int fbme_reduced_posterior_large(void* stream, int batch, int k, int n1, int n2, const int* support,
                                 const float* gamma, const double* gamma_sum, const float* noise, const void* bp,
                                 const void* dir1, const void* dir2, const void* mask_spectrum,
                                 float jitter_multiplier, void* work, size_t work_bytes, void* mean, float* var,
                                 float* dd, int* status) {
  if (batch < 1 || batch > 65535 || k < 1 || k > RP_KMAX || n1 < 1 || n2 < 1 || (long long)n1 * n2 < k)
    return FBME_ERR_ARG;
  if (!support || !gamma || !gamma_sum || !noise || !bp || !work || !mean || !var || !dd || !status)
    return FBME_ERR_ARG;
  if (!mask_spectrum && (!dir1 || !dir2)) return FBME_ERR_ARG;
  const RpLayout l = rp_layout(batch, k);
  if (work_bytes < l.total) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  char* w = (char*)work;
  double2* qs = (double2*)(w + l.qs);
  double2* ws = (double2*)(w + l.ws);
  double2* linv = (double2*)(w + l.linv);
  double2* spart = (double2*)(w + l.spart);
  double* dsc = (double*)(w + l.dsc);
  double2* rhs = (double2*)(w + l.rhs);
  double* rsc = (double*)(w + l.rsc);
  const int mp = (k + TS - 1) / TS * TS, nb = mp / TS;
  const long long n = (long long)n1 * n2, total = (long long)batch * n;
  const double2* d1 = (const double2*)dir1;
  const double2* d2 = (const double2*)dir2;
  const double2* mh = (const double2*)mask_spectrum;
  hipLaunchKernelGGL(k_rp_defaults, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, gamma, (float2*)mean,
                     var, dd, total);
  hipLaunchKernelGGL(k_rp_prep, dim3(batch), dim3(RP_KMAX), 0, st, k, n1, n2, support, gamma, gamma_sum, noise,
                     (const float2*)bp, d1, d2, mh, (double)jitter_multiplier, dsc, rhs, rsc);
  hipLaunchKernelGGL(k_rp_build, dim3((unsigned)(((long long)mp * mp + 255) / 256), batch), dim3(256), 0, st, k, mp,
                     n1, n2, support, gamma, d1, d2, mh, (const double*)dsc, (const double*)rsc, qs);
  for (int jb = 0; jb < nb; ++jb) {
    hipLaunchKernelGGL(k_tile_diag, dim3(batch), dim3(DN_THREADS), 0, st, qs, mp, jb, F32_TINY_DN, linv, status);
    const int rest = nb - jb - 1;
    if (rest > 0) {
      hipLaunchKernelGGL(k_tile_panel, dim3(rest, batch), dim3(DN_THREADS), 0, st, qs, mp, jb, (const double2*)linv);
      hipLaunchKernelGGL(k_tile_update, dim3(rest * (rest + 1) / 2, batch), dim3(DN_THREADS), 0, st, qs, mp, jb);
    }
  }
  for (int ib = 0; ib < nb; ++ib) {
    if (ib > 0)
      hipLaunchKernelGGL(k_tile_inverse_part, dim3(ib, DN_KS, batch), dim3(DN_THREADS), 0, st, (const double2*)qs,
                         mp, ib, (const double2*)ws, spart);
    hipLaunchKernelGGL(k_tile_inverse_row, dim3(ib + 1, batch), dim3(DN_THREADS), 0, st, mp, ib,
                       (const double2*)linv, (const double2*)spart, ws);
  }
  hipLaunchKernelGGL(k_rp_finish, dim3(batch), dim3(RP_KMAX), 0, st, k, mp, n, support, gamma, (const double*)dsc,
                     (const double2*)rhs, (const double2*)ws, (float2*)mean, var, dd);
  return (int)hipGetLastError();
}

}  // extern "C"
