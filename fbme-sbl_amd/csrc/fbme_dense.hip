// Exact dense posterior of the general solver, mode="dense" (gfx950):
// _posterior's dense branch (torch_sbl/solver.py:2962-2990).
//
//   * Q = noise I + A Gamma A^H over the observed samples, from the Fourier
//     lags of gamma (operators.py:282-297): c[l] = sum_n gamma_n
//     exp(-2 pi i l.n / N) at the (2 M1 - 1)(2 M2 - 1) sample-index
//     differences, a separable FP64 DFT;
//   * Cholesky with the GPU's one deterministic loading jitter * mean|diag|
//     (solver.py:2898-2920), then the triangular inverse W = L^-1, FP64, one
//     workgroup per scene (left-looking, the pivot row staged in LDS);
//   * Q^-1 y = W^H (W y);
//   * the data diagonal diag(A^H Q^-1 A)_n = sum_k |(A^H conj(w_k))_n|^2 over
//     the rows w_k of W (a_n^H Q^-1 a_n = ||W a_n||^2; the reference
//     accumulates Q^-1 along its Fourier lags, operators.py:418-442, equally
//     exact), and A^H Q^-1 y, from separable FP64 prefix transforms of those
//     rows (twiddles from per-block tables, phases reduced mod N).
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
// (obs[r] = sample index m1 * M2 + m2 of row r; NULL: all samples).
__global__ void k_dense_build(const double2* __restrict__ lags, const int* __restrict__ obs, int m, int m1, int m2,
                              const float* __restrict__ noise, double2* __restrict__ q) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= (long long)m * m) return;
  const int r = (int)(e / m), c = (int)(e - (long long)r * m);
  const int sr = obs ? obs[r] : r, sc = obs ? obs[c] : c;
  const int d1 = sr / m2 - sc / m2, d2 = sr % m2 - sc % m2;
  const long long li = (long long)b * (2 * m1 - 1) * (2 * m2 - 1) + (long long)(d1 + m1 - 1) * (2 * m2 - 1) +
                       (d2 + m2 - 1);
  double2 v = lags[li];
  if (r == c) v.x += (double)noise[b];
  q[(long long)b * m * m + e] = v;
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

// In place: Q + jitter * mean|Re diag Q| I = L L^H (left-looking, row j of
// L staged in LDS), then W = L^-1 (row by row, row i of L staged in LDS);
// the upper triangle is zeroed.  One workgroup per scene.
__global__ void __launch_bounds__(DN_THREADS)
k_dense_factor(double2* __restrict__ qs, int m, double jitter, double tiny, int* __restrict__ status) {
  __shared__ double2 row[DN_MMAX];
  __shared__ double red[DN_THREADS];
  __shared__ double piv_s;
  __shared__ int bad;
  const int t = threadIdx.x, b = blockIdx.x;
  double2* q = qs + (long long)b * m * m;
  double s = 0.0;
  for (int i = t; i < m; i += DN_THREADS) s += fabs(q[(long long)i * m + i].x);
  const double load = jitter * fmax(block_sum_dn(s, red) / m, tiny);
  if (t == 0) bad = 0;
  for (int i = t; i < m; i += DN_THREADS) q[(long long)i * m + i].x += load;
  __syncthreads();
  for (int j = 0; j < m; ++j) {
    for (int k = t; k < j; k += DN_THREADS) row[k] = q[(long long)j * m + k];
    __syncthreads();
    double d = 0.0;
    for (int k = t; k < j; k += DN_THREADS) d += row[k].x * row[k].x + row[k].y * row[k].y;
    d = block_sum_dn(d, red);
    if (t == 0) {
      const double v = q[(long long)j * m + j].x - d;
      if (!(v > 0.0) || !isfinite(v)) bad = 1;
      piv_s = sqrt(fmax(v, tiny));
      q[(long long)j * m + j] = make_double2(piv_s, 0.0);
    }
    __syncthreads();
    const double ip = 1.0 / piv_s;
    for (int i = j + 1 + t; i < m; i += DN_THREADS) {
      const double2* ri = q + (long long)i * m;
      double ax = ri[j].x, ay = ri[j].y;
      for (int k = 0; k < j; ++k) {
        // a -= L[i][k] conj(L[j][k])
        const double2 u = ri[k], w = row[k];
        ax -= u.x * w.x + u.y * w.y;
        ay -= u.y * w.x - u.x * w.y;
      }
      q[(long long)i * m + j] = make_double2(ax * ip, ay * ip);
    }
    __syncthreads();
  }
  // W = L^-1: W[i][c] = -(sum_{k=c}^{i-1} L[i][k] W[k][c]) / L[i][i]
  for (int i = 0; i < m; ++i) {
    for (int k = t; k <= i; k += DN_THREADS) row[k] = q[(long long)i * m + k];
    __syncthreads();
    const double il = 1.0 / row[i].x;
    for (int c = t; c < i; c += DN_THREADS) {
      double ax = 0.0, ay = 0.0;
      for (int k = c; k < i; ++k) {
        const double2 l = row[k], w = q[(long long)k * m + c];
        ax += l.x * w.x - l.y * w.y;
        ay += l.x * w.y + l.y * w.x;
      }
      q[(long long)i * m + c] = make_double2(-ax * il, -ay * il);
    }
    if (t == 0) q[(long long)i * m + i] = make_double2(il, 0.0);
    __syncthreads();
  }
  for (long long e = t; e < (long long)m * m; e += DN_THREADS) {
    const int r = (int)(e / m), c = (int)(e - (long long)r * m);
    if (c > r) q[e] = make_double2(0.0, 0.0);
  }
  if (t == 0 && bad && status) status[b] = 1;
}

// Rows of the transforms, FP64 in the full sample layout (zeros on the
// unobserved samples): rows 0..m-1 = conj(W[k, :]), row m = z = W^H (W y).
__global__ void __launch_bounds__(DN_THREADS)
k_dense_apply(const double2* __restrict__ ws, const float2* __restrict__ y, const int* __restrict__ obs, int m,
              int mfull, double2* __restrict__ rows) {
  __shared__ double2 u[DN_MMAX];
  const int t = threadIdx.x, b = blockIdx.x;
  const double2* w = ws + (long long)b * m * m;
  const float2* yb = y + (long long)b * mfull;
  double2* rb = rows + (long long)b * (m + 1) * mfull;
  for (long long e = t; e < (long long)(m + 1) * mfull; e += DN_THREADS) rb[e] = make_double2(0.0, 0.0);
  for (int i = t; i < m; i += DN_THREADS) {
    double ax = 0.0, ay = 0.0;
    for (int k = 0; k <= i; ++k) {
      const double2 a = w[(long long)i * m + k];
      const float2 v = yb[obs ? obs[k] : k];
      ax += a.x * v.x - a.y * v.y;
      ay += a.x * v.y + a.y * v.x;
    }
    u[i] = make_double2(ax, ay);
  }
  __syncthreads();
  for (int k = t; k < m; k += DN_THREADS) {
    double ax = 0.0, ay = 0.0;
    for (int i = k; i < m; ++i) {
      const double2 a = w[(long long)i * m + k];   // conj(W[i][k]) u_i
      ax += a.x * u[i].x + a.y * u[i].y;
      ay += a.x * u[i].y - a.y * u[i].x;
    }
    rb[(long long)m * mfull + (obs ? obs[k] : k)] = make_double2(ax, ay);
  }
  for (long long e = t; e < (long long)m * m; e += DN_THREADS) {
    const int r = (int)(e / m), c = (int)(e - (long long)r * m);
    const double2 a = w[e];
    rb[(long long)r * mfull + (obs ? obs[c] : c)] = make_double2(a.x, -a.y);
  }
}

// t1[b, r, m1, n2] = sum_m2 rows[b, r, m1 M2 + m2] exp(+2 pi i m2 n2 / N2).
__global__ void __launch_bounds__(DN_THREADS)
k_dense_t1(const double2* __restrict__ rows, int nrows, int m1, int m2, int n2, double2* __restrict__ t1) {
  __shared__ double2 tw[1024];
  for (int i = threadIdx.x; i < n2; i += DN_THREADS) {
    double sn, cs;
    sincospi(2.0 * (double)i / n2, &sn, &cs);
    tw[i] = make_double2(cs, sn);
  }
  __syncthreads();
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * DN_THREADS + threadIdx.x;
  if (e >= (long long)nrows * m1 * n2) return;
  const int c = (int)(e % n2);
  const long long rr = e / n2;                       // r * m1 + row
  const double2* src = rows + ((long long)b * nrows * m1 + rr) * m2;
  double ax = 0.0, ay = 0.0;
  int ph = 0;
  for (int k = 0; k < m2; ++k) {
    const double2 v = src[k], w = tw[ph];
    ax += v.x * w.x - v.y * w.y;
    ay += v.x * w.y + v.y * w.x;
    ph += c;
    if (ph >= n2) ph -= n2;
  }
  t1[(long long)b * nrows * m1 * n2 + e] = make_double2(ax, ay);
}

// Per grid point n = (n1, n2): u_r = sum_m1 t1[b, r, m1, n2] exp(+2 pi i m1 n1 / N1);
// dd = sum_{r<m} |u_r|^2, lifted = u_m; then mean = gamma lifted, variance,
// data diagonal and relevance partials (solver.py:2985-2990, 3016-3036).
__global__ void __launch_bounds__(DN_THREADS)
k_dense_moments(const double2* __restrict__ t1, int m, int m1, int n1, int n2, const float* __restrict__ gamma,
                float2* __restrict__ mean, float* __restrict__ var, float* __restrict__ dd_out,
                double* __restrict__ part) {
  __shared__ double2 tw[1024];
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
    double acc = 0.0, lx = 0.0, ly = 0.0;
    for (int r = 0; r <= m; ++r) {
      double ux = 0.0, uy = 0.0;
      int ph = 0;
      for (int k = 0; k < m1; ++k) {
        const double2 v = tb[((long long)r * m1 + k) * n2 + c2], w = tw[ph];
        ux += v.x * w.x - v.y * w.y;
        uy += v.x * w.y + v.y * w.x;
        ph += r1;
        if (ph >= n1) ph -= n1;
      }
      if (r < m) {
        acc += ux * ux + uy * uy;
      } else {
        lx = ux;
        ly = uy;
      }
    }
    const long long g = (long long)b * n + i;
    const double gm = gamma[g];
    mean[g] = make_float2((float)(gm * lx), (float)(gm * ly));
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
  size_t t, lags, q, rows, t1, total;
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
  l.q = take(16 * b * m * m);
  l.rows = take(16 * b * (m + 1) * mf);
  l.t1 = take(16 * b * (m + 1) * (size_t)m1 * n2);
  l.total = off;
  (void)n;
  return l;
}

}  // namespace

extern "C" {

size_t fbme_dense_workspace_bytes(int batch, int n1, int n2, int m1, int m2, int observed) {
  if (batch < 1 || observed < 1 || observed > DN_MMAX) return 0;
  return dense_layout(batch, n1, n2, m1, m2, observed).total;
}

int fbme_dense_posterior(void* stream, int batch, int n1, int n2, int m1, int m2, int observed, const int* obs_idx,
                         const void* y, const float* gamma, const float* noise, double jitter, double tiny,
                         void* work, void* mean, float* var, float* dd, double* part, double* rel_sum,
                         int* status) {
  if (batch < 1 || m1 < 1 || m2 < 1 || m1 > n1 || m2 > n2 || observed < 1 || observed > m1 * m2 ||
      observed > DN_MMAX || !work)
    return FBME_ERR_ARG;
  if (!obs_idx && observed != m1 * m2) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int m = observed, n = n1 * n2, mf = m1 * m2;
  const DenseLayout l = dense_layout(batch, n1, n2, m1, m2, m);
  char* w = (char*)work;
  double2* t = (double2*)(w + l.t);
  double2* lags = (double2*)(w + l.lags);
  double2* q = (double2*)(w + l.q);
  double2* rows = (double2*)(w + l.rows);
  double2* t1 = (double2*)(w + l.t1);
  const long long rows_e = (long long)n1 * (2 * m2 - 1), lag_e = (long long)(2 * m1 - 1) * (2 * m2 - 1);
  hipLaunchKernelGGL(k_lag_rows, dim3((unsigned)((rows_e + 255) / 256), batch), dim3(256), 0, st, gamma, n1, n2, m2,
                     t);
  hipLaunchKernelGGL(k_lag_cols, dim3((unsigned)((lag_e + 255) / 256), batch), dim3(256), 0, st, t, n1, m1, m2,
                     lags);
  hipLaunchKernelGGL(k_dense_build, dim3((unsigned)(((long long)m * m + 255) / 256), batch), dim3(256), 0, st, lags,
                     obs_idx, m, m1, m2, noise, q);
  hipLaunchKernelGGL(k_dense_factor, dim3(batch), dim3(DN_THREADS), 0, st, q, m, jitter, tiny, status);
  hipLaunchKernelGGL(k_dense_apply, dim3(batch), dim3(DN_THREADS), 0, st, q, (const float2*)y, obs_idx, m, mf,
                     rows);
  const long long t1_e = (long long)(m + 1) * m1 * n2;
  hipLaunchKernelGGL(k_dense_t1, dim3((unsigned)((t1_e + DN_THREADS - 1) / DN_THREADS), batch), dim3(DN_THREADS), 0,
                     st, rows, m + 1, m1, m2, n2, t1);
  const int blocks = (n + DN_THREADS - 1) / DN_THREADS;
  hipLaunchKernelGGL(k_dense_moments, dim3(blocks, batch), dim3(DN_THREADS), 0, st, t1, m, m1, n1, n2, gamma,
                     (float2*)mean, var, dd, part);
  hipLaunchKernelGGL(k_dense_seg, dim3(batch), dim3(64), 0, st, (const double*)part, blocks, rel_sum);
  return (int)hipGetLastError();
}

}  // extern "C"
