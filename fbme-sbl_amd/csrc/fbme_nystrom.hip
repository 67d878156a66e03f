// Randomised Nystrom preconditioner of the general solver (gfx950):
// _make_nystrom_preconditioner / _apply_nystrom_preconditioner of
// torch_sbl/solver.py:5851-5976.
//
// The reference orthonormalises the test matrix Omega (QR), forms
// Y = A Gamma A^H Q, core = Q^H Y, factor = Y L^-H (L the jitter-loaded
// Cholesky factor of core) and takes the SVD of the factor.  Everything but
// the two M x r products is r x r (r <= 32), so here:
//   * the r x r Grams G = Omega^H Omega, C = Omega^H Y and W = Y^H Y of the
//     unorthonormalised Omega and Y = A Gamma A^H Omega come from one pass
//     over the 2 r sample vectors (FP64 sums; the covariance products are the
//     solver's own matvec passes with a zero noise);
//   * with G = L_G L_G^H (Q = Omega L_G^-H), core = L_G^-1 C L_G^-H and the
//     factor is Y T with T = (L_G L_c)^-H, so factor^H factor = T^H W T.  Its
//     eigenpairs (cyclic complex Jacobi, FP64, one workgroup) give the
//     eigenvalues lambda = s^2 and U = Y T V diag(lambda)^-1/2, formed by one
//     more pass over Y;
//   * the apply v + U ((w - 1) o U^H v), w = (lambda_min + bg) / (bg +
//     lambda), is the stored-column Woodbury apply of fbme_pcg.hip with
//     S^-1 = diag(1 - w).
#include "fbme_device.h"

using namespace fbme;

namespace {

constexpr int NR = 32;            // largest rank
constexpr int NL = NR + 1;        // padded LDS row
constexpr int GB = 32;            // Gram blocks per system
constexpr int GT = 64;            // samples per Gram tile
constexpr double EPS32 = 1.1920928955078125e-07;
constexpr double TINY32 = 1.17549435082228750797e-38;

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zcmul(double2 a, double2 b) {   // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

// Omega [b, 32, M]: columns k < ri the atoms idx[b, k] (exact phases),
// ri <= k < r the Gaussian vectors rnd [b, r - ri, M], the rest zero.
__global__ void k_nys_omega(const int* __restrict__ idx, const float2* __restrict__ rnd, int n1, int n2, int m1,
                            int m2, int r, int ri, float2* __restrict__ omega) {
  const int b = blockIdx.y;
  const long long m = (long long)m1 * m2;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int s1 = (int)(i / m2), s2 = (int)(i - (long long)s1 * m2);
  float2* out = omega + (long long)b * NR * m + i;
  for (int k = 0; k < NR; ++k) {
    float2 v = make_float2(0.f, 0.f);
    if (k < ri) {
      const int g = idx[b * NR + k];
      const int k1 = g / n2, k2 = g - k1 * n2;
      const double ph = (double)(((long long)s1 * k1) % n1) / n1 + (double)(((long long)s2 * k2) % n2) / n2;
      double sn, cs;
      sincospi(-2.0 * ph, &sn, &cs);
      v = make_float2((float)cs, (float)sn);
    } else if (k < r) {
      v = rnd[((long long)b * (r - ri) + (k - ri)) * m + i];
    }
    out[(long long)k * m] = v;
  }
}

// Partial Grams over the tiles blockIdx.x, blockIdx.x + GB, ... of system b:
// part[b, block, g, i, j] for g = 0 (Omega^H Omega), 1 (Omega^H Y), 2 (Y^H Y).
// Thread t owns row i = t / 8 and columns j = 4 (t % 8) .. + 3.
__global__ void __launch_bounds__(FBME_BLOCK)
k_nys_grams(const float2* __restrict__ omega, const float2* __restrict__ y, long long m,
            double2* __restrict__ part) {
  __shared__ float2 so[NR][GT + 1];
  __shared__ float2 sy[NR][GT + 1];
  const int b = blockIdx.y;
  const int t = threadIdx.x;
  const int i = t >> 3, j0 = (t & 7) * 4;
  const float2* ob = omega + (long long)b * NR * m;
  const float2* yb = y + (long long)b * NR * m;
  double2 acc[3][4];
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[g][q] = make_double2(0.0, 0.0);
  const long long tiles = (m + GT - 1) / GT;
  for (long long tl = blockIdx.x; tl < tiles; tl += GB) {
    const long long base = tl * GT;
    __syncthreads();
    for (int e = t; e < NR * GT; e += FBME_BLOCK) {
      const int k = e / GT, c = e - k * GT;
      const bool live = base + c < m;
      so[k][c] = live ? ob[(long long)k * m + base + c] : make_float2(0.f, 0.f);
      sy[k][c] = live ? yb[(long long)k * m + base + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int c = 0; c < GT; ++c) {
      const float2 oi = so[i][c], yi = sy[i][c];
      const double2 a = make_double2(oi.x, oi.y), ay = make_double2(yi.x, yi.y);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 oj = so[j0 + q][c], yj = sy[j0 + q][c];
        const double2 bo = make_double2(oj.x, oj.y), by = make_double2(yj.x, yj.y);
        const double2 p0 = zcmul(a, bo), p1 = zcmul(a, by), p2 = zcmul(ay, by);
        acc[0][q].x += p0.x;
        acc[0][q].y += p0.y;
        acc[1][q].x += p1.x;
        acc[1][q].y += p1.y;
        acc[2][q].x += p2.x;
        acc[2][q].y += p2.y;
      }
    }
  }
  double2* out = part + ((long long)b * GB + blockIdx.x) * 3 * NR * NR;
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int q = 0; q < 4; ++q) out[(long long)g * NR * NR + i * NR + j0 + q] = acc[g][q];
}

// In-place lower Cholesky of the Hermitian r x r matrix a; false on a
// pivot that is not positive (or not finite).
__device__ bool chol_lds(double2 (*a)[NL], int r, int* flag) {
  const int t = threadIdx.x;
  if (t == 0) *flag = 0;
  __syncthreads();
  for (int j = 0; j < r; ++j) {
    if (t == 0) {
      const double d = a[j][j].x;
      if (!(d > 0.0) || !isfinite(d)) *flag = 1;
      a[j][j] = make_double2(sqrt(fmax(d, 0.0)), 0.0);
    }
    __syncthreads();
    const double piv = a[j][j].x;
    for (int i = j + 1 + t; i < r; i += FBME_BLOCK) {
      const double s = piv > 0.0 ? 1.0 / piv : 0.0;
      a[i][j] = make_double2(a[i][j].x * s, a[i][j].y * s);
    }
    __syncthreads();
    const int rem = r - j - 1;
    for (int e = t; e < rem * rem; e += FBME_BLOCK) {
      const int i = j + 1 + e / rem, q = j + 1 + e % rem;
      if (q <= i) {
        const double2 u = zmul(a[i][j], zconj(a[q][j]));
        a[i][q] = make_double2(a[i][q].x - u.x, a[i][q].y - u.y);
      }
    }
    __syncthreads();
  }
  for (int e = t; e < r * r; e += FBME_BLOCK) {   // zero the strict upper part
    const int i = e / r, q = e % r;
    if (q > i) a[i][q] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  return *flag == 0;
}

// x <- L^-1 x for the r columns of x (thread j < r solves column j).
__device__ void lower_solve(const double2 (*l)[NL], double2 (*x)[NL], int r) {
  const int j = threadIdx.x;
  if (j < r) {
    for (int i = 0; i < r; ++i) {
      double2 s = x[i][j];
      for (int p = 0; p < i; ++p) {
        const double2 u = zmul(l[i][p], x[p][j]);
        s.x -= u.x;
        s.y -= u.y;
      }
      const double d = l[i][i].x;
      x[i][j] = make_double2(s.x / d, s.y / d);
    }
  }
  __syncthreads();
}

// c = a b^H (bh) or a b, r x r.
__device__ void matmul(const double2 (*a)[NL], const double2 (*b)[NL], double2 (*c)[NL], int r, bool bh) {
  for (int e = threadIdx.x; e < r * r; e += FBME_BLOCK) {
    const int i = e / r, j = e % r;
    double2 s = make_double2(0.0, 0.0);
    for (int p = 0; p < r; ++p) {
      const double2 u = bh ? zmul(a[i][p], zconj(b[j][p])) : zmul(a[i][p], b[p][j]);
      s.x += u.x;
      s.y += u.y;
    }
    c[i][j] = s;
  }
  __syncthreads();
}

// The r x r phase of one system: Grams -> (lambda, Z, S^-1 = diag(1 - w)).
__global__ void __launch_bounds__(FBME_BLOCK)
k_nys_small(const double2* __restrict__ part, int r, long long m, const double* __restrict__ gamma_sum,
            const float* __restrict__ noise, float2* __restrict__ zout, float2* __restrict__ cinv,
            float* __restrict__ lam_out, float* __restrict__ bg_out) {
  __shared__ double2 ga[NR][NL], ca[NR][NL], wa[NR][NL], aa[NR][NL];
  __shared__ double lam[NR], red[FBME_BLOCK];
  __shared__ double prm[NR / 2][3];       // c, s, phase angle of each pair's rotation
  __shared__ double2 pe[NR / 2];
  __shared__ int pp[NR / 2], pq[NR / 2], rk[NR];
  __shared__ int flag, ok;
  const int b = blockIdx.x, t = threadIdx.x;
  // Grams, summed over the blocks in a fixed order
  const double2* pb = part + (long long)b * GB * 3 * NR * NR;
  for (int e = t; e < r * r; e += FBME_BLOCK) {
    const int i = e / r, j = e % r;
    double2 s0 = make_double2(0.0, 0.0), s1 = s0, s2 = s0;
    for (int q = 0; q < GB; ++q) {
      const double2* p = pb + (long long)q * 3 * NR * NR + i * NR + j;
      s0.x += p[0].x;
      s0.y += p[0].y;
      s1.x += p[NR * NR].x;
      s1.y += p[NR * NR].y;
      s2.x += p[2 * NR * NR].x;
      s2.y += p[2 * NR * NR].y;
    }
    ga[i][j] = s0;
    ca[i][j] = s1;
    wa[i][j] = s2;
  }
  if (t == 0) ok = 1;
  __syncthreads();
  // G = L_G L_G^H (ga); X = L_G^-1 C (ca); core = (L_G^-1 X^H)^H (aa)
  if (!chol_lds(ga, r, &flag) && t == 0) ok = 0;
  __syncthreads();
  lower_solve(ga, ca, r);
  for (int e = t; e < r * r; e += FBME_BLOCK) {
    const int i = e / r, j = e % r;
    aa[i][j] = zconj(ca[j][i]);
  }
  __syncthreads();
  lower_solve(ga, aa, r);   // aa = L_G^-1 X^H = core^H
  // symmetrised core (the Hermitian part) and its jitter-loaded factor
  // (10 eps_f32 mean|diag|, solver.py:5925-5932 via _stable_cholesky_raw)
  for (int e = t; e < r * r; e += FBME_BLOCK) {
    const int i = e / r, j = e % r;
    if (j <= i) {
      const double2 u = aa[i][j], v = aa[j][i];   // core = aa^H: core[i][j] = conj(aa[j][i])
      const double2 h = make_double2(0.5 * (v.x + u.x), 0.5 * (-v.y + u.y));
      ca[i][j] = h;
      ca[j][i] = zconj(h);
    }
  }
  __syncthreads();
  red[t] = (t < r) ? fabs(ca[t][t].x) : 0.0;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int i = 0; i < r; ++i) s += red[i];
    red[FBME_BLOCK - 1] = 10.0 * EPS32 * fmax(s / r, TINY32);
  }
  __syncthreads();
  const double jit = red[FBME_BLOCK - 1];
  if (t < r) ca[t][t].x += jit;
  __syncthreads();
  if (!chol_lds(ca, r, &flag) && t == 0) ok = 0;
  __syncthreads();
  // K = L_G L_c (aa), Kinv = K^-1 (ga), T = Kinv^H
  matmul(ga, ca, aa, r, false);
  for (int e = t; e < r * r; e += FBME_BLOCK) {
    const int i = e / r, j = e % r;
    ga[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  lower_solve(aa, ga, r);
  // H = Kinv W Kinv^H (wa), Hermitian
  matmul(ga, wa, aa, r, false);
  matmul(aa, ga, wa, r, true);
  for (int e = t; e < r * r; e += FBME_BLOCK) {
    const int i = e / r, j = e % r;
    if (j < i) {
      const double2 u = wa[i][j], v = wa[j][i];
      const double2 h = make_double2(0.5 * (u.x + v.x), 0.5 * (u.y - v.y));
      wa[i][j] = h;
      wa[j][i] = zconj(h);
    } else if (j == i) {
      wa[i][i].y = 0.0;
    }
  }
  // eigenvectors V (aa) by cyclic Jacobi over the round-robin pairing of
  // n = r (+1 when odd; the extra index is a zero row and column)
  const int n = r + (r & 1), np = n / 2;
  for (int e = t; e < n * n; e += FBME_BLOCK) {
    const int i = e / n, j = e % n;
    aa[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    if (i >= r || j >= r) wa[i][j] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int sweep = 0; sweep < 20; ++sweep) {
    double off = 0.0, all = 0.0;
    for (int e = t; e < n * n; e += FBME_BLOCK) {
      const int i = e / n, j = e % n;
      const double a2 = wa[i][j].x * wa[i][j].x + wa[i][j].y * wa[i][j].y;
      all += a2;
      if (i != j) off += a2;
    }
    red[t] = off;
    __syncthreads();
    for (int s = FBME_BLOCK / 2; s > 0; s >>= 1) {
      if (t < s) red[t] += red[t + s];
      __syncthreads();
    }
    const double off_sum = red[0];
    __syncthreads();
    red[t] = all;
    __syncthreads();
    for (int s = FBME_BLOCK / 2; s > 0; s >>= 1) {
      if (t < s) red[t] += red[t + s];
      __syncthreads();
    }
    const double all_sum = red[0];
    __syncthreads();
    if (!(off_sum > 1e-30 * all_sum)) break;   // also ends on NaN
    for (int step = 0; step < n - 1; ++step) {
      if (t < np) {
        // round-robin: position 0 fixed, the others rotate by `step`
        const int a0 = t, a1 = n - 1 - t;
        const int p = a0 == 0 ? 0 : ((a0 - 1 + step) % (n - 1)) + 1;
        const int q = ((a1 - 1 + step) % (n - 1)) + 1;
        pp[t] = p;
        pq[t] = q;
        const double2 h = wa[p][q];
        const double d = sqrt(h.x * h.x + h.y * h.y);
        if (d > 1e-300) {
          const double zeta = (wa[q][q].x - wa[p][p].x) / (2.0 * d);
          const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + tt * tt);
          prm[t][0] = c;
          prm[t][1] = tt * c;
          pe[t] = make_double2(h.x / d, h.y / d);   // e^{i phi}
          prm[t][2] = 1.0;
        } else {
          prm[t][2] = 0.0;
        }
      }
      __syncthreads();
      // columns of H and V: col_p <- c col_p - s e* col_q, col_q <- s col_p + c e* col_q
      for (int e = t; e < 2 * np * n; e += FBME_BLOCK) {
        const int which = e / (np * n), rem = e - which * np * n;
        const int k = rem / n, i = rem % n;
        if (prm[k][2] == 0.0) continue;
        double2 (*mat)[NL] = which ? aa : wa;
        const double c = prm[k][0], s = prm[k][1];
        const double2 ec = zconj(pe[k]);
        const int p = pp[k], q = pq[k];
        const double2 xp = mat[i][p], xq = mat[i][q];
        const double2 eq = zmul(ec, xq);
        mat[i][p] = make_double2(c * xp.x - s * eq.x, c * xp.y - s * eq.y);
        mat[i][q] = make_double2(s * xp.x + c * eq.x, s * xp.y + c * eq.y);
      }
      __syncthreads();
      // rows of H: row_p <- c row_p - s e row_q, row_q <- s row_p + c e row_q
      for (int e = t; e < np * n; e += FBME_BLOCK) {
        const int k = e / n, j = e % n;
        if (prm[k][2] == 0.0) continue;
        const double c = prm[k][0], s = prm[k][1];
        const int p = pp[k], q = pq[k];
        const double2 xp = wa[p][j], xq = wa[q][j];
        const double2 eq = zmul(pe[k], xq);
        wa[p][j] = make_double2(c * xp.x - s * eq.x, c * xp.y - s * eq.y);
        wa[q][j] = make_double2(s * xp.x + c * eq.x, s * xp.y + c * eq.y);
      }
      __syncthreads();
      if (t < np && prm[t][2] != 0.0) {
        wa[pp[t]][pq[t]] = make_double2(0.0, 0.0);
        wa[pq[t]][pp[t]] = make_double2(0.0, 0.0);
        wa[pp[t]][pp[t]].y = 0.0;
        wa[pq[t]][pq[t]].y = 0.0;
      }
      __syncthreads();
    }
  }
  // eigenvalues sorted descending (ties to the lower index)
  if (t < r) lam[t] = wa[t][t].x;
  __syncthreads();
  if (t < r) {
    int k = 0;
    for (int j = 0; j < r; ++j) k += (lam[j] > lam[t]) || (lam[j] == lam[t] && j < t);
    rk[t] = k;
  }
  __syncthreads();
  // background: noise + the trace per sample the approximation misses
  __shared__ double bgs, lmin;
  if (t == 0) {
    double s = 0.0, mn = INFINITY;
    for (int i = 0; i < r; ++i) {
      s += lam[i];
      mn = fmin(mn, lam[i]);
    }
    const double resid = fmax(gamma_sum[b] - s / (double)m, 0.0);
    bgs = fmax((double)noise[b] + resid, TINY32);
    lmin = mn;
    if (!ok) bgs = fmax((double)noise[b] + gamma_sum[b], TINY32);
  }
  __syncthreads();
  // Z[:, rank(i)] = Kinv^H V[:, i] / sqrt(lambda_i); S^-1 = diag(1 - w)
  float2* zb = zout + (long long)b * NR * NR;
  float2* cb = cinv + (long long)b * NR * NR;
  for (int e = t; e < NR * NR; e += FBME_BLOCK) {
    zb[e] = make_float2(0.f, 0.f);
    cb[e] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  if (ok) {
    for (int e = t; e < r * r; e += FBME_BLOCK) {
      const int j = e / r, i = e % r;
      const double li = lam[i];
      double2 s = make_double2(0.0, 0.0);
      if (li > TINY32) {
        for (int p = 0; p < r; ++p) {
          const double2 u = zmul(zconj(ga[p][j]), aa[p][i]);
          s.x += u.x;
          s.y += u.y;
        }
        const double sc = 1.0 / sqrt(li);
        s = make_double2(s.x * sc, s.y * sc);
      }
      zb[j * NR + rk[i]] = make_float2((float)s.x, (float)s.y);
    }
    if (t < r) {
      const double w = (lmin + bgs) / (bgs + lam[t]);
      cb[rk[t] * NR + rk[t]] = make_float2((float)(1.0 - w), 0.f);
    }
  }
  if (t < NR) lam_out[(long long)b * NR + (t < r ? rk[t] : t)] = (t < r && ok) ? (float)lam[t] : 0.f;
  if (t == 0) bg_out[b] = (float)bgs;
}

// U [b, 32, M] = Y Z: u[k, i] = sum_j y[j, i] z[j, k] (FP64 sums).
__global__ void __launch_bounds__(FBME_BLOCK)
k_nys_combine(const float2* __restrict__ y, const float2* __restrict__ z, long long m, int r, float2* __restrict__ u) {
  __shared__ float2 zs[NR][NR];
  const int b = blockIdx.y;
  for (int e = threadIdx.x; e < NR * NR; e += FBME_BLOCK) zs[e / NR][e % NR] = z[(long long)b * NR * NR + e];
  __syncthreads();
  const long long i = (long long)blockIdx.x * FBME_BLOCK + threadIdx.x;
  if (i >= m) return;
  const float2* yb = y + (long long)b * NR * m + i;
  float2 yv[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) yv[j] = j < r ? yb[(long long)j * m] : make_float2(0.f, 0.f);
  float2* ub = u + (long long)b * NR * m + i;
  for (int k = 0; k < NR; ++k) {
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const float2 a = yv[j], w = zs[j][k];
      sx += (double)a.x * w.x - (double)a.y * w.y;
      sy += (double)a.x * w.y + (double)a.y * w.x;
    }
    ub[(long long)k * m] = make_float2((float)sx, (float)sy);
  }
}

}  // namespace

extern "C" {

size_t fbme_nystrom_part_bytes(int batch) {
  return batch < 1 ? 0 : (size_t)batch * GB * 3 * NR * NR * sizeof(double2);
}

int fbme_nystrom_omega(void* stream, int batch, int n1, int n2, int m1, int m2, int rank, const int* idx,
                       const void* rnd, void* omega) {
  if (batch < 1 || n1 < 1 || n2 < 1 || m1 < 1 || m2 < 1 || rank < 1 || rank > NR || !idx || !omega)
    return FBME_ERR_ARG;
  const int ri = rank / 2 > 1 ? rank / 2 : 1;
  if (rank > ri && !rnd) return FBME_ERR_ARG;
  const long long m = (long long)m1 * m2;
  hipLaunchKernelGGL(k_nys_omega, dim3((unsigned)((m + FBME_BLOCK - 1) / FBME_BLOCK), batch), dim3(FBME_BLOCK), 0,
                     (hipStream_t)stream, idx, (const float2*)rnd, n1, n2, m1, m2, rank, ri, (float2*)omega);
  return (int)hipGetLastError();
}

int fbme_nystrom_setup(void* stream, int batch, long long m, int rank, const void* omega, const void* y,
                       const double* gamma_sum, const float* noise, void* part, void* z, void* cinv, float* lambda,
                       float* bg) {
  if (batch < 1 || m < 1 || rank < 1 || rank > NR || !omega || !y || !gamma_sum || !noise || !part || !z ||
      !cinv || !lambda || !bg)
    return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_nys_grams, dim3(GB, batch), dim3(FBME_BLOCK), 0, st, (const float2*)omega, (const float2*)y,
                     m, (double2*)part);
  hipLaunchKernelGGL(k_nys_small, dim3(batch), dim3(FBME_BLOCK), 0, st, (const double2*)part, rank, m, gamma_sum,
                     noise, (float2*)z, (float2*)cinv, lambda, bg);
  return (int)hipGetLastError();
}

int fbme_nystrom_vectors(void* stream, int batch, long long m, int rank, const void* y, const void* z, void* u) {
  if (batch < 1 || m < 1 || rank < 1 || rank > NR || !y || !z || !u) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_nys_combine, dim3((unsigned)((m + FBME_BLOCK - 1) / FBME_BLOCK), batch), dim3(FBME_BLOCK), 0,
                     (hipStream_t)stream, (const float2*)y, (const float2*)z, m, rank, (float2*)u);
  return (int)hipGetLastError();
}

}  // extern "C"
