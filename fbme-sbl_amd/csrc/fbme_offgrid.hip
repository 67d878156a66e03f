// Off-grid refinement of on-grid SBL peaks (gfx950): refine_offgrid_fourier
// of torch_sbl/offgrid.py:70-226 for 2-D grids (1-D grids run as 1 x N), one
// workgroup per scene with the whole Gauss-Newton loop inside the kernel.
//
// Nothing of size M x K is formed.  For atoms a_k[m] = exp(-2 pi i (m1 p1/N1
// + m2 p2/N2)) at continuous positions p_k:
//   * D^H y and the weighted sums sum_m (m_d / N_d) conj(a_k) y are separable
//     non-uniform DFTs of the data (one pass over y per evaluation);
//   * every inner product of atoms and their position derivatives is a
//     product of 1-D moment sums E_p(delta) = sum_{j<M} (j/N)^p e^{2 pi i j
//     delta / N}, p = 0, 1, 2, over the position differences (closed-form
//     Gram, D^H J and J^H J of the variable-projection Jacobian);
//   * ||y - D a||^2 = ||y||^2 - 2 Re(a^H D^H y) + a^H G a.
// Linear algebra (K <= 64 atoms, 2K <= 128 position coordinates) runs in
// FP64: in LDS for K <= 32, in a caller's global workspace (388 KiB per
// scene) for 32 < K <= 64; both run the same code in the same order, so a
// scene refined either way gives the same bits.  Candidate peaks come from a block radix select of the top
// max(64, 16 Kmax) magnitudes, sorted, then greedy periodic non-maximum
// suppression (offgrid.py:229-295) and the extended BIC model order
// (offgrid.py:318-361, orthogonalisation by the Cholesky factor of the Gram,
// which spans the same projections as the reference's QR).
#include <algorithm>

#include "fbme_device.h"

using namespace fbme;

namespace {

constexpr int OG_THREADS = 256;
constexpr int OG_KLDS = 32;       // refined atoms whose matrices sit in LDS
constexpr int OG_KMAX = 64;       // refined atoms with a global workspace
constexpr int OG_BICMAX = 64;     // candidates entering the model-order choice
constexpr int OG_CAND = 1024;     // sorted candidates
constexpr double OG_PI = 3.14159265358979323846;

__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zcmul(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cis(double turns) {   // exp(2 pi i turns)
  double s, c;
  sincospi(2.0 * turns, &s, &c);
  return make_double2(c, s);
}

struct OgArgs {
  int n1, n2, m1, m2;
  int component_count;   // <= 0: model order by BIC
  int max_components;
  float relative_threshold, minimum_separation;
  int max_iter;
  int dims;              // 1: a 1-D grid run as 1 x N (N1 = M1 = 1), 2: a 2-D grid
  double tolerance, max_step, damping, ridge, eps, tiny;
};

// LDS of one scene.  The candidate sort, the BIC system and the refinement
// matrices are alive in different phases and share storage.
struct OgLds {
  union {
    struct {
      unsigned long long keys[OG_CAND];
      unsigned hist[256];
    } sel;
    struct {
      double2 g[OG_BICMAX][OG_BICMAX + 1];
    } bic;
    struct {
      double2 g[OG_KLDS][OG_KLDS + 1];       // Gram of the atoms
      double2 l[OG_KLDS][OG_KLDS + 1];       // Cholesky factor of the loaded Gram
      double2 h[OG_KLDS][2 * OG_KLDS + 1];   // D^H J -> L^-1 D^H J -> loaded-Gram^-1 D^H J
      double nm[2 * OG_KLDS][2 * OG_KLDS + 1];   // normal matrix / its factor
    } gn;
  } u;
  double2 s0[OG_BICMAX], s1[OG_KMAX], s2[OG_KMAX];   // D^H y and the weighted sums
  double2 amp[OG_KMAX], tamp[OG_KMAX], z[OG_BICMAX];
  double2 part[OG_THREADS][3];
  double pos[OG_BICMAX][2], tpos[OG_KMAX][2], init_pos[OG_BICMAX][2];
  double step[2 * OG_KMAX], rhs[2 * OG_KMAX], diag[2 * OG_KMAX];
  double red[OG_THREADS];
  double yy;                 // ||y||^2
  double obj, tobj, lambda, reg;
  int sel_idx[OG_BICMAX];
  int nsel, k, flag, count;
  unsigned thresh;
  int need, taken;
};

// The refinement matrices of K <= KM atoms (layout of OgLds::u.gn): in LDS
// for KM = 32, in the scene's slice of the workspace for KM = 64.
template <int KM>
struct GnMats {
  double2 (*g)[KM + 1];
  double2 (*l)[KM + 1];
  double2 (*h)[2 * KM + 1];
  double (*nm)[2 * KM + 1];
};

template <int KM>
constexpr size_t gn_bytes() {
  return sizeof(double2) * (2 * KM * (KM + 1) + KM * (2 * KM + 1)) + sizeof(double) * 2 * KM * (2 * KM + 1);
}
// workspace bytes per scene (a multiple of 256)
constexpr size_t OG_WS_BYTES = (gn_bytes<OG_KMAX>() + 255) / 256 * 256;

// ------------------------------------------------------------- block helpers
__device__ double block_sum_d(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = OG_THREADS / 2; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// In-place Cholesky of the Hermitian K x K matrix a (lower factor); returns
// false if a pivot is not positive.
template <int LD>
__device__ bool chol_z(double2 (*a)[LD], int k, int* flag) {
  const int t = threadIdx.x;
  if (t == 0) *flag = 0;
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (t == 0) {
      const double d = a[j][j].x;
      if (!(d > 0.0) || !isfinite(d)) *flag = 1;
      a[j][j] = make_double2(sqrt(fmax(d, 0.0)), 0.0);
    }
    __syncthreads();
    const double piv = a[j][j].x;
    for (int i = j + 1 + t; i < k; i += OG_THREADS) a[i][j] = zscale(a[i][j], piv > 0.0 ? 1.0 / piv : 0.0);
    __syncthreads();
    const int rem = k - j - 1;
    for (int e = t; e < rem * rem; e += OG_THREADS) {
      const int i = j + 1 + e / rem, q = j + 1 + e % rem;
      if (q <= i) a[i][q] = zsub(a[i][q], zmul(a[i][j], zconj(a[q][j])));
    }
    __syncthreads();
  }
  return *flag == 0;
}

// Real symmetric version.
template <int LD>
__device__ bool chol_r(double (*a)[LD], int k, int* flag) {
  const int t = threadIdx.x;
  if (t == 0) *flag = 0;
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (t == 0) {
      const double d = a[j][j];
      if (!(d > 0.0) || !isfinite(d)) *flag = 1;
      a[j][j] = sqrt(fmax(d, 0.0));
    }
    __syncthreads();
    const double piv = a[j][j];
    for (int i = j + 1 + t; i < k; i += OG_THREADS) a[i][j] = piv > 0.0 ? a[i][j] / piv : 0.0;
    __syncthreads();
    const int rem = k - j - 1;
    for (int e = t; e < rem * rem; e += OG_THREADS) {
      const int i = j + 1 + e / rem, q = j + 1 + e % rem;
      if (q <= i) a[i][q] -= a[i][j] * a[q][j];
    }
    __syncthreads();
  }
  return *flag == 0;
}

// x <- L^-1 x for one complex vector (thread 0; K <= 64).
template <int LD>
__device__ void fwd_z(const double2 (*l)[LD], double2* x, int k) {
  for (int i = 0; i < k; ++i) {
    double2 s = x[i];
    for (int j = 0; j < i; ++j) s = zsub(s, zmul(l[i][j], x[j]));
    x[i] = zscale(s, 1.0 / l[i][i].x);
  }
}
// x <- L^-H x (thread 0).
template <int LD>
__device__ void bwd_z(const double2 (*l)[LD], double2* x, int k) {
  for (int i = k - 1; i >= 0; --i) {
    double2 s = x[i];
    for (int j = i + 1; j < k; ++j) s = zsub(s, zcmul(l[j][i], x[j]));
    x[i] = zscale(s, 1.0 / l[i][i].x);
  }
}

// ---------------------------------------------------- data sums and moments
// s0[k] = sum_m conj(a_k[m]) y[m], s1[k] / s2[k] the same weighted by
// m1 / N1 and m2 / N2 (only with `weighted`).  Thread (k, part) sums rows
// m1 = part, part + P, ...; partials reduced in a fixed order.
__device__ __forceinline__ double2 load_c(const float2* p, long long i) {
  const float2 v = p[i];
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ double2 load_c(const double2* p, long long i) { return p[i]; }

template <typename C>
__device__ void data_sums(const C* __restrict__ y, const double (*pos)[2], int k0, int k1, const OgArgs& a,
                          bool weighted, OgLds& L) {
  const int t = threadIdx.x;
  const int nk = k1 - k0;
  const int parts = min(32, OG_THREADS / max(nk, 1));
  const int k = k0 + t / parts, part = t % parts;
  double2 a0 = make_double2(0.0, 0.0), a1 = a0, a2 = a0;
  if (t < nk * parts) {
    const double p1 = pos[k][0], p2 = pos[k][1];
    const double2 w2 = cis(p2 / a.n2);
    for (int r = part; r < a.m1; r += parts) {
      double2 e = make_double2(1.0, 0.0), tr = make_double2(0.0, 0.0), ur = tr;
      const C* row = y + (long long)r * a.m2;
      for (int c = 0; c < a.m2; ++c) {
        const double2 term = zmul(e, load_c(row, c));
        tr = zadd(tr, term);
        if (weighted) ur = zadd(ur, zscale(term, (double)c));
        e = zmul(e, w2);
      }
      const double2 e1 = cis(fmod((double)r * p1, (double)a.n1) / a.n1);
      const double2 v0 = zmul(e1, tr);
      a0 = zadd(a0, v0);
      if (weighted) {
        a1 = zadd(a1, zscale(v0, (double)r / a.n1));
        a2 = zadd(a2, zscale(zmul(e1, ur), 1.0 / a.n2));
      }
    }
  }
  L.part[t][0] = a0;
  L.part[t][1] = a1;
  L.part[t][2] = a2;
  __syncthreads();
  if (t < nk) {
    double2 s0 = make_double2(0.0, 0.0), s1 = s0, s2 = s0;
    for (int q = 0; q < parts; ++q) {
      s0 = zadd(s0, L.part[t * parts + q][0]);
      s1 = zadd(s1, L.part[t * parts + q][1]);
      s2 = zadd(s2, L.part[t * parts + q][2]);
    }
    L.s0[k0 + t] = s0;
    if (weighted) {
      L.s1[k0 + t] = s1;
      L.s2[k0 + t] = s2;
    }
  }
  __syncthreads();
}

// 1-D moments E_p = sum_{j<M} (j/N)^p exp(2 pi i j delta / N), p = 0..P.
template <int P>
__device__ void moments(double delta, int m, int n, double2* e) {
  const double2 w = cis(delta / n);
  double2 z = make_double2(1.0, 0.0);
#pragma unroll
  for (int p = 0; p <= P; ++p) e[p] = make_double2(0.0, 0.0);
  for (int j = 0; j < m; ++j) {
    e[0] = zadd(e[0], z);
    if (P >= 1) e[1] = zadd(e[1], zscale(z, (double)j));
    if (P >= 2) e[2] = zadd(e[2], zscale(z, (double)j * j));
    z = zmul(z, w);
  }
  if (P >= 1) e[1] = zscale(e[1], 1.0 / n);
  if (P >= 2) e[2] = zscale(e[2], 1.0 / ((double)n * n));
}

// Gram g[k][l] = sum_m conj(a_k) a_l of atoms at pos (K x K, both halves).
template <int LD>
__device__ void gram(const double (*pos)[2], int k, const OgArgs& a, double2 (*g)[LD]) {
  const int t = threadIdx.x;
  for (int e = t; e < k * k; e += OG_THREADS) {
    const int r = e / k, q = e % k;
    if (q < r) continue;
    double2 e1[1], e2[1];
    moments<0>(pos[r][0] - pos[q][0], a.m1, a.n1, e1);
    moments<0>(pos[r][1] - pos[q][1], a.m2, a.n2, e2);
    const double2 v = zmul(e1[0], e2[0]);
    g[r][q] = v;
    g[q][r] = zconj(v);
  }
  __syncthreads();
}

// Amplitudes of the ridge-loaded least squares at L.tpos (offgrid.py:479-499)
// and the objective ||y - D a||^2: needs L.s0 at the same positions.
// Leaves the Gram in M.g and the loaded factor in M.l.
template <int KM>
__device__ bool fit(const double (*pos)[2], int k, const OgArgs& a, OgLds& L, const GnMats<KM>& M, double2* amp,
                    double* obj) {
  const int t = threadIdx.x;
  gram<KM + 1>(pos, k, a, M.g);
  if (t == 0) {
    double md = 0.0;
    for (int i = 0; i < k; ++i) md += M.g[i][i].x;
    L.reg = fmax(a.ridge, a.eps) * fmax(md / k, a.tiny);
  }
  __syncthreads();
  for (int e = t; e < k * k; e += OG_THREADS) {
    const int r = e / k, q = e % k;
    double2 v = M.g[r][q];
    if (r == q) v.x += L.reg;
    M.l[r][q] = v;
  }
  __syncthreads();
  const bool ok = chol_z<KM + 1>(M.l, k, &L.flag);
  if (t == 0) {
    for (int i = 0; i < k; ++i) amp[i] = L.s0[i];
    fwd_z<KM + 1>(M.l, amp, k);
    bwd_z<KM + 1>(M.l, amp, k);
    // ||y||^2 - 2 Re(a^H D^H y) + a^H G a
    double cross = 0.0, quad = 0.0;
    for (int i = 0; i < k; ++i) {
      cross += zcmul(amp[i], L.s0[i]).x;
      double2 ga = make_double2(0.0, 0.0);
      for (int j = 0; j < k; ++j) ga = zadd(ga, zmul(M.g[i][j], amp[j]));
      quad += zcmul(amp[i], ga).x;
    }
    *obj = fmax(L.yy - 2.0 * cross + quad, 0.0);
  }
  __syncthreads();
  return ok;
}

// Normal matrix Re(Jp^H Jp) and rhs Re(Jp^H r) of the projected Jacobian at
// L.pos with amplitudes L.amp (offgrid.py:502-535, 413-419).  Needs L.s0 /
// s1 / s2, the Gram and the loaded factor of fit() at L.pos.
template <int KM>
__device__ void normal_eq(int k, const OgArgs& a, OgLds& L, const GnMats<KM>& M) {
  const int t = threadIdx.x;
  const int n2k = 2 * k;
  const double tp = 2.0 * OG_PI;
  // D^H J into h, Re(J^H J) into nm
  for (int e = t; e < k * k; e += OG_THREADS) {
    const int r = e / k, q = e % k;
    double2 e1[3], e2[3];
    moments<2>(L.pos[r][0] - L.pos[q][0], a.m1, a.n1, e1);
    moments<2>(L.pos[r][1] - L.pos[q][1], a.m2, a.n2, e2);
    // J_{q,d}[m] = -2 pi i (m_d / N_d) a_q[m] amp_q
    const double2 mi = make_double2(0.0, -tp);
    const double2 aq = L.amp[q];
    M.h[r][2 * q] = zmul(zmul(mi, aq), zmul(e1[1], e2[0]));
    M.h[r][2 * q + 1] = zmul(zmul(mi, aq), zmul(e1[0], e2[1]));
    const double2 ca = zscale(zcmul(L.amp[r], aq), tp * tp);
    M.nm[2 * r][2 * q] = zmul(ca, zmul(e1[2], e2[0])).x;
    M.nm[2 * r][2 * q + 1] = zmul(ca, zmul(e1[1], e2[1])).x;
    M.nm[2 * r + 1][2 * q] = zmul(ca, zmul(e1[1], e2[1])).x;
    M.nm[2 * r + 1][2 * q + 1] = zmul(ca, zmul(e1[0], e2[2])).x;
  }
  __syncthreads();
  // J^H r = J^H y - H^H amp  (r = y - D amp)
  if (t < n2k) {
    const int q = t >> 1, d = t & 1;
    const double2 sy = d ? L.s2[q] : L.s1[q];
    double2 v = zmul(make_double2(0.0, tp), zcmul(L.amp[q], sy));   // 2 pi i conj(amp_q) S_d
    for (int l = 0; l < k; ++l) v = zsub(v, zcmul(M.h[l][t], L.amp[l]));
    L.rhs[t] = v.x;
  }
  __syncthreads();
  // X = L^-1 H (column per thread)
  if (t < n2k) {
    for (int i = 0; i < k; ++i) {
      double2 s = M.h[i][t];
      for (int j = 0; j < i; ++j) s = zsub(s, zmul(M.l[i][j], M.h[j][t]));
      M.h[i][t] = zscale(s, 1.0 / M.l[i][i].x);
    }
  }
  __syncthreads();
  // nm -= Re(X^H X)
  for (int e = t; e < n2k * n2k; e += OG_THREADS) {
    const int r = e / n2k, q = e % n2k;
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += zcmul(M.h[i][r], M.h[i][q]).x;
    M.nm[r][q] -= acc;
  }
  __syncthreads();
  // C = L^-H X
  if (t < n2k) {
    for (int i = k - 1; i >= 0; --i) {
      double2 s = M.h[i][t];
      for (int j = i + 1; j < k; ++j) s = zsub(s, zcmul(M.l[j][i], M.h[j][t]));
      M.h[i][t] = zscale(s, 1.0 / M.l[i][i].x);
    }
  }
  __syncthreads();
  // nm -= reg Re(C^H C); rhs -= Re(C^H D^H r), D^H r = s0 - G amp
  for (int e = t; e < n2k * n2k; e += OG_THREADS) {
    const int r = e / n2k, q = e % n2k;
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += zcmul(M.h[i][r], M.h[i][q]).x;
    M.nm[r][q] -= L.reg * acc;
  }
  if (t < k) {
    double2 v = L.s0[t];
    for (int j = 0; j < k; ++j) v = zsub(v, zmul(M.g[t][j], L.amp[j]));
    L.z[t] = v;
  }
  __syncthreads();
  if (t < n2k) {
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += zcmul(M.h[i][t], L.z[i]).x;
    L.rhs[t] -= acc;
  }
  __syncthreads();
}

// Solve (nm + lambda diag(max(diag nm, eps))) step = rhs; nm is overwritten
// by the factor.  Returns false on a non-positive pivot.
template <int KM>
__device__ bool solve_step(int n, const OgArgs& a, OgLds& L, const GnMats<KM>& M) {
  const int t = threadIdx.x;
  if (t < n) L.diag[t] = fmax(M.nm[t][t], a.eps);
  __syncthreads();
  if (t < n) M.nm[t][t] += L.lambda * L.diag[t];
  __syncthreads();
  const bool ok = chol_r<2 * KM + 1>(M.nm, n, &L.flag);
  if (t == 0) {
    for (int i = 0; i < n; ++i) {
      double s = L.rhs[i];
      for (int j = 0; j < i; ++j) s -= M.nm[i][j] * L.step[j];
      L.step[i] = s / M.nm[i][i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = L.step[i];
      for (int j = i + 1; j < n; ++j) s -= M.nm[j][i] * L.step[j];
      L.step[i] = s / M.nm[i][i];
    }
  }
  __syncthreads();
  return ok;
}

__device__ __forceinline__ unsigned mag_key(float2 v) {
  return __float_as_uint(sqrtf(v.x * v.x + v.y * v.y));
}
__device__ __forceinline__ unsigned mag_key(double2 v) {
  return __float_as_uint((float)sqrt(v.x * v.x + v.y * v.y));
}

// Projected Gauss--Newton from the selected positions L.pos (k atoms, the
// same in every thread), the position standard deviations and the outputs
// (offgrid.py:364-476, 538-564), with the K x K and 2K x 2K matrices in M.
template <int KM, typename C>
__device__ void refine(const C* __restrict__ y, const OgArgs& a, OgLds& L, const GnMats<KM>& M, int k, int b,
                       long long m, int kmax_out, double* __restrict__ out_init, double* __restrict__ out_pos,
                       double2* __restrict__ out_amp,
                       double* __restrict__ out_std, double* __restrict__ out_norm, int* __restrict__ out_count) {
  const int t = threadIdx.x;
  // ----------------------------------------------- projected Gauss--Newton
  data_sums(y, L.pos, 0, k, a, true, L);
  // The ridge loading keeps the Gram positive definite.  A failed pivot of
  // the normal equations (no finite step exists) or a NaN step ends the
  // iteration, as the non-finite test after the clamp does in the reference
  // (offgrid.py:407-409).
  fit<KM>(L.pos, k, a, L, M, L.amp, &L.obj);
  if (t == 0) {
    L.lambda = a.damping;
    out_norm[2 * b] = sqrt(L.obj);
  }
  __syncthreads();
  int iteration = 0, converged = 0;
  const int n2k = 2 * k;
  for (int it = 1; it <= a.max_iter; ++it) {
    iteration = it;
    normal_eq<KM>(k, a, L, M);
    const bool solved = solve_step<KM>(n2k, a, L, M);
    if (t == 0) {
      int finite = solved ? 1 : 0;
      for (int i = 0; i < n2k; ++i) {
        // torch.clamp semantics: NaN stays NaN (and ends the iteration),
        // +-Inf clamps to +-max_step (fmin/fmax alone would drop the NaN).
        const double v = L.step[i];
        L.step[i] = isnan(v) ? v : fmin(fmax(v, -a.max_step), a.max_step);
        finite &= isfinite(L.step[i]) ? 1 : 0;
      }
      L.flag = finite;
    }
    __syncthreads();
    if (!L.flag) break;
    const double previous = L.obj;
    double scale_taken = 0.0;
    bool accepted = false;
    for (int ls = 0; ls < 5 && !accepted; ++ls) {
      const double scale = 1.0 / (double)(1 << ls);
      if (t < k) {
        for (int d = 0; d < 2; ++d) {
          const double nd = d ? a.n2 : a.n1;
          double v = fmod(L.pos[t][d] + scale * L.step[2 * t + d], nd);
          if (v < 0.0) v += nd;
          L.tpos[t][d] = v;
        }
      }
      __syncthreads();
      data_sums(y, L.tpos, 0, k, a, false, L);
      fit<KM>(L.tpos, k, a, L, M, L.tamp, &L.tobj);
      // one decision for the whole workgroup (L.obj changes below)
      if (t == 0) L.flag = L.tobj < L.obj ? 1 : 0;
      __syncthreads();
      const bool better = L.flag != 0;
      __syncthreads();
      if (better) {
        accepted = true;
        scale_taken = scale;
        if (t < k) {
          L.pos[t][0] = L.tpos[t][0];
          L.pos[t][1] = L.tpos[t][1];
          L.amp[t] = L.tamp[t];
        }
        if (t == 0) {
          L.obj = L.tobj;
          L.lambda = fmax(L.lambda * 0.5, 1e-8);
        }
      }
      __syncthreads();
    }
    if (!accepted) {
      if (t == 0) L.lambda = fmin(L.lambda * 10.0, 1e8);
      __syncthreads();
      // restore the factor of the accepted positions for the next Jacobian
      data_sums(y, L.pos, 0, k, a, true, L);
      fit<KM>(L.pos, k, a, L, M, L.amp, &L.obj);
      if (L.lambda >= 1e8) {
        converged = 1;
        break;
      }
      continue;
    }
    // weighted sums and factor at the accepted positions
    data_sums(y, L.pos, 0, k, a, true, L);
    fit<KM>(L.pos, k, a, L, M, L.amp, &L.obj);
    double maxstep = 0.0;
    for (int i = 0; i < n2k; ++i) maxstep = fmax(maxstep, fabs(scale_taken * L.step[i]));
    const double rel = (previous - L.obj) / fmax(previous, a.eps);
    if (maxstep <= a.tolerance || rel <= a.tolerance * a.tolerance) {
      converged = 1;
      break;
    }
  }
  // ------------------------------------------- position standard deviation
  normal_eq<KM>(k, a, L, M);
  if (a.dims == 1) {
    // 1-D grids: the axis-0 coordinates (N1 = M1 = 1) have zero Jacobian
    // columns; the Fisher matrix of the axis-1 coordinates is the block the
    // identity on axis 0 leaves (its std is not reported)
    for (int e = t; e < n2k * n2k; e += OG_THREADS) {
      const int r = e / n2k, q = e % n2k;
      if (!(r & 1) || !(q & 1)) M.nm[r][q] = (r == q) ? 1.0 : 0.0;
    }
    __syncthreads();
  }
  // cov = 0.5 noise Fisher^-1, noise = ||r||^2 / max(M - K, 1)
  const double half_noise = 0.5 * L.obj / (double)max((int)m - k, 1);
  const bool fisher_ok = chol_r<2 * KM + 1>(M.nm, n2k, &L.flag);
  // diag of the inverse: || L^-1 e_i ||^2, one column per thread, the
  // columns kept in the storage of h (free here; 2K x 2K doubles fit in it)
  if (t < n2k) {
    double* v = &M.h[0][0].x;
    for (int i = 0; i < n2k; ++i) {
      double s = (i == t) ? 1.0 : 0.0;
      for (int j = t; j < i; ++j) s -= M.nm[i][j] * v[j * n2k + t];
      v[i * n2k + t] = i < t ? 0.0 : s / M.nm[i][i];
    }
    double acc = 0.0;
    for (int i = t; i < n2k; ++i) acc += v[i * n2k + t] * v[i * n2k + t];
    L.diag[t] = fisher_ok ? sqrt(fmax(half_noise * acc, 0.0)) : NAN;
  }
  __syncthreads();
  if (t < kmax_out) {
    const long long o = (long long)b * kmax_out + t;
    const bool live = t < k;
    out_init[2 * o] = live ? L.init_pos[t][0] : NAN;
    out_init[2 * o + 1] = live ? L.init_pos[t][1] : NAN;
    out_pos[2 * o] = live ? L.pos[t][0] : NAN;
    out_pos[2 * o + 1] = live ? L.pos[t][1] : NAN;
    out_amp[o] = live ? L.amp[t] : make_double2(NAN, NAN);
    out_std[2 * o] = live ? L.diag[2 * t] : NAN;
    out_std[2 * o + 1] = live ? L.diag[2 * t + 1] : NAN;
  }
  if (t == 0) {
    out_norm[2 * b + 1] = sqrt(L.obj);
    out_count[3 * b] = k;
    out_count[3 * b + 1] = iteration;
    out_count[3 * b + 2] = converged;
  }
}

// C: float2 (complex64 data) or double2 (complex128 data).
template <typename C>
__global__ void __launch_bounds__(OG_THREADS)
k_offgrid(const C* __restrict__ data, const C* __restrict__ coef, OgArgs a,
          double* __restrict__ out_init, double* __restrict__ out_pos, double2* __restrict__ out_amp,
          double* __restrict__ out_std, double* __restrict__ out_norm, int* __restrict__ out_count,
          char* __restrict__ ws) {
  __shared__ OgLds L;
  const int b = blockIdx.x, t = threadIdx.x;
  const long long n = (long long)a.n1 * a.n2, m = (long long)a.m1 * a.m2;
  const C* y = data + b * m;
  const C* x = coef + b * n;
  const int kmax_out = a.component_count > 0 ? a.component_count : a.max_components;
  // ---------------------------------------------- top-L magnitudes (radix select)
  const int lsel = (int)min<long long>(n, max(64, 16 * a.max_components));
  {
    unsigned prefix = 0u, pmask = 0u;
    int remaining = lsel;
    for (int shift = 24; shift >= 0; shift -= 8) {
      for (int i = t; i < 256; i += OG_THREADS) L.u.sel.hist[i] = 0u;
      __syncthreads();
      for (long long i = t; i < n; i += OG_THREADS) {
        const unsigned key = mag_key(x[i]);
        if ((key & pmask) == prefix) atomicAdd(&L.u.sel.hist[(key >> shift) & 255u], 1u);
      }
      __syncthreads();
      if (t == 0) {
        int acc = 0, digit = 0;
        for (int d = 255; d >= 0; --d) {
          const int c = (int)L.u.sel.hist[d];
          if (acc + c >= remaining) {
            digit = d;
            break;
          }
          acc += c;
        }
        L.need = remaining - acc;
        L.thresh = (unsigned)digit;
      }
      __syncthreads();
      remaining = L.need;
      prefix |= L.thresh << shift;
      pmask |= 255u << shift;
      __syncthreads();
    }
    // collect keys above the threshold, then the lowest-index ties
    if (t == 0) {
      L.count = 0;
      L.taken = 0;
    }
    __syncthreads();
    const unsigned thr = prefix;
    for (long long base = 0; base < n; base += OG_THREADS) {
      const long long i = base + t;
      const unsigned key = i < n ? mag_key(x[i]) : 0u;
      const bool gt = i < n && key > thr;
      const bool eq = i < n && key == thr;
      if (gt) {
        const int slot = atomicAdd(&L.count, 1);
        L.u.sel.keys[slot] = ((unsigned long long)key << 32) | (0xFFFFFFFFu - (unsigned)i);
      }
      // ties in index order: rank within this chunk by a scan of the flags
      L.red[t] = eq ? 1.0 : 0.0;
      __syncthreads();
      if (t == 0) {
        double run = 0.0;
        for (int q = 0; q < OG_THREADS; ++q) {
          const double f = L.red[q];
          L.red[q] = run;
          run += f;
        }
      }
      __syncthreads();
      const int rank = L.taken + (int)L.red[t];
      if (eq && rank < remaining) {
        const int slot = atomicAdd(&L.count, 1);
        L.u.sel.keys[slot] = ((unsigned long long)key << 32) | (0xFFFFFFFFu - (unsigned)i);
      }
      __syncthreads();
      if (t == OG_THREADS - 1) L.taken = rank + (eq ? 1 : 0);
      __syncthreads();
    }
    for (int i = lsel + t; i < OG_CAND; i += OG_THREADS) L.u.sel.keys[i] = 0ull;
    __syncthreads();
    // bitonic sort, descending
    for (int size = 2; size <= OG_CAND; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int i = t; i < OG_CAND; i += OG_THREADS) {
          const int j = i ^ stride;
          if (j > i) {
            const bool desc = (i & size) == 0;
            const unsigned long long u = L.u.sel.keys[i], v = L.u.sel.keys[j];
            if (desc ? (u < v) : (u > v)) {
              L.u.sel.keys[i] = v;
              L.u.sel.keys[j] = u;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  // ------------------------------------------------ non-maximum suppression
  if (t == 0) {
    const float v0 = __uint_as_float((unsigned)(L.u.sel.keys[0] >> 32));
    const float cutoff = a.relative_threshold * v0;
    const int target = a.component_count > 0 ? a.component_count : a.max_components;
    int ns = 0;
    for (int off = 0; off < lsel && ns < target; ++off) {
      const unsigned long long kv = L.u.sel.keys[off];
      const float v = __uint_as_float((unsigned)(kv >> 32));
      if (a.component_count <= 0 && ns > 0 && v < cutoff) break;
      const int idx = (int)(0xFFFFFFFFu - (unsigned)(kv & 0xFFFFFFFFull));
      const double c1 = (double)(idx / a.n2), c2 = (double)(idx % a.n2);
      bool close = false;
      for (int q = 0; q < ns && !close; ++q) {
        double d1 = fabs(L.pos[q][0] - c1), d2 = fabs(L.pos[q][1] - c2);
        d1 = fmin(d1, a.n1 - d1);
        d2 = fmin(d2, a.n2 - d2);
        close = (float)sqrt(d1 * d1 + d2 * d2) < a.minimum_separation;
      }
      if (close) continue;
      L.pos[ns][0] = c1;
      L.pos[ns][1] = c2;
      L.sel_idx[ns] = idx;
      ++ns;
    }
    if (a.component_count > 0) {
      for (int off = 0; off < lsel && ns < a.component_count; ++off) {
        const int idx = (int)(0xFFFFFFFFu - (unsigned)(L.u.sel.keys[off] & 0xFFFFFFFFull));
        bool have = false;
        for (int q = 0; q < ns; ++q) have |= L.sel_idx[q] == idx;
        if (have) continue;
        L.pos[ns][0] = (double)(idx / a.n2);
        L.pos[ns][1] = (double)(idx % a.n2);
        L.sel_idx[ns] = idx;
        ++ns;
      }
    }
    if (ns == 0) {
      const int idx = (int)(0xFFFFFFFFu - (unsigned)(L.u.sel.keys[0] & 0xFFFFFFFFull));
      L.pos[0][0] = (double)(idx / a.n2);
      L.pos[0][1] = (double)(idx % a.n2);
      ns = 1;
    }
    L.nsel = ns;
  }
  __syncthreads();
  // ||y||^2
  {
    double acc = 0.0;
    for (long long i = t; i < m; i += OG_THREADS) {
      const double2 v = load_c(y, i);
      acc += v.x * v.x + v.y * v.y;
    }
    const double s = block_sum_d(acc, L.red);
    if (t == 0) L.yy = s;
  }
  __syncthreads();
  // ------------------------------------------------------- BIC model order
  int k = L.nsel;
  if (a.component_count <= 0 && k > 1) {
    for (int k0 = 0; k0 < k; k0 += OG_KLDS) data_sums(y, L.pos, k0, min(k, k0 + OG_KLDS), a, false, L);
    gram<OG_BICMAX + 1>(L.pos, k, a, L.u.bic.g);
    chol_z<OG_BICMAX + 1>(L.u.bic.g, k, &L.flag);
    if (t == 0) {
      for (int i = 0; i < k; ++i) L.z[i] = L.s0[i];
      fwd_z<OG_BICMAX + 1>(L.u.bic.g, L.z, k);
      const double mm = (double)m;
      double cum = 0.0, best = INFINITY;
      int arg = 0;
      for (int i = 0; i < k; ++i) {
        cum += L.z[i].x * L.z[i].x + L.z[i].y * L.z[i].y;
        const double resid = fmax(L.yy - cum, a.eps);
        const double cnt = (double)(i + 1);
        // one complex amplitude and `dims` position coordinates per component
        const double bic = mm * log(resid / mm) + (2.0 + a.dims) * cnt * log(mm) + 2.0 * cnt * log((double)n);
        if (bic < best) {
          best = bic;
          arg = i;
        }
      }
      L.nsel = arg + 1;
    }
    __syncthreads();
    k = L.nsel;
  }
  if (t < k) {
    L.init_pos[t][0] = L.pos[t][0];
    L.init_pos[t][1] = L.pos[t][1];
  }
  __syncthreads();
  if (k <= OG_KLDS) {
    const GnMats<OG_KLDS> M{L.u.gn.g, L.u.gn.l, L.u.gn.h, L.u.gn.nm};
    refine<OG_KLDS>(y, a, L, M, k, b, m, kmax_out, out_init, out_pos, out_amp, out_std, out_norm,
                     out_count);
  } else if (ws != nullptr) {
    char* base = ws + (size_t)b * OG_WS_BYTES;
    GnMats<OG_KMAX> M;
    M.g = reinterpret_cast<double2 (*)[OG_KMAX + 1]>(base);
    M.l = M.g + OG_KMAX;
    M.h = reinterpret_cast<double2 (*)[2 * OG_KMAX + 1]>(M.l + OG_KMAX);
    M.nm = reinterpret_cast<double (*)[2 * OG_KMAX + 1]>(M.h + OG_KMAX);
    refine<OG_KMAX>(y, a, L, M, k, b, m, kmax_out, out_init, out_pos, out_amp, out_std, out_norm,
                     out_count);
  } else if (t == 0) {
    out_count[3 * b] = k;
    out_count[3 * b + 1] = -1;   // more atoms than LDS holds and no workspace
    out_count[3 * b + 2] = 0;
  }
}

}  // namespace

extern "C" {

size_t fbme_offgrid_workspace_bytes(int batch) { return batch < 1 ? 0 : (size_t)batch * OG_WS_BYTES; }

int fbme_offgrid_refine_ws(void* stream, int batch, int n1, int n2, int m1, int m2, const void* data,
                           const void* coefficients, int complex128, int component_count, int max_components,
                           float relative_threshold, float minimum_separation, int max_iter, double tolerance,
                           double max_step, double damping, double ridge, double eps, double tiny,
                           double* initial_positions, double* positions, void* amplitudes, double* position_std,
                           double* norms, int* counts, int dims, void* workspace, size_t workspace_bytes) {
  if (batch < 1 || n1 < 1 || n2 < 1 || m1 < 1 || m2 < 1 || m1 > n1 || m2 > n2) return FBME_ERR_ARG;
  if (dims != 1 && dims != 2) return FBME_ERR_ARG;
  if (dims == 1 && (n1 != 1 || m1 != 1)) return FBME_ERR_ARG;
  if ((long long)m1 * m2 < 2 || (long long)n1 * n2 > (1ll << 28)) return FBME_ERR_ARG;
  if (!data || !coefficients || !initial_positions || !positions || !amplitudes || !position_std || !norms ||
      !counts)
    return FBME_ERR_ARG;
  if (workspace && (workspace_bytes < (size_t)batch * OG_WS_BYTES || (reinterpret_cast<size_t>(workspace) & 15u)))
    return FBME_ERR_ARG;
  if (max_components < 1 || max_components > OG_BICMAX || component_count > (workspace ? OG_KMAX : OG_KLDS))
    return FBME_ERR_ARG;
  if (max_iter < 1 || !(tolerance > 0.0) || !(max_step > 0.0) || !(damping > 0.0) || ridge < 0.0) return FBME_ERR_ARG;
  OgArgs a;
  a.n1 = n1;
  a.n2 = n2;
  a.m1 = m1;
  a.m2 = m2;
  a.component_count = component_count;
  a.max_components = max_components;
  a.relative_threshold = relative_threshold;
  a.minimum_separation = minimum_separation;
  a.max_iter = max_iter;
  a.dims = dims;
  a.tolerance = tolerance;
  a.max_step = max_step;
  a.damping = damping;
  a.ridge = ridge;
  a.eps = eps;
  a.tiny = tiny;
  if (complex128)
    hipLaunchKernelGGL(k_offgrid<double2>, dim3(batch), dim3(OG_THREADS), 0, (hipStream_t)stream,
                       (const double2*)data, (const double2*)coefficients, a, initial_positions, positions,
                       (double2*)amplitudes, position_std, norms, counts, (char*)workspace);
  else
    hipLaunchKernelGGL(k_offgrid<float2>, dim3(batch), dim3(OG_THREADS), 0, (hipStream_t)stream,
                       (const float2*)data, (const float2*)coefficients, a, initial_positions, positions,
                       (double2*)amplitudes, position_std, norms, counts, (char*)workspace);
  return (int)hipGetLastError();
}

int fbme_offgrid_refine_dims(void* stream, int batch, int n1, int n2, int m1, int m2, const void* data,
                             const void* coefficients, int complex128, int component_count, int max_components,
                             float relative_threshold, float minimum_separation, int max_iter, double tolerance,
                             double max_step, double damping, double ridge, double eps, double tiny,
                             double* initial_positions, double* positions, void* amplitudes, double* position_std,
                             double* norms, int* counts, int dims) {
  return fbme_offgrid_refine_ws(stream, batch, n1, n2, m1, m2, data, coefficients, complex128, component_count,
                                max_components, relative_threshold, minimum_separation, max_iter, tolerance,
                                max_step, damping, ridge, eps, tiny, initial_positions, positions, amplitudes,
                                position_std, norms, counts, dims, nullptr, 0);
}

int fbme_offgrid_refine(void* stream, int batch, int n1, int n2, int m1, int m2, const void* data,
                        const void* coefficients, int complex128, int component_count, int max_components,
                        float relative_threshold, float minimum_separation, int max_iter, double tolerance,
                        double max_step, double damping, double ridge, double eps, double tiny,
                        double* initial_positions, double* positions, void* amplitudes, double* position_std,
                        double* norms, int* counts) {
  return fbme_offgrid_refine_dims(stream, batch, n1, n2, m1, m2, data, coefficients, complex128, component_count,
                                  max_components, relative_threshold, minimum_separation, max_iter, tolerance,
                                  max_step, damping, ridge, eps, tiny, initial_positions, positions, amplitudes,
                                  position_std, norms, counts, 2);
}

}  // extern "C"
