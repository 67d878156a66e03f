// Covariance matvec of the general solver on the 256 x 256 grid (gfx950).
//
// Q d = noise d + A Gamma A^H d (the operator torch_sbl/operators.py:328-355
// applies through the BTTB circulant embedding) evaluated through the grid
// instead: three passes on the register-resident 256-point FFT of
// fbme_fft16.h, with the probes batched in the lane dimension like the BTTB
// path.  Per lane (d and out [M1, M2], intermediate [256, M2]):
//   1. column pass of A^H: mid[n1, c] = sum_{m1 < M1} e^{+2 pi i m1 n1 / 256} d[m1, c];
//   2. row pass: the row's unnormalised inverse transform (M2 -> 256, zero
//      padded), times gamma[n1, :], forward transform, first M2 outputs back
//      into mid (in place);
//   3. column pass of A: out[k1, c] = noise d[k1, c] + sum_{n1} e^{-2 pi i n1 k1 / 256} mid[n1, c], k1 < M1.
// HBM traffic per lane: 16 M1 M2 + 32 * 256 M2 bytes (+ gamma rows, shared
// by the lanes of a scene through L2) against 16 M + 32 M1 E2 (+ the
// embedded spectrum) for the BTTB passes on the E1 x E2 = 320 x 320
// embedding, and no per-iteration spectrum to build.
#include "fbme_device.h"
#include "fbme_fft16.h"

using namespace fbme;

namespace {

constexpr int GC = 16;             // columns per column-pass workgroup
constexpr int XS = 16 * 17 + 8;    // LDS slots of one 256-point transpose

__global__ void __launch_bounds__(16 * GC)
k_gop_col_adjoint(const float2* __restrict__ d, float2* __restrict__ mid, int m1, int m2,
                  const float2* __restrict__ twtab) {
  __shared__ float2 xs[16 * 17 * GC];   // column c's element i at xs[i * GC + c]
  __shared__ float2 tw[256];
  const int c = threadIdx.x % GC, t = threadIdx.x / GC;
  const int lane = blockIdx.y;
  const int col = blockIdx.x * GC + c;
  const bool act = col < m2;
  load_tw256(tw, twtab);
  const float2* src = d + (long long)lane * m1 * m2;
  float2 v[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int i = t + 16 * j;
    v[j] = (act && i < m1) ? src[(long long)i * m2 + col] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft256<true, GC>(v, tw, xs + c, t);
  if (act) {
    float2* dst = mid + (long long)lane * 256 * m2 + col;
#pragma unroll
    for (int n = 0; n < 16; ++n) dst[(long long)(t + 16 * n) * m2] = v[n];
  }
}

__global__ void __launch_bounds__(64)
k_gop_row(float2* __restrict__ mid, const float* __restrict__ gamma, int m2, int rhs,
          const float2* __restrict__ twtab) {
  __shared__ float2 xs[4][XS];
  __shared__ float2 tw[256];
  const int g = threadIdx.x >> 4, t = threadIdx.x & 15;
  const int lane = blockIdx.y;
  const int b = lane / rhs;
  const int row = blockIdx.x * 4 + g;
  load_tw256(tw, twtab);
  float2* line = mid + ((long long)lane * 256 + row) * m2;
  const float* gr = gamma + ((long long)b * 256 + row) * 256;
  float2 v[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int c = t + 16 * j;
    v[j] = (c < m2) ? line[c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft256<true>(v, tw, xs[g], t);
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const float gv = gr[t + 16 * n];
    v[n] = make_float2(v[n].x * gv, v[n].y * gv);
  }
  fft256<false>(v, tw, xs[g], t);
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int k = t + 16 * n;
    if (k < m2) line[k] = v[n];
  }
}

// DOT: also part[lane, block] = sum Re(conj(d) out) over the block's
// outputs (FP64, fixed order), the PCG curvature <d, Q d>.
template <bool DOT>
__global__ void __launch_bounds__(16 * GC)
k_gop_col_forward(const float2* __restrict__ mid, const float2* __restrict__ d, const float* __restrict__ noise,
                  float2* __restrict__ out, int m1, int m2, int rhs, const float2* __restrict__ twtab,
                  double* __restrict__ part) {
  __shared__ float2 xs[16 * 17 * GC];
  __shared__ float2 tw[256];
  __shared__ double red[8];
  const int c = threadIdx.x % GC, t = threadIdx.x / GC;
  const int lane = blockIdx.y;
  const int col = blockIdx.x * GC + c;
  const bool act = col < m2;
  load_tw256(tw, twtab);
  const float2* src = mid + (long long)lane * 256 * m2 + col;
  float2 v[16], dl[16];
  const long long base = (long long)lane * m1 * m2 + col;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int k = t + 16 * j;
    v[j] = act ? src[(long long)k * m2] : make_float2(0.f, 0.f);
    dl[j] = (act && k < m1) ? d[base + (long long)k * m2] : make_float2(0.f, 0.f);   // loaded before the transform
  }
  __syncthreads();
  fft256<false, GC>(v, tw, xs + c, t);
  const float nv = noise[lane / rhs];
  double acc = 0.0;
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int k = t + 16 * n;
    if (act && k < m1) {
      const float2 dv = dl[n];
      const float2 o = make_float2(v[n].x + nv * dv.x, v[n].y + nv * dv.y);
      out[base + (long long)k * m2] = o;
      if (DOT) acc += (double)dv.x * o.x + (double)dv.y * o.y;
    }
  }
  if (DOT) {
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[(long long)lane * gridDim.x + blockIdx.x] = acc;
  }
}

// out[l] = sum_i part[l, i], one wavefront per lane (fixed shuffle tree).
__global__ void __launch_bounds__(64) k_gop_finish(const double* __restrict__ part, int blocks, int lanes,
                                                   double* __restrict__ out) {
  const int l = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < blocks; i += 64) s += part[(long long)l * blocks + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[l] = s;
}

// ------------------------------------------ low-rank preconditioner, N1 = 256
// F[m, k] = U[m1, k] V[m2, k] with U[m1, k] = e^{-2 pi i m1 k1 / 256}: the
// projection F^H v and the synthesis F c factor through the 256-point column
// transform.  With the atoms' distinct rows as "slots":
//   F^H v [k] = sum_{m2} conj(V[m2, k]) W[slot(k), m2],
//     W[s, m2] = sum_{m1} e^{+2 pi i m1 row(s) / 256} v[m1, m2]   (column IFFT, slot rows kept);
//   (F c)[m1, m2] = sum_s e^{-2 pi i m1 row(s) / 256} T[s, m2],
//     T[s, m2] = sum_{k in s} V[m2, k] c[k]       (slot rows formed in LDS, then their column FFT).
// map[b] = slot of each atom [32], row of each slot [32], slot count [1],
// slot of each grid row (or -1) [256], atoms ordered by (slot, index) [32],
// first position of each slot in that order [33].
constexpr int RANK = 32;
constexpr int MAP_INTS = 32 + 32 + 1 + 256 + 32 + 33;
constexpr int MAP_PERM = 321, MAP_START = 353;

// One wavefront per scene, lane k = atom k: slots are the distinct rows in
// order of first occurrence among the atoms (ballots instead of a serial
// scan); the atom order within a slot is increasing k.
__global__ void __launch_bounds__(64)
k_lrf_map(const int* __restrict__ idx, int n2, int* __restrict__ map) {
  const int b = blockIdx.x;
  const int k = threadIdx.x;
  int* m = map + (long long)b * MAP_INTS;
  for (int r = k; r < 256; r += 64) m[65 + r] = -1;
  const bool atom = k < RANK;
  const int row = atom ? idx[b * RANK + k] / n2 : -1;
  int first = k;
  for (int j = RANK - 1; j >= 0; --j) {
    const int rj = __shfl(row, j, 64);
    if (atom && rj == row) first = j;
  }
  const bool is_first = atom && first == k;
  const unsigned long long firsts = __ballot(is_first);
  const int slot_of_first = __popcll(firsts & ((1ull << k) - 1ull));
  const int slot = __shfl(slot_of_first, atom ? first : 0, 64);   // slot of the first atom with this row
  const int ns = __popcll(firsts);
  // rank of atom k among the atoms of its slot, and the slot's first position
  int before = 0, within = 0;
  for (int j = 0; j < RANK; ++j) {
    const int sj = __shfl(slot, j, 64);
    if (atom) {
      before += sj < slot;
      within += (sj == slot) && (j < k);
    }
  }
  __syncthreads();   // the row map is cleared before it is filled
  if (atom) {
    m[k] = slot;
    m[MAP_PERM + before + within] = k;
    if (is_first) {
      m[32 + slot] = row;
      m[65 + row] = slot;
      m[MAP_START + slot] = before;
    }
    if (k >= ns) m[32 + k] = -1;
  }
  if (k >= ns && k <= RANK) m[MAP_START + k] = RANK;
  if (k == 0) m[64] = ns;
}

// Projection onto the atoms, one workgroup per (16 columns, lane): the
// column IFFT of v[lane, :, c] gives W[s, c] on the slot rows (kept in LDS),
// then part[lane, block, k] = sum_{c in block} conj(V[c, k]) W[slot(k), c]
// in FP64 over the block's columns in order (the blocks are summed by
// k_lr_corr).
__global__ void __launch_bounds__(16 * GC)
k_lrf_col_project(const float2* __restrict__ v, const float2* __restrict__ tab, const int* __restrict__ map, int m1,
                  int m2, int rhs, const float2* __restrict__ twtab, double2* __restrict__ part) {
  // xs is the transform's transpose buffer; before and after the transform
  // it holds the slot rows ws [RANK][GC] and the group sums red [8][RANK]
  __shared__ float2 xs[16 * 17 * GC];
  __shared__ float2 tw[256];
  __shared__ int rowslot[256];
  float2(*ws)[GC] = reinterpret_cast<float2(*)[GC]>(xs);
  double2(*red)[RANK] = reinterpret_cast<double2(*)[RANK]>(xs + RANK * GC);
  const int c = threadIdx.x % GC, t = threadIdx.x / GC;
  const int lane = blockIdx.y;
  const int b = lane / rhs;
  const int col = blockIdx.x * GC + c;
  const bool act = col < m2;
  const int* mb = map + (long long)b * MAP_INTS;
  const float2* V = tab + ((long long)b * (m1 + m2) + m1) * RANK;
  const float2* src = v + (long long)lane * m1 * m2;
  float2 x[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int i = t + 16 * j;
    x[j] = (act && i < m1) ? src[(long long)i * m2 + col] : make_float2(0.f, 0.f);
  }
  // thread (g, k) reduces atom k over columns 2g, 2g + 1 of the block
  const int rk = threadIdx.x & (RANK - 1), rg = threadIdx.x >> 5;
  const int slot_k = mb[rk];
  float2 vk[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c2 = blockIdx.x * GC + 2 * rg + q;
    vk[q] = (c2 < m2) ? V[(long long)c2 * RANK + rk] : make_float2(0.f, 0.f);
  }
  rowslot[threadIdx.x] = mb[65 + threadIdx.x];
  load_tw256(tw, twtab);
  __syncthreads();
  fft256<true, GC>(x, tw, xs + c, t);
  __syncthreads();   // the transpose buffer is free again
  for (int e = threadIdx.x; e < RANK * GC; e += 16 * GC) ws[e / GC][e % GC] = make_float2(0.f, 0.f);
  __syncthreads();
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int sl = rowslot[t + 16 * n];
    if (sl >= 0) ws[sl][c] = x[n];
  }
  __syncthreads();
  {
    double ax = 0.0, ay = 0.0;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const float2 vv = vk[q], ww = ws[slot_k][2 * rg + q];
      ax += (double)(vv.x * ww.x + vv.y * ww.y);
      ay += (double)(vv.x * ww.y - vv.y * ww.x);
    }
    red[rg][rk] = make_double2(ax, ay);
  }
  __syncthreads();
  if (threadIdx.x < RANK) {
    const int k = threadIdx.x;
    double2 sum = red[0][k];
#pragma unroll
    for (int q = 1; q < 8; ++q) {
      sum.x += red[q][k].x;
      sum.y += red[q][k].y;
    }
    part[((long long)lane * gridDim.x + blockIdx.x) * RANK + k] = sum;
  }
}

// out[lane, m1, c] = v / bg - sum_s e^{-2 pi i m1 row(s) / 256} T[s, c] with
// T[s, c] = sum_{k in s} V[c, k] corr[lane, k] formed in LDS for the
// workgroup's columns (each atom feeds one slot; fixed summation order),
// then the column FFT of the slot rows; part[lane, block] = sum Re(conj(v) out).
// DIR: z is not stored; the pass writes the PCG direction d = z + beta d
// (0 for a lane whose update is not valid), beta = nrho / rho with nrho
// already formed from the projection (k_lr_corr), the lane logic of
// k_pcg_direction_parts.
template <bool DIR>
__global__ void __launch_bounds__(16 * GC)
k_lrf_col_apply(const float2* __restrict__ corr, const float2* __restrict__ tab, const int* __restrict__ map,
                const float2* __restrict__ v, const float* __restrict__ bg, float2* __restrict__ out, int m1,
                int m2, int rhs, const float2* __restrict__ twtab, double* __restrict__ part,
                const double* __restrict__ rho, const double* __restrict__ curv, const double* __restrict__ nrho,
                const int* __restrict__ active) {
  // xs is the transform's transpose buffer; before the transform it holds
  // the phase-table slice vsm [GC][RANK] and the slot rows ts [RANK][GC]
  __shared__ float2 xs[16 * 17 * GC];
  __shared__ float2 tw[256];
  __shared__ double red[8];
  __shared__ int rowslot[256];
  __shared__ int perm[RANK];
  __shared__ int start[RANK + 1];
  __shared__ float2 cs[RANK];
  float2(*vsm)[RANK] = reinterpret_cast<float2(*)[RANK]>(xs);
  float2(*ts)[GC] = reinterpret_cast<float2(*)[GC]>(xs + GC * RANK);
  const int c = threadIdx.x % GC, t = threadIdx.x / GC;
  const int lane = blockIdx.y;
  const int b = lane / rhs;
  const int col = blockIdx.x * GC + c;
  const bool act = col < m2;
  const int* mb = map + (long long)b * MAP_INTS;
  const float2* V = tab + ((long long)b * (m1 + m2) + m1) * RANK;
  rowslot[threadIdx.x] = mb[65 + threadIdx.x];
  if (threadIdx.x < RANK) {
    perm[threadIdx.x] = mb[MAP_PERM + threadIdx.x];
    cs[threadIdx.x] = corr[(long long)lane * RANK + threadIdx.x];
  }
  if (threadIdx.x <= RANK) start[threadIdx.x] = mb[MAP_START + threadIdx.x];
  // the input column (and the direction) are loaded first: their latency
  // overlaps the slot-row build
  const long long base = (long long)lane * m1 * m2 + col;
  float2 vin[16], dl[16];
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int k = t + 16 * n;
    vin[n] = (act && k < m1) ? v[base + (long long)k * m2] : make_float2(0.f, 0.f);
    if (DIR) dl[n] = (act && k < m1) ? out[base + (long long)k * m2] : make_float2(0.f, 0.f);
  }
  for (int e = threadIdx.x; e < GC * RANK; e += 16 * GC) {
    const int cc = e / RANK, k = e - cc * RANK;
    const int c2 = blockIdx.x * GC + cc;
    vsm[cc][k] = (c2 < m2) ? V[(long long)c2 * RANK + k] : make_float2(0.f, 0.f);
  }
  load_tw256(tw, twtab);
  __syncthreads();
  const int ns = mb[64];
  for (int e = threadIdx.x; e < RANK * GC; e += 16 * GC) {
    const int sl = e / GC, cc = e - sl * GC;
    float ax = 0.f, ay = 0.f;
    if (sl < ns) {
      for (int j = start[sl]; j < start[sl + 1]; ++j) {
        const int k = perm[j];
        const float2 vv = vsm[cc][k], w = cs[k];
        ax += vv.x * w.x - vv.y * w.y;
        ay += vv.x * w.y + vv.y * w.x;
      }
    }
    ts[sl][cc] = make_float2(ax, ay);
  }
  __syncthreads();
  float2 x[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int sl = rowslot[t + 16 * j];
    x[j] = (act && sl >= 0) ? ts[sl][c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft256<false, GC>(x, tw, xs + c, t);
  const float ib = 1.0f / bg[b];
  float beta = 0.f;
  bool vu = false;
  if (DIR) {
    const double nr = nrho[lane], rh = rho[lane], cv = curv[lane];
    const double tiny = 1.17549435082228750797e-38;
    vu = active[lane] != 0 && isfinite(cv) && isfinite(rh) && cv > tiny && rh > tiny && isfinite(nr) && nr > tiny;
    beta = vu ? (float)(nr / fmax(rh, tiny)) : 0.f;
  }
  double acc = 0.0;
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int k = t + 16 * n;
    if (act && k < m1) {
      const float2 vv = vin[n];
      const float2 o = make_float2(vv.x * ib - x[n].x, vv.y * ib - x[n].y);
      if (DIR) {
        out[base + (long long)k * m2] =
            vu ? make_float2(o.x + beta * dl[n].x, o.y + beta * dl[n].y) : make_float2(0.f, 0.f);
      } else {
        out[base + (long long)k * m2] = o;
        if (part) acc += (double)vv.x * o.x + (double)vv.y * o.y;
      }
    }
  }
  if (!DIR && part) {
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[(long long)lane * gridDim.x + blockIdx.x] = acc;
  }
}

}  // namespace

// k_lr_corr of fbme_pcg.hip: corr = S^-1 (proj / bg) / bg from chunk partials.
int fbme_lr_corr_launch(int lanes, int rhs, int chunks, const void* part, const void* sinv, const float* bg,
                        void* corr, hipStream_t st, const double* rr_part, int rr_blocks, double* nrho);
int fbme_pcg_commit_launch(int lanes, double* rho, const double* curv, const double* nrho, int* active,
                           hipStream_t st);

extern "C" {

int fbme_lowrank_map(void* stream, int batch, int n1, int n2, const int* idx, int* map) {
  if (batch < 1 || n1 != 256 || n2 < 1) return FBME_ERR_ARG;
  hipLaunchKernelGGL(k_lrf_map, dim3(batch), dim3(64), 0, (hipStream_t)stream, idx, n2, map);
  return (int)hipGetLastError();
}

int fbme_lowrank_apply_fft_dot(void* stream, int lanes, int rhs, int m1, int m2, const void* factors,
                               const void* sinv, const float* bg, const int* map, const void* v, void* out,
                               void* wbuf, void* part, void* corr, const void* tw256, double* dot_part,
                               double* dot) {
  if (lanes < 1 || rhs < 1 || lanes % rhs != 0 || m1 < 1 || m2 < 1 || m1 > 256 || v == out) return FBME_ERR_ARG;
  if (dot != nullptr && dot_part == nullptr) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const dim3 cg((m2 + GC - 1) / GC, lanes);
  const float2* tab = (const float2*)factors;
  (void)wbuf;
  hipLaunchKernelGGL(k_lrf_col_project, cg, dim3(16 * GC), 0, st, (const float2*)v, tab, map, m1, m2, rhs,
                     (const float2*)tw256, (double2*)part);
  int rc = fbme_lr_corr_launch(lanes, rhs, (int)cg.x, part, sinv, bg, corr, st, nullptr, 0, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_lrf_col_apply<false>, cg, dim3(16 * GC), 0, st, (const float2*)corr, tab, map,
                     (const float2*)v, bg, (float2*)out, m1, m2, rhs, (const float2*)tw256, dot_part,
                     (const double*)nullptr, (const double*)nullptr, (const double*)nullptr, (const int*)nullptr);
  if (dot)
    hipLaunchKernelGGL(k_gop_finish, dim3(lanes), dim3(64), 0, st, (const double*)dot_part, (int)cg.x, lanes, dot);
  return (int)hipGetLastError();
}

int fbme_lowrank_apply_fft_direction(void* stream, int lanes, int rhs, int m1, int m2, const void* factors,
                                     const void* sinv, const float* bg, const int* map, const void* r, void* d,
                                     void* part, void* corr, const void* tw256, const double* rr_part,
                                     int rr_blocks, double* rho, const double* curv, double* nrho, int* active) {
  if (lanes < 1 || rhs < 1 || lanes % rhs != 0 || m1 < 1 || m2 < 1 || m1 > 256 || r == d || rr_blocks < 1 ||
      !rr_part || !rho || !curv || !nrho || !active)
    return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const dim3 cg((m2 + GC - 1) / GC, lanes);
  const float2* tab = (const float2*)factors;
  hipLaunchKernelGGL(k_lrf_col_project, cg, dim3(16 * GC), 0, st, (const float2*)r, tab, map, m1, m2, rhs,
                     (const float2*)tw256, (double2*)part);
  int rc = fbme_lr_corr_launch(lanes, rhs, (int)cg.x, part, sinv, bg, corr, st, rr_part, rr_blocks, nrho);
  if (rc) return rc;
  hipLaunchKernelGGL(k_lrf_col_apply<true>, cg, dim3(16 * GC), 0, st, (const float2*)corr, tab, map,
                     (const float2*)r, bg, (float2*)d, m1, m2, rhs, (const float2*)tw256, (double*)nullptr,
                     (const double*)rho, curv, (const double*)nrho, (const int*)active);
  rc = (int)hipGetLastError();
  if (rc) return rc;
  return fbme_pcg_commit_launch(lanes, rho, curv, nrho, active, st);
}

int fbme_grid_matvec_dot(void* stream, int lanes, int rhs, int m1, int m2, const float* gamma, const float* noise,
                         const void* d, void* out, void* mid, const void* tw256, double* part, double* dot) {
  if (lanes < 1 || rhs < 1 || lanes % rhs != 0 || m1 < 1 || m2 < 1 || m1 > 256 || m2 > 256 || d == out)
    return FBME_ERR_ARG;
  if (dot != nullptr && part == nullptr) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const dim3 cg((m2 + GC - 1) / GC, lanes);
  hipLaunchKernelGGL(k_gop_col_adjoint, cg, dim3(16 * GC), 0, st, (const float2*)d, (float2*)mid, m1, m2,
                     (const float2*)tw256);
  hipLaunchKernelGGL(k_gop_row, dim3(64, lanes), dim3(64), 0, st, (float2*)mid, gamma, m2, rhs,
                     (const float2*)tw256);
  if (part) {
    hipLaunchKernelGGL(k_gop_col_forward<true>, cg, dim3(16 * GC), 0, st, (const float2*)mid, (const float2*)d, noise,
                       (float2*)out, m1, m2, rhs, (const float2*)tw256, part);
    if (dot)
      hipLaunchKernelGGL(k_gop_finish, dim3(lanes), dim3(64), 0, st, (const double*)part, (int)cg.x, lanes, dot);
  } else {
    hipLaunchKernelGGL(k_gop_col_forward<false>, cg, dim3(16 * GC), 0, st, (const float2*)mid, (const float2*)d,
                       noise, (float2*)out, m1, m2, rhs, (const float2*)tw256, (double*)nullptr);
  }
  return (int)hipGetLastError();
}

int fbme_grid_matvec(void* stream, int lanes, int rhs, int m1, int m2, const float* gamma, const float* noise,
                     const void* d, void* out, void* mid, const void* tw256) {
  return fbme_grid_matvec_dot(stream, lanes, rhs, m1, m2, gamma, noise, d, out, mid, tw256, nullptr, nullptr);
}

}  // extern "C"
