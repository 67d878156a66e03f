// Native plan runtime of the FBME hot path.
//
// A plan is built once for a list of scene shapes (uniform or uneven):
//   * per-scene descriptors (radix codes of the FFT lengths, tile sizes,
//     buffer offsets, partial-sum slots, candidate slots, Dirichlet tables),
//   * work tables of the column / row / tile passes, ordered so that the
//     scenes still iterating at UAMP step t always form a prefix (scenes are
//     sorted by their iteration budget T_u = ceil(0.8 N / M) floored at 3,
//     torch_sbl/solver.py:1287-1309), which gives per-scene convergence masks
//     without any extra table or device-side branching,
//   * a layout of one caller-provided workspace (the caller's allocator owns
//     every byte, so nothing allocates inside a hipGraph capture).
// fbme_plan_run() enqueues the complete fixed stage sequence on one stream;
// it performs no allocation, no synchronisation and no memset, so the whole
// call can be captured into a graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "fbme_hip.h"
#include "fbme_types.h"

extern "C" {
int fbme_launch_col_adjoint(const FbmeScene*, const FbmeWork*, int, int, int, const void*, const unsigned char*,
                            const void*, const void*, void*, void*, const double*, double*, int, int, int, float,
                            const void*, hipStream_t);
int fbme_launch_row(const FbmeScene*, const FbmeWork*, int, int, int, void*, void*, float*, void*, float*,
                    const double*, double*, int, int, float, float, float, float, float, float, uint32_t*, int*,
                    int, uint32_t*, const void*, hipStream_t);
int fbme_launch_col_forward(const FbmeScene*, const FbmeWork*, int, int, int, const void*, const void*,
                            const unsigned char*, const void*, void*, void*, const double*, double*, int, int, float,
                            const void*, hipStream_t);
int fbme_launch_mask_bits16(const unsigned char*, int, int, void*, hipStream_t);
int fbme_launch_scalars(const FbmeScene*, int, double*, const double*, int, int, const unsigned char*, float,
                        float, float, float, float, float, int, float, hipStream_t);
size_t fbme_polish_state_bytes(void);
size_t fbme_polish_result_bytes(void);
int fbme_launch_polish(const FbmeScene*, int, int, const void*, const double*, const void*, const void*,
                       const uint32_t*,
                       const int*, const uint32_t*, const int*, void*, void*, float*, float*, float*, float*,
                       long long*, float*, int*, const float*, const FbmePlanOptions*, hipStream_t);
int fbme_launch_gamma_sum(const FbmeScene*, int, const float*, double*, hipStream_t);
int fbme_launch_exchange_score(const FbmeScene*, const FbmeWork*, int, int, const void*, const void*, const void*,
                               const void*, uint32_t*, int*, int, hipStream_t);
int fbme_launch_output_uamp(const FbmeScene*, const FbmeWork*, int, int, const double*, const void*, const float*,
                            const float*, void*, float*, float*, unsigned char*, float, int, float*, float*, float*,
                            hipStream_t);
int fbme_launch_output(const FbmeScene*, const FbmeWork*, int, int, const double*, const void*, void*, float*,
                       float*, unsigned char*, float, int, const float*, hipStream_t);
int fbme_launch_truncation(const FbmeScene*, int, long long, long long, const float*, const void*, const double*,
                           float*, const void*, float*, int, hipStream_t);
}

namespace {

struct Region {
  size_t off = 0, bytes = 0;
};

}  // namespace

struct fbme_plan {
  FbmePlanOptions opt;
  int nscenes = 0;
  int max_iters = 0;
  std::vector<FbmeScene> scenes;            // sorted by iteration budget (desc)
  std::vector<FbmeWork> col_work, row_work, tile_work;
  // Work lists hold the LDS-Stockham items first and the register 16 x 16
  // items (scenes with a 256-point pass axis) after col_split / row_split.
  int col_split = 0, row_split = 0;
  std::vector<int> col_prefix, row_prefix, scene_prefix;  // per UAMP step t (index t-1)
  std::vector<int> col16_prefix, row16_prefix;
  std::vector<double> dir;                  // Dirichlet tables, interleaved (re, im)
  std::vector<float> twid;                  // twiddle tables, interleaved (re, im)
  std::vector<long long> grid_off, samp_off; // caller order
  long long total_grid = 0, total_samp = 0, total_mid = 0;
  int total_slots = 0, total_cand = 0, total_xcand = 0;
  int ept_col = 16, ept_row = 16;           // elements per thread of the passes
  float threshold = 0.f;
  // workspace layout
  Region r_scenes, r_col, r_row, r_tile, r_dir, r_tw, r_sc, r_part, r_mid, r_msg, r_pred, r_bp, r_gamma, r_mean, r_var,
      r_ckey, r_cidx, r_xkey, r_xidx, r_state, r_result, r_score;
  // observation masks: spectra of the masked scenes and the scratch of
  // their computation; the caller's mask bytes (packed like the data)
  Region r_mh, r_mhwork, r_mbits;
  // truncation diagnostic: tail variances [nscenes, N] and the two work
  // buffers of the BTTB spectrum [nscenes, max(N, E1 E2)] complex
  Region r_ttail, r_twa, r_twb;
  int te1 = 0, te2 = 0, ttw_n2 = 0, ttw_e2 = 0;
  int masked = 0;                           // scenes with an observation mask
  std::vector<int> mh_leaders;              // caller indices owning a spectrum
  const unsigned char* mask = nullptr;
  size_t workspace_bytes = 0;
  char* ws = nullptr;
  int bound = 0;
  // A pass whose scenes split between the LDS-Stockham list and the
  // register 16 x 16 list runs the two as parallel branches: the register
  // list on `side`, forked from and joined to the caller's stream by events
  // (a graph capture records the fork and join as edges).  Both lists touch
  // disjoint scenes and partial-sum slots.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  ~fbme_plan() {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
  }
};

namespace {

unsigned radix_code(int len, int prefer8, std::string* err) {
  if (len < 1 || len > 1024) {
    if (err) *err = "FFT length " + std::to_string(len) + " outside [1, 1024]";
    return 0xFFFFFFFFu;
  }
  std::vector<int> r;
  int rem = len;
  while (rem % 5 == 0) { r.push_back(5); rem /= 5; }
  while (rem % 3 == 0) { r.push_back(3); rem /= 3; }
  if (prefer8) {
    while (rem % 8 == 0) { r.push_back(8); rem /= 8; }
  }
  while (rem % 4 == 0) { r.push_back(4); rem /= 4; }
  while (rem % 2 == 0) { r.push_back(2); rem /= 2; }
  if (rem != 1) {
    if (err) *err = "FFT length " + std::to_string(len) + " has a prime factor other than 2, 3, 5";
    return 0xFFFFFFFFu;
  }
  if (r.size() > 7) {
    if (err) *err = "FFT length " + std::to_string(len) + " needs too many radix stages";
    return 0xFFFFFFFFu;
  }
  unsigned code = 0;
  for (size_t i = 0; i < r.size(); ++i) code |= (unsigned)r[i] << (4 * i);
  return code;  // len == 1 gives 0 stages (identity)
}

// D[d] = sum_{j<m} exp(2 pi i j d / n), exact closed form with integer
// argument reduction.
void dirichlet(int m, int n, std::vector<double>& out) {
  const double pi = 3.14159265358979323846;
  for (int d = 0; d < n; ++d) {
    if (d == 0) {
      out.push_back((double)m);
      out.push_back(0.0);
      continue;
    }
    const long long two_n = 2LL * n;
    const long long a = ((long long)d * (m - 1)) % two_n;      // phase pi a / n
    const long long b = ((long long)d * m) % two_n;            // sin(pi b / n)
    const double ratio = std::sin(pi * (double)b / n) / std::sin(pi * (double)d / n);
    out.push_back(ratio * std::cos(pi * (double)a / n));
    out.push_back(ratio * std::sin(pi * (double)a / n));
  }
}

int iteration_budget(long long n, long long m, const FbmePlanOptions& o) {
  // min(max_iter, max(min_iters, ceil(offset + scale N / M)))
  // (solver.py:1280-1309; paper preset offset 0, scale 0.8, floor 3)
  const double v = std::ceil((double)o.iteration_offset + (double)o.iteration_scale * (double)n / (double)m);
  long long b = (long long)v;
  b = std::max<long long>(o.min_iters, b);
  b = std::min<long long>(o.max_iters, b);
  return (int)b;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

extern "C" {

int fbme_abi_version(void) { return FBME_ABI_VERSION; }

int fbme_dirichlet_table(int m, int n, double* out) {
  if (m < 1 || n < 1 || m > n || !out) return FBME_ERR_ARG;
  std::vector<double> t;
  t.reserve(2 * (size_t)n);
  dirichlet(m, n, t);
  std::memcpy(out, t.data(), sizeof(double) * t.size());
  return 0;
}

void fbme_default_options(FbmePlanOptions* o, int stable) {
  o->uamp_damping = 0.05f;
  o->damping = 0.15f;
  o->min_variance = 1e-12f;
  o->max_precision = 1e12f;
  o->blend = 0.1f;
  o->noise_fraction = 1e-2f;
  o->margin_factor = 64.0f;
  o->jitter_multiplier = 8.0f;
  o->false_alarm = 0.5f;
  o->polish_size = 32;
  o->polish_iterations = stable ? 3 : 2;
  o->exchange_size = stable ? 12 : 0;
  o->diagnostics = stable ? 1 : 0;
  o->evidence_noise = 0;
  o->row_lines = 4;
  o->col_lines = 8;
  o->iteration_offset = 0.0f;
  o->iteration_scale = 0.8f;
  o->min_iters = 3;
  o->max_iters = 200;
  o->hyper_shape = 0.0f;
  o->hyper_rate = 0.0f;
  o->noise_shape = 0.0f;
  o->noise_rate = 0.0f;
  o->polish_damping = 0.0f;
  o->fft_kernel = 0;
  o->truncation_diagnostic = 0;
  o->prune_tail = 1;
  o->reinitialize = 1;
  o->evidence_gate = 1;
  o->fixed_noise = 0.0f;
}

fbme_plan* fbme_plan_create(int nscenes, const int* shapes, const FbmePlanOptions* opt, char* err, int errlen) {
  return fbme_plan_create_masked(nscenes, shapes, nullptr, nullptr, opt, err, errlen);
}

fbme_plan* fbme_plan_create_masked(int nscenes, const int* shapes, const int* observed, const int* mask_group,
                                   const FbmePlanOptions* opt, char* err, int errlen) {
  auto fail = [&](const std::string& msg) -> fbme_plan* {
    if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", msg.c_str());
    return nullptr;
  };
  if (nscenes < 1) return fail("a plan needs at least one scene");
  if (!opt) return fail("options pointer is null");
  if (opt->polish_size < 0 || opt->polish_size > FBME_KMAX)
    return fail("polish_size must be in [0, " + std::to_string(FBME_KMAX) + "] (0: no polish)");
  if (opt->polish_iterations < 1 || opt->polish_iterations > 8)
    return fail("polish_iterations must be in [1, 8]");
  if (opt->exchange_size < 0 || opt->exchange_size > 32) return fail("exchange_size must be in [0, 32]");
  if (opt->exchange_size > 0 && opt->polish_iterations < 2)
    return fail("residual exchange requires at least two polish evaluations");
  if (opt->min_iters < 1 || opt->max_iters < 1) return fail("iteration limits must be positive");
  if (opt->truncation_diagnostic && !opt->prune_tail)
    return fail("the truncation diagnostic certifies the zero-tail (pruned) model only");
  if (opt->exchange_size > 0 && (!opt->prune_tail || !opt->evidence_gate))
    return fail("the residual exchange runs on the pruned, evidence-gated polish only");
  if (opt->diagnostics && !opt->prune_tail)
    return fail("the fixed-budget diagnostics run on the pruned polish only");
  if (!(opt->fixed_noise >= 0.0f) || !std::isfinite(opt->fixed_noise))
    return fail("fixed_noise must be 0 (estimated) or a positive noise variance");
  fbme_plan* p = new fbme_plan();
  p->opt = *opt;
  p->nscenes = nscenes;
  std::vector<FbmeScene> sc(nscenes);
  std::string e;
  for (int i = 0; i < nscenes; ++i) {
    const int n1 = shapes[4 * i], n2 = shapes[4 * i + 1], m1 = shapes[4 * i + 2], m2 = shapes[4 * i + 3];
    if (m1 < 1 || m2 < 1 || m1 > n1 || m2 > n2) {
      delete p;
      return fail("scene " + std::to_string(i) + ": need 0 < M <= N on both axes");
    }
    FbmeScene& s = sc[i];
    std::memset(&s, 0, sizeof(s));
    s.n1 = n1; s.n2 = n2; s.m1 = m1; s.m2 = m2;
    s.rad1 = radix_code(n1, 1, &e);
    s.rad2 = radix_code(n2, 1, &e);
    if (s.rad1 == 0xFFFFFFFFu || s.rad2 == 0xFFFFFFFFu) {
      delete p;
      return fail("scene " + std::to_string(i) + ": " + e);
    }
    const long long n = (long long)n1 * n2, m = (long long)m1 * m2;
    const long long mo = observed ? observed[i] : m;
    if (mo < 1 || mo > m) {
      delete p;
      return fail("scene " + std::to_string(i) + ": observed samples must be in [1, M1 M2]");
    }
    s.m_obs = (int)mo;
    s.mh_off = -1;
    s.mbits_off = -1;
    s.iters = iteration_budget(n, mo, *opt);
    s.k = (int)std::min<long long>(opt->polish_size, n);
    s.orig = i;
  }
  // caller-order offsets
  p->grid_off.resize(nscenes);
  p->samp_off.resize(nscenes);
  long long go = 0, so = 0, mo = 0;
  for (int i = 0; i < nscenes; ++i) {
    p->grid_off[i] = go;
    p->samp_off[i] = so;
    sc[i].grid_off = go;
    sc[i].samp_off = so;
    sc[i].mid_off = mo;
    go += (long long)sc[i].n1 * sc[i].n2;
    so += (long long)sc[i].m1 * sc[i].m2;
    mo += (long long)sc[i].n1 * sc[i].m2;
  }
  p->total_grid = go;
  p->total_samp = so;
  p->total_mid = mo;
  // mask spectra [N1*N2] (double2) of the masked scenes, caller order; a
  // scene whose mask_group names an earlier scene of the same shape shares
  // that scene's mask and spectrum
  long long mh_total = 0, mh_work = 0, mbits_total = 0;
  for (int i = 0; i < nscenes; ++i) {
    if (sc[i].m_obs == sc[i].m1 * sc[i].m2) continue;
    ++p->masked;
    const int g = mask_group ? mask_group[i] : i;
    if (g < 0 || g > i) {
      delete p;
      return fail("scene " + std::to_string(i) + ": mask_group must name this or an earlier scene");
    }
    if (g != i) {
      const FbmeScene& a = sc[g];
      if (a.n1 != sc[i].n1 || a.n2 != sc[i].n2 || a.m1 != sc[i].m1 || a.m2 != sc[i].m2 ||
          a.m_obs != sc[i].m_obs || a.mh_off < 0) {
        delete p;
        return fail("scene " + std::to_string(i) + ": mask_group scene differs in shape or observed count");
      }
      sc[i].mh_off = a.mh_off;
      sc[i].mbits_off = a.mbits_off;
      continue;
    }
    sc[i].mh_off = mh_total;
    if (sc[i].n1 == 256) {   // the register column passes read the mask as 16-bit words
      sc[i].mbits_off = (int)mbits_total;
      mbits_total += (long long)sc[i].m2 * 16;
    }
    p->mh_leaders.push_back(i);
    mh_total += (long long)sc[i].n1 * sc[i].n2;
    mh_work = std::max<long long>(mh_work, (long long)sc[i].m1 * sc[i].n2);
  }
  // sort by iteration budget, descending (stable -> deterministic)
  std::vector<int> order(nscenes);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sc[a].iters > sc[b].iters; });
  p->scenes.resize(nscenes);
  for (int i = 0; i < nscenes; ++i) p->scenes[i] = sc[order[i]];
  // Dirichlet tables per distinct (m, n) pair
  std::vector<std::pair<int, int>> keys;
  std::vector<int> koff;
  auto table = [&](int m, int n) -> int {
    for (size_t i = 0; i < keys.size(); ++i)
      if (keys[i].first == m && keys[i].second == n) return koff[i];
    const int off = (int)(p->dir.size() / 2);
    dirichlet(m, n, p->dir);
    keys.push_back({m, n});
    koff.push_back(off);
    return off;
  };
  // twiddle tables per distinct FFT length, tw[i] = exp(-2 pi i * i / L)
  std::vector<int> tw_len, tw_off;
  auto twiddle = [&](int len) -> int {
    for (size_t i = 0; i < tw_len.size(); ++i)
      if (tw_len[i] == len) return tw_off[i];
    const int off = (int)(p->twid.size() / 2);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < len; ++i) {
      // exact integer reduction of the angle to [0, 2 pi)
      const double a = 2.0 * pi * (double)i / (double)len;
      p->twid.push_back((float)std::cos(a));
      p->twid.push_back((float)-std::sin(a));
    }
    tw_len.push_back(len);
    tw_off.push_back(off);
    return off;
  };
  // FFT kernel variant of the column (length N1) and row (length N2) passes.
  if (opt->fft_kernel < 0 || opt->fft_kernel > 2) {
    delete p;
    return fail("fft_kernel must be 0 (auto), 1 (LDS Stockham) or 2 (register 16 x 16)");
  }
  // Per scene: the register kernels take every pass whose axis has length
  // 256 (the row pass also needs its four-row items to keep the candidate
  // merge within one tile); the rest use the LDS Stockham kernels.
  auto r16_col_ok = [](const FbmeScene& s) { return s.n1 == 256; };
  auto r16_row_ok = [](const FbmeScene& s) {
    return s.n2 == 256 && ((s.n1 + FBME_R16_ROWS - 1) / FBME_R16_ROWS) * s.k <= FBME_TILE_MAX;
  };
  if (opt->fft_kernel == 2) {
    for (const FbmeScene& s : p->scenes) {
      if (!r16_col_ok(s) || !r16_row_ok(s)) {
        delete p;
        return fail("fft_kernel 2 needs N1 = N2 = 256 in every scene");
      }
    }
  }
  const bool use_r16 = opt->fft_kernel != 1;
  std::vector<FbmeWork> col16, row16;
  // work tables
  int slots = 0, cand = 0, xcand = 0;
  const int exch_cfg = opt->exchange_size;
  for (int si = 0; si < nscenes; ++si) {
    FbmeScene& s = p->scenes[si];
    s.dir1_off = table(s.m1, s.n1);
    s.dir2_off = table(s.m2, s.n2);
    s.tw1_off = twiddle(s.n1);
    s.tw2_off = twiddle(s.n2);
    const bool r16_col = use_r16 && r16_col_ok(s);
    const bool r16_row = use_r16 && r16_row_ok(s);
    const int cl = r16_col ? std::min(s.m2, FBME_R16_COLS)
                           : std::max(1, std::min({s.m2, FBME_TILE_MAX / s.n1, std::max(1, opt->col_lines)}));
    // Row tiles hold rl rows of n2 + 1 (padded) complex values in LDS, so
    // rl * (n2 + 1) <= FBME_TILE_MAX + 64 requires rl <= 64 as well.
    const int rl_max = std::max(1, std::min({s.n1, FBME_TILE_MAX / s.n2, 64}));
    int rl = r16_row ? FBME_R16_ROWS : std::max(1, std::min(rl_max, std::max(1, opt->row_lines)));
    // keep the candidate merge within one tile of LDS
    while (((s.n1 + rl - 1) / rl) * s.k > FBME_TILE_MAX && rl < rl_max) ++rl;
    s.col_items = (s.m2 + cl - 1) / cl;
    s.row_items = (s.n1 + rl - 1) / rl;
    if (s.row_items * s.k > FBME_TILE_MAX) {
      delete p;
      return fail("grid too large for the candidate merge of the polish selection");
    }
    s.col_part = slots;
    slots += s.col_items;
    s.row_part = slots;
    slots += s.row_items;
    s.cand_off = cand;
    cand += s.row_items * s.k;
    const long long n = (long long)s.n1 * s.n2;
    s.xtile_items = (int)((n + FBME_XTILE - 1) / FBME_XTILE);
    const int exch = std::min<long long>(exch_cfg, std::min<long long>(std::max(s.k - 1, 0), n - s.k));
    s.xcand_off = xcand;
    xcand += s.xtile_items * std::max(exch, 0);
    if (exch > 0 && s.xtile_items * exch > FBME_TILE_MAX) {
      delete p;
      return fail("grid too large for the exchange candidate merge");
    }
    for (int c = 0; c < s.col_items; ++c)
      (r16_col ? col16 : p->col_work).push_back({si, c * cl, std::min(cl, s.m2 - c * cl), c});
    for (int r = 0; r < s.row_items; ++r)
      (r16_row ? row16 : p->row_work).push_back({si, r * rl, std::min(rl, s.n1 - r * rl), r});
    for (int t = 0; t < s.xtile_items; ++t) {
      const long long first = (long long)t * FBME_XTILE;
      p->tile_work.push_back({si, (int)first, (int)std::min<long long>(FBME_XTILE, n - first), t});
    }
  }
  // Truncation diagnostic: the BTTB embedding of the full sample block
  // (operators.py:10-20, 100-105: first 2/3/5/7-smooth length >= 2M - 1 per
  // axis; an observation mask does not enter it, operators.py:299-326) on
  // the FFT passes of the general solver, for one shared shape.
  if (opt->truncation_diagnostic) {
    const FbmeScene& a = p->scenes[0];
    for (const FbmeScene& s2 : p->scenes) {
      if (s2.n1 != a.n1 || s2.n2 != a.n2 || s2.m1 != a.m1 || s2.m2 != a.m2) {
        delete p;
        return fail("the truncation diagnostic needs one grid and sample shape for every scene");
      }
    }
    auto smooth = [](int v) {
      for (int c = v;; ++c) {
        int r = c;
        for (int f : {2, 3, 5, 7})
          while (r % f == 0) r /= f;
        if (r == 1) return c;
      }
    };
    p->te1 = smooth(2 * a.m1 - 1);
    p->te2 = smooth(2 * a.m2 - 1);
    if (radix_code(p->te1, 1, &e) == 0xFFFFFFFFu || radix_code(p->te2, 1, &e) == 0xFFFFFFFFu) {
      delete p;
      return fail("truncation diagnostic: embedding " + std::to_string(p->te1) + " x " + std::to_string(p->te2) +
                  ": " + e);
    }
    p->ttw_n2 = twiddle(a.n2);
    p->ttw_e2 = twiddle(p->te2);
  }
  auto ept_for = [](int elems) { return elems <= 1024 ? 4 : elems <= 2048 ? 8 : 16; };
  int max_col = 1, max_row = 1;
  for (const FbmeWork& w : p->col_work) max_col = std::max(max_col, w.count * p->scenes[w.scene].n1);
  for (const FbmeWork& w : p->row_work) max_row = std::max(max_row, w.count * p->scenes[w.scene].n2);
  p->ept_col = ept_for(max_col);
  p->ept_row = ept_for(max_row);
  p->col_split = (int)p->col_work.size();
  p->row_split = (int)p->row_work.size();
  p->col_work.insert(p->col_work.end(), col16.begin(), col16.end());
  p->row_work.insert(p->row_work.end(), row16.begin(), row16.end());
  p->total_slots = slots;
  p->total_cand = cand;
  p->total_xcand = std::max(xcand, 1);
  p->max_iters = p->scenes[0].iters;
  // prefix counts per UAMP step
  for (int t = 1; t <= p->max_iters; ++t) {
    int ns = 0;
    while (ns < nscenes && p->scenes[ns].iters >= t) ++ns;
    // scenes are sorted, so the live items of each list form a prefix of it
    auto live = [&](const std::vector<FbmeWork>& v, int lo, int hi) {
      int n = 0;
      while (lo + n < hi && v[lo + n].scene < ns) ++n;
      return n;
    };
    p->scene_prefix.push_back(ns);
    p->col_prefix.push_back(live(p->col_work, 0, p->col_split));
    p->row_prefix.push_back(live(p->row_work, 0, p->row_split));
    p->col16_prefix.push_back(live(p->col_work, p->col_split, (int)p->col_work.size()));
    p->row16_prefix.push_back(live(p->row_work, p->row_split, (int)p->row_work.size()));
  }
  // support-mask threshold log(N / budget) is per scene in general; the
  // locked plans use one grid, uneven batches get the mask per scene shape
  // through the same formula evaluated on the host for the first scene.
  p->threshold = 0.f;
  // workspace layout
  size_t off = 0;
  auto reg = [&](Region& r, size_t bytes) {
    off = align_up(off, 256);
    r.off = off;
    r.bytes = bytes;
    off += bytes;
  };
  reg(p->r_scenes, sizeof(FbmeScene) * nscenes);
  reg(p->r_col, sizeof(FbmeWork) * p->col_work.size());
  reg(p->r_row, sizeof(FbmeWork) * p->row_work.size());
  reg(p->r_tile, sizeof(FbmeWork) * p->tile_work.size());
  reg(p->r_dir, sizeof(double) * p->dir.size());
  reg(p->r_tw, sizeof(float) * p->twid.size());
  reg(p->r_sc, sizeof(double) * SC_COUNT * nscenes);
  reg(p->r_part, sizeof(double) * PT_COUNT * (size_t)slots);
  reg(p->r_mid, 8 * (size_t)p->total_mid);
  reg(p->r_msg, 8 * (size_t)p->total_samp);
  reg(p->r_pred, 8 * (size_t)p->total_samp);
  reg(p->r_bp, 8 * (size_t)p->total_grid);
  reg(p->r_gamma, 4 * (size_t)p->total_grid);
  reg(p->r_mean, 8 * (size_t)p->total_grid);
  reg(p->r_var, 4 * (size_t)p->total_grid);
  reg(p->r_ckey, 4 * (size_t)cand);
  reg(p->r_cidx, 4 * (size_t)cand);
  reg(p->r_xkey, 4 * (size_t)p->total_xcand);
  reg(p->r_xidx, 4 * (size_t)p->total_xcand);
  reg(p->r_state, fbme_polish_state_bytes() * nscenes);
  reg(p->r_result, fbme_polish_result_bytes() * nscenes);
  // polish scores of the register-FFT row tiles (selected by a separate pass)
  long long score_cells = 0;
  for (size_t i = p->row_split; i < p->row_work.size(); ++i)
    score_cells = std::max<long long>(score_cells, p->scenes[p->row_work[i].scene].grid_off +
                                                       (long long)p->scenes[p->row_work[i].scene].n1 * 256);
  reg(p->r_score, 4 * (size_t)score_cells);
  reg(p->r_mh, 16 * (size_t)mh_total);
  reg(p->r_mhwork, 16 * (size_t)mh_work);
  reg(p->r_mbits, 2 * (size_t)mbits_total);
  if (opt->truncation_diagnostic) {
    const long long n = (long long)p->scenes[0].n1 * p->scenes[0].n2;
    const long long big = std::max<long long>(n, (long long)p->te1 * p->te2);
    reg(p->r_ttail, 4 * (size_t)(nscenes * n));
    reg(p->r_twa, 8 * (size_t)(nscenes * big));
    reg(p->r_twb, 8 * (size_t)(nscenes * big));
  }
  p->workspace_bytes = align_up(off, 256);
  return p;
}

void fbme_plan_destroy(fbme_plan* p) { delete p; }

long long fbme_plan_query(const fbme_plan* p, int what) {
  if (!p) return -1;
  switch (what) {
    case FBME_Q_WORKSPACE_BYTES: return (long long)p->workspace_bytes;
    case FBME_Q_TOTAL_GRID: return p->total_grid;
    case FBME_Q_TOTAL_SAMPLES: return p->total_samp;
    case FBME_Q_MAX_ITERS: return p->max_iters;
    case FBME_Q_NSCENES: return p->nscenes;
    case FBME_Q_EVALUATIONS: return p->opt.polish_iterations;
    case FBME_Q_LAUNCHES: {
      long long n = p->row_split > 0 ? 3 : 2;  // init
      for (int t = 1; t <= p->max_iters; ++t) n += (t > 1) ? 4 : 3;
      n += 2 + 1;       // final E-step + polish
      const bool exch = p->opt.exchange_size > 0 && p->total_xcand > 1;
      if (exch) n += 2;
      n += 1;           // output
      if (p->opt.truncation_diagnostic) n += 9;
      if (!p->opt.prune_tail) n += 1;
      return n;
    }
    case FBME_Q_R16_COL_ITEMS: return (long long)p->col_work.size() - p->col_split;
    case FBME_Q_R16_ROW_ITEMS: return (long long)p->row_work.size() - p->row_split;
    default: return -1;
  }
}

int fbme_plan_scene_info(const fbme_plan* p, int scene, long long* info /* [8] */) {
  if (!p || scene < 0 || scene >= p->nscenes || !info) return FBME_ERR_ARG;
  for (const FbmeScene& s : p->scenes) {
    if (s.orig != scene) continue;
    info[0] = s.grid_off;
    info[1] = s.samp_off;
    info[2] = s.iters;
    info[3] = s.k;
    info[4] = s.rad1;
    info[5] = s.rad2;
    info[6] = s.col_items;
    info[7] = s.row_items;
    return 0;
  }
  return FBME_ERR_ARG;
}

int fbme_plan_bind(fbme_plan* p, void* workspace, void* stream) {
  if (!p || !workspace) return FBME_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  p->ws = (char*)workspace;
  auto up = [&](const Region& r, const void* src) -> hipError_t {
    if (r.bytes == 0) return hipSuccess;
    return hipMemcpyAsync(p->ws + r.off, src, r.bytes, hipMemcpyHostToDevice, st);
  };
  hipError_t e = up(p->r_scenes, p->scenes.data());
  if (e == hipSuccess) e = up(p->r_col, p->col_work.data());
  if (e == hipSuccess) e = up(p->r_row, p->row_work.data());
  if (e == hipSuccess) e = up(p->r_tile, p->tile_work.data());
  if (e == hipSuccess) e = up(p->r_dir, p->dir.data());
  if (e == hipSuccess) e = up(p->r_tw, p->twid.data());
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  // side branch of passes split between both kernel lists (default priority)
  const bool split_col = p->col_split > 0 && p->col_split < (int)p->col_work.size();
  const bool split_row = p->row_split > 0 && p->row_split < (int)p->row_work.size();
  if (e == hipSuccess && !p->side && (split_col || split_row))
    e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
  if (e == hipSuccess && p->side && !p->ev_fork) e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess && p->side && !p->ev_join) e = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) return FBME_ERR_HIP;
  p->bound = 1;
  return 0;
}

int fbme_plan_set_mask(fbme_plan* p, const unsigned char* mask, void* stream) {
  if (!p || !p->bound || !mask || !p->masked) return FBME_ERR_ARG;
  char* w = p->ws;
  for (const FbmeScene& s : p->scenes) {
    if (s.mh_off < 0 || std::find(p->mh_leaders.begin(), p->mh_leaders.end(), s.orig) == p->mh_leaders.end())
      continue;
    int rc = fbme_mask_spectrum(stream, s.n1, s.n2, s.m1, s.m2, mask + s.samp_off,
                                w + p->r_mh.off + 16 * (size_t)s.mh_off, w + p->r_mhwork.off);
    if (!rc && s.mbits_off >= 0)
      rc = fbme_launch_mask_bits16(mask + s.samp_off, s.m1, s.m2, w + p->r_mbits.off + 2 * (size_t)s.mbits_off,
                                   (hipStream_t)stream);
    if (rc) return rc;
  }
  p->mask = mask;
  return 0;
}

namespace {

// One phase of a plan's stage sequence on stream st: 0 the initialisation,
// 1 UAMP iteration t_step (track: also the relative change of the step into
// SC_CHANGE), 2 the final E-step, polish (or, without a polish, the UAMP
// posterior) and outputs.  fbme_plan_run runs 0, 1 for t = 1 .. max_iters
// and 2.
int plan_phase(const fbme_plan* p, const FbmeIO* io, hipStream_t st, int phase, int t_step, int track) {
  if (!p || !io || !p->bound) return FBME_ERR_ARG;
  if (p->masked && !p->mask) return FBME_ERR_ARG;
  if (!io->y || !io->x || !io->posterior_variance || !io->precision || !io->noise || !io->dof || !io->nle ||
      !io->trace)
    return FBME_ERR_ARG;
  if (p->opt.truncation_diagnostic && !io->truncation) return FBME_ERR_ARG;
  char* w = p->ws;
  const FbmeScene* scenes = (const FbmeScene*)(w + p->r_scenes.off);
  const FbmeWork* colw = (const FbmeWork*)(w + p->r_col.off);
  const FbmeWork* roww = (const FbmeWork*)(w + p->r_row.off);
  const FbmeWork* tilew = (const FbmeWork*)(w + p->r_tile.off);
  const void* dir = w + p->r_dir.off;
  const void* mh = p->masked ? (const void*)(w + p->r_mh.off) : nullptr;
  const unsigned char* mask = p->masked ? p->mask : nullptr;
  const void* mbits = (p->masked && p->r_mbits.bytes) ? (const void*)(w + p->r_mbits.off) : nullptr;
  const void* tw = w + p->r_tw.off;
  double* sc = (double*)(w + p->r_sc.off);
  double* part = (double*)(w + p->r_part.off);
  void* mid = w + p->r_mid.off;
  void* msg = w + p->r_msg.off;
  void* pred = w + p->r_pred.off;
  void* bp = w + p->r_bp.off;
  float* gamma = (float*)(w + p->r_gamma.off);
  void* mean = w + p->r_mean.off;
  float* var = (float*)(w + p->r_var.off);
  uint32_t* ckey = (uint32_t*)(w + p->r_ckey.off);
  int* cidx = (int*)(w + p->r_cidx.off);
  uint32_t* xkey = (uint32_t*)(w + p->r_xkey.off);
  int* xidx = (int*)(w + p->r_xidx.off);
  void* states = w + p->r_state.off;
  void* results = w + p->r_result.off;
  const FbmePlanOptions& o = p->opt;
  const int ps = p->total_slots;
  const int ncol = (int)p->col_work.size(), nrow = (int)p->row_work.size(), ntile = (int)p->tile_work.size();
  const float hs = o.hyper_shape, hr = o.hyper_rate;
  int rc = 0;
#define FBME_TRY(x)            \
  do {                         \
    rc = (x);                  \
    if (rc) return FBME_ERR_HIP; \
  } while (0)
  // Each pass launches its LDS-Stockham list and its register 16 x 16 list.
  const FbmeWork* colw16 = colw + p->col_split;
  const FbmeWork* roww16 = roww + p->row_split;
  const int ncol16 = ncol - p->col_split, nrow16 = nrow - p->row_split;
  // Runs a pass's two kernel lists: in parallel branches when both have
  // work (the register list on the side stream, between a fork and a join
  // event), else in order on st.
  auto split = [&](int ng, int n16, auto&& lds, auto&& r16) -> int {
    if (ng > 0 && n16 > 0 && p->side) {
      if (hipEventRecord(p->ev_fork, st) != hipSuccess || hipStreamWaitEvent(p->side, p->ev_fork, 0) != hipSuccess)
        return FBME_ERR_HIP;
      int r = lds(st);
      if (!r) r = r16(p->side);
      if (hipEventRecord(p->ev_join, p->side) != hipSuccess || hipStreamWaitEvent(st, p->ev_join, 0) != hipSuccess)
        return FBME_ERR_HIP;
      return r;
    }
    int r = lds(st);
    return r ? r : r16(st);
  };
  auto col_adjoint = [&](int ng, int n16, int mode, int store) {
    return split(
        ng, n16,
        [&](hipStream_t q) {
          return fbme_launch_col_adjoint(scenes, colw, ng, ng, p->ept_col, io->y, mask, mbits, pred, msg, mid, sc,
                                         part, ps, mode, store, o.uamp_damping, tw, q);
        },
        [&](hipStream_t q) {
          return fbme_launch_col_adjoint(scenes, colw16, n16, n16, FBME_EPT_R16, io->y, mask, mbits, pred, msg, mid,
                                         sc, part, ps, mode, store, o.uamp_damping, tw, q);
        });
  };
  uint32_t* score = p->r_score.bytes ? (uint32_t*)(w + p->r_score.off) : nullptr;
  auto row = [&](int ng, int n16, int mode, uint32_t* key, int* idx, int store) {
    return split(
        ng, n16,
        [&](hipStream_t q) {
          return fbme_launch_row(scenes, roww, ng, ng, p->ept_row, mid, bp, gamma, mean, var, sc, part, ps, mode,
                                 o.uamp_damping, o.damping, o.max_precision, hs, hr, o.blend, key, idx, store,
                                 nullptr, tw, q);
        },
        [&](hipStream_t q) {
          return fbme_launch_row(scenes, roww16, n16, n16, FBME_EPT_R16, mid, bp, gamma, mean, var, sc, part, ps,
                                 mode, o.uamp_damping, o.damping, o.max_precision, hs, hr, o.blend, key, idx, store,
                                 score, tw, q);
        });
  };
  auto col_forward = [&](int ng, int n16, int first) {
    return split(
        ng, n16,
        [&](hipStream_t q) {
          return fbme_launch_col_forward(scenes, colw, ng, ng, p->ept_col, mid, io->y, mask, mbits, pred, msg, sc,
                                         part, ps, first, o.uamp_damping, tw, q);
        },
        [&](hipStream_t q) {
          return fbme_launch_col_forward(scenes, colw16, n16, n16, FBME_EPT_R16, mid, io->y, mask, mbits, pred, msg,
                                         sc, part, ps, first, o.uamp_damping, tw, q);
        });
  };
  if (phase == 0) {
    // Initialisation: A^H y, power, floor, sigma^2_0, sum gamma_0.
    FBME_TRY(col_adjoint(p->col_split, ncol16, 0, 0));
    // Register-FFT rows form A^H y inside their first UAMP step (every scene
    // runs at least one: min_iters >= 1); the LDS kernels write it here.
    FBME_TRY(row(p->row_split, 0, 0, nullptr, nullptr, 0));
    FBME_TRY(fbme_launch_scalars(scenes, p->nscenes, sc, part, ps, 0, nullptr, o.min_variance, o.noise_fraction,
                                 o.blend, o.damping, o.noise_shape, o.noise_rate, o.evidence_noise, o.fixed_noise,
                                 st));
    return 0;
  }
  if (phase == 1) {
    // UAMP iteration t over the scenes whose budget T_u reaches t
    const int t = t_step;
    if (t < 1 || t > p->max_iters) return FBME_ERR_ARG;
    const int nc = p->col_prefix[t - 1], nr = p->row_prefix[t - 1], ns = p->scene_prefix[t - 1];
    const int nc16 = p->col16_prefix[t - 1], nr16 = p->row16_prefix[t - 1];
    if (t > 1) FBME_TRY(col_adjoint(nc, nc16, 1, 1));
    FBME_TRY(row(nr, nr16, (t == 1 ? 1 : 2) + (track ? 3 : 0), nullptr, nullptr, 0));
    FBME_TRY(col_forward(nc, nc16, t == 1 ? 1 : 0));
    FBME_TRY(fbme_launch_scalars(scenes, ns, sc, part, ps, track ? 2 : 1, nullptr, o.min_variance, o.noise_fraction,
                                 o.blend, o.damping, o.noise_shape, o.noise_rate, o.evidence_noise, o.fixed_noise,
                                 st));
    return 0;
  }
  if (phase != 2) return FBME_ERR_ARG;
  // Final consistency E-step (no prediction).
  FBME_TRY(col_adjoint(p->col_split, ncol16, 1, 0));
  if (o.polish_size == 0) {
    // No polish: the UAMP posterior of the final hyperparameters is the
    // result (solver.py:1907-1925, 2081-2126).
    FBME_TRY(row(p->row_split, nrow16, 3, nullptr, nullptr, 1));
    FBME_TRY(fbme_launch_scalars(scenes, p->nscenes, sc, part, ps, 3, nullptr, o.min_variance, o.noise_fraction,
                                 o.blend, o.damping, o.noise_shape, o.noise_rate, o.evidence_noise, o.fixed_noise,
                                 st));
    FBME_TRY(fbme_launch_output_uamp(scenes, tilew, ntile, ntile, sc, mean, var, gamma, io->x, io->posterior_variance,
                                     io->precision, io->support_mask, o.false_alarm,
                                     (io->support_mask && o.false_alarm > 0.f) ? 1 : 0, io->noise, io->dof, io->nle,
                                     st));
    return 0;
  }
  // per-tile top-k candidates of the polish selection
  FBME_TRY(row(p->row_split, nrow16, 3, ckey, cidx, io->store_uamp ? 1 : 0));
  // Polish (the un-pruned model first sums the variances for its tail loading).
  if (!o.prune_tail) FBME_TRY(fbme_launch_gamma_sum(scenes, p->nscenes, gamma, sc, st));
  FBME_TRY(fbme_launch_polish(scenes, p->nscenes, 0, bp, sc, dir, mh, ckey, cidx, xkey, xidx, states, results, io->noise,
                              io->dof, io->nle, io->trace, io->exchanges, io->diag, io->status, gamma, &o, st));
  const bool exch = o.exchange_size > 0 && p->total_xcand > 1;
  if (exch) {
    FBME_TRY(fbme_launch_exchange_score(scenes, tilew, ntile, ntile, bp, dir, mh, states, xkey, xidx, o.exchange_size,
                                        st));
    FBME_TRY(fbme_launch_polish(scenes, p->nscenes, 1, bp, sc, dir, mh, ckey, cidx, xkey, xidx, states, results, io->noise,
                                io->dof, io->nle, io->trace, io->exchanges, io->diag, io->status, gamma, &o, st));
  }
  FBME_TRY(fbme_launch_output(scenes, tilew, ntile, ntile, sc, results, io->x, io->posterior_variance, io->precision,
                              io->support_mask, o.false_alarm, (io->support_mask && o.false_alarm > 0.f) ? 1 : 0,
                              o.prune_tail ? nullptr : gamma, st));
  if (o.truncation_diagnostic) {
    // max |BTTB spectrum of the tail variances| / noise (solver.py:3945-3958)
    const FbmeScene& a = p->scenes[0];
    const long long n = (long long)a.n1 * a.n2, e = (long long)p->te1 * p->te2;
    float* tail = (float*)(w + p->r_ttail.off);
    void* wa = w + p->r_twa.off;
    void* wb = w + p->r_twb.off;
    const float2* twb = (const float2*)tw;
    FBME_TRY(fbme_launch_truncation(scenes, p->nscenes, n, e, gamma, states, sc, tail, nullptr, nullptr, 0, st));
    // the spectrum lands in wb (free again after the embedding pass)
    FBME_TRY(fbme_gamma_spectrum(st, p->nscenes, a.n1, a.n2, a.m1, a.m2, p->te1, p->te2, fbme_radix_code(a.n1), fbme_radix_code(a.n2),
                                 fbme_radix_code(p->te1), fbme_radix_code(p->te2), tail, nullptr, wa, wb, wb,
                                 twb + p->ttw_n2, twb + p->ttw_n2, twb + p->ttw_e2, twb + p->ttw_e2));
    FBME_TRY(fbme_launch_truncation(scenes, p->nscenes, n, e, gamma, states, sc, tail, wb, io->truncation, 1, st));
  }
#undef FBME_TRY
  return 0;
}

}  // namespace

int fbme_plan_run(const fbme_plan* p, const FbmeIO* io, float support_threshold, void* stream) {
  (void)support_threshold;
  hipStream_t st = (hipStream_t)stream;
  int rc = plan_phase(p, io, st, 0, 0, 0);
  for (int t = 1; !rc && t <= p->max_iters; ++t) rc = plan_phase(p, io, st, 1, t, 0);
  return rc ? rc : plan_phase(p, io, st, 2, 0, 0);
}

int fbme_plan_begin(const fbme_plan* p, const FbmeIO* io, void* stream) {
  return plan_phase(p, io, (hipStream_t)stream, 0, 0, 0);
}

int fbme_plan_iterate(const fbme_plan* p, int t, int track_change, const FbmeIO* io, void* stream) {
  return plan_phase(p, io, (hipStream_t)stream, 1, t, track_change ? 1 : 0);
}

int fbme_plan_finish(const fbme_plan* p, const FbmeIO* io, void* stream) {
  return plan_phase(p, io, (hipStream_t)stream, 2, 0, 0);
}

int fbme_plan_run_stage(const fbme_plan* p, int stage, const FbmeIO* io, void* stream) {
  if (!p || !io || !p->bound || !io->y) return FBME_ERR_ARG;
  if (p->masked && !p->mask) return FBME_ERR_ARG;
  const unsigned char* mask = p->masked ? p->mask : nullptr;
  hipStream_t st = (hipStream_t)stream;
  char* w = p->ws;
  const void* mbits = (p->masked && p->r_mbits.bytes) ? (const void*)(w + p->r_mbits.off) : nullptr;
  const FbmeScene* scenes = (const FbmeScene*)(w + p->r_scenes.off);
  const FbmeWork* colw = (const FbmeWork*)(w + p->r_col.off);
  const FbmeWork* roww = (const FbmeWork*)(w + p->r_row.off);
  const FbmeWork* tilew = (const FbmeWork*)(w + p->r_tile.off);
  double* sc = (double*)(w + p->r_sc.off);
  double* part = (double*)(w + p->r_part.off);
  const void* tw = w + p->r_tw.off;
  const FbmePlanOptions& o = p->opt;
  const int ps = p->total_slots;
  const int ncol = (int)p->col_work.size(), nrow = (int)p->row_work.size(), ntile = (int)p->tile_work.size();
  void* mid = w + p->r_mid.off;
  void* msg = w + p->r_msg.off;
  void* pred = w + p->r_pred.off;
  void* bp = w + p->r_bp.off;
  float* gamma = (float*)(w + p->r_gamma.off);
  void* mean = w + p->r_mean.off;
  float* var = (float*)(w + p->r_var.off);
  const FbmeWork* colw16 = colw + p->col_split;
  const FbmeWork* roww16 = roww + p->row_split;
  const int ng_col = p->col_split, n16_col = ncol - p->col_split;
  const int ng_row = p->row_split, n16_row = nrow - p->row_split;
  switch (stage) {
    case FBME_STAGE_COL_ADJOINT:
      return (fbme_launch_col_adjoint(scenes, colw, ng_col, ng_col, p->ept_col, io->y, mask, mbits, pred, msg, mid, sc,
                                      part, ps, 1, 1, o.uamp_damping, tw, st) ||
              fbme_launch_col_adjoint(scenes, colw16, n16_col, n16_col, FBME_EPT_R16, io->y, mask, mbits, pred, msg,
                                      mid, sc, part, ps, 1, 1, o.uamp_damping, tw, st))
                 ? FBME_ERR_HIP
                 : 0;
    case FBME_STAGE_ROW:
      return (fbme_launch_row(scenes, roww, ng_row, ng_row, p->ept_row, mid, bp, gamma, mean, var, sc, part, ps, 2,
                              o.uamp_damping, o.damping, o.max_precision, o.hyper_shape, o.hyper_rate, o.blend, nullptr,
                              nullptr, 0, nullptr, tw, st) ||
              fbme_launch_row(scenes, roww16, n16_row, n16_row, FBME_EPT_R16, mid, bp, gamma, mean, var, sc, part, ps, 2,
                              o.uamp_damping, o.damping, o.max_precision, o.hyper_shape, o.hyper_rate, o.blend, nullptr,
                              nullptr, 0, nullptr, tw, st))
                 ? FBME_ERR_HIP
                 : 0;
    case FBME_STAGE_COL_FORWARD:
      return (fbme_launch_col_forward(scenes, colw, ng_col, ng_col, p->ept_col, mid, io->y, mask, mbits, pred, msg, sc,
                                      part, ps, 0, o.uamp_damping, tw, st) ||
              fbme_launch_col_forward(scenes, colw16, n16_col, n16_col, FBME_EPT_R16, mid, io->y, mask, mbits, pred,
                                      msg, sc, part, ps, 0, o.uamp_damping, tw, st))
                 ? FBME_ERR_HIP
                 : 0;
    case FBME_STAGE_OUTPUT:
      return fbme_launch_output(scenes, tilew, ntile, ntile, sc, w + p->r_result.off, io->x, io->posterior_variance,
                                io->precision, io->support_mask, o.false_alarm,
                                (io->support_mask && o.false_alarm > 0.f) ? 1 : 0, o.prune_tail ? nullptr : gamma,
                                st) ? FBME_ERR_HIP : 0;
    default:
      return FBME_ERR_ARG;
  }
}

int fbme_plan_debug_pointer(const fbme_plan* p, int which, void** out) {
  if (!p || !out || !p->bound) return FBME_ERR_ARG;
  const Region* r = nullptr;
  switch (which) {
    case 0: r = &p->r_gamma; break;
    case 1: r = &p->r_mean; break;
    case 2: r = &p->r_var; break;
    case 3: r = &p->r_bp; break;
    case 4: r = &p->r_sc; break;
    case 5: r = &p->r_pred; break;
    case 6: r = &p->r_msg; break;
    default: return FBME_ERR_ARG;
  }
  *out = p->ws + r->off;
  return 0;
}

}  // extern "C"
