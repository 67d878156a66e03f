/* C ABI of the MI355X-native FBME hot path (libfbme_hip.so, gfx950).
 *
 * The reference has no FFI: its hot path is Python calling Triton kernels and
 * torch.fft.  Each entry point below names the reference interface it
 * replaces (file:line, relative to the reference repository root).  All
 * pointers are device pointers unless stated; every buffer is owned by the
 * caller; nothing here allocates device memory, synchronises (except
 * fbme_plan_bind) or issues memsets, so every call can be captured into a
 * hipGraph on the given stream.  Return value 0 means success.
 *
 * Complex data is interleaved complex64 (float2).  2-D arrays are row-major:
 * grid index n = n1 * N2 + n2, sample index m = m1 * M2 + m2.
 */
#ifndef FBME_HIP_H
#define FBME_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FBME_ABI_VERSION 9

#define FBME_ERR_ARG 1
#define FBME_ERR_HIP 2

/* Numerics of the locked plans: SBLConfig.paper_hybrid_uamp
 * (torch_sbl/solver.py:198-280) with the Fast overrides
 * (torch_sbl_fastest_stable/profiles.py:63-70) or Stable overrides
 * (torch_sbl_fastest_stable/profiles.py:110-115). */
typedef struct FbmePlanOptions {
  float uamp_damping;      /* 0.05 UAMP message/denoise damping            */
  float damping;           /* 0.15 ARD and noise damping inside the sweep  */
  float min_variance;      /* 1e-12 variance floor relative to data power  */
  float max_precision;     /* 1e12                                         */
  float blend;             /* 0.1 backprojection initialisation blend      */
  float noise_fraction;    /* 1e-2 initial noise fraction                  */
  float margin_factor;     /* 64 evidence-gate margin in float32 eps       */
  float jitter_multiplier; /* 8 Cholesky loading in float32 eps            */
  float false_alarm;       /* 0.5 support-mask budget (<= 0: no mask)      */
  int polish_size;         /* 32 active-set capacity (<= 32; 0: no polish) */
  int polish_iterations;   /* 2 Fast / 3 Stable evaluations                */
  int exchange_size;       /* 0 Fast / 12 Stable residual exchange         */
  int diagnostics;         /* Stable fixed-budget diagnostics              */
  int evidence_noise;      /* 0 EM noise update, 1 evidence form           */
  int row_lines;           /* rows per row-pass work item (tuning)         */
  int col_lines;           /* columns per column-pass work item (tuning)   */
  float iteration_offset;  /* 0   T_u = clamp(ceil(off + scale N/M),       */
  float iteration_scale;   /* 0.8       min_iters, max_iters)              */
  int min_iters;           /* 3   max(min_iter, uamp_min_iter)             */
  int max_iters;           /* 200 max_iter                                 */
  float hyper_shape;       /* 0 ARD hyperprior a                           */
  float hyper_rate;        /* 0 ARD hyperprior b                           */
  float noise_shape;       /* 0 noise hyperprior c                         */
  float noise_rate;        /* 0 noise hyperprior d                         */
  float polish_damping;    /* 0 damping of the polish EM updates           */
  int fft_kernel;          /* 0 auto, 1 LDS Stockham passes, 2 register
                              16 x 16 passes (requires N1 or N2 = 256)     */
  int truncation_diagnostic; /* 0: off; 1: FbmeIO.truncation gets the
                              hard-pruning certificate of
                              _active_truncation_delta_bound
                              (solver.py:3927-3958, 4034-4050); needs one
                              grid and sample shape for every scene        */
  int prune_tail;          /* 1 zero-tail reduced model (locked plans); 0
                              diffuse tail loading nu = noise + sum of the
                              tail variances (solver.py:4087-4093,
                              3627-3633, 4342-4352)                        */
  int reinitialize;        /* 1 gamma_S from the UAMP second moments
                              (solver.py:4094-4106); 0 keep the UAMP gamma */
  int evidence_gate;       /* 1 keep the best evidence (solver.py:4168-4298);
                              0 take every evaluation                      */
  float fixed_noise;       /* 0: noise estimated; > 0: fixed_noise_variance
                              (solver.py:1151-1156, 1747-1750, 4358-4361,
                              3457-3458): sigma^2_0 = fixed, every update
                              max(fixed, floor), no noise residual         */
} FbmePlanOptions;

/* Outputs of one plan replay; per-scene arrays are in the caller's scene
 * order.  Mirrors the SBLResult fields the locked plans fill
 * (torch_sbl/solver.py:720-792, 2154-2393). */
typedef struct FbmeIO {
  const void* y;                /* [total_samples] complex64 observations       */
  void* x;                      /* [total_grid] complex64 reconstruction        */
  float* posterior_variance;    /* [total_grid]                                 */
  float* precision;             /* [total_grid]                                 */
  unsigned char* support_mask;  /* [total_grid] bool, optional                  */
  float* noise;                 /* [nscenes] noise variance                     */
  float* dof;                   /* [nscenes] effective dof                      */
  float* nle;                   /* [nscenes] negative log evidence              */
  float* trace;                 /* [nscenes, evaluations] accepted evidence     */
  long long* exchanges;         /* [nscenes] accepted residual exchanges, opt.  */
  float* diag;                  /* [nscenes, 4] gamma res., noise res.,
                                   objective decrease, monotone flag, optional */
  int* status;                  /* [nscenes] non-zero on Cholesky breakdown     */
  int store_uamp;               /* keep the final UAMP E-step in the workspace  */
  float* truncation;            /* [nscenes] truncation_delta_bound, required
                                   when options.truncation_diagnostic is set   */
} FbmeIO;

typedef struct fbme_plan fbme_plan;

enum {
  FBME_Q_WORKSPACE_BYTES = 0,
  FBME_Q_TOTAL_GRID = 1,
  FBME_Q_TOTAL_SAMPLES = 2,
  FBME_Q_MAX_ITERS = 3,
  FBME_Q_NSCENES = 4,
  FBME_Q_EVALUATIONS = 5,
  FBME_Q_LAUNCHES = 6,
  FBME_Q_R16_COL_ITEMS = 7,   /* column work items on the register 16 x 16 passes */
  FBME_Q_R16_ROW_ITEMS = 8    /* row work items on the register 16 x 16 passes    */
};

int fbme_abi_version(void);

/* Plain HBM streams, one pass with `unroll` (1, 2, 4 or 8) 16-byte
 * accesses in flight per thread: the ceiling the kernels' bandwidth is
 * compared with, and the known-size kernels that calibrate hardware byte
 * counters.  No counterpart in the reference (measurement only).  bytes: a
 * multiple of 16; pointers 16-byte aligned; copy buffers disjoint. */
int fbme_stream_copy(void* stream, const void* src, void* dst, long long bytes, int unroll);
int fbme_stream_fill(void* stream, void* dst, long long bytes, float value, int unroll);

/* Defaults of the Fast (stable = 0) or Stable (stable = 1) profile. */
void fbme_default_options(FbmePlanOptions* options, int stable);

/* Build a plan for `nscenes` scenes, shapes[4*i .. 4*i+3] = N1, N2, M1, M2.
 * Replaces the shape-fixing part of CUDAGraphSBLPlan.__init__
 * (torch_sbl/cuda_graph.py:22-62) and the operator cache
 * (torch_sbl/solver.py:844-857).  Uneven shapes are allowed.  On failure
 * returns NULL and writes a message into err. */
fbme_plan* fbme_plan_create(int nscenes, const int* shapes, const FbmePlanOptions* options, char* err,
                            int errlen);
/* fbme_plan_create with observation masks: observed[i] is the number of
 * samples scene i keeps (M1 M2 for an unmasked scene).  mask_group[i] (or
 * NULL: every scene its own) names this or an earlier scene of the same
 * shape whose mask scene i repeats, so the two share one mask spectrum.
 * The UAMP budget uses N / observed (solver.py:1296-1309 on the restricted
 * operator); the mask bytes themselves are attached after fbme_plan_bind
 * by fbme_plan_set_mask.  The reference builds the restricted operator per
 * call (solver.py:986-1024, operators.py:57-75). */
fbme_plan* fbme_plan_create_masked(int nscenes, const int* shapes, const int* observed, const int* mask_group,
                                   const FbmePlanOptions* options, char* err, int errlen);
/* Attach the observation masks of a masked plan: `mask` is a device buffer
 * of FBME_Q_TOTAL_SAMPLES bytes packed like the data (nonzero = observed),
 * read by every later run, so it must stay allocated and unchanged while
 * the plan runs.  Also computes each masked scene's mask spectrum into the
 * workspace (enqueued on `stream`; call before graph capture). */
int fbme_plan_set_mask(fbme_plan* plan, const unsigned char* mask, void* stream);
void fbme_plan_destroy(fbme_plan* plan);
long long fbme_plan_query(const fbme_plan* plan, int what);
/* info[0..7] = grid offset, sample offset, T_u, K, radix codes (2),
 * column items, row items of caller scene `scene`. */
int fbme_plan_scene_info(const fbme_plan* plan, int scene, long long* info);
/* Attach a caller-allocated workspace of FBME_Q_WORKSPACE_BYTES and upload
 * the plan tables (synchronises `stream`; call before graph capture). */
int fbme_plan_bind(fbme_plan* plan, void* workspace, void* stream);
/* Enqueue one complete solve: the captured body of fast_sbl under the locked
 * presets (torch_sbl/solver.py:925-2393 via cuda_graph.py:76-83).
 * `support_threshold` is unused (the mask threshold log(N / 0.5) is
 * evaluated per scene, solver.py:2108-2126). */
int fbme_plan_run(const fbme_plan* plan, const FbmeIO* io, float support_threshold, void* stream);
/* fbme_plan_run in three calls, for callers that decide between iterations
 * (the eager convergence checks of solver.py:1800-1904): fbme_plan_begin
 * (initialisation), fbme_plan_iterate for t = 1, 2, ... (UAMP iteration t
 * over the scenes whose budget reaches t; with track_change the step also
 * leaves ||mu_t - mu_t-1|| / max(||mu_t||, eps) per scene in the plan's
 * scalars, slot 8 of 9 doubles per scene, fbme_plan_debug_pointer 4), and
 * fbme_plan_finish (final E-step, polish or, with polish_size 0, the UAMP
 * posterior, and outputs).  Stopping before the budget finishes every
 * scene at that iteration. */
int fbme_plan_begin(const fbme_plan* plan, const FbmeIO* io, void* stream);
int fbme_plan_iterate(const fbme_plan* plan, int t, int track_change, const FbmeIO* io, void* stream);
int fbme_plan_finish(const fbme_plan* plan, const FbmeIO* io, void* stream);
/* Enqueue one pass of the plan on its current state, for per-kernel timing
 * (no reference counterpart; the reference's passes are separate Triton and
 * cuFFT launches inside one captured graph, torch_sbl/solver.py:2726-2870).
 * Valid after one fbme_plan_run. */
enum {
  FBME_STAGE_COL_ADJOINT = 0, /* message update + column pass of A^H */
  FBME_STAGE_ROW = 1,         /* row pass: A^H, denoise + ARD, A     */
  FBME_STAGE_COL_FORWARD = 2, /* column pass of A + residual          */
  FBME_STAGE_OUTPUT = 3       /* full-grid outputs                    */
};
int fbme_plan_run_stage(const fbme_plan* plan, int stage, const FbmeIO* io, void* stream);
/* Device pointers of internal state for tests: 0 gamma, 1 mean, 2 variance,
 * 3 backprojection, 4 per-scene scalars (double[8]), 5 residual y - A mu,
 * 6 message. */
int fbme_plan_debug_pointer(const fbme_plan* plan, int which, void** out);

/* Prefix DFT operator on a uniform batch.
 * fbme_prefix_forward replaces PrefixFourierOperator.forward_flat
 * (torch_sbl/operators.py:237-253): y[b] = fft2(x[b])[:M1, :M2].
 * fbme_prefix_adjoint replaces PrefixFourierOperator.adjoint_flat
 * (torch_sbl/operators.py:255-280): x[b] = N * ifft2(pad(v[b]), s=(N1, N2)).
 * `mid` is a workspace of batch * N1 * M2 complex64. */
int fbme_prefix_forward(void* stream, int batch, int n1, int n2, int m1, int m2, const void* x, void* y,
                        void* mid);
int fbme_prefix_adjoint(void* stream, int batch, int n1, int n2, int m1, int m2, const void* v, void* x,
                        void* mid);

/* Elementwise UAMP kernels, one per Triton wrapper of
 * torch_sbl/triton_kernels.py.  `partial` workspaces hold
 * batch * ceil(len / 256) doubles and make every reduction deterministic. */
/* uamp_message_update (triton_kernels.py:476-547) */
int fbme_uamp_message_update(void* stream, int batch, int m, const void* y, const void* prediction,
                             const void* old_message, const float* data_variance, const float* noise_variance,
                             void* message, float* pseudo_variance, float damping);
/* uamp_denoise (triton_kernels.py:550-667); prior_floor may be NULL (no ARD
 * update, output_prior unused). */
int fbme_uamp_denoise(void* stream, int batch, int n, const float* prior_variance, const float* pseudo_variance,
                      const void* lifted_message, const void* old_mean, const float* old_variance, void* mean,
                      float* variance, float* effective_dof, float* variance_sum, const float* prior_floor,
                      float* output_prior, float damping, float tiny, float hyper_shape, float hyper_rate,
                      float prior_damping, float prior_maximum, double* partial);
/* uamp_prior_update (triton_kernels.py:670-736) */
int fbme_uamp_prior_update(void* stream, int batch, int n, const void* mean, const float* posterior_variance,
                           const float* old_prior, const float* prior_floor, float* output, float hyper_shape,
                           float hyper_rate, float damping, float maximum);
/* complex_residual_energy (triton_kernels.py:739-803) */
int fbme_complex_residual_energy(void* stream, int batch, int m, const void* data, const void* prediction,
                                 float* output, double* partial);
/* uamp_noise_update (triton_kernels.py:806-904) */
int fbme_uamp_noise_update(void* stream, int batch, int m, const void* data, const void* prediction,
                           const float* old_noise, const float* effective_dof, const float* variance_floor,
                           float* output, float noise_shape, float noise_rate, float damping, int evidence_update,
                           float epsilon, double* partial);

/* ------------------------------------------------------------------------
 * General solver, mode="cg": covariance-free PCG + Hutchinson
 * (torch_sbl/solver.py:1315-1904 and 2023-2080).  Vectors of the 1 + P
 * right-hand sides of scene b are lanes b * R + r, each M complex64
 * contiguous.  Workspace sizes are stated per call; every pointer is device
 * memory unless noted; all calls are asynchronous on `stream`.
 * ------------------------------------------------------------------------ */

/* Radix digit code of an FFT length (factors 2, 3, 4, 5, 8), or 0xFFFFFFFF
 * when the length has another prime factor or exceeds 2048.  Host only. */
unsigned fbme_radix_code(int n);
/* Column FFT of [batch, rows_in (zero padded to n), cols] -> first rows_out
 * rows of the length-n (inverse when `inverse`) transform, n <= 2048. */
int fbme_column_fft(void* stream, int batch, int n, int rows_in, int rows_out, int cols, const void* in, void* out,
                    int inverse);
/* Host table out[2 * n] (double re, im): D[d] = sum_{a < m} exp(2 pi i a d / n),
 * the Dirichlet kernel of the Gram closed form (operators.py:444-467 columns,
 * Gram of the prefix rows). */
int fbme_dirichlet_table(int m, int n, double* out);
/* exp(-2 pi i k / len), k < len <= 2048, complex64 (FFT twiddles). */
int fbme_twiddle_table(void* stream, int len, void* out);
/* BTTB covariance spectrum of Q = noise I + A Gamma A^H on the circulant
 * embedding E1 x E2 (operators.py:299-326), N1, N2 <= 1024 and
 * 2 M - 1 <= E <= 2048 per axis.  work_a, work_b: batch * max(N1 N2, E1 E2)
 * complex64 each; spectrum: batch * E1 * E2. */
int fbme_gamma_spectrum(void* stream, int batch, int n1, int n2, int m1, int m2, int e1, int e2, unsigned code1,
                        unsigned code2, unsigned codee1, unsigned codee2, const float* gamma, const float* noise,
                        void* work_a, void* work_b, void* spectrum, const void* tw_n1, const void* tw_n2,
                        const void* tw_e1, const void* tw_e2);
/* Strang circulant (BCCB) preconditioner of Q on the M1 x M2 sample grid
 * (circulant_preconditioner_spectrum / apply_circulant_preconditioner,
 * operators.py:373-416; cg_preconditioner="circulant",
 * solver.py:5738-5744, 5779-5780).  eig [batch, M1 M2] float: the column of
 * signed lags (k <= M / 2 ? k : k - M) of fft2(gamma), transformed, plus
 * noise, floored at max(eps_f32 max|eig|, noise).  work_a, work_b: batch *
 * max(N1 N2, M1 M2) complex64 each.  apply: out = ifft2(fft2(v) / eig) for
 * `lanes` lanes, `rhs` per scene; work: lanes * M1 M2 complex64; v != out. */
int fbme_circulant_spectrum(void* stream, int batch, int n1, int n2, int m1, int m2, unsigned code_n2,
                            unsigned code_m2, const float* gamma, const float* noise, void* work_a, void* work_b,
                            float* eig, const void* tw_n2, const void* tw_m2);
int fbme_circulant_apply(void* stream, int lanes, int rhs, int m1, int m2, unsigned code_m2, const float* eig,
                         const void* v, void* out, void* work, const void* tw_m2);
/* out = Q d for `lanes` lanes, `rhs` lanes per scene (operators.py:328-355),
 * E1, E2 <= 2048.  work_a, work_b: lanes * M1 * E2 complex64 each. */
int fbme_toeplitz_matvec(void* stream, int lanes, int rhs, int m1, int m2, int e1, int e2, unsigned code_e1,
                         unsigned code_e2, const void* spectrum, const void* d, void* out, void* work_a,
                         void* work_b, const void* tw_e1, const void* tw_e2);
/* fbme_toeplitz_matvec with explicit workgroup tiles (same arithmetic, same
 * results): row_tile rows per workgroup of the two E2-point row passes,
 * col_tile columns per workgroup of the E1-point column pass; 0 takes the
 * largest tile LDS holds, larger values are capped to it.  Tuning hook of
 * the same operator (operators.py:328-355); fbme_toeplitz_tiles reports the
 * caps. */
int fbme_toeplitz_matvec_tiled(void* stream, int lanes, int rhs, int m1, int m2, int e1, int e2, unsigned code_e1,
                               unsigned code_e2, const void* spectrum, const void* d, void* out, void* work_a,
                               void* work_b, const void* tw_e1, const void* tw_e2, int row_tile, int col_tile);
int fbme_toeplitz_tiles(int e1, int e2, int* row_tile_max, int* col_tile_max);
/* out = noise d + A Gamma A^H d for 256 x 256 grids: the same covariance
 * matvec (operators.py:328-355) evaluated through the grid with three
 * register 256-point FFT passes instead of the BTTB embedding.  lanes =
 * batch * rhs; d, out [lanes, M1, M2] (distinct); gamma [batch, 65536];
 * noise [batch]; mid: lanes * 256 * M2 complex64; tw256:
 * fbme_twiddle_table(256).  Requires N1 = N2 = 256 and M1, M2 <= 256. */
int fbme_grid_matvec(void* stream, int lanes, int rhs, int m1, int m2, const float* gamma, const float* noise,
                     const void* d, void* out, void* mid, const void* tw256);
/* fbme_grid_matvec that also leaves dot[lane] = Re <d, out> (FP64), the PCG
 * curvature of solver.py:5645-5655, from the last pass's registers.  part:
 * lanes * ceil(M2 / 16) doubles. */
int fbme_grid_matvec_dot(void* stream, int lanes, int rhs, int m1, int m2, const float* gamma, const float* noise,
                         const void* d, void* out, void* mid, const void* tw256, double* part, double* dot);
/* PCG update and direction that take their inner product as the per-block
 * partials of the producing pass (fbme_grid_matvec_dot /
 * fbme_lowrank_apply_fft_dot with part but dot = NULL: no reduction launch)
 * and fold the validity commit into the rho commit; same lane logic as
 * fbme_pcg_update + fbme_pcg_direction (solver.py:5645-5719).
 * fbme_pcg_update_parts stores curv[lane]; fbme_pcg_direction_parts stores
 * nrho[lane] and commits rho / active (restart = 0).  fbme_pcg_commit_valid
 * is the validity commit alone, for a convergence check after
 * fbme_pcg_update_parts. */
int fbme_pcg_update_parts(void* stream, int lanes, int m, void* x, void* r, const void* d, const void* q,
                          const double* rho, const double* part, int nparts, double* curv, const int* active);
int fbme_pcg_direction_parts(void* stream, int lanes, int m, void* d, const void* z, double* rho,
                             const double* curv, const double* part, int nparts, double* nrho, int* active);
/* fbme_pcg_update_parts that also leaves rr_part[lane, block] = ||r||^2 over
 * each 256-sample block after the update (FP64), the input of
 * fbme_lowrank_apply_fft_direction.  rr_part: lanes * ceil(m / 256) doubles. */
int fbme_pcg_update_parts_rr(void* stream, int lanes, int m, void* x, void* r, const void* d, const void* q,
                             const double* rho, const double* part, int nparts, double* curv, const int* active,
                             double* rr_part);
int fbme_pcg_commit_valid(void* stream, int lanes, const double* rho, const double* curv, int* active);
/* 32 largest values of each x[b, :n] sorted descending, ties to the lower
 * index (the stable argsort of solver.py:5797-5800).  cand: scratch of
 * fbme_topk32_scratch_bytes(batch, n) bytes (per-tile candidates, then one
 * level of merged candidates per 64x reduction); 32 <= n <= 2^26. */
size_t fbme_topk32_scratch_bytes(int batch, long long n);
int fbme_topk32(void* stream, int batch, long long n, const float* x, void* cand, int* idx, float* val);
/* k <= 256 largest values of each x[b, :n] sorted descending, ties to the
 * lower index: the active-set selection of the general solver
 * (torch.topk in _select_active_indices, solver.py:3184-3216).  cand:
 * fbme_topk_scratch_bytes(batch, n, k) bytes; k <= n <= 2^26. */
size_t fbme_topk_scratch_bytes(int batch, long long n, int k);
int fbme_topk(void* stream, int batch, long long n, int k, const float* x, void* cand, int* idx, float* val);
/* fbme_topk32 without a workspace argument and without any allocation: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The candidates live in a fixed per-GPU device array of
 * FBME_TOPK32_NOALLOC_SCRATCH_BYTES, held until the process ends; inputs
 * needing more (fbme_topk32_scratch_bytes(batch, n) larger) return
 * FBME_ERR_ARG.  All calls on one GPU share that array: calls that overlap on
 * one GPU (concurrent streams or threads) race and must be serialised by the
 * caller.  The first call on a device resolves the array's address there
 * (FBME_ERR_HIP if that fails). */
#define FBME_TOPK32_NOALLOC_SCRATCH_BYTES (4u << 20)
int fbme_topk32_noalloc(void* stream, int batch, long long n, const float* x, int* idx, float* val);
/* Low-rank preconditioner (_make_low_rank_preconditioner,
 * solver.py:5786-5833).  The factor matrix F = columns * sqrt(gamma_k)
 * [M, 32] is separable and kept as its two phase tables: factors [batch,
 * M1 + M2, 32] complex64 = U [M1, 32] (row phases) then V [M2, 32] (column
 * phases times sqrt(gamma_k)), F[m1 M2 + m2, k] = U[m1, k] V[m2, k].
 * sinv [batch, 32, 32] = (I + F^H F / bg)^-1, bg = max(noise + omitted,
 * eps * sum gamma).  dir1, dir2: fbme_dirichlet_table(M1, N1), (M2, N2) on
 * the device.  status[b] = 1 when the Cholesky factorisation fails. */
int fbme_lowrank_setup(void* stream, int batch, int n1, int n2, int m1, int m2, const int* idx, const float* val,
                       const double* gamma_sum, const float* noise, const void* dir1, const void* dir2,
                       void* factors, void* sinv, float* bg, int* status);
/* out = v / bg - F S^-1 F^H v / bg^2 (_apply_low_rank_preconditioner,
 * solver.py:5836-5848), F generated from the phase tables of
 * fbme_lowrank_setup once per sample for all `rhs` lanes of a scene.
 * part: lanes * 64 * 64 doubles; corr: lanes * 32 complex64. */
int fbme_lowrank_apply(void* stream, int lanes, int rhs, int m1, int m2, const void* factors, const void* sinv,
                       const float* bg, const void* v, void* out, void* part, void* corr);
/* fbme_lowrank_apply that also leaves dot[lane] = Re <v, out> (FP64), the
 * PCG inner product <r, M^-1 r> of solver.py:5700-5708.  dot_part: lanes *
 * ceil(M1 M2 / 256) doubles. */
int fbme_lowrank_apply_dot(void* stream, int lanes, int rhs, int m1, int m2, const void* factors, const void* sinv,
                           const float* bg, const void* v, void* out, void* part, void* corr, double* dot_part,
                           double* dot);
/* out = vbase / bg - F S^-1 F^H vproj / bg^2: the low-rank apply with the
 * projection taken from a second vector (vproj = vbase is
 * fbme_lowrank_apply).  The grid-diagonal preconditioner below uses it with
 * bg = 1 and S^-1 = C^-1 to form r - F C^-1 F^H B^-1 r (Woodbury form of
 * (B + F F^H)^-1, the structure of _make_low_rank_preconditioner,
 * solver.py:5786-5848, with its isotropic background replaced).  part, corr
 * as fbme_lowrank_apply. */
int fbme_lowrank_apply_split(void* stream, int lanes, int rhs, int m1, int m2, const void* factors,
                             const void* sinv, const float* bg, const void* vproj, const void* vbase, void* out,
                             void* part, void* corr);
/* out = v - U S^-1 U^H v with U [batch, 32, M] complex64 stored columns
 * (lanes of `rhs` per scene share them), S^-1 [batch, 32, 32] complex64 and
 * ones: batch floats 1.0 (the Woodbury apply of fbme_lowrank_apply_split
 * with stored columns).  With dot, dot[lane] = Re <v, out> (FP64).  part:
 * lanes * ceil(M / 1024) * 32 double2; corr: lanes * 32 complex64; dot_part:
 * lanes * ceil(M / 256) doubles.  v != out. */
int fbme_stored_apply(void* stream, int lanes, int rhs, long long m, const void* ucols, const void* cinv,
                      const float* ones, const void* v, void* out, void* part, void* corr, double* dot_part,
                      double* dot);
/* Randomised Nystrom preconditioner (cg_preconditioner="nystrom",
 * _make_nystrom_preconditioner / _apply_nystrom_preconditioner,
 * solver.py:5851-5976), rank r <= 32, ri = max(1, r / 2) informed atoms.
 * fbme_nystrom_omega: the test matrix Omega [batch, 32, M1 M2] complex64:
 * columns k < ri the atoms idx[b, k] (fbme_topk32 order), ri <= k < r the
 * Gaussian vectors rnd [batch, r - ri, M] (NULL when r = ri), the rest 0.
 * The caller then forms Y = A Gamma A^H Omega (the matvec with a zero
 * noise).  fbme_nystrom_setup: from the r x r Grams of Omega and Y (FP64;
 * part: fbme_nystrom_part_bytes(batch)) the eigenpairs of the Nystrom
 * approximation Y (Omega^H Y)^-1 Y^H (orthonormalised test matrix,
 * 10 eps_f32 mean|diag| Cholesky loading, complex Jacobi): z [batch, 32, 32]
 * complex64 (U = Y z), cinv = diag(1 - w) [batch, 32, 32] complex64 with
 * w = (lambda_min + bg) / (bg + lambda), lambda [batch, 32] (descending,
 * zero past r), bg [batch] = max(noise + max(gamma_sum - sum lambda / M, 0),
 * tiny).  A failed factorisation leaves z = cinv = 0 (the apply is then the
 * identity, the reference's scalar fallback up to a scale PCG ignores).
 * fbme_nystrom_vectors: U = Y z [batch, 32, M] for fbme_stored_apply with
 * cinv, whose out = v + U ((w - 1) o U^H v) is the reference's apply. */
size_t fbme_nystrom_part_bytes(int batch);
int fbme_nystrom_omega(void* stream, int batch, int n1, int n2, int m1, int m2, int rank, const int* idx,
                       const void* rnd, void* omega);
int fbme_nystrom_setup(void* stream, int batch, long long m, int rank, const void* omega, const void* y,
                       const double* gamma_sum, const float* noise, void* part, void* z, void* cinv, float* lambda,
                       float* bg);
int fbme_nystrom_vectors(void* stream, int batch, long long m, int rank, const void* y, const void* z, void* u);
/* Grid-diagonal background of the low-rank preconditioner (no counterpart in
 * the reference, whose background is the scalar noise + omitted variance of
 * solver.py:5813-5819): B^-1 = A diag(w) A^H, Hermitian positive definite for
 * any w > 0, with w[b, n] = exp(clamp(o)) / (N (noise_b + N gamma_T[b, n])),
 * gamma_T = gamma without the 32 exact atoms idx [batch, 32] (from
 * fbme_topk32).  logscale o [batch, n] (NULL: 0) is the learned per-cell
 * correction, clamped to +-clamp.  B^-1 v is a matvec with (gamma, noise) =
 * (w, 0): fbme_grid_matvec on 256 x 256 grids, else fbme_toeplitz_matvec on
 * the spectrum fbme_gamma_spectrum forms from w. */
int fbme_gdiag_weights(void* stream, int batch, long long n, const float* gamma, const float* noise, const int* idx,
                       const float* logscale, float clamp, float* w);
/* The 32 factor columns of fbme_lowrank_setup as lanes: out [batch, 32, M1 M2]
 * complex64, out[b, k, m] = exp(-2 pi i (m1 k1 / N1 + m2 k2 / N2)) sqrt(gamma_k). */
int fbme_lowrank_columns(void* stream, int batch, int m1, int m2, const void* factors, void* out);
/* C^-1 [batch, 32, 32] complex64 of C = I + F^H B^-1 F, bf = B^-1 F
 * [batch, 32, M1 M2] (fbme_lowrank_columns through the B^-1 matvec); FP64
 * sums and Cholesky inverse, status[b] = 1 on a non-positive pivot.
 * part: batch * 32 * 64 * 32 double2. */
int fbme_gdiag_system(void* stream, int batch, int m1, int m2, const void* factors, const void* bf, void* part,
                      void* cinv, int* status);
/* out = (B + F F^H)^-1 v = b - U C^-1 F^H b, given b = B^-1 v (the B^-1
 * matvec), bf = U = B^-1 F and cinv from fbme_gdiag_system; ones: batch
 * floats 1.0.  With dot, dot[lane] = Re <v, out> (FP64).  part, corr,
 * dot_part as fbme_lowrank_apply_dot. */
int fbme_gdiag_apply(void* stream, int lanes, int rhs, int m1, int m2, const void* factors, const void* cinv,
                     const float* ones, const void* bf, const void* b, const void* v, void* out, void* part,
                     void* corr, double* dot_part, double* dot);
/* The same preconditioner for N1 = 256 grids through the 256-point column
 * transform: the projection is the column IFFT of v restricted to the
 * atoms' distinct rows, reduced onto the 32 atoms in the same pass; the
 * correction is a row synthesis on those rows, then one column FFT (same
 * operator as fbme_lowrank_apply, solver.py:5836-5848, with about a quarter
 * of the flops).  fbme_lowrank_map builds map [batch, 386] ints from the
 * atoms of fbme_topk32 (slot of each atom, row of each slot, slot count,
 * slot of each grid row, atoms ordered by slot, first position of each
 * slot).  wbuf: unused (may be NULL); part: lanes * ceil(M2 / 16) * 32
 * double2;
 * dot_part: lanes * ceil(M2 / 16) doubles; tw256: fbme_twiddle_table(256). */
int fbme_lowrank_map(void* stream, int batch, int n1, int n2, const int* idx, int* map);
int fbme_lowrank_apply_fft_dot(void* stream, int lanes, int rhs, int m1, int m2, const void* factors,
                               const void* sinv, const float* bg, const int* map, const void* v, void* out,
                               void* wbuf, void* part, void* corr, const void* tw256, double* dot_part,
                               double* dot);
/* The second half of a PCG iteration on N1 = 256 grids without storing z
 * (replaces fbme_lowrank_apply_fft_dot + fbme_pcg_direction_parts,
 * solver.py:5708-5720): nrho = <r, z> = ||r||^2 / bg - Re(proj^H corr) from
 * rr_part (fbme_pcg_update_parts_rr, rr_blocks per lane) and the projection,
 * then the synthesis pass writes d = z + (nrho / rho) d in place (0 for a
 * lane whose update is not valid), then rho / active are committed as
 * fbme_pcg_direction_parts does.  r and d distinct; part and corr as in
 * fbme_lowrank_apply_fft_dot. */
int fbme_lowrank_apply_fft_direction(void* stream, int lanes, int rhs, int m1, int m2, const void* factors,
                                     const void* sinv, const float* bg, const int* map, const void* r, void* d,
                                     void* part, void* corr, const void* tw256, const double* rr_part,
                                     int rr_blocks, double* rho, const double* curv, double* nrho, int* active);
/* out[l] = sum Re(conj(a_l) b_l) in FP64 (the PCG inner products of
 * _pcg, solver.py:5582-5723).  part: lanes * 64 doubles. */
int fbme_lane_dot(void* stream, int lanes, int m, const void* a, const void* b, double* part, double* out);
/* _pcg step (solver.py:5644-5662): valid = active & finite & curv, rho >
 * tiny; x += alpha d, r -= alpha q; active = valid. */
int fbme_pcg_update(void* stream, int lanes, int m, void* x, void* r, const void* d, const void* q,
                    const double* rho, const double* curv, int* active);
/* _pcg direction (solver.py:5692-5719): d = z + beta d (restart: d = z),
 * rho = nrho and active = active & nrho > tiny. */
int fbme_pcg_direction(void* stream, int lanes, int m, void* d, const void* z, double* rho, const double* nrho,
                       int* active, int restart);
/* out = a - b over n complex64 (true residual, solver.py:5663-5671). */
int fbme_sub(void* stream, long long n, const void* a, const void* b, void* out);
/* _pcg residual test (solver.py:5672-5690): rel = sqrt(rr) / max(sqrt(init),
 * eps); active = rel > tol (fixed budget: active &= rel > tol and the best
 * relative residual is tracked in best_rel / improved). */
int fbme_pcg_check(void* stream, int lanes, const double* rr, const double* init, int* active, double* best_rel,
                   int* improved, float tol, int fixed);
/* dst_l = src_l where mask[l] (best iterate of a fixed budget). */
int fbme_masked_copy(void* stream, int lanes, int m, void* dst, const void* src, const int* mask);
/* Hutchinson posterior (_posterior cg branch, solver.py:2991-3036):
 * dd = max(mean_p Re(lifted_p conj(A^H p)), 0), mean = gamma lifted_0,
 * var = max(gamma - gamma^2 dd, 0), rel_sum[b] = sum clamp(gamma dd, 0, 1).
 * part: batch * ceil(N / 256) doubles. */
int fbme_hutchinson(void* stream, int batch, int n, int probes, const void* lifted, const void* lprobes,
                    const float* gamma, void* mean, float* var, float* dd, double* part, double* rel_sum);
/* fbme_hutchinson averaging over the first `used` of the `probes` probe
 * lanes (lifted keeps 1 + probes lanes per scene): the adaptive probe
 * counts of _cg_schedule (solver.py:6086-6118).  1 <= used <= probes. */
int fbme_hutchinson_used(void* stream, int batch, int n, int probes, int used, const void* lifted,
                         const void* lprobes, const float* gamma, void* mean, float* var, float* dd, double* part,
                         double* rel_sum);
/* Pattern-coupled prior with a fixed coupling (SBLConfig.spatial_coupling,
 * solver.py:1100-1150, 1489-1590, 2396-2720) on [batch, N1 x N2] fields:
 * precision eta = pattern + coupling * (sum of the four axial neighbours of
 * pattern; circular != 0 wraps, else zero outside), gamma = max(1 /
 * clamp(eta, 1 / max_prec, max_prec), floor[b]).  fbme_spatial_init forms
 * pattern = (1 / gamma) / (1 + coupling * degree) from the initial gamma,
 * then gamma.  fbme_spatial_update replaces the MacKay update: second =
 * |mean|^2 + var (scratch [batch, N]), pattern <- keep pattern + (1 - keep)
 * clamp((1 + hyper_shape) / (second + coupling N(second) + hyper_rate)),
 * then gamma. */
int fbme_spatial_init(void* stream, int batch, int n1, int n2, float coupling, int circular, float max_prec,
                      const float* floor_, float* gamma, float* pattern);
int fbme_spatial_update(void* stream, int batch, int n1, int n2, float coupling, int circular, float hyper_shape,
                        float hyper_rate, float max_prec, float keep, const void* mean, const float* var,
                        const float* floor_, float* second, float* pattern, float* gamma);
/* Relevance projection (_project_relevance, solver.py:3156-3181), MacKay
 * update (1669-1677) and damping with floor (1776-1798); mask (may be NULL)
 * keeps gamma outside an active set.  dof[b] = projected relevance sum. */
int fbme_cg_em(void* stream, int batch, int n, int m, float* gamma, const void* mean, const float* dd,
               const double* rel_sum, const float* floor_, const unsigned char* mask, float hyper_shape,
               float hyper_rate, float max_prec, float keep, int update, double* part, double* dof,
               float* dof_f);
/* Initial power, floor, noise and backprojection gamma
 * (solver.py:1063-1089, 1151-1156); ye = ||y||^2, bpe = ||A^H y||^2. */
int fbme_cg_init(void* stream, int batch, int n, int m, const void* bp, const double* ye, const double* bpe,
                 float min_variance, float noise_fraction, float blend, float* power, float* floor_, float* noise,
                 float* gamma);
/* out[b] = sum clamp(gamma dd, 0, 1) (solver.py:3156-3164). */
int fbme_relevance_sum(void* stream, int batch, int n, const float* gamma, const float* dd, double* part,
                       double* out);
/* out = 1 / in (precision output, solver.py:2089). */
int fbme_reciprocal(void* stream, long long n, const float* in, float* out);
/* out[b] = (sum gamma, ||mean - prev||^2, ||mean||^2) in FP64 (relative
 * change, solver.py:1813-1826); mean and prev may be NULL. */
int fbme_grid_sums(void* stream, int batch, int n, const float* gamma, const void* mean, const void* prev,
                   double* part, double* out);
/* Reduced active-set posterior with diffuse-tail loading nu = noise +
 * omitted variance (_active_posterior, solver.py:3897-3924 with
 * 3591-3779), K <= 32: mean, var, dd on the support; elsewhere mean = 0,
 * var = gamma, dd = 0. */
int fbme_reduced_posterior(void* stream, int batch, int k, int n1, int n2, int m1, int m2, const int* support,
                           const float* gamma, const double* gamma_sum, const float* noise, const void* bp,
                           const double* ye, const void* dir1, const void* dir2, float jitter_multiplier,
                           void* mean, float* var, float* dd, int* status);

/* ------------------------------------------------------------------------
 * Observation masks of the general solver: sbl_2d(..., observation_mask=)
 * in mode="cg" (torch_sbl/solver.py:986-1024).  The reference restricts the
 * prefix operator to the observed samples (PrefixFourierOperator.restrict /
 * expand, torch_sbl/operators.py:57-75 and 473-496).  These kernels keep
 * every vector at M1 M2 samples with zeros on the unobserved ones, so the
 * restricted operator is the full one followed by the mask; `mask` is one
 * byte per sample (nonzero = observed), shared by the batch.
 * ------------------------------------------------------------------------ */
/* spectrum[d1, d2] = sum over observed (m1, m2) of
 * exp(+2 pi i (m1 d1 / N1 + m2 d2 / N2)), complex128 [N1, N2]: the Gram entry
 * a_k^H a_l of two atoms on the observed rows is
 * spectrum[(k1 - l1) mod N1, (k2 - l2) mod N2]; for a full mask it is the
 * product of the fbme_dirichlet_table tables.  work: M1 * N2 complex128.
 * (The reference forms the same Gram from the restricted columns,
 * solver.py:5786-5848 and 3591-3779.) */
int fbme_mask_spectrum(void* stream, int n1, int n2, int m1, int m2, const unsigned char* mask, void* spectrum,
                       void* work);
/* out[lane, i] *= mask[i] for every lane, the restriction of Q d or M^-1 v
 * to the observed samples (operators.py:57-75); with dot (and a), also
 * dot[lane] = Re <a, out> (FP64), as fbme_lane_dot.  part: lanes * 64
 * doubles. */
int fbme_mask_lanes_dot(void* stream, int lanes, int m, const unsigned char* mask, const void* a, void* out,
                        double* part, double* dot);
/* fbme_lowrank_setup with the factor Gram taken on the observed rows
 * (_build_low_rank_preconditioner on the restricted operator,
 * solver.py:5786-5848); mask_spectrum from fbme_mask_spectrum (NULL: the
 * Dirichlet tables dir1, dir2 as in fbme_lowrank_setup). */
int fbme_lowrank_setup_masked(void* stream, int batch, int n1, int n2, int m1, int m2, const int* idx,
                              const float* val, const double* gamma_sum, const float* noise, const void* dir1,
                              const void* dir2, const void* mask_spectrum, void* factors, void* sinv, float* bg,
                              int* status);
/* fbme_reduced_posterior on the observed samples (solver.py:3591-3779 on the
 * restricted operator): Gram from mask_spectrum (NULL: Dirichlet tables)
 * and `observed` samples in the evidence's log-determinant. */
int fbme_reduced_posterior_masked(void* stream, int batch, int k, int n1, int n2, int observed, const int* support,
                                  const float* gamma, const double* gamma_sum, const float* noise, const void* bp,
                                  const double* ye, const void* dir1, const void* dir2, const void* mask_spectrum,
                                  float jitter_multiplier, void* mean, float* var, float* dd, int* status);
/* fbme_reduced_posterior_masked for supports of 33 to 256 atoms (the
 * general solver's active set up to active_set_max_size = 256,
 * solver.py:3184-3216, 1316-1341, 3591-3779): the FP64 system padded to a
 * multiple of 32 and factored / inverted on 32 x 32 tiles.  mask_spectrum
 * (or NULL with the Dirichlet tables dir1, dir2) gives the Gram; gamma_sum:
 * [batch] FP64 sum of gamma; work: fbme_reduced_posterior_large_bytes(batch,
 * k) bytes; status[b] = 1 on a non-positive pivot. */
size_t fbme_reduced_posterior_large_bytes(int batch, int k);
int fbme_reduced_posterior_large(void* stream, int batch, int k, int n1, int n2, const int* support,
                                 const float* gamma, const double* gamma_sum, const float* noise, const void* bp,
                                 const void* dir1, const void* dir2, const void* mask_spectrum,
                                 float jitter_multiplier, void* work, size_t work_bytes, void* mean, float* var,
                                 float* dd, int* status);
/* ------------------------------------------------------------------------
 * Exact dense posterior of the general solver, mode="dense": _posterior's
 * dense branch (torch_sbl/solver.py:2962-2990) with Q = noise I + A Gamma
 * A^H built from the Fourier lags of gamma (operators.py:282-297), the GPU
 * Cholesky loading jitter * mean|diag| (solver.py:2898-2920; jitter =
 * cuda_cholesky_jitter_multiplier * eps) and the data diagonal
 * diag(A^H Q^-1 A) (operators.py:418-442), in FP64: tiled 32 x 32 Cholesky
 * and triangular inverse over many workgroups per scene.  1 <= batch <=
 * 65535; observed <= 2048 samples; obs_idx [observed] int32 lists their
 * distinct sample indices in [0, M1 M2) (NULL: all M1 M2).  y is
 * [B, M1 M2] complex64 in the full sample layout.  Writes mean [B, N]
 * complex64, var and dd [B, N] float32, rel_sum[b] = sum clamp(gamma dd,
 * 0, 1); part: B * ceil(N / 256) doubles.  status[b] (zeroed by the caller,
 * may be NULL): 1 on a failed pivot, 2 when an obs_idx entry lies outside
 * [0, M1 M2) (that sample is left out).  work: work_bytes >=
 * fbme_dense_workspace_bytes(batch, n1, n2, m1, m2, observed), which
 * returns 0 for arguments the posterior rejects. */
size_t fbme_dense_workspace_bytes(int batch, int n1, int n2, int m1, int m2, int observed);
int fbme_dense_posterior(void* stream, int batch, int n1, int n2, int m1, int m2, int observed, const int* obs_idx,
                         const void* y, const float* gamma, const float* noise, double jitter, double tiny,
                         void* work, size_t work_bytes, void* mean, float* var, float* dd, double* part,
                         double* rel_sum, int* status);

/* ------------------------------------------------------------------------
 * Off-grid refinement: refine_offgrid_fourier (torch_sbl/offgrid.py:70-226)
 * for 2-D grids (1-D: fbme_offgrid_refine_dims).  One workgroup per scene
 * runs the whole refinement:
 * top max(64, 16 max_components) magnitudes of the coefficients, periodic
 * non-maximum suppression (offgrid.py:229-295), the extended BIC model order
 * when component_count <= 0 (offgrid.py:318-361), projected Gauss-Newton
 * with Levenberg damping, trust bound and backtracking (offgrid.py:364-476)
 * and the position standard deviations (offgrid.py:538-564), all in FP64.
 * data [B, M1*M2] and coefficients [B, N1*N2] are complex64 (complex128 when
 * `complex128` is nonzero).  Limits: M1 M2 >= 2, max_components <= 64, at
 * most 32 refined atoms (counts[b*3+1] = -1 when the model order exceeds
 * 32; fbme_offgrid_refine_ws refines up to 64); every pointer is required
 * (FBME_ERR_ARG otherwise).
 * Outputs, with Kout = component_count > 0 ? component_count :
 * max_components rows per scene (rows past the scene's count are NaN):
 * initial_positions / positions / position_std [B, Kout, 2] double,
 * amplitudes [B, Kout] complex128, norms [B, 2] (initial, final residual
 * norm), counts [B, 3] (atoms, iterations, converged).  eps / tiny are the
 * data precision's (the reference's finfo of the real dtype). */
int fbme_offgrid_refine(void* stream, int batch, int n1, int n2, int m1, int m2, const void* data,
                        const void* coefficients, int complex128, int component_count, int max_components,
                        float relative_threshold, float minimum_separation, int max_iter, double tolerance,
                        double max_step, double damping, double ridge, double eps, double tiny,
                        double* initial_positions, double* positions, void* amplitudes, double* position_std,
                        double* norms, int* counts);
/* fbme_offgrid_refine for 1-D or 2-D grids (refine_offgrid_fourier with
 * len(grid_shape) = dims, offgrid.py:70-226): dims = 1 runs a grid of N points
 * sampled at its first M as n1 = m1 = 1, n2 = N, m2 = M; the model order
 * counts one position coordinate per component (offgrid.py:349-350) and the
 * axis-0 outputs (positions 0, std NaN-free but meaningless) are dropped by
 * the caller.  dims = 2 is fbme_offgrid_refine. */
int fbme_offgrid_refine_dims(void* stream, int batch, int n1, int n2, int m1, int m2, const void* data,
                             const void* coefficients, int complex128, int component_count, int max_components,
                             float relative_threshold, float minimum_separation, int max_iter, double tolerance,
                             double max_step, double damping, double ridge, double eps, double tiny,
                             double* initial_positions, double* positions, void* amplitudes, double* position_std,
                             double* norms, int* counts, int dims);
/* fbme_offgrid_refine_dims with up to 64 refined atoms (the reference's
 * default max_components = 64 with a model order above 32, or
 * component_count <= 64).  workspace: fbme_offgrid_workspace_bytes(batch)
 * bytes of device memory, 16-byte aligned, scratch of the K x K and 2K x 2K
 * refinement matrices of scenes with K > 32 (scenes with K <= 32 keep them in
 * LDS and do not touch it; both give the same bits for the same K).  A null
 * workspace is fbme_offgrid_refine_dims. */
size_t fbme_offgrid_workspace_bytes(int batch);
int fbme_offgrid_refine_ws(void* stream, int batch, int n1, int n2, int m1, int m2, const void* data,
                           const void* coefficients, int complex128, int component_count, int max_components,
                           float relative_threshold, float minimum_separation, int max_iter, double tolerance,
                           double max_step, double damping, double ridge, double eps, double tiny,
                           double* initial_positions, double* positions, void* amplitudes, double* position_std,
                           double* norms, int* counts, int dims, void* workspace, size_t workspace_bytes);

/* ------------------------------------------------------------------------
 * Joint sparsity of the general solver: SBLConfig(joint_sparsity=True)
 * (torch_sbl/solver.py:1090-1099, 1157-1159, 1653-1666, 1769-1774,
 * 3039-3153).  The B snapshots share one variance per grid cell, kept as B
 * identical rows of gamma [B, N]; with shared noise the posterior is one
 * PCG system whose lanes are the B snapshots followed by the P probes.
 * ------------------------------------------------------------------------ */
/* After fbme_cg_init: gamma[b] = max(mean over snapshots of the
 * backprojection initialisation, max_b floor_b) for every b
 * (solver.py:1071-1099); with shared_noise also noise[b] = mean_b noise_b
 * (solver.py:1157-1159). */
int fbme_joint_init(void* stream, int batch, int n, const void* bp, const double* bpe, const float* power,
                    const float* floor_, float blend, int shared_noise, float* gamma, float* noise);
/* Shared-noise posterior moments (_shared_mmv_posterior,
 * solver.py:3120-3153): lifted holds B + probes lanes of A^H Q^-1 [y, p];
 * mean[b] = gamma lifted[b], and the Hutchinson diagonal dd, variance and
 * relevance sum (rel_sum[b]) written for every snapshot.  part: B *
 * ceil(N / 256) doubles. */
int fbme_hutchinson_shared(void* stream, int batch, int n, int probes, const void* lifted, const void* lprobes,
                           const float* gamma, void* mean, float* var, float* dd, double* part, double* rel_sum);
/* Projected relevance per snapshot (dof) and the complex M-SBL EM update of
 * the shared variances (solver.py:1653-1666, 1776-1798); part: B *
 * ceil(N / 256) doubles. */
int fbme_cg_em_joint(void* stream, int batch, int n, int m, float* gamma, const void* mean, const float* var,
                     const float* dd, const double* rel_sum, const float* floor_, float hyper_shape, float hyper_rate,
                     float max_prec, float keep, int update, double* part, double* dof, float* dof_f);
/* Shared noise update (solver.py:1737-1774): each snapshot's update over its
 * observed samples (mask may be NULL), floored at its floor, averaged,
 * floored at the largest floor and damped; written for every snapshot. */
int fbme_joint_noise_update(void* stream, int batch, int m, const unsigned char* mask, int observed,
                            const void* data, const void* prediction, const float* old_noise,
                            const float* effective_dof, const float* variance_floor, float* output, float noise_shape,
                            float noise_rate, float damping, int evidence_update, float epsilon, double* partial);
/* fbme_uamp_noise_update over the observed samples: the residual energy of
 * the samples with mask[i] nonzero (mask NULL: all) and `observed` in place
 * of M in the update (solver.py:1737-1768 on the restricted operator). */
int fbme_uamp_noise_update_masked(void* stream, int batch, int m, const unsigned char* mask, int observed,
                                  const void* data, const void* prediction, const float* old_noise,
                                  const float* effective_dof, const float* variance_floor, float* output,
                                  float noise_shape, float noise_rate, float damping, int evidence_update,
                                  float epsilon, double* partial);
/* SBLConfig.fixed_noise_variance of the general solver (solver.py:1151-1156,
 * 1747-1750, 1765-1770, 1788-1798): noise[b] = value at the start
 * (variance_floor NULL) and max(value, variance_floor[b]) in place of every
 * noise update; with `shared` (joint sparsity with a shared noise) the
 * largest floor of the batch, max(value, max_b variance_floor[b]), for every
 * snapshot.  value > 0. */
int fbme_fixed_noise(void* stream, int batch, float value, const float* variance_floor, float* noise, int shared);

/* ------------------------------------------------------------------------
 * Learned warm start of the general solver (no counterpart in the
 * reference, which has no learned component): a 4-layer 3x3 convolutional
 * network, channels 16 -> 32 -> 32 -> 32 -> 16 (input channels 0, 1 =
 * log|A^H y|^2 and log gamma, output channel 0 = per-cell log-variance
 * correction), BF16 weights and activations on MFMA with FP32 accumulation,
 * trained on the GPU by these kernels.  Activations are [B, H, W, C] bf16;
 * parameters fp32, per layer W[9][ci][co] then b[co].
 * ------------------------------------------------------------------------ */
long long fbme_cnn_param_count(void);
long long fbme_cnn_weight_count(void);
/* bf16 forward (wf) and backward (wb) weights from the fp32 parameters,
 * fbme_cnn_weight_count() values each. */
int fbme_cnn_prepare(void* stream, const float* params, void* wf, void* wb);
/* x [B, n, 16] bf16 from the backprojection bp [B, n] complex64 and gamma
 * [B, n]; bpe = sum |bp|^2 and gsum = sum gamma per scene (FP64). */
int fbme_cnn_features(void* stream, int batch, int n, const void* bp, const float* gamma, const double* bpe,
                      const double* gsum, void* x);
/* Inputs of the preconditioner network: channels 0, 1 as fbme_cnn_features,
 * channel 2 = log(1 + N gamma / noise) / 8 (noise [batch]). */
int fbme_cnn_features_precond(void* stream, int batch, int n, const void* bp, const float* gamma,
                              const double* bpe, const double* gsum, const float* noise, void* x);
/* dz4 [count, 16] bf16 from a loss gradient g [count] at the network output
 * (channel 0), for losses formed outside fbme_cnn_loss. */
int fbme_cnn_output_grad(void* stream, long long count, const float* g, void* dz4);
/* out [B, H, W] fp32; act: 3 * B * H * W * 32 bf16 hidden activations. */
int fbme_cnn_forward(void* stream, int batch, int h, int w, const void* x, const void* wf, const float* params,
                     void* act, float* out);
long long fbme_cnn_grad_rows(int h);
/* fp32 partial-sum workspace of fbme_cnn_backward. */
long long fbme_cnn_part_floats(int batch, int h, int w);
/* Mean squared error of out against target over `count` cells, its gradient
 * dz4 [count, 16] bf16, loss (FP64) via part (ceil(count / 256) doubles). */
int fbme_cnn_loss(void* stream, long long count, const float* out, const float* target, void* dz4, double* part,
                  double* loss);
/* Parameter gradients (fp32, parameter layout) from the forward
 * activations; dz: 2 * B * H * W * 32 bf16 workspace. */
int fbme_cnn_backward(void* stream, int batch, int h, int w, const void* x, const void* act, const void* wb,
                      const void* dz4, void* dz, float* part, float* grads);
/* Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected at `step` >= 1). */
int fbme_cnn_adam(void* stream, long long n, float* params, const float* grads, float* m, float* v, float lr,
                  int step);
/* gamma <- max(gamma exp(clamp(out, +-clampv)), floor[b]). */
int fbme_cnn_apply_warm_start(void* stream, int batch, int n, float* gamma, const float* out, const float* floor_,
                              float clampv);

#ifdef __cplusplus
}
#endif

#endif /* FBME_HIP_H */
