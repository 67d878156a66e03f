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

#define FBME_ABI_VERSION 1

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
  int polish_size;         /* 32 active-set capacity (<= 32)               */
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
} FbmeIO;

typedef struct fbme_plan fbme_plan;

enum {
  FBME_Q_WORKSPACE_BYTES = 0,
  FBME_Q_TOTAL_GRID = 1,
  FBME_Q_TOTAL_SAMPLES = 2,
  FBME_Q_MAX_ITERS = 3,
  FBME_Q_NSCENES = 4,
  FBME_Q_EVALUATIONS = 5,
  FBME_Q_LAUNCHES = 6
};

int fbme_abi_version(void);

/* Defaults of the Fast (stable = 0) or Stable (stable = 1) profile. */
void fbme_default_options(FbmePlanOptions* options, int stable);

/* Build a plan for `nscenes` scenes, shapes[4*i .. 4*i+3] = N1, N2, M1, M2.
 * Replaces the shape-fixing part of CUDAGraphSBLPlan.__init__
 * (torch_sbl/cuda_graph.py:22-62) and the operator cache
 * (torch_sbl/solver.py:844-857).  Uneven shapes are allowed.  On failure
 * returns NULL and writes a message into err. */
fbme_plan* fbme_plan_create(int nscenes, const int* shapes, const FbmePlanOptions* options, char* err,
                            int errlen);
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
 * 3 backprojection, 4 per-scene scalars (double[8]), 5 prediction,
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

#ifdef __cplusplus
}
#endif

#endif /* FBME_HIP_H */
