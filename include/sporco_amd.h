/*
 * sporco_amd.h - C ABI of the MI355X (gfx950) convolutional BPDN ADMM solvers.
 *
 * Drop-in replacement of the C interface that the reference's Cython layer
 * binds (sporco_cuda/cbpdn.pyx:11-53, sporco_cuda/util.pyx:3-7).  Every
 * function returns a status code instead of calling exit() (the reference's
 * error convention, sporco_cuda/src/error_check.cuh:16-21 and
 * sporco_cuda/src/utils.cu:106-124), never changes the caller's current
 * device, and never keeps a pointer to caller memory after it returns.
 *
 * Array layouts are the reference's C order: D (M, Kh, Kw), S (R, C),
 * Y (M, R, C), L1 weights (1), (M) or (M, R, C), mask (R, C) int32.
 */
#ifndef SPORCO_AMD_H
#define SPORCO_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference has none: every failure exits the process). */
typedef enum scc_status {
  SCC_OK = 0,
  SCC_ERR_INVALID = 1,     /* bad sizes, options or pointers */
  SCC_ERR_NOMEM = 2,       /* device allocation failed */
  SCC_ERR_NODEVICE = 3,    /* no HIP device, or device index out of range */
  SCC_ERR_HIP = 4,         /* a HIP runtime call or kernel launch failed */
  SCC_ERR_UNSUPPORTED = 5  /* size outside what the kernels implement */
} scc_status;

/*
 * Algorithm options and problem sizes.  Same fields, names and meaning as
 * the reference's struct AlgOpt (sporco_cuda/src/algopt.h:7-48), so that the
 * Cython marshalling (cbpdn.pyx:57-79, 158-201, 297-349) carries over field
 * by field.  Differences: weight pointers are const; RhoRsdlTarget <= 0
 * means "derive from lambda" (sporco's None) instead of being overwritten
 * unconditionally (cbpdn.cu:159-162).
 */
typedef struct scc_algopt {
  int MaxMainIter;
  float rho;
  float RhoRsdlRatio;
  float RhoScaling;
  float RhoRsdlTarget;
  float RelaxParam;
  float AbsStopTol;
  float RelStopTol;
  const float *L1Weight;
  int L1_WEIGHT_M_SIZE;
  int L1_WEIGHT_ROW_SIZE;
  int L1_WEIGHT_COL_SIZE;
  const int *Weight; /* AMS mask (R, C), used when nWeight > 1 */
  int nWeight;
  int Verbose;
  int NonNegCoef;
  int NoBndryCross;
  int AuxVarObj;
  int HighMemSolve;
  int AutoRho;
  int AutoRhoScaling;
  int AutoRhoPeriod;
  int StdResiduals;
  int IMG_ROW_SIZE;
  int IMG_COL_SIZE;
  int DICT_M_SIZE;
  int DICT_ROW_SIZE;
  int DICT_COL_SIZE;
  const float *GrdWeight;
  int WEIGHT_SIZE;
  int device;
} scc_algopt;

/* Outcome of the last solve on the calling thread. */
typedef struct scc_info {
  int iterations;   /* ADMM iterations run */
  float rho;        /* final penalty parameter */
  float r, s;       /* final primal / dual residuals (normalised unless
                       StdResiduals) */
  float epri, edua; /* final stopping thresholds */
  double setup_ms;  /* device time of setup (spectra of D and S) */
  double solve_ms;  /* device time of the ADMM iterations */
} scc_info;

/* Replaces clear_opts (sporco_cuda/src/utils.cu:66-98): sets the -1
 * sentinels.  Unlike the reference it allocates nothing. */
void scc_clear_opts(scc_algopt *opt);

/* Replaces default_opts (sporco_cuda/src/utils.cu:12-64): fills sentinels
 * with defaults (RelaxParam 1.8, RelStopTol 1e-4, MaxMainIter 250, ...). */
void scc_default_opts(scc_algopt *opt);

/* Replaces cuda_wrapper_CBPDN (sporco_cuda/src/cbpdn.h:10,
 * sporco_cuda/src/cbpdn.cu:86-616): plain CBPDN, and CBPDN with an AMS
 * mask when opt->nWeight > 1 (filter 0 of D is then the impulse).
 * Host pointers; Y (M, R, C) is overwritten.  opt is read-only. */
int scc_cbpdn(const float *D, const float *S, float lambda,
              const scc_algopt *opt, float *Y);

/* Replaces cuda_wrapper_CBPDN_GR (sporco_cuda/src/cbpdn_grd.h:10-11,
 * sporco_cuda/src/cbpdn_grd.cu:97-658): gradient-regularised CBPDN, masked
 * when opt->nWeight > 1. */
int scc_cbpdn_grd(const float *D, const float *S, float lambda, float mu,
                  const scc_algopt *opt, float *Y);

/* Device-pointer forms of the two solvers for callers whose data already
 * lives in HBM.  All pointers in opt (L1Weight, Weight, GrdWeight) and the
 * array arguments are device pointers; work is queued on `stream` (a
 * hipStream_t, NULL = default stream) and the call returns after the
 * solve has completed.  No counterpart in the reference, whose C entry
 * points take host arrays only (cbpdn.cu:249-260, 575). */
int scc_cbpdn_device(const float *D, const float *S, float lambda, float mu,
                     int grad, const scc_algopt *opt, float *Y, void *stream);

/* Bytes of the fixed device array that holds the workspace of
 * scc_cbpdn_device_noalloc (one copy per GPU, held until the process
 * ends): 128 MiB. */
#define SCC_NOALLOC_SCRATCH_BYTES 134217728

/* Allocation-free form of scc_cbpdn_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * that prepares HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and stream semantics as
 * scc_cbpdn_device, and the same kernels; it calls no hipMalloc, hipFree
 * or other allocating routine.  The solver state lives in Y, every other
 * buffer (filter and work spectra, U, the image spectrum and small tables:
 * about 16 M Hc + 4 M N + 12 Hc bytes, Hc = R (C/2 + 1), N = R C) in a
 * fixed device array of SCC_NOALLOC_SCRATCH_BYTES per GPU; a problem that
 * needs more returns SCC_ERR_UNSUPPORTED before any device work (e.g.
 * 256 x 256 with 32 filters needs 25 MB, 512 x 512 with 32 filters 101 MB,
 * 512 x 512 with 64 filters does not fit).  Verbose returns
 * SCC_ERR_UNSUPPORTED (its objective buffers would need allocation);
 * per-iteration history and stage timing are not recorded.  Every call on
 * a GPU shares that GPU's array, which keeps the last call's contents:
 * calls that overlapped on one GPU would race, so the library serialises
 * them per GPU (each returns after its solve has completed). */
int scc_cbpdn_device_noalloc(const float *D, const float *S, float lambda,
                             float mu, int grad, const scc_algopt *opt,
                             float *Y, void *stream);

/* Info and error text of the last call on this thread (no reference
 * counterpart: the reference prints and exits). */
int scc_last_info(scc_info *info);
const char *scc_last_error(void);

/* Per-iteration history (r, s, rho) of the last solve on this thread, when
 * enabled with scc_set_history(1).  Writes up to cap triples; returns the
 * number of iterations recorded.  No reference counterpart. */
void scc_set_history(int enabled);
int scc_last_history(float *rsr, int cap);

/* Objective after every iteration of later solves on this thread (always
 * when Verbose): rows of (Fnc, DFid, Regl1, Regl2grd) with
 *   DFid = 1/2 || sum_m d_m * z_m - s ||^2 (1/2 || W (...) ||^2, impulse
 *          map excluded, for the masked variants),
 *   Regl1 = || w z ||_1,  Regl2grd = 1/2 sum_m w^g_m || grad z_m ||^2,
 *   Fnc = DFid + lambda Regl1 (+ mu Regl2grd),
 * z = X, or Y with AuxVarObj (sporco's definitions; the reference's
 * objective, cbpdn.cu:438-517, is not followed).  scc_last_objective
 * copies up to cap rows and returns the number of iterations. */
void scc_set_objective(int enabled);
int scc_last_objective(float *rows, int cap);

/* Per-stage device timing of later solves on this thread (HIP events on
 * the solve stream around each kernel of every iteration).  Stages:
 * 0 forward column FFT, 1 X-solve, 2 inverse column FFT, 3 row kernel
 * (inverse row FFT + update + forward row FFT), 4 residual reduction and
 * control, 5 forward row FFT after a rho change.  scc_last_stage_times
 * fills total milliseconds and launch counts for up to n stages and
 * returns the number of stages.  No reference counterpart. */
void scc_set_stage_timing(int enabled);
int scc_last_stage_times(double *ms_total, int *launches, int n);

/* Kernel choice of later solves on this thread: 1 (default) the kernels
 * compiled for fixed transform lengths where one exists (fused row kernels
 * for even widths that are multiples of 8 up to 1632 with the plain update:
 * no mask, no full-array weight, NonNegCoef and NoBndryCross off; column
 * transforms for heights that are multiples of 8 from 32 to 1024), 0 the
 * run-time-plan kernels everywhere.  Both compute the same expressions in
 * the same order; the switch exists to compare them.  No reference
 * counterpart. */
void scc_set_fixed_kernels(int on);

/* Graph replay of repeated fixed-iteration solves on this thread (default
 * on): a solve with no stopping tolerance, no Verbose output, objective or
 * stage timing whose configuration (sizes, parameters, buffers, weight and
 * mask pointers, iteration count) equals the previous such solve in the
 * same cached workspace runs its iterations from a captured HIP graph
 * (kernel nodes only), captured on the second call and replayed on later
 * ones; the results are those of direct launches.  scc_last_solve_graph
 * returns 1 when the last solve on this thread ran from a graph.  No
 * reference counterpart (the reference launches every kernel from the
 * host and synchronises twice per iteration, cbpdn.cu:318-567). */
void scc_set_solve_graph(int enabled);
int scc_last_solve_graph(void);

/* Differentiable unrolled plain CBPDN over a batch of B images of one
 * power-of-two size (sides 256..2048) sharing a dictionary: n_iter ADMM
 * iterations with fixed rho, relaxation alpha, scalar lambda, L1 weight 1.
 * All arrays are device pointers: D (M, K, K), S (B, R, C), Y (B, M, R, C),
 * ckpt of scc_unrolled_ckpt_floats(...) floats ((Y, U) every ckpt_every
 * iterations, and the tape of the last three segments: X-hat and the
 * soft-threshold support of each of their iterations); ckpt may be NULL in
 * the forward for an inference-only solve.  The backward takes dL/dY and
 * writes dL/dD (M, K, K), dL/dS (B, R, C) and dL/dlambda (one float, host
 * pointer), recomputing the earlier segments from their checkpoints and
 * running the hand-written adjoint kernels.  No
 * counterpart in the reference (extension of the same iteration,
 * cbpdn.cu:318-567). */
size_t scc_unrolled_ckpt_floats(int M, int B, int R, int C, int n_iter,
                                int ckpt_every);
int scc_unrolled_forward(const float *D, int M, int K, const float *S, int B,
                         int R, int C, float lambda, float rho, float alpha,
                         int n_iter, int ckpt_every, float *Y, float *ckpt,
                         void *stream);
int scc_unrolled_backward(const float *D, int M, int K, const float *S, int B,
                          int R, int C, float lambda, float rho, float alpha,
                          int n_iter, int ckpt_every, const float *ckpt,
                          const float *gY, float *gD, float *gS,
                          float *glambda, void *stream);
/* The same forward, writing V = alpha X + (1 - alpha) Y + U of every
 * iteration to V (n_iter * B * M * R * C floats, device): the soft-threshold
 * supports 1[|V| > lambda / rho] the forward took, for checking the adjoint
 * on the branch of the piecewise-linear iteration the GPU computed. */
int scc_unrolled_tape(const float *D, int M, int K, const float *S, int B,
                      int R, int C, float lambda, float rho, float alpha,
                      int n_iter, float *V, void *stream);
const char *scc_unrolled_last_error(void);

/* Reconstruction out_b = sum_m d_m * x_{b,m} (circular convolution with the
 * filters zero-padded at the origin) of B coefficient sets, and its adjoint
 * (gD = dL/dD (M, Kh, Kw), gX = dL/dX (B, M, R, C) from gout = dL/dout
 * (B, R, C); either output may be NULL).  Device pointers; D (M, Kh, Kw),
 * X (B, M, R, C), out (B, R, C).  No counterpart in the reference; it
 * replaces sporco's fft.fftconv use in examples/big_cbpdn.py:69. */
int scc_reconstruct(const float *D, int M, int Kh, int Kw, const float *X,
                    int B, int R, int C, float *out, void *stream);
int scc_reconstruct_backward(const float *D, int M, int Kh, int Kw,
                             const float *X, int B, int R, int C,
                             const float *gout, float *gD, float *gX,
                             void *stream);

/* Batched plain CBPDN: nimg images of any sizes sharing one dictionary D
 * (M, Kh, Kw), each with its own lambda, solved together.  Every ADMM stage
 * is one launch over a work list that spans all images, or every image runs
 * the single-image kernels of its size as a chain of launches, the chains
 * spread over up to 4 branches; either way the iteration loop is replayed
 * from a captured HIP graph.  The mode, panel width, rows per row block,
 * lanes per frequency of the solve and the grid shape are chosen by timing
 * candidates on the batch itself the first time a batch signature (device,
 * M, image sizes) is seen.  Per image the result equals scc_cbpdn
 * on that image (same iteration, residuals, stopping test and AutoRho,
 * applied per image); options are read from opt as there, except the size
 * fields, which come from the arguments.  L1Weight may hold 1 or M entries;
 * masks (nWeight > 1) are rejected.  D, images[i].S (R, C) and images[i].Y
 * (M, R, C) are device pointers; images and results are host arrays;
 * images[i].rho > 0 sets that image's initial rho, otherwise opt->rho.  No
 * counterpart in the reference, whose only batch form is a host loop over
 * cuda_wrapper_CBPDN (cbpdn.h:10). */
typedef struct {
  int R, C;
  const float *S;
  float *Y;
  float lambda;
  float rho;
} scc_batch_image;

typedef struct {
  int iterations;
  float rho, r, s, epri, edua;
} scc_batch_result;

typedef struct {
  int col_lds;       /* LDS budget of a column block (bytes) */
  int row_lds;       /* LDS budget of a row block (bytes) */
  int solve_lanes;   /* lanes per frequency in the X-solve */
  int solve_maps;    /* maps per lane in the X-solve */
  int persistent;    /* 1: occupancy-sized grids, 0: one block per item */
  int graph;         /* the iterations ran from a captured graph */
  double tune_ms;    /* autotuning time of this call (0 when cached) */
  double setup_ms;   /* spectra of D and the images (device time) */
  double solve_ms;   /* iterations (device time) */
  int branches;      /* 0: work-list kernels; k > 0: per-image kernel
                        chains on k graph branches */
} scc_batch_config;

int scc_batched_cbpdn(const float *D, int M, int Kh, int Kw, int nimg,
                      const scc_batch_image *images, const scc_algopt *opt,
                      scc_batch_result *results, void *stream);
/* Tuning policy of later calls: 1 autotune new signatures (default), 0 use
 * the fixed default configuration. */
int scc_batched_set_autotune(int on);
/* Mode of later calls: k < 0 the tuned mode (default), 0 work-list
 * kernels, k > 0 per-image kernel chains on min(k, 4) graph branches. */
int scc_batched_set_branches(int k);
/* Graph policy: 1 capture the iterations in a HIP graph (default), 0 launch
 * them directly. */
int scc_batched_set_graph(int on);
int scc_batched_last_config(scc_batch_config *cfg);

/* Device queries: replace get_device_count, get_current_device,
 * set_current_device, get_memory_info, get_device_name
 * (sporco_cuda/src/utils.cu:146-175, bound at util.pyx:3-7). */
int scc_get_device_count(void);
int scc_get_current_device(int *dev);
int scc_set_current_device(int dev);
int scc_get_memory_info(size_t *free_bytes, size_t *total_bytes);
const char *scc_get_device_name(int dev);

/* Release cached plans and device buffers (no reference counterpart; the
 * reference allocates and frees per call, cbpdn.cu:186-215, 583-615). */
void scc_release_cache(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* SPORCO_AMD_H */
