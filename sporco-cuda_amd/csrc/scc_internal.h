// Internal types shared by the HIP kernels and the host driver.
//
// Frequency-domain data is kept in a "panel" layout: the half spectrum of a
// map, R rows x H = C/2+1 columns, is cut into panels of TC adjacent columns;
// each panel is stored row-major (R x width) and panels follow each other.
// A column transform therefore reads one contiguous block, and a row
// transform touches TC*8 contiguous bytes per panel per row.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace scc {

constexpr int kMaxStages = 24;

// Mixed-radix Stockham plan for one transform length.
struct FftPlan {
  int L;                    // transform length (complex points)
  int nstages;
  int radix[kMaxStages];
  int tw_stride;            // twiddle table holds W_{L*tw_stride}^t
};

struct Geom {
  int R, C, M;              // image rows, columns, number of maps
  int H;                    // C/2 + 1
  int Lr;                   // row complex FFT length: C/2 (C even) or C
  int even;                 // C even: rows use the half-length packing
  int TC;                   // panel width
  int P;                    // number of panels
  int wlast;                // width of the last panel
  long long N;              // R*C
  long long Hc;             // R*H
  int lgR;                  // log2(R) when R is a power of two, else -1
};

// Offset of (r, c) of the half spectrum inside one map's panel layout.
__host__ __device__ inline long long panel_off(const Geom& g, int r, int c) {
  int p = c / g.TC;
  int w = (p == g.P - 1) ? g.wlast : g.TC;
  return (long long)p * g.TC * g.R + (long long)r * w + (c - p * g.TC);
}

// Device-resident solver state read and written by the kernels, so that an
// iteration never needs the host.
struct Ctrl {
  float rho;       // current penalty parameter
  float rsf;       // factor of the last rho update (U is divided by it)
  int done;        // stopping test passed: every later kernel returns early
  int iter;        // iterations completed
  float r, s, epri, edua;
  float nX, nY, nU;
  float pad_;
};

// Per-solve constants of the update and control kernels.
struct Params {
  float lambda;
  float mu;
  float alpha;          // RelaxParam
  float abs_tol, rel_tol;
  float tau;            // AutoRho Scaling
  float rsdl_ratio;     // AutoRho RsdlRatio
  float xi;             // AutoRho RsdlTarget
  float l1_scalar;
  float gw_scalar;
  double sqrt_nc;       // sqrt(number of coefficients)
  int std_residuals;
  int autorho, autoscaling, period;
  int nonneg, nobndry, Kh, Kw;
  int m0;               // 1 when map 0 is the AMS impulse map
  int l1_mode;          // 0 scalar, 1 per map, 2 full (M,R,C) array
  int grad;             // gradient regularisation on
  int gw_mode;          // 0 scalar, 1 per map
  int track;            // record per-iteration history
  // 1: the state array (Buffers::Y) holds V = Xr + U of the last update
  // instead of Y and U, which are Y = shrink(V), U = V - Y; the row kernel
  // then reads and writes one real array per coefficient instead of two.
  // An update-only row launch (AutoRho step) writes Y and U, and the
  // forward row transform of kFwdYUtoV turns them back into V.  Valid
  // when every threshold (lambda/rho) w is >= 0: then shrink(0) = 0, so V =
  // 0 is the initial Y = U = 0, and V' = Y + U/rsf decodes to (Y, U/rsf)
  // after a rho update.
  int vstate;
};

// Row-kernel modes
enum RowMode {
  kRowFused = 0,    // inverse row FFT + update + residuals + forward row FFT
  kRowUpdate = 1,   // inverse row FFT + update + residuals only
};
enum RowFwdMode {
  kFwdReal = 0,     // forward row FFT of a real (rows, C) array
  kFwdYU = 1,       // forward row FFT of Y - U/rsf (U rescaled in place)
  kFwdYUtoV = 2,    // the same transform; Y replaced by V = Y + U/rsf
};

struct Buffers {
  float* Y;
  float* U;
  float2* F;        // M maps, panel layout: A-hat / X-hat work spectrum
  float2* Df;       // M maps, panel layout: filter spectra
  float2* Sf;       // 1 map, panel layout: image spectrum
  float* ghg;       // 1 map, panel layout (GradReg only)
  const float* l1w; // per map (M) or full (M,R,C) L1 weights
  const float* gw;  // per map gradient weights (M)
  const int* mask;  // (R,C) AMS mask
  float* partials;  // 5 x nblocks residual partial sums
  double* blocksum; // 5 per reduction block (first reduction pass)
  Ctrl* ctrl;
  float* hist;      // 3 x max_iter history (r, s, rho), optional
  float2* twR;      // W_R^t, t < R
  float2* twC;      // W_C^t, t < C
  float* vout;      // optional copy of V = Xr + U (unrolled solve tape)
  unsigned* tbits;  // optional support bits of V (unrolled solve tape)
  float2* xtape;    // optional copy of X-hat after the solve (same layout)
  const float* Yin; // optional source of the row kernel's Y (else Y)
  const float* Uin; // optional source of the row kernel's U (else U)
};

// Kernel launchers (kernels.hip).
hipError_t launch_pad(const float* D, float* out, int M, int Kh, int Kw,
                      int R, int C, hipStream_t st);
// iter: the iteration of a kFwdYUtoV launch (its conversion is skipped
// when the solve stopped at an earlier iteration).
hipError_t launch_row_fwd(const Geom& g, int nmaps, const FftPlan& rp,
                          int mode, const float* src, float* Y, float* U,
                          float2* F, const float2* twC, const Ctrl* ctrl,
                          hipStream_t st, int iter = 0);
// parts / nparts / bsum: first pass of a deferred residual reduction (see
// launch_solve), as in the power-of-two column kernel.
hipError_t launch_col(const Geom& g, int nmaps, const FftPlan& cp,
                      float2* F, const float2* twR, int inverse,
                      const Ctrl* ctrl, hipStream_t st,
                      const float* parts = nullptr, int nparts = 0,
                      double* bsum = nullptr);
// ctl_iter > 0: block 0 first finishes the deferred reduction of that
// iteration from nbsum block sums (only with the register solve, M <= 256).
hipError_t launch_solve(const Geom& g, const Buffers& b, const Params& p,
                        hipStream_t st, int ctl_iter = 0, int nbsum = 0);
hipError_t launch_row(const Geom& g, const FftPlan& rp, const Buffers& b,
                      const Params& p, int mode, int nblocks_out[1],
                      hipStream_t st);
hipError_t launch_reduce(const Buffers& b, const Params& p, int nblocks,
                         int iter, int max_iter, hipStream_t st);
int row_blocks(const Geom& g);
// Fused row kernels with a compile-time row length (rows_ct.hip): whether
// Lr has one, and its launch (even width, plain update path, wavefront
// rows; same results as launch_row's kernel).
bool row_ct_supported(int LR);
// Column transforms with a compile-time length (cols_ct.h): whether R has
// one, and its launch (the arguments of launch_col's kernel).
bool col_ct_supported(int R);
hipError_t launch_col_ct(const Geom& g, int nmaps, float2* F,
                         const float2* twR, int inverse, const Ctrl* ctrl,
                         const float* parts, int nparts, double* bsum,
                         int nred, size_t shm, hipStream_t st);
// Whether launch_row and launch_col use them (this thread; default on).
void set_row_ct(bool on);
bool row_ct_enabled();
hipError_t launch_row_ct(const Geom& g, const Buffers& b, const Params& p,
                         int nr, int nb, size_t shm, hipStream_t st);
// flag[0] = 1 if any of the n device values is negative, else 0.
hipError_t launch_any_negative(const float* w, long long n, int* flag,
                               hipStream_t st);
// flag[0] = 1 if any of the n mask entries is neither 0 nor 1, else 0.
hipError_t launch_any_nonbinary(const int* mask, long long n, int* flag,
                                hipStream_t st);
// Zero fill of bytes / 4 32-bit words at p (a kernel, not a memset).
hipError_t launch_zero(void* p, size_t bytes, hipStream_t st);
// Twiddle table tw[t] = exp(-2 pi i t / L), t < L, from double-precision
// cos / sin rounded to float.
hipError_t launch_twiddles(float2* tw, int L, hipStream_t st);
// Gradient-regularisation spectrum 4 - 2 cos(2 pi r / R) - 2 cos(2 pi c / C)
// (double precision, rounded to float) of every half-spectrum frequency, in
// the panel layout of g.
hipError_t launch_ghg(const Geom& g, float* ghg, hipStream_t st);
// Y = shrink(V) of a state array holding V (Params::vstate) into out.
hipError_t launch_state_to_y(const Geom& g, const Buffers& b,
                             const Params& p, float* out, hipStream_t st);

// Objective kernels (objective.hip): per-frequency fidelity / gradient
// sums over the spectra G of the objective variable (E != nullptr: store the
// residual spectrum instead), L1 sums over the variable Z, masked spatial
// fidelity, and the totals (Fnc, DFid, Regl1, Regl2g) into out / hist.
int obj_blocks(const Geom& g);
hipError_t launch_obj_freq(const Geom& g, const Buffers& b, const Params& p,
                           const float2* G, float2* E, double* part, int iter,
                           hipStream_t st);
hipError_t launch_obj_l1(const Geom& g, const Params& p, const float* Z,
                         const float* l1w, const Ctrl* ctrl, double* part,
                         int nblk, int iter, hipStream_t st);
hipError_t launch_obj_mask(const Geom& g, const float* e, const int* mask,
                           const Ctrl* ctrl, double* part, int nblk, int iter,
                           hipStream_t st);
hipError_t launch_obj_final(const double* pf, int nf, const double* pl,
                            int nl, const double* pm, int nm, const Params& p,
                            double invN, const Ctrl* ctrl, float* out,
                            float* hist, int iter, hipStream_t st);

// Host helpers of the driver (api.hip): the thread's last-error message
// returned by scc_last_error, and the mixed-radix plan of one FFT length.
int set_error(int code, const std::string& msg);
// The same without building a string: "what: detail" (detail may be NULL).
int set_error(int code, const char* what, const char* detail);
bool make_plan(int L, int tw_stride, FftPlan* pl, bool pfa);
struct Pow2Config;
// Geometry of an (R, C) image with M maps; *pow2 tells whether the
// register-resident power-of-two kernels apply (their config in *p2).
int make_geom(int R, int C, int M, Geom* g, bool* pow2, Pow2Config* p2);
// Inverse real row transforms of nmaps maps of half spectra F (panel
// layout, destroyed) into out (nmaps, R, C), times scale.
hipError_t launch_row_inv(const Geom& g, int nmaps, const FftPlan& rp,
                          float2* F, float* out, const float2* twC,
                          float scale, hipStream_t st);

}  // namespace scc
