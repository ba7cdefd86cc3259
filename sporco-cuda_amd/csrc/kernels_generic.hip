// Generic-size kernels of one CBPDN ADMM iteration for gfx950.
//
// These kernels accept any image size (odd widths, prime factors of any
// size): the row and column transforms are a mixed-radix Stockham FFT in LDS
// with ping-pong buffers, specialised butterflies for radix 2 and 4 and a
// direct DFT butterfly for every other factor.  Power-of-two images take the
// register-resident kernels in kernels_pow2.hip instead.
//
// Stage map of one iteration (reference cbpdn.cu:318-567):
//   k_col (forward)   column FFTs: row spectra of Y-U -> 2-D spectrum
//   k_solve           per-frequency Sherman-Morrison X-solve (in place)
//   k_col (inverse)   column inverse FFTs
//   k_row             row inverse FFT -> X; relax, shrink, dual update,
//                     residual partial sums; Y-U -> row forward FFT
//   k_reduce          residual norms, stopping test, AutoRho (device side)

#include <algorithm>
#include <cmath>

#include "generic_common.h"
#include "scc_pow2.h"

namespace scc {

// Rows per row-kernel block: kRowWaves wavefronts of one or two rows each
// (wavefront-per-row kernels), or as many as fit a 60 KiB LDS budget
// (block-wide kernels, rows longer than the wavefront buffers allow).
static inline int row_nr(const Geom& g) {
  if (!row_wave_fits(g)) return row_nr_budget(g, 61440);
  const long long nrows = (long long)g.M * g.R;
  return nrows >= 2LL * kRowWaves * 2048 ? 2 * kRowWaves : kRowWaves;
}

static inline size_t row_lds(const Geom& g, int nr) {
  return row_wave_fits(g) ? (size_t)row_wave_lds(g)
                          : (size_t)2 * nr * gen::gld(g.Lr) * sizeof(float2);
}

// --------------------------------------------------------------------------
// Filter padding (cuda_Pad_Dict, cbpdn_kernels.cu:602-615): origin at [0,0].
// --------------------------------------------------------------------------
__global__ void k_pad(const float* __restrict__ D, float* __restrict__ out,
                      int M, int Kh, int Kw, int R, int C) {
  const long long n = (long long)M * R * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long m = i / ((long long)R * C);
    const long long rc = i - m * R * C;
    const int r = (int)(rc / C);
    const int c = (int)(rc - (long long)r * C);
    out[i] = (r < Kh && c < Kw) ? D[(m * Kh + r) * Kw + c] : 0.f;
  }
}

// --------------------------------------------------------------------------
// Forward row transform of a real array (setup) or of Y - U/rsf (AutoRho).
// --------------------------------------------------------------------------
template <bool WAVE>
__global__ __launch_bounds__(256) void k_row_fwd_generic(
    Geom g, long long nrows, FftPlan rp, int mode, const float* src,
    float* Y, float* U, float2* F, const float2* __restrict__ twC,
    const Ctrl* ctrl, int nr, int iter) {
  // The V conversion also runs when the stopping test passed at this very
  // iteration (its row kernel wrote Y and U), so that the state array holds
  // V whatever iteration the solve stopped at.
  if (mode == kFwdYU && ctrl->done) return;
  if (mode == kFwdYUtoV && ctrl->done && ctrl->iter != iter) return;
  extern __shared__ float2 lds[];
  const float rsf = (mode != kFwdReal) ? ctrl->rsf : 1.f;
  if constexpr (WAVE)
    row_fwd_body_wave(g, nrows, rp, mode, src, Y, U, F,
                      row_wave_table(g, lds, twC), rsf, nr, blockIdx.x, lds);
  else
    row_fwd_body(g, nrows, rp, mode, src, Y, U, F, twC, rsf, nr, blockIdx.x,
                 lds);
}

// Inverse real row transform of half spectra into real rows, times scale.
__global__ __launch_bounds__(256) void k_row_inv_generic(
    Geom g, long long nrows, FftPlan rp, const float2* F, float* out,
    const float2* __restrict__ twC, float scale, int nr) {
  extern __shared__ float2 lds[];
  row_inv_body(g, nrows, rp, F, out, twC, scale, nr, blockIdx.x, lds);
}

// --------------------------------------------------------------------------
// Column transforms, one panel of one map per block, in place.
// --------------------------------------------------------------------------
// With nparts > 0 the first nred blocks first add slices of the previous
// row kernel's residual partials (first pass of a deferred reduction).
__global__ __launch_bounds__(256) void k_col_generic(
    Geom g, FftPlan cp, float2* F, const float2* __restrict__ twR,
    float dir, const Ctrl* ctrl, const float* __restrict__ parts, int nparts,
    double* bsum, int nred) {
  if (ctrl != nullptr && ctrl->done) return;
  extern __shared__ float2 lds[];
  __shared__ double red[5][16];
  const int lin = blockIdx.y * gridDim.x + blockIdx.x;
  if (nparts > 0 && lin < nred)
    partials_slice(parts, nparts, bsum, lin, nred, red);
  // Forward: maps last to first, so that the row kernel's most recent
  // output is read while it is still in the last-level cache.
  const int m = dir > 0.f ? (int)gridDim.y - 1 - (int)blockIdx.y
                          : (int)blockIdx.y;
  col_body(g, cp, F, twR, dir, blockIdx.x, m, lds);
}

// --------------------------------------------------------------------------
// Per-frequency Sherman-Morrison solve, in place on F (one thread per
// frequency).  b_m = conj(D_m) S + rho A_m; delta_m = rho (plain) or
// w_m mu GHG + rho (GradReg);
//   x_m = (b_m - conj(D_m) [sum_j D_j b_j/delta_j] / [1 + sum_j |D_j|^2/delta_j]) / delta_m
// (cbpdn_kernels.cu:15-117 and 1198-1337 with C from 457-526 / 1457-1562;
// the coefficients are formed from the current rho, never stored).
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_solve_generic(Geom g, Buffers b,
                                                        Params p) {
  if (b.ctrl->done) return;
  const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (f >= g.Hc) return;
  const float rho = b.ctrl->rho;
  const float2 sh = b.Sf[f];
  const float gh = p.grad ? p.mu * b.ghg[f] : 0.f;
  float2 num = make_float2(0.f, 0.f);
  float den = 0.f;
  for (int m = 0; m < g.M; ++m) {
    const long long o = m * g.Hc + f;
    const float2 d = b.Df[o];
    const float2 a = b.F[o];
    const float2 bb = cadd(cmul(cconj(d), sh), cscale(a, rho));
    float inv;
    if (p.grad)
      inv = 1.f / ((p.gw_mode ? b.gw[m] : p.gw_scalar) * gh + rho);
    else
      inv = 1.f / rho;
    num = cadd(num, cscale(cmul(d, bb), inv));
    den += (d.x * d.x + d.y * d.y) * inv;
  }
  const float2 cb = cscale(num, 1.f / (1.f + den));
  for (int m = 0; m < g.M; ++m) {
    const long long o = m * g.Hc + f;
    const float2 d = b.Df[o];
    const float2 a = b.F[o];
    const float2 bb = cadd(cmul(cconj(d), sh), cscale(a, rho));
    float delta;
    if (p.grad)
      delta = (p.gw_mode ? b.gw[m] : p.gw_scalar) * gh + rho;
    else
      delta = rho;
    const float2 x = csub(bb, cmul(cconj(d), cb));
    b.F[o] = make_float2(x.x / delta, x.y / delta);
  }
}

// --------------------------------------------------------------------------
// Row kernel: inverse row FFT -> X, fused relax + shrink + AMS + dual update
// + residual partial sums, then (fused mode) forward row FFT of Y - U.
// --------------------------------------------------------------------------
// 4 waves per SIMD (at most 128 VGPRs, a few spilled): 8-12 % faster than
// the unconstrained allocation's 3 on uneven sizes.
template <bool WAVE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void
k_row_generic(Geom g, FftPlan rp,
                                                      Buffers b, Params p,
                                                      int mode, int nr) {
  if (b.ctrl->done) return;
  extern __shared__ float2 lds[];
  __shared__ float red[5][4];
  if constexpr (WAVE)
    row_body_wave(g, rp, b, p, mode, nr, blockIdx.x, gridDim.x, lds,
                  row_wave_table(g, lds, b.twC), red);
  else
    row_body(g, rp, b, p, mode, nr, blockIdx.x, gridDim.x, lds, red);
}

// --------------------------------------------------------------------------
// Y = shrink(V) of every coefficient of a state array holding V
// (Params::vstate), with the final rho; out may be the state array.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_state_to_y(Geom g, Buffers b,
                                                    Params p, float* out) {
  const float lr = p.lambda / b.ctrl->rho;
  const long long n = (long long)g.M * g.N;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const int m = (int)(i / g.N);
    const long long rc = i - (long long)m * g.N;
    const int r = (int)(rc / g.C);
    const int c = (int)(rc - (long long)r * g.C);
    out[i] = shrink_coef(g, p, b, b.Y[i], lr, m, r, c, i);
  }
}

// --------------------------------------------------------------------------
// Residual norms, stopping test and AutoRho.  Two deterministic passes:
// k_reduce_blocks sums contiguous chunks of the partial sums in double (one
// result per block), k_reduce_final adds the block results in index order
// and runs the control logic.  The kernel boundary between them makes the
// block results visible on every XCD.
// --------------------------------------------------------------------------
constexpr int kReduceBlocks = 64;  // <= 64: one lane each in the final pass
constexpr int kReduceThreads = 1024;

__global__ __launch_bounds__(kReduceThreads) void k_reduce_blocks(
    Buffers b, int nparts) {
  if (b.ctrl->done) return;
  __shared__ double red[5][kReduceThreads / 64];
  const int chunk = (nparts + gridDim.x - 1) / gridDim.x;
  const int i0 = blockIdx.x * chunk;
  const int i1 = min(nparts, i0 + chunk);
  double acc[5] = {0, 0, 0, 0, 0};
  const int tid = threadIdx.x;
  for (int i = i0 + tid; i < i1; i += kReduceThreads)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[j] += (double)b.partials[j * nparts + i];
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    double v = acc[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[j][wv] = v;
  }
  __syncthreads();
  if (tid < 5) {
    double v = 0;
    for (int w = 0; w < kReduceThreads / 64; ++w) v += red[tid][w];
    b.blocksum[blockIdx.x * 5 + tid] = v;
  }
}

// One wavefront: lane k loads block k's five sums, a fixed shuffle tree
// adds them (same order every call), lane 0 runs the control logic.
__global__ __launch_bounds__(64) void k_reduce_final(Buffers b, Params p,
                                                     int nblk, int iter) {
  if (b.ctrl->done) return;
  const int k = threadIdx.x;
  double s5[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    double v = k < nblk ? b.blocksum[k * 5 + j] : 0.0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    s5[j] = v;
  }
  if (k == 0) control_update(b, p, s5, iter);
}

// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
int row_blocks(const Geom& g) {
  const int nr = row_nr(g);
  const long long nrows = (long long)g.M * g.R;
  return (int)((nrows + nr - 1) / nr);
}

hipError_t launch_pad(const float* D, float* out, int M, int Kh, int Kw,
                      int R, int C, hipStream_t st) {
  const long long n = (long long)M * R * C;
  long long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_pad, dim3((unsigned)nb), dim3(256), 0, st, D, out, M,
                     Kh, Kw, R, C);
  return hipGetLastError();
}

hipError_t launch_row_fwd(const Geom& g, int nmaps, const FftPlan& rp,
                          int mode, const float* src, float* Y, float* U,
                          float2* F, const float2* twC, const Ctrl* ctrl,
                          hipStream_t st, int iter) {
  const int nr = row_nr(g);
  const long long nrows = (long long)nmaps * g.R;
  const unsigned nb = (unsigned)((nrows + nr - 1) / nr);
  const size_t shm = row_lds(g, nr);
  if (row_wave_fits(g))
    hipLaunchKernelGGL(k_row_fwd_generic<true>, dim3(nb), dim3(256), shm, st,
                       g, nrows, rp, mode, src, Y, U, F, twC, ctrl, nr, iter);
  else
    hipLaunchKernelGGL(k_row_fwd_generic<false>, dim3(nb), dim3(256), shm, st,
                       g, nrows, rp, mode, src, Y, U, F, twC, ctrl, nr, iter);
  return hipGetLastError();
}

hipError_t launch_row_inv(const Geom& g, int nmaps, const FftPlan& rp,
                          float2* F, float* out, const float2* twC,
                          float scale, hipStream_t st) {
  const int nr = row_nr_budget(g, 61440);
  const long long nrows = (long long)nmaps * g.R;
  const unsigned nb = (unsigned)((nrows + nr - 1) / nr);
  const size_t shm = (size_t)2 * nr * gen::gld(g.Lr) * sizeof(float2);
  hipLaunchKernelGGL(k_row_inv_generic, dim3(nb), dim3(256), shm, st, g,
                     nrows, rp, F, out, twC, scale, nr);
  return hipGetLastError();
}

hipError_t launch_col(const Geom& g, int nmaps, const FftPlan& cp, float2* F,
                      const float2* twR, int inverse, const Ctrl* ctrl,
                      hipStream_t st, const float* parts, int nparts,
                      double* bsum) {
  const size_t shm = (size_t)2 * g.TC * gen::gld(g.R) * sizeof(float2);
  const int nred = nparts > 0 ? pow2_defer_blocks(g, nmaps) : 0;
  if (row_ct_enabled() && col_ct_supported(g.R))
    return launch_col_ct(g, nmaps, F, twR, inverse, ctrl, parts, nparts, bsum,
                         nred, shm, st);
  hipLaunchKernelGGL(k_col_generic, dim3(g.P, nmaps), dim3(256), shm, st, g,
                     cp, F, twR, inverse ? -1.f : 1.f, ctrl, parts, nparts,
                     bsum, nred);
  return hipGetLastError();
}

hipError_t launch_solve(const Geom& g, const Buffers& b, const Params& p,
                        hipStream_t st, int ctl_iter, int nbsum) {
  if (solve_regs_fits(g.M)) return solve_regs(g, b, p, st, ctl_iter, nbsum);
  if (ctl_iter > 0) return hipErrorInvalidValue;  // no deferral here
  const unsigned nb = (unsigned)((g.Hc + 255) / 256);
  hipLaunchKernelGGL(k_solve_generic, dim3(nb), dim3(256), 0, st, g, b, p);
  return hipGetLastError();
}

hipError_t launch_row(const Geom& g, const FftPlan& rp, const Buffers& b,
                      const Params& p, int mode, int nblocks_out[1],
                      hipStream_t st) {
  const int nr = row_nr(g);
  const int nb = row_blocks(g);
  const size_t shm = row_lds(g, nr);
  nblocks_out[0] = nb;
  if (mode == kRowFused && g.even && p.vstate && p.m0 == 0 &&
      p.l1_mode != 2 && !p.nonneg && !p.nobndry && row_wave_fits(g) &&
      row_ct_enabled() && row_ct_supported(g.Lr))
    return launch_row_ct(g, b, p, nr, nb, shm, st);
  if (row_wave_fits(g))
    hipLaunchKernelGGL(k_row_generic<true>, dim3(nb), dim3(256), shm, st, g,
                       rp, b, p, mode, nr);
  else
    hipLaunchKernelGGL(k_row_generic<false>, dim3(nb), dim3(256), shm, st, g,
                       rp, b, p, mode, nr);
  nblocks_out[0] = nb;
  return hipGetLastError();
}

// flag[0] = 1 when any of the n values is negative (flag cleared first by
// the single-block launch below, or preset by the caller).
__global__ __launch_bounds__(256) void k_any_negative(const float* w,
                                                      long long n, int* flag) {
  int neg = 0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    neg |= w[i] < 0.f;
  if (__any(neg) && (threadIdx.x & 63) == 0) flag[0] = 1;
}

__global__ void k_clear_flag(int* flag) { flag[0] = 0; }

__global__ __launch_bounds__(256) void k_any_nonbinary(const int* mask,
                                                       long long n,
                                                       int* flag) {
  int bad = 0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    bad |= (unsigned)mask[i] > 1u;
  if (__any(bad) && (threadIdx.x & 63) == 0) flag[0] = 1;
}

hipError_t launch_any_nonbinary(const int* mask, long long n, int* flag,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, st, flag);
  const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_any_nonbinary, dim3(nb), dim3(256), 0, st, mask, n,
                     flag);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_twiddles(float2* tw, int L) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= L) return;
  const double a = -2.0 * M_PI * (double)t / (double)L;
  tw[t] = make_float2((float)cos(a), (float)sin(a));
}

hipError_t launch_twiddles(float2* tw, int L, hipStream_t st) {
  hipLaunchKernelGGL(k_twiddles, dim3((L + 255) / 256), dim3(256), 0, st, tw,
                     L);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_ghg(Geom g, float* ghg) {
  const long long f = blockIdx.x * 256LL + threadIdx.x;
  if (f >= g.Hc) return;
  const int r = (int)(f / g.H);
  const int c = (int)(f - (long long)r * g.H);
  const double v = 4.0 - 2.0 * cos(2.0 * M_PI * r / g.R) -
                   2.0 * cos(2.0 * M_PI * c / g.C);
  ghg[panel_off(g, r, c)] = (float)v;
}

hipError_t launch_ghg(const Geom& g, float* ghg, hipStream_t st) {
  hipLaunchKernelGGL(k_ghg, dim3((unsigned)((g.Hc + 255) / 256)), dim3(256),
                     0, st, g, ghg);
  return hipGetLastError();
}

// Zero fill of n 32-bit words (a kernel rather than a memset node, so that
// the clear can sit in a captured graph).  Four words per store where the
// pointer is 16-byte aligned.
__global__ __launch_bounds__(256) void k_zero(unsigned* p, long long n) {
  const long long stride = (long long)gridDim.x * 256;
  long long i = blockIdx.x * 256LL + threadIdx.x;
  if (((uintptr_t)p & 15) == 0) {
    uint4* q = reinterpret_cast<uint4*>(p);
    const long long n4 = n >> 2;
    for (long long k = i; k < n4; k += stride) q[k] = make_uint4(0, 0, 0, 0);
    for (long long k = (n4 << 2) + i; k < n; k += stride) p[k] = 0u;
  } else {
    for (long long k = i; k < n; k += stride) p[k] = 0u;
  }
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t st) {
  const long long n = (long long)(bytes / 4);
  if (n == 0) return hipSuccess;
  const unsigned nb =
      (unsigned)std::min<long long>((n / 4 + 255) / 256 + 1, 8192);
  hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, st, (unsigned*)p, n);
  return hipGetLastError();
}

hipError_t launch_any_negative(const float* w, long long n, int* flag,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, st, flag);
  const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_any_negative, dim3(nb), dim3(256), 0, st, w, n, flag);
  return hipGetLastError();
}

hipError_t launch_state_to_y(const Geom& g, const Buffers& b,
                             const Params& p, float* out, hipStream_t st) {
  const long long n = (long long)g.M * g.N;
  const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_state_to_y, dim3(nb), dim3(256), 0, st, g, b, p, out);
  return hipGetLastError();
}

hipError_t launch_reduce(const Buffers& b, const Params& p, int nblocks,
                         int iter, int /*max_iter*/, hipStream_t st) {
  const int nblk = nblocks < kReduceBlocks ? 1 : kReduceBlocks;
  hipLaunchKernelGGL(k_reduce_blocks, dim3(nblk), dim3(kReduceThreads), 0, st,
                     b, nblocks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, st, b, p, nblk,
                     iter);
  return hipGetLastError();
}

}  // namespace scc
