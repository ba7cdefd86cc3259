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

#include "scc_internal.h"
#include "scc_device.h"

namespace scc {

// --------------------------------------------------------------------------
// Mixed-radix Stockham FFT over nt transforms of length pl.L held in LDS at
// A + t*ld.  Ping-pongs between A and B and returns the buffer holding the
// result.  dir = +1 forward (W = exp(-2 pi i / L)), -1 inverse (unnormalised).
// Every thread of the block must call it (it contains block barriers).
// --------------------------------------------------------------------------
__device__ float2* fft_generic(float2* A, float2* B, int nt, int ld,
                               const FftPlan& pl,
                               const float2* __restrict__ tw, float dir,
                               int tid, int nthr) {
  const int L = pl.L;
  const int st = pl.tw_stride;
  int Ns = 1;
  for (int s = 0; s < pl.nstages; ++s) {
    const int Rx = pl.radix[s];
    const int nb = L / Rx;
    const int tstep = L / (Ns * Rx);
    const int total = nt * nb;
    for (int idx = tid; idx < total; idx += nthr) {
      const int t = idx / nb;
      const int j = idx - t * nb;
      const float2* in = A + t * ld;
      float2* out = B + t * ld;
      const int k = j % Ns;
      const int ob = (j - k) * Rx + k;
      if (Rx == 2) {
        float2 x0 = in[j];
        float2 x1 = in[j + nb];
        if (k) x1 = cmul(x1, twiddle(tw, k * tstep * st, dir));
        out[ob] = cadd(x0, x1);
        out[ob + Ns] = csub(x0, x1);
      } else if (Rx == 4) {
        float2 x0 = in[j];
        float2 x1 = in[j + nb];
        float2 x2 = in[j + 2 * nb];
        float2 x3 = in[j + 3 * nb];
        if (k) {
          x1 = cmul(x1, twiddle(tw, k * tstep * st, dir));
          x2 = cmul(x2, twiddle(tw, 2 * k * tstep * st, dir));
          x3 = cmul(x3, twiddle(tw, 3 * k * tstep * st, dir));
        }
        float2 s02 = cadd(x0, x2), d02 = csub(x0, x2);
        float2 s13 = cadd(x1, x3), d13 = csub(x1, x3);
        float2 jd = mul_mi(d13, dir);  // -i*dir*d13
        out[ob] = cadd(s02, s13);
        out[ob + Ns] = cadd(d02, jd);
        out[ob + 2 * Ns] = csub(s02, s13);
        out[ob + 3 * Ns] = csub(d02, jd);
      } else {
        for (int q = 0; q < Rx; ++q) {
          const int cq = k * tstep + q * nb;  // < L
          float2 acc = make_float2(0.f, 0.f);
          int e = 0;
          for (int r = 0; r < Rx; ++r) {
            acc = cadd(acc, cmul(in[j + r * nb], twiddle(tw, e * st, dir)));
            e += cq;
            if (e >= L) e -= L;
          }
          out[ob + q * Ns] = acc;
        }
      }
    }
    __syncthreads();
    float2* tmp = A;
    A = B;
    B = tmp;
    Ns *= Rx;
  }
  return A;
}

// Element i of the (nt rows x H columns) half-spectrum block of rows
// row0..row0+nt-1, enumerated panel by panel so that consecutive i touch
// consecutive addresses when the rows are consecutive rows of one map.
__device__ inline void panel_index(const Geom& g, int nt, int i, int& t,
                                   int& c) {
  const int full = (g.P - 1) * nt * g.TC;
  if (i < full) {
    const int per = nt * g.TC;
    const int p = i / per;
    const int rem = i - p * per;
    t = rem / g.TC;
    c = p * g.TC + (rem - t * g.TC);
  } else {
    const int rem = i - full;
    t = rem / g.wlast;
    c = (g.P - 1) * g.TC + (rem - t * g.wlast);
  }
}

__device__ inline long long spec_addr(const Geom& g, long long row, int c) {
  const long long m = row / g.R;
  const int r = (int)(row - m * g.R);
  return m * g.Hc + panel_off(g, r, c);
}

// Inverse-real-FFT pre-processing of one block of half spectra (in B) into
// the complex input of the length-Lr inverse transform (in A).
__device__ void c2r_prep(const Geom& g, const float2* B, float2* A, int nt,
                         int ld, const float2* __restrict__ twC, int tid,
                         int nthr) {
  if (g.even) {
    const int L = g.Lr;
    for (int i = tid; i < nt * L; i += nthr) {
      const int t = i / L;
      const int k = i - t * L;
      float2 xk = B[t * ld + k];
      float2 xn = B[t * ld + L - k];
      if (k == 0) {  // DC and Nyquist are real (C2R ignores their imag)
        xk.y = 0.f;
        xn.y = 0.f;
      }
      xn.y = -xn.y;
      const float2 s = cadd(xk, xn);
      const float2 wd = cmul(cconj(twC[k]), csub(xk, xn));
      A[t * ld + k] = make_float2(s.x - wd.y, s.y + wd.x);  // s + i*wd
    }
  } else {
    const int C = g.C;
    for (int i = tid; i < nt * C; i += nthr) {
      const int t = i / C;
      const int k = i - t * C;
      float2 v = (k < g.H) ? B[t * ld + k] : cconj(B[t * ld + C - k]);
      if (k == 0) v.y = 0.f;
      A[t * ld + k] = v;
    }
  }
}

// Forward-real-FFT post-processing of the length-Lr transform (in Z) and
// store of the half spectra to F in panel layout.
__device__ void r2c_store(const Geom& g, const float2* Z, int nt, int ld,
                          long long row0, float2* __restrict__ F,
                          const float2* __restrict__ twC, int tid, int nthr) {
  const int L = g.Lr;
  for (int i = tid; i < nt * g.H; i += nthr) {
    int t, k;
    panel_index(g, nt, i, t, k);
    float2 X;
    if (g.even) {
      const float2 zk = Z[t * ld + (k == L ? 0 : k)];
      const float2 zn = cconj(Z[t * ld + (k == 0 ? 0 : L - k)]);
      const float2 s = cadd(zk, zn);
      const float2 wd = cmul(twC[k], csub(zk, zn));
      X = make_float2(0.5f * (s.x + wd.y), 0.5f * (s.y - wd.x));
    } else {
      X = Z[t * ld + k];
    }
    F[spec_addr(g, row0 + t, k)] = X;
  }
}

__device__ void spec_load(const Geom& g, const float2* __restrict__ F,
                          float2* B, int nt, int ld, long long row0, int tid,
                          int nthr) {
  for (int i = tid; i < nt * g.H; i += nthr) {
    int t, c;
    panel_index(g, nt, i, t, c);
    B[t * ld + c] = F[spec_addr(g, row0 + t, c)];
  }
}

static inline int row_nr(const Geom& g) {
  const int ld = g.Lr + 1;
  int nr = 61440 / (2 * ld * 8);
  if (nr > 16) nr = 16;
  if (nr < 1) nr = 1;
  return nr;
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
__global__ __launch_bounds__(256) void k_row_fwd_generic(
    Geom g, long long nrows, FftPlan rp, int mode, const float* src,
    float* Y, float* U, float2* F, const float2* __restrict__ twC,
    const Ctrl* ctrl, int nr) {
  if (mode == kFwdYU && ctrl->done) return;
  extern __shared__ float2 lds[];
  const int ld = g.Lr + 1;
  float2* A = lds;
  float2* B = lds + nr * ld;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long row0 = (long long)blockIdx.x * nr;
  const int nt = (int)min((long long)nr, nrows - row0);
  const float rsf = (mode == kFwdYU) ? ctrl->rsf : 1.f;
  for (int i = tid; i < nt * g.C; i += nthr) {
    const int t = i / g.C;
    const int n = i - t * g.C;
    const long long gi = (row0 + t) * g.C + n;
    float a;
    if (mode == kFwdReal) {
      a = src[gi];
    } else {
      float u = U[gi];
      if (rsf != 1.f) {
        u = u / rsf;
        U[gi] = u;
      }
      a = Y[gi] - u;
    }
    if (g.even)
      reinterpret_cast<float*>(A + t * ld)[n] = a;
    else
      A[t * ld + n] = make_float2(a, 0.f);
  }
  __syncthreads();
  const float2* Z = fft_generic(A, B, nt, ld, rp, twC, 1.f, tid, nthr);
  r2c_store(g, Z, nt, ld, row0, F, twC, tid, nthr);
}

// --------------------------------------------------------------------------
// Column transforms, one panel of one map per block, in place.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_col_generic(
    Geom g, FftPlan cp, float2* F, const float2* __restrict__ twR,
    float dir, const Ctrl* ctrl) {
  if (ctrl != nullptr && ctrl->done) return;
  extern __shared__ float2 lds[];
  const int p = blockIdx.x;
  const long long m = blockIdx.y;
  const int w = (p == g.P - 1) ? g.wlast : g.TC;
  const int ld = g.R + 1;
  float2* A = lds;
  float2* B = lds + w * ld;
  float2* base = F + m * g.Hc + (long long)p * g.TC * g.R;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int n = g.R * w;
  for (int i = tid; i < n; i += nthr) {
    const int r = i / w;
    const int c = i - r * w;
    A[c * ld + r] = base[i];
  }
  __syncthreads();
  const float2* Z = fft_generic(A, B, w, ld, cp, twR, dir, tid, nthr);
  for (int i = tid; i < n; i += nthr) {
    const int r = i / w;
    const int c = i - r * w;
    base[i] = Z[c * ld + r];
  }
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
__global__ __launch_bounds__(256) void k_row_generic(Geom g, FftPlan rp,
                                                      Buffers b, Params p,
                                                      int mode, int nr) {
  if (b.ctrl->done) return;
  extern __shared__ float2 lds[];
  __shared__ float red[5][4];
  const int ld = g.Lr + 1;
  float2* A = lds;
  float2* B = lds + nr * ld;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long nrows = (long long)g.M * g.R;
  const long long row0 = (long long)blockIdx.x * nr;
  const int nt = (int)min((long long)nr, nrows - row0);
  const float rho = b.ctrl->rho;

  spec_load(g, b.F, B, nt, ld, row0, tid, nthr);
  __syncthreads();
  c2r_prep(g, B, A, nt, ld, b.twC, tid, nthr);
  __syncthreads();
  float2* P = fft_generic(A, B, nt, ld, rp, b.twC, -1.f, tid, nthr);
  float2* Q = (P == A) ? B : A;

  UpdateSums acc;
  const float invN = 1.0f / (float)g.N;
  const float lr = p.lambda / rho;
  for (int i = tid; i < nt * g.C; i += nthr) {
    const int t = i / g.C;
    const int n = i - t * g.C;
    const long long row = row0 + t;
    const int m = (int)(row / g.R);
    const int r = (int)(row - (long long)m * g.R);
    const float xv = g.even ? reinterpret_cast<const float*>(P + t * ld)[n]
                            : P[t * ld + n].x;
    const long long gi = row * g.C + n;
    float a = update_point(g, p, b, xv * invN, lr, m, r, n, gi, acc);
    if (mode == kRowFused) {
      if (g.even)
        reinterpret_cast<float*>(P + t * ld)[n] = a;
      else
        P[t * ld + n] = make_float2(a, 0.f);
    }
  }
  block_sums_store(acc, red, b.partials, gridDim.x, blockIdx.x, tid);
  if (mode == kRowFused) {
    __syncthreads();
    const float2* Z = fft_generic(P, Q, nt, ld, rp, b.twC, 1.f, tid, nthr);
    r2c_store(g, Z, nt, ld, row0, b.F, b.twC, tid, nthr);
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
                          hipStream_t st) {
  const int nr = row_nr(g);
  const long long nrows = (long long)nmaps * g.R;
  const unsigned nb = (unsigned)((nrows + nr - 1) / nr);
  const size_t shm = (size_t)2 * nr * (g.Lr + 1) * sizeof(float2);
  hipLaunchKernelGGL(k_row_fwd_generic, dim3(nb), dim3(256), shm, st, g,
                     nrows, rp, mode, src, Y, U, F, twC, ctrl, nr);
  return hipGetLastError();
}

hipError_t launch_col(const Geom& g, int nmaps, const FftPlan& cp, float2* F,
                      const float2* twR, int inverse, const Ctrl* ctrl,
                      hipStream_t st) {
  const size_t shm = (size_t)2 * g.TC * (g.R + 1) * sizeof(float2);
  hipLaunchKernelGGL(k_col_generic, dim3(g.P, nmaps), dim3(256), shm, st, g,
                     cp, F, twR, inverse ? -1.f : 1.f, ctrl);
  return hipGetLastError();
}

hipError_t launch_solve(const Geom& g, const Buffers& b, const Params& p,
                        hipStream_t st) {
  const unsigned nb = (unsigned)((g.Hc + 255) / 256);
  hipLaunchKernelGGL(k_solve_generic, dim3(nb), dim3(256), 0, st, g, b, p);
  return hipGetLastError();
}

hipError_t launch_row(const Geom& g, const FftPlan& rp, const Buffers& b,
                      const Params& p, int mode, int nblocks_out[1],
                      hipStream_t st) {
  const int nr = row_nr(g);
  const int nb = row_blocks(g);
  const size_t shm = (size_t)2 * nr * (g.Lr + 1) * sizeof(float2);
  hipLaunchKernelGGL(k_row_generic, dim3(nb), dim3(256), shm, st, g, rp, b,
                     p, mode, nr);
  nblocks_out[0] = nb;
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
