// Differentiable unrolled plain CBPDN (power-of-two images): forward with
// iterate checkpoints, and the hand-written adjoint of every stage.
//
// Forward: the solver's own kernels over B images of one size sharing the
// dictionary (column and row kernels run over B*M maps, the solve takes the
// image index), fixed rho, no stopping test; (Y, U) is saved every k
// iterations, and the tape of the last three segments is kept with the
// checkpoints.  The tape of an iteration is X-hat after the solve (written
// by the solve kernel) and the soft-threshold support of V = Xr + U as 4
// bits per coefficient pair (|V| > t and sign V, written by the row
// kernel).  Backward, per checkpoint segment from last to first: a segment
// without a kept tape is recomputed from its checkpoint with a tape, then
// every iteration is reversed:
//   k_adj_row1   shrink backward (support mask), relaxation backward, and the
//                forward row transform of Xbar = alpha Vbar
//   k_col_pow2   forward column transform            (adjoint of the inverse FFT)
//   k_adj_solve  Sherman-Morrison solve of the adjoint right-hand side, the
//                dictionary- and image-spectrum gradients, rho Bbar / w
//   k_col_pow2   inverse column transform            (adjoint of the forward FFT)
//   k_adj_row2   inverse row transform, Ybar += Abar, Ubar -= Abar
// Inside a segment k_adj_row21 fuses k_adj_row2 of one iteration with
// k_adj_row1 of the one before it, so Ybar and Ubar are read and written
// once per iteration.
// The FFT adjoints reuse the forward kernels; the half-spectrum column
// multiplicity w (1 for DC / Nyquist, 2 otherwise) and 1/N are applied in
// k_adj_solve.  Lambda's gradient is a deterministic FP64 sum of
// per-wavefront partials.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sporco_amd.h"
#include "pow2_common.h"
#include "scc_pow2.h"

namespace scc {

// --------------------------------------------------------------------------
// Adjoint kernels
// --------------------------------------------------------------------------

// Shrink / dual / relaxation backward of one lane's slots, in place on
// (yb, ub), from its support bits: Vbar = Ubar + (Ybar - Ubar) 1[|V| > t];
// tbar += (Ubar - Ybar) sign(V) 1[|V| > t]; Xbar = a Vbar (returned in z);
// Ybar <- (1 - a) Vbar; Ubar <- Vbar.
template <int E>
__device__ __forceinline__ void shrink_relax_bwd(
    float2 (&yb)[E], float2 (&ub)[E], float2 (&z)[E],
    const unsigned (&tb)[TapeWords<E>::W], float alpha, float& lacc) {
  const float a1 = 1.f - alpha;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const unsigned q = (tb[e / 8] >> (4 * (e % 8))) & 15u;
    const bool m0 = q & 1u, m1 = q & 4u;
    const float vb0 = m0 ? yb[e].x : ub[e].x;
    const float vb1 = m1 ? yb[e].y : ub[e].y;
    if (m0) lacc += (q & 2u) ? yb[e].x - ub[e].x : ub[e].x - yb[e].x;
    if (m1) lacc += (q & 8u) ? yb[e].y - ub[e].y : ub[e].y - yb[e].y;
    yb[e] = make_float2(a1 * vb0, a1 * vb1);
    ub[e] = make_float2(vb0, vb1);
    z[e] = make_float2(alpha * vb0, alpha * vb1);
  }
}

// Shrink / dual / relaxation backward on one row (support bits of the
// iteration in tbits; Ybar read from Ybin when given, else Yb), then the
// forward real FFT of Xbar into the row spectra of F; per-wavefront
// partial sums of tbar in lpart.
template <int LR, int TC>
__global__ __launch_bounds__(256) void k_adj_row1(Geom g, const float* Ybin,
                                                  float* Yb, float* Ub,
                                                  const unsigned* tbits,
                                                  float alpha, float2* F,
                                                  const float2* twC,
                                                  float* lpart) {
  using RC = RowCfg<LR>;
  constexpr int E = RC::E, T = RC::T, RPW = RC::RPW;
  constexpr int TW = TapeWords<E>::W;
  constexpr int C = 2 * LR;
  constexpr int LP = reg::lds_len(LR);
  __shared__ float2 lds[4 * RPW * LP];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int rw = lane / T;
  const int t = lane & (T - 1);
  const int partner = rw * T + ((T - t) & (T - 1));
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
  const int row0 = wave * RPW;
  const int R = g.R;
  const int m = row0 >> g.lgR;
  const int r = (row0 & (R - 1)) + rw;
  float2* L = lds + (wv * RPW + rw) * LP;
  const long long base = (long long)row0 * C;
  float2* Y2 = reinterpret_cast<float2*>(Yb + base);
  float2* U2 = reinterpret_cast<float2*>(Ub + base);
  const int lo = rw * LR + t;
  unsigned tb[TW];
  const unsigned* tw = tbits + ((long long)(row0 + rw) * T + t) * TW;
#pragma unroll
  for (int w = 0; w < TW; ++w) tb[w] = tw[w];
  // Ybar is read from Ybin when given (dL/dY of the last iterate).
  const float2* Y2i =
      Ybin ? reinterpret_cast<const float2*>(Ybin + base) : Y2;
  float2 yb[E], ub[E], z[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    yb[e] = Y2i[lo + T * e];
    ub[e] = U2[lo + T * e];
  }
  float lacc = 0.f;
  shrink_relax_bwd<E>(yb, ub, z, tb, alpha, lacc);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    Y2[lo + T * e] = yb[e];
    U2[lo + T * e] = ub[e];
  }
  const float ls = wave_sum(lacc);
  if (lane == 0) lpart[wave] = ls;
  reg::fft_wave<LR, E, 1>(z, t, L, twC, 2);
  store_spectrum<LR, TC>(z, t, partner, r, R, F + (long long)m * g.Hc, twC);
}

// k_adj_row2 of one iteration fused with k_adj_row1 of the previous one:
// Abar = inverse row FFT of F; (Ybar, Ubar) += (Abar, -Abar) in registers;
// then the shrink / relaxation backward with the previous iteration's
// support bits and the forward row FFT of Xbar back into F (each row's
// spectrum is read and written by the same wavefront).
template <int LR, int TC>
__global__ __launch_bounds__(256) void k_adj_row21(Geom g, float2* F,
                                                   float* Yb, float* Ub,
                                                   const unsigned* tbits,
                                                   float alpha,
                                                   const float2* twC,
                                                   float* lpart) {
  using RC = RowCfg<LR>;
  constexpr int E = RC::E, T = RC::T, RPW = RC::RPW;
  constexpr int TW = TapeWords<E>::W;
  constexpr int C = 2 * LR;
  constexpr int LP = reg::lds_len(LR);
  __shared__ float2 lds[4 * RPW * LP];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int rw = lane / T;
  const int t = lane & (T - 1);
  const int partner = rw * T + ((T - t) & (T - 1));
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
  const int row0 = wave * RPW;
  const int R = g.R;
  const int m = row0 >> g.lgR;
  const int r = (row0 & (R - 1)) + rw;
  float2* L = lds + (wv * RPW + rw) * LP;
  float2* Fm = F + (long long)m * g.Hc;
  const long long base = (long long)row0 * C;
  float2* Y2 = reinterpret_cast<float2*>(Yb + base);
  float2* U2 = reinterpret_cast<float2*>(Ub + base);
  const int lo = rw * LR + t;
  float2 z[E], yb[E], ub[E];
  load_spectrum_c2r<LR, TC>(z, t, partner, r, R, Fm, twC);
  unsigned tb[TW];
  const unsigned* tw = tbits + ((long long)(row0 + rw) * T + t) * TW;
#pragma unroll
  for (int w = 0; w < TW; ++w) tb[w] = tw[w];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    yb[e] = Y2[lo + T * e];
    ub[e] = U2[lo + T * e];
  }
  reg::fft_wave<LR, E, -1>(z, t, L, twC, 2);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    yb[e] = cadd(yb[e], z[e]);
    ub[e] = csub(ub[e], z[e]);
  }
  float lacc = 0.f;
  shrink_relax_bwd<E>(yb, ub, z, tb, alpha, lacc);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    Y2[lo + T * e] = yb[e];
    U2[lo + T * e] = ub[e];
  }
  const float ls = wave_sum(lacc);
  if (lane == 0) lpart[wave] = ls;
  reg::wave_sync();
  reg::fft_wave<LR, E, 1>(z, t, L, twC, 2);
  store_spectrum<LR, TC>(z, t, partner, r, R, Fm, twC);
}

// Adjoint X-solve.  Per frequency, with D the dictionary spectra, Xh the
// forward X-hat of the iteration (tape) and Xhbar = w F / N:
//   Bbar = (rho I + D^H D)^-1 Xhbar                  (Sherman-Morrison)
//   Dbar_m += -conj(Xh_m) sum_j D_j Bbar_j - conj(Bbar_m) sum_j D_j Xh_j
//             + conj(Bbar_m) S
//   Sbar   += sum_j D_j Bbar_j
//   F_m     = rho Bbar_m / w
// G lanes per frequency, MPT maps per lane.  A thread handles one
// frequency for a group of up to kImgGroup images (blockIdx.y), looped
// inside with Dbar accumulated in registers and added to that group's own
// accumulator Dhb + group M Hc (summed over groups, in order, by
// k_sum_groups after the last iteration).  Every load of an image step (F,
// X-hat, S, the running Sbar) is issued before its arithmetic, so a step
// waits for memory once.
constexpr int kImgGroup = 16;

template <int G, int MPT>
__global__ __launch_bounds__(256) void k_adj_solve(
    Geom g, int nimg, float2* F, const float2* __restrict__ Xh,
    const float2* __restrict__ Df, const float2* __restrict__ Sf,
    float2* Dhb, float2* Shb, float rho, float invN) {
  constexpr int FPW = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int gi = lane / FPW;
  const int fl = lane - gi * FPW;
  const long long f =
      ((long long)blockIdx.x * (blockDim.x >> 6) + wv) * FPW + fl;
  const bool valid = f < g.Hc;
  const int M = g.M;
  const float irho = 1.f / rho;
  const float wt = valid ? col_mult(g, panel_col(g, f)) : 1.f;
  const float sin_ = wt * invN;
  const float sout = rho / wt;
  float2 dd[MPT], dacc[MPT];
  float sig = 0.f;
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = gi * MPT + i;
    dacc[i] = make_float2(0.f, 0.f);
    dd[i] = make_float2(0.f, 0.f);
    if (valid && m < M) {
      dd[i] = Df[(long long)m * g.Hc + f];
      sig += dd[i].x * dd[i].x + dd[i].y * dd[i].y;
    }
  }
#pragma unroll
  for (int o = FPW; o < 64; o <<= 1) sig += __shfl_xor(sig, o, 64);
  const float isig = 1.f / (rho + sig);
  const int i0 = blockIdx.y * kImgGroup;
  const int i1 = min(nimg, i0 + kImgGroup);
  for (int img = i0; img < i1; ++img) {
    float2* Fi = F + (long long)img * M * g.Hc;
    const float2* Xi = Xh + (long long)img * M * g.Hc;
    float2 xb[MPT], xh[MPT];
    float2 sh = make_float2(0.f, 0.f), sb = make_float2(0.f, 0.f);
    if (valid) {
      sh = Sf[(long long)img * g.Hc + f];
      if (gi == 0) sb = Shb[(long long)img * g.Hc + f];
    }
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      const int m = gi * MPT + i;
      xb[i] = make_float2(0.f, 0.f);
      xh[i] = make_float2(0.f, 0.f);
      if (valid && m < M) {
        const long long o = (long long)m * g.Hc + f;
        xb[i] = Fi[o];
        xh[i] = Xi[o];
      }
    }
    float2 num = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      xb[i] = cscale(xb[i], sin_);
      num = cadd(num, cmul(dd[i], xb[i]));
    }
#pragma unroll
    for (int o = FPW; o < 64; o <<= 1) {
      num.x += __shfl_xor(num.x, o, 64);
      num.y += __shfl_xor(num.y, o, 64);
    }
    const float2 cb = cscale(num, isig);
    float2 sDB = make_float2(0.f, 0.f), sDX = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      xb[i] = cscale(csub(xb[i], cmul(cconj(dd[i]), cb)), irho);  // Bbar
      sDB = cadd(sDB, cmul(dd[i], xb[i]));
      sDX = cadd(sDX, cmul(dd[i], xh[i]));
    }
#pragma unroll
    for (int o = FPW; o < 64; o <<= 1) {
      sDB.x += __shfl_xor(sDB.x, o, 64);
      sDB.y += __shfl_xor(sDB.y, o, 64);
      sDX.x += __shfl_xor(sDX.x, o, 64);
      sDX.y += __shfl_xor(sDX.y, o, 64);
    }
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      const int m = gi * MPT + i;
      if (valid && m < M) {
        const float2 cbb = cconj(xb[i]);
        dacc[i] = cadd(dacc[i],
                       csub(cmul(cbb, csub(sh, sDX)), cmul(cconj(xh[i]), sDB)));
        Fi[(long long)m * g.Hc + f] = cscale(xb[i], sout);
      }
    }
    if (valid && gi == 0) Shb[(long long)img * g.Hc + f] = cadd(sb, sDB);
  }
  float2* Dg = Dhb + (long long)blockIdx.y * M * g.Hc;
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = gi * MPT + i;
    if (valid && m < M) {
      float2* d = Dg + (long long)m * g.Hc + f;
      *d = cadd(*d, dacc[i]);
    }
  }
}

// Dhb[0] += Dhb[1] + ... + Dhb[ngrp - 1] (n complex values each, in order).
__global__ void k_sum_groups(float2* Dhb, int ngrp, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    float2 a = Dhb[i];
    for (int k = 1; k < ngrp; ++k) a = cadd(a, Dhb[(long long)k * n + i]);
    Dhb[i] = a;
  }
}

// Unnormalised inverse real transform of every row's half spectrum, added
// to Ybar and subtracted from Ubar.
template <int LR, int TC>
__global__ __launch_bounds__(256) void k_adj_row2(Geom g, const float2* F,
                                                  float* Yb, float* Ub,
                                                  const float2* twC) {
  using RC = RowCfg<LR>;
  constexpr int E = RC::E, T = RC::T, RPW = RC::RPW;
  constexpr int C = 2 * LR;
  constexpr int LP = reg::lds_len(LR);
  __shared__ float2 lds[4 * RPW * LP];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int rw = lane / T;
  const int t = lane & (T - 1);
  const int partner = rw * T + ((T - t) & (T - 1));
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
  const int row0 = wave * RPW;
  const int R = g.R;
  const int m = row0 >> g.lgR;
  const int r = (row0 & (R - 1)) + rw;
  float2* L = lds + (wv * RPW + rw) * LP;
  const long long base = (long long)row0 * C;
  const int lo = rw * LR + t;
  float2 z[E];
  load_spectrum_c2r<LR, TC>(z, t, partner, r, R, F + (long long)m * g.Hc,
                            twC);
  reg::fft_wave<LR, E, -1>(z, t, L, twC, 2);
  float2* Y2 = reinterpret_cast<float2*>(Yb + base);
  float2* U2 = reinterpret_cast<float2*>(Ub + base);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float2 y = Y2[lo + T * e], u = U2[lo + T * e];
    Y2[lo + T * e] = cadd(y, z[e]);
    U2[lo + T * e] = csub(u, z[e]);
  }
}

// F /= column multiplicity, over nmaps maps of Hc values.
__global__ void k_scale_cols(Geom g, float2* F, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long f = i % g.Hc;
    F[i] = cscale(F[i], 1.f / col_mult(g, panel_col(g, f)));
  }
}

// dst (M, K, K) = src (M, R, C) restricted to the top-left K x K corner.
__global__ void k_crop(const float* src, float* dst, int M, int K, int R,
                       int C) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * K * K) return;
  const int m = i / (K * K), a = (i / K) % K, c = i % K;
  dst[i] = src[((long long)m * R + a) * C + c];
}

// Sum of n partials in two fixed-order FP64 passes: 64 blocks each add a
// contiguous chunk (k_lsum_blocks), one wavefront adds the 64 results
// (k_lsum_final) into *acc.
constexpr int kLsumBlocks = 64;

__global__ __launch_bounds__(1024) void k_lsum_blocks(const float* part,
                                                      long long n,
                                                      double* bsum) {
  __shared__ double red[16];
  const long long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long long i0 = blockIdx.x * chunk;
  const long long i1 = min(n, i0 + chunk);
  double v = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x)
    v += (double)part[i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    bsum[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(64) void k_lsum_final(const double* bsum,
                                                   double* acc) {
  double v = bsum[threadIdx.x];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (threadIdx.x == 0) *acc = v;
}

// --------------------------------------------------------------------------
// Launch helpers
// --------------------------------------------------------------------------
namespace {

#define SCC_UNR_CASES(X)                                               \
  X(128, 16) X(128, 8) X(128, 4) X(256, 16) X(256, 8) X(256, 4) X(512, 16) \
      X(512, 8) X(512, 4) X(1024, 16) X(1024, 8) X(1024, 4)

hipError_t adj_row1(const Geom& gbm, const Pow2Config& cfg, const float* Ybin,
                    float* Yb, float* Ub, const unsigned* tbits, float alpha,
                    float2* F, const float2* twC, float* lpart, int* nparts,
                    hipStream_t st) {
  const int LR = gbm.C / 2;
#define X(lr, tc)                                                          \
  if (LR == lr && cfg.TC == tc) {                                          \
    const unsigned nb = (unsigned)((long long)gbm.M * gbm.R /              \
                                   RowCfg<lr>::RPB);                        \
    hipLaunchKernelGGL((k_adj_row1<lr, tc>), dim3(nb), dim3(256), 0, st,   \
                       gbm, Ybin, Yb, Ub, tbits, alpha, F, twC, lpart);    \
    *nparts = (int)nb * 4;                                                 \
    return hipGetLastError();                                              \
  }
  SCC_UNR_CASES(X)
#undef X
  return hipErrorNotSupported;
}

hipError_t adj_row21(const Geom& gbm, const Pow2Config& cfg, float2* F,
                     float* Yb, float* Ub, const unsigned* tbits, float alpha,
                     const float2* twC, float* lpart, int* nparts,
                     hipStream_t st) {
  const int LR = gbm.C / 2;
#define X(lr, tc)                                                          \
  if (LR == lr && cfg.TC == tc) {                                          \
    const unsigned nb = (unsigned)((long long)gbm.M * gbm.R /              \
                                   RowCfg<lr>::RPB);                        \
    hipLaunchKernelGGL((k_adj_row21<lr, tc>), dim3(nb), dim3(256), 0, st,  \
                       gbm, F, Yb, Ub, tbits, alpha, twC, lpart);          \
    *nparts = (int)nb * 4;                                                 \
    return hipGetLastError();                                              \
  }
  SCC_UNR_CASES(X)
#undef X
  return hipErrorNotSupported;
}

// mode 0: Ybar += irfft, Ubar -= irfft; mode 1: out = irfft (unscaled).
hipError_t adj_row2(const Geom& gbm, const Pow2Config& cfg, float2* F,
                    float* Yb, float* Ub, float* out, const float2* twC,
                    int mode, hipStream_t st) {
  if (mode == 1) return pow2_row_inv(gbm, cfg, gbm.M, F, out, twC, 1.f, st);
  const int LR = gbm.C / 2;
#define X(lr, tc)                                                           \
  if (LR == lr && cfg.TC == tc) {                                           \
    const unsigned nb = (unsigned)((long long)gbm.M * gbm.R /               \
                                   RowCfg<lr>::RPB);                         \
    hipLaunchKernelGGL((k_adj_row2<lr, tc>), dim3(nb), dim3(256), 0, st,    \
                       gbm, F, Yb, Ub, twC);                                \
    return hipGetLastError();                                               \
  }
  SCC_UNR_CASES(X)
#undef X
  return hipErrorNotSupported;
}

// The adjoint solve with the maps split over the block's 4 wavefronts
// (used for 32 < M <= 128, where k_adj_solve needs 8 or more lanes per
// frequency and its reads cover half a 128-byte line of a map): a block
// takes 16 consecutive frequencies, wavefront w the maps (4 w + gi) MPT + i
// of lane group gi, so every read covers a whole line.  The per-frequency
// sums (sigma, then per image num, then sDB and sDX) meet in LDS and are
// added in a fixed order; every thread passes every barrier.
template <int MPT>
__global__ __launch_bounds__(256) void k_adj_solve_split(
    Geom g, int nimg, float2* F, const float2* __restrict__ Xh,
    const float2* __restrict__ Df, const float2* __restrict__ Sf,
    float2* Dhb, float2* Shb, float rho, float invN) {
  __shared__ float rn[4][16][2];  // sigma, then num per image
  __shared__ float rs[4][16][4];  // sDB, sDX per image
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int gi = lane >> 4;
  const int fl = lane & 15;
  const long long f = (long long)blockIdx.x * 16 + fl;
  const bool valid = f < g.Hc;
  const int M = g.M;
  const int m0 = (wv * 4 + gi) * MPT;
  const float irho = 1.f / rho;
  const float wt = valid ? col_mult(g, panel_col(g, f)) : 1.f;
  const float sin_ = wt * invN;
  const float sout = rho / wt;
  float2 dd[MPT], dacc[MPT];
  float sig = 0.f;
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = m0 + i;
    dacc[i] = make_float2(0.f, 0.f);
    dd[i] = make_float2(0.f, 0.f);
    if (valid && m < M) {
      dd[i] = Df[(long long)m * g.Hc + f];
      sig += dd[i].x * dd[i].x + dd[i].y * dd[i].y;
    }
  }
  sig += __shfl_xor(sig, 16, 64);
  sig += __shfl_xor(sig, 32, 64);
  if (gi == 0) rn[wv][fl][0] = sig;
  __syncthreads();
  sig = ((rn[0][fl][0] + rn[1][fl][0]) + rn[2][fl][0]) + rn[3][fl][0];
  const float isig = 1.f / (rho + sig);
  __syncthreads();  // sigma read before the first num is written
  const int i0 = blockIdx.y * kImgGroup;
  const int i1 = min(nimg, i0 + kImgGroup);
  for (int img = i0; img < i1; ++img) {
    float2* Fi = F + (long long)img * M * g.Hc;
    const float2* Xi = Xh + (long long)img * M * g.Hc;
    float2 xb[MPT], xh[MPT];
    float2 sh = make_float2(0.f, 0.f), sb = make_float2(0.f, 0.f);
    if (valid) {
      sh = Sf[(long long)img * g.Hc + f];
      if (wv == 0 && gi == 0) sb = Shb[(long long)img * g.Hc + f];
    }
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      const int m = m0 + i;
      xb[i] = make_float2(0.f, 0.f);
      xh[i] = make_float2(0.f, 0.f);
      if (valid && m < M) {
        const long long o = (long long)m * g.Hc + f;
        xb[i] = Fi[o];
        xh[i] = Xi[o];
      }
    }
    float2 num = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      xb[i] = cscale(xb[i], sin_);
      num = cadd(num, cmul(dd[i], xb[i]));
    }
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      num.x += __shfl_xor(num.x, o, 64);
      num.y += __shfl_xor(num.y, o, 64);
    }
    if (gi == 0) {
      rn[wv][fl][0] = num.x;
      rn[wv][fl][1] = num.y;
    }
    __syncthreads();
    num = make_float2(
        ((rn[0][fl][0] + rn[1][fl][0]) + rn[2][fl][0]) + rn[3][fl][0],
        ((rn[0][fl][1] + rn[1][fl][1]) + rn[2][fl][1]) + rn[3][fl][1]);
    const float2 cb = cscale(num, isig);
    float2 sDB = make_float2(0.f, 0.f), sDX = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      xb[i] = cscale(csub(xb[i], cmul(cconj(dd[i]), cb)), irho);  // Bbar
      sDB = cadd(sDB, cmul(dd[i], xb[i]));
      sDX = cadd(sDX, cmul(dd[i], xh[i]));
    }
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      sDB.x += __shfl_xor(sDB.x, o, 64);
      sDB.y += __shfl_xor(sDB.y, o, 64);
      sDX.x += __shfl_xor(sDX.x, o, 64);
      sDX.y += __shfl_xor(sDX.y, o, 64);
    }
    if (gi == 0) {
      rs[wv][fl][0] = sDB.x;
      rs[wv][fl][1] = sDB.y;
      rs[wv][fl][2] = sDX.x;
      rs[wv][fl][3] = sDX.y;
    }
    // Also orders this image's num reads before the next image's writes.
    __syncthreads();
    sDB = make_float2(
        ((rs[0][fl][0] + rs[1][fl][0]) + rs[2][fl][0]) + rs[3][fl][0],
        ((rs[0][fl][1] + rs[1][fl][1]) + rs[2][fl][1]) + rs[3][fl][1]);
    sDX = make_float2(
        ((rs[0][fl][2] + rs[1][fl][2]) + rs[2][fl][2]) + rs[3][fl][2],
        ((rs[0][fl][3] + rs[1][fl][3]) + rs[2][fl][3]) + rs[3][fl][3]);
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      const int m = m0 + i;
      if (valid && m < M) {
        const float2 cbb = cconj(xb[i]);
        dacc[i] = cadd(dacc[i],
                       csub(cmul(cbb, csub(sh, sDX)), cmul(cconj(xh[i]), sDB)));
        Fi[(long long)m * g.Hc + f] = cscale(xb[i], sout);
      }
    }
    if (valid && wv == 0 && gi == 0)
      Shb[(long long)img * g.Hc + f] = cadd(sb, sDB);
    // The sDB / sDX reads above finish before the next image writes rs:
    // the next image's num barrier lies between them.
  }
  float2* Dg = Dhb + (long long)blockIdx.y * M * g.Hc;
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = m0 + i;
    if (valid && m < M) {
      float2* d = Dg + (long long)m * g.Hc + f;
      *d = cadd(*d, dacc[i]);
    }
  }
}

int img_groups(int nimg) { return (nimg + kImgGroup - 1) / kImgGroup; }

hipError_t adj_solve(const Geom& g, int nimg, float2* F, const float2* Xh,
                     const float2* Df, const float2* Sf, float2* Dhb,
                     float2* Shb, float rho, hipStream_t st) {
  const float invN0 = 1.f / (float)g.N;
  if (g.M > 32 && g.M <= 128) {
    const dim3 nbs((unsigned)((g.Hc + 15) / 16), (unsigned)img_groups(nimg));
#define S(MP)                                                               \
  if (g.M <= 16 * MP) {                                                    \
    hipLaunchKernelGGL((k_adj_solve_split<MP>), nbs, dim3(256), 0, st, g,   \
                       nimg, F, Xh, Df, Sf, Dhb, Shb, rho, invN0);         \
    return hipGetLastError();                                              \
  }
    S(4) S(8)
#undef S
  }
  int G = 1;
  while (G * 8 < g.M) G *= 2;
  const long long fpb = 4LL * (64 / G);
  const dim3 nb((unsigned)((g.Hc + fpb - 1) / fpb),
                (unsigned)img_groups(nimg));
  const float invN = 1.f / (float)g.N;
#define X(gg)                                                                \
  if (G == gg) {                                                             \
    hipLaunchKernelGGL((k_adj_solve<gg, 8>), nb, dim3(256), 0, st, g, nimg,  \
                       F, Xh, Df, Sf, Dhb, Shb, rho, invN);                  \
    return hipGetLastError();                                                \
  }
  X(1) X(2) X(4) X(8) X(16) X(32)
#undef X
  return hipErrorNotSupported;
}

hipError_t scale_cols(const Geom& g, float2* F, int nmaps, hipStream_t st) {
  const long long n = (long long)nmaps * g.Hc;
  const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_scale_cols, dim3(nb), dim3(256), 0, st, g, F, n);
  return hipGetLastError();
}

thread_local std::string u_err;

int ufail(int code, const std::string& m) {
  u_err = m;
  return code;
}

#define UHIP(call)                                                          \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return ufail(e_ == hipErrorOutOfMemory ? SCC_ERR_NOMEM : SCC_ERR_HIP, \
                   std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

struct UBuf {
  void* p = nullptr;
  size_t n = 0;
  ~UBuf() {
    if (p) (void)hipFree(p);
  }
  int get(size_t bytes) {
    if (p && n >= bytes) return SCC_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (bytes == 0) return SCC_OK;
    if (hipMalloc(&p, bytes) != hipSuccess)
      return ufail(SCC_ERR_NOMEM, "hipMalloc failed (" +
                                      std::to_string(bytes) + " bytes)");
    n = bytes;
    return SCC_OK;
  }
};

// Buffers of one (device, R, C, M, B, k) configuration.
struct UWs {
  int device = -1, R = 0, C = 0, M = 0, B = 0, k = 0;
  Geom g{}, gbm{};
  Pow2Config cfg{};
  UBuf F, Df, Sf, Y, U, part, bsum, ctrl, twR, twC;
  UBuf tXh, tbits, Yb, Ub, Dhb, Shb, lacc, tmp, lparts;
  UBuf Fr;  // work spectra of the backward's recompute
  size_t nr = 0, nc = 0, nbits = 0;  // per-iteration tape sizes
};

std::mutex u_mu;
std::unique_ptr<UWs> u_ws;

int check_shapes(int M, int K, int B, int R, int C, int n_iter, int k,
                 float rho) {
  if (M < 1 || K < 1 || B < 1 || K > R || K > C)
    return ufail(SCC_ERR_INVALID, "bad sizes");
  if (n_iter < 1 || k < 1) return ufail(SCC_ERR_INVALID, "bad iteration count");
  if (!(rho > 0.f)) return ufail(SCC_ERR_INVALID, "rho must be positive");
  Pow2Config cfg;
  if (!pow2_select(R, C, M, &cfg))
    return ufail(SCC_ERR_UNSUPPORTED,
                 "unrolled solve needs power-of-two sides in 256..2048 and "
                 "M <= 256");
  return SCC_OK;
}

int get_uws(int R, int C, int M, int B, int k, UWs** out) {
  int dev = 0;
  UHIP(hipGetDevice(&dev));
  if (u_ws && u_ws->device == dev && u_ws->R == R && u_ws->C == C &&
      u_ws->M == M && u_ws->B == B && u_ws->k == k) {
    *out = u_ws.get();
    return SCC_OK;
  }
  u_ws.reset();
  auto w = std::make_unique<UWs>();
  w->device = dev;
  w->R = R;
  w->C = C;
  w->M = M;
  w->B = B;
  w->k = k;
  pow2_select(R, C, M, &w->cfg);
  Geom& g = w->g;
  g.R = R;
  g.C = C;
  g.M = M;
  g.H = C / 2 + 1;
  g.even = 1;
  g.Lr = C / 2;
  g.N = (long long)R * C;
  g.Hc = (long long)R * g.H;
  g.TC = w->cfg.TC;
  g.P = (g.H + g.TC - 1) / g.TC;
  g.wlast = g.H - (g.P - 1) * g.TC;
  for (g.lgR = 0; (1 << g.lgR) < R; ++g.lgR) {
  }
  w->gbm = g;
  w->gbm.M = B * M;
  const size_t nr = (size_t)B * M * g.N, nc = (size_t)B * M * g.Hc;
  const size_t nbits = (size_t)B * M * R * tape_words_per_row(C);
  w->nr = nr;
  w->nc = nc;
  w->nbits = nbits;
  const size_t nwaves = (size_t)pow2_row_blocks(w->gbm, w->cfg);
  int rc;
  if ((rc = w->F.get(nc * 8)) || (rc = w->Df.get((size_t)M * g.Hc * 8)) ||
      (rc = w->Sf.get((size_t)B * g.Hc * 8)) || (rc = w->Y.get(nr * 4)) ||
      (rc = w->U.get(nr * 4)) || (rc = w->part.get(nwaves * 5 * 4)) ||
      (rc = w->bsum.get(64 * 5 * 8)) || (rc = w->ctrl.get(sizeof(Ctrl))) ||
      (rc = w->twR.get((size_t)R * 8)) || (rc = w->twC.get((size_t)C * 8)) ||
      (rc = w->tXh.get((size_t)k * nc * 8)) ||
      (rc = w->tbits.get((size_t)k * nbits * 4)) ||
      (rc = w->Yb.get(nr * 4)) ||
      (rc = w->Ub.get(nr * 4)) ||
      (rc = w->Dhb.get((size_t)img_groups(B) * M * g.Hc * 8)) ||
      (rc = w->Shb.get((size_t)B * g.Hc * 8)) || (rc = w->lacc.get(8)) ||
      (rc = w->tmp.get((size_t)M * g.N * 4)))
    return rc;
  std::vector<float2> tw;
  for (int pass = 0; pass < 2; ++pass) {
    const int L = pass == 0 ? R : C;
    tw.resize(L);
    for (int t = 0; t < L; ++t) {
      const double a = -2.0 * M_PI * (double)t / (double)L;
      tw[t] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    UHIP(hipMemcpy(pass == 0 ? w->twR.p : w->twC.p, tw.data(), (size_t)L * 8,
                   hipMemcpyHostToDevice));
  }
  *out = w.get();
  u_ws = std::move(w);
  return SCC_OK;
}

Buffers make_buffers(UWs* w) {
  Buffers b{};
  b.Y = (float*)w->Y.p;
  b.U = (float*)w->U.p;
  b.F = (float2*)w->F.p;
  b.Df = (float2*)w->Df.p;
  b.Sf = (float2*)w->Sf.p;
  b.partials = (float*)w->part.p;
  b.blocksum = (double*)w->bsum.p;
  b.ctrl = (Ctrl*)w->ctrl.p;
  b.twR = (float2*)w->twR.p;
  b.twC = (float2*)w->twC.p;
  return b;
}

Params make_params(float lambda, float alpha) {
  Params p{};
  p.lambda = lambda;
  p.alpha = alpha;
  p.l1_mode = 0;
  p.l1_scalar = 1.f;
  p.gw_scalar = 1.f;
  return p;
}

// Spectra of the dictionary (shared) and of the B images; control block.
int setup(UWs* w, const float* D, int K, const float* S, float rho,
          hipStream_t st) {
  const Geom& g = w->g;
  Buffers b = make_buffers(w);
  UHIP(launch_pad(D, (float*)w->tmp.p, w->M, K, K, w->R, w->C, st));
  UHIP(pow2_row_fwd(g, w->cfg, w->M, kFwdReal, (const float*)w->tmp.p,
                    nullptr, nullptr, b.Df, b.twC, nullptr, st));
  UHIP(pow2_col(g, w->cfg, w->M, b.Df, b.twR, 0, nullptr, st));
  UHIP(pow2_row_fwd(g, w->cfg, w->B, kFwdReal, S, nullptr, nullptr, b.Sf,
                    b.twC, nullptr, st));
  UHIP(pow2_col(g, w->cfg, w->B, b.Sf, b.twR, 0, nullptr, st));
  Ctrl c0{};
  c0.rho = rho;
  c0.rsf = 1.f;
  UHIP(hipMemcpyAsync(b.ctrl, &c0, sizeof(Ctrl), hipMemcpyHostToDevice, st));
  UHIP(hipStreamSynchronize(st));  // c0 lives on this stack frame
  return SCC_OK;
}

// One forward iteration over the batch; F (or Fw when given) holds the row
// spectra of Y - U.  With tape pointers the solve also writes X-hat to xtape and the row
// kernel the support bits of V to tbits (or V itself to vout).  The row
// kernel reads (Y, U) from (Yin, Uin) when given and writes them to
// (Yout, Uout) when given, else in place in the workspace.
struct IterIO {
  const float* Yin = nullptr;
  const float* Uin = nullptr;
  float* Yout = nullptr;
  float* Uout = nullptr;
};

int fwd_iter(UWs* w, const Params& p, float2* xtape, unsigned* tbits,
             float* vout, const IterIO& io, hipStream_t st,
             float2* Fw = nullptr) {
  Buffers b = make_buffers(w);
  if (Fw != nullptr) b.F = Fw;
  b.xtape = xtape;
  b.tbits = tbits;
  b.vout = vout;
  const int BM = w->B * w->M;
  int nparts = 0;
  UHIP(pow2_col(w->g, w->cfg, BM, b.F, b.twR, 0, b.ctrl, st));
  UHIP(pow2_solve(w->g, w->cfg, b, p, st, w->B));
  b.xtape = nullptr;
  UHIP(pow2_col(w->g, w->cfg, BM, b.F, b.twR, 1, b.ctrl, st));
  // Inputs default to the workspace, also when the outputs are redirected.
  b.Yin = io.Yin ? io.Yin : b.Y;
  b.Uin = io.Uin ? io.Uin : b.U;
  if (io.Yout) b.Y = io.Yout;
  if (io.Uout) b.U = io.Uout;
  UHIP(pow2_row(w->gbm, w->cfg, b, p, kRowFused, &nparts, st));
  return SCC_OK;
}

// Checkpoint buffer: (Y, U) at the start of every segment, then the tape
// of the last kTapeSegs segments (X-hat of each of their iterations, then
// the support bits), which the backward pass reads instead of recomputing
// them.  Memory stays O(segments x state + kTapeSegs x k x tape).
constexpr int kTapeSegs = 3;

struct CkptLayout {
  int nseg;
  int it0_tape, cnt_tape;  // first taped iteration, number of them
  size_t tail;             // float offset of the tape
};

CkptLayout ckpt_layout(int n_iter, int k, size_t nr) {
  CkptLayout c;
  c.nseg = (n_iter + k - 1) / k;
  c.it0_tape = std::max(0, c.nseg - kTapeSegs) * k;
  c.cnt_tape = n_iter - c.it0_tape;
  c.tail = (size_t)c.nseg * 2 * nr;
  return c;
}

// Tape of iteration it0_tape + j.
float2* tail_xh(const UWs* w, float* ckpt, const CkptLayout& c, int j) {
  return reinterpret_cast<float2*>(ckpt + c.tail) + (size_t)j * w->nc;
}

unsigned* tail_bits(const UWs* w, float* ckpt, const CkptLayout& c, int j) {
  return reinterpret_cast<unsigned*>(ckpt + c.tail +
                                     (size_t)c.cnt_tape * 2 * w->nc) +
         (size_t)j * w->nbits;
}

}  // namespace
}  // namespace scc

using namespace scc;

extern "C" {

size_t scc_unrolled_ckpt_floats(int M, int B, int R, int C, int n_iter,
                                int ckpt_every) {
  if (M < 1 || B < 1 || R < 2 || C < 2 || n_iter < 1 || ckpt_every < 1)
    return 0;
  const size_t nr = (size_t)B * M * R * C;
  const size_t nc = (size_t)B * M * R * (C / 2 + 1);
  const size_t nbits = (size_t)B * M * R * tape_words_per_row(C);
  const CkptLayout c = ckpt_layout(n_iter, ckpt_every, nr);
  return c.tail + (size_t)c.cnt_tape * (2 * nc + nbits);
}

int scc_unrolled_forward(const float* D, int M, int K, const float* S, int B,
                         int R, int C, float lambda, float rho, float alpha,
                         int n_iter, int ckpt_every, float* Y, float* ckpt,
                         void* stream) {
  int rc = check_shapes(M, K, B, R, C, n_iter, ckpt_every, rho);
  if (rc) return rc;
  if (!D || !S || !Y) return ufail(SCC_ERR_INVALID, "NULL pointer");
  std::lock_guard<std::mutex> lk(u_mu);
  UWs* w;
  if ((rc = get_uws(R, C, M, B, ckpt_every, &w))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = setup(w, D, K, S, rho, st))) return rc;
  const size_t nr = w->nr;
  const Params p = make_params(lambda, alpha);
  const CkptLayout c = ckpt_layout(n_iter, ckpt_every, nr);
  // Iteration i reads (Y, U) from its segment's checkpoint slot when it
  // starts a segment and writes them into the next slot when it ends one,
  // so no copies are made; the last iteration writes Y to the output.
  auto slot = [&](int sg) { return ckpt + (size_t)sg * 2 * nr; };
  float* y0 = ckpt ? slot(0) : (float*)w->Y.p;
  UHIP(hipMemsetAsync(y0, 0, nr * 4, st));
  UHIP(hipMemsetAsync(ckpt ? y0 + nr : (float*)w->U.p, 0, nr * 4, st));
  UHIP(hipMemsetAsync(w->F.p, 0, w->nc * 8, st));
  for (int i = 0; i < n_iter; ++i) {
    IterIO io;
    const bool last = i == n_iter - 1;
    if (ckpt && i % ckpt_every == 0) {
      io.Yin = slot(i / ckpt_every);
      io.Uin = io.Yin + nr;
    }
    if (ckpt && !last && (i + 1) % ckpt_every == 0) {
      io.Yout = slot((i + 1) / ckpt_every);
      io.Uout = io.Yout + nr;
    }
    if (last) io.Yout = Y;
    const int j = i - c.it0_tape;  // slot in the checkpoint buffer's tape
    const bool tape = ckpt && j >= 0;
    if ((rc = fwd_iter(w, p, tape ? tail_xh(w, ckpt, c, j) : nullptr,
                       tape ? tail_bits(w, ckpt, c, j) : nullptr, nullptr,
                       io, st)))
      return rc;
  }
  UHIP(hipStreamSynchronize(st));
  return SCC_OK;
}

int scc_unrolled_backward(const float* D, int M, int K, const float* S, int B,
                          int R, int C, float lambda, float rho, float alpha,
                          int n_iter, int ckpt_every, const float* ckpt,
                          const float* gY, float* gD, float* gS,
                          float* glambda, void* stream) {
  int rc = check_shapes(M, K, B, R, C, n_iter, ckpt_every, rho);
  if (rc) return rc;
  if (!D || !S || !ckpt || !gY || !gD || !gS || !glambda)
    return ufail(SCC_ERR_INVALID, "NULL pointer");
  std::lock_guard<std::mutex> lk(u_mu);
  UWs* w;
  if ((rc = get_uws(R, C, M, B, ckpt_every, &w))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = setup(w, D, K, S, rho, st))) return rc;
  const Geom& g = w->g;
  const size_t nr = w->nr;
  const int BM = B * M;
  Buffers b = make_buffers(w);
  const Params p = make_params(lambda, alpha);
  float* Yb = (float*)w->Yb.p;
  float* Ub = (float*)w->Ub.p;
  float2* Dhb = (float2*)w->Dhb.p;
  float2* Shb = (float2*)w->Shb.p;
  double* lacc = (double*)w->lacc.p;
  float* ck = const_cast<float*>(ckpt);  // the tape is only read here
  const CkptLayout c = ckpt_layout(n_iter, ckpt_every, nr);
  UHIP(hipMemsetAsync(Ub, 0, nr * 4, st));
  const int ngrp = img_groups(B);
  UHIP(hipMemsetAsync(Dhb, 0, (size_t)ngrp * M * g.Hc * 8, st));
  UHIP(hipMemsetAsync(Shb, 0, (size_t)B * g.Hc * 8, st));
  // Per-wavefront partial sums of lambda-bar, one slice per iteration,
  // added up once after the last iteration.
  const size_t nparts = (size_t)pow2_row_blocks(w->gbm, w->cfg);
  if ((rc = w->lparts.get((size_t)n_iter * nparts * 4))) return rc;
  float* lparts = (float*)w->lparts.p;
  int slot = 0;
  int np = 0;
  // Tape of iteration j of segment sg: the last kTapeSegs segments' lives
  // with the checkpoints, the others' in the workspace (recomputed one
  // segment at a time).
  auto seg_cnt = [&](int sg) {
    return std::min(n_iter, (sg + 1) * ckpt_every) - sg * ckpt_every;
  };
  auto taped = [&](int sg) { return sg * ckpt_every >= c.it0_tape; };
  auto xh = [&](int sg, int j) {
    return taped(sg) ? tail_xh(w, ck, c, sg * ckpt_every + j - c.it0_tape)
                     : (float2*)w->tXh.p + (size_t)j * w->nc;
  };
  auto bits = [&](int sg, int j) {
    return taped(sg) ? tail_bits(w, ck, c, sg * ckpt_every + j - c.it0_tape)
                     : (unsigned*)w->tbits.p + (size_t)j * w->nbits;
  };
  // Recompute segment sg from its checkpoint (read in place: rsf is 1, so
  // the row transform of Y - U does not write U) into its own spectra Fr,
  // recording the tape.
  auto recompute = [&](int sg) -> int {
    float* cy = ck + (size_t)sg * 2 * nr;
    float2* Fr = (float2*)w->Fr.p;
    UHIP(pow2_row_fwd(w->gbm, w->cfg, BM, kFwdYU, nullptr, cy, cy + nr, Fr,
                      b.twC, b.ctrl, st));
    for (int j = 0; j < seg_cnt(sg); ++j) {
      IterIO io;
      if (j == 0) {
        io.Yin = cy;
        io.Uin = cy + nr;
      }
      int r2;
      if ((r2 = fwd_iter(w, p, xh(sg, j), bits(sg, j), nullptr, io, st, Fr)))
        return r2;
    }
    return SCC_OK;
  };
  if (!taped(0) && (rc = w->Fr.get(w->nc * 8))) return rc;
  // dL/dY enters through the first shrink backward.
  UHIP(adj_row1(w->gbm, w->cfg, gY, Yb, Ub, bits(c.nseg - 1, seg_cnt(c.nseg - 1) - 1),
                alpha, b.F, b.twC, lparts + (size_t)slot++ * nparts, &np, st));
  for (int sg = c.nseg - 1; sg >= 0; --sg) {
    for (int j = seg_cnt(sg) - 1; j >= 0; --j) {
      UHIP(pow2_col(g, w->cfg, BM, b.F, b.twR, 0, nullptr, st));
      UHIP(adj_solve(g, B, b.F, xh(sg, j), b.Df, b.Sf, Dhb, Shb, rho, st));
      UHIP(pow2_col(g, w->cfg, BM, b.F, b.twR, 1, nullptr, st));
      if (j > 0) {
        UHIP(adj_row21(w->gbm, w->cfg, b.F, Yb, Ub, bits(sg, j - 1), alpha,
                       b.twC, lparts + (size_t)slot++ * nparts, &np, st));
      } else if (sg > 0) {
        // The previous segment's tape, then the fused rows across the
        // segment boundary.
        if (!taped(sg - 1) && (rc = recompute(sg - 1))) return rc;
        UHIP(adj_row21(w->gbm, w->cfg, b.F, Yb, Ub,
                       bits(sg - 1, seg_cnt(sg - 1) - 1), alpha, b.twC,
                       lparts + (size_t)slot++ * nparts, &np, st));
      } else {
        UHIP(adj_row2(w->gbm, w->cfg, b.F, Yb, Ub, nullptr, b.twC, 0, st));
      }
    }
  }
  if (ngrp > 1) {
    const long long n = (long long)M * g.Hc;
    hipLaunchKernelGGL(k_sum_groups,
                       dim3((unsigned)std::min<long long>((n + 255) / 256,
                                                          65536)),
                       dim3(256), 0, st, Dhb, ngrp, n);
    UHIP(hipGetLastError());
  }
  // Image gradient: N irfft2(Sbar / w); dictionary: crop(N irfft2(Dbar / w)).
  Geom gB = g;
  gB.M = B;
  UHIP(scale_cols(g, Shb, B, st));
  UHIP(pow2_col(g, w->cfg, B, Shb, b.twR, 1, nullptr, st));
  UHIP(adj_row2(gB, w->cfg, Shb, nullptr, nullptr, gS, b.twC, 1, st));
  UHIP(scale_cols(g, Dhb, M, st));
  UHIP(pow2_col(g, w->cfg, M, Dhb, b.twR, 1, nullptr, st));
  UHIP(adj_row2(g, w->cfg, Dhb, nullptr, nullptr, (float*)w->tmp.p, b.twC, 1,
                st));
  const int nk = M * K * K;
  hipLaunchKernelGGL(k_crop, dim3((nk + 255) / 256), dim3(256), 0, st,
                     (const float*)w->tmp.p, gD, M, K, R, C);
  UHIP(hipGetLastError());
  hipLaunchKernelGGL(k_lsum_blocks, dim3(kLsumBlocks), dim3(1024), 0, st,
                     (const float*)lparts, (long long)slot * (long long)nparts,
                     b.blocksum);
  hipLaunchKernelGGL(k_lsum_final, dim3(1), dim3(64), 0, st,
                     (const double*)b.blocksum, lacc);
  UHIP(hipGetLastError());
  double l = 0.0;
  UHIP(hipMemcpyAsync(&l, lacc, 8, hipMemcpyDeviceToHost, st));
  UHIP(hipStreamSynchronize(st));
  *glambda = (float)(l / rho);
  return SCC_OK;
}

int scc_unrolled_tape(const float* D, int M, int K, const float* S, int B,
                      int R, int C, float lambda, float rho, float alpha,
                      int n_iter, float* V, void* stream) {
  int rc = check_shapes(M, K, B, R, C, n_iter, 1, rho);
  if (rc) return rc;
  if (!D || !S || !V) return ufail(SCC_ERR_INVALID, "NULL pointer");
  std::lock_guard<std::mutex> lk(u_mu);
  UWs* w;
  if ((rc = get_uws(R, C, M, B, 1, &w))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = setup(w, D, K, S, rho, st))) return rc;
  const size_t nr = w->nr;
  UHIP(hipMemsetAsync(w->Y.p, 0, nr * 4, st));
  UHIP(hipMemsetAsync(w->U.p, 0, nr * 4, st));
  UHIP(hipMemsetAsync(w->F.p, 0, w->nc * 8, st));
  const Params p = make_params(lambda, alpha);
  for (int i = 0; i < n_iter; ++i)
    if ((rc = fwd_iter(w, p, nullptr, nullptr, V + (size_t)i * nr, IterIO(),
                       st)))
      return rc;
  UHIP(hipStreamSynchronize(st));
  return SCC_OK;
}

const char* scc_unrolled_last_error(void) { return u_err.c_str(); }

}  // extern "C"
