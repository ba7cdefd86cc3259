// Register-resident kernels for power-of-two images (R, C in 256..2048).
//
// Same stage map as kernels_generic.hip, with every transform in registers
// (csrc/fft_reg.h) and compile-time sizes:
//   k_col_pow2    one panel (TC columns x R rows, contiguous in HBM) per
//                 block; E = 16 points per thread; LDS only for the stage
//                 exchanges.
//   k_solve_pow2  G lanes per frequency, MPT maps per lane held in
//                 registers, so A-hat and D-hat are read exactly once; the G
//                 partial sums are combined with cross-lane shuffles.
//   k_row_pow2    NR rows per block: spectrum -> inverse real FFT (half-
//                 length complex FFT + post-twiddle) -> relax / shrink / AMS
//                 / dual update / residual sums with Y and U prefetched as
//                 float4 -> Y - U -> forward real FFT -> spectrum.

#include "fft_reg.h"
#include "pow2_common.h"
#include "scc_device.h"
#include "scc_pow2.h"

namespace scc {

namespace {

constexpr int kColE = 16;  // points per thread in the column transforms

// One element of the fused update on a value x (= X, already / N).
// Y from V (general update rules).
__device__ __forceinline__ float shrink1(const Params& p, float v, float thr,
                                         int amsw, bool ams, bool zero_bndry) {
  float y;
  if (ams) {
    y = (float)(1 - amsw) * v;
  } else {
    y = copysignf(fmaxf(fabsf(v) - thr, 0.f), v);
    if (p.nonneg) y = fmaxf(y, 0.f);
    if (zero_bndry) y = 0.f;
  }
  return y;
}

__device__ __forceinline__ void upd1(const Params& p, float x, float yo,
                                     float uo, float thr, int amsw,
                                     bool ams, bool zero_bndry, float& yn,
                                     float& un, float& vv, UpdateSums& acc) {
  const float xr = (p.alpha == 1.f) ? x : p.alpha * x + (1.f - p.alpha) * yo;
  const float v = xr + uo;
  yn = shrink1(p, v, thr, amsw, ams, zero_bndry);
  un = v - yn;
  vv = v;
  acc.x += x * x;
  acc.y += yn * yn;
  acc.u += un * un;
  const float dxy = x - yn;
  acc.r += dxy * dxy;
  const float dy = yo - yn;
  acc.s += dy * dy;
}


}  // namespace

// --------------------------------------------------------------------------
// Column transforms: one panel (TC columns x R rows, row-major, contiguous
// in HBM) per block, E = 16 points per thread.  With nparts > 0 the first
// nred blocks (in dispatch order) also run the first pass of the previous
// row kernel's residual reduction (partials_slice) before their transforms.
// --------------------------------------------------------------------------
template <int R, int TC, int DIR>
__global__ __launch_bounds__(TC*(R / kColE)) void k_col_pow2(
    Geom g, float2* __restrict__ F, const float2* __restrict__ twR,
    const Ctrl* ctrl, const float* __restrict__ parts, int nparts,
    double* bsum, int nred) {
  constexpr int E = kColE;
  constexpr int T = R / E;
  // Column stride congruent to 32 / TC (mod 32) in 8-byte words: every
  // stage's exchange then hits each LDS bank pair exactly twice per
  // wavefront access (the minimum for 64 lanes of 8 bytes).
  constexpr int L0 = reg::lds_len(R);
  constexpr int LP = L0 + ((32 / TC - L0 % 32) + 32) % 32;
  __shared__ float2 lds[TC * LP];
  __shared__ double red[5][16];
  if (ctrl != nullptr && ctrl->done) return;
  // The previous row kernel's partials are complete at launch, so the
  // reducing blocks add their slices first and overlap the others' FFTs.
  const int lin = blockIdx.y * gridDim.x + blockIdx.x;
  if (nparts > 0 && lin < nred)
    partials_slice(parts, nparts, bsum, lin, nred, red);
  const int p = blockIdx.x;
  // The forward transform takes the maps last to first: the row kernel
  // before it writes them first to last, so its most recent output, still
  // in the last-level cache, is read before this kernel's own traffic
  // evicts it (105.7 -> 95.7 us at 1024 x 1024 x 64).
  const long long m =
      DIR > 0 ? (long long)gridDim.y - 1 - blockIdx.y : blockIdx.y;
  const int w = (p == g.P - 1) ? g.wlast : TC;
  const int c = threadIdx.x % TC;
  const int t = threadIdx.x / TC;
  float2* base = F + m * g.Hc + (long long)p * TC * R;
  float2* L = lds + c * LP;
  float2 v[E];
  const bool on = c < w;
  if (on) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = base[(t + T * e) * w + c];
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = make_float2(0.f, 0.f);
  }
  reg::fft_block<R, E, DIR>(v, t, L, twR, 1);
  if (on) {
#pragma unroll
    for (int e = 0; e < E; ++e) base[(t + T * e) * w + c] = v[e];
  }
}

// --------------------------------------------------------------------------
// Sherman-Morrison X-solve: G lanes per frequency, MPT maps per lane.
// --------------------------------------------------------------------------
// With ctl_iter > 0, block 0 first completes the deferred residual
// reduction of iteration ctl_iter (nbsum block sums from the column kernel,
// added in a fixed order) and runs the control update; that iteration was
// not an AutoRho step, so rho is unchanged for the other blocks.
// NTD: the filter spectra are read with non-temporal loads (see
// pow2_solve).
// Deferred reduction of iteration ctl_iter (see k_solve_pow2), run by
// block (0, 0) of a 256-thread solve launch.
__device__ __forceinline__ void solve_prologue(const Buffers& b,
                                               const Params& p, int ctl_iter,
                                               int nbsum) {
  if (ctl_iter > 0 && blockIdx.x == 0 && blockIdx.y == 0 && !b.ctrl->done) {
    __shared__ double red[5][4];
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nbsum; i += 256)
#pragma unroll
      for (int j = 0; j < 5; ++j) acc[j] += b.blocksum[5 * i + j];
    const int ln = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      double v = acc[j];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
      if (ln == 0) red[j][w] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double s5[5];
#pragma unroll
      for (int j = 0; j < 5; ++j)
        s5[j] = ((red[j][0] + red[j][1]) + red[j][2]) + red[j][3];
      control_update(b, p, s5, ctl_iter);
    }
    __syncthreads();
  }
}

template <int G, int MPT, bool GRAD, bool NTD>
__global__ __launch_bounds__(256) void k_solve_pow2(Geom g, Buffers b,
                                                    Params p, int ctl_iter,
                                                    int nbsum) {
  solve_prologue(b, p, ctl_iter, nbsum);
  if (b.ctrl->done) return;
  constexpr int FPW = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int gi = lane / FPW;
  const int fl = lane - gi * FPW;
  const long long f =
      ((long long)blockIdx.x * (blockDim.x >> 6) + wv) * FPW + fl;
  const bool valid = f < g.Hc;
  const float rho = b.ctrl->rho;
  const float irho = 1.f / rho;
  const int M = g.M;
  // Images of a batch (same size, shared dictionary): blockIdx.y.
  float2* const Fi = b.F + (long long)blockIdx.y * M * g.Hc;
  // Optional copy of X-hat (the unrolled solver's tape), same layout.
  float2* const Xt =
      b.xtape ? b.xtape + (long long)blockIdx.y * M * g.Hc : nullptr;
  float2 sh = make_float2(0.f, 0.f);
  float gh = 0.f;
  if (valid) {
    sh = b.Sf[(long long)blockIdx.y * g.Hc + f];
    if (GRAD) gh = p.mu * b.ghg[f];
  }
  // b_m and D_m stay in registers between the two passes; 1/delta_m too
  // when it varies with the map (GradReg), so no division is repeated.
  float2 bb[MPT], dd[MPT];
  float id[GRAD ? MPT : 1];
  float2 num = make_float2(0.f, 0.f);
  float den = 0.f;
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = gi * MPT + i;
    if (valid && m < M) {
      const long long o = (long long)m * g.Hc + f;
      const float2 d = NTD ? load_once(b.Df + o) : b.Df[o];
      const float2 a = Fi[o];
      const float2 x = cadd(cmul(cconj(d), sh), cscale(a, rho));
      float inv = irho;
      if (GRAD) {
        inv = 1.f / ((p.gw_mode ? b.gw[m] : p.gw_scalar) * gh + rho);
        id[i] = inv;
      }
      num = cadd(num, cscale(cmul(d, x), inv));
      den += (d.x * d.x + d.y * d.y) * inv;
      bb[i] = x;
      dd[i] = d;
    }
  }
#pragma unroll
  for (int o = FPW; o < 64; o <<= 1) {
    num.x += __shfl_xor(num.x, o, 64);
    num.y += __shfl_xor(num.y, o, 64);
    den += __shfl_xor(den, o, 64);
  }
  const float2 cb = cscale(num, 1.f / (1.f + den));
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = gi * MPT + i;
    if (valid && m < M) {
      const float inv = GRAD ? id[i] : irho;
      const float2 x = cscale(csub(bb[i], cmul(cconj(dd[i]), cb)), inv);
      Fi[(long long)m * g.Hc + f] = x;
      if (Xt != nullptr) Xt[(long long)m * g.Hc + f] = x;
    }
  }
}

// --------------------------------------------------------------------------
// The same solve for 64 < M <= 256 with the maps split over the block's 4
// wavefronts instead of over more lanes: a block takes 16 consecutive
// frequencies, wavefront w the maps (4 w + gi) MPT + i of lane group gi, so
// every load reads 16 consecutive frequencies (a whole 128-byte line) of a
// map with MPT maps per lane in registers.  The 4 wavefronts' partial sums
// meet in LDS and are added in a fixed order.
// --------------------------------------------------------------------------
template <int MPT, bool GRAD, bool NTD>
__global__ __launch_bounds__(256) void k_solve_split(Geom g, Buffers b,
                                                     Params p, int ctl_iter,
                                                     int nbsum) {
  __shared__ float part[4][16][3];
  __shared__ int sdone;
  solve_prologue(b, p, ctl_iter, nbsum);
  // The stop flag is read once per block: block 0 may set it while this
  // block runs, and every thread of the block must reach the barrier below.
  if (threadIdx.x == 0) sdone = b.ctrl->done;
  __syncthreads();
  if (sdone) return;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int gi = lane >> 4;
  const int fl = lane & 15;
  const long long f = (long long)blockIdx.x * 16 + fl;
  const bool valid = f < g.Hc;
  const float rho = b.ctrl->rho;
  const float irho = 1.f / rho;
  const int M = g.M;
  float2* const Fi = b.F + (long long)blockIdx.y * M * g.Hc;
  float2* const Xt =
      b.xtape ? b.xtape + (long long)blockIdx.y * M * g.Hc : nullptr;
  float2 sh = make_float2(0.f, 0.f);
  float gh = 0.f;
  if (valid) {
    sh = b.Sf[(long long)blockIdx.y * g.Hc + f];
    if (GRAD) gh = p.mu * b.ghg[f];
  }
  const int m0 = (wv * 4 + gi) * MPT;
  float2 bb[MPT], dd[MPT];
  float id[GRAD ? MPT : 1];
  float2 num = make_float2(0.f, 0.f);
  float den = 0.f;
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = m0 + i;
    if (valid && m < M) {
      const long long o = (long long)m * g.Hc + f;
      const float2 d = NTD ? load_once(b.Df + o) : b.Df[o];
      const float2 a = Fi[o];
      const float2 x = cadd(cmul(cconj(d), sh), cscale(a, rho));
      float inv = irho;
      if (GRAD) {
        inv = 1.f / ((p.gw_mode ? b.gw[m] : p.gw_scalar) * gh + rho);
        id[i] = inv;
      }
      num = cadd(num, cscale(cmul(d, x), inv));
      den += (d.x * d.x + d.y * d.y) * inv;
      bb[i] = x;
      dd[i] = d;
    }
  }
#pragma unroll
  for (int o = 16; o < 64; o <<= 1) {
    num.x += __shfl_xor(num.x, o, 64);
    num.y += __shfl_xor(num.y, o, 64);
    den += __shfl_xor(den, o, 64);
  }
  if (gi == 0) {
    part[wv][fl][0] = num.x;
    part[wv][fl][1] = num.y;
    part[wv][fl][2] = den;
  }
  __syncthreads();
  num = make_float2(((part[0][fl][0] + part[1][fl][0]) + part[2][fl][0]) +
                        part[3][fl][0],
                    ((part[0][fl][1] + part[1][fl][1]) + part[2][fl][1]) +
                        part[3][fl][1]);
  den = ((part[0][fl][2] + part[1][fl][2]) + part[2][fl][2]) + part[3][fl][2];
  const float2 cb = cscale(num, 1.f / (1.f + den));
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = m0 + i;
    if (valid && m < M) {
      const float inv = GRAD ? id[i] : irho;
      const float2 x = cscale(csub(bb[i], cmul(cconj(dd[i]), cb)), inv);
      Fi[(long long)m * g.Hc + f] = x;
      if (Xt != nullptr) Xt[(long long)m * g.Hc + f] = x;
    }
  }
}

// --------------------------------------------------------------------------
// Row kernels: every row belongs to one wavefront (RPW rows per wavefront,
// T lanes per row, E points per lane), so a row never waits on a block
// barrier.  The half spectrum is read and written in slot order (lane t
// holds columns t + T e); the real-FFT pre/post-processing pairs column k
// with column L - k, which lives in lane (T - t) mod T, slot E - 1 - e
// (lane 0: its own slot E - e, or column L), fetched with a lane shuffle.
// The last FFT stage leaves its outputs in slot order, which is also the
// order of the real row as float2 pairs, so the update and the forward
// transform of Y - U need no LDS pass at all.
// --------------------------------------------------------------------------
// TAPE (plain update only) records the unrolled solver's tape: 1 the
// support bits of V (|V| > t, sign V) to b.tbits, 2 V itself to b.vout.
// VS: the state array holds V (Params::vstate); the old Y and U are decoded
// from it, and a fused launch stores the new V (an update-only launch
// stores Y and U).
template <int LR, int TC, bool GENERAL, int TAPE, bool VS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    GENERAL ? 1 : RowCfg<LR>::WAVES))) void
k_row_pow2(Geom g, Buffers b, Params p, int mode) {
  using RC = RowCfg<LR>;
  constexpr int E = RC::E, T = RC::T, RPW = RC::RPW;
  // The state is prefetched with the spectrum where registers allow it, and
  // for the plain update on V also with 16 points per lane (two wavefronts
  // per SIMD either way: 1024-point rows 4 % faster).
  constexpr bool PF = RC::PREFETCH || (VS && !GENERAL && TAPE == 0);
  constexpr int C = 2 * LR;
  constexpr int LP = reg::lds_len(LR);
  __shared__ float2 lds[4 * RPW * LP];
  if (b.ctrl->done) return;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int rw = lane / T;
  const int t = lane & (T - 1);
  const int partner = rw * T + ((T - t) & (T - 1));
  // Wave-uniform bases (the RPW rows of a wavefront share one map) keep
  // every per-lane address a 32-bit offset from a scalar base.
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
  const int row0 = wave * RPW;
  const int R = g.R;
  const int m = row0 >> g.lgR;
  const int r = (row0 & (R - 1)) + rw;
  const int row = row0 + rw;
  float2* L = lds + (wv * RPW + rw) * LP;
  float2* Fm = b.F + (long long)m * g.Hc;
  const long long rowoff = (long long)row * C;
  float2* Y2 = reinterpret_cast<float2*>(b.Y + (long long)row0 * C);
  float2* U2 = reinterpret_cast<float2*>(b.U + (long long)row0 * C);
  // Y and U are read from Yin / Uin when given (a checkpoint), written to
  // Y / U.
  const float2* Y2i = b.Yin ? reinterpret_cast<const float2*>(
                                  b.Yin + (long long)row0 * C)
                            : Y2;
  const float2* U2i = b.Uin ? reinterpret_cast<const float2*>(
                                  b.Uin + (long long)row0 * C)
                            : U2;
  const int lo = rw * LR + t;  // this lane's first float2 of its row

  // Half spectrum in slot order, column L on lane 0; Y and U prefetched.
  float2 v[E], y2[E], u2[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = spec_slot<TC, T, E>(Fm, t, e, r, R);
  float2 xL = make_float2(0.f, 0.f);
  if (t == 0) xL = Fm[(LR / TC) * TC * R + r];
  if (PF) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      y2[e] = Y2i[lo + T * e];
      if (!VS) u2[e] = U2i[lo + T * e];
    }
  }
  // Inverse real-FFT pre-processing:
  // Z'[k] = (X[k] + conj X[L-k]) + i W_C^-k (X[k] - conj X[L-k]).
  float2 z[E];
  const float2 wt = b.twC[t];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float2 sh = shfl2(v[E - 1 - e], partner);
    float2 xn = (t == 0) ? (e == 0 ? xL : v[(E - e) % E]) : sh;
    float2 xk = v[e];
    if (t == 0 && e == 0) {  // DC and Nyquist are real
      xk.y = 0.f;
      xn.y = 0.f;
    }
    xn.y = -xn.y;
    const float2 s = cadd(xk, xn);
    const float2 wd = cmul(cconj(reg::w_slot<E>(wt, e)), csub(xk, xn));
    z[e] = make_float2(s.x - wd.y, s.y + wd.x);
  }
  reg::fft_wave<LR, E, -1>(z, t, L, b.twC, 2);

  // Fused update on the real row: slot e holds x[2k], x[2k+1] (times N).
  const float rho = b.ctrl->rho;
  const float invN = 1.0f / (float)g.N;
  const float lr = p.lambda / rho;
  const bool ams = GENERAL && m < p.m0;
  const float wmap = p.l1_mode == 0 ? p.l1_scalar
                     : (p.l1_mode == 1 && !ams ? b.l1w[m] : 0.f);
  const bool bndry_row = p.nobndry && r > R - p.Kh;
  UpdateSums acc;
  if (!PF) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      y2[e] = Y2i[lo + T * e];
      if (!VS) u2[e] = U2i[lo + T * e];
    }
  }
  const bool store_v = VS && mode == kRowFused;
  if (GENERAL) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + T * e;
      const int n0 = 2 * k;
      float th0 = lr * wmap, th1 = lr * wmap;
      int w0 = 0, w1 = 0;
      if (ams) {
        const int2 w2 =
            reinterpret_cast<const int2*>(b.mask + (long long)r * C)[k];
        w0 = w2.x;
        w1 = w2.y;
      } else if (p.l1_mode == 2) {
        const float2 w2 = reinterpret_cast<const float2*>(b.l1w + rowoff)[k];
        th0 = lr * w2.x;
        th1 = lr * w2.y;
      }
      const bool zb0 = p.nobndry && (bndry_row || n0 > C - p.Kw);
      const bool zb1 = p.nobndry && (bndry_row || n0 + 1 > C - p.Kw);
      float yo0 = y2[e].x, yo1 = y2[e].y, uo0, uo1;
      if (VS) {  // y2 holds V
        yo0 = shrink1(p, y2[e].x, th0, w0, ams, zb0);
        yo1 = shrink1(p, y2[e].y, th1, w1, ams, zb1);
        uo0 = y2[e].x - yo0;
        uo1 = y2[e].y - yo1;
      } else {
        uo0 = u2[e].x;
        uo1 = u2[e].y;
      }
      float yn0, un0, yn1, un1, vv0, vv1;
      upd1(p, z[e].x * invN, yo0, uo0, th0, w0, ams, zb0, yn0, un0, vv0, acc);
      upd1(p, z[e].y * invN, yo1, uo1, th1, w1, ams, zb1, yn1, un1, vv1, acc);
      if (store_v) {
        Y2[lo + T * e] = make_float2(vv0, vv1);
      } else {
        Y2[lo + T * e] = make_float2(yn0, yn1);
        U2[lo + T * e] = make_float2(un0, un1);
      }
      z[e] = make_float2(yn0 - un0, yn1 - un1);
    }
  } else {
    // Plain update: Xr = a X + (1 - a) Yold, V = Xr + U,
    // Y = copysign(max(|V| - t, 0), V), U = V - Y.
    const float th = lr * wmap, mth = -th;
    const float a = p.alpha, a1 = 1.f - p.alpha;
    constexpr int TW = TapeWords<E>::W;
    unsigned tb[TW];
#pragma unroll
    for (int w = 0; w < TW; ++w) tb[w] = 0u;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float x0 = z[e].x * invN, x1 = z[e].y * invN;
      float yo0 = y2[e].x, yo1 = y2[e].y, uo0, uo1;
      if (VS) {  // y2 holds V; th >= 0 (split_u)
        uo0 = split_u(y2[e].x, th, mth);
        uo1 = split_u(y2[e].y, th, mth);
        yo0 = y2[e].x - uo0;
        yo1 = y2[e].y - uo1;
      } else {
        uo0 = u2[e].x;
        uo1 = u2[e].y;
      }
      const float v0 = fmaf(a, x0, fmaf(a1, yo0, uo0));
      const float v1 = fmaf(a, x1, fmaf(a1, yo1, uo1));
      float yn0, yn1, un0, un1;
      if (VS) {
        un0 = split_u(v0, th, mth);
        un1 = split_u(v1, th, mth);
        yn0 = v0 - un0;
        yn1 = v1 - un1;
      } else {
        yn0 = copysignf(fmaxf(fabsf(v0) - th, 0.f), v0);
        yn1 = copysignf(fmaxf(fabsf(v1) - th, 0.f), v1);
        un0 = v0 - yn0;
        un1 = v1 - yn1;
      }
      // Support |V| > t is Y != 0, and there sign V = sign Y.
      if (TAPE == 1)
        tb[e / 8] |= ((unsigned)(yn0 != 0.f) | ((unsigned)(yn0 < 0.f) << 1) |
                      ((unsigned)(yn1 != 0.f) << 2) |
                      ((unsigned)(yn1 < 0.f) << 3))
                     << (4 * (e % 8));
      if (TAPE == 2)
        reinterpret_cast<float2*>(b.vout + (long long)row0 * C)[lo + T * e] =
            make_float2(v0, v1);
      acc.x = fmaf(x0, x0, fmaf(x1, x1, acc.x));
      acc.y = fmaf(yn0, yn0, fmaf(yn1, yn1, acc.y));
      acc.u = fmaf(un0, un0, fmaf(un1, un1, acc.u));
      const float d0 = x0 - yn0, d1 = x1 - yn1;
      acc.r = fmaf(d0, d0, fmaf(d1, d1, acc.r));
      const float s0 = yo0 - yn0, s1 = yo1 - yn1;
      acc.s = fmaf(s0, s0, fmaf(s1, s1, acc.s));
      if (store_v) {
        Y2[lo + T * e] = make_float2(v0, v1);
      } else {
        Y2[lo + T * e] = make_float2(yn0, yn1);
        U2[lo + T * e] = make_float2(un0, un1);
      }
      z[e] = make_float2(yn0 - un0, yn1 - un1);
    }
    if (TAPE == 1) {
      unsigned* tw = b.tbits + ((long long)row * T + t) * TW;
#pragma unroll
      for (int w = 0; w < TW; ++w) tw[w] = tb[w];
    }
  }
  // Per-wavefront partial sums (no block barrier).
  {
    float sv[5] = {acc.x, acc.y, acc.u, acc.r, acc.s};
    const int nwaves = gridDim.x * 4;
    wave_sum5(sv);
    if (lane == 0)
#pragma unroll
      for (int j = 0; j < 5; ++j) b.partials[j * nwaves + wave] = sv[j];
  }
  if (mode != kRowFused) return;
  reg::fft_wave<LR, E, 1>(z, t, L, b.twC, 2);
  store_spectrum<LR, TC>(z, t, partner, r, R, Fm, b.twC);
}

// Forward row transform of a real array (setup) or of Y - U / rsf.
template <int LR, int TC>
__global__ __launch_bounds__(256) void k_row_fwd_pow2(
    Geom g, int mode, const float* __restrict__ src, float* Y, float* U,
    float2* __restrict__ F, const float2* __restrict__ twC, const Ctrl* ctrl,
    int iter) {
  using RC = RowCfg<LR>;
  constexpr int E = RC::E, T = RC::T, RPW = RC::RPW;
  constexpr int C = 2 * LR;
  constexpr int LP = reg::lds_len(LR);
  __shared__ float2 lds[4 * RPW * LP];
  // The V conversion also runs when the stopping test passed at this very
  // iteration (its row kernel wrote Y and U), so that the state array holds
  // V whatever iteration the solve stopped at.
  if (mode == kFwdYU && ctrl->done) return;
  if (mode == kFwdYUtoV && ctrl->done && ctrl->iter != iter) return;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int rw = lane / T;
  const int t = lane & (T - 1);
  const int partner = rw * T + ((T - t) & (T - 1));
  const int row0 = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv) * RPW;
  const int R = g.R;
  const int m = row0 >> g.lgR;
  const int r = (row0 & (R - 1)) + rw;
  float2* L = lds + (wv * RPW + rw) * LP;
  const long long rowoff = (long long)row0 * C;
  const int lo = rw * LR + t;
  float2 z[E];
  if (mode == kFwdReal) {
    const float2* S2 = reinterpret_cast<const float2*>(src + rowoff);
#pragma unroll
    for (int e = 0; e < E; ++e) z[e] = S2[lo + T * e];
  } else {
    const float rsf = ctrl->rsf;
    float2* Y2 = reinterpret_cast<float2*>(Y + rowoff);
    float2* U2 = reinterpret_cast<float2*>(U + rowoff);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float2 y = Y2[lo + T * e];
      float2 u = U2[lo + T * e];
      if (rsf != 1.f) u = make_float2(u.x / rsf, u.y / rsf);
      if (mode == kFwdYUtoV)  // V with Y = shrink(V), U / rsf = V - Y
        Y2[lo + T * e] = make_float2(y.x + u.x, y.y + u.y);
      else if (rsf != 1.f)
        U2[lo + T * e] = u;
      z[e] = make_float2(y.x - u.x, y.y - u.y);
    }
  }
  reg::fft_wave<LR, E, 1>(z, t, L, twC, 2);
  store_spectrum<LR, TC>(z, t, partner, r, R, F + (long long)m * g.Hc, twC);
}

// --------------------------------------------------------------------------
// Dispatch
// --------------------------------------------------------------------------
namespace {

bool is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

// Panel width: 16 columns up to R = 512; for longer columns fewer, so that
// a column block keeps 512 threads and two blocks fit in one CU's LDS.
int col_tc(int R) { return R <= 512 ? 16 : (R == 1024 ? 8 : 4); }

// Inverse real row transform of half spectra (panel layout) into real
// rows, times scale: one or more rows per wavefront as in the row kernel.
template <int LR, int TC>
__global__ __launch_bounds__(256) void k_row_inv_pow2(Geom g,
                                                      const float2* F,
                                                      float* out,
                                                      const float2* twC,
                                                      float scale) {
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
  float2 z[E];
  load_spectrum_c2r<LR, TC>(z, t, partner, r, R, F + (long long)m * g.Hc,
                            twC);
  reg::fft_wave<LR, E, -1>(z, t, L, twC, 2);
  float2* O2 = reinterpret_cast<float2*>(out + (long long)row0 * C);
  const int lo = rw * LR + t;
#pragma unroll
  for (int e = 0; e < E; ++e) O2[lo + T * e] = cscale(z[e], scale);
}

template <int LR, int TC>
hipError_t row_launch(const Geom& g, const Buffers& b, const Params& p,
                      int mode, int* nblocks, hipStream_t st) {
  const unsigned nb = (unsigned)((long long)g.M * g.R / RowCfg<LR>::RPB);
  const bool general = p.m0 || p.l1_mode == 2 || p.nonneg || p.nobndry;
  if (general && p.vstate)
    hipLaunchKernelGGL((k_row_pow2<LR, TC, true, 0, true>), dim3(nb),
                       dim3(256), 0, st, g, b, p, mode);
  else if (general)
    hipLaunchKernelGGL((k_row_pow2<LR, TC, true, 0, false>), dim3(nb),
                       dim3(256), 0, st, g, b, p, mode);
  else if (b.tbits != nullptr)
    hipLaunchKernelGGL((k_row_pow2<LR, TC, false, 1, false>), dim3(nb),
                       dim3(256), 0, st, g, b, p, mode);
  else if (b.vout != nullptr)
    hipLaunchKernelGGL((k_row_pow2<LR, TC, false, 2, false>), dim3(nb),
                       dim3(256), 0, st, g, b, p, mode);
  else if (p.vstate)
    hipLaunchKernelGGL((k_row_pow2<LR, TC, false, 0, true>), dim3(nb),
                       dim3(256), 0, st, g, b, p, mode);
  else
    hipLaunchKernelGGL((k_row_pow2<LR, TC, false, 0, false>), dim3(nb),
                       dim3(256), 0, st, g, b, p, mode);
  *nblocks = (int)nb * 4;  // partial sums per wavefront
  return hipGetLastError();
}

template <int LR, int TC>
hipError_t row_fwd_launch(const Geom& g, int nmaps, int mode, const float* src,
                          float* Y, float* U, float2* F, const float2* twC,
                          const Ctrl* ctrl, hipStream_t st, int iter) {
  const unsigned nb = (unsigned)((long long)nmaps * g.R / RowCfg<LR>::RPB);
  hipLaunchKernelGGL((k_row_fwd_pow2<LR, TC>), dim3(nb), dim3(256), 0, st, g,
                     mode, src, Y, U, F, twC, ctrl, iter);
  return hipGetLastError();
}

// (C/2, TC) combinations: C in 256..2048, TC from R.
#define SCC_ROW_CASES(X)                                      \
  X(128, 16) X(128, 8) X(128, 4) X(256, 16) X(256, 8) X(256, 4) \
  X(512, 16) X(512, 8) X(512, 4) X(1024, 16) X(1024, 8) X(1024, 4)

}  // namespace

bool pow2_select(int R, int C, int M, Pow2Config* cfg) {
  if (!is_pow2(R) || !is_pow2(C)) return false;
  if (R < 256 || R > 2048 || C < 256 || C > 2048) return false;
  if (M < 1 || M > 256) return false;
  cfg->TC = col_tc(R);
  cfg->rows = C / 2 <= 512 ? 64 / ((C / 2) / 8) : 1;  // rows per wavefront
  int G = 1;
  while (G * 16 < M) G *= 2;
  cfg->variant = G;
  return true;
}

int pow2_defer_blocks(const Geom& g, int nmaps) {
  return std::min(64, g.P * nmaps);
}

int pow2_row_blocks(const Geom& g, const Pow2Config& cfg) {
  return (int)((long long)g.M * g.R / cfg.rows);  // one partial per wavefront
}

hipError_t pow2_row(const Geom& g, const Pow2Config& cfg, const Buffers& b,
                    const Params& p, int mode, int* nblocks, hipStream_t st) {
  const int LR = g.C / 2;
#define X(lr, tc)                                                   \
  if (LR == lr && cfg.TC == tc)                                     \
    return row_launch<lr, tc>(g, b, p, mode, nblocks, st);
  SCC_ROW_CASES(X)
#undef X
  return hipErrorNotSupported;
}

hipError_t pow2_row_fwd(const Geom& g, const Pow2Config& cfg, int nmaps,
                        int mode, const float* src, float* Y, float* U,
                        float2* F, const float2* twC, const Ctrl* ctrl,
                        hipStream_t st, int iter) {
  const int LR = g.C / 2;
#define X(lr, tc)                                                           \
  if (LR == lr && cfg.TC == tc)                                             \
    return row_fwd_launch<lr, tc>(g, nmaps, mode, src, Y, U, F, twC, ctrl,  \
                                  st, iter);
  SCC_ROW_CASES(X)
#undef X
  return hipErrorNotSupported;
}

hipError_t pow2_row_inv(const Geom& g, const Pow2Config& cfg, int nmaps,
                        const float2* F, float* out, const float2* twC,
                        float scale, hipStream_t st) {
  const int LR = g.C / 2;
#define X(lr, tc)                                                           \
  if (LR == lr && cfg.TC == tc) {                                           \
    const unsigned nb =                                                     \
        (unsigned)((long long)nmaps * g.R / RowCfg<lr>::RPB);               \
    hipLaunchKernelGGL((k_row_inv_pow2<lr, tc>), dim3(nb), dim3(256), 0,    \
                       st, g, F, out, twC, scale);                          \
    return hipGetLastError();                                               \
  }
  SCC_ROW_CASES(X)
#undef X
  return hipErrorNotSupported;
}

hipError_t pow2_col(const Geom& g, const Pow2Config& cfg, int nmaps,
                    float2* F, const float2* twR, int inverse,
                    const Ctrl* ctrl, hipStream_t st, const float* parts,
                    int nparts, double* bsum) {
  const dim3 grid(g.P, nmaps);
  const int nred = pow2_defer_blocks(g, nmaps);
#define X(r, tc)                                                          \
  if (g.R == r && cfg.TC == tc) {                                         \
    const dim3 blk(tc * (r / kColE));                                     \
    if (inverse)                                                          \
      hipLaunchKernelGGL((k_col_pow2<r, tc, -1>), grid, blk, 0, st, g, F, \
                         twR, ctrl, parts, nparts, bsum, nred);           \
    else                                                                  \
      hipLaunchKernelGGL((k_col_pow2<r, tc, 1>), grid, blk, 0, st, g, F,  \
                         twR, ctrl, parts, nparts, bsum, nred);           \
    return hipGetLastError();                                             \
  }
  X(256, 16) X(512, 16) X(1024, 8) X(2048, 4)
#undef X
  return hipErrorNotSupported;
}

// The solve is independent of the image size and the spectrum layout
// (frequency f of map m at m Hc + f, GHG at f), so the generic-size solver
// uses it too for M <= 256: G lanes per frequency, 16 maps per lane.
bool solve_regs_fits(int M) { return M >= 1 && M <= 256; }

hipError_t solve_regs(const Geom& g, const Buffers& b, const Params& p,
                      hipStream_t st, int ctl_iter, int nbsum) {
  Pow2Config cfg{};
  int G = 1;
  while (G * 16 < g.M) G *= 2;
  cfg.variant = G;
  return pow2_solve(g, cfg, b, p, st, 1, ctl_iter, nbsum);
}

hipError_t pow2_solve(const Geom& g, const Pow2Config& cfg, const Buffers& b,
                      const Params& p, hipStream_t st, int nimg, int ctl_iter,
                      int nbsum) {
  const int G = cfg.variant;
  const long long fpb = 4LL * (64 / G);  // frequencies per 256-thread block
  const dim3 nb((unsigned)((g.Hc + fpb - 1) / fpb), (unsigned)nimg);
  // The filter spectra bypass the last-level cache (256 MB) when one image
  // reads them once per launch and the iteration's arrays (state, work and
  // filter spectra) exceed that cache: the work spectrum then stays there
  // for the next kernels (X-solve 161 -> 142 us and inverse columns 94 ->
  // 89 us at 1024 x 1024 x 64; rows 320 -> 296 us at 616 x 744).  Smaller
  // problems, and batches of images sharing the spectra, keep them cached
  // across iterations or images and read them normally (non-temporal loads
  // measured 10-20 % slower solves there).
  const double iter_bytes = 4.0 * g.M * (double)g.N + 16.0 * g.M * (double)g.Hc;
  const bool ntd = nimg == 1 && iter_bytes > 256.0 * 1024 * 1024;
#define X(gg, NTD)                                                          \
  if (G == gg && ntd == NTD) {                                             \
    if (p.grad)                                                            \
      hipLaunchKernelGGL((k_solve_pow2<gg, 16, true, NTD>), dim3(nb),       \
                         dim3(256), 0, st, g, b, p, ctl_iter, nbsum);      \
    else                                                                   \
      hipLaunchKernelGGL((k_solve_pow2<gg, 16, false, NTD>), dim3(nb),      \
                         dim3(256), 0, st, g, b, p, ctl_iter, nbsum);      \
    return hipGetLastError();                                              \
  }
  // 64 < M <= 256: maps split over the block's wavefronts (whole-line
  // reads), 8 or 16 maps per lane.
  if (G == 8 || G == 16) {
    const dim3 nbs((unsigned)((g.Hc + 15) / 16), (unsigned)nimg);
#define S(MP, NTD)                                                          \
  if (G == MP && ntd == NTD) {                                             \
    if (p.grad)                                                            \
      hipLaunchKernelGGL((k_solve_split<MP, true, NTD>), dim3(nbs),         \
                         dim3(256), 0, st, g, b, p, ctl_iter, nbsum);      \
    else                                                                   \
      hipLaunchKernelGGL((k_solve_split<MP, false, NTD>), dim3(nbs),        \
                         dim3(256), 0, st, g, b, p, ctl_iter, nbsum);      \
    return hipGetLastError();                                              \
  }
    S(8, false) S(8, true) S(16, false) S(16, true)
#undef S
  }
  X(1, false) X(2, false) X(4, false) X(8, false) X(16, false)
  X(1, true) X(2, true) X(4, true) X(8, true) X(16, true)
#undef X
  return hipErrorNotSupported;
}

}  // namespace scc
