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
#include "scc_device.h"
#include "scc_pow2.h"

namespace scc {

namespace {

constexpr int kRowE = 8;  // points per thread in the row transforms
constexpr int kColE = 16;  // points per thread in the column transforms

// One element of the fused update on a value x (= X, already / N).
__device__ __forceinline__ void upd1(const Params& p, float x, float yo,
                                     float uo, float thr, int amsw,
                                     bool ams, bool zero_bndry, float& yn,
                                     float& un, UpdateSums& acc) {
  const float xr = (p.alpha == 1.f) ? x : p.alpha * x + (1.f - p.alpha) * yo;
  const float v = xr + uo;
  if (ams) {
    yn = (float)(1 - amsw) * v;
  } else {
    const float a = fabsf(v) - thr;
    yn = copysignf(fmaxf(a, 0.f), v);
    if (p.nonneg) yn = fmaxf(yn, 0.f);
    if (zero_bndry) yn = 0.f;
  }
  un = v - yn;
  acc.x += x * x;
  acc.y += yn * yn;
  acc.u += un * un;
  const float dxy = x - yn;
  acc.r += dxy * dxy;
  const float dy = yo - yn;
  acc.s += dy * dy;
}

// Global address of element i (panel order) of the NR-row block starting at
// row r0 of map base `mb`, and the (row, column) it holds.
template <int LR, int NR, int TC>
__device__ __forceinline__ long long row_spec_addr(int i, int R, int r0,
                                                   int& rr, int& c) {
  constexpr int P0 = LR / TC;  // full panels; the last one is column LR
  constexpr int PER = NR * TC;
  if (i < P0 * PER) {
    const int pp = i / PER;
    const int rem = i - pp * PER;
    rr = rem / TC;
    c = pp * TC + (rem - rr * TC);
    return (long long)pp * TC * R + (long long)r0 * TC + rem;
  }
  rr = i - P0 * PER;
  c = LR;
  return (long long)P0 * TC * R + r0 + rr;
}

}  // namespace

// --------------------------------------------------------------------------
// Column transforms
// --------------------------------------------------------------------------
template <int R, int TC, int DIR>
__global__ __launch_bounds__(TC*(R / kColE)) void k_col_pow2(
    Geom g, float2* __restrict__ F, const float2* __restrict__ twR,
    const Ctrl* ctrl) {
  constexpr int E = kColE;
  constexpr int T = R / E;
  constexpr int LP = reg::lds_len(R);
  __shared__ float2 lds[TC * LP];
  if (ctrl != nullptr && ctrl->done) return;
  const int p = blockIdx.x;
  const long long m = blockIdx.y;
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
  reg::fft_dev<R, E, DIR>(v, t, L, twR, 1);
  if (on) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int r = t + T * e;
      base[r * w + c] = L[reg::lpad(r)];
    }
  }
}

// --------------------------------------------------------------------------
// Sherman-Morrison X-solve: G lanes per frequency, MPT maps per lane.
// --------------------------------------------------------------------------
template <int G, int MPT>
__global__ __launch_bounds__(256) void k_solve_pow2(Geom g, Buffers b,
                                                    Params p) {
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
  const int M = g.M;
  float2 sh = make_float2(0.f, 0.f);
  float gh = 0.f;
  if (valid) {
    sh = b.Sf[f];
    if (p.grad) gh = p.mu * b.ghg[f];
  }
  float2 bb[MPT], dd[MPT];
  float2 num = make_float2(0.f, 0.f);
  float den = 0.f;
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = gi * MPT + i;
    if (valid && m < M) {
      const long long o = (long long)m * g.Hc + f;
      const float2 d = b.Df[o];
      const float2 a = b.F[o];
      const float2 x = cadd(cmul(cconj(d), sh), cscale(a, rho));
      float inv;
      if (p.grad)
        inv = 1.f / ((p.gw_mode ? b.gw[m] : p.gw_scalar) * gh + rho);
      else
        inv = 1.f / rho;
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
      float delta = rho;
      if (p.grad) delta = (p.gw_mode ? b.gw[m] : p.gw_scalar) * gh + rho;
      const float2 x = csub(bb[i], cmul(cconj(dd[i]), cb));
      b.F[(long long)m * g.Hc + f] = make_float2(x.x / delta, x.y / delta);
    }
  }
}

// --------------------------------------------------------------------------
// Row kernel
// --------------------------------------------------------------------------
template <int LR, int NR, int TC>
__global__ __launch_bounds__(NR*(LR / kRowE)) void k_row_pow2(Geom g,
                                                              Buffers b,
                                                              Params p,
                                                              int mode) {
  constexpr int E = kRowE;
  constexpr int T = LR / E;
  constexpr int NT = NR * T;
  constexpr int C = 2 * LR;
  constexpr int H = LR + 1;
  constexpr int LP = reg::lds_len(LR);
  constexpr int NW = NT / 64;
  constexpr int NF4 = E / 2;  // float4 of Y / U per thread
  __shared__ float2 lds[NR * LP];
  __shared__ float red[5][NW];
  if (b.ctrl->done) return;
  const int tid = threadIdx.x;
  const int rl = tid / T;
  const int t = tid - rl * T;
  const int R = g.R;
  const long long grow0 = (long long)blockIdx.x * NR;
  const int m = (int)(grow0 / R);  // rows of a block share the map
  const int r0 = (int)(grow0 - (long long)m * R);
  const int r = r0 + rl;
  const long long rowoff = (grow0 + rl) * C;
  float2* L = lds + rl * LP;
  float2* Fm = b.F + (long long)m * g.Hc;

  // Prefetch Y and U of this thread's float4 columns.
  float4 y4[NF4], u4[NF4];
  const float4* Yr = reinterpret_cast<const float4*>(b.Y + rowoff);
  const float4* Ur = reinterpret_cast<const float4*>(b.U + rowoff);
#pragma unroll
  for (int i = 0; i < NF4; ++i) {
    y4[i] = Yr[t + T * i];
    u4[i] = Ur[t + T * i];
  }
  // 1. Half spectra of the NR rows into LDS.
  for (int i = tid; i < NR * H; i += NT) {
    int rr, c;
    const long long a = row_spec_addr<LR, NR, TC>(i, R, r0, rr, c);
    lds[rr * LP + reg::lpad(c)] = Fm[a];
  }
  __syncthreads();
  // 2. Inverse real-FFT pre-processing into this thread's slots:
  //    Z'[k] = (X[k] + conj X[L-k]) + i W_C^-k (X[k] - conj X[L-k]).
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + T * e;
    float2 xk = L[reg::lpad(k)];
    float2 xn = L[reg::lpad(LR - k)];
    if (k == 0) {
      xk.y = 0.f;
      xn.y = 0.f;
    }
    xn.y = -xn.y;
    const float2 s = cadd(xk, xn);
    const float2 wd = cmul(cconj(b.twC[k]), csub(xk, xn));
    v[e] = make_float2(s.x - wd.y, s.y + wd.x);
  }
  __syncthreads();
  reg::fft_dev<LR, E, -1>(v, t, L, b.twC, 2);

  // 4. Fused update on X = z / N (z interleaves x[2n], x[2n+1]).
  const float rho = b.ctrl->rho;
  const float invN = 1.0f / (float)g.N;
  const float lr = p.lambda / rho;
  const bool ams = m < p.m0;
  const float wmap = p.l1_mode == 0 ? p.l1_scalar
                     : (p.l1_mode == 1 && !ams ? b.l1w[m] : 0.f);
  const bool bndry_row = p.nobndry && r > R - p.Kh;
  UpdateSums acc;
  float4* Yw = reinterpret_cast<float4*>(b.Y + rowoff);
  float4* Uw = reinterpret_cast<float4*>(b.U + rowoff);
#pragma unroll
  for (int i = 0; i < NF4; ++i) {
    const int f4 = t + T * i;
    const int n0 = 4 * f4;
    const float2 z0 = L[reg::lpad(2 * f4)];
    const float2 z1 = L[reg::lpad(2 * f4 + 1)];
    const float xs[4] = {z0.x * invN, z0.y * invN, z1.x * invN, z1.y * invN};
    const float ys[4] = {y4[i].x, y4[i].y, y4[i].z, y4[i].w};
    const float us[4] = {u4[i].x, u4[i].y, u4[i].z, u4[i].w};
    float th[4] = {lr * wmap, lr * wmap, lr * wmap, lr * wmap};
    int mw[4] = {0, 0, 0, 0};
    if (ams) {
      const int4 w4 = reinterpret_cast<const int4*>(b.mask + (long long)r * C)[f4];
      mw[0] = w4.x;
      mw[1] = w4.y;
      mw[2] = w4.z;
      mw[3] = w4.w;
    } else if (p.l1_mode == 2) {
      const float4 w4 = reinterpret_cast<const float4*>(b.l1w + rowoff)[f4];
      th[0] = lr * w4.x;
      th[1] = lr * w4.y;
      th[2] = lr * w4.z;
      th[3] = lr * w4.w;
    }
    float yn[4], un[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool zb = p.nobndry && (bndry_row || n0 + j > C - p.Kw);
      upd1(p, xs[j], ys[j], us[j], th[j], mw[j], ams, zb, yn[j], un[j], acc);
    }
    Yw[f4] = make_float4(yn[0], yn[1], yn[2], yn[3]);
    Uw[f4] = make_float4(un[0], un[1], un[2], un[3]);
    L[reg::lpad(2 * f4)] = make_float2(yn[0] - un[0], yn[1] - un[1]);
    L[reg::lpad(2 * f4 + 1)] = make_float2(yn[2] - un[2], yn[3] - un[3]);
  }
  block_sums_store<NW>(acc, red, b.partials, gridDim.x, blockIdx.x, tid);
  if (mode != kRowFused) return;
  // 5. Forward transform of Y - U (the block barrier of block_sums_store
  //    orders the LDS writes above before these reads).
  reg::slots_read<LR, E>(v, t, L);
  __syncthreads();
  reg::fft_dev<LR, E, 1>(v, t, L, b.twC, 2);
  // 6. Real-FFT post-processing and store in panel order:
  //    X[k] = (Z[k] + conj Z[L-k]) / 2 - i W_C^k (Z[k] - conj Z[L-k]) / 2.
  for (int i = tid; i < NR * H; i += NT) {
    int rr, c;
    const long long a = row_spec_addr<LR, NR, TC>(i, R, r0, rr, c);
    const float2* Z = lds + rr * LP;
    const float2 zk = Z[reg::lpad(c == LR ? 0 : c)];
    const float2 zn = cconj(Z[reg::lpad(c == 0 ? 0 : LR - c)]);
    const float2 s = cadd(zk, zn);
    const float2 wd = cmul(b.twC[c], csub(zk, zn));
    Fm[a] = make_float2(0.5f * (s.x + wd.y), 0.5f * (s.y - wd.x));
  }
}

// Forward row transform of a real array (setup) or of Y - U / rsf.
template <int LR, int NR, int TC>
__global__ __launch_bounds__(NR*(LR / kRowE)) void k_row_fwd_pow2(
    Geom g, int mode, const float* __restrict__ src, float* Y, float* U,
    float2* __restrict__ F, const float2* __restrict__ twC,
    const Ctrl* ctrl) {
  constexpr int E = kRowE;
  constexpr int T = LR / E;
  constexpr int NT = NR * T;
  constexpr int C = 2 * LR;
  constexpr int H = LR + 1;
  constexpr int LP = reg::lds_len(LR);
  constexpr int NF4 = E / 2;
  __shared__ float2 lds[NR * LP];
  if (mode == kFwdYU && ctrl->done) return;
  const int tid = threadIdx.x;
  const int rl = tid / T;
  const int t = tid - rl * T;
  const int R = g.R;
  const long long grow0 = (long long)blockIdx.x * NR;
  const int m = (int)(grow0 / R);
  const int r0 = (int)(grow0 - (long long)m * R);
  const long long rowoff = (grow0 + rl) * C;
  float2* L = lds + rl * LP;
  const float rsf = (mode == kFwdYU) ? ctrl->rsf : 1.f;
#pragma unroll
  for (int i = 0; i < NF4; ++i) {
    const int f4 = t + T * i;
    float4 a;
    if (mode == kFwdReal) {
      a = reinterpret_cast<const float4*>(src + rowoff)[f4];
    } else {
      const float4 y = reinterpret_cast<const float4*>(Y + rowoff)[f4];
      float4 u = reinterpret_cast<const float4*>(U + rowoff)[f4];
      if (rsf != 1.f) {
        u = make_float4(u.x / rsf, u.y / rsf, u.z / rsf, u.w / rsf);
        reinterpret_cast<float4*>(U + rowoff)[f4] = u;
      }
      a = make_float4(y.x - u.x, y.y - u.y, y.z - u.z, y.w - u.w);
    }
    L[reg::lpad(2 * f4)] = make_float2(a.x, a.y);
    L[reg::lpad(2 * f4 + 1)] = make_float2(a.z, a.w);
  }
  __syncthreads();
  float2 v[E];
  reg::slots_read<LR, E>(v, t, L);
  __syncthreads();
  reg::fft_dev<LR, E, 1>(v, t, L, twC, 2);
  float2* Fm = F + (long long)m * g.Hc;
  for (int i = tid; i < NR * H; i += NT) {
    int rr, c;
    const long long a = row_spec_addr<LR, NR, TC>(i, R, r0, rr, c);
    const float2* Z = lds + rr * LP;
    const float2 zk = Z[reg::lpad(c == LR ? 0 : c)];
    const float2 zn = cconj(Z[reg::lpad(c == 0 ? 0 : LR - c)]);
    const float2 s = cadd(zk, zn);
    const float2 wd = cmul(twC[c], csub(zk, zn));
    Fm[a] = make_float2(0.5f * (s.x + wd.y), 0.5f * (s.y - wd.x));
  }
}

// --------------------------------------------------------------------------
// Dispatch
// --------------------------------------------------------------------------
namespace {

bool is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

int col_tc(int R) {
  return R <= 512 ? 16 : (R == 1024 ? 8 : 4);
}
// Rows per row-kernel block: 256 threads of kRowE points each.
int row_nr(int C) { return 256 / ((C / 2) / kRowE); }

template <int LR, int NR, int TC>
hipError_t row_launch(const Geom& g, const Buffers& b, const Params& p,
                      int mode, int* nblocks, hipStream_t st) {
  const unsigned nb = (unsigned)((long long)g.M * g.R / NR);
  hipLaunchKernelGGL((k_row_pow2<LR, NR, TC>), dim3(nb), dim3(NR * LR / kRowE),
                     0, st, g, b, p, mode);
  *nblocks = (int)nb;
  return hipGetLastError();
}

template <int LR, int NR, int TC>
hipError_t row_fwd_launch(const Geom& g, int nmaps, int mode, const float* src,
                          float* Y, float* U, float2* F, const float2* twC,
                          const Ctrl* ctrl, hipStream_t st) {
  const unsigned nb = (unsigned)((long long)nmaps * g.R / NR);
  hipLaunchKernelGGL((k_row_fwd_pow2<LR, NR, TC>), dim3(nb),
                     dim3(NR * LR / kRowE), 0, st, g, mode, src, Y, U, F, twC,
                     ctrl);
  return hipGetLastError();
}

// (C, TC) combinations: C in 256..2048, TC from R.
#define SCC_ROW_CASES(X)              \
  X(128, 16, 16) X(128, 16, 8) X(128, 16, 4) \
  X(256, 8, 16) X(256, 8, 8) X(256, 8, 4)    \
  X(512, 4, 16) X(512, 4, 8) X(512, 4, 4)    \
  X(1024, 2, 16) X(1024, 2, 8) X(1024, 2, 4)

}  // namespace

bool pow2_select(int R, int C, int M, Pow2Config* cfg) {
  if (!is_pow2(R) || !is_pow2(C)) return false;
  if (R < 256 || R > 2048 || C < 256 || C > 2048) return false;
  if (M < 1 || M > 256) return false;
  cfg->TC = col_tc(R);
  cfg->rows = row_nr(C);
  int G = 1;
  while (G * 16 < M) G *= 2;
  cfg->variant = G;
  return true;
}

int pow2_row_blocks(const Geom& g, const Pow2Config& cfg) {
  return (int)((long long)g.M * g.R / cfg.rows);
}

hipError_t pow2_row(const Geom& g, const Pow2Config& cfg, const Buffers& b,
                    const Params& p, int mode, int* nblocks, hipStream_t st) {
  const int LR = g.C / 2;
#define X(lr, nr, tc)                                               \
  if (LR == lr && cfg.TC == tc)                                     \
    return row_launch<lr, nr, tc>(g, b, p, mode, nblocks, st);
  SCC_ROW_CASES(X)
#undef X
  return hipErrorNotSupported;
}

hipError_t pow2_row_fwd(const Geom& g, const Pow2Config& cfg, int nmaps,
                        int mode, const float* src, float* Y, float* U,
                        float2* F, const float2* twC, const Ctrl* ctrl,
                        hipStream_t st) {
  const int LR = g.C / 2;
#define X(lr, nr, tc)                                                       \
  if (LR == lr && cfg.TC == tc)                                             \
    return row_fwd_launch<lr, nr, tc>(g, nmaps, mode, src, Y, U, F, twC,    \
                                      ctrl, st);
  SCC_ROW_CASES(X)
#undef X
  return hipErrorNotSupported;
}

hipError_t pow2_col(const Geom& g, const Pow2Config& cfg, int nmaps,
                    float2* F, const float2* twR, int inverse,
                    const Ctrl* ctrl, hipStream_t st) {
  const dim3 grid(g.P, nmaps);
#define X(r, tc)                                                          \
  if (g.R == r && cfg.TC == tc) {                                         \
    const dim3 blk(tc * (r / kColE));                                     \
    if (inverse)                                                          \
      hipLaunchKernelGGL((k_col_pow2<r, tc, -1>), grid, blk, 0, st, g, F, \
                         twR, ctrl);                                      \
    else                                                                  \
      hipLaunchKernelGGL((k_col_pow2<r, tc, 1>), grid, blk, 0, st, g, F,  \
                         twR, ctrl);                                      \
    return hipGetLastError();                                             \
  }
  X(256, 16) X(512, 16) X(1024, 8) X(2048, 4)
#undef X
  return hipErrorNotSupported;
}

hipError_t pow2_solve(const Geom& g, const Pow2Config& cfg, const Buffers& b,
                      const Params& p, hipStream_t st) {
  const int G = cfg.variant;
  const long long fpb = 4LL * (64 / G);  // frequencies per 256-thread block
  const unsigned nb = (unsigned)((g.Hc + fpb - 1) / fpb);
#define X(gg)                                                                \
  if (G == gg) {                                                             \
    hipLaunchKernelGGL((k_solve_pow2<gg, 16>), dim3(nb), dim3(256), 0, st, g, \
                       b, p);                                                \
    return hipGetLastError();                                                \
  }
  X(1) X(2) X(4) X(8) X(16)
#undef X
  return hipErrorNotSupported;
}

}  // namespace scc
