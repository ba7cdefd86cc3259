// Differentiable unrolled plain CBPDN (power-of-two images): forward with
// iterate checkpoints, and the hand-written adjoint of every stage.
//
// Forward: the solver's own kernels over B images of one size sharing the
// dictionary (column and row kernels run over B*M maps, the solve takes the
// image index), fixed rho, no stopping test; (Y, U) is saved every k
// iterations.  Backward, per checkpoint segment from last to first: the
// segment is recomputed from its checkpoint with a tape (X-hat after each
// solve, V = Xr + U from each row kernel), then every iteration is reversed:
//   k_adj_row1   shrink backward (support mask), relaxation backward, and the
//                forward row transform of Xbar = alpha Vbar
//   k_col_pow2   forward column transform            (adjoint of the inverse FFT)
//   k_adj_solve  Sherman-Morrison solve of the adjoint right-hand side, the
//                dictionary- and image-spectrum gradients, rho Bbar / w
//   k_col_pow2   inverse column transform            (adjoint of the forward FFT)
//   k_adj_row2   inverse row transform, Ybar += Abar, Ubar -= Abar
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

// Shrink / dual / relaxation backward on one row, then the forward real FFT
// of Xbar into the row spectra of F.  Vbar = Ubar + (Ybar - Ubar) 1[|V| > t];
// tbar += (Ubar - Ybar) sign(V) 1[|V| > t]; Xbar = a Vbar;
// Ybar <- (1 - a) Vbar; Ubar <- Vbar.
template <int LR, int TC>
__global__ __launch_bounds__(256) void k_adj_row1(Geom g, float* Yb, float* Ub,
                                                  const float* V, float thr,
                                                  float alpha, float2* F,
                                                  const float2* twC,
                                                  float* lpart) {
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
  float2* Y2 = reinterpret_cast<float2*>(Yb + base);
  float2* U2 = reinterpret_cast<float2*>(Ub + base);
  const float2* V2 = reinterpret_cast<const float2*>(V + base);
  const int lo = rw * LR + t;
  const float a1 = 1.f - alpha;
  float lacc = 0.f;
  float2 z[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float2 v = V2[lo + T * e];
    const float2 yb = Y2[lo + T * e];
    const float2 ub = U2[lo + T * e];
    const bool m0 = fabsf(v.x) > thr, m1 = fabsf(v.y) > thr;
    const float vb0 = m0 ? yb.x : ub.x;
    const float vb1 = m1 ? yb.y : ub.y;
    if (m0) lacc += (ub.x - yb.x) * (v.x < 0.f ? -1.f : 1.f);
    if (m1) lacc += (ub.y - yb.y) * (v.y < 0.f ? -1.f : 1.f);
    Y2[lo + T * e] = make_float2(a1 * vb0, a1 * vb1);
    U2[lo + T * e] = make_float2(vb0, vb1);
    z[e] = make_float2(alpha * vb0, alpha * vb1);
  }
  const float ls = wave_sum(lacc);
  if (lane == 0) lpart[wave] = ls;
  reg::fft_wave<LR, E, 1>(z, t, L, twC, 2);
  store_spectrum<LR, TC>(z, t, partner, r, R, F + (long long)m * g.Hc, twC);
}

// Adjoint X-solve.  Per frequency, with D the dictionary spectra, Xh the
// forward X-hat of the iteration (tape) and Xhbar = w F / N:
//   Bbar = (rho I + D^H D)^-1 Xhbar                  (Sherman-Morrison)
//   Dbar_m += -conj(Xh_m) sum_j D_j Bbar_j - conj(Bbar_m) sum_j D_j Xh_j
//             + conj(Bbar_m) S
//   Sbar   += sum_j D_j Bbar_j
//   F_m     = rho Bbar_m / w
// G lanes per frequency, MPT maps per lane; the images of the batch are
// looped inside so that Dbar is read and written once per launch.
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
  for (int img = 0; img < nimg; ++img) {
    float2* Fi = F + (long long)img * M * g.Hc;
    const float2* Xi = Xh + (long long)img * M * g.Hc;
    float2 xb[MPT], xh[MPT];
    float2 num = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      const int m = gi * MPT + i;
      xb[i] = make_float2(0.f, 0.f);
      xh[i] = make_float2(0.f, 0.f);
      if (valid && m < M) {
        const long long o = (long long)m * g.Hc + f;
        xb[i] = cscale(Fi[o], sin_);
        xh[i] = Xi[o];
        num = cadd(num, cmul(dd[i], xb[i]));
      }
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
    const float2 sh =
        valid ? Sf[(long long)img * g.Hc + f] : make_float2(0.f, 0.f);
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
    if (valid && gi == 0) {
      float2* s = Shb + (long long)img * g.Hc + f;
      *s = cadd(*s, sDB);
    }
  }
#pragma unroll
  for (int i = 0; i < MPT; ++i) {
    const int m = gi * MPT + i;
    if (valid && m < M) {
      float2* d = Dhb + (long long)m * g.Hc + f;
      *d = cadd(*d, dacc[i]);
    }
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

// acc += sum of n partials (one block, fixed order, FP64).
__global__ __launch_bounds__(1024) void k_sum_parts(const float* part, int n,
                                                    double* acc) {
  __shared__ double red[16];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += (double)part[i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    *acc += s;
  }
}

// --------------------------------------------------------------------------
// Launch helpers
// --------------------------------------------------------------------------
namespace {

#define SCC_UNR_CASES(X)                                               \
  X(128, 16) X(128, 8) X(128, 4) X(256, 16) X(256, 8) X(256, 4) X(512, 16) \
      X(512, 8) X(512, 4) X(1024, 16) X(1024, 8) X(1024, 4)

hipError_t adj_row1(const Geom& gbm, const Pow2Config& cfg, float* Yb,
                    float* Ub, const float* V, float thr, float alpha,
                    float2* F, const float2* twC, float* lpart, int* nparts,
                    hipStream_t st) {
  const int LR = gbm.C / 2;
#define X(lr, tc)                                                          \
  if (LR == lr && cfg.TC == tc) {                                          \
    const unsigned nb = (unsigned)((long long)gbm.M * gbm.R /              \
                                   RowCfg<lr>::RPB);                        \
    hipLaunchKernelGGL((k_adj_row1<lr, tc>), dim3(nb), dim3(256), 0, st,   \
                       gbm, Yb, Ub, V, thr, alpha, F, twC, lpart);         \
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

hipError_t adj_solve(const Geom& g, int nimg, float2* F, const float2* Xh,
                     const float2* Df, const float2* Sf, float2* Dhb,
                     float2* Shb, float rho, hipStream_t st) {
  int G = 1;
  while (G * 8 < g.M) G *= 2;
  const long long fpb = 4LL * (64 / G);
  const unsigned nb = (unsigned)((g.Hc + fpb - 1) / fpb);
  const float invN = 1.f / (float)g.N;
#define X(gg)                                                                \
  if (G == gg) {                                                             \
    hipLaunchKernelGGL((k_adj_solve<gg, 8>), dim3(nb), dim3(256), 0, st, g,  \
                       nimg, F, Xh, Df, Sf, Dhb, Shb, rho, invN);            \
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
  UBuf tXh, tV, Yb, Ub, Dhb, Shb, lacc, tmp;
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
  const size_t nwaves = (size_t)pow2_row_blocks(w->gbm, w->cfg);
  int rc;
  if ((rc = w->F.get(nc * 8)) || (rc = w->Df.get((size_t)M * g.Hc * 8)) ||
      (rc = w->Sf.get((size_t)B * g.Hc * 8)) || (rc = w->Y.get(nr * 4)) ||
      (rc = w->U.get(nr * 4)) || (rc = w->part.get(nwaves * 5 * 4)) ||
      (rc = w->bsum.get(64 * 5 * 8)) || (rc = w->ctrl.get(sizeof(Ctrl))) ||
      (rc = w->twR.get((size_t)R * 8)) || (rc = w->twC.get((size_t)C * 8)) ||
      (rc = w->tXh.get((size_t)k * nc * 8)) ||
      (rc = w->tV.get((size_t)k * nr * 4)) || (rc = w->Yb.get(nr * 4)) ||
      (rc = w->Ub.get(nr * 4)) || (rc = w->Dhb.get((size_t)M * g.Hc * 8)) ||
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

// One forward iteration over the batch; F holds the row spectra of Y - U.
// With a tape slot, X-hat after the solve and V are recorded.
int fwd_iter(UWs* w, const Params& p, float2* tXh, float* tV,
             hipStream_t st) {
  Buffers b = make_buffers(w);
  b.vout = tV;
  const int BM = w->B * w->M;
  int nparts = 0;
  UHIP(pow2_col(w->g, w->cfg, BM, b.F, b.twR, 0, b.ctrl, st));
  UHIP(pow2_solve(w->g, w->cfg, b, p, st, w->B));
  if (tXh != nullptr)
    UHIP(hipMemcpyAsync(tXh, b.F, (size_t)BM * w->g.Hc * 8,
                        hipMemcpyDeviceToDevice, st));
  UHIP(pow2_col(w->g, w->cfg, BM, b.F, b.twR, 1, b.ctrl, st));
  UHIP(pow2_row(w->gbm, w->cfg, b, p, kRowFused, &nparts, st));
  return SCC_OK;
}

}  // namespace
}  // namespace scc

using namespace scc;

extern "C" {

size_t scc_unrolled_ckpt_floats(int M, int B, int R, int C, int n_iter,
                                int ckpt_every) {
  if (M < 1 || B < 1 || R < 1 || C < 1 || n_iter < 1 || ckpt_every < 1)
    return 0;
  const size_t nseg = (size_t)(n_iter + ckpt_every - 1) / ckpt_every;
  return nseg * 2 * (size_t)B * M * R * C;
}

int scc_unrolled_forward(const float* D, int M, int K, const float* S, int B,
                         int R, int C, float lambda, float rho, float alpha,
                         int n_iter, int ckpt_every, float* Y, float* ckpt,
                         void* stream) {
  int rc = check_shapes(M, K, B, R, C, n_iter, ckpt_every, rho);
  if (rc) return rc;
  if (!D || !S || !Y || !ckpt) return ufail(SCC_ERR_INVALID, "NULL pointer");
  std::lock_guard<std::mutex> lk(u_mu);
  UWs* w;
  if ((rc = get_uws(R, C, M, B, ckpt_every, &w))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = setup(w, D, K, S, rho, st))) return rc;
  const size_t nr = (size_t)B * M * w->g.N;
  UHIP(hipMemsetAsync(w->Y.p, 0, nr * 4, st));
  UHIP(hipMemsetAsync(w->U.p, 0, nr * 4, st));
  UHIP(hipMemsetAsync(w->F.p, 0, (size_t)B * M * w->g.Hc * 8, st));
  const Params p = make_params(lambda, alpha);
  for (int i = 0; i < n_iter; ++i) {
    if (i % ckpt_every == 0) {
      float* c = ckpt + (size_t)(i / ckpt_every) * 2 * nr;
      UHIP(hipMemcpyAsync(c, w->Y.p, nr * 4, hipMemcpyDeviceToDevice, st));
      UHIP(hipMemcpyAsync(c + nr, w->U.p, nr * 4, hipMemcpyDeviceToDevice,
                          st));
    }
    if ((rc = fwd_iter(w, p, nullptr, nullptr, st))) return rc;
  }
  UHIP(hipMemcpyAsync(Y, w->Y.p, nr * 4, hipMemcpyDeviceToDevice, st));
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
  const size_t nr = (size_t)B * M * g.N;
  const size_t nc = (size_t)B * M * g.Hc;
  const int BM = B * M;
  Buffers b = make_buffers(w);
  const Params p = make_params(lambda, alpha);
  float* Yb = (float*)w->Yb.p;
  float* Ub = (float*)w->Ub.p;
  float2* Dhb = (float2*)w->Dhb.p;
  float2* Shb = (float2*)w->Shb.p;
  double* lacc = (double*)w->lacc.p;
  UHIP(hipMemcpyAsync(Yb, gY, nr * 4, hipMemcpyDeviceToDevice, st));
  UHIP(hipMemsetAsync(Ub, 0, nr * 4, st));
  UHIP(hipMemsetAsync(Dhb, 0, (size_t)M * g.Hc * 8, st));
  UHIP(hipMemsetAsync(Shb, 0, (size_t)B * g.Hc * 8, st));
  UHIP(hipMemsetAsync(lacc, 0, 8, st));
  const float thr = lambda / rho;
  const int nseg = (n_iter + ckpt_every - 1) / ckpt_every;
  for (int s = nseg - 1; s >= 0; --s) {
    const int it0 = s * ckpt_every;
    const int cnt = std::min(n_iter, it0 + ckpt_every) - it0;
    // Restore the segment's start state and its first row spectra.
    const float* c = ckpt + (size_t)s * 2 * nr;
    UHIP(hipMemcpyAsync(b.Y, c, nr * 4, hipMemcpyDeviceToDevice, st));
    UHIP(hipMemcpyAsync(b.U, c + nr, nr * 4, hipMemcpyDeviceToDevice, st));
    UHIP(pow2_row_fwd(w->gbm, w->cfg, BM, kFwdYU, nullptr, b.Y, b.U, b.F,
                      b.twC, b.ctrl, st));
    for (int j = 0; j < cnt; ++j)
      if ((rc = fwd_iter(w, p, (float2*)w->tXh.p + (size_t)j * nc,
                         (float*)w->tV.p + (size_t)j * nr, st)))
        return rc;
    for (int j = cnt - 1; j >= 0; --j) {
      int nparts = 0;
      UHIP(adj_row1(w->gbm, w->cfg, Yb, Ub, (float*)w->tV.p + (size_t)j * nr,
                    thr, alpha, b.F, b.twC, b.partials, &nparts, st));
      hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, st, b.partials,
                         nparts, lacc);
      UHIP(hipGetLastError());
      UHIP(pow2_col(g, w->cfg, BM, b.F, b.twR, 0, nullptr, st));
      UHIP(adj_solve(g, B, b.F, (float2*)w->tXh.p + (size_t)j * nc, b.Df,
                     b.Sf, Dhb, Shb, rho, st));
      UHIP(pow2_col(g, w->cfg, BM, b.F, b.twR, 1, nullptr, st));
      UHIP(adj_row2(w->gbm, w->cfg, b.F, Yb, Ub, nullptr, b.twC, 0, st));
    }
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
  const size_t nr = (size_t)B * M * w->g.N;
  UHIP(hipMemsetAsync(w->Y.p, 0, nr * 4, st));
  UHIP(hipMemsetAsync(w->U.p, 0, nr * 4, st));
  UHIP(hipMemsetAsync(w->F.p, 0, (size_t)B * M * w->g.Hc * 8, st));
  const Params p = make_params(lambda, alpha);
  for (int i = 0; i < n_iter; ++i)
    if ((rc = fwd_iter(w, p, nullptr, V + (size_t)i * nr, st))) return rc;
  UHIP(hipStreamSynchronize(st));
  return SCC_OK;
}

const char* scc_unrolled_last_error(void) { return u_err.c_str(); }

}  // extern "C"
