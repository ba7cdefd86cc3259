// Reconstruction operator  s_b = sum_m d_m * x_{b,m}  (circular convolution,
// filters zero-padded at the origin) and its adjoint, on the GPU.  This is
// sporco's fft.fftconv use in the reference's examples
// (examples/big_cbpdn.py:69, the data term of the objective): both FFT
// kernel families are reused, the spectra are combined per frequency.
//
//   forward   R_b   = sum_m D_m X_{b,m}                 (half spectra)
//   adjoint   gX_bm = conj(D_m) G_b,  gD_m = sum_b conj(X_{b,m}) G_b
// followed by inverse transforms scaled by 1/N; gD is cropped to the
// filter support (the adjoint of the zero-padding).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sporco_amd.h"
#include "scc_device.h"
#include "scc_internal.h"
#include "scc_pow2.h"

namespace scc {

// Rf[b] = sum_m Df[m] Xf[b, m] per frequency; grid (frequencies, images).
__global__ __launch_bounds__(256) void k_rc_mul(long long Hc, int M,
                                                const float2* __restrict__ Df,
                                                const float2* __restrict__ Xf,
                                                float2* __restrict__ Rf) {
  const long long f = blockIdx.x * 256LL + threadIdx.x;
  if (f >= Hc) return;
  const long long b = blockIdx.y;
  const float2* X = Xf + b * M * Hc;
  float2 acc = make_float2(0.f, 0.f);
  for (int m = 0; m < M; ++m)
    acc = cadd(acc, cmul(Df[(long long)m * Hc + f], X[(long long)m * Hc + f]));
  Rf[b * Hc + f] = acc;
}

// gXf[b, m] = conj(Df[m]) Gf[b] (in place over Xf when gX is wanted),
// gDf[m] = sum_b conj(Xf[b, m]) Gf[b]; grid (frequencies, maps).
__global__ __launch_bounds__(256) void k_rc_adj(long long Hc, int M, int B,
                                                const float2* __restrict__ Df,
                                                float2* Xf,
                                                const float2* __restrict__ Gf,
                                                float2* __restrict__ gDf,
                                                int want_gx) {
  const long long f = blockIdx.x * 256LL + threadIdx.x;
  if (f >= Hc) return;
  const int m = blockIdx.y;
  const float2 d = cconj(Df[(long long)m * Hc + f]);
  float2 acc = make_float2(0.f, 0.f);
  for (int b = 0; b < B; ++b) {
    const long long o = ((long long)b * M + m) * Hc + f;
    const float2 g = Gf[(long long)b * Hc + f];
    acc = cadd(acc, cmul(cconj(Xf[o]), g));
    if (want_gx) Xf[o] = cmul(d, g);
  }
  if (gDf != nullptr) gDf[(long long)m * Hc + f] = acc;
}

// dst (M, Kh, Kw) = src (M, R, C) restricted to the top-left corner.
__global__ void k_rc_crop(const float* src, float* dst, int M, int Kh, int Kw,
                          int R, int C) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * Kh * Kw) return;
  const int m = i / (Kh * Kw), a = (i / Kw) % Kh, c = i % Kw;
  dst[i] = src[((long long)m * R + a) * C + c];
}

namespace {

struct RBuf {
  void* p = nullptr;
  size_t n = 0;
  ~RBuf() {
    if (p) (void)hipFree(p);
  }
  int get(size_t bytes) {
    if (p && n >= bytes) return SCC_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (bytes == 0) return SCC_OK;
    if (hipMalloc(&p, bytes) != hipSuccess)
      return set_error(SCC_ERR_NOMEM, "hipMalloc of " +
                                          std::to_string(bytes) + " bytes");
    n = bytes;
    return SCC_OK;
  }
};

#define RHIP(call)                                                         \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return set_error(                                                    \
          e_ == hipErrorOutOfMemory ? SCC_ERR_NOMEM : SCC_ERR_HIP,         \
          std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)

// Buffers and transforms of one (device, R, C, M, B).
struct RWs {
  int device = -1, R = 0, C = 0, M = 0, B = 0;
  Geom g{};
  bool pow2 = false;
  Pow2Config p2{};
  FftPlan rp{}, cp{};
  RBuf Df, Xf, Gf, gDf, tmp, twR, twC;
};

std::mutex r_mu;
std::unique_ptr<RWs> r_ws;

int get_rws(int R, int C, int M, int B, RWs** out) {
  int dev = 0;
  RHIP(hipGetDevice(&dev));
  if (r_ws && r_ws->device == dev && r_ws->R == R && r_ws->C == C &&
      r_ws->M == M && r_ws->B == B) {
    *out = r_ws.get();
    return SCC_OK;
  }
  r_ws.reset();
  auto w = std::make_unique<RWs>();
  w->device = dev;
  w->R = R;
  w->C = C;
  w->M = M;
  w->B = B;
  make_geom(R, C, M, &w->g, &w->pow2, &w->p2);
  const Geom& g = w->g;
  if (!make_plan(g.Lr, g.even ? 2 : 1, &w->rp, false) ||
      !make_plan(R, 1, &w->cp, true))
    return set_error(SCC_ERR_UNSUPPORTED, "FFT length has too many factors");
  const size_t hc = (size_t)g.Hc;
  int rc;
  if ((rc = w->Df.get((size_t)M * hc * 8)) ||
      (rc = w->Xf.get((size_t)B * M * hc * 8)) ||
      (rc = w->Gf.get((size_t)B * hc * 8)) ||
      (rc = w->gDf.get((size_t)M * hc * 8)) ||
      (rc = w->tmp.get((size_t)M * g.N * 4)) ||
      (rc = w->twR.get((size_t)R * 8)) || (rc = w->twC.get((size_t)C * 8)))
    return rc;
  std::vector<float2> tw;
  for (int pass = 0; pass < 2; ++pass) {
    const int L = pass == 0 ? R : C;
    tw.resize(L);
    for (int t = 0; t < L; ++t) {
      const double a = -2.0 * M_PI * (double)t / (double)L;
      tw[t] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    RHIP(hipMemcpy(pass == 0 ? w->twR.p : w->twC.p, tw.data(), (size_t)L * 8,
                   hipMemcpyHostToDevice));
  }
  *out = w.get();
  r_ws = std::move(w);
  return SCC_OK;
}

// Forward 2-D real transform of nmaps real maps into spectra F.
int fwd2(RWs* w, int nmaps, const float* src, float2* F, hipStream_t st) {
  const Geom& g = w->g;
  const float2* twR = (const float2*)w->twR.p;
  const float2* twC = (const float2*)w->twC.p;
  if (w->pow2) {
    RHIP(pow2_row_fwd(g, w->p2, nmaps, kFwdReal, src, nullptr, nullptr, F,
                      twC, nullptr, st));
    RHIP(pow2_col(g, w->p2, nmaps, F, twR, 0, nullptr, st));
  } else {
    RHIP(launch_row_fwd(g, nmaps, w->rp, kFwdReal, src, nullptr, nullptr, F,
                        twC, nullptr, st));
    RHIP(launch_col(g, nmaps, w->cp, F, twR, 0, nullptr, st));
  }
  return SCC_OK;
}

// Inverse 2-D real transform of nmaps spectra (destroyed) into out / N.
int inv2(RWs* w, int nmaps, float2* F, float* out, hipStream_t st) {
  const Geom& g = w->g;
  const float2* twR = (const float2*)w->twR.p;
  const float2* twC = (const float2*)w->twC.p;
  const float s = 1.f / (float)g.N;
  if (w->pow2) {
    RHIP(pow2_col(g, w->p2, nmaps, F, twR, 1, nullptr, st));
    RHIP(pow2_row_inv(g, w->p2, nmaps, F, out, twC, s, st));
  } else {
    RHIP(launch_col(g, nmaps, w->cp, F, twR, 1, nullptr, st));
    RHIP(launch_row_inv(g, nmaps, w->rp, F, out, twC, s, st));
  }
  return SCC_OK;
}

int check(const float* D, int M, int Kh, int Kw, const float* X, int B, int R,
          int C) {
  if (D == nullptr || X == nullptr)
    return set_error(SCC_ERR_INVALID, "NULL array argument");
  if (M < 1 || Kh < 1 || Kw < 1 || B < 1 || R < 1 || C < 1 || Kh > R ||
      Kw > C)
    return set_error(SCC_ERR_INVALID, "bad sizes");
  if (R > 4095 || (C % 2 == 0 ? C > 7678 : C > 3839) ||
      (long long)B * M * R * C > (1LL << 36))
    return set_error(SCC_ERR_UNSUPPORTED, "image too large");
  return SCC_OK;
}

// Spectra of the padded filters and of X.
int spectra(RWs* w, const float* D, int Kh, int Kw, const float* X,
            hipStream_t st) {
  RHIP(launch_pad(D, (float*)w->tmp.p, w->M, Kh, Kw, w->R, w->C, st));
  int rc;
  if ((rc = fwd2(w, w->M, (const float*)w->tmp.p, (float2*)w->Df.p, st)))
    return rc;
  return fwd2(w, w->B * w->M, X, (float2*)w->Xf.p, st);
}

}  // namespace
}  // namespace scc

using namespace scc;

extern "C" {

int scc_reconstruct(const float* D, int M, int Kh, int Kw, const float* X,
                    int B, int R, int C, float* out, void* stream) {
  int rc = check(D, M, Kh, Kw, X, B, R, C);
  if (rc) return rc;
  if (out == nullptr) return set_error(SCC_ERR_INVALID, "NULL output");
  std::lock_guard<std::mutex> lk(r_mu);
  RWs* w;
  if ((rc = get_rws(R, C, M, B, &w))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = spectra(w, D, Kh, Kw, X, st))) return rc;
  const Geom& g = w->g;
  hipLaunchKernelGGL(k_rc_mul, dim3((unsigned)((g.Hc + 255) / 256), B),
                     dim3(256), 0, st, g.Hc, M, (const float2*)w->Df.p,
                     (const float2*)w->Xf.p, (float2*)w->Gf.p);
  RHIP(hipGetLastError());
  if ((rc = inv2(w, B, (float2*)w->Gf.p, out, st))) return rc;
  RHIP(hipStreamSynchronize(st));
  return SCC_OK;
}

int scc_reconstruct_backward(const float* D, int M, int Kh, int Kw,
                             const float* X, int B, int R, int C,
                             const float* gout, float* gD, float* gX,
                             void* stream) {
  int rc = check(D, M, Kh, Kw, X, B, R, C);
  if (rc) return rc;
  if (gout == nullptr) return set_error(SCC_ERR_INVALID, "NULL gradient");
  std::lock_guard<std::mutex> lk(r_mu);
  RWs* w;
  if ((rc = get_rws(R, C, M, B, &w))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = spectra(w, D, Kh, Kw, X, st))) return rc;
  if ((rc = fwd2(w, B, gout, (float2*)w->Gf.p, st))) return rc;
  const Geom& g = w->g;
  hipLaunchKernelGGL(k_rc_adj, dim3((unsigned)((g.Hc + 255) / 256), M),
                     dim3(256), 0, st, g.Hc, M, B, (const float2*)w->Df.p,
                     (float2*)w->Xf.p, (const float2*)w->Gf.p,
                     gD ? (float2*)w->gDf.p : nullptr, gX ? 1 : 0);
  RHIP(hipGetLastError());
  if (gX != nullptr && (rc = inv2(w, B * M, (float2*)w->Xf.p, gX, st)))
    return rc;
  if (gD != nullptr) {
    if ((rc = inv2(w, M, (float2*)w->gDf.p, (float*)w->tmp.p, st))) return rc;
    const int nk = M * Kh * Kw;
    hipLaunchKernelGGL(k_rc_crop, dim3((nk + 255) / 256), dim3(256), 0, st,
                       (const float*)w->tmp.p, gD, M, Kh, Kw, R, C);
    RHIP(hipGetLastError());
  }
  RHIP(hipStreamSynchronize(st));
  return SCC_OK;
}

}  // extern "C"
