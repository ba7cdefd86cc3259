// Objective of the CBPDN problems, evaluated on request after an iteration
// (Verbose, or objective tracking):
//   DFid   = 1/2 || sum_m d_m * z_m - s ||^2        (plain, GradReg)
//            1/2 || W (sum_{m>=1} d_m * z_m - s) ||^2   (mask: map 0 is
//            the impulse map of the additive mask simulation)
//   Regl1  = || w (.) z ||_1                         (maps m >= m0)
//   Regl2g = 1/2 sum_m wg_m || grad z_m ||^2           (GradReg, circular
//            first differences: |G|^2 = 4 - 2 cos - 2 cos per frequency)
//   Fnc    = DFid + lambda Regl1 (+ mu Regl2g)
// with z = X (AuxVarObj off) or Y (AuxVarObj on), sporco's definitions
// (the reference's own objective kernels, cbpdn_kernels.cu:913-1135 with
// the host code cbpdn.cu:438-517, normalise the fidelity term wrongly and
// read a stale spectrum; these are not followed).
//
// The fidelity and gradient terms are half-spectrum sums with Parseval's
// weights (1/N, columns other than DC / Nyquist counted twice); the masked
// fidelity needs the spatial residual, so its spectrum is stored and
// transformed back by the caller.  Partial sums are double, per block, and
// added in a fixed order by k_obj_final.

#include "scc_device.h"
#include "scc_internal.h"

namespace scc {

namespace {

template <int NW>
__device__ __forceinline__ double block_sum_d(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NW; ++w) s += red[w];
  __syncthreads();
  return s;
}

// The objective of iteration `iter` is skipped once the solve stopped at
// an earlier iteration.
__device__ __forceinline__ bool obj_skip(const Ctrl* c, int iter) {
  return c->done && c->iter < iter;
}

}  // namespace

__global__ __launch_bounds__(256) void k_obj_freq(Geom g, Buffers b, Params p,
                                                  const float2* __restrict__ G,
                                                  float2* E, double* part,
                                                  int iter) {
  if (obj_skip(b.ctrl, iter)) return;
  __shared__ double red[4];
  const long long f = blockIdx.x * 256LL + threadIdx.x;
  double dfd = 0.0, grd = 0.0;
  if (f < g.Hc) {
    const float w = col_mult(g, panel_col(g, f));
    float2 e = make_float2(-b.Sf[f].x, -b.Sf[f].y);
    float gs = 0.f;
    for (int m = p.m0; m < g.M; ++m) {
      const long long o = (long long)m * g.Hc + f;
      const float2 x = G[o];
      e = cadd(e, cmul(b.Df[o], x));
      if (p.grad)
        gs += (p.gw_mode ? b.gw[m] : p.gw_scalar) * (x.x * x.x + x.y * x.y);
    }
    if (E != nullptr)
      E[f] = e;
    else
      dfd = (double)w * ((double)e.x * e.x + (double)e.y * e.y);
    if (p.grad) grd = (double)w * b.ghg[f] * gs;
  }
  const double sd = block_sum_d<4>(dfd, red);
  const double sg = block_sum_d<4>(grd, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd;
    part[2 * blockIdx.x + 1] = sg;
  }
}

// sum_{m >= m0} sum_n |w z|, grid-stride, one partial per block.
__global__ __launch_bounds__(256) void k_obj_l1(Geom g, Params p,
                                                const float* __restrict__ Z,
                                                const float* l1w,
                                                const Ctrl* ctrl,
                                                double* part, int iter) {
  if (obj_skip(ctrl, iter)) return;
  __shared__ double red[4];
  const long long n0 = (long long)p.m0 * g.N;
  const long long n1 = (long long)g.M * g.N;
  double acc = 0.0;
  for (long long i = n0 + blockIdx.x * 256LL + threadIdx.x; i < n1;
       i += (long long)gridDim.x * 256) {
    float w = p.l1_scalar;
    if (p.l1_mode == 1)
      w = l1w[i / g.N];
    else if (p.l1_mode == 2)
      w = l1w[i];
    acc += fabsf(w * Z[i]);
  }
  const double s = block_sum_d<4>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// sum_n (W_n e_n)^2 of the spatial residual e, one partial per block.
__global__ __launch_bounds__(256) void k_obj_mask(Geom g,
                                                  const float* __restrict__ e,
                                                  const int* __restrict__ mask,
                                                  const Ctrl* ctrl,
                                                  double* part, int iter) {
  if (obj_skip(ctrl, iter)) return;
  __shared__ double red[4];
  double acc = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < g.N;
       i += (long long)gridDim.x * 256) {
    const double v = (double)mask[i] * e[i];
    acc += v * v;
  }
  const double s = block_sum_d<4>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Totals (one wavefront, fixed order): out = (Fnc, DFid, Regl1, Regl2g),
// also stored as row iter - 1 of hist when hist is given.
__global__ __launch_bounds__(64) void k_obj_final(
    const double* pf, int nf, const double* pl, int nl, const double* pm,
    int nm, Params p, double invN, const Ctrl* ctrl, float* out, float* hist,
    int iter) {
  if (obj_skip(ctrl, iter)) return;
  double d = 0.0, gr = 0.0, l = 0.0, mk = 0.0;
  for (int i = threadIdx.x; i < nf; i += 64) {
    d += pf[2 * i];
    gr += pf[2 * i + 1];
  }
  for (int i = threadIdx.x; i < nl; i += 64) l += pl[i];
  for (int i = threadIdx.x; i < nm; i += 64) mk += pm[i];
  for (int o = 32; o > 0; o >>= 1) {
    d += __shfl_down(d, o, 64);
    gr += __shfl_down(gr, o, 64);
    l += __shfl_down(l, o, 64);
    mk += __shfl_down(mk, o, 64);
  }
  if (threadIdx.x == 0) {
    const double dfd = nm > 0 ? 0.5 * mk : 0.5 * d * invN;
    const double rg = 0.5 * gr * invN;
    const double fnc = dfd + (double)p.lambda * l +
                       (p.grad ? (double)p.mu * rg : 0.0);
    const float v[4] = {(float)fnc, (float)dfd, (float)l, (float)rg};
    for (int k = 0; k < 4; ++k) {
      out[k] = v[k];
      if (hist != nullptr) hist[4 * (iter - 1) + k] = v[k];
    }
  }
}

int obj_blocks(const Geom& g) { return (int)((g.Hc + 255) / 256); }

hipError_t launch_obj_freq(const Geom& g, const Buffers& b, const Params& p,
                           const float2* G, float2* E, double* part, int iter,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_obj_freq, dim3(obj_blocks(g)), dim3(256), 0, st, g, b,
                     p, G, E, part, iter);
  return hipGetLastError();
}

hipError_t launch_obj_l1(const Geom& g, const Params& p, const float* Z,
                         const float* l1w, const Ctrl* ctrl, double* part,
                         int nblk, int iter, hipStream_t st) {
  hipLaunchKernelGGL(k_obj_l1, dim3(nblk), dim3(256), 0, st, g, p, Z, l1w,
                     ctrl, part, iter);
  return hipGetLastError();
}

hipError_t launch_obj_mask(const Geom& g, const float* e, const int* mask,
                           const Ctrl* ctrl, double* part, int nblk, int iter,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_obj_mask, dim3(nblk), dim3(256), 0, st, g, e, mask,
                     ctrl, part, iter);
  return hipGetLastError();
}

hipError_t launch_obj_final(const double* pf, int nf, const double* pl,
                            int nl, const double* pm, int nm, const Params& p,
                            double invN, const Ctrl* ctrl, float* out,
                            float* hist, int iter, hipStream_t st) {
  hipLaunchKernelGGL(k_obj_final, dim3(1), dim3(64), 0, st, pf, nf, pl, nl,
                     pm, nm, p, invN, ctrl, out, hist, iter);
  return hipGetLastError();
}

}  // namespace scc
