// Device helpers shared by the generic and power-of-two kernels: complex
// arithmetic, the fused per-coefficient update, block reductions of the
// residual partial sums and the device-side stopping / AutoRho logic.
#pragma once

#include "scc_internal.h"

namespace scc {

// The split of V by a soft threshold t >= 0 into U = V - shrink(V, t) and
// Y = shrink(V, t) = copysign(max(|V| - t, 0), V): U is V clamped to
// [-t, t] (one v_med3_f32) and Y = V - U.  Where |V| > t both forms take
// the same subtraction |V| - t, so they agree bit for bit, except that Y is
// +0 rather than copysign(0, V) where |V| <= t (no sum, product, stored
// state or support test tells the two zeros apart).  Two instructions
// instead of four per value.  Only for t >= 0 (the one-array state).
__device__ __forceinline__ float split_u(float v, float t, float mt) {
  return __builtin_amdgcn_fmed3f(v, mt, t);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) {
  return make_float2(a.x, -a.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}
// Non-temporal load of a value that is read once per pass and not again
// before much other traffic (the filter spectra in the X-solve of large
// problems): it does not displace data that the next kernel reads back from
// the last-level cache.
__device__ __forceinline__ float2 load_once(const float2* p) {
  return __builtin_bit_cast(
      float2, __builtin_nontemporal_load(
                  reinterpret_cast<const unsigned long long*>(p)));
}
// -i * dir * a
__device__ __forceinline__ float2 mul_mi(float2 a, float dir) {
  return make_float2(a.y * dir, -a.x * dir);
}
// Table entry exp(-2 pi i e / L), conjugated for the inverse (dir = -1).
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw,
                                          int e, float dir) {
  float2 w = tw[e];
  w.y *= dir;
  return w;
}

// Column index of flat half-spectrum offset f (< Hc) in the panel layout.
__device__ __forceinline__ int panel_col(const Geom& g, long long f) {
  const int pr = g.TC * g.R;
  const int p = (int)(f / pr);
  const int o = (int)(f - (long long)p * pr);
  const int w = (p == g.P - 1) ? g.wlast : g.TC;
  return p * g.TC + o % w;
}

// Multiplicity of a half-spectrum column in the full spectrum: 1 for DC
// (and Nyquist when C is even), 2 for the others.
__device__ __forceinline__ float col_mult(const Geom& g, int c) {
  return (c == 0 || (g.even && c == g.C / 2)) ? 1.f : 2.f;
}

struct UpdateSums {
  float x = 0.f, y = 0.f, u = 0.f, r = 0.f, s = 0.f;
};

// One coefficient of the fused Y/U update for X = x (already divided by N):
//   Xr = alpha X + (1 - alpha) Yold                (cbpdn_kernels.cu:657-689)
//   V  = Xr + U
//   Y  = copysign(max(|V| - (lambda/rho) w, 0), V)  (cbpdn_kernels.cu:129-311)
//   AMS map 0: Y = (1 - W) V                       (cbpdn_kernels.cu:1137-1155)
//   U  = V - Y
// plus the five residual sums of cbpdn_kernels.cu:695-783 (X unrelaxed,
// Yprv = Yold, so no copy of the previous Y is kept).  Returns Y - U, the
// input of the next iteration's forward transform.
// Y of coefficient (m, r, n) from V = Xr + U: the soft threshold
// (lambda/rho) w with the NonNegCoef / NoBndryCross rules, or (1 - W) V on
// the AMS impulse map.
__device__ __forceinline__ float shrink_coef(const Geom& g, const Params& p,
                                             const Buffers& b, float v,
                                             float lr, int m, int r, int n,
                                             long long gi) {
  float yn;
  if (m < p.m0) {
    yn = (float)(1 - b.mask[(long long)r * g.C + n]) * v;
  } else {
    const float w = p.l1_mode == 0 ? p.l1_scalar
                                   : (p.l1_mode == 1 ? b.l1w[m] : b.l1w[gi]);
    const float a = fabsf(v) - lr * w;
    yn = copysignf(fmaxf(a, 0.f), v);
    if (p.nonneg) yn = fmaxf(yn, 0.f);
    if (p.nobndry && (r > g.R - p.Kh || n > g.C - p.Kw)) yn = 0.f;
  }
  return yn;
}

// store_v (state V, fused row launch): the new V is stored instead of Y and
// U.  With p.vstate the old Y and U are decoded from the stored V.
__device__ __forceinline__ float update_point(const Geom& g, const Params& p,
                                              const Buffers& b, float x,
                                              float lr, int m, int r, int n,
                                              long long gi, UpdateSums& acc,
                                              bool store_v = false) {
  float yo, uo;
  if (p.vstate) {
    const float vo = b.Y[gi];
    yo = shrink_coef(g, p, b, vo, lr, m, r, n, gi);
    uo = vo - yo;
  } else {
    yo = b.Y[gi];
    uo = b.U[gi];
  }
  const float xr = (p.alpha == 1.f) ? x : p.alpha * x + (1.f - p.alpha) * yo;
  const float v = xr + uo;
  const float yn = shrink_coef(g, p, b, v, lr, m, r, n, gi);
  const float un = v - yn;
  if (store_v) {
    b.Y[gi] = v;
  } else {
    b.Y[gi] = yn;
    b.U[gi] = un;
  }
  acc.x += x * x;
  acc.y += yn * yn;
  acc.u += un * un;
  const float dxy = x - yn;
  acc.r += dxy * dxy;
  const float dy = yo - yn;
  acc.s += dy * dy;
  return yn - un;
}

__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// wave_sum of five values at once (the same additions in the same order
// for each, so the results equal five wave_sum calls): the five lane
// shuffles of a step are issued together and wait once.
__device__ __forceinline__ void wave_sum5(float (&v)[5]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    float s[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) s[j] = __shfl_down(v[j], o, 64);
#pragma unroll
    for (int j = 0; j < 5; ++j) v[j] += s[j];
  }
}

// Reduce the five sums over the block (fixed order, deterministic) and
// store them to partials[j * nblocks + bid].  red must hold 5 x nwaves.
template <int NW>
__device__ __forceinline__ void block_sums_store(const UpdateSums& acc,
                                                 float (*red)[NW],
                                                 float* partials, int nblocks,
                                                 int bid, int tid) {
  const int lane = tid & 63, wv = tid >> 6;
  float v[5] = {acc.x, acc.y, acc.u, acc.r, acc.s};
  wave_sum5(v);
  if (lane == 0)
    for (int j = 0; j < 5; ++j) red[j][wv] = v[j];
  __syncthreads();
  if (tid < 5) {
    float s = 0.f;
    for (int w = 0; w < NW; ++w) s += red[tid][w];
    partials[tid * nblocks + bid] = s;
  }
}

// First pass of a deferred residual reduction, run by the blocks of the
// next column kernel: block b of nb adds its slice of the per-wavefront
// partial sums (5 arrays of nparts floats) in double and stores the five
// results to bsum[5 b + j].  Fixed order, so the sums are deterministic.
__device__ __forceinline__ void partials_slice(const float* __restrict__ parts,
                                               int nparts, double* bsum,
                                               int b, int nb,
                                               double (*red)[16]) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int chunk = (nparts + nb - 1) / nb;
  const int i0 = b * chunk;
  const int i1 = min(nparts, i0 + chunk);
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = i0 + tid; i < i1; i += nthr)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[j] += (double)parts[j * nparts + i];
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    double v = acc[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[j][wv] = v;
  }
  __syncthreads();
  if (tid < 5) {
    double v = 0.0;
    for (int w = 0; w < (nthr >> 6); ++w) v += red[tid][w];
    bsum[5 * b + tid] = v;
  }
}

// Residuals, stopping test and AutoRho from the five global sums
// (reference host code cbpdn.cu:405-436 and 521-559, with sporco's
// sqrt(Nc) absolute tolerance term and guarded normalisers).  Runs on one
// thread.
__device__ inline void control_update(const Buffers& b, const Params& p,
                                      const double* s5, int iter) {
  Ctrl* c = b.ctrl;
  const double rho = c->rho;
  const double nX = sqrt(s5[0]), nY = sqrt(s5[1]), nU = sqrt(s5[2]);
  double r = sqrt(s5[3]);
  double s = rho * sqrt(s5[4]);
  double rn = fmax(nX, nY), sn = rho * nU, epri, edua;
  if (p.std_residuals) {
    epri = p.sqrt_nc * p.abs_tol + rn * p.rel_tol;
    edua = p.sqrt_nc * p.abs_tol + sn * p.rel_tol;
  } else {
    if (rn == 0.0) rn = 1.0;
    if (sn == 0.0) sn = 1.0;
    r /= rn;
    s /= sn;
    epri = p.sqrt_nc * p.abs_tol / rn + p.rel_tol;
    edua = p.sqrt_nc * p.abs_tol / sn + p.rel_tol;
  }
  c->iter = iter;
  c->r = (float)r;
  c->s = (float)s;
  c->epri = (float)epri;
  c->edua = (float)edua;
  c->nX = (float)nX;
  c->nY = (float)nY;
  c->nU = (float)nU;
  if (p.track && b.hist != nullptr) {
    b.hist[3 * (iter - 1) + 0] = (float)r;
    b.hist[3 * (iter - 1) + 1] = (float)s;
    b.hist[3 * (iter - 1) + 2] = (float)rho;
  }
  float rsf_out = 1.f;
  if (p.autorho && iter > 1 && iter % p.period == 0) {
    double rhomlt;
    if (p.autoscaling) {
      if (s == 0.0 || r == 0.0) {
        rhomlt = p.tau;
      } else {
        const double q = r / (s * p.xi);
        rhomlt = sqrt(q >= 1.0 ? q : 1.0 / q);
        if (rhomlt > p.tau) rhomlt = p.tau;
      }
    } else {
      rhomlt = p.tau;
    }
    double rsf = 1.0;
    if (r > p.xi * p.rsdl_ratio * s)
      rsf = rhomlt;
    else if (s > (p.rsdl_ratio / p.xi) * r)
      rsf = 1.0 / rhomlt;
    rsf_out = (float)rsf;
    c->rho = c->rho * rsf_out;
  }
  c->rsf = rsf_out;
  if (r < epri && s < edua) c->done = 1;
}

}  // namespace scc
