// Helpers shared by the power-of-two row kernels of the solver
// (kernels_pow2.hip) and of its adjoint (unrolled.hip).
#pragma once

#include <type_traits>

#include "fft_reg.h"
#include "scc_device.h"

namespace scc {

// Every row belongs to one wavefront (RPW rows per wavefront, T lanes per
// row, E points per lane).  The half spectrum is read and written in slot
// order (lane t holds columns t + T e); the real-FFT pre/post-processing
// pairs column k with column L - k, which lives in lane (T - t) mod T, slot
// E - 1 - e (lane 0: its own slot E - e, or column L), fetched with a lane
// shuffle.
template <int LR>
struct RowCfg {
  static constexpr int E = LR <= 512 ? 8 : 16;
  static constexpr int T = LR / E;       // lanes per row
  static constexpr int RPW = 64 / T;     // rows per wavefront
  static constexpr int RPB = 4 * RPW;    // rows per 256-thread block
  // Y and U are loaded with the spectrum (latency hidden behind the inverse
  // transform) where registers allow it.
  static constexpr bool PREFETCH = E <= 8;
  // Occupancy target for the register allocator (4 waves per SIMD fits the
  // E = 8 plain update without spills).
  static constexpr int WAVES = E <= 8 ? 4 : 2;
};

// Soft-threshold support tape of the unrolled solver: 4 bits per lane slot
// e (one float pair v0, v1): bit 0 |v0| > t, bit 1 v0 < 0 (meaningful where
// bit 0 is set), bit 2 |v1| > t, bit 3 v1 < 0.  Lane t of row `row` keeps TapeWords<E>::W words at
// bits[(row * T + t) * W + w]; slot e is in word e / 8 at bit 4 (e % 8).
template <int E>
struct TapeWords {
  static constexpr int W = (4 * E + 31) / 32;
};

// 32-bit words of the support tape per row of C columns (host side).
inline int tape_words_per_row(int C) {
  const int LR = C / 2;
  const int E = LR <= 512 ? 8 : 16;
  return (LR / E) * ((4 * E + 31) / 32);
}

__device__ __forceinline__ float2 shfl2(float2 v, int src) {
  return make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// Offset of half-spectrum column k of row r in a map (panel layout with
// full panels of TC columns and a last panel holding column LR).
template <int TC>
__device__ __forceinline__ int spec_off(int k, int r, int R) {
  return (k / TC) * (TC * R) + r * TC + (k % TC);
}

// Slot e of lane t (column t + T e, TC dividing T) of row r of a map's half
// spectrum.  Column t + T e sits (T / TC) e panels after column t, so with
// E <= 8 slots the address is the lane's 32-bit byte offset (one register
// for all slots; a map's spectrum is far below 4 GiB) from a uniform base
// per slot, with no per-element arithmetic.  With 16 slots the element's
// own index measured faster (2 % on 1024-point rows) and is kept.
template <int TC, int T, int E, typename F2>
__device__ __forceinline__ F2& spec_slot(F2* Fm, int t, int e, int r, int R) {
  static_assert(T % TC == 0, "panel width must divide the lanes per row");
  if constexpr (E <= 8) {
    using B = typename std::conditional<std::is_const<F2>::value,
                                        const char, char>::type;
    const unsigned bo = (unsigned)spec_off<TC>(t, r, R) * 8u;
    return *reinterpret_cast<F2*>(
        reinterpret_cast<B*>(Fm + (long long)e * T * R) + bo);
  } else {
    return Fm[spec_off<TC>(t + T * e, r, R)];
  }
}

// X[k] from the forward transform Z of the packed real row:
// X[k] = (Z[k] + conj Z[L-k]) / 2 - i W_C^k (Z[k] - conj Z[L-k]) / 2.
__device__ __forceinline__ float2 r2c_post(float2 zk, float2 znc, float2 w) {
  const float2 s = cadd(zk, znc);
  const float2 wd = cmul(w, csub(zk, znc));
  return make_float2(0.5f * (s.x + wd.y), 0.5f * (s.y - wd.x));
}

template <int LR, int TC>
__device__ __forceinline__ void store_spectrum(const float2 (&z)[RowCfg<LR>::E],
                                               int t, int partner, int r,
                                               int R, float2* Fm,
                                               const float2* twC) {
  using RC = RowCfg<LR>;
  constexpr int E = RC::E, T = RC::T;
  const float2 wt = twC[t];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float2 sh = shfl2(z[E - 1 - e], partner);
    const float2 zn = (t == 0) ? (e == 0 ? z[0] : z[(E - e) % E]) : sh;
    spec_slot<TC, T, E>(Fm, t, e, r, R) =
        r2c_post(z[e], cconj(zn), reg::w_slot<E>(wt, e));
  }
  if (t == 0)  // column L: Re Z[0] - Im Z[0]
    Fm[(LR / TC) * TC * R + r] = make_float2(z[0].x - z[0].y, 0.f);
}

// GENERAL selects the update with an AMS mask, a full weight array,
// NonNegCoef or NoBndryCross; without it the threshold is one value per row.

// Load the half spectrum of row r (slot order, column L on lane 0) and form
// the input of the length-LR inverse transform of the packed real row:
// Z'[k] = (X[k] + conj X[L-k]) + i W_C^-k (X[k] - conj X[L-k]).
template <int LR, int TC>
__device__ __forceinline__ void load_spectrum_c2r(float2 (&z)[RowCfg<LR>::E],
                                                  int t, int partner, int r,
                                                  int R, const float2* Fm,
                                                  const float2* twC) {
  using RC = RowCfg<LR>;
  constexpr int E = RC::E, T = RC::T;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = spec_slot<TC, T, E>(Fm, t, e, r, R);
  float2 xL = make_float2(0.f, 0.f);
  if (t == 0) xL = Fm[(LR / TC) * TC * R + r];
  const float2 wt = twC[t];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float2 sh = shfl2(v[E - 1 - e], partner);
    float2 xn = (t == 0) ? (e == 0 ? xL : v[(E - e) % E]) : sh;
    float2 xk = v[e];
    if (t == 0 && e == 0) {
      xk.y = 0.f;
      xn.y = 0.f;
    }
    xn.y = -xn.y;
    const float2 s = cadd(xk, xn);
    const float2 wd = cmul(cconj(reg::w_slot<E>(wt, e)), csub(xk, xn));
    z[e] = make_float2(s.x - wd.y, s.y + wd.x);
  }
}

}  // namespace scc
