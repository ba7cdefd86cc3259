// Fused row kernel of the generic sizes with the row length a compile-time
// constant, instantiated for every row length Lr = 4 q, kCtQMin <= q <=
// kCtQMax (image widths that are multiples of 8 from 32 to 1632, rows that
// fit the wavefront kernel's LDS), in kCtParts translation units
// (rows_ct_p*.hip).
//
// Same work, same arithmetic and the same order of operations as the
// wavefront-per-row kernel with a run-time plan (generic_common.h
// row_body_wave, plain update path): the stages are the same gen::stage_*
// functions on the plan gen::plan_radices makes, but with the length, the
// radices, the butterfly counts and the stage strides known to the
// compiler, so the per-item quotients, the stage dispatch and the address
// arithmetic fold into constants and the work-item loops have fixed trip
// counts.
#pragma once

#include "generic_common.h"

namespace scc {

constexpr int kCtQMin = 4, kCtQMax = 204, kCtParts = 8;

namespace {

struct CtPlan {
  int n;
  int r[kMaxStages];
};

template <int L>
struct RowPlan {
  static constexpr CtPlan make() {
    CtPlan a{};
    a.n = gen::plan_radices(L, a.r, kMaxStages, false);
    return a;
  }
  static constexpr CtPlan p = make();
};

// Output pairs per work item of a direct-DFT stage of prime P with nb
// butterflies on one wavefront: the item rounds times the per-input cost
// (two loads and the pair sums, then 8 instructions per output pair),
// smallest for at most 4 pairs (more cost registers the kernel does not
// have).  E.g. 4 x 89, 4 x 67 and 12 x 31 take 3 pairs: one round of items
// whose inner loop is a quarter shorter than with 4.
constexpr int ct_prime_pairs(int nb, int P) {
  const int h = (P - 1) / 2;
  int best = 4;
  long long best_cost = -1;
  for (int q = 1; q <= 4; ++q) {
    const int items = nb * ((h + q - 1) / q);
    const long long cost = (long long)((items + 63) / 64) * h * (8 + 8 * q);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = q;
    }
  }
  return best;
}

// Forward (DIR = 1) or inverse transform of one row of length L held in
// LDS at A (ping-pong with B) by one wavefront; returns the buffer with the
// result.  tw: the row twiddle table W_{2L}^t (stride 2 for W_L).
template <int DIR, int L, int S = 0, int NS = 1>
__device__ __forceinline__ float2* fft_ct(float2* A, float2* B,
                                          const float2* tw, int lane) {
  constexpr CtPlan P = RowPlan<L>::p;
  if constexpr (S >= P.n) {
    return A;
  } else {
    constexpr int RX = P.r[S];
    const gen::GStage st = gen::make_stage(L, RX, NS);
    if constexpr (RX >= 17) {  // direct DFT stage (gen::prime_stage)
      if constexpr (NS > 1) {
        gen::stage_prime_twiddle<DIR>(A, 0, 1, st, tw, 2, lane, 64);
        reg::wave_sync();
      }
      gen::stage_prime<DIR, ct_prime_pairs(L / RX, RX)>(
          A, B, gen::gio_lds(0), 1, st, tw, 2, lane, 64);
    } else {
      gen::stage_reg<RX, DIR>(A, B, gen::gio_lds(0), 1, st, tw, 2, lane, 64);
    }
    reg::wave_sync();
    return fft_ct<DIR, L, S + 1, NS * RX>(B, A, tw, lane);
  }
}

// Fused row kernel, even width C = 2 LR, V state, one threshold per row (no
// mask, per-map or scalar L1 weight, no NonNegCoef / NoBndryCross): the
// plain path of row_body_wave.
template <int LR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void
k_row_ct(Geom g, Buffers b, Params p, int nr) {
  if (b.ctrl->done) return;
  extern __shared__ __align__(16) float2 lds[];
  __shared__ float red[5][4];
  constexpr int L = LR, C = 2 * LR, H = LR + 1, ld = LR + 1;
  const float2* twL = row_wave_table(g, lds, b.twC);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float2* A = lds + wv * 2 * ld;
  float2* B = A + ld;
  const long long nrows = (long long)g.M * g.R;
  const float rho = b.ctrl->rho;
  const float invN = 1.0f / (float)g.N;
  const float lr = p.lambda / rho;
  const int tsh = __ffs(g.TC) - 1;
  const bool alpha1 = p.alpha == 1.f;
  const long long blk = blockIdx.x;
  // Half-spectrum addressing in 32 bits (panel_off_sh's layout): column
  // c of row r sits at c + (c >> tsh) K + r TC, K = TC (R - 1), in the
  // full panels (c < c0 = (P - 1) TC) and at (P - 1) TC R + r wlast + c - c0
  // in the last one.  Lane 0 takes the Nyquist column L beside column 0,
  // so the loops run L / 64 rounds.
  const unsigned K = (unsigned)(g.TC * (g.R - 1));
  const int c0 = (g.P - 1) * g.TC;
  const unsigned lastp = (unsigned)(c0 * g.R);
  UpdateSums acc;
  for (int s = 0; s < nr / kRowWaves; ++s) {
    const long long row = wave_row(g, blk, wv, s, nr, nrows);
    if (row >= nrows) break;
    const int m = (int)(row / g.R);
    const int r = (int)(row - (long long)m * g.R);
    float2* Fm = b.F + (long long)m * g.Hc;
    // Half spectrum row -> B, inverse real-FFT pre-processing -> A
    // (wave_c2r_load).
    const unsigned rT = (unsigned)(r * g.TC);
    const unsigned rL = lastp + (unsigned)(r * g.wlast) - (unsigned)c0;
    auto at = [&](int c) {
      return c < c0 ? (unsigned)c + (unsigned)(c >> tsh) * K + rT
                    : rL + (unsigned)c;
    };
#pragma unroll
    for (int c = lane; c < L; c += 64) B[c] = Fm[at(c)];
    if (lane == 0) B[L] = Fm[at(L)];
    // The row's state, loaded with the spectrum (its latency then overlaps
    // the pre-processing and the inverse transform: rows 2-7 % faster)
    // where at most 6 update rounds keep it in registers without spills;
    // longer rows load it in the update loop.
    const long long gi0 = row * C;
    float2* V2 = reinterpret_cast<float2*>(b.Y + gi0);
    constexpr int NV = (L + 63) / 64;
    constexpr bool VPRE = NV <= 6;
    float2 vpre[VPRE ? NV : 1];
    if constexpr (VPRE) {
#pragma unroll
      for (int i = 0; i < NV; ++i)
        if (lane + 64 * i < L) vpre[i] = V2[lane + 64 * i];
    }
    reg::wave_sync();
#pragma unroll 1
    for (int k = lane; k < L; k += 64) {
      float2 xk = B[k];
      float2 xn = B[L - k];
      if (k == 0) {
        xk.y = 0.f;
        xn.y = 0.f;
      }
      xn.y = -xn.y;
      const float2 sm = cadd(xk, xn);
      const float2 wd = cmul(cconj(twL[k]), csub(xk, xn));
      A[k] = make_float2(sm.x - wd.y, sm.y + wd.x);
    }
    reg::wave_sync();
    float2* P = fft_ct<-1, L>(A, B, twL, lane);
    float2* Q = (P == A) ? B : A;
    const float th = lr * (p.l1_mode == 0 ? p.l1_scalar : b.l1w[m]);
    auto update = [&](int k, float2 vo) {
      const float2 xz = P[k];
      const float x0 = xz.x * invN, x1 = xz.y * invN;
      const float uo0 = split_u(vo.x, th, -th), uo1 = split_u(vo.y, th, -th);
      const float yo0 = vo.x - uo0, yo1 = vo.y - uo1;
      const float xr0 = alpha1 ? x0 : p.alpha * x0 + (1.f - p.alpha) * yo0;
      const float xr1 = alpha1 ? x1 : p.alpha * x1 + (1.f - p.alpha) * yo1;
      const float v0 = xr0 + uo0, v1 = xr1 + uo1;
      const float un0 = split_u(v0, th, -th), un1 = split_u(v1, th, -th);
      const float yn0 = v0 - un0, yn1 = v1 - un1;
      acc.x += x0 * x0;
      acc.y += yn0 * yn0;
      acc.u += un0 * un0;
      acc.r += (x0 - yn0) * (x0 - yn0);
      acc.s += (yo0 - yn0) * (yo0 - yn0);
      acc.x += x1 * x1;
      acc.y += yn1 * yn1;
      acc.u += un1 * un1;
      acc.r += (x1 - yn1) * (x1 - yn1);
      acc.s += (yo1 - yn1) * (yo1 - yn1);
      V2[k] = make_float2(v0, v1);
      P[k] = make_float2(yn0 - un0, yn1 - un1);
    };
    if constexpr (VPRE) {
#pragma unroll
      for (int i = 0; i < NV; ++i)
        if (lane + 64 * i < L) update(lane + 64 * i, vpre[i]);
    } else {
#pragma unroll 1
      for (int k = lane; k < L; k += 64) update(k, V2[k]);
    }
    reg::wave_sync();
    const float2* Z = fft_ct<1, L>(P, Q, twL, lane);
    // Forward real-FFT post-processing and store (wave_r2c_store).
    auto post = [&](int k) {
      const float2 zk = Z[k == L ? 0 : k];
      const float2 zn = cconj(Z[k == 0 ? 0 : L - k]);
      const float2 sm = cadd(zk, zn);
      const float2 wd = cmul(twL[k], csub(zk, zn));
      return make_float2(0.5f * (sm.x + wd.y), 0.5f * (sm.y - wd.x));
    };
#pragma unroll 1
    for (int k = lane; k < L; k += 64) Fm[at(k)] = post(k);
    if (lane == 0) Fm[at(L)] = post(L);
    reg::wave_sync();  // this row's LDS reads before the next row's writes
  }
  block_sums_store(acc, red, b.partials, gridDim.x, (int)blk, threadIdx.x);
}

// The kernel of row length lr, or null, among the lengths 4 q with
// q = kCtQMin + PART (mod kCtParts), q <= kCtQMax: one translation unit per
// PART, so that the instantiations compile in parallel.
template <int PART, int Q = kCtQMin + PART>
const void* ct_pick(int lr) {
  if constexpr (Q > kCtQMax) {
    return nullptr;
  } else {
    if (lr == 4 * Q) return reinterpret_cast<const void*>(&k_row_ct<4 * Q>);
    return ct_pick<PART, Q + kCtParts>(lr);
  }
}

}  // namespace
}  // namespace scc
