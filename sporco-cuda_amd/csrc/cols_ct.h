// Column transforms of the generic sizes with the column length a
// compile-time constant, instantiated for every length R = 8 q,
// kCtColQMin <= q <= kCtColQMax, in kCtParts translation units
// (cols_ct_p*.hip).
//
// Same work and arithmetic as k_col_generic's direct path (generic_common.h
// col_direct): one panel of w columns of one map per block, the first stage
// reading its butterflies from the panel, the middle ones between two LDS
// buffers, the last writing back, on the plan gen::plan_radices makes with
// the prime-factor radices; the length, radices, butterfly counts and
// strides are constants, and each direct-DFT stage takes the number of
// output pairs per work item that needs the fewest item rounds.
#pragma once

#include "generic_common.h"
#include "rows_ct.h"

namespace scc {

constexpr int kCtColQMin = 4, kCtColQMax = 128;

namespace {

template <int R>
struct ColPlan {
  static constexpr CtPlan make() {
    CtPlan a{};
    a.n = gen::plan_radices(R, a.r, kMaxStages, true);
    return a;
  }
  static constexpr CtPlan p = make();
};

// Panel width of the column transforms of length R (generic_common.h
// col_panel_width with the 32 KiB budget, for rows wide enough).
constexpr int ct_col_panel(int R) {
  int cap = 32768 / (16 * (R + 1)), tc = 1;
  while (tc * 2 <= cap) tc *= 2;
  return tc;
}

// Output pairs per work item of a direct-DFT stage of prime P with nb
// butterflies per column, w columns and 256 threads (see ct_prime_pairs).
constexpr int ct_col_prime_pairs(int nb, int P, int w) {
  const int h = (P - 1) / 2;
  int best = 4;
  long long best_cost = -1;
  for (int q = 1; q <= 4; ++q) {
    const int items = w * nb * ((h + q - 1) / q);
    const long long cost = (long long)((items + 255) / 256) * h * (8 + 8 * q);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = q;
    }
  }
  return best;
}

// Stage S of the transforms of the w columns at base (row-major panel):
// src holds the previous stage's output in LDS (dst is free).
template <int DIR, int R, int S = 0, int NS = 1>
__device__ __forceinline__ void col_ct(float2* base, float2* src,
                                       float2* dst, int w, const gen::GIo& gin,
                                       const gen::GIo& gout,
                                       const float2* tw, int tid) {
  constexpr CtPlan P = ColPlan<R>::p;
  constexpr int RX = P.r[S];
  constexpr bool first = S == 0, last = S == P.n - 1;
  constexpr int ld = R + 1;
  const gen::GStage st = gen::make_stage(R, RX, NS);
  if constexpr (!first && RX >= 17 && NS > 1) {
    gen::stage_prime_twiddle<DIR>(src, ld, w, st, tw, 1, tid, 256);
    __syncthreads();
  }
  const float2* in = first ? base : src;
  float2* out = last ? base : (first ? src : dst);
  const gen::GIo io = first ? gin : (last ? gout : gen::gio_lds(ld));
  if constexpr (RX >= 17) {
    constexpr int q = ct_col_prime_pairs(R / RX, RX, ct_col_panel(R));
    gen::stage_prime<DIR, q>(in, out, io, w, st, tw, 1, tid, 256);
  } else {
    gen::stage_reg<RX, DIR>(in, out, io, w, st, tw, 1, tid, 256);
  }
  if constexpr (!last) {
    __syncthreads();
    if constexpr (first)
      col_ct<DIR, R, S + 1, NS * RX>(base, src, dst, w, gin, gout, tw, tid);
    else
      col_ct<DIR, R, S + 1, NS * RX>(base, dst, src, w, gin, gout, tw, tid);
  }
}

// k_col_generic with the direct path of length R for any panel width
// (items take the columns fastest when w is a power of two, else the
// butterflies).
template <int R, int DIR>
__global__ __launch_bounds__(256) void k_col_ct(
    Geom g, float2* F, const float2* __restrict__ twR, const Ctrl* ctrl,
    const float* __restrict__ parts, int nparts, double* bsum, int nred) {
  static_assert(ColPlan<R>::p.n >= 2, "direct path needs two stages");
  if (ctrl != nullptr && ctrl->done) return;
  extern __shared__ __align__(16) float2 lds[];
  __shared__ double red[5][16];
  const int lin = blockIdx.y * gridDim.x + blockIdx.x;
  if (nparts > 0 && lin < nred)
    partials_slice(parts, nparts, bsum, lin, nred, red);
  const int m = DIR > 0 ? (int)gridDim.y - 1 - (int)blockIdx.y
                        : (int)blockIdx.y;
  const int p = blockIdx.x;
  const int w = (p == g.P - 1) ? g.wlast : g.TC;
  constexpr int ld = R + 1;
  float2* base = F + (long long)m * g.Hc + (long long)p * g.TC * R;
  const int tsh = (w & (w - 1)) == 0 ? __ffs(w) - 1 : -1;
  gen::GIo gin;  // panel -> LDS
  gin.lda = 1;
  gin.esa = w;
  gin.ldb = ld;
  gin.esb = 1;
  gin.tsh = tsh;
  gen::GIo gout;  // LDS -> panel
  gout.lda = ld;
  gout.esa = 1;
  gout.ldb = 1;
  gout.esb = w;
  gout.tsh = tsh;
  col_ct<DIR, R>(base, lds, lds + w * ld, w, gin, gout, twR, threadIdx.x);
}

// The kernel of column length r and direction, or null, among the lengths
// 8 q with q = kCtColQMin + PART (mod kCtParts), q <= kCtColQMax.
template <int PART, int Q = kCtColQMin + PART>
const void* ct_col_pick(int r, int dir) {
  if constexpr (Q > kCtColQMax) {
    return nullptr;
  } else {
    if (r == 8 * Q)
      return dir > 0 ? reinterpret_cast<const void*>(&k_col_ct<8 * Q, 1>)
                     : reinterpret_cast<const void*>(&k_col_ct<8 * Q, -1>);
    return ct_col_pick<PART, Q + kCtParts>(r, dir);
  }
}

}  // namespace
}  // namespace scc
