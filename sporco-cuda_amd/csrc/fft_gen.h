// Mixed-radix Stockham FFT with a run-time plan, for the generic-size
// kernels (any image size).  nt transforms of length L live in LDS at
// A + t * ld, element i at A + t * ld + gpad(i) (= i); a stage reads buffer A and
// writes buffer B (ping-pong).  Stage s of radix Rx with Ns = product of the
// earlier radices runs nb = L / Rx butterflies per transform; butterfly j
// (k = j mod Ns) reads elements j + r nb (r < Rx), multiplies them by the
// stage twiddles W_L^{r k L / (Ns Rx)}, takes their length-Rx DFT and writes
// output q to (j - k) Rx + k + q Ns.
//
// Butterflies: radix 2, 4, 8 as split-radix register DFTs with constant
// twiddles; odd primes 3 .. 13 as register DFTs with constant coefficients;
// larger primes as a direct DFT whose work items compute kQB output pairs
// of one butterfly with recurrent phases (all odd primes pair the inputs
// as x_r +- x_{P-r}).
//
// The per-stage functions are __host__ __device__ and loop over work items
// tid, tid + nthr, ..., so that the CPU check runs them for every item in
// turn with the barriers between stages replaced by the end of the loop.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_reg.h"

namespace scc {
namespace gen {

using reg::c_add;
using reg::c_mul;
using reg::c_sub;

// LDS index of element i.  Unpadded: with run-time plans the index
// arithmetic of a padded layout (one spare slot per 16 values) cost more
// than the bank conflicts it avoided (measured 5-13 % slower row and 3-11 %
// slower column transforms on uneven sizes).
__host__ __device__ __forceinline__ constexpr int gpad(int i) { return i; }
// Complex slots of one buffer holding elements 0 .. L (inclusive: the
// half-spectrum rows of even widths have L + 1 entries).
__host__ __device__ __forceinline__ constexpr int gld(int L) { return L + 1; }

// Image sides whose generic column and row transforms fit the LDS of one
// block: two buffers of a column in 64 KiB, of a row in 60 KiB.
__host__ __device__ inline bool sides_fit_lds(int R, int C) {
  const int Lr = (C % 2 == 0) ? C / 2 : C;
  return R >= 1 && C >= 1 && 16 * gld(R) <= 65536 && 16 * gld(Lr) <= 61440;
}

// Widest power-of-two panel (<= H columns) whose block-wide column
// transform (two buffers per column) fits `budget` bytes of LDS (at
// least one column).
__host__ __device__ inline int col_panel_width(int R, int H, int budget) {
  int cap = budget / (16 * gld(R));
  if (cap > H) cap = H;
  int TC = 1;
  while (TC * 2 <= cap) TC *= 2;
  return TC;
}

// Reciprocal of a divisor for the quotient estimates below: the hardware
// reciprocal (within 1 ulp) on the device, where an IEEE division would
// cost a ten-instruction sequence at every stage of every transform.
__host__ __device__ __forceinline__ float quot_recip(int d) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf((float)d);
#else
  return 1.f / (float)d;
#endif
}

// Quotient by a divisor fixed over a loop, for operands below 2^22: with
// the reciprocal within 2^-23 relative, the float estimate is within one of
// the quotient (its error is below (a / d) 2^-22 < 1), corrected once each
// way.
struct GDiv {
  int d;
  float inv;
  __host__ __device__ __forceinline__ explicit GDiv(int dd)
      : d(dd), inv(quot_recip(dd)) {}
  __host__ __device__ __forceinline__ int div(int a) const {
    int q = (int)((float)a * inv);
    const int r = a - q * d;
    if (r < 0) --q;
    if (r >= d) ++q;
    return q;
  }
};

// cos and sin of 2 pi k / P for the odd primes with register butterflies,
// k = 1 .. (P - 1) / 2.
__host__ __device__ constexpr float odd_cos(int P, int k) {
  return P == 3    ? -0.5f
         : P == 5  ? (k == 1 ? 0.30901699437494745f : -0.80901699437494734f)
         : P == 7  ? (k == 1   ? 0.62348980185873359f
                      : k == 2 ? -0.22252093395631434f
                               : -0.90096886790241903f)
         : P == 11 ? (k == 1   ? 0.84125353283118121f
                      : k == 2 ? 0.41541501300188644f
                      : k == 3 ? -0.142314838273285f
                      : k == 4 ? -0.65486073394528499f
                               : -0.95949297361449737f)
                   : (k == 1   ? 0.88545602565320991f
                      : k == 2 ? 0.56806474673115592f
                      : k == 3 ? 0.12053668025532301f
                      : k == 4 ? -0.35460488704253545f
                      : k == 5 ? -0.74851074817110119f
                               : -0.97094181742605201f);
}
__host__ __device__ constexpr float odd_sin(int P, int k) {
  return P == 3    ? 0.86602540378443871f
         : P == 5  ? (k == 1 ? 0.95105651629515353f : 0.58778525229247325f)
         : P == 7  ? (k == 1   ? 0.7818314824680298f
                      : k == 2 ? 0.97492791218182362f
                               : 0.43388373911755823f)
         : P == 11 ? (k == 1   ? 0.54064081745559756f
                      : k == 2 ? 0.90963199535451833f
                      : k == 3 ? 0.9898214418809328f
                      : k == 4 ? 0.75574957435425827f
                               : 0.28173255684142967f)
                   : (k == 1   ? 0.46472317204376851f
                      : k == 2 ? 0.82298386589365635f
                      : k == 3 ? 0.99270887409805397f
                      : k == 4 ? 0.93501624268541483f
                      : k == 5 ? 0.66312265824079519f
                               : 0.23931566428755768f);
}

// In-place length-P DFT, P odd prime <= 13, forward (DIR = 1:
// W = exp(-2 pi i / P)) or unnormalised inverse (DIR = -1).  With
// a_r = x_r + x_{P-r}, b_r = x_r - x_{P-r} and theta = 2 pi q r / P:
//   X_q     = x_0 + sum_r a_r cos(theta) - DIR i sum_r b_r sin(theta)
//   X_{P-q} = x_0 + sum_r a_r cos(theta) + DIR i sum_r b_r sin(theta).
template <int P, int DIR>
__host__ __device__ __forceinline__ void dft_odd(float2* x) {
  constexpr int H = (P - 1) / 2;
  float2 a[H], b[H];
#pragma unroll
  for (int r = 1; r <= H; ++r) {
    a[r - 1] = c_add(x[r], x[P - r]);
    b[r - 1] = c_sub(x[r], x[P - r]);
  }
  float2 y0 = x[0];
#pragma unroll
  for (int r = 0; r < H; ++r) y0 = c_add(y0, a[r]);
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    float2 t = x[0];
    float2 u = make_float2(0.f, 0.f);
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int e = (q * r) % P;
      const int ef = e <= H ? e : P - e;
      const float c = odd_cos(P, ef);
      const float s = e <= H ? odd_sin(P, ef) : -odd_sin(P, ef);
      t.x += c * a[r - 1].x;
      t.y += c * a[r - 1].y;
      u.x += s * b[r - 1].x;
      u.y += s * b[r - 1].y;
    }
    // DIR i u = DIR (-u.y, u.x)
    const float2 iu = make_float2(-(float)DIR * u.y, (float)DIR * u.x);
    x[q] = c_sub(t, iu);
    x[P - q] = c_add(t, iu);
  }
  x[0] = y0;
}

// Inverse of a modulo m (a, m coprime, m small).
__host__ __device__ constexpr int inv_mod(int a, int m) {
  for (int i = 1; i < m; ++i)
    if ((a * i) % m == 1) return i;
  return 1;
}

// In-place length A B DFT, A in {2, 4}, B an odd prime <= 7, by the prime-
// factor (Good-Thomas) index maps: with n = (B n1 + A n2) mod AB and k the
// residue with k = k1 (mod A), k = k2 (mod B), W_AB^{n k} = W_A^{n1 k1}
// W_B^{n2 k2}, so A B-point DFTs followed by B A-point DFTs need no
// twiddles between them (all indices are compile-time register names).
template <int A, int B, int DIR>
__host__ __device__ __forceinline__ void dft_pfa(float2* x) {
  constexpr int N = A * B;
  constexpr int eB = inv_mod(B % A, A), eA = inv_mod(A % B, B);
  float2 t[A][B];
#pragma unroll
  for (int n1 = 0; n1 < A; ++n1)
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) t[n1][n2] = x[(B * n1 + A * n2) % N];
#pragma unroll
  for (int n1 = 0; n1 < A; ++n1) dft_odd<B, DIR>(t[n1]);
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) {
    float2 u[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) u[n1] = t[n1][k2];
    reg::Dft<A, DIR>::run(u);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1)
      x[(B * eB * k1 + A * eA * k2) % N] = u[k1];
  }
}

template <int RX, int DIR>
__host__ __device__ __forceinline__ void bfly(float2* x) {
  if constexpr (RX == 2 || RX == 4 || RX == 8 || RX == 16)
    reg::Dft<RX, DIR>::run(x);
  else if constexpr (RX == 6)
    dft_pfa<2, 3, DIR>(x);
  else if constexpr (RX == 10)
    dft_pfa<2, 5, DIR>(x);
  else if constexpr (RX == 12)
    dft_pfa<4, 3, DIR>(x);
  else if constexpr (RX == 14)
    dft_pfa<2, 7, DIR>(x);
  else
    dft_odd<RX, DIR>(x);
}

// W_L^e from the table W_{L tws}^t, conjugated for the inverse.
template <int DIR>
__host__ __device__ __forceinline__ float2 tw_at(const float2* tw, int e,
                                                  int tws) {
  float2 w = tw[e * tws];
  if (DIR < 0) w.y = -w.y;
  return w;
}

// Shape of one stage.
struct GStage {
  int L, Rx, nb, Ns, tstep;
  bool p2;  // Ns is a power of two: k = j & (Ns - 1)
};

__host__ __device__ __forceinline__ GStage make_stage(int L, int Rx, int Ns) {
  GStage s;
  s.L = L;
  s.Rx = Rx;
  s.nb = L / Rx;
  s.Ns = Ns;
  s.tstep = L / (Ns * Rx);
  s.p2 = (Ns & (Ns - 1)) == 0;
  return s;
}

// Where a stage reads and writes: element i of transform t at
// A + t * lda + i * esa (and B + t * ldb + i * esb for the outputs).  The
// LDS buffers have (ld, 1); a column panel in global memory (w columns of R
// rows, row-major) has (1, w).  tsh >= 0: work items take the transforms
// fastest (t = idx & (nt - 1), nt = 1 << tsh), so that consecutive threads
// touch adjacent columns of the panel; tsh < 0: the butterflies fastest.
struct GIo {
  int lda, esa, ldb, esb, tsh;
};

__host__ __device__ __forceinline__ GIo gio_lds(int ld) {
  GIo io;
  io.lda = ld;
  io.esa = 1;
  io.ldb = ld;
  io.esb = 1;
  io.tsh = -1;
  return io;
}

// Register-butterfly stage (RX <= 14).
template <int RX, int DIR>
__host__ __device__ __forceinline__ void stage_reg(const float2* A, float2* B,
                                                   const GIo& io, int nt,
                                                   const GStage& s,
                                                   const float2* tw, int tws,
                                                   int tid, int nthr) {
  const int nb = s.nb, Ns = s.Ns;
  const GDiv dnb(nb), dns(Ns);
  const int total = nt * nb;
#pragma unroll 1
  for (int idx = tid; idx < total; idx += nthr) {
    int t = 0, j = idx;
    if (io.tsh >= 0) {
      t = idx & (nt - 1);
      j = idx >> io.tsh;
    } else if (nt > 1) {
      t = dnb.div(idx);
      j = idx - t * nb;
    }
    const int k = s.p2 ? (j & (Ns - 1)) : j - dns.div(j) * Ns;
    const float2* in = A + t * io.lda;
    float2* out = B + t * io.ldb;
    float2 x[RX];
#pragma unroll
    for (int r = 0; r < RX; ++r) x[r] = in[gpad(j + r * nb) * io.esa];
    if (k) {
      // W_L^{e1 r} = (W_L^{e1})^r: one table read, the powers by products
      // of lower powers (at most log2(RX) + 1 products deep).
      float2 pw[RX];
      pw[1] = tw_at<DIR>(tw, k * s.tstep, tws);
#pragma unroll
      for (int r = 2; r < RX; ++r) pw[r] = c_mul(pw[r / 2], pw[r - r / 2]);
#pragma unroll
      for (int r = 1; r < RX; ++r) x[r] = c_mul(x[r], pw[r]);
    }
    bfly<RX, DIR>(x);
    const int ob = (j - k) * RX + k;
#pragma unroll
    for (int q = 0; q < RX; ++q) out[gpad(ob + q * Ns) * io.esb] = x[q];
  }
}

// Prime radix P >= 17 (any odd radix the register butterflies do not
// cover), in two passes with a barrier between them.
//
// Pass 1 (stage_prime_twiddle, only when Ns > 1): the stage twiddles in
// place, x_n *= W_L^{r e1} for n = j + r nb, e1 = (j mod Ns) tstep, one
// product per element instead of one per element and work item.
//
// Pass 2 (stage_prime): outputs paired as in dft_odd: with the (twiddled)
// inputs x_r, a_r = x_r + x_{P-r}, b_r = x_r - x_{P-r} (r = 1 .. h,
// h = (P - 1) / 2) and z = W_P^{q r} (conjugated for the inverse),
//   t_q = x_0 + sum_r Re(z) a_r,  v_q = sum_r Im(z) b_r,
//   X_q = t_q + i v_q,  X_{P-q} = t_q - i v_q,  X_0 = x_0 + sum_r a_r.
// A work item computes the pairs q0 .. q0 + kQB - 1 of butterfly j (the
// item with q0 = 1 also X_0).  The phases advance by one complex product
// per input pair and are re-read from the table every 16 pairs, which
// bounds their rounding drift.
// True for the radices stage() sends to stage_prime (those without a
// register butterfly).
__host__ __device__ __forceinline__ bool prime_stage(const GStage& s) {
  const int r = s.Rx;
  return !(r == 2 || r == 3 || r == 4 || r == 5 || r == 6 || r == 7 ||
           r == 8 || r == 10 || r == 11 || r == 12 || r == 13 || r == 14);
}

template <int DIR>
__host__ __device__ __forceinline__ void stage_prime_twiddle(
    float2* A, int ld, int nt, const GStage& s, const float2* tw, int tws,
    int tid, int nthr) {
  const int L = s.L, nb = s.nb, Ns = s.Ns;
  const GDiv dL(L), dnb(nb), dns(Ns);
  const int total = nt * L;
#pragma unroll 1
  for (int idx = tid; idx < total; idx += nthr) {
    int t = 0, n = idx;
    if (nt > 1) {
      t = dL.div(idx);
      n = idx - t * L;
    }
    const int r = dnb.div(n);
    const int j = n - r * nb;
    const int k = s.p2 ? (j & (Ns - 1)) : j - dns.div(j) * Ns;
    if (r > 0 && k > 0) {
      float2* x = A + t * ld + gpad(n);
      *x = c_mul(*x, tw_at<DIR>(tw, r * k * s.tstep, tws));
    }
  }
}

// kQB: output pairs per work item (4 with run-time plans; the kernels with
// compile-time lengths pick it per stage).
template <int DIR, int kQB = 4>
__host__ __device__ __forceinline__ void stage_prime(const float2* A,
                                                     float2* B, const GIo& io,
                                                     int nt, const GStage& s,
                                                     const float2* tw, int tws,
                                                     int tid, int nthr) {
  const int Rx = s.Rx, nb = s.nb, Ns = s.Ns;
  const int h = (Rx - 1) / 2;
  const int nq = (h + kQB - 1) / kQB;
  const int per = nb * nq;
  const GDiv dper(per), dnb(nb), dns(Ns);
  const int total = nt * per;
#pragma unroll 1
  for (int idx = tid; idx < total; idx += nthr) {
    int t = 0, rem = idx;
    if (io.tsh >= 0) {
      t = idx & (nt - 1);
      rem = idx >> io.tsh;
    } else if (nt > 1) {
      t = dper.div(idx);
      rem = idx - t * per;
    }
    const int qb = dnb.div(rem);
    const int j = rem - qb * nb;
    const int k = s.p2 ? (j & (Ns - 1)) : j - dns.div(j) * Ns;
    const int q0 = 1 + qb * kQB;
    const float2* in = A + t * io.lda;
    const int ea = io.esa;
    const float2 x0 = in[gpad(j) * ea];
    float2 y0 = x0;
    float2 tq[kQB], vq[kQB], z[kQB], stp[kQB];
    int qq[kQB];
#pragma unroll
    for (int q = 0; q < kQB; ++q) {
      qq[q] = (q0 + q <= h) ? q0 + q : h;
      tq[q] = x0;
      vq[q] = make_float2(0.f, 0.f);
      stp[q] = tw_at<DIR>(tw, qq[q] * nb, tws);
      z[q] = stp[q];
    }
#pragma unroll 1
    for (int r = 1; r <= h; ++r) {
      const float2 xa = in[gpad(j + r * nb) * ea];
      const float2 xb = in[gpad(j + (Rx - r) * nb) * ea];
      const float2 a = c_add(xa, xb);
      const float2 b = c_sub(xa, xb);
      y0 = c_add(y0, a);
#pragma unroll
      for (int q = 0; q < kQB; ++q) {
        tq[q].x += z[q].x * a.x;
        tq[q].y += z[q].x * a.y;
        vq[q].x += z[q].y * b.x;
        vq[q].y += z[q].y * b.y;
      }
      if (((r + 1) & 15) == 0) {
#pragma unroll
        for (int q = 0; q < kQB; ++q)
          z[q] = tw_at<DIR>(tw, ((qq[q] * (r + 1)) % Rx) * nb, tws);
      } else {
#pragma unroll
        for (int q = 0; q < kQB; ++q) z[q] = c_mul(z[q], stp[q]);
      }
    }
    float2* out = B + t * io.ldb;
    const int eb = io.esb;
    const int ob = (j - k) * Rx + k;
    if (qb == 0) out[gpad(ob) * eb] = y0;
#pragma unroll
    for (int q = 0; q < kQB; ++q)
      if (q0 + q <= h) {
        const float2 iv = make_float2(-vq[q].y, vq[q].x);
        out[gpad(ob + (q0 + q) * Ns) * eb] = c_add(tq[q], iv);
        out[gpad(ob + (Rx - q0 - q) * Ns) * eb] = c_sub(tq[q], iv);
      }
  }
}

// Radix decomposition of a length L >= 1 into at most `max` stages, in
// order; returns the number of stages (0 for L = 1), or -1 if more are
// needed.  The power-of-two part 2^a runs as radix-8 stages; with `pfa` a
// remainder of 2 or 4 merges with the smallest odd prime into one prime-
// factor stage (6, 10, 12 or 14) when there is one, else it runs as one
// radix-4 or radix-2 stage (8 x 2 as 4 x 4); the remaining odd primes
// follow in ascending order.  Fewer, wider stages mean fewer passes over
// LDS and fewer partly idle item rounds (measured: 768-point columns 12 %
// faster); the row kernels, whose extra stage code cost more than the
// merge gained, take plans without them.
constexpr int plan_radices(int L, int* radix, int max, bool pfa = true) {
  int ns = 0;
  auto push = [&](int r) {
    if (ns >= max) return false;
    radix[ns++] = r;
    return true;
  };
  int n = L, a = 0;
  while (n % 2 == 0) {
    n /= 2;
    ++a;
  }
  const int e8 = a / 3;
  const int rest = 1 << (a % 3);
  int p0 = 0;  // smallest odd prime factor
  for (int p = 3; p <= n && !p0; p += 2)
    if (n % p == 0) p0 = p;
  if (pfa && rest > 1 && p0 > 0 && rest * p0 <= 14) {
    for (int i = 0; i < e8; ++i)
      if (!push(8)) return -1;
    if (!push(rest * p0)) return -1;
    n /= p0;
  } else {
    for (int i = 0; i < e8 - (rest == 2 && e8 > 0 ? 1 : 0); ++i)
      if (!push(8)) return -1;
    if (rest == 2 && e8 > 0) {
      if (!push(4) || !push(4)) return -1;
    } else if (rest > 1) {
      if (!push(rest)) return -1;
    }
  }
  for (int p = 3; n > 1; p += 2)
    while (n % p == 0) {
      if (!push(p)) return -1;
      n /= p;
    }
  return ns;
}

// Stage pre-pass: the in-place twiddles of a prime stage (prime_stage)
// with Ns > 1; the caller runs it on every thread, then a barrier, then
// stage().  No-op for the other stages (stage_pre_needed false).
__host__ __device__ __forceinline__ bool stage_pre_needed(const GStage& s) {
  return prime_stage(s) && s.Ns > 1;
}
template <int DIR>
__host__ __device__ inline void stage_pre(float2* A, int ld, int nt,
                                          const GStage& s, const float2* tw,
                                          int tws, int tid, int nthr) {
  if (stage_pre_needed(s))
    stage_prime_twiddle<DIR>(A, ld, nt, s, tw, tws, tid, nthr);
}

// One stage of radix Rx: A -> B with the addressing io.  PFA: also the
// prime-factor radices 6, 10, 12 and 14 (plans from plan_radices(..., true)).
template <int DIR, bool PFA = true>
__host__ __device__ inline void stage_io(const float2* A, float2* B,
                                         const GIo& io, int nt,
                                         const GStage& s, const float2* tw,
                                         int tws, int tid, int nthr) {
  if constexpr (PFA) {
    switch (s.Rx) {
      case 6: stage_reg<6, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); return;
      case 10: stage_reg<10, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); return;
      case 12: stage_reg<12, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); return;
      case 14: stage_reg<14, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); return;
      default: break;
    }
  }
  switch (s.Rx) {
    case 2: stage_reg<2, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
    case 4: stage_reg<4, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
    case 8: stage_reg<8, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
    case 3: stage_reg<3, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
    case 5: stage_reg<5, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
    case 7: stage_reg<7, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
    case 11: stage_reg<11, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
    case 13: stage_reg<13, DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
    default: stage_prime<DIR>(A, B, io, nt, s, tw, tws, tid, nthr); break;
  }
}

// One stage between two LDS buffers of stride ld.
template <int DIR, bool PFA = true>
__host__ __device__ inline void stage(const float2* A, float2* B, int ld,
                                      int nt, const GStage& s,
                                      const float2* tw, int tws, int tid,
                                      int nthr) {
  stage_io<DIR, PFA>(A, B, gio_lds(ld), nt, s, tw, tws, tid, nthr);
}

}  // namespace gen
}  // namespace scc
