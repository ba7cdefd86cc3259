// Device code of the generic-size kernels, shared by the single-image
// kernels (kernels_generic.hip) and the work-list kernels of the batched
// solver (batched.hip): the mixed-radix Stockham FFT in LDS, the panel-layout
// spectrum loads and stores, and the body of every stage with its block
// index and dynamic LDS passed in.
#pragma once

#include "scc_internal.h"
#include "scc_device.h"

namespace scc {

// Division by a value that is fixed over a loop, without the integer
// divide sequence: a float estimate of the quotient, corrected to the exact
// one (the estimate is off by at most one for the operands used here,
// a < 2^22).
struct FastDiv {
  int d;
  float inv;
  __device__ __forceinline__ explicit FastDiv(int dd)
      : d(dd), inv(1.f / (float)dd) {}
  __device__ __forceinline__ int div(int a) const {
    int q = (int)((float)a * inv);
    int r = a - q * d;
    while (r < 0) {
      --q;
      r += d;
    }
    while (r >= d) {
      ++q;
      r -= d;
    }
    return q;
  }
};

// Radix-RX butterfly j of a Stockham stage from registers: inputs
// in[r nb] (r < RX) times the stage twiddles W_L^{k tstep r}, a length-RX
// DFT with W_RX^{q r} = W_L^{(q r mod RX) nb}, outputs out[q Ns].
template <int RX>
__device__ __forceinline__ void small_dft(const float2* in, float2* out,
                                          int nb, int Ns, int k, int tstep,
                                          const float2* __restrict__ tw,
                                          int st, float dir) {
  float2 x[RX];
#pragma unroll
  for (int r = 0; r < RX; ++r) x[r] = in[r * nb];
  if (k) {
#pragma unroll
    for (int r = 1; r < RX; ++r)
      x[r] = cmul(x[r], twiddle(tw, k * tstep * r * st, dir));
  }
#pragma unroll
  for (int q = 0; q < RX; ++q) {
    float2 acc = x[0];
#pragma unroll
    for (int r = 1; r < RX; ++r)
      acc = cadd(acc, cmul(x[r], twiddle(tw, ((q * r) % RX) * nb * st, dir)));
    out[q * Ns] = acc;
  }
}

// --------------------------------------------------------------------------
// Mixed-radix Stockham FFT over nt transforms of length pl.L held in LDS at
// A + t*ld.  Ping-pongs between A and B and returns the buffer holding the
// result.  dir = +1 forward (W = exp(-2 pi i / L)), -1 inverse (unnormalised).
// Every thread of the block must call it (it contains block barriers).
// Radices 2 and 4 and the small odd primes 3, 5, 7 run one butterfly per
// thread from registers; larger prime radices spread the outputs of a
// butterfly over several threads (so that a radix-p stage keeps all threads
// busy): output q of butterfly j is
//   sum_r (in[j + r nb] W_L^{r k tstep}) W_Rx^{q r},  k = j mod Ns.
// --------------------------------------------------------------------------
__device__ float2* fft_generic(float2* A, float2* B, int nt, int ld,
                               const FftPlan& pl,
                               const float2* __restrict__ tw, float dir,
                               int tid, int nthr) {
  const int L = pl.L;
  const int st = pl.tw_stride;
  int Ns = 1;
  for (int s = 0; s < pl.nstages; ++s) {
    const int Rx = pl.radix[s];
    const int nb = L / Rx;
    const int tstep = L / (Ns * Rx);
    const FastDiv dnb(nb), dns(Ns);
    if (Rx <= 7) {
      const int total = nt * nb;
      for (int idx = tid; idx < total; idx += nthr) {
        const int t = dnb.div(idx);
        const int j = idx - t * nb;
        const float2* in = A + t * ld + j;
        const int k = j - dns.div(j) * Ns;
        float2* out = B + t * ld + (j - k) * Rx + k;
        if (Rx == 2) {
          float2 x0 = in[0];
          float2 x1 = in[nb];
          if (k) x1 = cmul(x1, twiddle(tw, k * tstep * st, dir));
          out[0] = cadd(x0, x1);
          out[Ns] = csub(x0, x1);
        } else if (Rx == 4) {
          float2 x0 = in[0];
          float2 x1 = in[nb];
          float2 x2 = in[2 * nb];
          float2 x3 = in[3 * nb];
          if (k) {
            x1 = cmul(x1, twiddle(tw, k * tstep * st, dir));
            x2 = cmul(x2, twiddle(tw, 2 * k * tstep * st, dir));
            x3 = cmul(x3, twiddle(tw, 3 * k * tstep * st, dir));
          }
          float2 s02 = cadd(x0, x2), d02 = csub(x0, x2);
          float2 s13 = cadd(x1, x3), d13 = csub(x1, x3);
          float2 jd = mul_mi(d13, dir);  // -i*dir*d13
          out[0] = cadd(s02, s13);
          out[Ns] = cadd(d02, jd);
          out[2 * Ns] = csub(s02, s13);
          out[3 * Ns] = csub(d02, jd);
        } else if (Rx == 3) {
          small_dft<3>(in, out, nb, Ns, k, tstep, tw, st, dir);
        } else if (Rx == 5) {
          small_dft<5>(in, out, nb, Ns, k, tstep, tw, st, dir);
        } else {
          small_dft<7>(in, out, nb, Ns, k, tstep, tw, st, dir);
        }
      }
    } else {
      // Prime radix >= 11: a work item computes kQB outputs q0 .. q0 + kQB
      // - 1 of butterfly j.  Each input is read once and multiplied by its
      // stage twiddle W_L^{r k tstep}; the DFT phases W_Rx^{q r} advance by
      // one complex product per input (W_Rx^q) and are re-read from the
      // table every 16 inputs, which bounds their rounding drift.
      constexpr int kQB = 4;
      const int nq = (Rx + kQB - 1) / kQB;
      const FastDiv dblk(nb * nq), drx(Rx);
      const int total = nt * nb * nq;
      for (int idx = tid; idx < total; idx += nthr) {
        const int t = dblk.div(idx);
        const int rem = idx - t * nb * nq;
        const int qb = dnb.div(rem);
        const int j = rem - qb * nb;
        const int k = j - dns.div(j) * Ns;
        const int q0 = qb * kQB;
        const float2* in = A + t * ld + j;
        float2 acc[kQB], z[kQB], stp[kQB];
#pragma unroll
        for (int q = 0; q < kQB; ++q) {
          acc[q] = make_float2(0.f, 0.f);
          z[q] = make_float2(1.f, 0.f);
          const int qq = min(q0 + q, Rx - 1);
          stp[q] = twiddle(tw, qq * nb * st, dir);
        }
        for (int r = 0; r < Rx; ++r) {
          float2 x = in[r * nb];
          if (k && r) x = cmul(x, twiddle(tw, r * k * tstep * st, dir));
#pragma unroll
          for (int q = 0; q < kQB; ++q) acc[q] = cadd(acc[q], cmul(x, z[q]));
          if (((r + 1) & 15) == 0) {
#pragma unroll
            for (int q = 0; q < kQB; ++q) {
              const int e = min(q0 + q, Rx - 1) * (r + 1);
              z[q] = twiddle(tw, (e - drx.div(e) * Rx) * nb * st, dir);
            }
          } else {
#pragma unroll
            for (int q = 0; q < kQB; ++q) z[q] = cmul(z[q], stp[q]);
          }
        }
        float2* out = B + t * ld + (j - k) * Rx + k;
#pragma unroll
        for (int q = 0; q < kQB; ++q)
          if (q0 + q < Rx) out[(q0 + q) * Ns] = acc[q];
      }
    }
    __syncthreads();
    float2* tmp = A;
    A = B;
    B = tmp;
    Ns *= Rx;
  }
  return A;
}

// Element i of the (nt rows x H columns) half-spectrum block of rows
// row0..row0+nt-1, enumerated panel by panel so that consecutive i touch
// consecutive addresses when the rows are consecutive rows of one map.
struct PanelIndex {
  int full, per, TC, wlast, c0;
  FastDiv dper, dtc, dwl;
  __device__ __forceinline__ PanelIndex(const Geom& g, int nt)
      : full((g.P - 1) * nt * g.TC), per(nt * g.TC), TC(g.TC),
        wlast(g.wlast), c0((g.P - 1) * g.TC), dper(nt * g.TC), dtc(g.TC),
        dwl(g.wlast) {}
  __device__ __forceinline__ void operator()(int i, int& t, int& c) const {
    if (i < full) {
      const int p = dper.div(i);
      const int rem = i - p * per;
      t = dtc.div(rem);
      c = p * TC + (rem - t * TC);
    } else {
      const int rem = i - full;
      t = dwl.div(rem);
      c = c0 + (rem - t * wlast);
    }
  }
};

// Offset of (row, c) of the half spectra of consecutive maps (row counts
// rows of all maps); dR divides by g.R.
__device__ inline long long spec_addr(const Geom& g, const FastDiv& dR,
                                      int row, int c) {
  const int m = dR.div(row);
  const int r = row - m * g.R;
  return (long long)m * g.Hc + panel_off(g, r, c);
}

// Inverse-real-FFT pre-processing of one block of half spectra (in B) into
// the complex input of the length-Lr inverse transform (in A).
__device__ void c2r_prep(const Geom& g, const float2* B, float2* A, int nt,
                         int ld, const float2* __restrict__ twC, int tid,
                         int nthr) {
  if (g.even) {
    const int L = g.Lr;
    const FastDiv dl(L);
    for (int i = tid; i < nt * L; i += nthr) {
      const int t = dl.div(i);
      const int k = i - t * L;
      float2 xk = B[t * ld + k];
      float2 xn = B[t * ld + L - k];
      if (k == 0) {  // DC and Nyquist are real (C2R ignores their imag)
        xk.y = 0.f;
        xn.y = 0.f;
      }
      xn.y = -xn.y;
      const float2 s = cadd(xk, xn);
      const float2 wd = cmul(cconj(twC[k]), csub(xk, xn));
      A[t * ld + k] = make_float2(s.x - wd.y, s.y + wd.x);  // s + i*wd
    }
  } else {
    const int C = g.C;
    const FastDiv dc(C);
    for (int i = tid; i < nt * C; i += nthr) {
      const int t = dc.div(i);
      const int k = i - t * C;
      float2 v = (k < g.H) ? B[t * ld + k] : cconj(B[t * ld + C - k]);
      if (k == 0) v.y = 0.f;
      A[t * ld + k] = v;
    }
  }
}

// Forward-real-FFT post-processing of the length-Lr transform (in Z) and
// store of the half spectra to F in panel layout.
__device__ void r2c_store(const Geom& g, const float2* Z, int nt, int ld,
                          long long row0, float2* __restrict__ F,
                          const float2* __restrict__ twC, int tid, int nthr) {
  const int L = g.Lr;
  const PanelIndex pix(g, nt);
  const FastDiv dR(g.R);
  for (int i = tid; i < nt * g.H; i += nthr) {
    int t, k;
    pix(i, t, k);
    float2 X;
    if (g.even) {
      const float2 zk = Z[t * ld + (k == L ? 0 : k)];
      const float2 zn = cconj(Z[t * ld + (k == 0 ? 0 : L - k)]);
      const float2 s = cadd(zk, zn);
      const float2 wd = cmul(twC[k], csub(zk, zn));
      X = make_float2(0.5f * (s.x + wd.y), 0.5f * (s.y - wd.x));
    } else {
      X = Z[t * ld + k];
    }
    F[spec_addr(g, dR, (int)row0 + t, k)] = X;
  }
}

__device__ void spec_load(const Geom& g, const float2* __restrict__ F,
                          float2* B, int nt, int ld, long long row0, int tid,
                          int nthr) {
  const PanelIndex pix(g, nt);
  const FastDiv dR(g.R);
  for (int i = tid; i < nt * g.H; i += nthr) {
    int t, c;
    pix(i, t, c);
    B[t * ld + c] = F[spec_addr(g, dR, (int)row0 + t, c)];
  }
}

// Rows of one row-kernel block for an LDS budget of `budget` bytes (two
// ping-pong buffers of Lr + 1 complex values per row), at most 16.
inline int row_nr_budget(const Geom& g, int budget) {
  const int ld = g.Lr + 1;
  int nr = budget / (2 * ld * 8);
  if (nr > 16) nr = 16;
  if (nr < 1) nr = 1;
  return nr;
}

// Forward row transform of rows row0.. of a real array (setup) or of
// Y - U/rsf (AutoRho), nt rows, into F.
__device__ inline void row_fwd_body(const Geom& g, long long nrows,
                                    const FftPlan& rp, int mode,
                                    const float* src, float* Y, float* U,
                                    float2* F, const float2* __restrict__ twC,
                                    float rsf, int nr, long long blk,
                                    float2* lds) {
  const int ld = g.Lr + 1;
  float2* A = lds;
  float2* B = lds + nr * ld;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long row0 = blk * nr;
  const int nt = (int)min((long long)nr, nrows - row0);
  const FastDiv dc(g.C);
  for (int i = tid; i < nt * g.C; i += nthr) {
    const int t = dc.div(i);
    const int n = i - t * g.C;
    const long long gi = (row0 + t) * g.C + n;
    float a;
    if (mode == kFwdReal) {
      a = src[gi];
    } else {
      float u = U[gi];
      if (rsf != 1.f) {
        u = u / rsf;
        U[gi] = u;
      }
      a = Y[gi] - u;
    }
    if (g.even)
      reinterpret_cast<float*>(A + t * ld)[n] = a;
    else
      A[t * ld + n] = make_float2(a, 0.f);
  }
  __syncthreads();
  const float2* Z = fft_generic(A, B, nt, ld, rp, twC, 1.f, tid, nthr);
  r2c_store(g, Z, nt, ld, row0, F, twC, tid, nthr);
}

// Column transforms of panel p of map m, in place.
__device__ inline void col_body(const Geom& g, const FftPlan& cp, float2* F,
                                const float2* __restrict__ twR, float dir,
                                int p, long long m, float2* lds) {
  const int w = (p == g.P - 1) ? g.wlast : g.TC;
  const int ld = g.R + 1;
  float2* A = lds;
  float2* B = lds + w * ld;
  float2* base = F + m * g.Hc + (long long)p * g.TC * g.R;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int n = g.R * w;
  const FastDiv dw(w);
  for (int i = tid; i < n; i += nthr) {
    const int r = dw.div(i);
    const int c = i - r * w;
    A[c * ld + r] = base[i];
  }
  __syncthreads();
  const float2* Z = fft_generic(A, B, w, ld, cp, twR, dir, tid, nthr);
  for (int i = tid; i < n; i += nthr) {
    const int r = dw.div(i);
    const int c = i - r * w;
    base[i] = Z[c * ld + r];
  }
}

// Row kernel body for rows blk*nr..: inverse row FFT -> X, fused update and
// residual partial sums (stored at partials[j * nparts + blk]), then in
// fused mode the forward row FFT of Y - U.
__device__ inline void row_body(const Geom& g, const FftPlan& rp,
                                const Buffers& b, const Params& p, int mode,
                                int nr, long long blk, int nparts,
                                float2* lds, float (*red)[4]) {
  const int ld = g.Lr + 1;
  float2* A = lds;
  float2* B = lds + nr * ld;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long nrows = (long long)g.M * g.R;
  const long long row0 = blk * nr;
  const int nt = (int)min((long long)nr, nrows - row0);
  const float rho = b.ctrl->rho;

  spec_load(g, b.F, B, nt, ld, row0, tid, nthr);
  __syncthreads();
  c2r_prep(g, B, A, nt, ld, b.twC, tid, nthr);
  __syncthreads();
  float2* P = fft_generic(A, B, nt, ld, rp, b.twC, -1.f, tid, nthr);
  float2* Q = (P == A) ? B : A;

  UpdateSums acc;
  const float invN = 1.0f / (float)g.N;
  const float lr = p.lambda / rho;
  const FastDiv dc(g.C), dR(g.R);
  for (int i = tid; i < nt * g.C; i += nthr) {
    const int t = dc.div(i);
    const int n = i - t * g.C;
    const long long row = row0 + t;
    const int m = dR.div((int)row);
    const int r = (int)row - m * g.R;
    const float xv = g.even ? reinterpret_cast<const float*>(P + t * ld)[n]
                            : P[t * ld + n].x;
    const long long gi = row * g.C + n;
    float a = update_point(g, p, b, xv * invN, lr, m, r, n, gi, acc);
    if (mode == kRowFused) {
      if (g.even)
        reinterpret_cast<float*>(P + t * ld)[n] = a;
      else
        P[t * ld + n] = make_float2(a, 0.f);
    }
  }
  block_sums_store(acc, red, b.partials, nparts, (int)blk, tid);
  if (mode == kRowFused) {
    __syncthreads();
    const float2* Z = fft_generic(P, Q, nt, ld, rp, b.twC, 1.f, tid, nthr);
    r2c_store(g, Z, nt, ld, row0, b.F, b.twC, tid, nthr);
  }
}

// Inverse real row transform of rows blk*nr.. of the half spectra F into
// out (real rows of length C), times scale.
__device__ inline void row_inv_body(const Geom& g, long long nrows,
                                    const FftPlan& rp, const float2* F,
                                    float* out, const float2* __restrict__ twC,
                                    float scale, int nr, long long blk,
                                    float2* lds) {
  const int ld = g.Lr + 1;
  float2* A = lds;
  float2* B = lds + nr * ld;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long row0 = blk * nr;
  const int nt = (int)min((long long)nr, nrows - row0);
  spec_load(g, F, B, nt, ld, row0, tid, nthr);
  __syncthreads();
  c2r_prep(g, B, A, nt, ld, twC, tid, nthr);
  __syncthreads();
  const float2* P = fft_generic(A, B, nt, ld, rp, twC, -1.f, tid, nthr);
  const FastDiv dc(g.C);
  for (int i = tid; i < nt * g.C; i += nthr) {
    const int t = dc.div(i);
    const int n = i - t * g.C;
    const float xv = g.even ? reinterpret_cast<const float*>(P + t * ld)[n]
                            : P[t * ld + n].x;
    out[(row0 + t) * g.C + n] = xv * scale;
  }
}

}  // namespace scc
