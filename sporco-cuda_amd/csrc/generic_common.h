// Device code of the generic-size kernels, shared by the single-image
// kernels (kernels_generic.hip) and the work-list kernels of the batched
// solver (batched.hip): the mixed-radix Stockham FFT in LDS, the panel-layout
// spectrum loads and stores, and the body of every stage with its block
// index and dynamic LDS passed in.
#pragma once

#include "fft_gen.h"
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
      : d(dd), inv(gen::quot_recip(dd)) {}
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

// --------------------------------------------------------------------------
// Mixed-radix Stockham FFT over nt transforms of length pl.L held in LDS at
// A + t*ld (element index gen::gpad).  Ping-pongs between A and B and
// returns the buffer holding the result.  DIR = +1 forward
// (W = exp(-2 pi i / L)), -1 inverse (unnormalised).  fft_block: every
// thread of the block calls it (block barriers between stages); fft_wave:
// one transform per wavefront (nt = 1), wavefront ordering only.
// --------------------------------------------------------------------------
// The plan is read through readfirstlane: it may live in global memory (the
// batched descriptors), where the compiler cannot see that it is uniform,
// and a divergent radix would turn the stage dispatch into divergent code.
__device__ __forceinline__ int uni(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// PFA: the plan may hold the prime-factor radices 6, 10, 12, 14 (column
// plans; row plans are built without them, which keeps the row kernels'
// stage code smaller).
template <int DIR, bool PFA = false>
__device__ float2* fft_block(float2* A, float2* B, int nt, int ld,
                             const FftPlan& pl, const float2* __restrict__ tw,
                             int tid, int nthr) {
  const int nst = uni(pl.nstages), L = uni(pl.L), tws = uni(pl.tw_stride);
  int Ns = 1;
  for (int s = 0; s < nst; ++s) {
    const int Rx = uni(pl.radix[s]);
    const gen::GStage st = gen::make_stage(L, Rx, Ns);
    if (gen::stage_pre_needed(st)) {
      gen::stage_pre<DIR>(A, ld, nt, st, tw, tws, tid, nthr);
      __syncthreads();
    }
    gen::stage<DIR, PFA>(A, B, ld, nt, st, tw, tws, tid, nthr);
    __syncthreads();
    float2* tmp = A;
    A = B;
    B = tmp;
    Ns *= Rx;
  }
  return A;
}

template <int DIR>
__device__ float2* fft_wave(float2* A, float2* B, const FftPlan& pl,
                            const float2* __restrict__ tw, int lane) {
  const int nst = uni(pl.nstages), L = uni(pl.L), tws = uni(pl.tw_stride);
  int Ns = 1;
  for (int s = 0; s < nst; ++s) {
    const int Rx = uni(pl.radix[s]);
    const gen::GStage st = gen::make_stage(L, Rx, Ns);
    if (gen::stage_pre_needed(st)) {
      gen::stage_pre<DIR>(A, 0, 1, st, tw, tws, lane, 64);
      reg::wave_sync();
    }
    gen::stage<DIR, false>(A, B, 0, 1, st, tw, tws, lane, 64);
    reg::wave_sync();
    float2* tmp = A;
    A = B;
    B = tmp;
    Ns *= Rx;
  }
  return A;
}

__device__ __forceinline__ float uni(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
template <class T>
__device__ __forceinline__ T* uni(T* p) {
  const unsigned long long a = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  return reinterpret_cast<T*>(((unsigned long long)hi << 32) | lo);
}

// Geometry made wavefront-uniform (see uni).
__device__ __forceinline__ Geom geom_uni(const Geom& s) {
  Geom g;
  g.R = uni(s.R);
  g.C = uni(s.C);
  g.M = uni(s.M);
  g.H = uni(s.H);
  g.Lr = uni(s.Lr);
  g.even = uni(s.even);
  g.TC = uni(s.TC);
  g.P = uni(s.P);
  g.wlast = uni(s.wlast);
  g.N = (long long)g.R * g.C;
  g.Hc = (long long)g.R * g.H;
  g.lgR = uni(s.lgR);
  return g;
}

template <bool PFA = false>
__device__ inline float2* fft_generic(float2* A, float2* B, int nt, int ld,
                                      const FftPlan& pl,
                                      const float2* __restrict__ tw, float dir,
                                      int tid, int nthr) {
  return dir > 0.f ? fft_block<1, PFA>(A, B, nt, ld, pl, tw, tid, nthr)
                   : fft_block<-1, PFA>(A, B, nt, ld, pl, tw, tid, nthr);
}

// Wavefronts of the wavefront-per-row blocks.
constexpr int kRowWaves = 4;

// Real value n of a row held as complex pairs (even widths) in an LDS
// buffer.
__device__ __forceinline__ float& real_at(float2* row, int n) {
  return reinterpret_cast<float*>(row + gen::gpad(n >> 1))[n & 1];
}
__device__ __forceinline__ float real_at(const float2* row, int n) {
  return reinterpret_cast<const float*>(row + gen::gpad(n >> 1))[n & 1];
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
      float2 xk = B[t * ld + gen::gpad(k)];
      float2 xn = B[t * ld + gen::gpad(L - k)];
      if (k == 0) {  // DC and Nyquist are real (C2R ignores their imag)
        xk.y = 0.f;
        xn.y = 0.f;
      }
      xn.y = -xn.y;
      const float2 s = cadd(xk, xn);
      const float2 wd = cmul(cconj(twC[k]), csub(xk, xn));
      A[t * ld + gen::gpad(k)] = make_float2(s.x - wd.y, s.y + wd.x);  // s + i*wd
    }
  } else {
    const int C = g.C;
    const FastDiv dc(C);
    for (int i = tid; i < nt * C; i += nthr) {
      const int t = dc.div(i);
      const int k = i - t * C;
      float2 v = (k < g.H) ? B[t * ld + gen::gpad(k)]
                           : cconj(B[t * ld + gen::gpad(C - k)]);
      if (k == 0) v.y = 0.f;
      A[t * ld + gen::gpad(k)] = v;
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
      const float2 zk = Z[t * ld + gen::gpad(k == L ? 0 : k)];
      const float2 zn = cconj(Z[t * ld + gen::gpad(k == 0 ? 0 : L - k)]);
      const float2 s = cadd(zk, zn);
      const float2 wd = cmul(twC[k], csub(zk, zn));
      X = make_float2(0.5f * (s.x + wd.y), 0.5f * (s.y - wd.x));
    } else {
      X = Z[t * ld + gen::gpad(k)];
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
    B[t * ld + gen::gpad(c)] = F[spec_addr(g, dR, (int)row0 + t, c)];
  }
}

// Rows of one row-kernel block for an LDS budget of `budget` bytes (two
// ping-pong buffers of Lr + 1 complex values per row), at most 16.
inline int row_nr_budget(const Geom& g, int budget) {
  const int ld = gen::gld(g.Lr);
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
  const int ld = gen::gld(g.Lr);
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
      if (rsf != 1.f) u = u / rsf;
      const float y = Y[gi];
      if (mode == kFwdYUtoV)
        Y[gi] = y + u;  // V with Y = shrink(V), U / rsf = V - Y
      else if (rsf != 1.f)
        U[gi] = u;
      a = y - u;
    }
    if (g.even)
      real_at(A + t * ld, n) = a;
    else
      A[t * ld + gen::gpad(n)] = make_float2(a, 0.f);
  }
  __syncthreads();
  const float2* Z = fft_generic(A, B, nt, ld, rp, twC, 1.f, tid, nthr);
  r2c_store(g, Z, nt, ld, row0, F, twC, tid, nthr);
}

// Column transforms of a panel of w columns (w a power of two) straight
// from and to global memory: the first stage reads its butterflies from the
// panel (row-major, adjacent columns on adjacent threads) into LDS, the
// middle stages run between the LDS buffers A and B, and the last writes
// its outputs back into the panel, which saves the copy-in and copy-out
// passes through LDS and their barriers.  Needs at least two stages (the
// panel is read completely, behind a barrier, before it is written).
template <int DIR>
__device__ void col_direct(const FftPlan& cp, float2* base, float2* A,
                           float2* B, int w, int ld,
                           const float2* __restrict__ tw, int tid, int nthr) {
  const int nst = uni(cp.nstages), L = uni(cp.L), tws = uni(cp.tw_stride);
  const int tsh = __ffs(w) - 1;
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
  const gen::GIo gmid = gen::gio_lds(ld);
  float2* src = A;
  float2* dst = B;
  int Ns = 1;
  for (int s = 0; s < nst; ++s) {
    const int Rx = uni(cp.radix[s]);
    const gen::GStage st = gen::make_stage(L, Rx, Ns);
    if (s == 0) {
      gen::stage_io<DIR, true>(base, A, gin, w, st, tw, tws, tid, nthr);
    } else {
      if (gen::stage_pre_needed(st)) {
        gen::stage_pre<DIR>(src, ld, w, st, tw, tws, tid, nthr);
        __syncthreads();
      }
      if (s == nst - 1) {
        gen::stage_io<DIR, true>(src, base, gout, w, st, tw, tws, tid, nthr);
      } else {
        gen::stage_io<DIR, true>(src, dst, gmid, w, st, tw, tws, tid, nthr);
        float2* tmp = src;
        src = dst;
        dst = tmp;
      }
    }
    if (s < nst - 1) __syncthreads();
    Ns *= Rx;
  }
}

// Column transforms of panel p of map m, in place.
__device__ inline void col_body(const Geom& g, const FftPlan& cp, float2* F,
                                const float2* __restrict__ twR, float dir,
                                int p, long long m, float2* lds) {
  const int w = (p == g.P - 1) ? g.wlast : g.TC;
  const int ld = gen::gld(g.R);
  float2* A = lds;
  float2* B = lds + w * ld;
  float2* base = F + m * g.Hc + (long long)p * g.TC * g.R;
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (uni(cp.nstages) >= 2 && (w & (w - 1)) == 0) {
    if (dir > 0.f)
      col_direct<1>(cp, base, A, B, w, ld, twR, tid, nthr);
    else
      col_direct<-1>(cp, base, A, B, w, ld, twR, tid, nthr);
    return;
  }
  const int n = g.R * w;
  const FastDiv dw(w);
  for (int i = tid; i < n; i += nthr) {
    const int r = dw.div(i);
    const int c = i - r * w;
    A[c * ld + gen::gpad(r)] = base[i];
  }
  __syncthreads();
  const float2* Z = fft_generic<true>(A, B, w, ld, cp, twR, dir, tid, nthr);
  for (int i = tid; i < n; i += nthr) {
    const int r = dw.div(i);
    const int c = i - r * w;
    base[i] = Z[c * ld + gen::gpad(r)];
  }
}

// Row kernel body for rows blk*nr..: inverse row FFT -> X, fused update and
// residual partial sums (stored at partials[j * nparts + blk]), then in
// fused mode the forward row FFT of Y - U.
__device__ inline void row_body(const Geom& g, const FftPlan& rp,
                                const Buffers& b, const Params& p, int mode,
                                int nr, long long blk, int nparts,
                                float2* lds, float (*red)[4]) {
  const int ld = gen::gld(g.Lr);
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
    const float xv = g.even ? real_at(P + t * ld, n) : P[t * ld + gen::gpad(n)].x;
    const long long gi = row * g.C + n;
    float a = update_point(g, p, b, xv * invN, lr, m, r, n, gi, acc,
                           p.vstate && mode == kRowFused);
    if (mode == kRowFused) {
      if (g.even)
        real_at(P + t * ld, n) = a;
      else
        P[t * ld + gen::gpad(n)] = make_float2(a, 0.f);
    }
  }
  block_sums_store(acc, red, b.partials, nparts, (int)blk, tid);
  if (mode == kRowFused) {
    __syncthreads();
    const float2* Z = fft_generic(P, Q, nt, ld, rp, b.twC, 1.f, tid, nthr);
    r2c_store(g, Z, nt, ld, row0, b.F, b.twC, tid, nthr);
  }
}

// --------------------------------------------------------------------------
// Wavefront-per-row bodies: the block's kRowWaves wavefronts take nr rows
// (wave_row), kRowWaves at a time, each in its own pair of LDS buffers of
// gld(Lr) complex values, ordered by wavefront sync only.  The row twiddle
// table (C values W_C^t) follows the buffers in LDS.  Used when buffers and
// table fit 64 KiB.
// --------------------------------------------------------------------------

// Row of wavefront wv at step s (s < nr / kRowWaves) of row block blk.
// Blocks are dealt to the 8 XCDs round-robin (block b on XCD b mod 8), and
// with 2-column panels the half-spectrum entries of 8 consecutive rows share
// one 128-byte line.  In plain order a block's first kRowWaves rows use half
// of each line and its next ones the other half a whole row-time later, by
// when the line has left the L2 and is fetched again (measured: 1.55 x the
// algorithmic reads).  So with panels of at most 2 columns, in every group
// of 16 blocks, blocks b and b + 8 (same XCD, resident together) take the
// two halves of the same 2 kRowWaves rows at the same step (1.03 x at
// 768 x 768).  Wider panels (kRowWaves rows fill a line) and the rows past
// the last full group keep the plain order.
__device__ __forceinline__ long long wave_row(const Geom& g, long long blk,
                                              int wv, int s, int nr,
                                              long long nrows) {
  const long long span = 16LL * nr;
  const long long nsg = g.TC <= 2 ? nrows / span : 0;
  if (blk < 16 * nsg) {
    const int k = (int)(blk & 15);
    return (blk >> 4) * span + (long long)(k & 7) * 2 * nr +
           2 * kRowWaves * s + kRowWaves * (k >> 3) + wv;
  }
  return nsg * span + (blk - 16 * nsg) * nr + wv + (long long)kRowWaves * s;
}
__host__ __device__ inline int row_wave_lds(const Geom& g) {
  return (kRowWaves * 2 * gen::gld(g.Lr) + g.C) * 8;
}
__host__ __device__ inline bool row_wave_fits(const Geom& g) {
  return row_wave_lds(g) <= 65536;
}

// Copy of the row twiddle table into the LDS after the wavefront buffers;
// every thread of the block calls it (ends with a block barrier).
__device__ inline float2* row_wave_table(const Geom& g, float2* lds,
                                         const float2* __restrict__ twC) {
  float2* T = lds + kRowWaves * 2 * gen::gld(g.Lr);
  for (int i = threadIdx.x; i < g.C; i += blockDim.x) T[i] = twC[i];
  __syncthreads();
  return T;
}

// Offset of (r, c) in one map's panel layout; TC = 1 << tsh.
__device__ __forceinline__ long long panel_off_sh(const Geom& g, int tsh,
                                                  int r, int c) {
  const int p = c >> tsh;
  const int w = (p == g.P - 1) ? g.wlast : g.TC;
  return ((long long)p * g.R << tsh) + (long long)r * w + (c - (p << tsh));
}

// Half spectrum row (m, r) of Fm -> complex input of the inverse row
// transform in A (B receives the raw row).
__device__ __forceinline__ void wave_c2r_load(const Geom& g, int tsh,
                                              const float2* __restrict__ Fm,
                                              int r, float2* A, float2* B,
                                              const float2* __restrict__ twC,
                                              int lane) {
  for (int c = lane; c < g.H; c += 64)
    B[gen::gpad(c)] = Fm[panel_off_sh(g, tsh, r, c)];
  reg::wave_sync();
  if (g.even) {
    const int L = g.Lr;
    for (int k = lane; k < L; k += 64) {
      float2 xk = B[gen::gpad(k)];
      float2 xn = B[gen::gpad(L - k)];
      if (k == 0) {  // DC and Nyquist are real (C2R ignores their imag)
        xk.y = 0.f;
        xn.y = 0.f;
      }
      xn.y = -xn.y;
      const float2 s = cadd(xk, xn);
      const float2 wd = cmul(cconj(twC[k]), csub(xk, xn));
      A[gen::gpad(k)] = make_float2(s.x - wd.y, s.y + wd.x);  // s + i*wd
    }
  } else {
    const int C = g.C;
    for (int k = lane; k < C; k += 64) {
      float2 v = (k < g.H) ? B[gen::gpad(k)] : cconj(B[gen::gpad(C - k)]);
      if (k == 0) v.y = 0.f;
      A[gen::gpad(k)] = v;
    }
  }
  reg::wave_sync();
}

// Forward-real-FFT post-processing of the length-Lr transform Z and store
// of half spectrum row (m, r) to Fm.
__device__ __forceinline__ void wave_r2c_store(const Geom& g, int tsh,
                                               const float2* Z, float2* Fm,
                                               int r,
                                               const float2* __restrict__ twC,
                                               int lane) {
  const int L = g.Lr;
  for (int k = lane; k < g.H; k += 64) {
    float2 X;
    if (g.even) {
      const float2 zk = Z[gen::gpad(k == L ? 0 : k)];
      const float2 zn = cconj(Z[gen::gpad(k == 0 ? 0 : L - k)]);
      const float2 s = cadd(zk, zn);
      const float2 wd = cmul(twC[k], csub(zk, zn));
      X = make_float2(0.5f * (s.x + wd.y), 0.5f * (s.y - wd.x));
    } else {
      X = Z[gen::gpad(k)];
    }
    Fm[panel_off_sh(g, tsh, r, k)] = X;
  }
}

// Row kernel body, one wavefront per row: inverse row FFT -> X, fused
// update and residual partial sums (reduced over the block into
// partials[j * nparts + blk]), then in fused mode the forward row FFT of
// Y - U.  twL: the row twiddle table in LDS (row_wave_table).
__device__ inline void row_body_wave(const Geom& g, const FftPlan& rp,
                                     const Buffers& b, const Params& p,
                                     int mode, int nr, long long blk,
                                     int nparts, float2* lds,
                                     const float2* twL, float (*red)[4]) {
  const int ld = gen::gld(g.Lr);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float2* A = lds + wv * 2 * ld;
  float2* B = A + ld;
  const long long nrows = (long long)g.M * g.R;
  const float rho = b.ctrl->rho;
  const float invN = 1.0f / (float)g.N;
  const float lr = p.lambda / rho;
  const int tsh = __ffs(g.TC) - 1;
  const int C = g.C;
  // The common case (see the update loop below).
  const bool plain = g.even && p.vstate && p.m0 == 0 && p.l1_mode != 2 &&
                     !p.nonneg && !p.nobndry;
  const bool alpha1 = p.alpha == 1.f;
  UpdateSums acc;
  for (int s = 0; s < nr / kRowWaves; ++s) {
    const long long row = wave_row(g, blk, wv, s, nr, nrows);
    if (row >= nrows) break;
    const int m = (int)(row / g.R);
    const int r = (int)(row - (long long)m * g.R);
    float2* Fm = b.F + (long long)m * g.Hc;
    wave_c2r_load(g, tsh, Fm, r, A, B, twL, lane);
    float2* P = fft_wave<-1>(A, B, rp, twL, lane);
    float2* Q = (P == A) ? B : A;
    const long long gi0 = row * C;
    if (plain) {
      // Even width, V state, one threshold per row: the real pair
      // (2k, 2k + 1) is complex slot k, V is read and written as float2,
      // no per-element option tests (update_point's arithmetic).
      const float th = lr * (p.l1_mode == 0 ? p.l1_scalar : b.l1w[m]);
      float2* V2 = reinterpret_cast<float2*>(b.Y + gi0);
      float2* U2 = reinterpret_cast<float2*>(b.U + gi0);
      for (int k = lane; k < g.Lr; k += 64) {
        const float2 xz = P[gen::gpad(k)];
        const float2 vo = V2[k];
        const float x0 = xz.x * invN, x1 = xz.y * invN;
        const float uo0 = split_u(vo.x, th, -th);
        const float uo1 = split_u(vo.y, th, -th);
        const float yo0 = vo.x - uo0, yo1 = vo.y - uo1;
        const float xr0 = alpha1 ? x0 : p.alpha * x0 + (1.f - p.alpha) * yo0;
        const float xr1 = alpha1 ? x1 : p.alpha * x1 + (1.f - p.alpha) * yo1;
        const float v0 = xr0 + uo0, v1 = xr1 + uo1;
        const float un0 = split_u(v0, th, -th);
        const float un1 = split_u(v1, th, -th);
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
        if (mode == kRowFused) {
          V2[k] = make_float2(v0, v1);
          P[gen::gpad(k)] = make_float2(yn0 - un0, yn1 - un1);
        } else {
          V2[k] = make_float2(yn0, yn1);
          U2[k] = make_float2(un0, un1);
        }
      }
    } else {
      for (int n = lane; n < C; n += 64) {
        float* slot = g.even ? &real_at(P, n) : &P[gen::gpad(n)].x;
        const float a = update_point(g, p, b, *slot * invN, lr, m, r, n,
                                     gi0 + n, acc,
                                     p.vstate && mode == kRowFused);
        if (mode == kRowFused) {
          if (g.even)
            *slot = a;
          else
            P[gen::gpad(n)] = make_float2(a, 0.f);
        }
      }
    }
    if (mode == kRowFused) {
      reg::wave_sync();
      const float2* Z = fft_wave<1>(P, Q, rp, twL, lane);
      wave_r2c_store(g, tsh, Z, Fm, r, twL, lane);
    }
    reg::wave_sync();  // this row's LDS reads before the next row's writes
  }
  block_sums_store(acc, red, b.partials, nparts, (int)blk, threadIdx.x);
}

// Forward row transform, one wavefront per row, of the rows of block blk
// (wave_row) of a real array (setup) or of Y - U/rsf (AutoRho), into F.  twL: the row twiddle
// table in LDS.
__device__ inline void row_fwd_body_wave(const Geom& g, long long nrows,
                                         const FftPlan& rp, int mode,
                                         const float* src, float* Y, float* U,
                                         float2* F, const float2* twL,
                                         float rsf, int nr, long long blk,
                                         float2* lds) {
  const int ld = gen::gld(g.Lr);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float2* A = lds + wv * 2 * ld;
  float2* B = A + ld;
  const int tsh = __ffs(g.TC) - 1;
  const int C = g.C;
  for (int s = 0; s < nr / kRowWaves; ++s) {
    const long long row = wave_row(g, blk, wv, s, nr, nrows);
    if (row >= nrows) break;
    const int m = (int)(row / g.R);
    const int r = (int)(row - (long long)m * g.R);
    const long long gi0 = row * C;
    for (int n = lane; n < C; n += 64) {
      const long long gi = gi0 + n;
      float a;
      if (mode == kFwdReal) {
        a = src[gi];
      } else {
        float u = U[gi];
        if (rsf != 1.f) u = u / rsf;
        const float y = Y[gi];
        if (mode == kFwdYUtoV)
          Y[gi] = y + u;  // V with Y = shrink(V), U / rsf = V - Y
        else if (rsf != 1.f)
          U[gi] = u;
        a = y - u;
      }
      if (g.even)
        real_at(A, n) = a;
      else
        A[gen::gpad(n)] = make_float2(a, 0.f);
    }
    reg::wave_sync();
    const float2* Z = fft_wave<1>(A, B, rp, twL, lane);
    wave_r2c_store(g, tsh, Z, F + (long long)m * g.Hc, r, twL, lane);
    reg::wave_sync();
  }
}

// Inverse real row transform of rows blk*nr.. of the half spectra F into
// out (real rows of length C), times scale.
__device__ inline void row_inv_body(const Geom& g, long long nrows,
                                    const FftPlan& rp, const float2* F,
                                    float* out, const float2* __restrict__ twC,
                                    float scale, int nr, long long blk,
                                    float2* lds) {
  const int ld = gen::gld(g.Lr);
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
    const float xv = g.even ? real_at(P + t * ld, n) : P[t * ld + gen::gpad(n)].x;
    out[(row0 + t) * g.C + n] = xv * scale;
  }
}

}  // namespace scc
