// Register-resident Stockham FFT of power-of-two length L for gfx950.
//
// T = L / E threads cooperate on one transform; thread t holds the E
// elements t + T e (e < E) in registers.  Every stage uses radix RS <= E
// (RS divides E), so the inputs of a thread's butterflies are its own
// registers: butterfly j = t + T b reads elements j + r L/RS, i.e. slots
// b + r E/RS.  Outputs go to their natural Stockham positions in LDS and are
// read back into slots t + T e, so one stage costs one LDS write and one LDS
// read per element.  Stage radices: E while E divides what is left, then the
// remainder, e.g. L = 1024, E = 16 -> 16, 16, 4.
//
// The phases (compute, write, read) are __host__ __device__ so that the
// index and twiddle arithmetic is tested on the CPU with the barriers
// replaced by loops over threads.
#pragma once

#include <hip/hip_runtime.h>

namespace scc {
namespace reg {

__host__ __device__ __forceinline__ float2 c_add(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__host__ __device__ __forceinline__ float2 c_sub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__host__ __device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (-i)^DIR-ish: multiply by -i for the forward (DIR = 1) transform,
// by +i for the inverse (DIR = -1).
template <int DIR>
__host__ __device__ __forceinline__ float2 c_mul_mi(float2 a) {
  return DIR > 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// cos(2 pi m / 16), m = 0..15.
__host__ __device__ constexpr float kCos16(int m) {
  return m == 0 ? 1.0f
       : m == 1 ? 0.92387953251128674f
       : m == 2 ? 0.70710678118654752f
       : m == 3 ? 0.38268343236508979f
       : m == 4 ? 0.0f
       : m == 5 ? -0.38268343236508979f
       : m == 6 ? -0.70710678118654752f
       : m == 7 ? -0.92387953251128674f
       : m == 8 ? -1.0f
       : m == 9 ? -0.92387953251128674f
       : m == 10 ? -0.70710678118654752f
       : m == 11 ? -0.38268343236508979f
       : m == 12 ? 0.0f
       : m == 13 ? 0.38268343236508979f
       : m == 14 ? 0.70710678118654752f
       : 0.92387953251128674f;
}
// cos(2 pi m / 32), m = 0..31 (the even entries are kCos16(m / 2)).
__host__ __device__ constexpr float kCos32(int m) {
  return (m & 1) == 0 ? kCos16(m / 2)
       : (m == 1 || m == 31)  ? 0.98078528040323044f
       : (m == 3 || m == 29)  ? 0.83146961230254524f
       : (m == 5 || m == 27)  ? 0.55557023301960218f
       : (m == 7 || m == 25)  ? 0.19509032201612826f
       : (m == 9 || m == 23)  ? -0.19509032201612826f
       : (m == 11 || m == 21) ? -0.55557023301960218f
       : (m == 13 || m == 19) ? -0.83146961230254524f
                              : -0.98078528040323044f;  // 15, 17
}
// exp(-2 pi i DIR m / N) for N dividing 32.
template <int N, int DIR>
__host__ __device__ constexpr float2 w_const(int m) {
  return make_float2(kCos32(((m * (32 / N)) % 32)),
                     (float)(-DIR) * kCos32((((m * (32 / N)) % 32) + 24) % 32));
}

// In-place DFT of N = 2, 4, 8 or 16 points, forward (DIR = 1) or
// unnormalised inverse (DIR = -1).  N = 8, 16 use the split
// N = 4 * (N / 4): inner 4-point DFTs, constant twiddles, outer DFTs.
template <int N, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
  __host__ __device__ __forceinline__ static void run(float2* x) {
    const float2 a = x[0], b = x[1];
    x[0] = c_add(a, b);
    x[1] = c_sub(a, b);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __host__ __device__ __forceinline__ static void run(float2* x) {
    const float2 s02 = c_add(x[0], x[2]), d02 = c_sub(x[0], x[2]);
    const float2 s13 = c_add(x[1], x[3]), d13 = c_mul_mi<DIR>(c_sub(x[1], x[3]));
    x[0] = c_add(s02, s13);
    x[1] = c_add(d02, d13);
    x[2] = c_sub(s02, s13);
    x[3] = c_sub(d02, d13);
  }
};

template <int N, int DIR>
struct Dft {
  static_assert(N == 8 || N == 16, "radix");
  __host__ __device__ __forceinline__ static void run(float2* x) {
    constexpr int N2 = N / 4;  // n = N2 n1 + n2, k = k1 + 4 k2
    float2 y[N];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      float2 a[4];
#pragma unroll
      for (int n1 = 0; n1 < 4; ++n1) a[n1] = x[N2 * n1 + n2];
      Dft<4, DIR>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) {
        float2 v = a[k1];
        if (n2 * k1 != 0) v = c_mul(v, w_const<N, DIR>(n2 * k1));
        y[n2 * 4 + k1] = v;
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 b[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[n2] = y[n2 * 4 + k1];
      Dft<N2, DIR>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) x[k1 + 4 * k2] = b[k2];
    }
  }
};

// Padded LDS index: one spare slot after every 16 complex values keeps the
// strided Stockham writes free of bank conflicts.
__host__ __device__ __forceinline__ constexpr int lpad(int i) {
  return i + (i >> 4);
}
__host__ __device__ constexpr int lds_len(int L) { return L + (L >> 4) + 1; }

// lpad(a + b) = lpad(a) + lpad(b) for the index pairs of the exchanges
// below: b a multiple of a power of two P and a < P (slots: a = t < T,
// b = T e), or a = (j - k) RS + k with k < NS, (j - k) a multiple of NS and
// b = q NS (stage outputs).  In both cases a + b crosses no multiple of 16
// beyond those in b, so (a + b) >> 4 = (a >> 4) + (b >> 4).  With E <= 8
// points per thread the exchanges therefore address LDS as one per-thread
// base plus compile-time offsets (the offset field of the LDS instruction)
// instead of one address register per element: 7 % fewer VALU instructions
// and 3 % faster 512-point row kernels.  With 16 points per thread the
// per-element addresses measured faster (1024-point rows 6 %), so those
// transforms keep them.
template <int E>
struct LdsOffsetForm {
  static constexpr bool value = E <= 8;
};

template <int L, int E>
struct Plan {
  static_assert((L & (L - 1)) == 0 && (E & (E - 1)) == 0 && E <= L, "pow2");
  static constexpr int T = L / E;
  static constexpr int stages() {
    int n = 0, left = L;
    while (left > 1) {
      left /= (left >= E ? E : left);
      ++n;
    }
    return n;
  }
  static constexpr int radix(int s) {
    int left = L;
    for (int i = 0; i < s; ++i) left /= (left >= E ? E : left);
    return left >= E ? E : left;
  }
  static constexpr int ns(int s) {
    int n = 1;
    for (int i = 0; i < s; ++i) n *= radix(i);
    return n;
  }
};

// Twiddle W_L^e (table W_{L*tws}^t, t < L*tws), conjugated when DIR < 0.
template <int DIR>
__host__ __device__ __forceinline__ float2 tw_get(const float2* tw, int e,
                                                   int tws) {
  float2 w = tw[e * tws];
  if (DIR < 0) w.y = -w.y;
  return w;
}

// W_C^{t + T e} of a real row of C = 2 L points held as T lanes x E slots
// (C / T = 2 E): the lane's table entry W_C^t times the constant
// W_{2E}^e.
template <int E>
__host__ __device__ __forceinline__ float2 w_slot(float2 wt, int e) {
  return e == 0 ? wt : c_mul(wt, w_const<2 * E, 1>(e));
}

// Stage S: twiddle and butterflies on thread t's registers.
template <int L, int E, int S, int DIR>
__host__ __device__ __forceinline__ void stage_compute(float2 (&v)[E], int t,
                                                       const float2* tw,
                                                       int tws) {
  using P = Plan<L, E>;
  constexpr int RS = P::radix(S);
  constexpr int NS = P::ns(S);
  constexpr int NB = E / RS;
  constexpr int T = P::T;
  constexpr int STEP = L / (NS * RS);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    float2 x[RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) x[r] = v[b + r * NB];
    if (NS > 1) {
      // W^{k r STEP} = (W^{k STEP})^r: one table read, the powers by
      // products of lower powers (at most log2(RS) products deep).
      const int k = (t + T * b) & (NS - 1);
      float2 pw[RS];
      pw[1] = tw_get<DIR>(tw, k * STEP, tws);
#pragma unroll
      for (int r = 2; r < RS; ++r) pw[r] = c_mul(pw[r / 2], pw[r - r / 2]);
#pragma unroll
      for (int r = 1; r < RS; ++r) x[r] = c_mul(x[r], pw[r]);
    }
    Dft<RS, DIR>::run(x);
#pragma unroll
    for (int q = 0; q < RS; ++q) v[b + q * NB] = x[q];
  }
}

// Stage S outputs to their natural positions in the (padded) LDS buffer.
template <int L, int E, int S>
__host__ __device__ __forceinline__ void stage_write(const float2 (&v)[E],
                                                     int t, float2* lds) {
  using P = Plan<L, E>;
  constexpr int RS = P::radix(S);
  constexpr int NS = P::ns(S);
  constexpr int NB = E / RS;
  constexpr int T = P::T;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + T * b;
    const int k = j & (NS - 1);
    const int ob = (j - k) * RS + k;
    if constexpr (LdsOffsetForm<E>::value) {
      float2* base = lds + lpad(ob);
#pragma unroll
      for (int q = 0; q < RS; ++q) base[lpad(q * NS)] = v[b + q * NB];
    } else {
#pragma unroll
      for (int q = 0; q < RS; ++q) lds[lpad(ob + q * NS)] = v[b + q * NB];
    }
  }
}

// Slots t + T e from the LDS buffer (natural order).
template <int L, int E>
__host__ __device__ __forceinline__ void slots_read(float2 (&v)[E], int t,
                                                    const float2* lds) {
  constexpr int T = L / E;
  if constexpr (LdsOffsetForm<E>::value) {
    const float2* base = lds + lpad(t);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = base[lpad(T * e)];
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = lds[lpad(t + T * e)];
  }
}

template <int L, int E>
__host__ __device__ __forceinline__ void slots_write(const float2 (&v)[E],
                                                     int t, float2* lds) {
  constexpr int T = L / E;
  if constexpr (LdsOffsetForm<E>::value) {
    float2* base = lds + lpad(t);
#pragma unroll
    for (int e = 0; e < E; ++e) base[lpad(T * e)] = v[e];
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) lds[lpad(t + T * e)] = v[e];
  }
}

// The last stage needs no exchange: with NS = L / RS its butterfly j < NS
// writes elements j + q NS = t + T (b + q NB), i.e. the thread's own slots.
// So a transform whose last stage is kept in registers leaves element
// t + T e of the result in v[e].

// Ordering of LDS accesses between the lanes of one wavefront: LDS
// instructions of a wavefront execute in order, so only the compiler must
// be kept from moving accesses across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Transform whose T threads all belong to one wavefront: stage exchanges
// through this transform's private LDS region with wavefront ordering only
// (no block barrier).  Slots in, slots out (element t + T e in v[e]).
template <int L, int E, int DIR, int S = 0>
__device__ __forceinline__ void fft_wave(float2 (&v)[E], int t, float2* lds,
                                         const float2* tw, int tws) {
  stage_compute<L, E, S, DIR>(v, t, tw, tws);
  if constexpr (S + 1 < Plan<L, E>::stages()) {
    stage_write<L, E, S>(v, t, lds);
    wave_sync();
    slots_read<L, E>(v, t, lds);
    wave_sync();
    fft_wave<L, E, DIR, S + 1>(v, t, lds, tw, tws);
  }
}

// Transform spread over several wavefronts of the block: block barriers
// around every exchange.  Slots in, slots out.  Every thread of the block
// must call it.
template <int L, int E, int DIR, int S = 0>
__device__ __forceinline__ void fft_block(float2 (&v)[E], int t, float2* lds,
                                          const float2* tw, int tws) {
  stage_compute<L, E, S, DIR>(v, t, tw, tws);
  if constexpr (S + 1 < Plan<L, E>::stages()) {
    stage_write<L, E, S>(v, t, lds);
    __syncthreads();
    slots_read<L, E>(v, t, lds);
    __syncthreads();
    fft_block<L, E, DIR, S + 1>(v, t, lds, tw, tws);
  }
}

// Whole transform on the device: registers in (slots), result in natural
// order in lds on return (block barrier included).  Every thread of the
// block must call it.  Two barriers per stage: LDS is single-buffered.
template <int L, int E, int DIR, int S = 0>
__device__ __forceinline__ void fft_dev(float2 (&v)[E], int t, float2* lds,
                                        const float2* tw, int tws) {
  stage_compute<L, E, S, DIR>(v, t, tw, tws);
  stage_write<L, E, S>(v, t, lds);
  __syncthreads();
  if constexpr (S + 1 < Plan<L, E>::stages()) {
    slots_read<L, E>(v, t, lds);
    __syncthreads();
    fft_dev<L, E, DIR, S + 1>(v, t, lds, tw, tws);
  }
}

}  // namespace reg
}  // namespace scc
