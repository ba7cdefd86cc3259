// CPU check of the register-resident Stockham FFT phases (csrc/fft_reg.h):
// runs stage_compute / stage_write / slots_read for every thread in turn
// (the device's block barriers become loops) and compares the natural-order
// result with a direct double-precision DFT; also checks the real-FFT slot
// twiddles (lane table entry times a constant root of unity).  Prints one line per case and
// exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fft_reg.h"

using namespace scc::reg;

template <int L, int E, int DIR, int S>
void run_stages(std::vector<std::vector<float2>>& v, float2* lds,
                const float2* tw, int tws) {
  constexpr int T = L / E;
  for (int t = 0; t < T; ++t) {
    float2(&r)[E] = *reinterpret_cast<float2(*)[E]>(v[t].data());
    stage_compute<L, E, S, DIR>(r, t, tw, tws);
  }
  for (int t = 0; t < T; ++t) {
    const float2(&r)[E] = *reinterpret_cast<const float2(*)[E]>(v[t].data());
    stage_write<L, E, S>(r, t, lds);
  }
  if constexpr (S + 1 < Plan<L, E>::stages()) {
    for (int t = 0; t < T; ++t) {
      float2(&r)[E] = *reinterpret_cast<float2(*)[E]>(v[t].data());
      slots_read<L, E>(r, t, lds);
    }
    run_stages<L, E, DIR, S + 1>(v, lds, tw, tws);
  }
}

// Same stages with the last one kept in registers (fft_wave / fft_block).
template <int L, int E, int DIR, int S>
void run_stages_slots(std::vector<std::vector<float2>>& v, float2* lds,
                      const float2* tw, int tws) {
  constexpr int T = L / E;
  for (int t = 0; t < T; ++t) {
    float2(&r)[E] = *reinterpret_cast<float2(*)[E]>(v[t].data());
    stage_compute<L, E, S, DIR>(r, t, tw, tws);
  }
  if constexpr (S + 1 < Plan<L, E>::stages()) {
    for (int t = 0; t < T; ++t) {
      const float2(&r)[E] =
          *reinterpret_cast<const float2(*)[E]>(v[t].data());
      stage_write<L, E, S>(r, t, lds);
    }
    for (int t = 0; t < T; ++t) {
      float2(&r)[E] = *reinterpret_cast<float2(*)[E]>(v[t].data());
      slots_read<L, E>(r, t, lds);
    }
    run_stages_slots<L, E, DIR, S + 1>(v, lds, tw, tws);
  }
}

template <int L, int E, int DIR>
int check(int tws) {
  constexpr int T = L / E;
  const int Lt = L * tws;
  std::vector<float2> tw(Lt);
  for (int t = 0; t < Lt; ++t) {
    const double a = -2.0 * M_PI * t / Lt;
    tw[t] = make_float2((float)cos(a), (float)sin(a));
  }
  std::vector<double> xr(L), xi(L);
  srand(L * 7 + E + (DIR > 0));
  for (int n = 0; n < L; ++n) {
    xr[n] = rand() / (double)RAND_MAX - 0.5;
    xi[n] = rand() / (double)RAND_MAX - 0.5;
  }
  std::vector<std::vector<float2>> v(T, std::vector<float2>(E));
  for (int t = 0; t < T; ++t)
    for (int e = 0; e < E; ++e)
      v[t][e] = make_float2((float)xr[t + T * e], (float)xi[t + T * e]);
  std::vector<std::vector<float2>> v2 = v;
  std::vector<float2> lds(lds_len(L));
  run_stages<L, E, DIR, 0>(v, lds.data(), tw.data(), tws);
  std::vector<float2> lds2(lds_len(L));
  run_stages_slots<L, E, DIR, 0>(v2, lds2.data(), tw.data(), tws);
  double err = 0, ref = 0;
  for (int k = 0; k < L; ++k) {
    double sr = 0, si = 0;
    for (int n = 0; n < L; ++n) {
      const double a = -2.0 * M_PI * DIR * (double)((long long)n * k % L) / L;
      sr += xr[n] * cos(a) - xi[n] * sin(a);
      si += xr[n] * sin(a) + xi[n] * cos(a);
    }
    const float2 g = lds[lpad(k)];
    err = fmax(err, fabs(g.x - sr) + fabs(g.y - si));
    const float2 g2 = v2[k % T][k / T];
    err = fmax(err, fabs(g2.x - sr) + fabs(g2.y - si));
    ref = fmax(ref, fabs(sr) + fabs(si));
  }
  const double rel = err / ref;
  const bool ok = rel < 2e-6;
  printf("L=%d E=%d dir=%d tws=%d stages=%d rel_err=%.3e %s\n", L, E, DIR,
         tws, Plan<L, E>::stages(), rel, ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

// w_slot<E>(W_C^t, e) against W_C^{t + T e} in double (C = 2 L, T = L / E).
template <int L, int E>
int check_slot_twiddles() {
  constexpr int T = L / E, C = 2 * L;
  double err = 0;
  for (int t = 0; t < T; ++t) {
    const double a = -2.0 * M_PI * t / C;
    const float2 wt = make_float2((float)cos(a), (float)sin(a));
    for (int e = 0; e < E; ++e) {
      const double b = -2.0 * M_PI * (t + T * e) / C;
      const float2 w = w_slot<E>(wt, e);
      err = fmax(err, fabs(w.x - cos(b)) + fabs(w.y - sin(b)));
    }
  }
  const bool ok = err < 5e-7;
  printf("slot twiddles L=%d E=%d max_err=%.3e %s\n", L, E, err,
         ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

int main() {
  int bad = 0;
  bad += check_slot_twiddles<512, 8>();
  bad += check_slot_twiddles<128, 8>();
  bad += check_slot_twiddles<1024, 16>();
  bad += check_slot_twiddles<2048, 16>();
  bad += check<16, 16, 1>(1);
  bad += check<16, 4, -1>(1);
  bad += check<64, 8, 1>(2);
  bad += check<128, 16, -1>(1);
  bad += check<256, 16, 1>(1);
  bad += check<512, 8, 1>(2);
  bad += check<512, 8, -1>(2);
  bad += check<512, 16, -1>(1);
  bad += check<1024, 16, 1>(1);
  bad += check<1024, 16, -1>(2);
  bad += check<2048, 16, 1>(1);
  bad += check<2048, 16, -1>(1);
  bad += check<1024, 8, -1>(2);
  bad += check<256, 8, -1>(2);
  bad += check<128, 8, 1>(2);
  return bad;
}
