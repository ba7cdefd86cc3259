// CPU check of the run-time-plan Stockham FFT stages (csrc/fft_gen.h): runs
// gen::stage (after gen::stage_pre where a stage has one) for every work
// item in turn (the barriers become the ends of the item loops), for one transform per wavefront
// (nt = 1, 64 items at a time) and several per block (nt > 1, 256), and
// compares the result with a direct double-precision DFT.  Also checks the
// float-based divider against integer division and the plan builder
// (gen::plan_radices) on every length up to 8192.  Prints one line per
// case and exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fft_gen.h"

using namespace scc::gen;

static int check(int L, const std::vector<int>& radices, int tws, int dir,
                 int nt, int nthr) {
  int prod = 1;
  for (int r : radices) prod *= r;
  if (prod != L) {
    printf("bad plan for L=%d\n", L);
    return 1;
  }
  const int Lt = L * tws;
  std::vector<float2> tw(Lt);
  for (int t = 0; t < Lt; ++t) {
    const double a = -2.0 * M_PI * t / Lt;
    tw[t] = make_float2((float)cos(a), (float)sin(a));
  }
  const int ld = gld(L) + (nt > 1 ? 3 : 0);
  std::vector<float2> A(nt * ld, make_float2(1e30f, 1e30f)), B = A;
  std::vector<double> xr(nt * L), xi(nt * L);
  srand(L * 131 + nt * 7 + (dir > 0) + tws * 3);
  for (int t = 0; t < nt; ++t)
    for (int n = 0; n < L; ++n) {
      xr[t * L + n] = rand() / (double)RAND_MAX - 0.5;
      xi[t * L + n] = rand() / (double)RAND_MAX - 0.5;
      A[t * ld + gpad(n)] =
          make_float2((float)xr[t * L + n], (float)xi[t * L + n]);
    }
  float2* a = A.data();
  float2* b = B.data();
  int Ns = 1;
  for (int r : radices) {
    const GStage st = make_stage(L, r, Ns);
    if (stage_pre_needed(st))
      for (int tid = 0; tid < nthr; ++tid) {
        if (dir > 0)
          stage_pre<1>(a, ld, nt, st, tw.data(), tws, tid, nthr);
        else
          stage_pre<-1>(a, ld, nt, st, tw.data(), tws, tid, nthr);
      }
    for (int tid = 0; tid < nthr; ++tid) {
      if (dir > 0)
        stage<1>(a, b, ld, nt, st, tw.data(), tws, tid, nthr);
      else
        stage<-1>(a, b, ld, nt, st, tw.data(), tws, tid, nthr);
    }
    float2* tmp = a;
    a = b;
    b = tmp;
    Ns *= r;
  }
  double err = 0, ref = 0;
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < L; ++k) {
      double sr = 0, si = 0;
      for (int n = 0; n < L; ++n) {
        const double ang =
            -2.0 * M_PI * dir * (double)((long long)n * k % L) / L;
        sr += xr[t * L + n] * cos(ang) - xi[t * L + n] * sin(ang);
        si += xr[t * L + n] * sin(ang) + xi[t * L + n] * cos(ang);
      }
      const float2 g = a[t * ld + gpad(k)];
      err = fmax(err, fabs(g.x - sr) + fabs(g.y - si));
      ref = fmax(ref, fabs(sr) + fabs(si));
    }
  const double rel = err / ref;
  const bool ok = rel < 3e-6;
  printf("L=%d plan=", L);
  for (size_t i = 0; i < radices.size(); ++i)
    printf("%s%d", i ? "x" : "", radices[i]);
  printf(" dir=%d tws=%d nt=%d nthr=%d rel_err=%.3e %s\n", dir, tws, nt, nthr,
         rel, ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

// The column kernels' direct path (generic_common.h col_direct): w
// transforms of length L held as a row-major panel (element n of column t
// at P[n w + t]); the first stage reads the panel, the middle ones run
// between two LDS-like buffers, the last writes the panel in place.
static int check_panel(int L, const std::vector<int>& radices, int w,
                       int dir, int nthr) {
  const int tws = 1;
  std::vector<float2> tw(L);
  for (int t = 0; t < L; ++t) {
    const double a = -2.0 * M_PI * t / L;
    tw[t] = make_float2((float)cos(a), (float)sin(a));
  }
  const int ld = gld(L);
  std::vector<float2> P(L * w), A(w * ld, make_float2(1e30f, 1e30f)), B = A;
  std::vector<double> xr(w * L), xi(w * L);
  srand(L * 17 + w * 5 + (dir > 0));
  for (int n = 0; n < L; ++n)
    for (int t = 0; t < w; ++t) {
      xr[t * L + n] = rand() / (double)RAND_MAX - 0.5;
      xi[t * L + n] = rand() / (double)RAND_MAX - 0.5;
      P[n * w + t] = make_float2((float)xr[t * L + n], (float)xi[t * L + n]);
    }
  int tsh = 0;
  while ((1 << tsh) < w) ++tsh;
  GIo gin, gout;
  gin.lda = 1;
  gin.esa = w;
  gin.ldb = ld;
  gin.esb = 1;
  gin.tsh = tsh;
  gout.lda = ld;
  gout.esa = 1;
  gout.ldb = 1;
  gout.esb = w;
  gout.tsh = tsh;
  const GIo gmid = gio_lds(ld);
  float2* src = A.data();
  float2* dst = B.data();
  int Ns = 1;
  const int nst = (int)radices.size();
  for (int s = 0; s < nst; ++s) {
    const GStage st = make_stage(L, radices[s], Ns);
    if (s > 0 && stage_pre_needed(st))
      for (int tid = 0; tid < nthr; ++tid) {
        if (dir > 0)
          stage_pre<1>(src, ld, w, st, tw.data(), tws, tid, nthr);
        else
          stage_pre<-1>(src, ld, w, st, tw.data(), tws, tid, nthr);
      }
    const float2* in = s == 0 ? P.data() : src;
    float2* out = s == nst - 1 ? P.data() : (s == 0 ? A.data() : dst);
    const GIo& io = s == 0 ? gin : (s == nst - 1 ? gout : gmid);
    for (int tid = 0; tid < nthr; ++tid) {
      if (dir > 0)
        stage_io<1, true>(in, out, io, w, st, tw.data(), tws, tid, nthr);
      else
        stage_io<-1, true>(in, out, io, w, st, tw.data(), tws, tid, nthr);
    }
    if (s > 0 && s < nst - 1) {
      float2* tmp = src;
      src = dst;
      dst = tmp;
    }
    Ns *= radices[s];
  }
  double err = 0, ref = 0;
  for (int t = 0; t < w; ++t)
    for (int k = 0; k < L; ++k) {
      double sr = 0, si = 0;
      for (int n = 0; n < L; ++n) {
        const double ang =
            -2.0 * M_PI * dir * (double)((long long)n * k % L) / L;
        sr += xr[t * L + n] * cos(ang) - xi[t * L + n] * sin(ang);
        si += xr[t * L + n] * sin(ang) + xi[t * L + n] * cos(ang);
      }
      const float2 g = P[k * w + t];
      err = fmax(err, fabs(g.x - sr) + fabs(g.y - si));
      ref = fmax(ref, fabs(sr) + fabs(si));
    }
  const double rel = err / ref;
  const bool ok = rel < 3e-6;
  printf("panel L=%d w=%d stages=%d dir=%d rel_err=%.3e %s\n", L, w, nst,
         dir, rel, ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

static int check_div() {
  int bad = 0;
  for (int d = 1; d <= 4100; d += (d < 64 ? 1 : 13)) {
    const GDiv q(d);
    for (int a = 0; a < (1 << 22); a += (a < 70000 ? 1 : 9973))
      if (q.div(a) != a / d) {
        if (bad < 5) printf("GDiv(%d).div(%d) = %d\n", d, a, q.div(a));
        ++bad;
      }
    const int a = (1 << 22) - 1;
    if (q.div(a) != a / d) ++bad;
  }
  printf("GDiv %s\n", bad ? "FAIL" : "ok");
  return bad ? 1 : 0;
}

int main() {
  int bad = check_div();
  struct Case {
    int L;
    std::vector<int> r;
    int tws;
  };
  const std::vector<Case> cases = {
      {384, {8, 8, 2, 3}, 2},   {384, {8, 4, 4, 3}, 2}, {356, {4, 89}, 2},
      {64, {8, 8}, 1},          {96, {8, 4, 3}, 1},     {128, {8, 4, 4}, 1},
      {2, {2}, 1},              {4, {4}, 2},            {48, {4, 4, 3}, 1},
      {45, {3, 3, 5}, 1},       {105, {3, 5, 7}, 1},    {143, {11, 13}, 1},
      {264, {8, 3, 11}, 2},     {17, {17}, 1},          {31, {31}, 1},
      {232, {8, 29}, 2},        {712, {8, 89}, 1},      {664, {8, 83}, 2},
      {184, {8, 23}, 1},        {97, {97}, 1},          {200, {8, 5, 5}, 2},
      {1, {}, 1},               {384, {8, 8, 6}, 2},    {768, {8, 8, 12}, 1},
      {480, {8, 12, 5}, 1},     {6, {6}, 1},            {10, {10}, 2},
      {12, {12}, 1},            {14, {14}, 1},          {80, {8, 10}, 2},
      {112, {8, 14}, 1},        {420, {4, 3, 5, 7}, 1}, {420, {10, 6, 7}, 2},
      {336, {8, 6, 7}, 1},
  };
  for (const Case& c : cases)
    for (int dir : {1, -1}) {
      bad += check(c.L, c.r, c.tws, dir, 1, 64);
      bad += check(c.L, c.r, c.tws, dir, 3, 256);
    }
  // The plan builder: every length up to 8192 factors into radices the
  // stages support, whose product is the length, within 12 stages.
  // Without pfa (row plans) no radix is an even composite.
  int badplan = 0;
  for (int L = 1; L <= 8192; ++L)
    for (bool pfa : {true, false}) {
    int r[12];
    const int ns = plan_radices(L, r, 12, pfa);
    int prod = 1;
    bool ok = ns >= 0;
    for (int i = 0; ok && i < ns; ++i) {
      const int x = r[i];
      ok = x == 2 || x == 4 || x == 8 || (x >= 3 && x % 2 == 1) ||
           (pfa && (x == 6 || x == 10 || x == 12 || x == 14));
      prod *= x;
    }
    if (!ok || prod != L) {
      if (badplan < 5) printf("plan_radices(%d) wrong\n", L);
      ++badplan;
    }
  }
  printf("plan_radices %s\n", badplan ? "FAIL" : "ok");
  bad += badplan ? 1 : 0;
  // Transforms with the builder's own plans: row plans for every multiple
  // of 4 in the row lengths of images with sides in [128, 768], column
  // plans for a spread of column lengths.
  for (int L = 4; L <= 384; L += 4) {
    int r[12];
    const int ns = plan_radices(L, r, 12, false);
    bad += check(L, std::vector<int>(r, r + ns), 2, 1, 1, 64);
  }
  for (int L : {768, 760, 744, 720, 696, 672, 600, 504, 462, 390, 210, 150}) {
    int r[12];
    const int ns = plan_radices(L, r, 12, true);
    bad += check(L, std::vector<int>(r, r + ns), 1, -1, 1, 64);
    bad += check(L, std::vector<int>(r, r + ns), 1, 1, 2, 256);
  }
  // Direct panel path of the column kernels (two or more stages).
  for (int L : {768, 536, 616, 376, 24, 40, 712, 1000, 150}) {
    int r[12];
    const int ns = plan_radices(L, r, 12, true);
    if (ns < 2) continue;
    for (int w : {1, 2, 4})
      for (int dir : {1, -1})
        bad += check_panel(L, std::vector<int>(r, r + ns), w, dir, 256);
  }
  return bad;
}
