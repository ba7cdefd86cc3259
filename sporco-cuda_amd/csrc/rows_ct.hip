// Dispatch of the compile-time-length fused row kernels (rows_ct.h) and
// column transforms (cols_ct.h).
#include "cols_ct.h"
#include "rows_ct.h"

namespace scc {

const void* row_ct_kernel_p0(int lr);
const void* row_ct_kernel_p1(int lr);
const void* row_ct_kernel_p2(int lr);
const void* row_ct_kernel_p3(int lr);
const void* row_ct_kernel_p4(int lr);
const void* row_ct_kernel_p5(int lr);
const void* row_ct_kernel_p6(int lr);
const void* row_ct_kernel_p7(int lr);
const void* col_ct_kernel_p0(int r, int dir);
const void* col_ct_kernel_p1(int r, int dir);
const void* col_ct_kernel_p2(int r, int dir);
const void* col_ct_kernel_p3(int r, int dir);
const void* col_ct_kernel_p4(int r, int dir);
const void* col_ct_kernel_p5(int r, int dir);
const void* col_ct_kernel_p6(int r, int dir);
const void* col_ct_kernel_p7(int r, int dir);

namespace {

const void* row_ct_kernel(int lr) {
  static_assert(kCtParts == 8, "one dispatch entry per part");
  if (lr % 4 != 0 || lr < 4 * kCtQMin || lr > 4 * kCtQMax) return nullptr;
  switch ((lr / 4 - kCtQMin) % kCtParts) {
    case 0: return row_ct_kernel_p0(lr);
    case 1: return row_ct_kernel_p1(lr);
    case 2: return row_ct_kernel_p2(lr);
    case 3: return row_ct_kernel_p3(lr);
    case 4: return row_ct_kernel_p4(lr);
    case 5: return row_ct_kernel_p5(lr);
    case 6: return row_ct_kernel_p6(lr);
    default: return row_ct_kernel_p7(lr);
  }
}

const void* col_ct_kernel(int r, int dir) {
  if (r % 8 != 0 || r < 8 * kCtColQMin || r > 8 * kCtColQMax) return nullptr;
  switch ((r / 8 - kCtColQMin) % kCtParts) {
    case 0: return col_ct_kernel_p0(r, dir);
    case 1: return col_ct_kernel_p1(r, dir);
    case 2: return col_ct_kernel_p2(r, dir);
    case 3: return col_ct_kernel_p3(r, dir);
    case 4: return col_ct_kernel_p4(r, dir);
    case 5: return col_ct_kernel_p5(r, dir);
    case 6: return col_ct_kernel_p6(r, dir);
    default: return col_ct_kernel_p7(r, dir);
  }
}

thread_local bool g_row_ct = true;

}  // namespace

void set_row_ct(bool on) { g_row_ct = on; }
bool row_ct_enabled() { return g_row_ct; }

bool row_ct_supported(int LR) { return row_ct_kernel(LR) != nullptr; }

bool col_ct_supported(int R) { return col_ct_kernel(R, 1) != nullptr; }

hipError_t launch_col_ct(const Geom& g, int nmaps, float2* F,
                         const float2* twR, int inverse, const Ctrl* ctrl,
                         const float* parts, int nparts, double* bsum,
                         int nred, size_t shm, hipStream_t st) {
  const void* k = col_ct_kernel(g.R, inverse ? -1 : 1);
  if (k == nullptr) return hipErrorInvalidValue;
  Geom ga = g;
  float2* Fa = F;
  const float2* ta = twR;
  const Ctrl* ca = ctrl;
  const float* pa = parts;
  int npa = nparts;
  double* ba = bsum;
  int nra = nred;
  void* args[] = {&ga, &Fa, &ta, &ca, &pa, &npa, &ba, &nra};
  return hipLaunchKernel(k, dim3(g.P, nmaps), dim3(256), args, shm, st);
}

hipError_t launch_row_ct(const Geom& g, const Buffers& b, const Params& p,
                         int nr, int nb, size_t shm, hipStream_t st) {
  const void* k = row_ct_kernel(g.Lr);
  if (k == nullptr) return hipErrorInvalidValue;
  Geom ga = g;
  Buffers ba = b;
  Params pa = p;
  int nra = nr;
  void* args[] = {&ga, &ba, &pa, &nra};
  return hipLaunchKernel(k, dim3(nb), dim3(256), args, shm, st);
}

}  // namespace scc
