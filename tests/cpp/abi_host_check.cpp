// Host paths of the C ABI under AddressSanitizer (no GPU needed): option
// defaults, argument validation before any device call, the size check of
// the allocation-free entry point, the batched entry point's rejections,
// the no-device status, error strings and cache release.  Built by
// tests/test_abi_asan.py with the library's host code instrumented
// (hipcc -Xarch_host -fsanitize=address); prints FAIL lines and returns 1
// on a failed check, and ASan aborts on any invalid host memory access.
#include <cstdio>
#include <cstring>
#include <vector>

#include "sporco_amd.h"

static int g_fail = 0;
#define CHECK(cond)                                                  \
  do {                                                               \
    if (!(cond)) {                                                   \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);     \
      ++g_fail;                                                      \
    }                                                                \
  } while (0)

static scc_algopt base_opt(int R, int C, int M, int K, const float* l1) {
  scc_algopt o;
  scc_clear_opts(&o);
  o.IMG_ROW_SIZE = R;
  o.IMG_COL_SIZE = C;
  o.DICT_M_SIZE = M;
  o.DICT_ROW_SIZE = K;
  o.DICT_COL_SIZE = K;
  o.rho = 1.f;
  o.MaxMainIter = 5;
  o.L1Weight = l1;
  return o;
}

int main() {
  // Sentinels and defaults (utils.cu:12-98).
  scc_algopt o;
  scc_clear_opts(&o);
  CHECK(o.rho == -1 && o.MaxMainIter == -1 && o.device == -1);
  scc_default_opts(&o);
  CHECK(o.device == 0 && o.AutoRho == 1 && o.MaxMainIter == 250);
  CHECK(o.RelaxParam > 1.79f && o.RelaxParam < 1.81f);
  scc_clear_opts(nullptr);
  scc_default_opts(nullptr);

  const int R = 8, C = 7, M = 2, K = 3;
  std::vector<float> D(M * K * K, 1.f), S(R * C, 1.f), Y(M * R * C, 0.f);
  float one = 1.f;

  // Invalid input: a status and a message, before any device call.
  CHECK(scc_cbpdn(D.data(), S.data(), 0.1f, nullptr, Y.data()) ==
        SCC_ERR_INVALID);
  CHECK(std::strlen(scc_last_error()) > 0);
  scc_algopt b = base_opt(R, C, M, K, &one);
  CHECK(scc_cbpdn(nullptr, S.data(), 0.1f, &b, Y.data()) == SCC_ERR_INVALID);
  b.IMG_ROW_SIZE = 0;
  CHECK(scc_cbpdn(D.data(), S.data(), 0.1f, &b, Y.data()) == SCC_ERR_INVALID);
  b = base_opt(R, C, M, K, &one);
  b.DICT_ROW_SIZE = R + 1;
  CHECK(scc_cbpdn(D.data(), S.data(), 0.1f, &b, Y.data()) == SCC_ERR_INVALID);
  b = base_opt(R, C, M, K, &one);
  b.L1_WEIGHT_M_SIZE = 3;
  CHECK(scc_cbpdn(D.data(), S.data(), 0.1f, &b, Y.data()) == SCC_ERR_INVALID);
  b = base_opt(R, C, M, K, nullptr);
  CHECK(scc_cbpdn(D.data(), S.data(), 0.1f, &b, Y.data()) == SCC_ERR_INVALID);
  b = base_opt(R, C, M, K, &one);
  b.rho = 0.f;
  CHECK(scc_cbpdn_grd(D.data(), S.data(), 0.1f, 0.1f, &b, Y.data()) ==
        SCC_ERR_INVALID);
  b = base_opt(R, C, M, K, &one);
  b.nWeight = 5;
  CHECK(scc_cbpdn(D.data(), S.data(), 0.1f, &b, Y.data()) == SCC_ERR_INVALID);
  b = base_opt(5000, C, M, K, &one);
  CHECK(scc_cbpdn(D.data(), S.data(), 0.1f, &b, Y.data()) ==
        SCC_ERR_UNSUPPORTED);

  // Allocation-free entry point: a problem beyond the scratch array is
  // rejected on the host, and Verbose is not available.
  b = base_opt(512, 512, 64, 8, &one);
  CHECK(scc_cbpdn_device_noalloc(D.data(), S.data(), 0.1f, 0.f, 0, &b,
                                 Y.data(), nullptr) == SCC_ERR_UNSUPPORTED);
  CHECK(std::strstr(scc_last_error(), "scratch") != nullptr);
  b = base_opt(R, C, M, K, &one);
  b.Verbose = 1;
  CHECK(scc_cbpdn_device_noalloc(D.data(), S.data(), 0.1f, 0.f, 0, &b,
                                 Y.data(), nullptr) == SCC_ERR_UNSUPPORTED);

  // Batched entry point: masks and bad image sizes are rejected.
  scc_batch_image im[2];
  std::memset(im, 0, sizeof(im));
  im[0].R = R;
  im[0].C = C;
  im[0].S = S.data();
  im[0].Y = Y.data();
  im[0].lambda = 0.1f;
  im[1] = im[0];
  im[1].R = 0;
  scc_batch_result res[2];
  b = base_opt(R, C, M, K, &one);
  CHECK(scc_batched_cbpdn(D.data(), M, K, K, 2, im, &b, res, nullptr) ==
        SCC_ERR_INVALID);
  int mask[R * C];
  for (int i = 0; i < R * C; ++i) mask[i] = 1;
  b.Weight = mask;
  b.nWeight = R * C;
  CHECK(scc_batched_cbpdn(D.data(), M, K, K, 1, im, &b, res, nullptr) !=
        SCC_OK);

  // A valid call: solved when a device is present, else a status.
  b = base_opt(R, C, M, K, &one);
  const int n = scc_get_device_count();
  const int rc = scc_cbpdn(D.data(), S.data(), 0.1f, &b, Y.data());
  CHECK(n > 0 ? rc == SCC_OK : rc == SCC_ERR_NODEVICE);
  scc_info info;
  CHECK(scc_last_info(&info) == SCC_OK);
  CHECK(scc_last_info(nullptr) == SCC_ERR_INVALID);
  CHECK(scc_get_device_name(-1) == nullptr);
  size_t fb = 0, tb = 0;
  CHECK(scc_get_memory_info(nullptr, &tb) == SCC_ERR_INVALID);
  if (n == 0) CHECK(scc_get_memory_info(&fb, &tb) == SCC_ERR_NODEVICE);
  CHECK(scc_last_history(nullptr, 0) >= 0);
  CHECK(scc_last_objective(nullptr, 0) >= 0);
  scc_release_cache();
  std::printf("%s (%d failed)\n", g_fail ? "FAIL" : "ok", g_fail);
  return g_fail ? 1 : 0;
}
