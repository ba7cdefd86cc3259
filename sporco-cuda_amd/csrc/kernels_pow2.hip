// Power-of-two kernels: not yet enabled; every size takes the generic path.
#include "scc_pow2.h"

namespace scc {

bool pow2_select(int, int, int, Pow2Config*) { return false; }
int pow2_row_blocks(const Geom&, const Pow2Config&) { return 0; }
hipError_t pow2_row_fwd(const Geom&, const Pow2Config&, int, int,
                        const float*, float*, float*, float2*, const float2*,
                        const Ctrl*, hipStream_t) {
  return hipErrorNotSupported;
}
hipError_t pow2_col(const Geom&, const Pow2Config&, int, float2*,
                    const float2*, int, const Ctrl*, hipStream_t) {
  return hipErrorNotSupported;
}
hipError_t pow2_solve(const Geom&, const Pow2Config&, const Buffers&,
                      const Params&, hipStream_t) {
  return hipErrorNotSupported;
}
hipError_t pow2_row(const Geom&, const Pow2Config&, const Buffers&,
                    const Params&, int, int*, hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace scc
