// Compile-time-length fused row kernels, part 0 of kCtParts (rows_ct.h).
#include "rows_ct.h"

namespace scc {
const void* row_ct_kernel_p0(int lr) { return ct_pick<0>(lr); }
}  // namespace scc
