// Compile-time-length fused row kernels, part 2 of kCtParts (rows_ct.h).
#include "rows_ct.h"

namespace scc {
const void* row_ct_kernel_p2(int lr) { return ct_pick<2>(lr); }
}  // namespace scc
