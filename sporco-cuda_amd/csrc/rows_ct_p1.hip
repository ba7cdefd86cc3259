// Compile-time-length fused row kernels, part 1 of kCtParts (rows_ct.h).
#include "rows_ct.h"

namespace scc {
const void* row_ct_kernel_p1(int lr) { return ct_pick<1>(lr); }
}  // namespace scc
