// Compile-time-length fused row kernels, part 6 of kCtParts (rows_ct.h).
#include "rows_ct.h"

namespace scc {
const void* row_ct_kernel_p6(int lr) { return ct_pick<6>(lr); }
}  // namespace scc
