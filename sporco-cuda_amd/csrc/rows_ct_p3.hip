// Compile-time-length fused row kernels, part 3 of kCtParts (rows_ct.h).
#include "rows_ct.h"

namespace scc {
const void* row_ct_kernel_p3(int lr) { return ct_pick<3>(lr); }
}  // namespace scc
