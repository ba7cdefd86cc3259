// Compile-time-length fused row kernels, part 4 of kCtParts (rows_ct.h).
#include "rows_ct.h"

namespace scc {
const void* row_ct_kernel_p4(int lr) { return ct_pick<4>(lr); }
}  // namespace scc
