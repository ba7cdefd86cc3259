// Compile-time-length fused row kernels, part 7 of kCtParts (rows_ct.h).
#include "rows_ct.h"

namespace scc {
const void* row_ct_kernel_p7(int lr) { return ct_pick<7>(lr); }
}  // namespace scc
