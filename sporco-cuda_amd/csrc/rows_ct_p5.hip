// Compile-time-length fused row kernels, part 5 of kCtParts (rows_ct.h).
#include "rows_ct.h"

namespace scc {
const void* row_ct_kernel_p5(int lr) { return ct_pick<5>(lr); }
}  // namespace scc
