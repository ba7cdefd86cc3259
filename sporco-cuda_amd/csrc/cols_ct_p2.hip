// Compile-time-length column transforms, part 2 of kCtParts (cols_ct.h).
#include "cols_ct.h"

namespace scc {
const void* col_ct_kernel_p2(int r, int dir) { return ct_col_pick<2>(r, dir); }
}  // namespace scc
