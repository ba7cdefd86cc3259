// Compile-time-length column transforms, part 0 of kCtParts (cols_ct.h).
#include "cols_ct.h"

namespace scc {
const void* col_ct_kernel_p0(int r, int dir) { return ct_col_pick<0>(r, dir); }
}  // namespace scc
