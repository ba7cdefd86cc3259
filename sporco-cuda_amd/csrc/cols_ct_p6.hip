// Compile-time-length column transforms, part 6 of kCtParts (cols_ct.h).
#include "cols_ct.h"

namespace scc {
const void* col_ct_kernel_p6(int r, int dir) { return ct_col_pick<6>(r, dir); }
}  // namespace scc
