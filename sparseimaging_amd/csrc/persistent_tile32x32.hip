// persistent_tile32x32.hip -- the persistent solver's kernels for 32 x 32 gridding
// tiles (4 cells per thread), in a translation unit of their own.
#include "persistent_body.h"

namespace mf {
template const void* persistent_kernel<4>(int, bool, bool);
}  // namespace mf
