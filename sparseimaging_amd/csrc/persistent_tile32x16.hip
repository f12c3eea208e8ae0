// persistent_tile32x16.hip -- the persistent solver's kernels for 32 x 16 gridding
// tiles (2 cells per thread), in a translation unit of their own.
#include "persistent_body.h"

namespace mf {
template const void* persistent_kernel<2>(int, bool, bool);
}  // namespace mf
