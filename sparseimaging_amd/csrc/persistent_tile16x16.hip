// persistent_tile16x16.hip -- the persistent solver's kernels for 16 x 16 gridding
// tiles (1 cell per thread), in a translation unit of their own.
#include "persistent_body.h"

namespace mf {
template const void* persistent_kernel<1>(int, bool, bool);
}  // namespace mf
