// ffx_w2.hip -- instantiates the propagation drivers for 128-bit keys
// (2 words per mask).
#include "ffx_engine.h"

namespace ffx_host {
FFX_DISPATCH_DECL(, 2);
}  // namespace ffx_host
