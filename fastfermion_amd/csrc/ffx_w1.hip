// ffx_w1.hip -- instantiates the propagation drivers for 64-bit keys
// (1 word per mask).
#include "ffx_engine.h"

namespace ffx_host {
FFX_DISPATCH_DECL(, 1);
}  // namespace ffx_host
