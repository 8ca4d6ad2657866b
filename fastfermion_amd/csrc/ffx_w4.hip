// ffx_w4.hip -- instantiates the propagation drivers for 256-bit keys
// (4 words per mask).
#include "ffx_engine.h"

namespace ffx_host {
FFX_DISPATCH_DECL(, 4);
}  // namespace ffx_host
