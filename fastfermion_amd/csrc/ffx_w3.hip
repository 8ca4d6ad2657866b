// ffx_w3.hip -- instantiates the propagation drivers for 192-bit keys
// (3 words per mask).
#include "ffx_engine.h"

namespace ffx_host {
FFX_DISPATCH_DECL(, 3);
}  // namespace ffx_host
