// ffx_inst.hip -- explicit instantiation of the propagation drivers for one key
// width (FFX_INST_W = 1..4 words per mask), one coefficient arithmetic
// (FFX_INST_P: 0 FP64, 1 FP32, 2 FP32 compensated) and one algebra (FFX_INST_ALG:
// 0 Pauli, 1 Majorana; the FP64 objects also hold the contribution logs).
// fastfermion_amd/build.py compiles this file once per combination, so the
// kernels of the 24 combinations compile in parallel.
#include "ffx_engine.h"

#if !defined(FFX_INST_W) || !defined(FFX_INST_P) || !defined(FFX_INST_ALG)
#error "compile with -DFFX_INST_W=<1..4> -DFFX_INST_P=<0..2> -DFFX_INST_ALG=<0|1>"
#endif

#if FFX_INST_P == 0
#define FFX_INST_CP F64
#elif FFX_INST_P == 1
#define FFX_INST_CP F32
#else
#define FFX_INST_CP F32C
#endif

namespace ffx_host {
#if FFX_INST_ALG == 0
FFX_PAULI_INST(, FFX_INST_W, FFX_INST_CP);
#if FFX_INST_P == 0
FFX_LOG_INST(, PauliAlg<FFX_INST_W>);
#endif
#else
FFX_MAJORANA_INST(, FFX_INST_W, FFX_INST_CP);
#if FFX_INST_P == 0
FFX_LOG_INST(, MajoranaAlg<FFX_INST_W>);
#endif
#endif
}  // namespace ffx_host
