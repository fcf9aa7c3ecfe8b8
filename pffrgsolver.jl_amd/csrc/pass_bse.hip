// Vertex-pass kernel instantiation for P_BSE (FP64 and FP32 per-element passes, FP32-mixed shared-node pass).
#include "pffrg_pass.h"
#include "pffrg_shared.h"

namespace pffrg {
PFFRG_PASS_INSTANCES(, P_BSE)
PFFRG_SHARED_INSTANCES(, P_BSE)
}  // namespace pffrg
