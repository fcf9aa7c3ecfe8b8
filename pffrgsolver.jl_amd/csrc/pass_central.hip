// Vertex-pass kernel instantiation for P_CENTRAL (FP64 and FP32 per-element passes, FP32-mixed shared-node pass).
#include "pffrg_pass.h"
#include "pffrg_shared.h"

namespace pffrg {
PFFRG_PASS_INSTANCES(, P_CENTRAL)
PFFRG_SHARED_INSTANCES(, P_CENTRAL)
}  // namespace pffrg
