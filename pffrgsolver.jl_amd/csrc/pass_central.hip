// Vertex-pass kernel instantiation for P_CENTRAL (FP64 and FP32-mixed).
#include "pffrg_pass.h"

namespace pffrg {
PFFRG_PASS_INSTANCES(, P_CENTRAL)
}  // namespace pffrg
