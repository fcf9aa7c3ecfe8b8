// Vertex-pass kernel instantiation for P_BSE (FP64 and FP32-mixed).
#include "pffrg_pass.h"

namespace pffrg {
PFFRG_PASS_INSTANCES(, P_BSE)
}  // namespace pffrg
