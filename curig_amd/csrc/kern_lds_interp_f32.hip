// LDS path: interp kernels and launchers for float planes.
#include "kern_lds.h"

namespace curig {
CURIG_LDS_INTERP_INST(float)
}  // namespace curig
