// LDS path: interp kernels and launchers for double planes.
#include "kern_lds.h"

namespace curig {
CURIG_LDS_INTERP_INST(double)
}  // namespace curig
