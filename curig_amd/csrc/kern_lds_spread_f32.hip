// LDS path: spread kernels and launchers for float planes.
#include "kern_lds.h"

namespace curig {
CURIG_LDS_SPREAD_INST(float)
}  // namespace curig
