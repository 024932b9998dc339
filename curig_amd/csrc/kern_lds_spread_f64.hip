// LDS path: spread kernels and launchers for double planes.
#include "kern_lds.h"

namespace curig {
CURIG_LDS_SPREAD_INST(double)
}  // namespace curig
