// LDS path: cols kernels and launchers for double planes.
#include "kern_lds.h"

namespace curig {
CURIG_LDS_COLS_INST(double)
}  // namespace curig
