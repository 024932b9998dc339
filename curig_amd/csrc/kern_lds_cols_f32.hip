// LDS path: cols kernels and launchers for float planes.
#include "kern_lds.h"

namespace curig {
CURIG_LDS_COLS_INST(float)
}  // namespace curig
