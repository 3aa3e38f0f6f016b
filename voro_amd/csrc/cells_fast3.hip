// Cell kernel instantiations for the order-3 fast tier (own translation
// unit so the device code of the tiers compiles in parallel).
#include <hip/hip_runtime.h>

#include "vgpu_cells.h"

namespace vgpu {
VGPU_DEFINE_CELLS_LAUNCHER(launch_cells_fast3, Tier3)
}
