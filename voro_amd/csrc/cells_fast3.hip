// Cell kernel instantiations for the order-3 fast tier, at 2, 3 and 4
// waves per SIMD (own translation unit so the tiers compile in parallel).
#include <hip/hip_runtime.h>

#include "vgpu_cells.h"

namespace vgpu {
int launch_cells_fast3(bool poly, int occ, int waves, hipStream_t s, const RectView &v, int ncells, const int *list,
		       const int *list_count, const int *cellblk, int *counters, char *ws, size_t slab,
		       unsigned stamp_base, int per_wave, const CellOut &out) {
	if (poly) {
		if (occ >= 4) VGPU_CELLS_LAUNCH(Tier3, true, 4);
		else if (occ == 3) VGPU_CELLS_LAUNCH(Tier3, true, 3);
		else VGPU_CELLS_LAUNCH(Tier3, true, 2);
	} else {
		if (occ >= 4) VGPU_CELLS_LAUNCH(Tier3, false, 4);
		else if (occ == 3) VGPU_CELLS_LAUNCH(Tier3, false, 3);
		else VGPU_CELLS_LAUNCH(Tier3, false, 2);
	}
	return hipGetLastError() == hipSuccess ? 0 : 6;
}
}
