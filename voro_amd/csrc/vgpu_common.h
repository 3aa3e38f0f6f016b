// Shared definitions for the HIP engine: qualifiers, tier capacities and the
// lane-interleaved per-cell workspace.
//
// One Voronoi cell is built by one lane.  Its vertex/edge/neighbour tables
// live in a workspace slab owned by the lane's wavefront: every array element
// i of lane l sits at  slab + field_offset + (i*64 + l)*sizeof(T),  so when
// the 64 lanes of a wavefront touch the same index (vertex 0 at the start of
// every plane cut, the 8 box corners, the face walk) the access is one
// contiguous 64-element line per field.
#pragma once
#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VG_INL __host__ __device__ __forceinline__
#define VG_FN __host__ __device__ __forceinline__
#else
#define VG_INL inline
#define VG_FN inline
#endif

namespace vgpu {

// Tier capacities.  A cell that outgrows any of them sets its overflow flag
// and is recomputed from scratch by the next tier; results never depend on
// the tier that produced them.
struct TierFast {
	enum {
		MAXV = 128,    // vertices
		MAXO = 4,      // edges per vertex row (vertex order)
		MAXS = 128,    // delete / auxiliary / extra-search / marginal stacks
		MAXP = 32,     // order-1 and order-2 vertex pools
		W = 7,         // half-width of the block-mask window (blocks)
		MAXQ = 512,    // block queue (ring)
	};
};
struct TierWide {
	enum { MAXV = 1024, MAXO = 16, MAXS = 1024, MAXP = 256, W = 31, MAXQ = 8192 };
};
struct TierHuge {
	enum { MAXV = 4096, MAXO = 32, MAXS = 4096, MAXP = 1024, W = 63, MAXQ = 32768 };
};

// Byte layout of one lane-interleaved workspace slab for a tier.  ST is the
// interleave (64 on the GPU: one slab per wavefront; 1 in CPU builds).
template<class T_, int ST>
struct Layout {
	enum { MAXV = T_::MAXV, MAXO = T_::MAXO, MAXS = T_::MAXS, MAXP = T_::MAXP, W = T_::W, MAXQ = T_::MAXQ };
	enum { WD = 2 * W + 1, WBITS = WD * WD * WD, WWORDS = (WBITS + 31) / 32 };
	static constexpr size_t al(size_t x) { return (x + 255) & ~size_t(255); }
	static constexpr size_t o_x = 0;
	static constexpr size_t o_y = o_x + al((size_t)MAXV * ST * 8);
	static constexpr size_t o_z = o_y + al((size_t)MAXV * ST * 8);
	static constexpr size_t o_a = o_z + al((size_t)MAXV * ST * 8);
	static constexpr size_t o_mask = o_a + al((size_t)MAXV * ST * 8);
	static constexpr size_t o_ne = o_mask + al((size_t)MAXV * ST * 4);
	static constexpr size_t o_ed = o_ne + al((size_t)MAXV * MAXO * ST * 4);
	static constexpr size_t o_bp = o_ed + al((size_t)MAXV * MAXO * ST * 2);
	static constexpr size_t o_slot = o_bp + al((size_t)MAXV * ST * 2);
	static constexpr size_t o_ds = o_slot + al((size_t)MAXV * ST * 2);
	static constexpr size_t o_ds2 = o_ds + al((size_t)MAXS * ST * 2);
	static constexpr size_t o_xse = o_ds2 + al((size_t)MAXS * ST * 2);
	static constexpr size_t o_dm = o_xse + al((size_t)MAXS * ST * 2);
	static constexpr size_t o_pool1 = o_dm + al((size_t)MAXS * ST * 2);
	static constexpr size_t o_pool2 = o_pool1 + al((size_t)MAXP * ST * 2);
	static constexpr size_t o_rel = o_pool2 + al((size_t)MAXP * ST * 2);
	static constexpr size_t o_nu = o_rel + al((size_t)MAXV * MAXO * ST * 1);
	static constexpr size_t o_bm = o_nu + al((size_t)MAXV * ST * 1);
	static constexpr size_t o_bms = o_bm + al((size_t)WWORDS * ST * 4);
	static constexpr size_t o_qu = o_bms + al((size_t)WWORDS * ST * 4);
	static constexpr size_t bytes = o_qu + al((size_t)MAXQ * ST * 4);
	static constexpr size_t o_nbs = o_qu;   // output staging (reuses the ring)
	enum { NBCAP = MAXV };
};

// One vertex of the unit Voronoi cell of a periodic container, the start
// cell of every particle's cell there: doubled coordinates (as the cells
// store them), order, edges and edge relations.  64 bytes.
struct UnitVert {
	double x, y, z;
	int nu;
	short ed[8];
	signed char rel[8];
	int pad[3];
};

// Per-cell status bits.
enum : int {
	ST_EXISTS = 1,      // compute_cell returned true
	ST_OVERFLOW = 2,    // a tier capacity was exceeded: recompute in the next tier
	ST_FATAL = 4,       // a state the reference treats as a fatal error
};

}  // namespace vgpu
