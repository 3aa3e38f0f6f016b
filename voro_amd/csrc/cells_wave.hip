// The wavefront-cell tier: one cell per wavefront, vertex v in lane v
// (wcell_dev.h).  Cells it does not finish go to the per-lane tiers through
// the retry list in `out`.
#include <hip/hip_runtime.h>

#include "batch_dev.h"
#include "vgpu_cells.h"
#include "wave_cell_dev.h"
#include "wcell_dev.h"

namespace vgpu {

namespace {

// Persistent kernel, one wavefront per workgroup, over every cell or over the
// device list of cells whose count is at *list_count, `chunk` cells taken per
// counter increment.  The cells are split into
// one contiguous share per XCD (workgroups go to the 8 XCDs round robin) so
// that the cells in flight on an XCD read neighbouring particle blocks
// through its L2; a wavefront whose share is exhausted moves on to the next.
template<int K, bool POLY, int OCC, bool PRD>
__global__ void __launch_bounds__(64, OCC)
k_wcells(RectView v_, int ncells, const int *__restrict__ list, const int *__restrict__ list_count,
	 const int *__restrict__ cellblk, int *__restrict__ counters, char *__restrict__ ws, unsigned stamp_base,
	 int chunk, CellOut out) {
	typedef const __attribute__((address_space(4))) RectView CView;
	const int lane = threadIdx.x;
	typedef WLayoutT<K> L;
	WCellT<K> c;
	c.setup(ws + (size_t)blockIdx.x * L::bytes, lane, v_.max_len_sq);
	const int total = list_count ? *list_count : ncells;
	const int nshare = 8;
	const int my = blockIdx.x % nshare;
	const int share = (total + nshare - 1) / nshare;
	for (int h = 0; h < nshare; h++) {
		const int sh = (my + h) % nshare;
		const int lo = sh * share, hi = min(total, lo + share);
		for (;;) {
			int base = 0;
			if (lane == 0) base = atomicAdd(&counters[sh], chunk);
			base = __builtin_amdgcn_readfirstlane(base);
			if (lo + base >= hi) break;
			const int end = min(hi, lo + base + chunk);
			// The view (the first kernel argument, at offset 0 of the
			// argument segment) is read through a pointer the compiler
			// cannot hoist out of this loop: its fields are then reloaded
			// where needed instead of being held, and spilled, across the
			// candidate loop (scratch 124 -> 16 bytes per lane).
			int z;
			asm volatile("s_mov_b32 %0, 0" : "=s"(z));
			const RectView &v = *(const RectView *)((CView *)__builtin_amdgcn_kernarg_segment_ptr() + z);
			for (int idx = lo + base; idx < end; idx++) {
				const int cell = list ? list[idx] : idx;
				wave_cell<K, POLY, PRD>(c, v, cell, cellblk[cell], stamp_base + (unsigned)cell + 1u, out);
			}
		}
	}
}

// First tier of a batch of containers over its chunk queue (each entry up to
// 64 consecutive cells of one container): a wavefront takes one entry at a
// time and computes its cells one after another as wavefront cells (up to 64
// vertices).  The queue is split into one contiguous share per XCD, as in
// k_wcells; cells it does not finish go to out's retry list.  OCC waves per
// SIMD (4, 5 or 6).
template<bool POLY, int OCC>
__global__ void __launch_bounds__(64, OCC)
k_batch_wfirst(const RectView *__restrict__ views, const BatchChunk *__restrict__ chunks,
	       const int *__restrict__ nchunks, const int *__restrict__ cellblk, int *__restrict__ counters,
	       char *__restrict__ ws, size_t slab_bytes, const unsigned *__restrict__ stamp_dev, CellOut out) {
	const int lane = threadIdx.x;
	WCell c;
	c.setup(ws + (size_t)blockIdx.x * slab_bytes, lane, 0.0);
	const unsigned stamp_base = *stamp_dev;
	const int total = *nchunks;
	const int nshare = 8;
	const int my = blockIdx.x % nshare;
	const int share = (total + nshare - 1) / nshare;
	for (int h = 0; h < nshare; h++) {
		const int sh = (my + h) % nshare;
		const int lo = sh * share, hi = min(total, lo + share);
		for (;;) {
			int k = 0;
			if (lane == 0) k = atomicAdd(&counters[sh], 1);
			k = __builtin_amdgcn_readfirstlane(k);
			if (lo + k >= hi) break;
			const int cont = __builtin_amdgcn_readfirstlane(chunks[lo + k].cont);
			const int first = __builtin_amdgcn_readfirstlane(chunks[lo + k].first);
			const int count = __builtin_amdgcn_readfirstlane(chunks[lo + k].count);
			// The view is read through the constant address space: its fields
			// are then reloaded where needed instead of being held (and
			// spilled) across the cell loop (measured: 16 instead of 472 bytes
			// of scratch per lane).
			typedef const __attribute__((address_space(4))) RectView CView;
			const RectView &v = *(const RectView *)((CView *)views + cont);
			c.set_tol(v.max_len_sq);
			for (int t = 0; t < count; t++) {
				const int cell = first + t;
				wave_cell<1, POLY, false>(c, v, cell, cellblk[cell] - v.boff, stamp_base + (unsigned)cell + 1u,
							  out);
			}
		}
	}
}

}  // namespace

#define VGPU_WAVE_LAUNCH(K, P, O, R)                                                                            \
	hipLaunchKernelGGL((k_wcells<K, P, O, R>), dim3(waves), dim3(64), 0, s, v, ncells, list, list_count, cellblk, \
			   counters, ws, stamp_base, chunk, out)

int launch_cells_wave(int k, bool poly, bool prd, int occ, int waves, hipStream_t s, const RectView &v, int ncells,
		      const int *list, const int *list_count, const int *cellblk, int *counters, char *ws,
		      unsigned stamp_base, int chunk, const CellOut &out) {
	if (k == 4) {
		// the second overflow tier: few cells, four vertex sets
		if (prd) {
			if (poly) VGPU_WAVE_LAUNCH(4, true, 2, true);
			else VGPU_WAVE_LAUNCH(4, false, 2, true);
		} else {
			if (poly) VGPU_WAVE_LAUNCH(4, true, 2, false);
			else VGPU_WAVE_LAUNCH(4, false, 2, false);
		}
	} else if (k == 2) {
		// the overflow tier: few cells, registers for two vertex sets
		if (prd) {
			if (poly) VGPU_WAVE_LAUNCH(2, true, VG_K2_OCC, true);
			else VGPU_WAVE_LAUNCH(2, false, VG_K2_OCC, true);
		} else {
			if (poly) VGPU_WAVE_LAUNCH(2, true, VG_K2_OCC, false);
			else VGPU_WAVE_LAUNCH(2, false, VG_K2_OCC, false);
		}
	} else if (prd) {
		// periodic containers: 4, 5 or 6 waves per SIMD
		if (occ == 5) {
			if (poly) VGPU_WAVE_LAUNCH(1, true, 5, true);
			else VGPU_WAVE_LAUNCH(1, false, 5, true);
		} else if (occ >= 6) {
			if (poly) VGPU_WAVE_LAUNCH(1, true, 6, true);
			else VGPU_WAVE_LAUNCH(1, false, 6, true);
		} else {
			if (poly) VGPU_WAVE_LAUNCH(1, true, 4, true);
			else VGPU_WAVE_LAUNCH(1, false, 4, true);
		}
	} else if (occ >= 8) {
		if (poly) VGPU_WAVE_LAUNCH(1, true, 8, false);
		else VGPU_WAVE_LAUNCH(1, false, 8, false);
	} else if (occ == 6) {
		if (poly) VGPU_WAVE_LAUNCH(1, true, 6, false);
		else VGPU_WAVE_LAUNCH(1, false, 6, false);
	} else if (occ == 5) {
		if (poly) VGPU_WAVE_LAUNCH(1, true, 5, false);
		else VGPU_WAVE_LAUNCH(1, false, 5, false);
	} else if (occ <= 2) {
		if (poly) VGPU_WAVE_LAUNCH(1, true, 2, false);
		else VGPU_WAVE_LAUNCH(1, false, 2, false);
	} else if (occ == 3) {
		if (poly) VGPU_WAVE_LAUNCH(1, true, 3, false);
		else VGPU_WAVE_LAUNCH(1, false, 3, false);
	} else {
		if (poly) VGPU_WAVE_LAUNCH(1, true, 4, false);
		else VGPU_WAVE_LAUNCH(1, false, 4, false);
	}
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

#define VGPU_BATCH_WAVE_LAUNCH(P, O)                                                                       \
	hipLaunchKernelGGL((k_batch_wfirst<P, O>), dim3(waves), dim3(64), 0, s, views, chunks, nchunks, cellblk, \
			   counters, ws, slab, stamp_dev, out)

int launch_batch_wfirst(bool poly, int occ, int waves, hipStream_t s, const RectView *views, const BatchChunk *chunks,
			const int *nchunks, const int *cellblk, int *counters, char *ws, size_t slab,
			const unsigned *stamp_dev, const CellOut &out) {
	if (occ >= 6) {
		if (poly) VGPU_BATCH_WAVE_LAUNCH(true, 6);
		else VGPU_BATCH_WAVE_LAUNCH(false, 6);
	} else if (occ == 5) {
		if (poly) VGPU_BATCH_WAVE_LAUNCH(true, 5);
		else VGPU_BATCH_WAVE_LAUNCH(false, 5);
	} else {
		if (poly) VGPU_BATCH_WAVE_LAUNCH(true, 4);
		else VGPU_BATCH_WAVE_LAUNCH(false, 4);
	}
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

}  // namespace vgpu
