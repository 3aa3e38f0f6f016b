// The wavefront-cell tier: one cell per wavefront, vertex v in lane v
// (wcell_dev.h).  Cells it does not finish go to the per-lane tiers through
// the retry list in `out`.
#include <hip/hip_runtime.h>

#include "batch_dev.h"
#include "vgpu_cells.h"
#include "wcell_dev.h"

namespace vgpu {

namespace {

// Neighbour ids in face-walk order: id i in lane i & 63 of set i >> 6.
template<int K>
struct WSink {
	WCellT<K> *c;
	VG_INL void put(int i, int v) { if (i < (int)WLayoutT<K>::NBCAP) c->put_nb(i, v); }
};

// Copies a finished wavefront cell's tables into the geometry arena: each
// lane writes the vertices it holds (order 3: three edges and back pointers
// from the topology word); a cell that does not fit is counted in top[2].
template<int K>
__device__ __forceinline__ void wave_geo_write(WCellT<K> &c, const GeoSink &g, int cell, bool ok) {
	const int lane = threadIdx.x;
	if (!ok) {
		if (lane == 0) { g.vstart[cell] = 0; g.np[cell] = 0; g.esz[cell] = 0; }
		return;
	}
	const int p = c.p;
	int vs = 0, es = 0;
	if (lane == 0) { vs = atomicAdd(&g.top[0], p); es = atomicAdd(&g.top[1], 3 * p); }
	vs = __builtin_amdgcn_readfirstlane(vs);
	es = __builtin_amdgcn_readfirstlane(es);
	if ((long long)vs + p > g.vcap || (long long)es + 3 * p > g.ecap) {
		if (lane == 0) { g.vstart[cell] = -1; g.np[cell] = 0; g.esz[cell] = 0; atomicAdd(&g.top[2], 1); }
		return;
	}
	if (lane == 0) { g.vstart[cell] = vs; g.np[cell] = p; g.estart[cell] = es; g.esz[cell] = 3 * p; }
	eachk<K>([&](auto hc) __attribute__((always_inline)) {
		constexpr int h = decltype(hc)::value;
		const int v = 64 * h + lane;
		if (v < p) {
			const size_t a = (size_t)(vs + v);
			g.xyz[3 * a] = c.vx[h]; g.xyz[3 * a + 1] = c.vy[h]; g.xyz[3 * a + 2] = c.vz[h];
			g.nu[a] = 3;
			g.eoff[a] = es + 3 * v;
			const unsigned w = c.vt[h];
			for (int i = 0; i < 3; i++) {
				g.ed[es + 3 * v + i] = (short)WCellT<K>::twE(w, i);
				g.rel[es + 3 * v + i] = (unsigned char)WCellT<K>::twR(w, i);
			}
		}
	});
}

// One cell of the container with view v, block b of the view's numbering:
// compute_cell, then the outputs (lane 0 writes the scalars, every lane its
// neighbour ids) or, if the cell overflowed, an entry in out's retry list.
template<int K, bool POLY, bool PRD>
__device__ __forceinline__ void wave_cell(WCellT<K> &c, const RectView &v, int cell, int b, unsigned stamp,
					  const CellOut &out) {
	typedef WLayoutT<K> L;
	const int lane = threadIdx.x;
	const int s = cell - v.cstart[b];
	const int ci = b % v.nx, cj = (b / v.nx) % v.gy, ck = b / v.nxy;
	CellStats stt;
	const bool ok = compute_cell_rect<WCellT<K>, POLY, PRD>(c, v, b, s, ci, cj, ck, stamp, stt,
							  (unsigned long long)cell << 24);
	double vol = 0, cx = 0, cy = 0, cz = 0;
	int nf = 0;
	if (!c.ovf && ok) {
		WSink<K> sk{&c};
		c.outputs(vol, cx, cy, cz, nf, sk);
		if (nf > (int)L::NBCAP) c.ovf = true;
	}
	if (c.ovf) {
		if (lane == 0) {
			const int slot_ = atomicAdd(out.retry_count, 1);
			out.retry[slot_] = cell;
			out.status[cell] = ST_OVERFLOW;
		}
		return;
	}
	if (lane == 0) {
		out.vol[cell] = vol;
		out.cent[3 * cell] = cx; out.cent[3 * cell + 1] = cy; out.cent[3 * cell + 2] = cz;
		out.nfaces[cell] = nf;
		out.status[cell] = ok ? ST_EXISTS : 0;
		if (out.st_planes) {
			out.st_planes[cell] = (short)min(stt.n_planes, 32767);
			out.st_exit[cell] = (signed char)stt.wl_exit;
			out.st_tier[cell] = (signed char)out.tier;
		}
		if (PRD && out.touch) out.touch[cell] = (unsigned char)stt.img;
	}
	if (out.geo.top) wave_geo_write(c, out.geo, cell, ok);
	if (out.want_nbrs) {
		int off = 0;
		if (lane == 0 && nf > 0) off = atomicAdd(out.arena_top, nf);
		off = __builtin_amdgcn_readfirstlane(off);
		if (off + nf > out.arena_cap) {
			if (lane == 0) { out.nbr_start[cell] = -1; atomicAdd(out.fail_count, 1); }
		} else {
			if (lane == 0) out.nbr_start[cell] = off;
			eachk<K>([&](auto hc) __attribute__((always_inline)) {
				constexpr int h = decltype(hc)::value;
				if (64 * h + lane < nf) out.nbr[off + 64 * h + lane] = c.nb_lane(h);
			});
		}
	}
}

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
// k_wcells; cells it does not finish go to out's retry list.
template<bool POLY>
__global__ void __launch_bounds__(64, 4)
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

// Wavefront cells of a batch over the device list of cells whose count is at
// *list_count (the learned tier's leftovers), each cell with its container's
// view (bcont: the container of every block), one cell per counter
// increment, shared out per XCD as in k_wcells.
template<bool POLY>
__global__ void __launch_bounds__(64, 4)
k_batch_wlist(const RectView *__restrict__ views, const int *__restrict__ bcont, const int *__restrict__ list,
	      const int *__restrict__ list_count, const int *__restrict__ cellblk, int *__restrict__ counters,
	      char *__restrict__ ws, size_t slab_bytes, const unsigned *__restrict__ stamp_dev, CellOut out) {
	const int lane = threadIdx.x;
	WCell c;
	c.setup(ws + (size_t)blockIdx.x * slab_bytes, lane, 0.0);
	const unsigned stamp_base = *stamp_dev;
	const int total = *list_count;
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
			const int cell = __builtin_amdgcn_readfirstlane(list[lo + k]);
			const int gb = __builtin_amdgcn_readfirstlane(cellblk[cell]);
			const int cont = __builtin_amdgcn_readfirstlane(bcont[gb]);
			typedef const __attribute__((address_space(4))) RectView CView;
			const RectView &v = *(const RectView *)((CView *)views + cont);
			c.set_tol(v.max_len_sq);
			wave_cell<1, POLY, false>(c, v, cell, gb - v.boff, stamp_base + (unsigned)cell + 1u, out);
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
			if (poly) VGPU_WAVE_LAUNCH(2, true, 2, true);
			else VGPU_WAVE_LAUNCH(2, false, 2, true);
		} else {
			if (poly) VGPU_WAVE_LAUNCH(2, true, 2, false);
			else VGPU_WAVE_LAUNCH(2, false, 2, false);
		}
	} else if (prd) {
		if (poly) VGPU_WAVE_LAUNCH(1, true, 4, true);
		else VGPU_WAVE_LAUNCH(1, false, 4, true);
	} else if (occ >= 8) {
		if (poly) VGPU_WAVE_LAUNCH(1, true, 8, false);
		else VGPU_WAVE_LAUNCH(1, false, 8, false);
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

int launch_batch_wfirst(bool poly, int waves, hipStream_t s, const RectView *views, const BatchChunk *chunks,
			const int *nchunks, const int *cellblk, int *counters, char *ws, size_t slab,
			const unsigned *stamp_dev, const CellOut &out) {
	if (poly)
		hipLaunchKernelGGL((k_batch_wfirst<true>), dim3(waves), dim3(64), 0, s, views, chunks, nchunks, cellblk,
				   counters, ws, slab, stamp_dev, out);
	else
		hipLaunchKernelGGL((k_batch_wfirst<false>), dim3(waves), dim3(64), 0, s, views, chunks, nchunks, cellblk,
				   counters, ws, slab, stamp_dev, out);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_wlist(bool poly, int waves, hipStream_t s, const RectView *views, const int *bcont, const int *list,
		       const int *list_count, const int *cellblk, int *counters, char *ws, size_t slab,
		       const unsigned *stamp_dev, const CellOut &out) {
	if (poly)
		hipLaunchKernelGGL((k_batch_wlist<true>), dim3(waves), dim3(64), 0, s, views, bcont, list, list_count, cellblk,
				   counters, ws, slab, stamp_dev, out);
	else
		hipLaunchKernelGGL((k_batch_wlist<false>), dim3(waves), dim3(64), 0, s, views, bcont, list, list_count,
				   cellblk, counters, ws, slab, stamp_dev, out);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

}  // namespace vgpu
