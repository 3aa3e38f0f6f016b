// Batches of independent containers (many containers of uneven size in one
// pass): device-side set-up of the per-container tables, and the persistent
// work-queue cell kernels.
//
// All containers' particles are binned into one block array (container c
// owns blocks [boff[c], boff[c+1])), so every container's cells are
// contiguous, in its own c_loop_all order.  The first tier walks a queue of
// chunks -- up to CH consecutive cells of one container -- so that a
// wavefront's container, and with it the container's view, is uniform; the
// retry tiers walk lists of cells of any container, each lane reading its
// own container's view.  Every kernel takes its counts from device memory,
// so the whole pipeline can be captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include "batch_dev.h"
#include "grid_host.h"

namespace vgpu {

// Worklist radii of every container's block size (src/v_base.cc:23-69),
// one thread per (container, sub-region).
__global__ void k_batch_mrad(int ncont, const RectGeom *__restrict__ geoms, const unsigned *__restrict__ wl,
			     double *__restrict__ mrad) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= ncont * 64) return;
	const RectGeom g = geoms[t >> 6];
	worklist_radii_region(g.boxx, g.boxy, g.boxz, t & 63, wl, mrad + (size_t)(t >> 6) * 4096 + (size_t)(t & 63) * 64);
}

// Block index of every particle in its container (as k_locate), shifted to
// the batch's block numbering; largest radius per container.
__global__ void k_batch_locate(int n, const double *__restrict__ xyz, const double *__restrict__ r,
			       const RectGeom *__restrict__ geoms, const int *__restrict__ pcont,
			       const int *__restrict__ boff, int *__restrict__ blk, int *__restrict__ slot,
			       dbl4 *__restrict__ rem, int *__restrict__ cnt, unsigned long long *__restrict__ maxr_bits) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const int c = pcont[t];
	double x = xyz[3 * t], y = xyz[3 * t + 1], z = xyz[3 * t + 2];
	int b = locate_rect(geoms[c], x, y, z);
	if (b >= 0) b += boff[c];
	blk[t] = b;
	const double rr = r ? r[t] : 0.5;
	rem[t] = dbl4{x, y, z, rr};
	if (b >= 0) {
		slot[t] = atomicAdd(&cnt[b], 1);
		if (r && rr > 0) atomicMax(&maxr_bits[c], (unsigned long long)__double_as_longlong(rr));
	}
}

// Per container: its first cell (cfirst, with cfirst[ncont] = all cells),
// its number of chunks of ch cells (nchunk, entry ncont = 0 so that an
// exclusive scan gives offsets and total), and its largest radius into the
// view (radical containers).
__global__ void k_batch_counts(int ncont, int ch, const int *__restrict__ boff, const int *__restrict__ bstart,
			       const unsigned long long *__restrict__ maxr_bits, int poly, RectView *__restrict__ views,
			       int *__restrict__ cfirst, int *__restrict__ nchunk) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c > ncont) return;
	const int a = bstart[boff[c]];
	cfirst[c] = a;
	if (c == ncont) { nchunk[c] = 0; return; }
	const int e = bstart[boff[c + 1]];
	nchunk[c] = (e - a + ch - 1) / ch;
	views[c].max_radius = poly ? __longlong_as_double((long long)maxr_bits[c]) : 0.0;
}

// The chunk queue: container c's chunks at [coff[c], coff[c+1]); the total
// goes to *nchunks.
__global__ void k_batch_chunks(int ncont, int ch, const int *__restrict__ cfirst, const int *__restrict__ coff,
			       BatchChunk *__restrict__ chunks, int *__restrict__ nchunks) {
	const int c = blockIdx.x;
	if (c == 0 && threadIdx.x == 0) *nchunks = coff[ncont];
	if (c >= ncont) return;
	const int a = cfirst[c], e = cfirst[c + 1], k0 = coff[c];
	for (int k = threadIdx.x; k * ch < e - a; k += blockDim.x)
		chunks[k0 + k] = BatchChunk{c, a + k * ch, min(ch, e - a - k * ch), 0};
}

// The container of a cell, from its block: the last c with boff[c] <= block.
__device__ __forceinline__ int batch_container(const int *__restrict__ boff, int ncont, int block) {
	int lo = 0, hi = ncont - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (boff[mid] <= block) lo = mid; else hi = mid - 1;
	}
	return lo;
}

// One cell of the container with view v, with the outputs and the retry
// handling of the single-container kernel.
template<class CELL, bool POLY, bool LAST>
__device__ __forceinline__ void batch_cell(CELL &c, const RectView &v, int cell, int block, unsigned stamp,
					   const CellOut &out) {
	const int b = block - v.boff;
	const int s = cell - v.cstart[b];
	const int ci = b % v.nx, cj = (b / v.nx) % v.gy, ck = b / v.nxy;
	CellStats stt;
	bool ok = compute_cell_rect<CELL, POLY>(c, v, b, s, ci, cj, ck, stamp, stt);
	double vol = 0, cx = 0, cy = 0, cz = 0;
	int nf = 0;
	if (!c.ovf && ok) {
		SlabSink<CELL> sk{&c};
		c.outputs(vol, cx, cy, cz, nf, sk);
		if (nf > (int)CELL::L::NBCAP) c.ovf = true;
	}
	if (c.ovf) {
		if (LAST) {
			atomicAdd(out.fail_count, 1);
			out.status[cell] = c.fatal ? ST_FATAL : ST_OVERFLOW;
		} else {
			const int slot_ = atomicAdd(out.retry_count, 1);
			out.retry[slot_] = cell;
			out.status[cell] = ST_OVERFLOW;
		}
		return;
	}
	out.vol[cell] = vol;
	out.cent[3 * cell] = cx; out.cent[3 * cell + 1] = cy; out.cent[3 * cell + 2] = cz;
	out.nfaces[cell] = nf;
	out.status[cell] = ok ? ST_EXISTS : 0;
	if (out.want_nbrs) {
		const int off = nf > 0 ? atomicAdd(out.arena_top, nf) : 0;
		if (off + nf > out.arena_cap) { out.nbr_start[cell] = -1; atomicAdd(out.fail_count, 1); }
		else {
			out.nbr_start[cell] = off;
			for (int t = 0; t < nf; t++) out.nbr[off + t] = c.template A<int>(CELL::L::o_nbs, t);
		}
	}
	if (out.st_planes) {
		out.st_planes[cell] = (short)min(stt.n_planes, 32767);
		out.st_exit[cell] = (signed char)stt.wl_exit;
		out.st_tier[cell] = (signed char)out.tier;
	}
}

// First tier over the chunk queue.  The queue is split into one contiguous
// share per XCD (workgroups are dispatched to the 8 XCDs round robin), so
// that the chunks of a container, which read the same blocks, share an L2;
// a wavefront whose share is exhausted moves on to the next.
template<bool POLY, int OCC>
__global__ void __launch_bounds__(64, OCC)
k_batch_first(const RectView *__restrict__ views, const BatchChunk *__restrict__ chunks, const int *__restrict__ nchunks,
	      const int *__restrict__ cellblk, int *__restrict__ counters, char *__restrict__ ws, size_t slab_bytes,
	      const unsigned *__restrict__ stamp_dev, CellOut out) {
	typedef VCell3 CELL;
	const int lane = threadIdx.x;
	CELL c;
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
			const RectView &v = views[cont];
			c.set_tol(v.max_len_sq);
			if (lane < count) {
				const int cell = first + lane;
				batch_cell<CELL, POLY, false>(c, v, cell, cellblk[cell], stamp_base + (unsigned)cell + 1u, out);
			}
		}
	}
}

// Retry tiers over a list of cells of any container.  Each wavefront takes
// per_wave cells at a time, per_wave chosen from the list length on the
// device so that a short list of expensive cells spreads over all
// wavefronts.
template<class T_, bool POLY, bool LAST>
__global__ void __launch_bounds__(64, 2)
k_batch_retry(const RectView *__restrict__ views, const int *__restrict__ boff, int ncont, const int *__restrict__ list,
	      const int *__restrict__ list_count, const int *__restrict__ cellblk, int *__restrict__ counters,
	      char *__restrict__ ws, size_t slab_bytes, const unsigned *__restrict__ stamp_dev, CellOut out) {
	typedef typename TierCell<T_>::type CELL;
	const int lane = threadIdx.x;
	CELL c;
	c.setup(ws + (size_t)blockIdx.x * slab_bytes, lane, 0.0);
	const unsigned stamp_base = *stamp_dev;
	const int total = *list_count;
	const int per_wave = max(1, min(64, (total + (int)gridDim.x - 1) / (int)gridDim.x));
	for (;;) {
		int base = 0;
		if (lane == 0) base = atomicAdd(&counters[0], per_wave);
		base = __builtin_amdgcn_readfirstlane(base);
		if (base >= total) break;
		const int idx = base + lane;
		if (idx < total && lane < per_wave) {
			const int cell = list[idx];
			const int block = cellblk[cell];
			const RectView &v = views[batch_container(boff, ncont, block)];
			c.set_tol(v.max_len_sq);
			batch_cell<CELL, POLY, LAST>(c, v, cell, block, stamp_base + (unsigned)cell + 1u, out);
		}
	}
}

int launch_batch_mrad(hipStream_t s, int ncont, const RectGeom *geoms, const unsigned *wl, double *mrad) {
	hipLaunchKernelGGL(k_batch_mrad, dim3((ncont * 64 + 255) / 256), dim3(256), 0, s, ncont, geoms, wl, mrad);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_locate(hipStream_t s, int n, const double *xyz, const double *r, const RectGeom *geoms,
			const int *pcont, const int *boff, int *blk, int *slot, dbl4 *rem, int *cnt,
			unsigned long long *maxr_bits) {
	if (n > 0)
		hipLaunchKernelGGL(k_batch_locate, dim3((n + 255) / 256), dim3(256), 0, s, n, xyz, r, geoms, pcont, boff, blk,
				   slot, rem, cnt, maxr_bits);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_counts(hipStream_t s, int ncont, int ch, const int *boff, const int *bstart,
			const unsigned long long *maxr_bits, int poly, RectView *views, int *cfirst, int *nchunk) {
	hipLaunchKernelGGL(k_batch_counts, dim3((ncont + 1 + 255) / 256), dim3(256), 0, s, ncont, ch, boff, bstart,
			   maxr_bits, poly, views, cfirst, nchunk);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_chunks(hipStream_t s, int ncont, int ch, const int *cfirst, const int *coff, BatchChunk *chunks,
			int *nchunks) {
	hipLaunchKernelGGL(k_batch_chunks, dim3(ncont > 0 ? ncont : 1), dim3(64), 0, s, ncont, ch, cfirst, coff, chunks,
			   nchunks);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_first(bool poly, int occ, int waves, hipStream_t s, const RectView *views, const BatchChunk *chunks,
		       const int *nchunks, const int *cellblk, int *counters, char *ws, size_t slab,
		       const unsigned *stamp_dev, const CellOut &out) {
#define VGPU_BATCH_FIRST(P, O)                                                                                  \
	hipLaunchKernelGGL((k_batch_first<P, O>), dim3(waves), dim3(64), 0, s, views, chunks, nchunks, cellblk, counters, \
			   ws, slab, stamp_dev, out)
	if (poly) {
		if (occ >= 4) VGPU_BATCH_FIRST(true, 4);
		else if (occ == 3) VGPU_BATCH_FIRST(true, 3);
		else VGPU_BATCH_FIRST(true, 2);
	} else {
		if (occ >= 4) VGPU_BATCH_FIRST(false, 4);
		else if (occ == 3) VGPU_BATCH_FIRST(false, 3);
		else VGPU_BATCH_FIRST(false, 2);
	}
#undef VGPU_BATCH_FIRST
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_retry(int tier, bool poly, int waves, hipStream_t s, const RectView *views, const int *boff,
		       int ncont, const int *list, const int *list_count, const int *cellblk, int *counters, char *ws,
		       size_t slab, const unsigned *stamp_dev, const CellOut &out) {
#define VGPU_BATCH_RETRY(T, P, L)                                                                                  \
	hipLaunchKernelGGL((k_batch_retry<T, P, L>), dim3(waves), dim3(64), 0, s, views, boff, ncont, list, list_count, \
			   cellblk, counters, ws, slab, stamp_dev, out)
	if (tier == 0) {
		if (poly) VGPU_BATCH_RETRY(Tier3, true, false);
		else VGPU_BATCH_RETRY(Tier3, false, false);
	} else if (tier == 1) {
		if (poly) VGPU_BATCH_RETRY(TierWide, true, false);
		else VGPU_BATCH_RETRY(TierWide, false, false);
	} else {
		if (poly) VGPU_BATCH_RETRY(TierHuge, true, true);
		else VGPU_BATCH_RETRY(TierHuge, false, true);
	}
#undef VGPU_BATCH_RETRY
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

}  // namespace vgpu
