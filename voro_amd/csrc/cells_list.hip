// The learned path's kernels: candidate-list gathering (count, per-batch
// capacity, fill) and the list-driven order-3 cell kernel.
//
// Lists are stored per batch of 64 consecutive cells, entry k of the cell
// in lane l at [(batch_offset + k) * 64 + l], so that the cell kernel's 64
// lanes read entry k of their lists in one coalesced access.
#include <hip/hip_runtime.h>

#include "batch_dev.h"
#include "learn.h"
#include "list_dev.h"
#include "vgpu_cells.h"
#include "wave_cell_dev.h"
#include "wcell_dev.h"

namespace vgpu {

// Neighbour ids of a wavefront cell in face-walk order: id i in lane i.
struct WListSink {
	WCell *c;
	VG_INL void put(int i, int v) { if (i < (int)WLayout::NBCAP) c->put_nb(i, v); }
};

// Schedule: cells are processed in order of decreasing predicted cost
// (64 cost classes, any order within a class), so that the 64 cells a
// wavefront cuts in lockstep need similar work.  Position i of the schedule
// holds cell order[i]; list counts and batches follow the schedule.
// ncp (may be null): the cell count on the device, at most nc.
__global__ void k_order_hist(int nc, const int *__restrict__ ncp, const int *__restrict__ cellblk,
			     const unsigned char *__restrict__ cost, int per_cell, int *__restrict__ hist) {
	__shared__ int h[64];
	if (ncp) nc = min(nc, *ncp);
	if (threadIdx.x < 64) h[threadIdx.x] = 0;
	__syncthreads();
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c < nc) atomicAdd(&h[63 - cost[per_cell ? c : cellblk[c]]], 1);
	__syncthreads();
	if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_order_scan(int *__restrict__ hist) {
	if (threadIdx.x == 0) {
		int a = 0;
		for (int t = 0; t < 64; t++) { const int v = hist[t]; hist[t] = a; a += v; }
	}
}

// Slots are claimed per workgroup: a class's cells of one workgroup are
// counted in LDS and take one global atomic together.
__global__ void k_order_scatter(int nc, const int *__restrict__ ncp, const int *__restrict__ cellblk,
				const unsigned char *__restrict__ cost, int per_cell, int *__restrict__ fill,
				int *__restrict__ order) {
	__shared__ int h[64], base[64];
	if (ncp) nc = min(nc, *ncp);
	if (threadIdx.x < 64) h[threadIdx.x] = 0;
	__syncthreads();
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	int k = 0, r = 0;
	if (c < nc) {
		k = 63 - cost[per_cell ? c : cellblk[c]];
		r = atomicAdd(&h[k], 1);
	}
	__syncthreads();
	if (threadIdx.x < 64 && h[threadIdx.x]) base[threadIdx.x] = atomicAdd(&fill[threadIdx.x], h[threadIdx.x]);
	__syncthreads();
	if (c < nc) order[base[k] + r] = c;
}

template<bool PRD>
__global__ void k_list_count(RectView v, int nc, const int *__restrict__ cellblk, const int *__restrict__ order,
			     const unsigned char *__restrict__ depth, int per_cell, int *__restrict__ cnt) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nc) return;
	const int c = order[i];
	const int G = depth[per_cell ? c : cellblk[c]];
	cnt[i] = list_length(G);
}

// Capacity of each batch of 64 cells (the longest list in it); entry nb is
// set to 0 so that an exclusive scan gives the offsets and the total.
__global__ void k_batch_max(int nc, const int *__restrict__ cnt, int *__restrict__ bmax) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	const int nb = (nc + 63) / 64;
	int v = c < nc ? cnt[c] : 0;
	for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
	const int b = c >> 6;
	if ((threadIdx.x & 63) == 0 && b < nb) bmax[b] = v;
	if (c == 0) bmax[nb] = 0;
}

// csr: position i's entries contiguous from boff[i] (wavefront cells);
// otherwise interleaved per batch of 64 positions from row boff[i / 64].
template<bool PRD>
__global__ void k_list_fill(RectView v, int nc, const int *__restrict__ cellblk, const int *__restrict__ order,
			    const unsigned char *__restrict__ depth, int per_cell, const int *__restrict__ boff,
			    int csr, ListRec *__restrict__ LA) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nc) return;
	const int c = order[i];
	const int G = depth[per_cell ? c : cellblk[c]];
	if (csr) {
		const size_t base = (size_t)boff[i];
		gather_cell_list<PRD>(v, cellblk[c], c, G, LA + base, 1);
	} else {
		const size_t base = (size_t)boff[i >> 6] * 64 + (i & 63);
		gather_cell_list<PRD>(v, cellblk[c], c, G, LA + base, 64);
	}
}

// The contiguous lists (wavefront cells): one wavefront per cell, lane g
// writing record g, so that a list is written in full cache lines (one
// thread per cell wrote each 32-byte record on its own line: 4.7 GB of
// writes for 1.7 GB of records).  Four cells per workgroup.
template<bool PRD>
__global__ void __launch_bounds__(256) k_list_fill_lanes(RectView v, int nc, const int *__restrict__ cellblk,
							  const int *__restrict__ order, const unsigned char *__restrict__ depth,
							  int per_cell, const int *__restrict__ boff, ListRec *__restrict__ LA) {
	const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), g = (int)(threadIdx.x & 63);
	if (i >= nc) return;
	const int c = order[i];
	const int G = depth[per_cell ? c : cellblk[c]];
	if (G >= 64 || g > G) return;
	const ListCtx x = list_ctx<PRD>(v, cellblk[c], c);
	LA[(size_t)boff[i] + g] = list_record<PRD>(v, x, g, G);
}

// One wavefront per workgroup, a batch of 64 cells at a time (batches
// taken in order through an atomic counter).  Cells whose list ends before
// the radius test stops them, and cells that overflow the order-3 tier, go to
// the plain path's retry list; *fallbacks counts the former, *routed the
// cells given no list (predicted to reach the block search).  PRD: a
// periodic container (first visits recorded with the plain path's keys,
// image reads flagged in out.touch).
template<bool POLY, bool PRD, int OCC>
__global__ void __launch_bounds__(64, OCC)
k_cells_list(RectView v, int nc, const int *__restrict__ cellblk, const int *__restrict__ order, const int *__restrict__ cnt,
	     const int *__restrict__ boff, const ListRec *__restrict__ LA,
	     int *__restrict__ counters, char *__restrict__ ws, int *__restrict__ fallbacks, int *__restrict__ routed,
	     CellOut out) {
	const int lane = threadIdx.x;
	VCell3 c;
	c.setup(ws + (size_t)blockIdx.x * VCell3::L::bytes, lane, v.max_len_sq);
	const int nb = (nc + 63) / 64;
	for (;;) {
		int b = 0;
		if (lane == 0) b = atomicAdd(&counters[0], 1);
		b = __shfl(b, 0);
		if (b >= nb) break;
		const int pos = b * 64 + lane;
		if (pos >= nc) continue;
		const int cell = order[pos];
		const int ijk = cellblk[cell];
		const int s = cell - v.cstart[ijk];
		const size_t base = (size_t)boff[b] * 64 + lane;
		CellStats stt;
		const int r = compute_cell_list<VCell3, POLY, PRD>(c, v, ijk, s, LA + base, 64, cnt[pos],
								   stt, (unsigned long long)cell << 24);
		double vol = 0, cx = 0, cy = 0, cz = 0;
		int nf = 0;
		bool retry = r < 0 || c.ovf;
		if (!retry && r == 1) {
			SlabSink<VCell3> sk{&c};
			c.outputs(vol, cx, cy, cz, nf, sk);
			if (c.ovf || nf > (int)VCell3::L::NBCAP) retry = true;
		}
		if (retry) {
			if (r == -1) atomicAdd(fallbacks, 1);
			else if (r == -3) atomicAdd(routed, 1);
			const int slot_ = atomicAdd(out.retry_count, 1);
			out.retry[slot_] = cell;
			out.status[cell] = ST_OVERFLOW;
			continue;
		}
		out.vol[cell] = vol;
		out.cent[3 * cell] = cx; out.cent[3 * cell + 1] = cy; out.cent[3 * cell + 2] = cz;
		out.nfaces[cell] = nf;
		out.status[cell] = r == 1 ? ST_EXISTS : 0;
		if (out.want_nbrs) {
			const int off = nf > 0 ? atomicAdd(out.arena_top, nf) : 0;
			if (off + nf > out.arena_cap) { out.nbr_start[cell] = -1; atomicAdd(out.fail_count, 1); }
			else {
				out.nbr_start[cell] = off;
				for (int t = 0; t < nf; t++) out.nbr[off + t] = c.template A<int>(VCell3::L::o_nbs, t);
			}
		}
		if (out.st_planes) {
			out.st_planes[cell] = (short)min(stt.n_planes, 32767);
			out.st_exit[cell] = (signed char)stt.wl_exit;
			out.st_tier[cell] = (signed char)out.tier;
		}
		if (PRD && out.touch) out.touch[cell] = (unsigned char)stt.img;
	}
}

// One cell of the list-driven wavefront tier: compute_cell_list over the
// cell's list (nent records from LA), then the outputs, or an entry in
// out's retry list (overflow).  A cell whose list ends first (*fallbacks) is
// continued from the list's last step, and a cell that has no list
// (*routed) computed, by the plain algorithm at once in the same wavefront
// (block-mask stamp `stamp`), so that the deep cells -- scheduled first --
// overlap with the rest instead of running after them.  res: the wavefront's
// LDS record of where the list ended.
template<bool POLY, bool PRD>
__device__ __forceinline__ void wlist_cell(WCell &c, const RectView &v, int cell, int ijk, const ListRec *LA, int nent,
					   unsigned stamp, int *fallbacks, int *routed, const CellOut &out, ListResume &res) {
	const int lane = threadIdx.x;
	const int s = cell - v.cstart[ijk];
	CellStats stt;
	const int r = compute_cell_list<WCell, POLY, PRD>(c, v, ijk, s, LA, 1, nent, stt,
							  (unsigned long long)cell << 24, &res);
	double vol = 0, cx = 0, cy = 0, cz = 0;
	int nf = 0;
	if (r == -1 || r == -3) {
		if (lane == 0) atomicAdd(r == -1 ? fallbacks : routed, 1);
		if (r == -1) list_resume_fill<POLY, PRD>(v, ijk, s, LA, 1, nent, (unsigned long long)cell << 24, res);
		else res.g = -1;
		wave_cell<1, POLY, PRD>(c, v, cell, ijk, stamp, out, &res);
		return;
	}
	bool retry = r < 0 || c.ovf;
	if (!retry && r == 1) {
		WListSink sk{&c};
		c.outputs(vol, cx, cy, cz, nf, sk);
		if (c.ovf || nf > (int)WLayout::NBCAP) retry = true;
	}
	if (retry) {
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
		out.status[cell] = r == 1 ? ST_EXISTS : 0;
		if (out.st_planes) {
			out.st_planes[cell] = (short)min(stt.n_planes, 32767);
			out.st_exit[cell] = (signed char)stt.wl_exit;
			out.st_tier[cell] = (signed char)out.tier;
		}
		if (PRD && out.touch) out.touch[cell] = (unsigned char)stt.img;
	}
	if (out.want_nbrs) {
		int off = 0;
		if (lane == 0 && nf > 0) off = atomicAdd(out.arena_top, nf);
		off = __builtin_amdgcn_readfirstlane(off);
		if (off + nf > out.arena_cap) {
			if (lane == 0) { out.nbr_start[cell] = -1; atomicAdd(out.fail_count, 1); }
		} else {
			if (lane == 0) out.nbr_start[cell] = off;
			if (lane < nf) out.nbr[off + lane] = c.nb_lane(0);
		}
	}
}

// The same over the lists, one cell per wavefront (wcell_dev.h): the list
// of schedule position pos is contiguous from boff[pos], its entries read
// with uniform loads.  Positions are taken `chunk` at a time from one
// counter per XCD share.
template<bool POLY, bool PRD, int OCC>
__global__ void __launch_bounds__(64, OCC)
k_wcells_list(RectView v_, int nc, const int *__restrict__ cellblk, const int *__restrict__ order,
	      const int *__restrict__ cnt, const int *__restrict__ boff, const ListRec *__restrict__ LA,
	      int *__restrict__ counters, char *__restrict__ ws, unsigned stamp_base, int chunk,
	      int *__restrict__ fallbacks, int *__restrict__ routed, CellOut out) {
	typedef const __attribute__((address_space(4))) RectView CView;
	const int lane = threadIdx.x;
	__shared__ ListResume res;
	WCell c;
	c.setup(ws + (size_t)blockIdx.x * WLayout::bytes, lane, v_.max_len_sq);
	const int nshare = 8;
	const int my = blockIdx.x % nshare;
	const int share = (nc + nshare - 1) / nshare;
	for (int h = 0; h < nshare; h++) {
		const int sh = (my + h) % nshare;
		const int lo = sh * share, hi = min(nc, lo + share);
		for (;;) {
			int base = 0;
			if (lane == 0) base = atomicAdd(&counters[sh], chunk);
			base = __builtin_amdgcn_readfirstlane(base);
			if (lo + base >= hi) break;
			const int end = min(hi, lo + base + chunk);
			// the view read through the argument segment, as in k_wcells
			int z;
			asm volatile("s_mov_b32 %0, 0" : "=s"(z));
			const RectView &v = *(const RectView *)((CView *)__builtin_amdgcn_kernarg_segment_ptr() + z);
			for (int pos = lo + base; pos < end; pos++) {
				const int cell = order[pos];
				const int ijk = cellblk[cell];
				const size_t bs = (size_t)boff[pos];
				wlist_cell<POLY, PRD>(c, v, cell, ijk, LA + bs, cnt[pos], stamp_base + (unsigned)cell + 1u, fallbacks,
						      routed, out, res);
			}
		}
	}
}

// ---- batches of containers ----

// The container of every block of a batch (blocks of container c are
// boff[c] .. boff[c+1]-1).
__global__ void k_batch_bcont(int ncont, const int *__restrict__ boff, int *__restrict__ bcont) {
	const int c = blockIdx.x;
	if (c >= ncont) return;
	for (int t = boff[c] + threadIdx.x; t < boff[c + 1]; t += blockDim.x) bcont[t] = c;
}

// Block features of every block of a batch: its container's 27 neighbouring
// block counts over that container's mean count (learn.h's layout; input 27
// is 0), as a single container's blocks get them.
__global__ void k_batch_features(int nb, const int *__restrict__ bcont, const RectView *__restrict__ views,
				 float *__restrict__ X) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= nb) return;
	const RectView &v = views[bcont[q]];
	const int b = q - v.boff;
	const int nx = v.nx, ny = v.ny, nz = v.nz, nxyz = nx * ny * nz;
	const int ncells = v.cstart[nxyz] - v.cstart[0];
	const float inv_mean = ncells > 0 ? (float)nxyz / (float)ncells : 0.f;
	const int i = b % nx, j = (b / nx) % ny, k = b / (nx * ny);
	int t = 0;
	for (int dk = -1; dk <= 1; dk++)
		for (int dj = -1; dj <= 1; dj++)
			for (int di = -1; di <= 1; di++, t++) {
				const int a = i + di, c = j + dj, e = k + dk;
				float x = 0.f;
				if (a >= 0 && a < nx && c >= 0 && c < ny && e >= 0 && e < nz) {
					const int r = a + nx * (c + ny * e);
					x = (float)(v.bstart[r + 1] - v.bstart[r]) * inv_mean;
				}
				X[(size_t)q * learn::IN + t] = x;
			}
	X[(size_t)q * learn::IN + 27] = 0.f;
}

// Candidate lists of a batch's cells, schedule position i (cell order[i])
// by thread i, in one pass: the length (depth + 1 records), a
// wavefront-wide scan, one 64-bit bump of *top per wavefront, then the
// fill, one list at a time by the whole wavefront.  A list that would
// pass `cap` records is not written: its length is set to 0 and the cell
// takes the plain path (*top still grows by it, so the caller learns how
// much room everything needed).  loff[i] / lcnt[i]: the
// list's first record and length.  Sizes are read on the device, so the
// sequence needs no host round trip.
__global__ void __launch_bounds__(256)
k_batch_list_build(const RectView *__restrict__ views, const int *__restrict__ bcont, int n, const int *__restrict__ ncp,
		   const int *__restrict__ cellblk, const int *__restrict__ order, const unsigned char *__restrict__ depth,
		   long long cap, unsigned long long *__restrict__ top, int *__restrict__ lcnt, int *__restrict__ loff,
		   ListRec *__restrict__ LA) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int nc = min(n, *ncp);
	const int lane = threadIdx.x & 63;
	int cell = 0, gb = 0, cnt = 0;
	if (i < nc) {
		cell = order[i];
		gb = cellblk[cell];
		cnt = list_length(depth[gb]);
	}
	// inclusive scan over the wavefront
	int inc = cnt;
	for (int o = 1; o < 64; o <<= 1) {
		const int y = __shfl_up(inc, o);
		if (lane >= o) inc += y;
	}
	const int total = __shfl(inc, 63);
	unsigned long long base = 0;
	if (lane == 63 && total > 0) base = atomicAdd(top, (unsigned long long)total);
	base = __shfl(base, 63);
	const unsigned long long off = base + (unsigned long long)(inc - cnt);
	const bool ok = i < nc && cnt > 0 && off + (unsigned long long)cnt <= (unsigned long long)cap;
	if (i < n) {
		lcnt[i] = ok ? cnt : 0;
		loff[i] = ok ? (int)off : 0;
	}
	// The records: each list written by the whole wavefront, lane g writing
	// record g, so that every list is one run of full cache lines.
	for (unsigned long long m = __ballot(ok); m; m &= m - 1) {
		const int src = __builtin_ctzll(m);
		const int cj = __shfl(cell, src), gbj = __shfl(gb, src), G = __shfl(cnt, src) - 1;
		const unsigned long long oj = (unsigned long long)(unsigned)__shfl((int)(unsigned)off, src) |
					      ((unsigned long long)(unsigned)__shfl((int)(unsigned)(off >> 32), src) << 32);
		if (lane <= G) {
			const RectView &v = views[bcont[gbj]];
			const ListCtx x = list_ctx<false>(v, gbj - v.boff, cj);
			LA[oj + lane] = list_record<false>(v, x, lane, G);
		}
	}
}

// The list-driven wavefront tier over a batch's schedule (positions
// 0 .. *ncp-1), each cell with its container's view; positions are taken
// `chunk` at a time from one counter per XCD share.  OCC waves per SIMD (4,
// 5 or 6).
template<bool POLY, int OCC>
__global__ void __launch_bounds__(64, OCC)
k_batch_wcells_list(const RectView *__restrict__ views, const int *__restrict__ bcont, const int *__restrict__ ncp,
		    const int *__restrict__ cellblk, const int *__restrict__ order, const int *__restrict__ lcnt,
		    const int *__restrict__ loff, const ListRec *__restrict__ LA,
		    int *__restrict__ counters, char *__restrict__ ws, size_t slab_bytes, const unsigned *__restrict__ stamp_dev,
		    int chunk, int *__restrict__ fallbacks, int *__restrict__ routed, CellOut out) {
	const int lane = threadIdx.x;
	__shared__ ListResume res;
	WCell c;
	c.setup(ws + (size_t)blockIdx.x * slab_bytes, lane, 0.0);
	const unsigned stamp_base = *stamp_dev;
	const int nc = *ncp;
	const int nshare = 8;
	const int my = blockIdx.x % nshare;
	const int share = (nc + nshare - 1) / nshare;
	for (int h = 0; h < nshare; h++) {
		const int sh = (my + h) % nshare;
		const int lo = sh * share, hi = min(nc, lo + share);
		for (;;) {
			int base = 0;
			if (lane == 0) base = atomicAdd(&counters[sh], chunk);
			base = __builtin_amdgcn_readfirstlane(base);
			if (lo + base >= hi) break;
			const int end = min(hi, lo + base + chunk);
			for (int pos = lo + base; pos < end; pos++) {
				const int cell = __builtin_amdgcn_readfirstlane(order[pos]);
				const int gb = __builtin_amdgcn_readfirstlane(cellblk[cell]);
				const int cont = __builtin_amdgcn_readfirstlane(bcont[gb]);
				typedef const __attribute__((address_space(4))) RectView CView;
				const RectView &v = *(const RectView *)((CView *)views + cont);
				c.set_tol(v.max_len_sq);
				const size_t bs = (size_t)(unsigned)loff[pos];
				wlist_cell<POLY, false>(c, v, cell, gb - v.boff, LA + bs, lcnt[pos], stamp_base + (unsigned)cell + 1u,
							fallbacks, routed, out, res);
			}
		}
	}
}

int launch_batch_bcont(hipStream_t s, int ncont, const int *boff, int *bcont) {
	if (ncont > 0) hipLaunchKernelGGL(k_batch_bcont, dim3(ncont), dim3(256), 0, s, ncont, boff, bcont);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_features(hipStream_t s, int nb, const int *bcont, const RectView *views, float *X) {
	if (nb > 0) hipLaunchKernelGGL(k_batch_features, dim3((nb + 255) / 256), dim3(256), 0, s, nb, bcont, views, X);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_list_build(hipStream_t s, const RectView *views, const int *bcont, int n, const int *ncp,
			    const int *cellblk, const int *order, const unsigned char *depth, long long cap,
			    unsigned long long *top, int *lcnt, int *loff, ListRec *LA) {
	if (n > 0)
		hipLaunchKernelGGL(k_batch_list_build, dim3((n + 255) / 256), dim3(256), 0, s, views, bcont, n, ncp, cellblk,
				   order, depth, cap, top, lcnt, loff, LA);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_batch_wcells_list(bool poly, int occ, int waves, int chunk, hipStream_t s, const RectView *views,
			     const int *bcont, const int *ncp, const int *cellblk, const int *order, const int *lcnt,
			     const int *loff, const ListRec *LA, int *counters, char *ws, size_t slab,
			     const unsigned *stamp_dev, int *fallbacks, int *routed, const CellOut &out) {
#define VG_BWLIST(P, O)                                                                                          \
	hipLaunchKernelGGL((k_batch_wcells_list<P, O>), dim3(waves), dim3(64), 0, s, views, bcont, ncp, cellblk, order, \
			   lcnt, loff, LA, counters, ws, slab, stamp_dev, chunk, fallbacks, routed, out)
	if (occ >= 6) { if (poly) VG_BWLIST(true, 6); else VG_BWLIST(false, 6); }
	else if (occ == 5) { if (poly) VG_BWLIST(true, 5); else VG_BWLIST(false, 5); }
	else { if (poly) VG_BWLIST(true, 4); else VG_BWLIST(false, 4); }
#undef VG_BWLIST
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_cells_wlist(bool poly, bool prd, int occ, int waves, int chunk, hipStream_t s, const RectView &v, int nc,
		       const int *cellblk, const int *order, const int *cnt, const int *boff, const ListRec *LA,
		       int *counters, char *ws, unsigned stamp_base, int *fallbacks, int *routed, const CellOut &out) {
#define VG_WLIST(P, R, O)                                                                                           \
	hipLaunchKernelGGL((k_wcells_list<P, R, O>), dim3(waves), dim3(64), 0, s, v, nc, cellblk, order, cnt, boff, LA, \
			   counters, ws, stamp_base, chunk, fallbacks, routed, out)
	// 4, 5 or 6 waves per SIMD
	if (prd) {
		if (occ == 5) { if (poly) VG_WLIST(true, true, 5); else VG_WLIST(false, true, 5); }
		else if (occ >= 6) { if (poly) VG_WLIST(true, true, 6); else VG_WLIST(false, true, 6); }
		else { if (poly) VG_WLIST(true, true, 4); else VG_WLIST(false, true, 4); }
	} else if (occ == 5) { if (poly) VG_WLIST(true, false, 5); else VG_WLIST(false, false, 5); }
	else if (occ >= 6) { if (poly) VG_WLIST(true, false, 6); else VG_WLIST(false, false, 6); }
	else { if (poly) VG_WLIST(true, false, 4); else VG_WLIST(false, false, 4); }
#undef VG_WLIST
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_list_order(hipStream_t s, int nc, const int *cellblk, const unsigned char *cost, int per_cell, int *hist,
		      int *order, const int *ncp) {
	const int g = (nc + 255) / 256;
	hipLaunchKernelGGL(k_order_hist, dim3(g), dim3(256), 0, s, nc, ncp, cellblk, cost, per_cell, hist);
	hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(64), 0, s, hist);
	hipLaunchKernelGGL(k_order_scatter, dim3(g), dim3(256), 0, s, nc, ncp, cellblk, cost, per_cell, hist, order);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

// v.fv must be null: the gather only reads.
int launch_list_gather(bool prd, hipStream_t s, const RectView &v, int nc, const int *cellblk, const int *order,
		       const unsigned char *depth, int per_cell, int *cnt, int *bmax) {
	const int g = (nc + 255) / 256;
	if (v.fv) return 2;
	if (prd) hipLaunchKernelGGL(k_list_count<true>, dim3(g), dim3(256), 0, s, v, nc, cellblk, order, depth, per_cell, cnt);
	else hipLaunchKernelGGL(k_list_count<false>, dim3(g), dim3(256), 0, s, v, nc, cellblk, order, depth, per_cell, cnt);
	hipLaunchKernelGGL(k_batch_max, dim3(g), dim3(256), 0, s, nc, cnt, bmax);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_list_fill(bool prd, hipStream_t s, const RectView &v, int nc, const int *cellblk, const int *order,
		     const unsigned char *depth, int per_cell, const int *boff, bool csr, ListRec *LA) {
	if (v.fv) return 2;
	if (csr) {
		const unsigned blocks = (unsigned)(((long long)nc + 3) / 4);
		if (prd)
			hipLaunchKernelGGL(k_list_fill_lanes<true>, dim3(blocks), dim3(256), 0, s, v, nc, cellblk, order, depth,
					   per_cell, boff, LA);
		else
			hipLaunchKernelGGL(k_list_fill_lanes<false>, dim3(blocks), dim3(256), 0, s, v, nc, cellblk, order, depth,
					   per_cell, boff, LA);
	} else if (prd)
		hipLaunchKernelGGL(k_list_fill<true>, dim3((nc + 255) / 256), dim3(256), 0, s, v, nc, cellblk, order, depth,
				   per_cell, boff, 0, LA);
	else
		hipLaunchKernelGGL(k_list_fill<false>, dim3((nc + 255) / 256), dim3(256), 0, s, v, nc, cellblk, order, depth,
				   per_cell, boff, 0, LA);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_cells_list(bool poly, bool prd, int waves, hipStream_t s, const RectView &v, int nc, const int *cellblk,
		      const int *order, const int *cnt, const int *boff, const ListRec *LA,
		      int *counters, char *ws, int *fallbacks, int *routed, const CellOut &out) {
#define VG_LIST(P, R)                                                                                               \
	hipLaunchKernelGGL((k_cells_list<P, R, 2>), dim3(waves), dim3(64), 0, s, v, nc, cellblk, order, cnt, boff, LA, \
			   counters, ws, fallbacks, routed, out)
	if (prd) { if (poly) VG_LIST(true, true); else VG_LIST(false, true); }
	else { if (poly) VG_LIST(true, false); else VG_LIST(false, false); }
#undef VG_LIST
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

}  // namespace vgpu
