// Binning (stable counting sort by block), scans and neighbour compaction
// kernels.  Included by the API translation unit only.
#pragma once
#include "batch_dev.h"

namespace vgpu {

// Block index of every particle, as container_base::put_remap computes it
// (src/container.cc:174-191), the position remapped into the primary domain
// for periodic directions, and an (unordered) slot in its block.  Dropped
// particles get block -1.  The largest radius of a stored particle is
// folded into *maxr_bits (positive doubles order like their bit patterns).
__global__ void k_locate(int n, const double *__restrict__ xyz, const double *__restrict__ r, RectGeom g,
			 int *__restrict__ blk, int *__restrict__ slot, dbl4 *__restrict__ rem,
			 int *__restrict__ cnt, unsigned long long *__restrict__ maxr_bits) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	double x = xyz[3 * t], y = xyz[3 * t + 1], z = xyz[3 * t + 2];
	const int b = locate_rect(g, x, y, z);
	blk[t] = b;
	double rr = r ? r[t] : 0.5;
	rem[t] = dbl4{x, y, z, rr};
	if (b >= 0) {
		slot[t] = atomicAdd(&cnt[b], 1);
		if (r && rr > 0) atomicMax(maxr_bits, (unsigned long long)__double_as_longlong(rr));
	}
}

// Geometry needed to bin particles of a periodic container.
struct PrdBin {
	double bx, bxy, by, bxz, byz, bz;
	double xsp, ysp, zsp;
	int nx, ny, nz, ey, ez, oy;
};

__device__ __forceinline__ int step_div_d(int a, int b) { return a >= 0 ? a / b : -1 + (a + 1) / b; }

// Block index of every particle of a periodic container and its position
// remapped into the primary domain, as put_locate_block computes them
// (src/container_prd.cc:189-217): z first (shifting y and x by the lattice's
// off-diagonal terms), then y (shifting x), then x.  A coordinate whose
// block index would not fit an int is dropped (block -1).
__global__ void k_locate_prd(int n, const double *__restrict__ xyz, const double *__restrict__ r, PrdBin g,
			     int *__restrict__ blk, int *__restrict__ slot, dbl4 *__restrict__ rem,
			     int *__restrict__ cnt, unsigned long long *__restrict__ maxr_bits) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	double x = xyz[3 * t], y = xyz[3 * t + 1], z = xyz[3 * t + 2];
	int b = -1;
	const double lim = 1073741824.0;
	const double zs = z * g.zsp;
	if (fabs(zs) < lim) {
		int k = step_int_d(zs);
		if (k < 0 || k >= g.nz) {
			const int ak = step_div_d(k, g.nz);
			z -= ak * g.bz; y -= ak * g.byz; x -= ak * g.bxz; k -= ak * g.nz;
		}
		const double ys = y * g.ysp;
		if (fabs(ys) < lim) {
			int j = step_int_d(ys);
			if (j < 0 || j >= g.ny) {
				const int aj = step_div_d(j, g.ny);
				y -= aj * g.by; x -= aj * g.bxy; j -= aj * g.ny;
			}
			const double xs = x * g.xsp;
			if (fabs(xs) < lim) {
				int i = step_int_d(xs);
				if (i < 0 || i >= g.nx) {
					const int ai = step_div_d(i, g.nx);
					x -= ai * g.bx; i -= ai * g.nx;
				}
				j += g.ey; k += g.ez;
				b = i + g.nx * (j + g.oy * k);
			}
		}
	}
	blk[t] = b;
	double rr = r ? r[t] : 0.5;
	rem[t] = dbl4{x, y, z, rr};
	if (b >= 0) {
		slot[t] = atomicAdd(&cnt[b], 1);
		if (r && rr > 0) atomicMax(maxr_bits, (unsigned long long)__double_as_longlong(rr));
	}
}

// Block arrays of a periodic container with its image blocks, assembled on
// the device: the primary blocks' particles as binned, the image particles
// (built on the host in creation order, `rank` = position within its block)
// after them in block order.  cnt[b] = particles of block b (primary counts
// from the binning, image counts scattered), then an exclusive scan gives
// the block starts.
__global__ void k_prd_counts(int nblocks, const int *__restrict__ bstart, int *__restrict__ cnt) {
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b > nblocks) return;
	cnt[b] = b < nblocks ? bstart[b + 1] - bstart[b] : 0;
}
__global__ void k_prd_img_counts(int n, const int2 *__restrict__ bc, int *__restrict__ cnt) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const int2 q = bc[t];
	cnt[q.x] = q.y;
}
__global__ void k_prd_copy_primary(int nc, const int *__restrict__ cellblk, const int *__restrict__ bstart,
				   const dbl4 *__restrict__ P, const int *__restrict__ pid, const int *__restrict__ ball,
				   dbl4 *__restrict__ Pall, int *__restrict__ pidall) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= nc) return;
	const int b = cellblk[c];
	const int d = ball[b] + (c - bstart[b]);
	Pall[d] = P[c];
	pidall[d] = pid[c];
}
__global__ void k_prd_copy_images(int n, const int *__restrict__ reg, const int *__restrict__ rank,
				  const dbl4 *__restrict__ ip, const int *__restrict__ iid, const int *__restrict__ ball,
				  dbl4 *__restrict__ Pall, int *__restrict__ pidall) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const int d = ball[reg[t]] + rank[t];
	Pall[d] = ip[t];
	pidall[d] = iid[t];
}

// check_duplicate (src/common.cc:14-20), which the periodic containers'
// put() applies against every particle already in the block
// (src/container_prd.cc:94-112): flags a pair closer than 1e-5.
__global__ void k_duplicates(int nblocks, const dbl4 *__restrict__ P, const int *__restrict__ cellblk,
			     const int *__restrict__ bstart, int *__restrict__ flag) {
	int d = blockIdx.x * blockDim.x + threadIdx.x;
	if (d >= bstart[nblocks]) return;
	const dbl4 a = P[d];
	for (int u = bstart[cellblk[d]]; u < d; u++) {
		const dbl4 q = P[u];
		const double dx = q.x - a.x, dy = q.y - a.y, dz = q.z - a.z;
		if (dx * dx + dy * dy + dz * dz < 1e-10) { atomicOr(flag, 1); return; }
	}
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) p[t] = v;
}
__global__ void k_fill_u8(unsigned char *p, int n, unsigned char v) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) p[t] = v;
}

// Periodic containers' replay bookkeeping.  The image blocks (rows outside
// [ey, wy) x [ez, wz)) that some cell visited, with their first-visit keys,
// in any order (the host sorts them by key); entries beyond cap are only
// counted.
__global__ void k_prd_visited(int nb, const unsigned long long *__restrict__ fv, int nx, int oy, int ey, int wy,
			      int ez, int wz, unsigned long long *__restrict__ keys, int *__restrict__ blks, int cap,
			      int *__restrict__ count) {
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nb) return;
	const unsigned long long f = fv[b];
	if (f == ~0ull) return;
	const int j = (b / nx) % oy, k = b / (nx * oy);
	if (j >= ey && j < wy && k >= ez && k < wz) return;   // a primary block
	const int i = atomicAdd(count, 1);
	if (i < cap) { keys[i] = f; blks[i] = b; }
}
// The cells that read an image block in any pass so far (acc |= touch), as
// a list in any order.
__global__ void k_prd_touched(int nc, const unsigned char *__restrict__ touch, unsigned char *__restrict__ acc,
			      int *__restrict__ list, int *__restrict__ count) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= nc) return;
	const unsigned char a = acc[c] | touch[c];
	acc[c] = a;
	if (a) list[atomicAdd(count, 1)] = c;
}
// *flag = 1 if a and b differ anywhere.
__global__ void k_differ_u64(int n, const unsigned long long *__restrict__ a, const unsigned long long *__restrict__ b,
			     int *__restrict__ flag) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n && a[t] != b[t]) *flag = 1;
}

// Exclusive scan, tiles of 1024 ints by 256 threads; tile totals to part.
__global__ void k_scan_tiles(const int *__restrict__ in, int *__restrict__ out, int n, int *__restrict__ part) {
	__shared__ int sh[256];
	int base = blockIdx.x * 1024 + threadIdx.x * 4;
	int v[4], s = 0;
	for (int t = 0; t < 4; t++) { v[t] = base + t < n ? in[base + t] : 0; s += v[t]; }
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int off = 1; off < 256; off <<= 1) {
		int a = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
		__syncthreads();
		sh[threadIdx.x] += a;
		__syncthreads();
	}
	int run = sh[threadIdx.x] - s;
	for (int t = 0; t < 4; t++) { if (base + t < n) out[base + t] = run; run += v[t]; }
	if (threadIdx.x == 255) part[blockIdx.x] = sh[255];
}
__global__ void k_scan_add(int *__restrict__ out, int n, const int *__restrict__ part) {
	int t = blockIdx.x * 1024 + threadIdx.x * 4;
	int a = part[blockIdx.x];
	for (int u = 0; u < 4; u++) if (t + u < n) out[t + u] += a;
}

// Unordered placement into block segments.
__global__ void k_place(int n, const int *__restrict__ blk, const int *__restrict__ slot,
			const int *__restrict__ bstart, int *__restrict__ tmp) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	int b = blk[t];
	if (b >= 0) tmp[bstart[b] + slot[t]] = t;
}

// Stable order within each block: rank = number of particles of the same
// block that were put earlier (lower input index).  Writes the sorted
// particle, its id, its input index and its block.
__global__ void k_rank(int n, const int *__restrict__ blk, const int *__restrict__ bstart,
		       const int *__restrict__ tmp, const dbl4 *__restrict__ rem, const int *__restrict__ ids,
		       dbl4 *__restrict__ P, int *__restrict__ pid, int *__restrict__ src, int *__restrict__ cellblk) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	int b = blk[t];
	if (b < 0) return;
	int s = bstart[b], e = bstart[b + 1], rank = 0;
	for (int u = s; u < e; u++) rank += tmp[u] < t;
	int d = s + rank;
	P[d] = rem[t];
	pid[d] = ids[t];
	src[d] = t;
	cellblk[d] = b;
}

// Batches: the same two steps with the cell count read on the device
// (*ncells <= n; entries from *ncells on count zero faces).
// A cell whose ids did not fit the arena (start -1; the compute reports a
// memory error) contributes no entries, so the CSR never outgrows the arena.
__global__ void k_face_counts_dev(int n, const int *__restrict__ ncells, const int *__restrict__ status,
				  const int *__restrict__ nfaces, const int *__restrict__ start, int *__restrict__ cnt) {
	int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c > n) return;
	cnt[c] = (c < *ncells && (status[c] & ST_EXISTS) && start[c] >= 0) ? nfaces[c] : 0;
}
__global__ void k_gather_nbrs_dev(int n, const int *__restrict__ ncells, const int *__restrict__ status,
				  const int *__restrict__ nfaces, const int *__restrict__ start,
				  const int *__restrict__ arena, const int *__restrict__ off, int *__restrict__ csr) {
	int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= n || c >= *ncells) return;
	if (!(status[c] & ST_EXISTS) || start[c] < 0) return;
	int s = start[c], o = off[c], k = nfaces[c];
	for (int t = 0; t < k; t++) csr[o + t] = arena[s + t];
}

// Loop-order CSR of the neighbour lists: out[off[c] ..] = arena[start[c] ..].
__global__ void k_gather_nbrs(int ncells, const int *__restrict__ status, const int *__restrict__ nfaces,
			      const int *__restrict__ start, const int *__restrict__ arena, const int *__restrict__ off,
			      int *__restrict__ csr) {
	int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= ncells) return;
	if (!(status[c] & ST_EXISTS) || start[c] < 0) return;
	int s = start[c], o = off[c], k = nfaces[c];
	for (int t = 0; t < k; t++) csr[o + t] = arena[s + t];
}
__global__ void k_face_counts(int ncells, const int *__restrict__ status, const int *__restrict__ nfaces,
			      const int *__restrict__ start, int *__restrict__ cnt) {
	int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c > ncells) return;
	cnt[c] = (c < ncells && (status[c] & ST_EXISTS) && start[c] >= 0) ? nfaces[c] : 0;
}

}  // namespace vgpu
