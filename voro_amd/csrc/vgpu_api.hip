// C-ABI of the engine (include/voro_gpu.h): container handles, particle
// upload, the binning + cell pipeline, and result read-back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/voro_gpu.h"
#include "grid_host.h"
#include "learn.h"
#include "prd_host.h"
#include "vgpu_bin.h"
#include "wcell_dev.h"

using namespace vgpu;

namespace {

#define HIPCHK(x)                                                   \
	do {                                                        \
		hipError_t e_ = (x);                                \
		if (e_ != hipSuccess) { return VGPU_HIP_ERROR; }    \
	} while (0)

// Grow-only device buffer.
struct DBuf {
	void *p = nullptr;
	size_t n = 0;
	DBuf() = default;
	DBuf(const DBuf &) = delete;
	DBuf &operator=(const DBuf &) = delete;
	~DBuf() { release(); }
	int ensure(size_t bytes) {
		if (bytes <= n) return VGPU_OK;
		if (p) (void)hipFree(p);
		p = nullptr; n = 0;
		if (bytes == 0) return VGPU_OK;
		if (hipMalloc(&p, bytes) != hipSuccess) { p = nullptr; return VGPU_MEMORY_ERROR; }
		n = bytes;
		return VGPU_OK;
	}
	void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
	template<class T> T *as() const { return (T *)p; }
};

// Grow-only pinned host buffer (device-to-host copies at full speed, no
// per-call allocation or zeroing).
struct HBuf {
	void *p = nullptr;
	size_t n = 0;
	HBuf() = default;
	HBuf(const HBuf &) = delete;
	HBuf &operator=(const HBuf &) = delete;
	~HBuf() { release(); }
	int ensure(size_t bytes) {
		if (bytes <= n) return VGPU_OK;
		release();
		if (bytes == 0) return VGPU_OK;
		if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) { p = nullptr; return VGPU_MEMORY_ERROR; }
		n = bytes;
		return VGPU_OK;
	}
	void release() { if (p) (void)hipHostFree(p); p = nullptr; n = 0; }
	void swap(HBuf &o) { std::swap(p, o.p); std::swap(n, o.n); }
	template<class T> T *as() const { return (T *)p; }
};

// Lane-interleaved workspace slabs of one tier, with a front and back pad.
struct Workspace {
	DBuf buf;
	size_t slab = 0;
	int waves = 0;
	unsigned stamp_base = 0;
	static const size_t PAD = 1 << 20;
	int ensure(size_t slab_bytes, int nwaves) {
		if (slab_bytes == slab && nwaves <= waves) return VGPU_OK;
		buf.release();
		int st = buf.ensure(2 * PAD + slab_bytes * (size_t)nwaves);
		if (st) return st;
		// Block-mask stamps must start out invalid: zero everything once.
		if (hipMemset(buf.p, 0, buf.n) != hipSuccess) return VGPU_HIP_ERROR;
		slab = slab_bytes;
		waves = nwaves;
		stamp_base = 0;
		return VGPU_OK;
	}
	char *base() const { return (char *)buf.p + PAD; }
};

__global__ void k_fill_int(int *p, int n, int v) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) p[t] = v;
}

}  // namespace

struct vgpu_container {
	int kind = 0;
	bool poly = false;   // radical (power) variant: container_poly, container_periodic_poly
	bool prd = false;    // periodic lattice: container_periodic(_poly)
	PrdGeom pg;          // periodic geometry and unit cell
	int nblocks = 0;     // nx*ny*nz, or nx*oy*oz for a periodic container
	double geom[6] = {0, 0, 0, 0, 0, 0};
	int grid[3] = {1, 1, 1};
	int per[3] = {0, 0, 0};
	RectGeom g;
	double max_len_sq = 0;
	std::vector<double> walls;
	std::vector<int> h_ids;
	std::vector<double> h_xyz, h_r;
	bool have_dev_input = false;
	int n_in = 0;
	int ncells = 0;
	double max_radius = 0;
	bool computed = false;
	bool binned = false;     // block arrays current (binning done since the last put)
	unsigned last_flags = 0;
	int fast_waves = 0;
	int fast_occ = 2;
	int first_tier = 0;      // 0: wavefront cells, 1: order-3 cells, one per lane
	int wave_waves = 0;
	int wave_occ = 5;
	hipEvent_t ev[5];
	hipEvent_t ev_list = nullptr;   // learned tier: lists gathered, cell kernel starts
	double gather_ms = 0;
	bool events = false;
	double timings[4] = {0, 0, 0, 0};

	DBuf in_ids, in_xyz, in_r;
	DBuf blk, slot, rem, cnt, bstart, tmp, part, P, pid, src, cellblk;
	DBuf d_mrad, d_wl, d_walls, d_maxr, d_counters;
	DBuf vol, cent, nfaces, status, nbr_start, arena, csr_off, csr, ctr;
	DBuf retry0, retry1, retry2, retry3, retry4;
	DBuf fq_xyz, fq_found, fq_pid, fq_pos, fq_retry, fq_retry2;   // find_voronoi_cell queries
	DBuf fq_r, fq_vol, fq_nbr;                                   // ghost cells
	DBuf fq_touch, fq_list, img_okb;                             // periodic queries: image reads, reruns
	DBuf st_planes, st_exit, st_tier;
	// learned path
	bool have_pred = false;
	int pred_with_r = 0;
	float pred_margin = 0;
	int pred_route = 64;
	DBuf d_pred, d_feat, d_cfeat, d_depth, d_cost, d_order, d_hist, lcnt, bmax, boff, list_a;
	long long last_fallbacks = 0, last_entries = 0, last_routed = 0;
	Workspace ws_fast, ws_wide, ws_huge, ws_wave, ws_wave4;
	long long nbr_total = 0;
	bool csr_valid = false;
	// periodic containers: unit cell, block arrays with image blocks, first
	// visits, cells that read an image block
	DBuf d_unit, Pall, pidall, ball, fv, touch, d_flag;
	DBuf img_p, img_id, img_reg, img_rank, img_bc, img_cnt;   // image particles for the device assembly
	std::vector<int> h_rank;                                   // per-block creation counters (host, kept zero)
	// host copies of the primary blocks (the image creation reads them), the
	// first-visit keys of two passes and the image-read flags (pinned)
	HBuf hP, hpid, hcs, h_vkeys, h_vblks;
	DBuf d_prev, d_vkeys, d_vblks, d_tacc;   // device side of the replay (keys of the last pass, visited image blocks, touched cells)
	std::unique_ptr<ImageSet> h_images;   // the image blocks being built (storage kept)
	ImageSet &image_set(const dbl4 *P_, const int *pid_, const int *cstart_) {
		if (!h_images) h_images.reset(new ImageSet(pg, P_, pid_, cstart_));
		else h_images->reset(P_, pid_, cstart_);
		return *h_images;
	}
	int image_passes = 0;
	// the image blocks the reference would hold after the last compute and
	// queries since the binning, in their order of creation
	std::vector<int> prd_order;
	// neighbour arena size that held the last compute (grown when it overflows)
	long long acap_hint = 0;
	// polyhedron statistics (VGPU_F_GEOMETRY): the cell tables kept by the
	// tiers, then the loop-order statistics computed from them
	DBuf geo_top, geo_vstart, geo_np, geo_estart, geo_esz, geo_xyz, geo_nu, geo_eoff, geo_ed, geo_rel;
	DBuf go_voff, go_eoff, go_fvoff, go_nv, go_ne, go_nfv;
	DBuf go_vert, go_ord, go_edges, go_erel, go_farea, go_fperim, go_normal, go_forder, go_fverts;
	DBuf go_surface, go_edist, go_nedges, go_maxr2;
	long long geo_vcap_hint = 0, geo_ecap_hint = 0;
	long long geo_nv = 0, geo_ne = 0, geo_nfv = 0, geo_nf = 0;
	bool geo_valid = false;
};

namespace {

int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

// Most wavefronts a retry tier spreads its list over (a wide-tier slab is
// ~16 MB, a huge-tier slab ~64 MB; the slabs are allocated as lists need
// them).  Retried cells are few and expensive, so the list is spread thin.
enum { RETRY_WAVES_WIDE = 512, RETRY_WAVES_HUGE = 32 };

// Exclusive scan of n+1 ints (in[n] must be 0): out[n] = total.
// The partial-sum buffer is grown here unless `no_alloc` (graph capture:
// the caller has sized it with scan_part_bytes).
size_t scan_part_bytes(int n1) {
	const int tiles = nblk(n1, 1024);
	return 2 * ((size_t)tiles + (size_t)nblk(tiles, 1024) + 8) * sizeof(int);
}
int scan_excl(DBuf &part, int *in, int *out, int n1, hipStream_t s, bool no_alloc = false) {
	int tiles = nblk(n1, 1024);
	// part buffers for up to three levels
	size_t need = (size_t)tiles + (size_t)nblk(tiles, 1024) + 8;
	if (no_alloc) {
		if (part.n < 2 * need * sizeof(int)) return VGPU_INTERNAL_ERROR;
	} else {
		int st = part.ensure(2 * need * sizeof(int));
		if (st) return st;
	}
	int *p1 = part.as<int>();
	int *p1s = p1 + need;
	hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(256), 0, s, in, out, n1, p1);
	if (tiles > 1) {
		if (tiles > 1024 * 1024) return VGPU_INVALID_ARGUMENT;
		int t2 = nblk(tiles, 1024);
		int *p2 = p1 + tiles;
		int *p2s = p1s + tiles;
		hipLaunchKernelGGL(k_scan_tiles, dim3(t2), dim3(256), 0, s, p1, p1s, tiles, p2);
		if (t2 > 1) {
			int *p3 = p2 + t2;  // t2 <= 1024: one tile
			hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, s, p2, p2s, t2, p3);
			hipLaunchKernelGGL(k_scan_add, dim3(t2), dim3(256), 0, s, p1s, tiles, p2s);
		}
		hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(256), 0, s, out, n1, p1s);
	}
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

RectView make_view(vgpu_container *h) {
	RectView v;
	const RectGeom &g = h->g;
	v.ax = g.ax; v.bx = g.bx; v.ay = g.ay; v.by = g.by; v.az = g.az; v.bz = g.bz;
	v.nx = g.nx; v.ny = g.ny; v.nz = g.nz;
	v.gy = h->prd ? h->pg.oy : g.ny;
	v.nxy = g.nx * v.gy;
	v.boxx = g.boxx; v.boxy = g.boxy; v.boxz = g.boxz; v.xsp = g.xsp; v.ysp = g.ysp; v.zsp = g.zsp;
	v.xper = g.xper; v.yper = g.yper; v.zper = g.zper;
	v.hx = g.xper ? 2 * g.nx + 1 : g.nx; v.hy = g.yper ? 2 * g.ny + 1 : g.ny; v.hz = g.zper ? 2 * g.nz + 1 : g.nz;
	if (h->prd) { v.hx = 2 * g.nx + 1; v.hy = 2 * h->pg.ey + 1; v.hz = 2 * h->pg.ez + 1; }
	v.bxsq = g.boxx * g.boxx + g.boxy * g.boxy + g.boxz * g.boxz;
	v.max_len_sq = h->max_len_sq;
	v.max_radius = h->max_radius;
	v.P = h->P.as<dbl4>(); v.pid = h->pid.as<int>(); v.bstart = h->bstart.as<int>(); v.cstart = h->bstart.as<int>();
	v.mrad = h->d_mrad.as<double>(); v.wl = h->d_wl.as<unsigned>();
	v.walls = h->d_walls.as<double>(); v.nwalls = (int)(h->walls.size() / VGPU_WALL_STRIDE);
	const PrdGeom &pg = h->pg;
	v.bxy = h->prd ? pg.bxy : 0; v.bxz = h->prd ? pg.bxz : 0; v.byz = h->prd ? pg.byz : 0;
	v.ey = pg.ey; v.ez = pg.ez; v.wy = pg.wy; v.wz = pg.wz; v.oy = pg.oy; v.oz = pg.oz;
	v.unit = h->d_unit.as<UnitVert>(); v.unit_p = h->prd ? (int)pg.unit.size() : 0;
	v.fv = nullptr;
	v.img_ok = nullptr;
	v.boff = 0;
	return v;
}

template<class T_>
int launch_tier(vgpu_container *h, Workspace &ws, int waves, const RectView &v, int ncells, const int *list,
		const int *list_count, CellOut out, hipStream_t s) {
	typedef typename TierCell<T_>::type::L L;
	int st = ws.ensure(L::bytes, waves);
	if (st) return st;
	if ((unsigned long long)ws.stamp_base + (unsigned long long)ncells + 2 >= 0xffffffffull) {
		HIPCHK(hipMemset(ws.buf.p, 0, ws.buf.n));
		ws.stamp_base = 0;
	}
	int *ctr = h->d_counters.as<int>();
	int per_wave = 64;
	if (list_count) {
		// a retry list: spread its (few, expensive) cells over all wavefronts
		int cnt = 0;
		HIPCHK(hipMemcpyAsync(&cnt, list_count, 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (cnt == 0) return VGPU_OK;
		per_wave = std::max(1, std::min(64, (cnt + waves - 1) / waves));
		waves = std::max(1, std::min(waves, (cnt + per_wave - 1) / per_wave));
	}
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr, 8, 0);
	const bool poly = h->poly;
	const int ti = std::is_same<T_, Tier3>::value ? 0 : (T_::MAXV == TierWide::MAXV ? 1 : 2);
	if (h->prd) {
		if (launch_cells_prd(ti, poly, waves, s, v, ncells, list, list_count, h->cellblk.as<int>(), ctr, ws.base(),
				     ws.slab, ws.stamp_base, per_wave, out))
			return VGPU_HIP_ERROR;
	} else {
		cells_launcher fn = ti == 0 ? launch_cells_fast3 : (ti == 1 ? launch_cells_wide : launch_cells_huge);
		if (fn(poly, h->fast_occ, waves, s, v, ncells, list, list_count, h->cellblk.as<int>(), ctr, ws.base(), ws.slab,
		       ws.stamp_base, per_wave, out))
			return VGPU_HIP_ERROR;
	}
	ws.stamp_base += (unsigned)ncells + 1u;
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

// Persistent wavefronts of a batch's order-3 per-lane tier when it takes
// the wavefront-cell tier's leftovers (few cells).
const int BATCH_LEFTOVER_WAVES = 512;

int default_fast_waves(int occ) {
	int dev = 0, cus = 256;
	if (hipGetDevice(&dev) == hipSuccess) {
		hipDeviceProp_t pr;
		if (hipGetDeviceProperties(&pr, dev) == hipSuccess && pr.multiProcessorCount > 0) cus = pr.multiProcessorCount;
	}
	return cus * 4 * occ;
}

// A wavefront-cell tier (k vertices per lane) over every cell (list ==
// nullptr) or over a device list of cells whose count is at *list_count;
// leftovers go to out.retry.
int launch_wave(vgpu_container *h, int k, const RectView &v, int nc, const int *list, const int *list_count,
		CellOut out, hipStream_t s) {
	// periodic containers: compiled for 4, 5 and 6 waves per SIMD
	const int occ = k == 2 ? VG_K2_OCC : (k > 1 ? 2 : (h->prd ? std::min(6, std::max(4, h->wave_occ)) : h->wave_occ));
	int waves = (h->wave_waves > 0 && k == 1) ? h->wave_waves : default_fast_waves(occ);
	int cnt = nc;
	if (list_count) {
		HIPCHK(hipMemcpyAsync(&cnt, list_count, 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (cnt == 0) return VGPU_OK;
	}
	// four cells per counter increment, one for a list too short to fill
	// every wavefront that way
	const int chunk = cnt >= 4 * waves ? 4 : 1;
	waves = std::max(1, std::min(waves, (cnt + chunk - 1) / chunk));
	// the 254-vertex tier's slabs hold a wider block-search window
	Workspace &ws = k == 4 ? h->ws_wave4 : h->ws_wave;
	int st = ws.ensure(k == 4 ? WLayoutT<4>::bytes : WLayout::bytes, waves);
	if (st) return st;
	if ((unsigned long long)ws.stamp_base + (unsigned long long)nc + 2 >= 0xffffffffull) {
		HIPCHK(hipMemset(ws.buf.p, 0, ws.buf.n));
		ws.stamp_base = 0;
	}
	int *ctr = h->d_counters.as<int>();
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr, 8, 0);
	if (launch_cells_wave(k, h->poly, h->prd, occ, waves, s, v, nc, list, list_count, h->cellblk.as<int>(), ctr,
			      ws.base(), ws.stamp_base, chunk, out))
		return VGPU_HIP_ERROR;
	ws.stamp_base += (unsigned)nc + 1u;
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

// The predictor's block features of a container into h->d_feat; returns
// the number of blocks (samples) through *nsamp.  A periodic container's
// samples are its whole block grid (image rows included, zero features
// there), so that a cell's block index is its sample.
int container_features(vgpu_container *h, int nc, int *nsamp, hipStream_t s) {
	using namespace vgpu::learn;
	const int nreal = h->g.nx * h->g.ny * h->g.nz;
	const int nb = h->prd ? h->nblocks : nreal;
	int st;
	if ((st = h->d_feat.ensure((size_t)nb * IN * 4))) return st;
	const float inv_mean = (float)nreal / (float)std::max(nc, 1);
	if (h->prd) st = block_features_dev(h->g.nx, h->g.ny, h->g.nz, h->bstart.as<int>(), inv_mean, h->d_feat.as<float>(),
					    s, h->pg.oy, h->pg.oz, h->pg.ey, h->pg.ez, true);
	else st = block_features_dev(h->g.nx, h->g.ny, h->g.nz, h->bstart.as<int>(), inv_mean, h->d_feat.as<float>(), s);
	*nsamp = nb;
	return st;
}

// The learned tier: features, depth prediction, list gathering, the
// list-driven cell kernel.  Cells it does not finish go to retry0.  v may
// carry a periodic container's first-visit record (v.fv): the gather reads
// a copy without it, the list kernel records the visits.
int run_learned(vgpu_container *h, const RectView &v, int nc, int waves, CellOut out, hipStream_t s) {
	using namespace vgpu::learn;
	int st, nxyz = 0;
	int *ctr = h->ctr.as<int>();
	if ((st = container_features(h, nc, &nxyz, s))) return st;
	RectView vg = v;
	vg.fv = nullptr;
	const float *X = h->d_feat.as<float>();
	int ns = nxyz;
	if (h->pred_with_r) {
		if ((st = h->d_cfeat.ensure((size_t)nc * IN * 4))) return st;
		const float inv = h->max_radius > 0 ? (float)(1.0 / h->max_radius) : 0.f;
		if ((st = cell_features_dev(nc, h->cellblk.as<int>(), h->P.as<double>() + 3, inv, h->d_feat.as<float>(),
					    h->d_cfeat.as<float>(), s)))
			return st;
		X = h->d_cfeat.as<float>();
		ns = nc;
	}
	if ((st = h->d_depth.ensure((size_t)ns)) || (st = h->d_cost.ensure((size_t)ns))) return st;
	if ((st = mlp_depth_dev(h->d_pred.as<float>(), X, ns, h->pred_margin, h->pred_route, h->d_depth.as<unsigned char>(),
				h->d_cost.as<unsigned char>(), s)))
		return st;
	const int nb = nblk(nc, 64);
	const bool csr = h->first_tier == 0;   // wavefront cells: each cell's list contiguous
	if ((st = h->lcnt.ensure((size_t)(nc + 1) * 4)) || (st = h->bmax.ensure((size_t)(nb + 1) * 4)) ||
	    (st = h->boff.ensure((size_t)(std::max(nb, nc) + 1) * 4)) || (st = h->d_order.ensure((size_t)nc * 4)) ||
	    (st = h->d_hist.ensure(64 * 4)))
		return st;
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, h->d_hist.as<int>(), 64, 0);
	if (launch_list_order(s, nc, h->cellblk.as<int>(), h->d_cost.as<unsigned char>(), h->pred_with_r,
			      h->d_hist.as<int>(), h->d_order.as<int>()))
		return VGPU_HIP_ERROR;
	if (launch_list_gather(h->prd, s, vg, nc, h->cellblk.as<int>(), h->d_order.as<int>(), h->d_depth.as<unsigned char>(),
			       h->pred_with_r, h->lcnt.as<int>(), h->bmax.as<int>()))
		return VGPU_HIP_ERROR;
	size_t ent;
	if (csr) {
		// boff[pos] = entries before schedule position pos
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, h->lcnt.as<int>() + nc, 1, 0);
		if ((st = scan_excl(h->part, h->lcnt.as<int>(), h->boff.as<int>(), nc + 1, s))) return st;
		int tot = 0;
		HIPCHK(hipMemcpyAsync(&tot, h->boff.as<int>() + nc, 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (tot < 0) return VGPU_MEMORY_ERROR;
		ent = (size_t)std::max(tot, 1);
		h->last_entries = tot;
	} else {
		// boff[b] = rows before batch b of 64 positions (the longest list sets a batch's rows)
		if ((st = scan_excl(h->part, h->bmax.as<int>(), h->boff.as<int>(), nb + 1, s))) return st;
		int rows = 0;
		HIPCHK(hipMemcpyAsync(&rows, h->boff.as<int>() + nb, 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		ent = (size_t)std::max(rows, 1) * 64;
		h->last_entries = (long long)rows * 64;
	}
	if ((st = h->list_a.ensure(ent * sizeof(ListRec)))) return st;
	if (launch_list_fill(h->prd, s, vg, nc, h->cellblk.as<int>(), h->d_order.as<int>(), h->d_depth.as<unsigned char>(),
			     h->pred_with_r, h->boff.as<int>(), csr, h->list_a.as<ListRec>()))
		return VGPU_HIP_ERROR;
	if (!h->ev_list) HIPCHK(hipEventCreate(&h->ev_list));
	HIPCHK(hipEventRecord(h->ev_list, s));
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, h->d_counters.as<int>(), 8, 0);
	out.tier = 0; out.retry = h->retry0.as<int>(); out.retry_count = ctr + 4;
	if (h->first_tier == 0) {
		// one cell per wavefront
		// the list kernel at the wavefront tier's register budget (4, 5 or
		// 6 waves per SIMD)
		const int locc = h->wave_occ >= 6 ? 6 : (h->wave_occ == 5 ? 5 : 4);
		int w = h->wave_waves > 0 ? h->wave_waves : default_fast_waves(locc);
		const int chunk = nc >= 4 * w ? 4 : 1;
		w = std::max(1, std::min(w, (nc + chunk - 1) / chunk));
		Workspace &ws = h->ws_wave;
		if ((st = ws.ensure(WLayout::bytes, w))) return st;
		if ((unsigned long long)ws.stamp_base + (unsigned long long)nc + 2 >= 0xffffffffull) {
			HIPCHK(hipMemset(ws.buf.p, 0, ws.buf.n));
			ws.stamp_base = 0;
		}
		if (launch_cells_wlist(h->poly, h->prd, locc, w, chunk, s, v, nc, h->cellblk.as<int>(), h->d_order.as<int>(),
				       h->lcnt.as<int>(), h->boff.as<int>(), h->list_a.as<ListRec>(),
				       h->d_counters.as<int>(), ws.base(), ws.stamp_base, ctr + 5, ctr + 9, out))
			return VGPU_HIP_ERROR;
		ws.stamp_base += (unsigned)nc + 1u;
	} else {
		if ((st = h->ws_fast.ensure(VCell3::L::bytes, waves))) return st;
		if (launch_cells_list(h->poly, h->prd, waves, s, v, nc, h->cellblk.as<int>(), h->d_order.as<int>(),
				      h->lcnt.as<int>(), h->boff.as<int>(), h->list_a.as<ListRec>(),
				      h->d_counters.as<int>(), h->ws_fast.base(), ctr + 5, ctr + 9,
				      out))
			return VGPU_HIP_ERROR;
	}
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

// The plain tiers over every cell (list == nullptr) or over a device list
// of cells whose count is at *list_count: order-3, then wide, then huge
// over each tier's overflow.
// from2: the list's cells overflowed the 64-vertex wavefront cell already
// (the learned tier's leftovers): start with the 128-vertex one.
int run_tiers(vgpu_container *h, const RectView &v, int nc, int fw, const int *list, const int *list_count,
	      CellOut out, hipStream_t s, bool mark_ev2, int tier0 = 0, bool from2 = false) {
	int st;
	int *ctr = h->ctr.as<int>();
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr + 1, 2, 0);
	if (h->first_tier == 0) {
		// one cell per wavefront, up to 64 vertices; leftovers (count at
		// ctr[6]) to the same with up to 128, theirs (ctr[7]) up to 254, and
		// theirs (ctr[1]) to the wide per-lane tier
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr + 6, 3, 0);
		if (from2) {
			out.tier = tier0; out.retry = h->retry4.as<int>(); out.retry_count = ctr + 7;
			if ((st = launch_wave(h, 2, v, nc, list, list_count, out, s))) return st;
			if (mark_ev2) HIPCHK(hipEventRecord(h->ev[2], s));
		} else {
			out.tier = tier0; out.retry = h->retry3.as<int>(); out.retry_count = ctr + 6;
			if ((st = launch_wave(h, 1, v, nc, list, list_count, out, s))) return st;
			if (mark_ev2) HIPCHK(hipEventRecord(h->ev[2], s));
			out.tier = 1; out.retry = h->retry4.as<int>(); out.retry_count = ctr + 7;
			if ((st = launch_wave(h, 2, v, nc, h->retry3.as<int>(), ctr + 6, out, s))) return st;
		}
		// the 254-vertex tier's leftovers go to the wide per-lane tier: the
		// order-3 per-lane tier's capacities are all smaller, so it finishes
		// none of them
		out.retry = h->retry1.as<int>(); out.retry_count = ctr + 1;
		if ((st = launch_wave(h, 4, v, nc, h->retry4.as<int>(), ctr + 7, out, s))) return st;
	} else {
		out.tier = tier0; out.retry = h->retry1.as<int>(); out.retry_count = ctr + 1;
		if ((st = launch_tier<Tier3>(h, h->ws_fast, fw, v, nc, list, list_count, out, s))) return st;
		if (mark_ev2) HIPCHK(hipEventRecord(h->ev[2], s));
	}
	out.tier = 2; out.retry = h->retry2.as<int>(); out.retry_count = ctr + 2;
	if ((st = launch_tier<TierWide>(h, h->ws_wide, RETRY_WAVES_WIDE, v, nc, h->retry1.as<int>(), ctr + 1, out, s))) return st;
	out.tier = 3; out.retry = nullptr; out.retry_count = nullptr;
	return launch_tier<TierHuge>(h, h->ws_huge, RETRY_WAVES_HUGE, v, nc, h->retry2.as<int>(), ctr + 2, out, s);
}

// Uploads the block arrays of every block (primary and image) and points
// the view at them.
int upload_blocks(vgpu_container *h, const ImageSet &im, int nc, RectView &v, hipStream_t s) {
	const int nb = h->nblocks;
	const int ni = (int)im.reg.size();
	// position of every image particle within its block (creation order),
	// and the particle count of every image block
	if ((int)h->h_rank.size() != nb) h->h_rank.assign(nb, 0);
	std::vector<int> rank(std::max(ni, 1));
	std::vector<int2> bc;
	for (int t = 0; t < ni; t++) {
		const int r = im.reg[t];
		if (h->h_rank[r] == 0) bc.push_back(int2{r, 0});
		rank[t] = h->h_rank[r]++;
	}
	for (int2 &q : bc) { q.y = h->h_rank[q.x]; h->h_rank[q.x] = 0; }
	const int nbc = (int)bc.size();
	int st;
	if ((st = h->img_p.ensure((size_t)std::max(ni, 1) * sizeof(dbl4))) || (st = h->img_id.ensure((size_t)std::max(ni, 1) * 4)) ||
	    (st = h->img_reg.ensure((size_t)std::max(ni, 1) * 4)) || (st = h->img_rank.ensure((size_t)std::max(ni, 1) * 4)) ||
	    (st = h->img_bc.ensure((size_t)std::max(nbc, 1) * sizeof(int2))) ||
	    (st = h->img_cnt.ensure((size_t)(nb + 1) * 4)) || (st = h->Pall.ensure((size_t)std::max(nc + ni, 1) * sizeof(dbl4))) ||
	    (st = h->pidall.ensure((size_t)std::max(nc + ni, 1) * 4)) || (st = h->ball.ensure((size_t)(nb + 1) * 4)))
		return st;
	if (ni) {
		HIPCHK(hipMemcpyAsync(h->img_p.p, im.ip.data(), (size_t)ni * sizeof(dbl4), hipMemcpyHostToDevice, s));
		HIPCHK(hipMemcpyAsync(h->img_id.p, im.iid.data(), (size_t)ni * 4, hipMemcpyHostToDevice, s));
		HIPCHK(hipMemcpyAsync(h->img_reg.p, im.reg.data(), (size_t)ni * 4, hipMemcpyHostToDevice, s));
		HIPCHK(hipMemcpyAsync(h->img_rank.p, rank.data(), (size_t)ni * 4, hipMemcpyHostToDevice, s));
	}
	if (nbc) HIPCHK(hipMemcpyAsync(h->img_bc.p, bc.data(), (size_t)nbc * sizeof(int2), hipMemcpyHostToDevice, s));
	int *cnt = h->img_cnt.as<int>();
	hipLaunchKernelGGL(k_prd_counts, dim3(nblk(nb + 1, 256)), dim3(256), 0, s, nb, h->bstart.as<int>(), cnt);
	if (nbc) hipLaunchKernelGGL(k_prd_img_counts, dim3(nblk(nbc, 256)), dim3(256), 0, s, nbc, h->img_bc.as<int2>(), cnt);
	if ((st = scan_excl(h->part, cnt, h->ball.as<int>(), nb + 1, s))) return st;
	if (nc) hipLaunchKernelGGL(k_prd_copy_primary, dim3(nblk(nc, 256)), dim3(256), 0, s, nc, h->cellblk.as<int>(),
				   h->bstart.as<int>(), h->P.as<dbl4>(), h->pid.as<int>(), h->ball.as<int>(),
				   h->Pall.as<dbl4>(), h->pidall.as<int>());
	if (ni) hipLaunchKernelGGL(k_prd_copy_images, dim3(nblk(ni, 256)), dim3(256), 0, s, ni, h->img_reg.as<int>(),
				   h->img_rank.as<int>(), h->img_p.as<dbl4>(), h->img_id.as<int>(), h->ball.as<int>(),
				   h->Pall.as<dbl4>(), h->pidall.as<int>());
	// the host vectors are read by the copies above
	HIPCHK(hipStreamSynchronize(s));
	v.P = h->Pall.as<dbl4>(); v.pid = h->pidall.as<int>(); v.bstart = h->ball.as<int>();
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

// The cells of a periodic container.  Pass 1 computes every cell against
// the image blocks near the primary domain, built in create_all_images
// order, and records, per block, the first visit over the loop.  The image blocks are then rebuilt by creating
// them in that order -- the reference's lazy creation order -- and the cells
// that read an image block are recomputed, recording first visits again.
// When the recorded order reproduces itself, every cell has been computed
// against the reference's block arrays.
//
// learned: the first pass takes the learned tier first (its list kernel
// records first visits with the plain path's keys), the rest the plain tiers.
int run_prd_cells(vgpu_container *h, RectView &v, CellOut out, int nc, int fw, hipStream_t s, bool learned) {
	const PrdGeom &g = h->pg;
	const int nb = h->nblocks;
	int st;
	if ((st = h->hP.ensure((size_t)std::max(nc, 1) * sizeof(dbl4))) || (st = h->hpid.ensure((size_t)std::max(nc, 1) * 4)) ||
	    (st = h->hcs.ensure((size_t)(nb + 1) * 4)) || (st = h->d_prev.ensure((size_t)nb * 8)) ||
	    (st = h->d_tacc.ensure((size_t)std::max(nc, 1))))
		return st;
	const dbl4 *hP = h->hP.as<dbl4>();
	const int *hpid = h->hpid.as<int>(), *hcs = h->hcs.as<int>();
	HIPCHK(hipMemcpyAsync(h->hP.p, h->P.p, (size_t)nc * sizeof(dbl4), hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(h->hpid.p, h->pid.p, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(h->hcs.p, h->bstart.p, (size_t)(nb + 1) * 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	{
		ImageSet &im = h->image_set(hP, hpid, hcs);
		for (int b : prd_margin_images(g, PRD_FIRST_MARGIN)) im.create_block(b);
		if ((st = upload_blocks(h, im, nc, v, s))) return st;
	}
	if ((st = h->fv.ensure((size_t)nb * 8)) || (st = h->touch.ensure((size_t)nc))) return st;
	unsigned long long *fv = h->fv.as<unsigned long long>();
	hipLaunchKernelGGL(k_fill_u64, dim3(nblk(nb, 256)), dim3(256), 0, s, fv, nb, ~0ull);
	hipLaunchKernelGGL(k_fill_u8, dim3(nblk(nc, 256)), dim3(256), 0, s, h->touch.as<unsigned char>(), nc, 0);
	v.fv = fv;
	out.touch = h->touch.as<unsigned char>();
	if (learned) {
		if ((st = run_learned(h, v, nc, fw, out, s))) return st;
		HIPCHK(hipEventRecord(h->ev[2], s));
		if ((st = run_tiers(h, v, nc, fw, h->retry0.as<int>(), h->ctr.as<int>() + 4, out, s, false, 1,
				    h->first_tier == 0)))
			return st;
	} else if ((st = run_tiers(h, v, nc, fw, nullptr, nullptr, out, s, true))) return st;
	// The bookkeeping between passes stays on the device: the cells touched
	// so far (list for the next pass), the visited image blocks with their
	// keys (the host sorts them into the creation order), and whether a
	// pass's keys equal the previous one's.
	int *ctr = h->ctr.as<int>();
	unsigned char *tacc = h->d_tacc.as<unsigned char>();
	hipLaunchKernelGGL(k_fill_u8, dim3(nblk(nc, 256)), dim3(256), 0, s, tacc, nc, 0);
	h->image_passes = 1;
	for (int pass = 0; pass < 8; pass++) {
		// ctr[4]: cells for the next pass, ctr[10]: visited image blocks,
		// ctr[11]: the keys differ (ctr[5] keeps the learned tier's fallbacks)
		int cnt[2] = {0, 0};
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr + 4, 1, 0);
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr + 10, 2, 0);
		hipLaunchKernelGGL(k_prd_touched, dim3(nblk(nc, 256)), dim3(256), 0, s, nc, h->touch.as<unsigned char>(), tacc,
				   h->retry0.as<int>(), ctr + 4);
		int cap = (int)(h->d_vblks.n / 4);
		hipLaunchKernelGGL(k_prd_visited, dim3(nblk(nb, 256)), dim3(256), 0, s, nb, fv, g.nx, g.oy, g.ey, g.wy, g.ez,
				   g.wz, h->d_vkeys.as<unsigned long long>(), h->d_vblks.as<int>(), cap, ctr + 10);
		HIPCHK(hipMemcpyAsync(&cnt[0], ctr + 4, 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipMemcpyAsync(&cnt[1], ctr + 10, 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (cnt[0] == 0) return VGPU_OK;   // no cell read an image block
		const int nv = cnt[1];
		if (nv > cap) {
			// more visited image blocks than the buffers hold: grow, list again
			if ((st = h->d_vkeys.ensure((size_t)nv * 8)) || (st = h->d_vblks.ensure((size_t)nv * 4))) return st;
			cap = nv;
			hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr + 10, 1, 0);
			hipLaunchKernelGGL(k_prd_visited, dim3(nblk(nb, 256)), dim3(256), 0, s, nb, fv, g.nx, g.oy, g.ey, g.wy,
					   g.ez, g.wz, h->d_vkeys.as<unsigned long long>(), h->d_vblks.as<int>(), cap, ctr + 10);
		}
		if ((st = h->h_vkeys.ensure((size_t)std::max(nv, 1) * 8)) || (st = h->h_vblks.ensure((size_t)std::max(nv, 1) * 4)))
			return st;
		if (nv) {
			HIPCHK(hipMemcpyAsync(h->h_vkeys.p, h->d_vkeys.p, (size_t)nv * 8, hipMemcpyDeviceToHost, s));
			HIPCHK(hipMemcpyAsync(h->h_vblks.p, h->d_vblks.p, (size_t)nv * 4, hipMemcpyDeviceToHost, s));
		}
		// this pass's keys, to compare the next pass's with
		HIPCHK(hipMemcpyAsync(h->d_prev.p, fv, (size_t)nb * 8, hipMemcpyDeviceToDevice, s));
		HIPCHK(hipStreamSynchronize(s));
		// the creation order: visited image blocks by first-visit key (keys
		// are distinct: each names one visit of one cell)
		{
			const unsigned long long *kk = h->h_vkeys.as<unsigned long long>();
			const int *bb = h->h_vblks.as<int>();
			std::vector<int> idx(nv);
			for (int q = 0; q < nv; q++) idx[q] = q;
			std::sort(idx.begin(), idx.end(), [&](int a, int b) { return kk[a] < kk[b]; });
			h->prd_order.resize(nv);
			for (int q = 0; q < nv; q++) h->prd_order[q] = bb[idx[q]];
		}
		ImageSet &im = h->image_set(hP, hpid, hcs);
		for (int b : h->prd_order) im.create_block(b);
		if (im.late) return VGPU_INTERNAL_ERROR;
		if ((st = upload_blocks(h, im, nc, v, s))) return st;
		hipLaunchKernelGGL(k_fill_u64, dim3(nblk(nb, 256)), dim3(256), 0, s, fv, nb, ~0ull);
		hipLaunchKernelGGL(k_fill_u8, dim3(nblk(nc, 256)), dim3(256), 0, s, h->touch.as<unsigned char>(), nc, 0);
		if ((st = run_tiers(h, v, nc, fw, h->retry0.as<int>(), ctr + 4, out, s, false))) return st;
		int differ = 0;
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr + 11, 1, 0);
		hipLaunchKernelGGL(k_differ_u64, dim3(nblk(nb, 256)), dim3(256), 0, s, nb, fv,
				   h->d_prev.as<unsigned long long>(), ctr + 11);
		HIPCHK(hipMemcpyAsync(&differ, ctr + 11, 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		h->image_passes = pass + 2;
		if (!differ) return VGPU_OK;
	}
	h->prd_order.clear();
	return VGPU_INTERNAL_ERROR;   // the first-visit order did not settle
}

// Internal status: a capacity grew, compute again.
const int VGPU_RETRY = -100;

// The statistics of every cell from the tables the tiers kept: output sizes
// in loop order, their scans, then one thread per cell.  The face offsets
// are the neighbour CSR's (h->tmp after the CSR stage).
int geometry_pass(vgpu_container *h, const GeoSink &g, int nc, hipStream_t s) {
	const size_t C1 = (size_t)nc + 1;
	int st;
	if ((st = h->go_nv.ensure(C1 * 4)) || (st = h->go_ne.ensure(C1 * 4)) || (st = h->go_nfv.ensure(C1 * 4)) ||
	    (st = h->go_voff.ensure(C1 * 4)) || (st = h->go_eoff.ensure(C1 * 4)) || (st = h->go_fvoff.ensure(C1 * 4)))
		return st;
	if (launch_geo_sizes(s, nc, h->status.as<int>(), h->nfaces.as<int>(), g, h->go_nv.as<int>(), h->go_ne.as<int>(),
			     h->go_nfv.as<int>()))
		return VGPU_HIP_ERROR;
	if ((st = scan_excl(h->part, h->go_nv.as<int>(), h->go_voff.as<int>(), nc + 1, s)) ||
	    (st = scan_excl(h->part, h->go_ne.as<int>(), h->go_eoff.as<int>(), nc + 1, s)) ||
	    (st = scan_excl(h->part, h->go_nfv.as<int>(), h->go_fvoff.as<int>(), nc + 1, s)))
		return st;
	int tot[4];
	HIPCHK(hipMemcpyAsync(&tot[0], h->go_voff.as<int>() + nc, 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(&tot[1], h->go_eoff.as<int>() + nc, 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(&tot[2], h->go_fvoff.as<int>() + nc, 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(&tot[3], h->tmp.as<int>() + nc, 4, hipMemcpyDeviceToHost, s));   // faces: the CSR total
	HIPCHK(hipStreamSynchronize(s));
	const size_t nv = (size_t)std::max(tot[0], 1), ne = (size_t)std::max(tot[1], 1), nfv = (size_t)std::max(tot[2], 1);
	const size_t nf = (size_t)std::max(tot[3], 1);
	if ((st = h->go_vert.ensure(nv * 24)) || (st = h->go_ord.ensure(nv * 4)) || (st = h->go_edges.ensure(ne * 4)) ||
	    (st = h->go_erel.ensure(ne * 4)) || (st = h->go_farea.ensure(nf * 8)) || (st = h->go_fperim.ensure(nf * 8)) ||
	    (st = h->go_normal.ensure(nf * 24)) || (st = h->go_forder.ensure(nf * 4)) ||
	    (st = h->go_fverts.ensure(nfv * 4)) || (st = h->go_surface.ensure(C1 * 8)) ||
	    (st = h->go_edist.ensure(C1 * 8)) || (st = h->go_nedges.ensure(C1 * 4)) || (st = h->go_maxr2.ensure(C1 * 8)))
		return st;
	GeoStats o;
	o.status = h->status.as<int>(); o.face_off = h->tmp.as<int>(); o.vert_off = h->go_voff.as<int>();
	o.edge_off = h->go_eoff.as<int>(); o.fv_off = h->go_fvoff.as<int>();
	o.tol = 10.0 * 2.220446049250313e-16 * h->max_len_sq;
	o.vertices = h->go_vert.as<double>(); o.orders = h->go_ord.as<int>(); o.edges = h->go_edges.as<int>();
	o.edge_rel = h->go_erel.as<int>(); o.face_area = h->go_farea.as<double>(); o.face_perim = h->go_fperim.as<double>();
	o.normal = h->go_normal.as<double>(); o.face_order = h->go_forder.as<int>(); o.face_verts = h->go_fverts.as<int>();
	o.surface = h->go_surface.as<double>(); o.edge_dist = h->go_edist.as<double>(); o.n_edges = h->go_nedges.as<int>();
	o.maxr2 = h->go_maxr2.as<double>();
	if (launch_geo_stats(s, nc, g, o)) return VGPU_HIP_ERROR;
	HIPCHK(hipStreamSynchronize(s));
	h->geo_nv = tot[0]; h->geo_ne = tot[1]; h->geo_nfv = tot[2]; h->geo_nf = tot[3];
	h->geo_valid = true;
	return VGPU_OK;
}

int run_pipeline(vgpu_container *h, unsigned flags, hipStream_t s, bool bin_only = false) {
	const int nxyz = h->nblocks;
	int n = h->n_in;
	int st;
	if (!h->events) {
		for (int t = 0; t < 5; t++) HIPCHK(hipEventCreate(&h->ev[t]));
		h->events = true;
	}
	HIPCHK(hipEventRecord(h->ev[0], s));
	// upload pending host puts
	if (!h->have_dev_input) {
		if ((st = h->in_ids.ensure((size_t)std::max(n, 1) * 4))) return st;
		if ((st = h->in_xyz.ensure((size_t)std::max(n, 1) * 24))) return st;
		if (n) {
			HIPCHK(hipMemcpyAsync(h->in_ids.p, h->h_ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
			HIPCHK(hipMemcpyAsync(h->in_xyz.p, h->h_xyz.data(), (size_t)n * 24, hipMemcpyHostToDevice, s));
		}
		if (h->poly) {
			if ((st = h->in_r.ensure((size_t)std::max(n, 1) * 8))) return st;
			if (n) HIPCHK(hipMemcpyAsync(h->in_r.p, h->h_r.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
		}
	}
	const size_t N1 = (size_t)std::max(n, 1);
	if ((st = h->blk.ensure(N1 * 4)) || (st = h->slot.ensure(N1 * 4)) || (st = h->rem.ensure(N1 * 32)) ||
	    (st = h->cnt.ensure((size_t)(nxyz + 1) * 4)) || (st = h->bstart.ensure((size_t)(nxyz + 1) * 4)) ||
	    (st = h->tmp.ensure(N1 * 4)) || (st = h->P.ensure(N1 * 32)) || (st = h->pid.ensure(N1 * 4)) ||
	    (st = h->src.ensure(N1 * 4)) || (st = h->cellblk.ensure(N1 * 4)) || (st = h->d_maxr.ensure(8)) ||
	    (st = h->d_counters.ensure(64 * 4)))
		return st;
	hipLaunchKernelGGL(k_fill_int, dim3(nblk(nxyz + 1, 256)), dim3(256), 0, s, h->cnt.as<int>(), nxyz + 1, 0);
	HIPCHK(hipMemsetAsync(h->d_maxr.p, 0, 8, s));
	const double *rptr = h->poly ? h->in_r.as<double>() : nullptr;
	if (n && h->prd) {
		const PrdGeom &pg = h->pg;
		PrdBin pb{pg.bx, pg.bxy, pg.by, pg.bxz, pg.byz, pg.bz, pg.xsp, pg.ysp, pg.zsp,
			  pg.nx, pg.ny, pg.nz, pg.ey, pg.ez, pg.oy};
		hipLaunchKernelGGL(k_locate_prd, dim3(nblk(n, 256)), dim3(256), 0, s, n, h->in_xyz.as<double>(), rptr, pb,
				   h->blk.as<int>(), h->slot.as<int>(), h->rem.as<dbl4>(), h->cnt.as<int>(),
				   h->d_maxr.as<unsigned long long>());
	} else if (n) {
		hipLaunchKernelGGL(k_locate, dim3(nblk(n, 256)), dim3(256), 0, s, n, h->in_xyz.as<double>(), rptr, h->g,
				   h->blk.as<int>(), h->slot.as<int>(), h->rem.as<dbl4>(), h->cnt.as<int>(),
				   h->d_maxr.as<unsigned long long>());
	}
	if ((st = scan_excl(h->part, h->cnt.as<int>(), h->bstart.as<int>(), nxyz + 1, s))) return st;
	if (n) {
		hipLaunchKernelGGL(k_place, dim3(nblk(n, 256)), dim3(256), 0, s, n, h->blk.as<int>(), h->slot.as<int>(),
				   h->bstart.as<int>(), h->tmp.as<int>());
		hipLaunchKernelGGL(k_rank, dim3(nblk(n, 256)), dim3(256), 0, s, n, h->blk.as<int>(), h->bstart.as<int>(),
				   h->tmp.as<int>(), h->rem.as<dbl4>(), h->in_ids.as<int>(), h->P.as<dbl4>(), h->pid.as<int>(),
				   h->src.as<int>(), h->cellblk.as<int>());
	}
	HIPCHK(hipGetLastError());
	int nc = 0, dup = 0;
	unsigned long long mrb = 0;
	if (h->prd && n) {
		// periodic put() rejects a particle within 1e-5 of one in its block
		if ((st = h->d_flag.ensure(4))) return st;
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, h->d_flag.as<int>(), 1, 0);
		hipLaunchKernelGGL(k_duplicates, dim3(nblk(n, 256)), dim3(256), 0, s, nxyz, h->P.as<dbl4>(), h->cellblk.as<int>(),
				   h->bstart.as<int>(), h->d_flag.as<int>());
	}
	HIPCHK(hipMemcpyAsync(&nc, h->bstart.as<int>() + nxyz, 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(&mrb, h->d_maxr.p, 8, hipMemcpyDeviceToHost, s));
	if (h->prd && n) HIPCHK(hipMemcpyAsync(&dup, h->d_flag.p, 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	h->ncells = nc;
	double mr;
	std::memcpy(&mr, &mrb, 8);
	h->max_radius = h->poly ? mr : 0;
	if (dup) { h->computed = false; return VGPU_DUPLICATE; }
	HIPCHK(hipEventRecord(h->ev[1], s));
	h->binned = true;
	h->prd_order.clear();
	if (bin_only) return VGPU_OK;

	// outputs
	const size_t C1 = (size_t)std::max(nc, 1);
	const bool nbrs = (flags & VGPU_F_NEIGHBORS) != 0;
	// periodic containers recompute the cells near image blocks: room for both
	long long acap = nbrs ? std::min<long long>(2147483000LL, std::max<long long>(h->acap_hint, (h->prd ? 48LL : 24LL) * (long long)C1 + 4096))
			      : 1;
	if ((st = h->vol.ensure(C1 * 8)) || (st = h->cent.ensure(C1 * 24)) || (st = h->nfaces.ensure(C1 * 4)) ||
	    (st = h->status.ensure(C1 * 4)) || (st = h->nbr_start.ensure(C1 * 4)) || (st = h->arena.ensure((size_t)acap * 4)) ||
	    (st = h->ctr.ensure(16 * 4)) || (st = h->retry0.ensure(C1 * 4)) || (st = h->retry1.ensure(C1 * 4)) ||
	    (st = h->retry2.ensure(C1 * 4)) || (st = h->retry3.ensure(C1 * 4)) ||
	    (st = h->retry4.ensure(C1 * 4)))
		return st;
	const bool stats = (flags & VGPU_F_STATS) != 0;
	if (stats) {
		if ((st = h->st_planes.ensure(C1 * 2)) || (st = h->st_exit.ensure(C1)) || (st = h->st_tier.ensure(C1))) return st;
	}
	// [0] arena top [1..2] retry counts of tiers 1, 2 [3] failures [4] learned-tier retries [5] fallbacks
	// [6..7] wavefront-cell retry counts [9] learned: cells given no list
	// [10..11] periodic replay: visited image blocks, keys differ
	int *ctr = h->ctr.as<int>();
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr, 16, 0);
	RectView v = make_view(h);
	CellOut out;
	out.vol = h->vol.as<double>(); out.cent = h->cent.as<double>(); out.nfaces = h->nfaces.as<int>();
	out.status = h->status.as<int>(); out.nbr_start = h->nbr_start.as<int>(); out.nbr = h->arena.as<int>();
	out.arena_top = ctr; out.arena_cap = (int)acap; out.fail_count = ctr + 3;
	out.st_planes = stats ? h->st_planes.as<short>() : nullptr;
	out.st_exit = stats ? h->st_exit.as<signed char>() : nullptr;
	out.st_tier = stats ? h->st_tier.as<signed char>() : nullptr;
	out.want_nbrs = nbrs ? 1 : 0;
	out.touch = nullptr;
	const bool geometry = (flags & VGPU_F_GEOMETRY) != 0 && nbrs && (flags & VGPU_F_CSR);
	h->geo_valid = false;
	if (geometry) {
		// about 27 vertices per cell on random points; grown and computed
		// again if a cell does not fit
		const long long vcap = std::min<long long>(2147483000LL,
			std::max(h->geo_vcap_hint, (h->prd ? 64LL : 40LL) * (long long)C1 + 4096));
		const long long ecap = std::min<long long>(2147483000LL, std::max(h->geo_ecap_hint, 3 * vcap + 4096));
		if ((st = h->geo_top.ensure(4 * 4)) || (st = h->geo_vstart.ensure(C1 * 4)) || (st = h->geo_np.ensure(C1 * 4)) ||
		    (st = h->geo_estart.ensure(C1 * 4)) || (st = h->geo_esz.ensure(C1 * 4)) ||
		    (st = h->geo_xyz.ensure((size_t)vcap * 24)) || (st = h->geo_nu.ensure((size_t)vcap)) ||
		    (st = h->geo_eoff.ensure((size_t)vcap * 4)) || (st = h->geo_ed.ensure((size_t)ecap * 2)) ||
		    (st = h->geo_rel.ensure((size_t)ecap)))
			return st;
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, h->geo_top.as<int>(), 4, 0);
		GeoSink &g = out.geo;
		g.top = h->geo_top.as<int>(); g.vcap = (int)vcap; g.ecap = (int)ecap;
		g.vstart = h->geo_vstart.as<int>(); g.np = h->geo_np.as<int>(); g.estart = h->geo_estart.as<int>();
		g.esz = h->geo_esz.as<int>(); g.xyz = h->geo_xyz.as<double>(); g.nu = h->geo_nu.as<unsigned char>();
		g.eoff = h->geo_eoff.as<int>(); g.ed = h->geo_ed.as<short>(); g.rel = h->geo_rel.as<unsigned char>();
	}

	int fw = h->fast_waves > 0 ? h->fast_waves : default_fast_waves(h->fast_occ);
	fw = std::max(1, std::min(fw, nblk(std::max(nc, 1), 64)));
	// the cell tables are kept by the plain tiers only
	const bool learned = (flags & VGPU_F_LEARNED) != 0 && !geometry;
	h->last_fallbacks = 0;
	h->last_entries = 0;
	h->last_routed = 0;
	h->image_passes = 0;
	if (nc > 0 && h->prd) {
		if ((st = run_prd_cells(h, v, out, nc, fw, s, learned))) return st;
	} else if (nc > 0 && learned) {
		// tier 0: candidate lists cut at the predicted worklist depth; the
		// cells it does not finish (retry0, count ctr[4]) take the plain tiers
		if ((st = run_learned(h, v, nc, fw, out, s))) return st;
		HIPCHK(hipEventRecord(h->ev[2], s));
		if ((st = run_tiers(h, v, nc, fw, h->retry0.as<int>(), ctr + 4, out, s, false, 1, h->first_tier == 0)))
			return st;
	} else if (nc > 0) {
		// the plain tiers: wavefront cells (or order-3 per lane), wide, huge
		if ((st = run_tiers(h, v, nc, fw, nullptr, nullptr, out, s, true))) return st;
	} else {
		HIPCHK(hipEventRecord(h->ev[2], s));
	}
	HIPCHK(hipEventRecord(h->ev[3], s));
	h->csr_valid = false;
	if (nbrs && (flags & VGPU_F_CSR) && nc > 0) {
		if ((st = h->csr_off.ensure((C1 + 1) * 4))) return st;
		hipLaunchKernelGGL(k_face_counts, dim3(nblk(nc + 1, 256)), dim3(256), 0, s, nc, h->status.as<int>(),
				   h->nfaces.as<int>(), h->nbr_start.as<int>(), h->csr_off.as<int>());
		// scan in place into a second buffer region
		if ((st = h->tmp.ensure((C1 + 1) * 4))) return st;
		if ((st = scan_excl(h->part, h->csr_off.as<int>(), h->tmp.as<int>(), nc + 1, s))) return st;
		if ((st = h->csr.ensure((size_t)acap * 4))) return st;
		hipLaunchKernelGGL(k_gather_nbrs, dim3(nblk(nc, 256)), dim3(256), 0, s, nc, h->status.as<int>(),
				   h->nfaces.as<int>(), h->nbr_start.as<int>(), h->arena.as<int>(), h->tmp.as<int>(),
				   h->csr.as<int>());
		h->csr_valid = true;
	}
	HIPCHK(hipEventRecord(h->ev[4], s));
	int hc[10];
	HIPCHK(hipMemcpyAsync(hc, ctr, 40, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	// a full neighbour arena: grow it and compute again
	if (nbrs && (long long)(unsigned)hc[0] > acap) {
		h->acap_hint = (long long)(unsigned)hc[0] + (long long)(unsigned)hc[0] / 8 + 4096;
		h->computed = false;
		return VGPU_RETRY;
	}
	h->nbr_total = hc[0];
	if (geometry && nc > 0) {
		int gt[4];
		HIPCHK(hipMemcpyAsync(gt, h->geo_top.p, 16, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (gt[2] > 0) {
			// a cell's tables did not fit: room for everything asked for, and more
			h->geo_vcap_hint = (long long)(unsigned)gt[0] + (long long)(unsigned)gt[0] / 4 + 4096;
			h->geo_ecap_hint = (long long)(unsigned)gt[1] + (long long)(unsigned)gt[1] / 4 + 4096;
			h->computed = false;
			return VGPU_RETRY;
		}
		if (hc[3] == 0 && (st = geometry_pass(h, out.geo, nc, s))) return st;
	}
	float ms;
	for (int t = 0; t < 4; t++) {
		HIPCHK(hipEventElapsedTime(&ms, h->ev[t], h->ev[t + 1]));
		h->timings[t] = ms;
	}
	h->nbr_total = hc[0];
	h->gather_ms = 0;
	if (learned && nc > 0) {
		h->last_fallbacks = hc[5];
		h->last_routed = hc[9];
		HIPCHK(hipEventElapsedTime(&ms, h->ev[1], h->ev_list));
		h->gather_ms = ms;
	}
	// a cell no tier finished, or ids beyond the arena: no results to read
	h->computed = hc[3] == 0;
	h->last_flags = flags;
	if (hc[3] > 0) return VGPU_MEMORY_ERROR;
	return VGPU_OK;
}

}  // namespace


// ---------------------------------------------------------------------------
// Batches of containers (vgpu_batch_*): many independent rectangular
// containers computed in one pass over one set of device arrays.

struct vgpu_batch {
	int ncont = 0;
	bool poly = false;
	int n = 0;                       // particles put, over all containers
	int nblocks = 0;
	std::vector<int> h_boff, h_pfirst, h_nwalls;
	std::vector<int> h_cfirst;       // first cell of each container (after a compute)
	DBuf views, geoms, boff, pcont, pfirst, mrad, walls, d_wl, maxr, cfirst, nchunk, coff, chunks, nchunks, stamp;
	DBuf in_ids, in_xyz, in_r;
	DBuf blk, slot, rem, cnt, bstart, tmp, part, P, pid, src, cellblk;
	DBuf vol, cent, nfaces, status, nbr_start, arena, csr_off, csr, ctr, counters, retry1, retry2, retry3;
	DBuf st_planes, st_exit, st_tier;
	Workspace ws_fast, ws_wide, ws_huge, ws_wave;
	long long acap = 0;
	// learned tier: predictor (per block), its outputs, the schedule, the
	// candidate lists (capacity lcap entries, grown after a compute that
	// needed more), the container of every block
	DBuf d_pred, d_feat, d_depth, d_cost, d_hist, d_order, lcnt, loff, list_a, bcont, ltop;
	bool have_pred = false, bcont_ok = false;
	float pred_margin = 0;
	int pred_route = 64;
	long long lcap = 0;
	long long last_fallbacks = 0, last_routed = 0, last_entries = 0, last_need = 0;
	// launch configuration (autotuned): persistent wavefronts of the first
	// tier, its kind (0 wavefront cells, 2..4 one cell per lane at that many
	// waves per SIMD) and the cells per queue entry
	int waves = 0, occ = 2, chunk = 64;
	int waves_req = 0;               // persistent wavefronts asked for (0: as many as fit); waves: this compute's
	unsigned long long stamp_used = 0;   // stamps handed out since the workspaces were cleared
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	unsigned graph_flags = 0;
	hipStream_t gstream = nullptr;   // capture stream when the caller passes the null stream
	hipEvent_t ev[4];
	bool events = false;
	double timings[3] = {0, 0, 0};
	bool computed = false;
	unsigned last_flags = 0;
	long long nbr_total = 0;
	int fails = 0;
	void drop_graph() {
		if (exec) (void)hipGraphExecDestroy(exec);
		if (graph) (void)hipGraphDestroy(graph);
		exec = nullptr; graph = nullptr;
	}
};

namespace {

// Waves per SIMD of a batch's wavefront-cell tiers: first tier 5 or 6 asks
// for that many, the other settings take 4.
int batch_wave_occ(const vgpu_batch *b) { return b->occ >= 5 ? b->occ : 4; }

__global__ void k_batch_pcont(int ncont, const int *__restrict__ pfirst, int *__restrict__ pcont) {
	const int c = blockIdx.x;
	if (c >= ncont) return;
	for (int t = pfirst[c] + threadIdx.x; t < pfirst[c + 1]; t += blockDim.x) pcont[t] = c;
}

// Adds `inc` to the device stamp base (after each pass, so that the next
// pass -- a graph replay included -- uses fresh block-mask stamps).
__global__ void k_stamp_advance(unsigned *stamp, unsigned inc) {
	if (threadIdx.x == 0 && blockIdx.x == 0) *stamp += inc;
}

// Stamp bases per tier are read on the device, so a captured graph keeps
// working after replays.
__global__ void k_batch_stamp_args(const unsigned *stamp, unsigned *args, int n) {
	if (threadIdx.x == 0 && blockIdx.x == 0) for (int t = 0; t < 3; t++) args[t] = *stamp + (unsigned)t * (unsigned)(n + 1);
}

// Sizes every buffer a compute of this batch uses, so that the kernel
// sequence itself allocates nothing (and can be captured).
int batch_prepare(vgpu_batch *b, unsigned flags) {
	const size_t N1 = (size_t)std::max(b->n, 1), NB1 = (size_t)b->nblocks + 1, C1 = (size_t)b->ncont + 1;
	int st;
	const bool stats = (flags & VGPU_F_STATS) != 0;
	b->acap = std::min<long long>(2147483000LL, 24LL * (long long)N1 + 4096);
	if ((st = b->blk.ensure(N1 * 4)) || (st = b->slot.ensure(N1 * 4)) || (st = b->rem.ensure(N1 * 32)) ||
	    (st = b->cnt.ensure(NB1 * 4)) || (st = b->bstart.ensure(NB1 * 4)) || (st = b->tmp.ensure((N1 + 1) * 4)) ||
	    (st = b->P.ensure(N1 * 32)) || (st = b->pid.ensure(N1 * 4)) || (st = b->src.ensure(N1 * 4)) ||
	    (st = b->cellblk.ensure(N1 * 4)) || (st = b->maxr.ensure(C1 * 8)) || (st = b->cfirst.ensure(C1 * 4)) ||
	    (st = b->nchunk.ensure(C1 * 4)) || (st = b->coff.ensure(C1 * 4)) ||
	    (st = b->chunks.ensure((N1 + C1) * sizeof(BatchChunk))) || (st = b->nchunks.ensure(4)) ||
	    (st = b->stamp.ensure(16 * 4)) || (st = b->vol.ensure(N1 * 8)) || (st = b->cent.ensure(N1 * 24)) ||
	    (st = b->nfaces.ensure(N1 * 4)) || (st = b->status.ensure(N1 * 4)) || (st = b->nbr_start.ensure(N1 * 4)) ||
	    (st = b->arena.ensure((size_t)b->acap * 4)) || (st = b->csr_off.ensure((N1 + 1) * 4)) ||
	    (st = b->csr.ensure((size_t)b->acap * 4)) || (st = b->ctr.ensure(16 * 4)) || (st = b->counters.ensure(64 * 4)) ||
	    (st = b->retry1.ensure(N1 * 4)) || (st = b->retry2.ensure(N1 * 4)) || (st = b->retry3.ensure(N1 * 4)))
		return st;
	const size_t pb = std::max(std::max(scan_part_bytes((int)NB1), scan_part_bytes((int)C1)), scan_part_bytes((int)N1 + 1));
	if ((st = b->part.ensure(pb))) return st;
	if (stats && ((st = b->st_planes.ensure(N1 * 2)) || (st = b->st_exit.ensure(N1)) || (st = b->st_tier.ensure(N1))))
		return st;
	const bool learned = (flags & VGPU_F_LEARNED) != 0;
	if (learned || !b->bcont_ok) {
		if ((st = b->bcont.ensure(NB1 * 4))) return st;
		if (!b->bcont_ok) {
			if (launch_batch_bcont(nullptr, b->ncont, b->boff.as<int>(), b->bcont.as<int>())) return VGPU_HIP_ERROR;
			HIPCHK(hipStreamSynchronize(nullptr));
			b->bcont_ok = true;
		}
	}
	if (learned) {
		using namespace vgpu::learn;
		// a list holds depth + 1 <= 64 records; trained depths are about 40
		// worklist steps on random and lattice points
		if (b->lcap <= 0) b->lcap = std::min<long long>(2147483000LL, 48LL * (long long)N1 + 4096);
		if (b->list_a.n < (size_t)b->lcap * sizeof(ListRec))
			b->drop_graph();   // the list buffers move
		if ((st = b->d_feat.ensure(NB1 * IN * 4)) || (st = b->d_depth.ensure(NB1)) || (st = b->d_cost.ensure(NB1)) ||
		    (st = b->d_hist.ensure(64 * 4)) || (st = b->d_order.ensure(N1 * 4)) || (st = b->lcnt.ensure(N1 * 4)) ||
		    (st = b->loff.ensure(N1 * 4)) || (st = b->ltop.ensure(8)) ||
		    (st = b->list_a.ensure((size_t)b->lcap * sizeof(ListRec))))
			return st;
	}
	const bool wave = b->occ == 0 || b->occ >= 5 || learned;
	int w = b->waves_req > 0 ? b->waves_req : default_fast_waves(wave ? batch_wave_occ(b) : b->occ);
	w = std::max(1, std::min(w, (wave ? nblk((long long)N1, b->chunk) : nblk((long long)N1, 64)) + b->ncont));
	b->waves = w;
	typedef Layout<TierWide, 64> LW;
	typedef Layout<TierHuge, 64> LH;
	if ((wave && (st = b->ws_wave.ensure(WLayout::bytes, w))) ||
	    (st = b->ws_fast.ensure(VCell3::L::bytes, wave ? BATCH_LEFTOVER_WAVES : w)) ||
	    (st = b->ws_wide.ensure(LW::bytes, 64)) || (st = b->ws_huge.ensure(LH::bytes, 4)))
		return st;
	// stamps: three passes of n+1 per compute; clear the workspaces before
	// the 32-bit stamp space runs out
	const unsigned long long per = 3ull * ((unsigned long long)N1 + 1) + 3;
	if (b->stamp_used + per >= 0xffffffffull || b->stamp_used == 0) {
		Workspace *wss[4] = {&b->ws_fast, &b->ws_wide, &b->ws_huge, &b->ws_wave};
		for (Workspace *x : wss)
			if (x->buf.p) HIPCHK(hipMemset(x->buf.p, 0, x->buf.n));
		HIPCHK(hipMemset(b->stamp.p, 0, 16 * 4));
		b->stamp_used = 1;
		b->drop_graph();   // the graph's stamps restart from zero with the workspace
	}
	if (!b->events) {
		for (int t = 0; t < 4; t++) HIPCHK(hipEventCreate(&b->ev[t]));
		b->events = true;
	}
	return VGPU_OK;
}

// The kernel sequence of one batch compute (no allocation, no host sync).
int batch_enqueue(vgpu_batch *b, unsigned flags, hipStream_t s, bool timing) {
	int st;
	const int n = b->n, nb = b->nblocks, nc = b->ncont;
	const bool stats = (flags & VGPU_F_STATS) != 0, nbrs = (flags & VGPU_F_NEIGHBORS) != 0;
	if (timing) HIPCHK(hipEventRecord(b->ev[0], s));
	hipLaunchKernelGGL(k_fill_int, dim3(nblk(nb + 1, 256)), dim3(256), 0, s, b->cnt.as<int>(), nb + 1, 0);
	hipLaunchKernelGGL(k_fill_u64, dim3(nblk(nc, 256)), dim3(256), 0, s, b->maxr.as<unsigned long long>(), nc, 0ull);
	if (launch_batch_locate(s, n, b->in_xyz.as<double>(), b->poly ? b->in_r.as<double>() : nullptr,
				b->geoms.as<RectGeom>(), b->pcont.as<int>(), b->boff.as<int>(), b->blk.as<int>(),
				b->slot.as<int>(), b->rem.as<dbl4>(), b->cnt.as<int>(), b->maxr.as<unsigned long long>()))
		return VGPU_HIP_ERROR;
	if ((st = scan_excl(b->part, b->cnt.as<int>(), b->bstart.as<int>(), nb + 1, s, true))) return st;
	if (n) {
		hipLaunchKernelGGL(k_place, dim3(nblk(n, 256)), dim3(256), 0, s, n, b->blk.as<int>(), b->slot.as<int>(),
				   b->bstart.as<int>(), b->tmp.as<int>());
		hipLaunchKernelGGL(k_rank, dim3(nblk(n, 256)), dim3(256), 0, s, n, b->blk.as<int>(), b->bstart.as<int>(),
				   b->tmp.as<int>(), b->rem.as<dbl4>(), b->in_ids.as<int>(), b->P.as<dbl4>(), b->pid.as<int>(),
				   b->src.as<int>(), b->cellblk.as<int>());
	}
	if (launch_batch_counts(s, nc, b->chunk, b->boff.as<int>(), b->bstart.as<int>(), b->maxr.as<unsigned long long>(),
				b->poly ? 1 : 0, b->views.as<RectView>(), b->cfirst.as<int>(), b->nchunk.as<int>()))
		return VGPU_HIP_ERROR;
	if ((st = scan_excl(b->part, b->nchunk.as<int>(), b->coff.as<int>(), nc + 1, s, true))) return st;
	if (launch_batch_chunks(s, nc, b->chunk, b->cfirst.as<int>(), b->coff.as<int>(), b->chunks.as<BatchChunk>(),
				b->nchunks.as<int>()))
		return VGPU_HIP_ERROR;
	int *ctr = b->ctr.as<int>();
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr, 16, 0);
	if (timing) HIPCHK(hipEventRecord(b->ev[1], s));
	CellOut out;
	out.vol = b->vol.as<double>(); out.cent = b->cent.as<double>(); out.nfaces = b->nfaces.as<int>();
	out.status = b->status.as<int>(); out.nbr_start = b->nbr_start.as<int>(); out.nbr = b->arena.as<int>();
	out.arena_top = ctr; out.arena_cap = (int)b->acap; out.fail_count = ctr + 3;
	out.st_planes = stats ? b->st_planes.as<short>() : nullptr;
	out.st_exit = stats ? b->st_exit.as<signed char>() : nullptr;
	out.st_tier = stats ? b->st_tier.as<signed char>() : nullptr;
	out.touch = nullptr;
	out.want_nbrs = nbrs ? 1 : 0;
	// per-tier stamp bases from the device counter (ctr[8..10])
	hipLaunchKernelGGL(k_batch_stamp_args, dim3(1), dim3(64), 0, s, b->stamp.as<unsigned>(), (unsigned *)(ctr + 8), n);
	int *cnt8 = b->counters.as<int>();
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, cnt8, 8, 0);
	out.tier = 0; out.retry = b->retry1.as<int>(); out.retry_count = ctr + 1;
	// the wide tier's list: the first tier's leftovers, or (wavefront cells
	// first) those the order-3 per-lane tier leaves (ctr[5])
	const int *wide_list = b->retry1.as<int>(), *wide_count = ctr + 1;
	if (flags & VGPU_F_LEARNED) {
		// the learned tier: block features, predicted depths (and costs, the
		// schedule's order), the candidate lists, the list-driven wavefront
		// cells (cells whose list ends first, ctr[12], or that have none,
		// ctr[13], continue there as plain wavefront cells); overflows
		// (retry1) take the per-lane order-3 tier, and so on
		using namespace vgpu::learn;
		const int *ncp = b->bstart.as<int>() + nb;
		if (launch_batch_features(s, nb, b->bcont.as<int>(), b->views.as<RectView>(), b->d_feat.as<float>()))
			return VGPU_HIP_ERROR;
		if ((st = mlp_depth_dev(b->d_pred.as<float>(), b->d_feat.as<float>(), nb, b->pred_margin, b->pred_route,
					b->d_depth.as<unsigned char>(), b->d_cost.as<unsigned char>(), s)))
			return st;
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, b->d_hist.as<int>(), 64, 0);
		if (n > 0 && launch_list_order(s, n, b->cellblk.as<int>(), b->d_cost.as<unsigned char>(), 0,
					       b->d_hist.as<int>(), b->d_order.as<int>(), ncp))
			return VGPU_HIP_ERROR;
		hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, b->ltop.as<unsigned long long>(), 1, 0ull);
		if (launch_batch_list_build(s, b->views.as<RectView>(), b->bcont.as<int>(), n, ncp, b->cellblk.as<int>(),
					    b->d_order.as<int>(), b->d_depth.as<unsigned char>(), b->lcap,
					    b->ltop.as<unsigned long long>(), b->lcnt.as<int>(), b->loff.as<int>(),
					    b->list_a.as<ListRec>()))
			return VGPU_HIP_ERROR;
		// overflows (more than 64 vertices) to the per-lane order-3 tier
		out.tier = 0; out.retry = b->retry1.as<int>(); out.retry_count = ctr + 1;
		const int chunk = n >= 4 * b->waves ? 4 : 1;
		if (launch_batch_wcells_list(b->poly, batch_wave_occ(b), b->waves, chunk, s, b->views.as<RectView>(), b->bcont.as<int>(), ncp,
					     b->cellblk.as<int>(), b->d_order.as<int>(), b->lcnt.as<int>(), b->loff.as<int>(),
					     b->list_a.as<ListRec>(), cnt8, b->ws_wave.base(), b->ws_wave.slab,
					     (const unsigned *)(ctr + 8), ctr + 12, ctr + 13, out))
			return VGPU_HIP_ERROR;
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, cnt8, 8, 0);
		out.tier = 1; out.retry = b->retry3.as<int>(); out.retry_count = ctr + 5;
		if (launch_batch_retry(0, b->poly, BATCH_LEFTOVER_WAVES, s, b->views.as<RectView>(), b->boff.as<int>(), nc,
				       b->retry1.as<int>(), ctr + 1, b->cellblk.as<int>(), cnt8, b->ws_fast.base(),
				       b->ws_fast.slab, (const unsigned *)(ctr + 8), out))
			return VGPU_HIP_ERROR;
		wide_list = b->retry3.as<int>(); wide_count = ctr + 5;
	} else if (b->occ == 0 || b->occ >= 5) {
		if (launch_batch_wfirst(b->poly, batch_wave_occ(b), b->waves, s, b->views.as<RectView>(), b->chunks.as<BatchChunk>(),
					b->nchunks.as<int>(), b->cellblk.as<int>(), cnt8, b->ws_wave.base(), b->ws_wave.slab,
					(const unsigned *)(ctr + 8), out))
			return VGPU_HIP_ERROR;
		hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, cnt8, 8, 0);
		out.tier = 1; out.retry = b->retry3.as<int>(); out.retry_count = ctr + 5;
		if (launch_batch_retry(0, b->poly, BATCH_LEFTOVER_WAVES, s, b->views.as<RectView>(), b->boff.as<int>(), nc,
				       b->retry1.as<int>(), ctr + 1, b->cellblk.as<int>(), cnt8, b->ws_fast.base(),
				       b->ws_fast.slab, (const unsigned *)(ctr + 8), out))
			return VGPU_HIP_ERROR;
		wide_list = b->retry3.as<int>(); wide_count = ctr + 5;
	} else if (launch_batch_first(b->poly, b->occ, b->waves, s, b->views.as<RectView>(), b->chunks.as<BatchChunk>(),
				      b->nchunks.as<int>(), b->cellblk.as<int>(), cnt8, b->ws_fast.base(), b->ws_fast.slab,
				      (const unsigned *)(ctr + 8), out))
		return VGPU_HIP_ERROR;
	if (timing) HIPCHK(hipEventRecord(b->ev[2], s));
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, cnt8, 8, 0);
	out.tier = 2; out.retry = b->retry2.as<int>(); out.retry_count = ctr + 2;
	if (launch_batch_retry(1, b->poly, 64, s, b->views.as<RectView>(), b->boff.as<int>(), nc, wide_list, wide_count,
			       b->cellblk.as<int>(), cnt8, b->ws_wide.base(), b->ws_wide.slab, (const unsigned *)(ctr + 9),
			       out))
		return VGPU_HIP_ERROR;
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, cnt8, 8, 0);
	out.tier = 3; out.retry = nullptr; out.retry_count = nullptr;
	if (launch_batch_retry(2, b->poly, 4, s, b->views.as<RectView>(), b->boff.as<int>(), nc, b->retry2.as<int>(),
			       ctr + 2, b->cellblk.as<int>(), cnt8, b->ws_huge.base(), b->ws_huge.slab,
			       (const unsigned *)(ctr + 10), out))
		return VGPU_HIP_ERROR;
	hipLaunchKernelGGL(k_stamp_advance, dim3(1), dim3(64), 0, s, b->stamp.as<unsigned>(), 3u * (unsigned)(n + 1) + 3u);
	if (nbrs && n > 0) {
		// cells: bstart[nblocks] (particles outside their box are not cells)
		const int *ncells = b->bstart.as<int>() + nb;
		hipLaunchKernelGGL(k_face_counts_dev, dim3(nblk(n + 1, 256)), dim3(256), 0, s, n, ncells, b->status.as<int>(),
				   b->nfaces.as<int>(), b->nbr_start.as<int>(), b->csr_off.as<int>());
		if ((st = scan_excl(b->part, b->csr_off.as<int>(), b->tmp.as<int>(), n + 1, s, true))) return st;
		hipLaunchKernelGGL(k_gather_nbrs_dev, dim3(nblk(n, 256)), dim3(256), 0, s, n, ncells, b->status.as<int>(),
				   b->nfaces.as<int>(), b->nbr_start.as<int>(), b->arena.as<int>(), b->tmp.as<int>(),
				   b->csr.as<int>());
	}
	if (timing) HIPCHK(hipEventRecord(b->ev[3], s));
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

int batch_run(vgpu_batch *b, unsigned flags, hipStream_t s) {
	int st;
	if ((st = batch_prepare(b, flags))) return st;
	const bool graph = (flags & VGPU_F_GRAPH) != 0;
	if (graph && s == nullptr) {
		// the null stream cannot be captured: use the batch's own stream
		if (!b->gstream) HIPCHK(hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking));
		HIPCHK(hipDeviceSynchronize());
		s = b->gstream;
	}
	if (graph) {
		const unsigned key = flags & ~VGPU_F_GRAPH;
		if (b->exec && b->graph_flags != key) b->drop_graph();
		if (!b->exec) {
			HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
			int est = batch_enqueue(b, key, s, false);
			hipGraph_t g = nullptr;
			hipError_t ce = hipStreamEndCapture(s, &g);
			if (est) { if (g) (void)hipGraphDestroy(g); return est; }
			if (ce != hipSuccess) return VGPU_HIP_ERROR;
			b->graph = g;
			HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
			b->graph_flags = key;
		}
		HIPCHK(hipEventRecord(b->ev[0], s));
		HIPCHK(hipGraphLaunch(b->exec, s));
		HIPCHK(hipEventRecord(b->ev[3], s));
	} else {
		if ((st = batch_enqueue(b, flags, s, true))) return st;
	}
	b->stamp_used += 3ull * ((unsigned long long)std::max(b->n, 1) + 1) + 3;
	int hc[16];
	unsigned long long need = 0;
	HIPCHK(hipMemcpyAsync(hc, b->ctr.p, 64, hipMemcpyDeviceToHost, s));
	if (flags & VGPU_F_LEARNED) HIPCHK(hipMemcpyAsync(&need, b->ltop.p, 8, hipMemcpyDeviceToHost, s));
	b->h_cfirst.resize(b->ncont + 1);
	HIPCHK(hipMemcpyAsync(b->h_cfirst.data(), b->cfirst.p, (size_t)(b->ncont + 1) * 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	float ms = 0;
	if (graph) {
		HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[3]));
		b->timings[0] = 0; b->timings[1] = ms; b->timings[2] = 0;
	} else {
		for (int t = 0; t < 3; t++) {
			HIPCHK(hipEventElapsedTime(&ms, b->ev[t], b->ev[t + 1]));
			b->timings[t] = ms;
		}
	}
	b->nbr_total = hc[0];
	b->fails = hc[3];
	b->last_fallbacks = b->last_routed = b->last_entries = b->last_need = 0;
	if (flags & VGPU_F_LEARNED) {
		b->last_fallbacks = hc[12];
		b->last_routed = hc[13];
		b->last_need = (long long)need;
		b->last_entries = std::min((long long)need, b->lcap);
		// lists that did not fit took the plain path (results are the
		// same); give the next compute room for all of them
		if ((long long)need > b->lcap)
			b->lcap = std::min<long long>(2147483000LL, (long long)need + (long long)need / 8 + 4096);
	}
	b->computed = hc[3] == 0;   // no results to read after a failure
	b->last_flags = flags;
	return hc[3] > 0 ? VGPU_MEMORY_ERROR : VGPU_OK;
}

}  // namespace

extern "C" {

int vgpu_create(int kind, const double *geom, const int *grid, const int *periodic, int init_mem,
		vgpu_container **out) {
	(void)init_mem;
	if (!out || !geom || !grid) return VGPU_INVALID_ARGUMENT;
	*out = nullptr;
	if (kind < VGPU_CONTAINER || kind > VGPU_CONTAINER_PERIODIC_POLY) return VGPU_INVALID_ARGUMENT;
	if (grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0) return VGPU_INVALID_ARGUMENT;
	if ((long long)grid[0] * grid[1] * grid[2] > 1000000000LL) return VGPU_INVALID_ARGUMENT;
	const bool prd = kind >= VGPU_CONTAINER_PERIODIC;
	for (int t = 0; t < 6; t++) if (!std::isfinite(geom[t])) return VGPU_INVALID_ARGUMENT;
	if (prd) {
		if (!(geom[0] > 0) || !(geom[2] > 0) || !(geom[5] > 0)) return VGPU_INVALID_ARGUMENT;
	} else if (!(geom[1] > geom[0]) || !(geom[3] > geom[2]) || !(geom[5] > geom[4])) return VGPU_INVALID_ARGUMENT;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return VGPU_NO_DEVICE;
	vgpu_container *h = new vgpu_container();
	h->kind = kind;
	h->prd = prd;
	h->poly = kind == VGPU_CONTAINER_POLY || kind == VGPU_CONTAINER_PERIODIC_POLY;
	for (int t = 0; t < 6; t++) h->geom[t] = geom[t];
	for (int t = 0; t < 3; t++) { h->grid[t] = grid[t]; h->per[t] = (!prd && periodic) ? (periodic[t] != 0) : 0; }
	RectGeom &g = h->g;
	if (prd) {
		// geom = {bx,bxy,by,bxz,byz,bz}: unit cell, then the block grid
		// (src/container_prd.cc:30-37, src/unitcell.cc:26-85)
		if (!h->pg.build(geom[0], geom[1], geom[2], geom[3], geom[4], geom[5], grid[0], grid[1], grid[2])) {
			delete h;
			return VGPU_MEMORY_ERROR;
		}
		if ((long long)h->pg.nblocks() > 1000000000LL) { delete h; return VGPU_INVALID_ARGUMENT; }
		g.ax = 0; g.bx = geom[0]; g.ay = 0; g.by = geom[2]; g.az = 0; g.bz = geom[5];
		g.nx = grid[0]; g.ny = grid[1]; g.nz = grid[2];
		g.boxx = h->pg.boxx; g.boxy = h->pg.boxy; g.boxz = h->pg.boxz;
		g.xsp = h->pg.xsp; g.ysp = h->pg.ysp; g.zsp = h->pg.zsp;
		g.xper = g.yper = g.zper = 0;
		h->max_len_sq = h->pg.max_len_sq;
		h->nblocks = (int)h->pg.nblocks();
		const size_t ub = h->pg.unit.size() * sizeof(UnitVert);
		if (h->d_unit.ensure(ub) ||
		    hipMemcpy(h->d_unit.p, h->pg.unit.data(), ub, hipMemcpyHostToDevice) != hipSuccess) {
			delete h;
			return VGPU_HIP_ERROR;
		}
	} else {
		g.ax = geom[0]; g.bx = geom[1]; g.ay = geom[2]; g.by = geom[3]; g.az = geom[4]; g.bz = geom[5];
		g.nx = grid[0]; g.ny = grid[1]; g.nz = grid[2];
		// src/container.cc:30-36, src/v_base.cc:23-25
		g.boxx = (g.bx - g.ax) / g.nx; g.boxy = (g.by - g.ay) / g.ny; g.boxz = (g.bz - g.az) / g.nz;
		g.xsp = 1 / g.boxx; g.ysp = 1 / g.boxy; g.zsp = 1 / g.boxz;
		g.xper = h->per[0]; g.yper = h->per[1]; g.zper = h->per[2];
		h->max_len_sq = (g.bx - g.ax) * (g.bx - g.ax) * (g.xper ? 0.25 : 1) + (g.by - g.ay) * (g.by - g.ay) * (g.yper ? 0.25 : 1)
				+ (g.bz - g.az) * (g.bz - g.az) * (g.zper ? 0.25 : 1);
		h->nblocks = g.nx * g.ny * g.nz;
	}
	std::vector<double> mrad(4096);
	worklist_radii(g.boxx, g.boxy, g.boxz, mrad.data());
	if (h->d_mrad.ensure(4096 * 8) || h->d_wl.ensure(4096 * 4) || h->d_walls.ensure(8 * VGPU_WALL_STRIDE * 64)) {
		delete h;
		return VGPU_MEMORY_ERROR;
	}
	if (hipMemcpy(h->d_mrad.p, mrad.data(), 4096 * 8, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(h->d_wl.p, vgpu_worklist_table, 4096 * 4, hipMemcpyHostToDevice) != hipSuccess) {
		delete h;
		return VGPU_HIP_ERROR;
	}
	*out = h;
	return VGPU_OK;
}

int vgpu_destroy(vgpu_container *h) {
	if (!h) return VGPU_INVALID_ARGUMENT;
	DBuf *bufs[] = {&h->in_ids, &h->in_xyz, &h->in_r, &h->blk, &h->slot, &h->rem, &h->cnt, &h->bstart, &h->tmp,
			&h->part, &h->P, &h->pid, &h->src, &h->cellblk, &h->d_mrad, &h->d_wl, &h->d_walls, &h->d_maxr,
			&h->d_counters, &h->vol, &h->cent, &h->nfaces, &h->status, &h->nbr_start, &h->arena, &h->csr_off,
			&h->csr, &h->ctr, &h->retry0, &h->retry1, &h->retry2, &h->st_planes, &h->st_exit, &h->st_tier,
			&h->d_pred, &h->d_feat, &h->d_cfeat, &h->d_depth, &h->d_cost, &h->d_order, &h->d_hist, &h->lcnt,
			&h->bmax, &h->boff, &h->list_a, &h->d_unit, &h->Pall, &h->pidall, &h->ball, &h->fv,
			&h->touch, &h->d_flag, &h->retry3, &h->retry4, &h->img_p, &h->img_id,
			&h->img_reg, &h->img_rank, &h->img_bc, &h->img_cnt, &h->fq_xyz, &h->fq_found, &h->fq_pid,
			&h->fq_pos, &h->fq_retry, &h->fq_retry2, &h->fq_r, &h->fq_vol, &h->fq_nbr, &h->fq_touch,
			&h->fq_list, &h->img_okb, &h->d_prev, &h->d_vkeys, &h->d_vblks, &h->d_tacc};
	for (DBuf *b : bufs) b->release();
	h->ws_fast.buf.release(); h->ws_wide.buf.release(); h->ws_huge.buf.release(); h->ws_wave.buf.release();
	h->ws_wave4.buf.release();
	if (h->events) for (int t = 0; t < 5; t++) (void)hipEventDestroy(h->ev[t]);
	if (h->ev_list) (void)hipEventDestroy(h->ev_list);
	delete h;
	return VGPU_OK;
}

int vgpu_add_wall(vgpu_container *h, int type, const double *prm, int w_id) {
	if (!h || h->prd || !prm || type < 0 || type > 3) return VGPU_INVALID_ARGUMENT;
	if ((int)(h->walls.size() / VGPU_WALL_STRIDE) >= 64) return VGPU_MEMORY_ERROR;
	double w[VGPU_WALL_STRIDE] = {0};
	w[0] = type; w[1] = w_id;
	w[2] = prm[0]; w[3] = prm[1]; w[4] = prm[2];
	if (type <= 1) w[5] = prm[3];
	else {
		// the constants the reference's constructors derive (src/wall.hh:83-85, 104-107)
		const double xa = prm[3], ya = prm[4], za = prm[5];
		w[5] = xa; w[6] = ya; w[7] = za;
		w[8] = 1 / (xa * xa + ya * ya + za * za);
		if (type == 2) w[9] = prm[6];
		else { const double ang = prm[6]; w[9] = sin(ang); w[10] = cos(ang); w[11] = tan(ang); }
	}
	h->walls.insert(h->walls.end(), w, w + VGPU_WALL_STRIDE);
	if (hipMemcpy(h->d_walls.p, h->walls.data(), h->walls.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
		return VGPU_HIP_ERROR;
	h->computed = false;
	h->binned = false;
	return VGPU_OK;
}

int vgpu_add_wall_sphere(vgpu_container *h, double xc, double yc, double zc, double rc, int w_id) {
	const double prm[4] = {xc, yc, zc, rc};
	return vgpu_add_wall(h, 0, prm, w_id);
}

// container_base::point_inside (src/container.cc:484-487) with each wall's
// point_inside (src/wall.cc:17-19, 44-46, 55-60, 77-85).
int vgpu_point_inside(vgpu_container *h, double x, double y, double z, int *inside) {
	if (!h || !inside || h->prd) return VGPU_INVALID_ARGUMENT;
	const RectGeom &g = h->g;
	*inside = 0;
	if (x < g.ax || x > g.bx || y < g.ay || y > g.by || z < g.az || z > g.bz) return VGPU_OK;
	for (size_t o = 0; o < h->walls.size(); o += VGPU_WALL_STRIDE) {
		const double *w = h->walls.data() + o;
		const int type = (int)w[0];
		const double xc = w[2], yc = w[3], zc = w[4];
		bool in;
		if (type == 0) in = (x - xc) * (x - xc) + (y - yc) * (y - yc) + (z - zc) * (z - zc) < w[5] * w[5];
		else if (type == 1) in = x * xc + y * yc + z * zc < w[5];
		else {
			double xd = x - xc, yd = y - yc, zd = z - zc;
			double pa = (xd * w[5] + yd * w[6] + zd * w[7]) * w[8];
			xd -= w[5] * pa; yd -= w[6] * pa; zd -= w[7] * pa;
			if (type == 2) in = xd * xd + yd * yd + zd * zd < w[9] * w[9];
			else {
				pa *= w[11];
				if (pa < 0) in = false;
				else { pa *= pa; in = xd * xd + yd * yd + zd * zd < pa; }
			}
		}
		if (!in) return VGPU_OK;
	}
	*inside = 1;
	return VGPU_OK;
}

int vgpu_put(vgpu_container *h, int n, const int *ids, const double *xyz, const double *r) {
	if (!h || n < 0 || (n > 0 && (!ids || !xyz))) return VGPU_INVALID_ARGUMENT;
	if (h->poly && n > 0 && !r) return VGPU_INVALID_ARGUMENT;
	if (h->have_dev_input) {
		// switching back to host-side puts: start from the device contents
		h->h_ids.resize(h->n_in);
		h->h_xyz.resize((size_t)h->n_in * 3);
		h->h_r.resize(h->poly ? h->n_in : 0);
		if (h->n_in) {
			if (hipMemcpy(h->h_ids.data(), h->in_ids.p, (size_t)h->n_in * 4, hipMemcpyDeviceToHost) != hipSuccess ||
			    hipMemcpy(h->h_xyz.data(), h->in_xyz.p, (size_t)h->n_in * 24, hipMemcpyDeviceToHost) != hipSuccess)
				return VGPU_HIP_ERROR;
			if (h->poly &&
			    hipMemcpy(h->h_r.data(), h->in_r.p, (size_t)h->n_in * 8, hipMemcpyDeviceToHost) != hipSuccess)
				return VGPU_HIP_ERROR;
		}
		h->have_dev_input = false;
	}
	if ((long long)h->n_in + n > 2000000000LL / 32) return VGPU_MEMORY_ERROR;
	h->h_ids.insert(h->h_ids.end(), ids, ids + n);
	h->h_xyz.insert(h->h_xyz.end(), xyz, xyz + 3 * (size_t)n);
	if (h->poly) h->h_r.insert(h->h_r.end(), r, r + n);
	h->n_in += n;
	h->computed = false;
	h->binned = false;
	return VGPU_OK;
}

int vgpu_put_device(vgpu_container *h, int n, const int *d_ids, const double *d_xyz, const double *d_r) {
	if (!h || n < 0 || (n > 0 && (!d_ids || !d_xyz))) return VGPU_INVALID_ARGUMENT;
	if (h->poly && n > 0 && !d_r) return VGPU_INVALID_ARGUMENT;
	if (n > 2000000000 / 32) return VGPU_MEMORY_ERROR;
	int st;
	const size_t N1 = (size_t)std::max(n, 1);
	if ((st = h->in_ids.ensure(N1 * 4)) || (st = h->in_xyz.ensure(N1 * 24))) return st;
	if (n) {
		if (hipMemcpy(h->in_ids.p, d_ids, (size_t)n * 4, hipMemcpyDeviceToDevice) != hipSuccess ||
		    hipMemcpy(h->in_xyz.p, d_xyz, (size_t)n * 24, hipMemcpyDeviceToDevice) != hipSuccess)
			return VGPU_HIP_ERROR;
	}
	if (h->poly) {
		if ((st = h->in_r.ensure(N1 * 8))) return st;
		if (n && hipMemcpy(h->in_r.p, d_r, (size_t)n * 8, hipMemcpyDeviceToDevice) != hipSuccess) return VGPU_HIP_ERROR;
	}
	h->h_ids.clear(); h->h_xyz.clear(); h->h_r.clear();
	h->have_dev_input = true;
	h->n_in = n;
	h->computed = false;
	h->binned = false;
	return VGPU_OK;
}

int vgpu_clear(vgpu_container *h) {
	if (!h) return VGPU_INVALID_ARGUMENT;
	h->h_ids.clear(); h->h_xyz.clear(); h->h_r.clear();
	h->have_dev_input = false;
	h->n_in = 0;
	h->computed = false;
	h->binned = false;
	return VGPU_OK;
}

int vgpu_compute_all(vgpu_container *h, unsigned flags, void *stream) {
	if (!h || ((flags & VGPU_F_LEARNED) && !h->have_pred)) return VGPU_INVALID_ARGUMENT;
	// a cell kind = 2 | image << 3 | block << 4 must fit an int
	if ((flags & VGPU_F_LEARNED) && h->prd && h->nblocks >= (1 << 27)) return VGPU_INVALID_ARGUMENT;
	if (flags & VGPU_F_GEOMETRY) flags |= VGPU_F_NEIGHBORS | VGPU_F_CSR;
	// a capacity that overflowed (neighbour ids, cell tables) has been grown:
	// compute again (each retry at least doubles what was missing)
	int st = VGPU_RETRY;
	for (int attempt = 0; attempt < 6 && st == VGPU_RETRY; attempt++) st = run_pipeline(h, flags, (hipStream_t)stream);
	return st == VGPU_RETRY ? VGPU_MEMORY_ERROR : st;
}

namespace {

// The three find tiers (order-3, wide and huge mask windows) over n queries,
// or over the device list of them counted at *list_count; `failed` counts
// the queries no window held.  Periodic containers: touch as
// launch_find_cells.
int find_tiers(vgpu_container *h, const RectView &v, int n, const double *dq, const int *list, const int *list_count,
	       int *dfound, int *dpid, double *dpos, unsigned char *touch, hipStream_t s, int &failed) {
	const size_t N = (size_t)n;
	int st;
	failed = 0;
	if ((st = h->fq_retry.ensure(N * 4)) || (st = h->fq_retry2.ensure(N * 4)) || (st = h->d_counters.ensure(64 * 4)))
		return st;
	typedef Layout<TierWide, 64> LW;
	typedef Layout<TierHuge, 64> LH;
	const int w0 = std::max(1, std::min(default_fast_waves(2), nblk((long long)n, 64)));
	if ((st = h->ws_fast.ensure(VCell3::L::bytes, std::max(w0, h->ws_fast.waves))) ||
	    (st = h->ws_wide.ensure(LW::bytes, std::max(64, h->ws_wide.waves))) ||
	    (st = h->ws_huge.ensure(LH::bytes, std::max(4, h->ws_huge.waves))))
		return st;
	Workspace *wss[3] = {&h->ws_fast, &h->ws_wide, &h->ws_huge};
	for (Workspace *x : wss)
		if ((unsigned long long)x->stamp_base + N + 2 >= 0xffffffffull) {
			HIPCHK(hipMemset(x->buf.p, 0, x->buf.n));
			x->stamp_base = 0;
		}
	// [0] / [2] / [4] query counters of the three tiers, [1] / [3] their retry
	// counts, [5] queries no window held ([8] a caller's list count)
	int *ctr = h->d_counters.as<int>();
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr, 8, 0);
	const bool prd = h->prd;
	if (launch_find_cells(0, h->poly, prd, w0, s, v, n, dq, list, list_count, dfound, dpid, dpos,
			      h->fq_retry.as<int>(), ctr + 1, ctr, h->ws_fast.base(), h->ws_fast.slab,
			      h->ws_fast.stamp_base, touch))
		return VGPU_HIP_ERROR;
	h->ws_fast.stamp_base += (unsigned)n + 1u;
	// the few searches that left the order-3 window, with the wide one, then
	// the huge one (each launched only if it has queries)
	int retried = 0;
	HIPCHK(hipMemcpyAsync(&retried, ctr + 1, 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	if (retried > 0) {
		if (launch_find_cells(1, h->poly, prd, 64, s, v, n, dq, h->fq_retry.as<int>(), ctr + 1, dfound, dpid, dpos,
				      h->fq_retry2.as<int>(), ctr + 3, ctr + 2, h->ws_wide.base(), h->ws_wide.slab,
				      h->ws_wide.stamp_base, touch))
			return VGPU_HIP_ERROR;
		h->ws_wide.stamp_base += (unsigned)n + 1u;
		HIPCHK(hipMemcpyAsync(&retried, ctr + 3, 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
	}
	if (retried > 0) {
		if (launch_find_cells(2, h->poly, prd, 4, s, v, n, dq, h->fq_retry2.as<int>(), ctr + 3, dfound, dpid, dpos,
				      nullptr, ctr + 5, ctr + 4, h->ws_huge.base(), h->ws_huge.slab, h->ws_huge.stamp_base,
				      touch))
			return VGPU_HIP_ERROR;
		h->ws_huge.stamp_base += (unsigned)n + 1u;
		HIPCHK(hipMemcpyAsync(&failed, ctr + 5, 4, hipMemcpyDeviceToHost, s));
	}
	HIPCHK(hipStreamSynchronize(s));
	return VGPU_OK;
}

// Copies the binned primary blocks to the host (the image creation reads
// them there).
int prd_host_blocks(vgpu_container *h, hipStream_t s) {
	const int nc = h->ncells, nb = h->nblocks;
	int st;
	if ((st = h->hP.ensure((size_t)std::max(nc, 1) * sizeof(dbl4))) || (st = h->hpid.ensure((size_t)std::max(nc, 1) * 4)) ||
	    (st = h->hcs.ensure((size_t)(nb + 1) * 4)))
		return st;
	if (nc) {
		HIPCHK(hipMemcpyAsync(h->hP.p, h->P.p, (size_t)nc * sizeof(dbl4), hipMemcpyDeviceToHost, s));
		HIPCHK(hipMemcpyAsync(h->hpid.p, h->pid.p, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
	}
	HIPCHK(hipMemcpyAsync(h->hcs.p, h->bstart.p, (size_t)(nb + 1) * 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	return VGPU_OK;
}

// find_voronoi_cell on a periodic container.  The reference's queries create
// the image blocks they reach, one query after another, on top of those the
// last compute_all_cells (and earlier queries) created; the particle order
// inside an image block is its order of creation.  As for the cells
// (run_prd_cells), the order is found by a fixed point: the queries run
// against the blocks created so far plus those near the primary domain,
// recording first visits; the blocks they reached are then created in the
// order of those visits and the queries that read an image block run again,
// until the order reproduces itself.
int find_prd(vgpu_container *h, int n, const double *dq, int *dfound, int *dpid, double *dpos, hipStream_t s,
	     int &failed) {
	const PrdGeom &g = h->pg;
	const int nb = h->nblocks, nc = h->ncells;
	const size_t N = (size_t)n;
	int st;
	if ((st = prd_host_blocks(h, s))) return st;
	std::vector<unsigned char> prior(nb, 0);
	for (int b : h->prd_order) prior[b] = 1;
	std::vector<int> order = h->prd_order;
	for (int b : prd_margin_images(g, PRD_FIRST_MARGIN))
		if (!prior[b]) order.push_back(b);
	if ((st = h->fv.ensure((size_t)nb * 8)) || (st = h->fq_touch.ensure(N)) || (st = h->fq_list.ensure(N * 4)) ||
	    (st = h->d_counters.ensure(64 * 4)))
		return st;
	RectView v = make_view(h);
	unsigned long long *fv = h->fv.as<unsigned long long>();
	unsigned char *touch = h->fq_touch.as<unsigned char>();
	int *lc = h->d_counters.as<int>() + 8;
	const int *list = nullptr, *list_count = nullptr;
	std::vector<unsigned long long> keys(nb);
	std::vector<unsigned char> t(N);
	for (int pass = 0; pass < 8; pass++) {
		ImageSet im(g, h->hP.as<dbl4>(), h->hpid.as<int>(), h->hcs.as<int>());
		for (int b : order) im.create_block(b);
		if (im.late) return VGPU_INTERNAL_ERROR;
		if ((st = upload_blocks(h, im, nc, v, s))) return st;
		v.fv = fv;
		hipLaunchKernelGGL(k_fill_u64, dim3(nblk(nb, 256)), dim3(256), 0, s, fv, nb, ~0ull);
		hipLaunchKernelGGL(k_fill_u8, dim3(nblk(n, 256)), dim3(256), 0, s, touch, n, 0);
		if ((st = find_tiers(h, v, n, dq, list, list_count, dfound, dpid, dpos, touch, s, failed))) return st;
		HIPCHK(hipMemcpyAsync(keys.data(), fv, (size_t)nb * 8, hipMemcpyDeviceToHost, s));
		HIPCHK(hipMemcpyAsync(t.data(), touch, N, hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		std::vector<int> rerun;
		for (int q = 0; q < n; q++)
			if (t[q]) rerun.push_back(q);
		if (rerun.empty()) return VGPU_OK;   // no query read an image block: none is created
		std::vector<int> next = h->prd_order;
		for (int b : prd_visit_order(g, keys.data()))
			if (!prior[b]) next.push_back(b);
		if (next == order) {
			h->prd_order.swap(next);
			return VGPU_OK;
		}
		order.swap(next);
		const int m = (int)rerun.size();
		HIPCHK(hipMemcpyAsync(h->fq_list.p, rerun.data(), (size_t)m * 4, hipMemcpyHostToDevice, s));
		HIPCHK(hipMemcpyAsync(lc, &m, 4, hipMemcpyHostToDevice, s));
		list = h->fq_list.as<int>();
		list_count = lc;
	}
	return VGPU_INTERNAL_ERROR;   // the creation order did not settle
}

}  // namespace

int vgpu_find_voronoi_cells(vgpu_container *h, int n, const double *xyz, int *found, int *pid, double *pos,
			    int on_device) {
	if (!h || n < 0 || (n > 0 && (!xyz || !found || !pid || !pos))) return VGPU_INVALID_ARGUMENT;
	if (n == 0) return VGPU_OK;
	hipStream_t s = nullptr;
	int st;
	if (!h->binned && (st = run_pipeline(h, 0, s, true))) return st;
	const size_t N = (size_t)n;
	if (h->ncells == 0) {
		// no particle: no point has a cell (src/container.cc:260-261)
		if (on_device) HIPCHK(hipMemset(found, 0, N * 4));
		else std::memset(found, 0, N * 4);
		return VGPU_OK;
	}
	const double *dq = xyz;
	int *dfound = found, *dpid = pid;
	double *dpos = pos;
	if (!on_device) {
		if ((st = h->fq_xyz.ensure(N * 24)) || (st = h->fq_found.ensure(N * 4)) || (st = h->fq_pid.ensure(N * 4)) ||
		    (st = h->fq_pos.ensure(N * 24)))
			return st;
		HIPCHK(hipMemcpyAsync(h->fq_xyz.p, xyz, N * 24, hipMemcpyHostToDevice, s));
		dq = h->fq_xyz.as<double>(); dfound = h->fq_found.as<int>(); dpid = h->fq_pid.as<int>();
		dpos = h->fq_pos.as<double>();
	}
	int failed = 0;
	if (h->prd) st = find_prd(h, n, dq, dfound, dpid, dpos, s, failed);
	else st = find_tiers(h, make_view(h), n, dq, nullptr, nullptr, dfound, dpid, dpos, nullptr, s, failed);
	if (st) return st;
	if (!on_device) {
		HIPCHK(hipMemcpyAsync(found, dfound, N * 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipMemcpyAsync(pid, dpid, N * 4, hipMemcpyDeviceToHost, s));
		HIPCHK(hipMemcpyAsync(pos, dpos, N * 24, hipMemcpyDeviceToHost, s));
	}
	HIPCHK(hipStreamSynchronize(s));
	return failed > 0 ? VGPU_MEMORY_ERROR : VGPU_OK;
}

int vgpu_compute_ghost_cells(vgpu_container *h, int n, const double *xyz, const double *r, int *exists, double *vol,
			     double *cent, int *nfaces, int *nbr, int max_nbr) {
	if (!h || n < 0 || max_nbr < 1 || (n > 0 && (!xyz || !exists || !vol || !cent || !nfaces || !nbr)) ||
	    (h->poly && n > 0 && !r))
		return VGPU_INVALID_ARGUMENT;
	if (n == 0) return VGPU_OK;
	hipStream_t s = nullptr;
	int st;
	if (!h->binned && (st = run_pipeline(h, 0, s, true))) return st;
	const size_t N = (size_t)n;
	RectView v = make_view(h);
	if (h->prd) {
		// The block arrays as the reference holds them before the call: the
		// primary blocks and the image blocks created by the last compute and
		// queries.  A ghost point's search that needs another image block is
		// given up: the reference would create that block during the call and
		// copy the point itself into it, with an id it never sets.
		if ((st = prd_host_blocks(h, s))) return st;
		const int nb = h->nblocks;
		ImageSet im(h->pg, h->hP.as<dbl4>(), h->hpid.as<int>(), h->hcs.as<int>());
		std::vector<unsigned char> ok(nb, 0);
		for (int b = 0; b < nb; b++) ok[b] = h->pg.primary(b) ? 1 : 0;
		for (int b : h->prd_order) { im.create_block(b); ok[b] = 1; }
		if (im.late) return VGPU_INTERNAL_ERROR;
		if ((st = upload_blocks(h, im, h->ncells, v, s)) || (st = h->img_okb.ensure((size_t)nb))) return st;
		HIPCHK(hipMemcpy(h->img_okb.p, ok.data(), (size_t)nb, hipMemcpyHostToDevice));
		v.img_ok = h->img_okb.as<unsigned char>();
	}
	if ((st = h->fq_xyz.ensure(N * 24)) || (st = h->fq_r.ensure(N * 8)) || (st = h->fq_found.ensure(N * 4)) ||
	    (st = h->fq_vol.ensure(N * 8)) || (st = h->fq_pos.ensure(N * 24)) || (st = h->fq_pid.ensure(N * 4)) ||
	    (st = h->fq_nbr.ensure(N * (size_t)max_nbr * 4)) || (st = h->d_counters.ensure(64 * 4)))
		return st;
	HIPCHK(hipMemcpyAsync(h->fq_xyz.p, xyz, N * 24, hipMemcpyHostToDevice, s));
	if (h->poly) HIPCHK(hipMemcpyAsync(h->fq_r.p, r, N * 8, hipMemcpyHostToDevice, s));
	typedef Layout<TierWide, 64> LW;
	const int waves = std::max(1, std::min(64, nblk((long long)n, 64)));
	if ((st = h->ws_wide.ensure(LW::bytes, std::max(waves, h->ws_wide.waves)))) return st;
	if ((unsigned long long)h->ws_wide.stamp_base + N + 2 >= 0xffffffffull) {
		HIPCHK(hipMemset(h->ws_wide.buf.p, 0, h->ws_wide.buf.n));
		h->ws_wide.stamp_base = 0;
	}
	if ((st = h->fq_retry.ensure(N * 4))) return st;
	// [0] / [2] point counters of the wide and huge tiers, [1] the wide
	// tier's retries, [3] points no tier finished
	int *ctr = h->d_counters.as<int>();
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, ctr, 8, 0);
	if (launch_ghost_cells(0, h->poly, h->prd, waves, s, v, h->g, n, h->fq_xyz.as<double>(), h->fq_r.as<double>(), nullptr,
			       nullptr, h->fq_found.as<int>(), h->fq_vol.as<double>(), h->fq_pos.as<double>(),
			       h->fq_pid.as<int>(), h->fq_nbr.as<int>(), max_nbr, ctr, h->fq_retry.as<int>(), ctr + 1,
			       h->ws_wide.base(), h->ws_wide.slab, h->ws_wide.stamp_base))
		return VGPU_HIP_ERROR;
	h->ws_wide.stamp_base += (unsigned)n + 1u;
	int retried = 0;
	HIPCHK(hipMemcpyAsync(&retried, ctr + 1, 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	if (retried > 0) {
		// the points whose cell outgrew the wide cell, with the huge one
		typedef Layout<TierHuge, 64> LH;
		const int hw = std::max(1, std::min(4, nblk((long long)retried, 64)));
		if ((st = h->ws_huge.ensure(LH::bytes, std::max(hw, h->ws_huge.waves)))) return st;
		if ((unsigned long long)h->ws_huge.stamp_base + N + 2 >= 0xffffffffull) {
			HIPCHK(hipMemset(h->ws_huge.buf.p, 0, h->ws_huge.buf.n));
			h->ws_huge.stamp_base = 0;
		}
		if (launch_ghost_cells(1, h->poly, h->prd, hw, s, v, h->g, n, h->fq_xyz.as<double>(), h->fq_r.as<double>(),
				       h->fq_retry.as<int>(), ctr + 1, h->fq_found.as<int>(), h->fq_vol.as<double>(),
				       h->fq_pos.as<double>(), h->fq_pid.as<int>(), h->fq_nbr.as<int>(), max_nbr, ctr + 2,
				       nullptr, ctr + 3, h->ws_huge.base(), h->ws_huge.slab, h->ws_huge.stamp_base))
			return VGPU_HIP_ERROR;
		h->ws_huge.stamp_base += (unsigned)n + 1u;
	}
	int failed = 0;
	HIPCHK(hipMemcpyAsync(exists, h->fq_found.p, N * 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(vol, h->fq_vol.p, N * 8, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(cent, h->fq_pos.p, N * 24, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(nfaces, h->fq_pid.p, N * 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(nbr, h->fq_nbr.p, N * (size_t)max_nbr * 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipMemcpyAsync(&failed, ctr + 3, 4, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	return failed > 0 ? VGPU_MEMORY_ERROR : VGPU_OK;
}

int vgpu_get_put_index(vgpu_container *h, int *put_index) {
	if (!h || !put_index) return VGPU_INVALID_ARGUMENT;
	if (!h->binned) return VGPU_INTERNAL_ERROR;
	long long n = 0;
	int st = vgpu_num_cells(h, &n);
	if (st) return st;
	if (n > 0) HIPCHK(hipMemcpy(put_index, h->src.p, (size_t)n * 4, hipMemcpyDeviceToHost));
	return VGPU_OK;
}

int vgpu_kind(vgpu_container *h, int *kind) {
	if (!h || !kind) return VGPU_INVALID_ARGUMENT;
	*kind = h->kind;
	return VGPU_OK;
}

int vgpu_bin(vgpu_container *h) {
	if (!h) return VGPU_INVALID_ARGUMENT;
	if (h->binned) return VGPU_OK;
	return run_pipeline(h, 0, nullptr, true);
}

int vgpu_get_particles(vgpu_container *h, long long *n, int *id, double *pos, double *r) {
	if (!h) return VGPU_INVALID_ARGUMENT;
	if (!h->binned) return VGPU_INTERNAL_ERROR;
	const int nc = h->ncells;
	if (n) *n = nc;
	if (nc == 0 || (!id && !pos && !r)) return VGPU_OK;
	if (id) HIPCHK(hipMemcpy(id, h->pid.p, (size_t)nc * 4, hipMemcpyDeviceToHost));
	if (pos || r) {
		std::vector<dbl4> P(nc);
		HIPCHK(hipMemcpy(P.data(), h->P.p, (size_t)nc * 32, hipMemcpyDeviceToHost));
		for (int c = 0; c < nc; c++) {
			if (pos) { pos[3 * c] = P[c].x; pos[3 * c + 1] = P[c].y; pos[3 * c + 2] = P[c].z; }
			// the stored radius; default_radius (src/config.hh:90) for the plain kinds
			if (r) r[c] = h->poly ? P[c].w : 0.5;
		}
	}
	return VGPU_OK;
}

// c_loop_subset, src/c_loops.hh:218-258 and src/c_loops.cc:24-150: the
// loop's state machine over the stored block counts and positions, every
// step as the reference takes it (including its quirks: the first row
// starts at the wrapped index, later rows at the unwrapped one).
namespace {
struct SubsetLoop {
	enum Mode { no_check, sphere, box };
	Mode mode = no_check;
	int nx, ny, nz, nxy, nxyz;
	double ax, ay, az, sx, sy, sz, xsp, ysp, zsp;
	bool xper, yper, zper;
	const int *co;            // particles per block
	const int *bs;            // first loop index of each block
	const dbl4 *P;            // stored positions, loop order
	double px = 0, py = 0, pz = 0, apx = 0, apy = 0, apz = 0;
	double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
	int ai = 0, bi = 0, aj = 0, bj = 0, ak = 0, bk = 0;
	int ci = 0, cj = 0, ck = 0, di = 0, dj = 0, dk = 0, inc1 = 0, inc2 = 0;
	int i = 0, j = 0, k = 0, ijk = 0, q = 0;
	static int step_mod(int a, int b) { return a >= 0 ? a % b : b - 1 - (b - 1 - a) % b; }
	static int step_div(int a, int b) { return a >= 0 ? a / b : -1 + (a + 1) / b; }
	static int step_int(double a) { return a < 0 ? int(a) - 1 : int(a); }
	void setup_sphere(double vx, double vy, double vz, double r, bool bounds_test) {
		if (bounds_test) { mode = sphere; v0 = vx; v1 = vy; v2 = vz; v3 = r * r; } else mode = no_check;
		ai = step_int((vx - ax - r) * xsp);
		bi = step_int((vx - ax + r) * xsp);
		aj = step_int((vy - ay - r) * ysp);
		bj = step_int((vy - ay + r) * ysp);
		ak = step_int((vz - az - r) * zsp);
		bk = step_int((vz - az + r) * zsp);
		setup_common();
	}
	void setup_box(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax, bool bounds_test) {
		if (bounds_test) { mode = box; v0 = xmin; v1 = xmax; v2 = ymin; v3 = ymax; v4 = zmin; v5 = zmax; }
		else mode = no_check;
		ai = step_int((xmin - ax) * xsp);
		bi = step_int((xmax - ax) * xsp);
		aj = step_int((ymin - ay) * ysp);
		bj = step_int((ymax - ay) * ysp);
		ak = step_int((zmin - az) * zsp);
		bk = step_int((zmax - az) * zsp);
		setup_common();
	}
	void setup_intbox(int ai_, int bi_, int aj_, int bj_, int ak_, int bk_) {
		ai = ai_; bi = bi_; aj = aj_; bj = bj_; ak = ak_; bk = bk_;
		mode = no_check;
		setup_common();
	}
	void setup_common() {
		if (!xper) {
			if (ai < 0) { ai = 0; if (bi < 0) bi = 0; }
			if (bi >= nx) { bi = nx - 1; if (ai >= nx) ai = nx - 1; }
		}
		if (!yper) {
			if (aj < 0) { aj = 0; if (bj < 0) bj = 0; }
			if (bj >= ny) { bj = ny - 1; if (aj >= ny) aj = ny - 1; }
		}
		if (!zper) {
			if (ak < 0) { ak = 0; if (bk < 0) bk = 0; }
			if (bk >= nz) { bk = nz - 1; if (ak >= nz) ak = nz - 1; }
		}
		ci = ai; cj = aj; ck = ak;
		di = i = step_mod(ci, nx); apx = px = step_div(ci, nx) * sx;
		dj = j = step_mod(cj, ny); apy = py = step_div(cj, ny) * sy;
		dk = k = step_mod(ck, nz); apz = pz = step_div(ck, nz) * sz;
		inc1 = di - step_mod(bi, nx);
		inc2 = nx * (ny + dj - step_mod(bj, ny)) + inc1;
		inc1 += nx;
		ijk = di + nx * (dj + ny * dk);
		q = 0;
	}
	bool out_of_bounds() const {
		const dbl4 &pp = P[bs[ijk] + q];
		if (mode == sphere) {
			const double fx = pp.x + px - v0, fy = pp.y + py - v1, fz = pp.z + pz - v2;
			return fx * fx + fy * fy + fz * fz > v3;
		}
		double f = pp.x + px;
		if (f < v0 || f > v1) return true;
		f = pp.y + py;
		if (f < v2 || f > v3) return true;
		f = pp.z + pz;
		return f < v4 || f > v5;
	}
	bool next_block() {
		if (i < bi) {
			i++;
			if (ci < nx - 1) { ci++; ijk++; } else { ci = 0; ijk += 1 - nx; px += sx; }
			return true;
		} else if (j < bj) {
			i = ai; ci = di; px = apx; j++;
			if (cj < ny - 1) { cj++; ijk += inc1; } else { cj = 0; ijk += inc1 - nxy; py += sy; }
			return true;
		} else if (k < bk) {
			i = ai; ci = di; j = aj; cj = dj; px = apx; py = apy; k++;
			if (ck < nz - 1) { ck++; ijk += inc2; } else { ck = 0; ijk += inc2 - nxyz; pz += sz; }
			return true;
		}
		return false;
	}
	bool start() {
		while (co[ijk] == 0) { if (!next_block()) return false; }
		while (mode != no_check && out_of_bounds()) {
			q++;
			while (q >= co[ijk]) { q = 0; if (!next_block()) return false; }
		}
		return true;
	}
	bool inc() {
		do {
			q++;
			while (q >= co[ijk]) { q = 0; if (!next_block()) return false; }
		} while (mode != no_check && out_of_bounds());
		return true;
	}
};
}  // namespace

int vgpu_loop_subset(vgpu_container *h, int mode, const double *v, const int *iv, int bounds_test, long long *index,
		     int *ijk_pos, long long cap, long long *n) {
	if (!h || !n || h->prd || mode < 0 || mode > 2 || (mode < 2 && !v) || (mode == 2 && !iv)) return VGPU_INVALID_ARGUMENT;
	int st;
	if (!h->binned && (st = run_pipeline(h, 0, nullptr, true))) return st;
	const int nb = h->nblocks, nc = h->ncells;
	std::vector<int> bs(nb + 1), co(nb);
	std::vector<dbl4> P(std::max(nc, 1));
	HIPCHK(hipMemcpy(bs.data(), h->bstart.p, (size_t)(nb + 1) * 4, hipMemcpyDeviceToHost));
	if (nc) HIPCHK(hipMemcpy(P.data(), h->P.p, (size_t)nc * 32, hipMemcpyDeviceToHost));
	for (int b = 0; b < nb; b++) co[b] = bs[b + 1] - bs[b];
	const RectGeom &g = h->g;
	SubsetLoop L;
	L.nx = g.nx; L.ny = g.ny; L.nz = g.nz; L.nxy = g.nx * g.ny; L.nxyz = L.nxy * g.nz;
	L.ax = g.ax; L.ay = g.ay; L.az = g.az;
	L.sx = g.bx - g.ax; L.sy = g.by - g.ay; L.sz = g.bz - g.az;
	L.xsp = g.xsp; L.ysp = g.ysp; L.zsp = g.zsp;
	L.xper = g.xper; L.yper = g.yper; L.zper = g.zper;
	L.co = co.data(); L.bs = bs.data(); L.P = P.data();
	if (mode == 0) L.setup_sphere(v[0], v[1], v[2], v[3], bounds_test != 0);
	else if (mode == 1) L.setup_box(v[0], v[1], v[2], v[3], v[4], v[5], bounds_test != 0);
	else L.setup_intbox(iv[0], iv[1], iv[2], iv[3], iv[4], iv[5]);
	long long m = 0;
	if (nc > 0 && L.start()) do {
		if (m < cap) {
			if (index) index[m] = bs[L.ijk] + L.q;
			if (ijk_pos) { ijk_pos[3 * m] = L.i; ijk_pos[3 * m + 1] = L.j; ijk_pos[3 * m + 2] = L.k; }
		}
		m++;
	} while (L.inc());
	*n = m;
	return VGPU_OK;
}

int vgpu_num_cells(vgpu_container *h, long long *n) {
	if (!h || !n) return VGPU_INVALID_ARGUMENT;
	if (!h->computed) return VGPU_INVALID_ARGUMENT;
	*n = h->ncells;
	return VGPU_OK;
}

int vgpu_get_results(vgpu_container *h, int *id, int *ijk, int *q, int *exists, double *pos, double *vol,
		     double *centroid, int *nfaces, long long *nbr_off, int *nbr, long long nbr_cap,
		     long long *nbr_total) {
	if (!h || !h->computed) return VGPU_INVALID_ARGUMENT;
	const int nc = h->ncells;
	const int nxyz = h->nblocks;
	if (nc == 0) {
		if (nbr_off) nbr_off[0] = 0;
		if (nbr_total) *nbr_total = 0;
		return VGPU_OK;
	}
	std::vector<int> st(nc), nf(nc);
	HIPCHK(hipMemcpy(st.data(), h->status.p, (size_t)nc * 4, hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(nf.data(), h->nfaces.p, (size_t)nc * 4, hipMemcpyDeviceToHost));
	if (id) HIPCHK(hipMemcpy(id, h->pid.p, (size_t)nc * 4, hipMemcpyDeviceToHost));
	if (ijk || q) {
		std::vector<int> cb(nc), bs(nxyz + 1);
		HIPCHK(hipMemcpy(cb.data(), h->cellblk.p, (size_t)nc * 4, hipMemcpyDeviceToHost));
		HIPCHK(hipMemcpy(bs.data(), h->bstart.p, (size_t)(nxyz + 1) * 4, hipMemcpyDeviceToHost));
		for (int c = 0; c < nc; c++) {
			if (ijk) ijk[c] = cb[c];
			if (q) q[c] = c - bs[cb[c]];
		}
	}
	if (exists) for (int c = 0; c < nc; c++) exists[c] = (st[c] & ST_EXISTS) ? 1 : 0;
	if (pos) {
		std::vector<dbl4> P(nc);
		HIPCHK(hipMemcpy(P.data(), h->P.p, (size_t)nc * 32, hipMemcpyDeviceToHost));
		for (int c = 0; c < nc; c++) { pos[3 * c] = P[c].x; pos[3 * c + 1] = P[c].y; pos[3 * c + 2] = P[c].z; }
	}
	if (vol) HIPCHK(hipMemcpy(vol, h->vol.p, (size_t)nc * 8, hipMemcpyDeviceToHost));
	if (centroid) HIPCHK(hipMemcpy(centroid, h->cent.p, (size_t)nc * 24, hipMemcpyDeviceToHost));
	if (nfaces) for (int c = 0; c < nc; c++) nfaces[c] = (st[c] & ST_EXISTS) ? nf[c] : 0;
	long long tot = 0;
	std::vector<long long> off(nc + 1);
	for (int c = 0; c < nc; c++) { off[c] = tot; if (st[c] & ST_EXISTS) tot += nf[c]; }
	off[nc] = tot;
	if (nbr_off) std::memcpy(nbr_off, off.data(), (size_t)(nc + 1) * 8);
	if (nbr_total) *nbr_total = tot;
	if (nbr && (h->last_flags & VGPU_F_NEIGHBORS)) {
		if (h->csr_valid) {
			long long m = std::min(tot, nbr_cap);
			if (m > 0) HIPCHK(hipMemcpy(nbr, h->csr.p, (size_t)m * 4, hipMemcpyDeviceToHost));
		} else {
			std::vector<int> start(nc), arena((size_t)std::max<long long>(h->nbr_total, 1));
			HIPCHK(hipMemcpy(start.data(), h->nbr_start.p, (size_t)nc * 4, hipMemcpyDeviceToHost));
			if (h->nbr_total > 0)
				HIPCHK(hipMemcpy(arena.data(), h->arena.p, (size_t)h->nbr_total * 4, hipMemcpyDeviceToHost));
			for (int c = 0; c < nc; c++) {
				if (!(st[c] & ST_EXISTS)) continue;
				for (int t = 0; t < nf[c]; t++) {
					long long d = off[c] + t;
					if (d < nbr_cap) nbr[d] = arena[(size_t)start[c] + t];
				}
			}
		}
	}
	return VGPU_OK;
}

int vgpu_geometry_sizes(vgpu_container *h, long long *n_vertices, long long *n_edge_slots, long long *n_faces,
			long long *n_face_vertex_entries) {
	if (!h) return VGPU_INVALID_ARGUMENT;
	if (!h->computed || !h->geo_valid) return VGPU_INTERNAL_ERROR;
	if (n_vertices) *n_vertices = h->geo_nv;
	if (n_edge_slots) *n_edge_slots = h->geo_ne;
	if (n_faces) *n_faces = h->geo_nf;
	if (n_face_vertex_entries) *n_face_vertex_entries = h->geo_nfv;
	return VGPU_OK;
}

int vgpu_get_geometry(vgpu_container *h, long long *vert_off, double *vertices, int *vertex_orders, long long *edge_off,
		      int *edges, int *edge_rel, double *face_areas, double *face_perimeters, double *normals,
		      int *face_orders, long long *fv_off, int *face_vertices, double *surface_area,
		      double *total_edge_distance, int *number_of_edges, double *max_radius_squared) {
	if (!h) return VGPU_INVALID_ARGUMENT;
	if (!h->computed || !h->geo_valid) return VGPU_INTERNAL_ERROR;
	const int nc = h->ncells;
	hipStream_t s = nullptr;
	const size_t nv = (size_t)h->geo_nv, ne = (size_t)h->geo_ne, nfv = (size_t)h->geo_nfv, nf = (size_t)h->geo_nf;
	auto offs = [&](DBuf &d, long long *out) -> int {
		if (!out) return VGPU_OK;
		std::vector<int> t((size_t)nc + 1);
		if (hipMemcpy(t.data(), d.p, t.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return VGPU_HIP_ERROR;
		for (size_t i = 0; i < t.size(); i++) out[i] = t[i];
		return VGPU_OK;
	};
	int st;
	if ((st = offs(h->go_voff, vert_off)) || (st = offs(h->go_eoff, edge_off)) || (st = offs(h->go_fvoff, fv_off)))
		return st;
	struct Cp { void *dst; const DBuf *src; size_t bytes; } cps[] = {
		{vertices, &h->go_vert, nv * 24}, {vertex_orders, &h->go_ord, nv * 4}, {edges, &h->go_edges, ne * 4},
		{edge_rel, &h->go_erel, ne * 4}, {face_areas, &h->go_farea, nf * 8}, {face_perimeters, &h->go_fperim, nf * 8},
		{normals, &h->go_normal, nf * 24}, {face_orders, &h->go_forder, nf * 4}, {face_vertices, &h->go_fverts, nfv * 4},
		{surface_area, &h->go_surface, (size_t)nc * 8}, {total_edge_distance, &h->go_edist, (size_t)nc * 8},
		{number_of_edges, &h->go_nedges, (size_t)nc * 4}, {max_radius_squared, &h->go_maxr2, (size_t)nc * 8}};
	for (const Cp &c : cps)
		if (c.dst && c.bytes) HIPCHK(hipMemcpyAsync(c.dst, c.src->p, c.bytes, hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	return VGPU_OK;
}

int vgpu_get_stats(vgpu_container *h, int *n_planes, int *wl_exit, int *tier) {
	if (!h || !h->computed || !(h->last_flags & VGPU_F_STATS)) return VGPU_INVALID_ARGUMENT;
	const int nc = h->ncells;
	if (nc == 0) return VGPU_OK;
	std::vector<short> a(nc);
	std::vector<signed char> b(nc), c(nc);
	HIPCHK(hipMemcpy(a.data(), h->st_planes.p, (size_t)nc * 2, hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(b.data(), h->st_exit.p, (size_t)nc, hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(c.data(), h->st_tier.p, (size_t)nc, hipMemcpyDeviceToHost));
	for (int t = 0; t < nc; t++) {
		if (n_planes) n_planes[t] = a[t];
		if (wl_exit) wl_exit[t] = b[t];
		if (tier) tier[t] = c[t];
	}
	return VGPU_OK;
}

int vgpu_sum_volumes(vgpu_container *h, double *out) {
	if (!h || !out || !h->computed) return VGPU_INVALID_ARGUMENT;
	const int nc = h->ncells;
	std::vector<double> v(std::max(nc, 1));
	std::vector<int> st(std::max(nc, 1));
	if (nc) {
		HIPCHK(hipMemcpy(v.data(), h->vol.p, (size_t)nc * 8, hipMemcpyDeviceToHost));
		HIPCHK(hipMemcpy(st.data(), h->status.p, (size_t)nc * 4, hipMemcpyDeviceToHost));
	}
	// Sequential sum in loop order, as sum_cell_volumes() accumulates.
	double s = 0;
	for (int c = 0; c < nc; c++) if (st[c] & ST_EXISTS) s += v[c];
	*out = s;
	return VGPU_OK;
}

int vgpu_device_results(vgpu_container *h, const double **vol, const double **centroid, const int **nfaces,
			const int **status, const int **nbr_start, const int **nbr) {
	if (!h || !h->computed) return VGPU_INVALID_ARGUMENT;
	if (vol) *vol = h->vol.as<double>();
	if (centroid) *centroid = h->cent.as<double>();
	if (nfaces) *nfaces = h->nfaces.as<int>();
	if (status) *status = h->status.as<int>();
	if (nbr_start) *nbr_start = h->nbr_start.as<int>();
	if (nbr) *nbr = h->arena.as<int>();
	return VGPU_OK;
}

int vgpu_set_launch(vgpu_container *h, int fast_waves, int fast_occupancy) {
	if (!h || fast_waves < 0 || fast_occupancy < 0 || fast_occupancy > 4) return VGPU_INVALID_ARGUMENT;
	h->fast_waves = fast_waves;
	if (fast_occupancy > 0) h->fast_occ = fast_occupancy < 2 ? 2 : fast_occupancy;
	return VGPU_OK;
}

int vgpu_set_first_tier(vgpu_container *h, int tier, int waves, int occupancy) {
	if (!h || tier < 0 || tier > 1 || waves < 0 || occupancy < 0 || occupancy > 8 || occupancy == 1 ||
	    occupancy == 7)
		return VGPU_INVALID_ARGUMENT;
	h->first_tier = tier;
	h->wave_waves = waves;
	if (occupancy) h->wave_occ = occupancy;
	return VGPU_OK;
}

int vgpu_autotune(vgpu_container *h, unsigned flags, int reps, vgpu_tune_trial *trials, int max_trials,
		  int *n_trials) {
	if (!h || reps < 1 || max_trials < 0 || (max_trials > 0 && !trials) || (flags & VGPU_F_GEOMETRY))
		return VGPU_INVALID_ARGUMENT;
	const bool learned = (flags & VGPU_F_LEARNED) != 0;
	if (learned && !h->have_pred) return VGPU_INVALID_ARGUMENT;
	const int keep_tier = h->first_tier, keep_wocc = h->wave_occ, keep_focc = h->fast_occ;
	const float keep_margin = h->pred_margin;
	std::vector<vgpu_tune_trial> cand;
	// Times configuration c `reps` times (fastest kept); false on an error.
	int st = VGPU_OK;
	auto trial = [&](vgpu_tune_trial c) -> bool {
		h->first_tier = c.first_tier;
		if (c.first_tier == 0) h->wave_occ = c.occupancy;
		else h->fast_occ = c.occupancy;
		h->pred_margin = std::max(0.f, keep_margin + c.depth_offset);
		c.depth_offset = h->pred_margin - keep_margin;
		c.waves = c.first_tier == 0 ? (h->wave_waves > 0 ? h->wave_waves : default_fast_waves(c.occupancy))
					    : (h->fast_waves > 0 ? h->fast_waves : default_fast_waves(c.occupancy));
		double ms = 1e300;
		for (int r = 0; r < reps; r++) {
			if ((st = vgpu_compute_all(h, flags, nullptr))) return false;
			double tot = 0;
			for (int k = 0; k < 4; k++) tot += h->timings[k];
			ms = std::min(ms, tot);
		}
		c.ms = ms;
		cand.push_back(c);
		return true;
	};
	// The fastest, except that a configuration within 1% of it with fewer
	// waves per SIMD is preferred: more waves hold fewer registers each and
	// spill more to memory, for gains that small.
	auto best_of = [&]() {
		int f = 0;
		for (size_t t = 1; t < cand.size(); t++)
			if (cand[t].ms < cand[f].ms) f = (int)t;
		int bi = f;
		for (size_t t = 0; t < cand.size(); t++)
			if (cand[t].ms <= 1.01 * cand[f].ms && cand[t].first_tier == cand[f].first_tier &&
			    cand[t].depth_offset == cand[f].depth_offset && cand[t].occupancy < cand[bi].occupancy)
				bi = (int)t;
		return bi;
	};
	// Stage 1: the first tier -- one wavefront cell (vertex tables in
	// registers) at each compiled register budget (periodic containers: 4, 5
	// and 6 waves per SIMD) or, on the plain path, 64 order-3 cells per
	// wavefront at each budget.
	// Stage 2, learned path: the depth margin around the trained one at the
	// best budget.
	bool ok = true;
	for (int o : {3, 4, 5, 6})
		if (ok && !(h->prd && o == 3)) ok = trial(vgpu_tune_trial{0, o, 0, 0.f, 0.0});
	if (!learned)
		for (int o : {2, 3, 4})
			if (ok) ok = trial(vgpu_tune_trial{1, o, 0, 0.f, 0.0});
	if (ok && learned) {
		const vgpu_tune_trial b = cand[best_of()];
		for (float d : {-2.f, -1.f, 1.f, 2.f, 4.f})
			if (ok) ok = trial(vgpu_tune_trial{b.first_tier, b.occupancy, 0, d, 0.0});
	}
	if (!ok) {
		h->first_tier = keep_tier; h->wave_occ = keep_wocc; h->fast_occ = keep_focc;
		h->pred_margin = keep_margin;
		return st;
	}
	const int bi = best_of();
	const vgpu_tune_trial &b = cand[bi];
	h->first_tier = b.first_tier;
	h->wave_occ = b.first_tier == 0 ? b.occupancy : keep_wocc;
	h->fast_occ = b.first_tier == 1 ? b.occupancy : keep_focc;
	h->pred_margin = keep_margin + b.depth_offset;
	// the best first, then the rest in the order tried
	int m = 0;
	if (max_trials > 0) trials[m++] = b;
	for (size_t t = 0; t < cand.size() && m < max_trials; t++)
		if ((int)t != bi) trials[m++] = cand[t];
	if (n_trials) *n_trials = (int)cand.size();
	return VGPU_OK;
}

int vgpu_last_timings(vgpu_container *h, double *ms4) {
	if (!h || !ms4) return VGPU_INVALID_ARGUMENT;
	for (int t = 0; t < 4; t++) ms4[t] = h->timings[t];
	return VGPU_OK;
}

int vgpu_set_predictor(vgpu_container *h, vgpu_predictor *p) {
	if (!h) return VGPU_INVALID_ARGUMENT;
	if (!p) { h->have_pred = false; return VGPU_OK; }
	if (!p->trained || p->params.size() != (size_t)vgpu::learn::NP) return VGPU_INVALID_ARGUMENT;
	if (p->with_radius && !h->poly) return VGPU_INVALID_ARGUMENT;
	int st = h->d_pred.ensure(sizeof(float) * vgpu::learn::NP);
	if (st) return st;
	HIPCHK(hipMemcpy(h->d_pred.p, p->params.data(), sizeof(float) * vgpu::learn::NP, hipMemcpyHostToDevice));
	h->pred_margin = p->margin;
	h->pred_route = p->route;
	h->pred_with_r = p->with_radius;
	h->have_pred = true;
	return VGPU_OK;
}

int vgpu_learned_stats(vgpu_container *h, long long *fallbacks, long long *list_entries, double *gather_ms,
		       long long *routed) {
	if (!h) return VGPU_INVALID_ARGUMENT;
	if (routed) *routed = h->last_routed;
	if (gather_ms) *gather_ms = h->gather_ms;
	if (fallbacks) *fallbacks = h->last_fallbacks;
	if (list_entries) *list_entries = h->last_entries;
	return VGPU_OK;
}

// Training samples from the last vgpu_compute_all (run with VGPU_F_STATS):
// per block (or per cell, for a predictor that takes the radius) the
// features and the largest worklist exit and nplane count.
int vgpu_predictor_add_samples(vgpu_predictor *p, vgpu_container *h) {
	using namespace vgpu::learn;
	if (!p || !h || !h->computed || !(h->last_flags & VGPU_F_STATS)) return VGPU_INVALID_ARGUMENT;
	if (p->with_radius && !h->poly) return VGPU_INVALID_ARGUMENT;
	const int nc = h->ncells;
	if (nc <= 0) return VGPU_OK;
	hipStream_t s = nullptr;
	int st, nxyz = 0;
	if ((st = container_features(h, nc, &nxyz, s))) return st;
	std::vector<float> X;
	std::vector<int> lab;
	int ns;
	if (p->with_radius) {
		if ((st = h->d_cfeat.ensure((size_t)nc * IN * 4))) return st;
		const float inv = h->max_radius > 0 ? (float)(1.0 / h->max_radius) : 0.f;
		if ((st = cell_features_dev(nc, h->cellblk.as<int>(), h->P.as<double>() + 3, inv, h->d_feat.as<float>(),
					    h->d_cfeat.as<float>(), s)))
			return st;
		ns = nc;
		X.resize((size_t)nc * IN);
		HIPCHK(hipMemcpy(X.data(), h->d_cfeat.p, X.size() * 4, hipMemcpyDeviceToHost));
		std::vector<signed char> ex(nc);
		std::vector<short> pl(nc);
		HIPCHK(hipMemcpy(ex.data(), h->st_exit.p, (size_t)nc, hipMemcpyDeviceToHost));
		HIPCHK(hipMemcpy(pl.data(), h->st_planes.p, (size_t)nc * 2, hipMemcpyDeviceToHost));
		lab.resize((size_t)2 * nc);
		for (int c = 0; c < nc; c++) { lab[2 * c] = ex[c]; lab[2 * c + 1] = pl[c]; }
	} else {
		DBuf d_lab;
		if ((st = d_lab.ensure((size_t)nxyz * 8))) return st;
		if ((st = block_labels_dev(nc, h->cellblk.as<int>(), h->st_exit.as<signed char>(), h->st_planes.as<short>(),
					   nxyz, d_lab.as<int>(), s))) {
			d_lab.release();
			return st;
		}
		X.resize((size_t)nxyz * IN);
		lab.resize((size_t)2 * nxyz);
		std::vector<int> bs(nxyz + 1);
		bool ok = hipMemcpy(X.data(), h->d_feat.p, X.size() * 4, hipMemcpyDeviceToHost) == hipSuccess &&
			  hipMemcpy(lab.data(), d_lab.p, lab.size() * 4, hipMemcpyDeviceToHost) == hipSuccess &&
			  hipMemcpy(bs.data(), h->bstart.p, bs.size() * 4, hipMemcpyDeviceToHost) == hipSuccess;
		d_lab.release();
		if (!ok) return VGPU_HIP_ERROR;
		// only blocks that hold cells carry a label
		int m = 0;
		for (int b = 0; b < nxyz; b++) {
			if (bs[b + 1] == bs[b]) continue;
			std::memmove(&X[(size_t)m * IN], &X[(size_t)b * IN], IN * 4);
			lab[2 * m] = lab[2 * b];
			lab[2 * m + 1] = lab[2 * b + 1];
			m++;
		}
		ns = m;
	}
	for (int t = 0; t < ns; t++) {
		p->X.insert(p->X.end(), &X[(size_t)t * IN], &X[(size_t)t * IN] + IN);
		p->T.push_back((float)lab[2 * t] / 64.f);
		p->T.push_back(std::log1p((float)lab[2 * t + 1]) / 8.f);
	}
	return VGPU_OK;
}

// Training samples from the last vgpu_batch_compute (run with
// VGPU_F_STATS): per block of every container, its features and the largest
// worklist exit and nplane count of its cells (a block predictor only).
int vgpu_predictor_add_batch_samples(vgpu_predictor *p, vgpu_batch *b) {
	using namespace vgpu::learn;
	if (!p || !b || !b->computed || !(b->last_flags & VGPU_F_STATS) || p->with_radius) return VGPU_INVALID_ARGUMENT;
	const int nb = b->nblocks;
	if (nb <= 0 || b->n <= 0) return VGPU_OK;
	int st;
	if ((st = b->bcont.ensure(((size_t)nb + 1) * 4)) || (st = b->d_feat.ensure(((size_t)nb + 1) * IN * 4))) return st;
	if (!b->bcont_ok) {
		if (launch_batch_bcont(nullptr, b->ncont, b->boff.as<int>(), b->bcont.as<int>())) return VGPU_HIP_ERROR;
		b->bcont_ok = true;
	}
	if (launch_batch_features(nullptr, nb, b->bcont.as<int>(), b->views.as<RectView>(), b->d_feat.as<float>()))
		return VGPU_HIP_ERROR;
	std::vector<int> bs((size_t)nb + 1);
	HIPCHK(hipMemcpy(bs.data(), b->bstart.p, bs.size() * 4, hipMemcpyDeviceToHost));
	const int nc = bs[nb];
	DBuf d_lab;
	if ((st = d_lab.ensure((size_t)nb * 8))) return st;
	if ((st = block_labels_dev(nc, b->cellblk.as<int>(), b->st_exit.as<signed char>(), b->st_planes.as<short>(), nb,
				   d_lab.as<int>(), nullptr))) {
		d_lab.release();
		return st;
	}
	std::vector<float> X((size_t)nb * IN);
	std::vector<int> lab((size_t)2 * nb);
	const bool ok = hipMemcpy(X.data(), b->d_feat.p, X.size() * 4, hipMemcpyDeviceToHost) == hipSuccess &&
			hipMemcpy(lab.data(), d_lab.p, lab.size() * 4, hipMemcpyDeviceToHost) == hipSuccess;
	d_lab.release();
	if (!ok) return VGPU_HIP_ERROR;
	for (int q = 0; q < nb; q++) {
		if (bs[q + 1] == bs[q]) continue;   // only blocks that hold cells carry a label
		p->X.insert(p->X.end(), &X[(size_t)q * IN], &X[(size_t)q * IN] + IN);
		p->T.push_back((float)lab[2 * q] / 64.f);
		p->T.push_back(std::log1p((float)lab[2 * q + 1]) / 8.f);
	}
	return VGPU_OK;
}

int vgpu_predictor_clear_samples(vgpu_predictor *p) {
	if (!p) return VGPU_INVALID_ARGUMENT;
	p->X.clear();
	p->T.clear();
	return VGPU_OK;
}


// ---- batches of containers ----

int vgpu_batch_create(vgpu_container *const *hs, int count, vgpu_batch **out) {
	if (!out || count <= 0 || !hs) return VGPU_INVALID_ARGUMENT;
	*out = nullptr;
	const int kind = hs[0] ? hs[0]->kind : -1;
	if (kind != VGPU_CONTAINER && kind != VGPU_CONTAINER_POLY) return VGPU_INVALID_ARGUMENT;
	long long n = 0, nb = 0, nw = 0;
	for (int c = 0; c < count; c++) {
		if (!hs[c] || hs[c]->kind != kind) return VGPU_INVALID_ARGUMENT;
		n += hs[c]->n_in;
		nb += hs[c]->nblocks;
		nw += (long long)(hs[c]->walls.size() / VGPU_WALL_STRIDE);
	}
	if (n > 2000000000LL / 32 || nb + 1 > 2000000000LL) return VGPU_MEMORY_ERROR;
	vgpu_batch *b = new vgpu_batch();
	b->ncont = count;
	b->poly = kind == VGPU_CONTAINER_POLY;
	b->n = (int)n;
	b->nblocks = (int)nb;
	b->h_boff.assign(count + 1, 0);
	b->h_pfirst.assign(count + 1, 0);
	std::vector<int> woff(count + 1, 0);
	for (int c = 0; c < count; c++) {
		b->h_boff[c + 1] = b->h_boff[c] + hs[c]->nblocks;
		b->h_pfirst[c + 1] = b->h_pfirst[c] + hs[c]->n_in;
		woff[c + 1] = woff[c] + (int)(hs[c]->walls.size() / VGPU_WALL_STRIDE);
	}
	int st = VGPU_OK;
	auto fail = [&](int e) { vgpu_batch_destroy(b); return e; };
	const size_t N1 = (size_t)std::max(b->n, 1), C1 = (size_t)count + 1;
	if ((st = b->in_ids.ensure(N1 * 4)) || (st = b->in_xyz.ensure(N1 * 24)) || (st = b->in_r.ensure(N1 * 8)) ||
	    (st = b->geoms.ensure(count * sizeof(RectGeom))) || (st = b->boff.ensure(C1 * 4)) ||
	    (st = b->pfirst.ensure(C1 * 4)) || (st = b->pcont.ensure(N1 * 4)) ||
	    (st = b->mrad.ensure((size_t)count * 4096 * 8)) || (st = b->d_wl.ensure(4096 * 4)) ||
	    (st = b->walls.ensure((size_t)std::max<long long>(nw, 1) * VGPU_WALL_STRIDE * 8)) ||
	    (st = b->views.ensure(count * sizeof(RectView))))
		return fail(st);
	// particles, in container order
	std::vector<int> ids((size_t)b->n);
	std::vector<double> xyz((size_t)b->n * 3), r(b->poly ? (size_t)b->n : 0), wall((size_t)nw * VGPU_WALL_STRIDE);
	std::vector<RectGeom> geo(count);
	for (int c = 0; c < count; c++) {
		vgpu_container *h = hs[c];
		const size_t o = (size_t)b->h_pfirst[c], m = (size_t)h->n_in;
		if (h->have_dev_input) {
			if (m && (hipMemcpy(ids.data() + o, h->in_ids.p, m * 4, hipMemcpyDeviceToHost) != hipSuccess ||
				  hipMemcpy(xyz.data() + 3 * o, h->in_xyz.p, m * 24, hipMemcpyDeviceToHost) != hipSuccess ||
				  (b->poly && hipMemcpy(r.data() + o, h->in_r.p, m * 8, hipMemcpyDeviceToHost) != hipSuccess)))
				return fail(VGPU_HIP_ERROR);
		} else if (m) {
			std::memcpy(ids.data() + o, h->h_ids.data(), m * 4);
			std::memcpy(xyz.data() + 3 * o, h->h_xyz.data(), m * 24);
			if (b->poly) std::memcpy(r.data() + o, h->h_r.data(), m * 8);
		}
		if (!h->walls.empty())
			std::memcpy(wall.data() + (size_t)woff[c] * VGPU_WALL_STRIDE, h->walls.data(), h->walls.size() * 8);
		geo[c] = h->g;
	}
	if ((b->n && (hipMemcpy(b->in_ids.p, ids.data(), (size_t)b->n * 4, hipMemcpyHostToDevice) != hipSuccess ||
		      hipMemcpy(b->in_xyz.p, xyz.data(), (size_t)b->n * 24, hipMemcpyHostToDevice) != hipSuccess ||
		      (b->poly && hipMemcpy(b->in_r.p, r.data(), (size_t)b->n * 8, hipMemcpyHostToDevice) != hipSuccess))) ||
	    (nw && hipMemcpy(b->walls.p, wall.data(), wall.size() * 8, hipMemcpyHostToDevice) != hipSuccess) ||
	    hipMemcpy(b->geoms.p, geo.data(), count * sizeof(RectGeom), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(b->boff.p, b->h_boff.data(), C1 * 4, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(b->pfirst.p, b->h_pfirst.data(), C1 * 4, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(b->d_wl.p, vgpu_worklist_table, 4096 * 4, hipMemcpyHostToDevice) != hipSuccess)
		return fail(VGPU_HIP_ERROR);
	hipLaunchKernelGGL(k_batch_pcont, dim3(count), dim3(256), 0, nullptr, count, b->pfirst.as<int>(), b->pcont.as<int>());
	if (launch_batch_mrad(nullptr, count, b->geoms.as<RectGeom>(), b->d_wl.as<unsigned>(), b->mrad.as<double>()))
		return fail(VGPU_HIP_ERROR);
	// buffers (their addresses go into the views)
	if ((st = batch_prepare(b, VGPU_F_NEIGHBORS))) return fail(st);
	std::vector<RectView> vw(count);
	for (int c = 0; c < count; c++) {
		vgpu_container *h = hs[c];
		RectView v = make_view(h);
		v.P = b->P.as<dbl4>(); v.pid = b->pid.as<int>();
		v.bstart = b->bstart.as<int>() + b->h_boff[c]; v.cstart = v.bstart;
		v.mrad = b->mrad.as<double>() + (size_t)c * 4096; v.wl = b->d_wl.as<unsigned>();
		v.walls = b->walls.as<double>() + (size_t)woff[c] * VGPU_WALL_STRIDE; v.nwalls = woff[c + 1] - woff[c];
		v.max_radius = 0;
		v.unit = nullptr; v.unit_p = 0; v.fv = nullptr;
		v.boff = b->h_boff[c];
		vw[c] = v;
	}
	if (hipMemcpy(b->views.p, vw.data(), count * sizeof(RectView), hipMemcpyHostToDevice) != hipSuccess ||
	    hipDeviceSynchronize() != hipSuccess)
		return fail(VGPU_HIP_ERROR);
	*out = b;
	return VGPU_OK;
}

int vgpu_batch_destroy(vgpu_batch *b) {
	if (!b) return VGPU_INVALID_ARGUMENT;
	b->drop_graph();
	if (b->gstream) (void)hipStreamDestroy(b->gstream);
	DBuf *bufs[] = {&b->views, &b->geoms, &b->boff, &b->pcont, &b->pfirst, &b->mrad, &b->walls, &b->d_wl, &b->maxr,
			&b->cfirst, &b->nchunk, &b->coff, &b->chunks, &b->nchunks, &b->stamp, &b->in_ids, &b->in_xyz,
			&b->in_r, &b->blk, &b->slot, &b->rem, &b->cnt, &b->bstart, &b->tmp, &b->part, &b->P, &b->pid,
			&b->src, &b->cellblk, &b->vol, &b->cent, &b->nfaces, &b->status, &b->nbr_start, &b->arena,
			&b->csr_off, &b->csr, &b->ctr, &b->counters, &b->retry1, &b->retry2, &b->retry3, &b->st_planes,
			&b->st_exit, &b->st_tier, &b->d_pred, &b->d_feat, &b->d_depth, &b->d_cost, &b->d_hist,
			&b->d_order, &b->lcnt, &b->loff, &b->list_a, &b->bcont, &b->ltop};
	for (DBuf *x : bufs) x->release();
	b->ws_fast.buf.release(); b->ws_wide.buf.release(); b->ws_huge.buf.release(); b->ws_wave.buf.release();
	if (b->events) for (int t = 0; t < 4; t++) (void)hipEventDestroy(b->ev[t]);
	delete b;
	return VGPU_OK;
}

int vgpu_batch_compute(vgpu_batch *b, unsigned flags, void *stream) {
	if (!b || (flags & VGPU_F_GEOMETRY) || ((flags & VGPU_F_LEARNED) && !b->have_pred)) return VGPU_INVALID_ARGUMENT;
	return batch_run(b, flags, (hipStream_t)stream);
}

int vgpu_batch_set_predictor(vgpu_batch *b, vgpu_predictor *p) {
	if (!b) return VGPU_INVALID_ARGUMENT;
	b->drop_graph();   // margin and route are launch arguments of a captured graph
	if (!p) { b->have_pred = false; return VGPU_OK; }
	if (!p->trained || p->params.size() != (size_t)vgpu::learn::NP || p->with_radius) return VGPU_INVALID_ARGUMENT;
	int st = b->d_pred.ensure(sizeof(float) * vgpu::learn::NP);
	if (st) return st;
	HIPCHK(hipMemcpy(b->d_pred.p, p->params.data(), sizeof(float) * vgpu::learn::NP, hipMemcpyHostToDevice));
	b->pred_margin = p->margin;
	b->pred_route = p->route;
	b->have_pred = true;
	return VGPU_OK;
}

int vgpu_batch_learned_stats(vgpu_batch *b, long long *fallbacks, long long *routed, long long *list_entries,
			     long long *list_need) {
	if (!b) return VGPU_INVALID_ARGUMENT;
	if (fallbacks) *fallbacks = b->last_fallbacks;
	if (routed) *routed = b->last_routed;
	if (list_entries) *list_entries = b->last_entries;
	if (list_need) *list_need = b->last_need;
	return VGPU_OK;
}

int vgpu_batch_list_capacity(vgpu_batch *b, long long *get, const long long *set) {
	if (!b || (set && (*set < 1 || *set > 2147483000LL))) return VGPU_INVALID_ARGUMENT;
	if (set && *set != b->lcap) {
		b->drop_graph();
		b->lcap = *set;
		b->list_a.release();
	}
	if (get) *get = b->lcap;
	return VGPU_OK;
}

int vgpu_batch_set_positions(vgpu_batch *b, const double *xyz, int on_device) {
	if (!b || (!xyz && b->n > 0)) return VGPU_INVALID_ARGUMENT;
	if (b->n == 0) return VGPU_OK;
	HIPCHK(hipMemcpy(b->in_xyz.p, xyz, (size_t)b->n * 24, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
	b->computed = false;
	return VGPU_OK;
}

int vgpu_batch_num_cells(vgpu_batch *b, int index, long long *n) {
	if (!b || !n || !b->computed || index < 0 || index >= b->ncont) return VGPU_INVALID_ARGUMENT;
	*n = b->h_cfirst[index + 1] - b->h_cfirst[index];
	return VGPU_OK;
}

int vgpu_batch_get_results(vgpu_batch *b, int index, int *id, int *ijk, int *q, int *exists, double *pos, double *vol,
			   double *centroid, int *nfaces, long long *nbr_off, int *nbr, long long nbr_cap,
			   long long *nbr_total) {
	if (!b || !b->computed || index < 0 || index >= b->ncont) return VGPU_INVALID_ARGUMENT;
	const int cf = b->h_cfirst[index], nc = b->h_cfirst[index + 1] - cf;
	if (nc == 0) {
		if (nbr_off) nbr_off[0] = 0;
		if (nbr_total) *nbr_total = 0;
		return VGPU_OK;
	}
	std::vector<int> st(nc), nf(nc);
	HIPCHK(hipMemcpy(st.data(), b->status.as<int>() + cf, (size_t)nc * 4, hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(nf.data(), b->nfaces.as<int>() + cf, (size_t)nc * 4, hipMemcpyDeviceToHost));
	if (id) HIPCHK(hipMemcpy(id, b->pid.as<int>() + cf, (size_t)nc * 4, hipMemcpyDeviceToHost));
	if (ijk || q) {
		const int b0 = b->h_boff[index], nbk = b->h_boff[index + 1] - b0;
		std::vector<int> cb(nc), bs(nbk + 1);
		HIPCHK(hipMemcpy(cb.data(), b->cellblk.as<int>() + cf, (size_t)nc * 4, hipMemcpyDeviceToHost));
		HIPCHK(hipMemcpy(bs.data(), b->bstart.as<int>() + b0, (size_t)(nbk + 1) * 4, hipMemcpyDeviceToHost));
		for (int c = 0; c < nc; c++) {
			if (ijk) ijk[c] = cb[c] - b0;
			if (q) q[c] = cf + c - bs[cb[c] - b0];
		}
	}
	if (exists) for (int c = 0; c < nc; c++) exists[c] = (st[c] & ST_EXISTS) ? 1 : 0;
	if (pos) {
		std::vector<dbl4> P(nc);
		HIPCHK(hipMemcpy(P.data(), b->P.as<dbl4>() + cf, (size_t)nc * 32, hipMemcpyDeviceToHost));
		for (int c = 0; c < nc; c++) { pos[3 * c] = P[c].x; pos[3 * c + 1] = P[c].y; pos[3 * c + 2] = P[c].z; }
	}
	if (vol) HIPCHK(hipMemcpy(vol, b->vol.as<double>() + cf, (size_t)nc * 8, hipMemcpyDeviceToHost));
	if (centroid) HIPCHK(hipMemcpy(centroid, b->cent.as<double>() + (size_t)3 * cf, (size_t)nc * 24, hipMemcpyDeviceToHost));
	if (nfaces) for (int c = 0; c < nc; c++) nfaces[c] = (st[c] & ST_EXISTS) ? nf[c] : 0;
	long long tot = 0;
	std::vector<long long> off(nc + 1);
	for (int c = 0; c < nc; c++) { off[c] = tot; if (st[c] & ST_EXISTS) tot += nf[c]; }
	off[nc] = tot;
	if (nbr_off) std::memcpy(nbr_off, off.data(), (size_t)(nc + 1) * 8);
	if (nbr_total) *nbr_total = tot;
	if (nbr && (b->last_flags & VGPU_F_NEIGHBORS) && tot > 0) {
		int g0 = 0;
		HIPCHK(hipMemcpy(&g0, b->tmp.as<int>() + cf, 4, hipMemcpyDeviceToHost));
		const long long m = std::min(tot, nbr_cap);
		if (m > 0) HIPCHK(hipMemcpy(nbr, b->csr.as<int>() + g0, (size_t)m * 4, hipMemcpyDeviceToHost));
	}
	return VGPU_OK;
}

int vgpu_batch_sum_volumes(vgpu_batch *b, double *out) {
	if (!b || !out || !b->computed) return VGPU_INVALID_ARGUMENT;
	const int n = b->h_cfirst[b->ncont];
	std::vector<double> v(std::max(n, 1));
	std::vector<int> st(std::max(n, 1));
	if (n) {
		HIPCHK(hipMemcpy(v.data(), b->vol.p, (size_t)n * 8, hipMemcpyDeviceToHost));
		HIPCHK(hipMemcpy(st.data(), b->status.p, (size_t)n * 4, hipMemcpyDeviceToHost));
	}
	for (int c = 0; c < b->ncont; c++) {
		double s = 0;
		for (int t = b->h_cfirst[c]; t < b->h_cfirst[c + 1]; t++) if (st[t] & ST_EXISTS) s += v[t];
		out[c] = s;
	}
	return VGPU_OK;
}

int vgpu_batch_config(vgpu_batch *b, int *get3, const int *set3) {
	if (!b) return VGPU_INVALID_ARGUMENT;
	if (get3) { get3[0] = b->waves; get3[1] = b->occ; get3[2] = b->chunk; }
	if (set3) {
		if (set3[0] < 0 || set3[1] == 1 || set3[1] < 0 || set3[1] > 6 || set3[2] < 1 || set3[2] > 64)
			return VGPU_INVALID_ARGUMENT;
		b->waves_req = set3[0]; b->occ = set3[1]; b->chunk = set3[2];
		b->drop_graph();
	}
	return VGPU_OK;
}

// The project's autotuner for batches: times the compute (eager, `reps`
// runs each, best kept) for the per-lane first tier at every register budget
// (2, 3, 4 waves per SIMD) with 64 or 32 cells per queue entry and for the
// wavefront-cell first tier at 4, 5 and 6 waves per SIMD (16 cells per
// entry), each with as many persistent wavefronts as fit, and keeps the
// fastest.  Results do not depend
// on the choice.  (On FCC-like containers, whose cells are alike, the
// per-lane tier wins; on uneven cells the wavefront cells.)
int vgpu_batch_autotune(vgpu_batch *b, unsigned flags, int reps, int *best3, double *best_ms) {
	if (!b || reps < 1 || (flags & (VGPU_F_LEARNED | VGPU_F_GRAPH))) return VGPU_INVALID_ARGUMENT;
	const int cands[9][2] = {{2, 64}, {2, 32}, {3, 64}, {3, 32}, {4, 64}, {4, 32}, {0, 16}, {5, 16}, {6, 16}};
	double best = 1e300;
	int bo = 2, bc = 64;
	for (int oi = 0; oi < 9; oi++) {
		const int cfg[3] = {0, cands[oi][0], cands[oi][1]};
		int st = vgpu_batch_config(b, nullptr, cfg);
		if (st) return st;
		double t = 1e300;
		for (int r = 0; r < reps; r++) {
			if ((st = batch_run(b, flags, nullptr))) return st;
			t = std::min(t, b->timings[0] + b->timings[1] + b->timings[2]);
		}
		if (t < best) { best = t; bo = cands[oi][0]; bc = cands[oi][1]; }
	}
	const int cfg[3] = {0, bo, bc};
	int st = vgpu_batch_config(b, nullptr, cfg);
	if (st) return st;
	if (best3) { best3[0] = default_fast_waves(bo == 0 ? 4 : bo); best3[1] = bo; best3[2] = bc; }   // 5, 6: wavefront cells
	if (best_ms) *best_ms = best;
	return VGPU_OK;
}

int vgpu_batch_last_timings(vgpu_batch *b, double *ms3) {
	if (!b || !ms3) return VGPU_INVALID_ARGUMENT;
	for (int t = 0; t < 3; t++) ms3[t] = b->timings[t];
	return VGPU_OK;
}

const char *vgpu_status_string(int s) {
	switch (s) {
	case VGPU_OK: return "ok";
	case VGPU_FILE_ERROR: return "file error";
	case VGPU_MEMORY_ERROR: return "capacity exceeded";
	case VGPU_INTERNAL_ERROR: return "internal error";
	case VGPU_CMD_LINE_ERROR: return "command line error";
	case VGPU_INVALID_ARGUMENT: return "invalid argument";
	case VGPU_HIP_ERROR: return "HIP runtime error";
	case VGPU_DUPLICATE: return "duplicate particle";
	case VGPU_NO_DEVICE: return "no GPU device";
	default: return "unknown status";
	}
}

}  // extern "C"
