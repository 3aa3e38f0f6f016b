// CPU build of the engine's device code, for tests only.
//
// Compiles voro_amd/csrc/compute_dev.h with interleave 1 and drives it from a
// plain host loop (stable counting-sort binning, one cell at a time, tier
// escalation on overflow), so that the kernel logic can be checked against
// the oracle on hosts without a GPU.  It is not a product path: the C-ABI
// library never falls back to it.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../voro_amd/csrc/compute_dev.h"
#include "../../voro_amd/csrc/grid_host.h"
#include "../../voro_amd/csrc/list_dev.h"
#include "../../voro_amd/csrc/prd_host.h"
#include "../../voro_amd/csrc/wcell_dev.h"

using namespace vgpu;

extern "C" {
struct hx_spec {
	int kind; double geom[6]; int grid[3]; int periodic[3]; int init_mem; int n_spheres; const double *spheres;
	int n_walls; const double *walls;   // {type, w_id, constructor arguments}
};
struct hx_out {
	int *id; int *ijk; int *q; int *exists; double *vol; double *cent; int *nfaces;
	long long *nbr_off; int *nbr; long long nbr_cap; double *pos;
};
}

namespace {
struct VecSink {
	std::vector<int> *v;
	void put(int i, int x) { (void)i; v->push_back(x); }
};

template<class CELL, size_t BYTES>
struct Runner {
	std::vector<char> slab;
	CELL c;
	unsigned stamp = 0;
	explicit Runner(double mls) : slab(BYTES + 2 * 65536, 0) { c.setup(slab.data() + 65536, 0, mls); }
	template<bool POLY, bool PRD = false>
	int run(const RectView &v, int ijk, int s, int ci, int cj, int ck, double *vol, double *cent, int *nf,
		std::vector<int> &nb, CellStats &st, unsigned long long key0 = 0) {
		bool ok = compute_cell_rect<CELL, POLY, PRD>(c, v, ijk, s, ci, cj, ck, ++stamp, st, key0);
		if (c.ovf) return c.fatal ? -2 : -1;
		if (ok) {
			VecSink sk{&nb};
			c.outputs(*vol, cent[0], cent[1], cent[2], *nf, sk);
		}
		return ok ? 1 : 0;
	}
};

// 1: a wavefront cell whose list ends is continued from the list's last
// step by compute_cell_rect (as the device's list tier does); 0: -1 sends
// it on to the plain tiers, which compute it again.
static int list_resume = 1;

// The list-driven tier: gather the cell's candidate list at depth G, run
// compute_cell_list; -1 sends the cell on to the plain tiers (the order-3
// cell) or continues it (the wavefront cell, list_resume).
// PRD: a periodic container; the gather reads a copy of the view without
// the first-visit record, the cell records its visits with key0.
template<bool POLY, class CELL = VCell3, size_t BYTES = Layout3<>::lane_bytes, bool PRD = false>
struct ListRunner {
	std::vector<char> slab;
	CELL c;
	std::vector<ListRec> la;
	unsigned stamp = 0;
	explicit ListRunner(double mls) : slab(BYTES + 2 * 65536, 0) { c.setup(slab.data() + 65536, 0, mls); }
	int run(const RectView &v, const int *cellblk, int cell, int G, int ijk, int s, double *vol, double *cent, int *nf,
		std::vector<int> &nb, CellStats &st, long long *fallbacks, unsigned long long key0 = 0) {
		RectView vg = v;
		vg.fv = nullptr;
		la.resize(64);
		const int n = gather_cell_list<PRD>(vg, cellblk[cell], cell, G, la.data(), 1);
		ListResume rs;
		int r = compute_cell_list<CELL, POLY, PRD>(c, v, ijk, s, la.data(), 1, n, st, key0, &rs);
		if (r == -1) {
			(*fallbacks)++;
			if (!std::is_same<CELL, WCell>::value || !list_resume) return -1;
			list_resume_fill<POLY, PRD>(v, ijk, s, la.data(), 1, n, key0, rs);
			const int ci = ijk % v.nx, cj = (ijk / v.nx) % v.gy, ck = ijk / v.nxy;
			const bool ok = compute_cell_rect<CELL, POLY, PRD>(c, v, ijk, s, ci, cj, ck, ++stamp, st, key0, nullptr, 0,
									   &rs);
			if (c.ovf) return -1;
			r = ok ? 1 : 0;
		}
		if (r < 0 || c.ovf) return -1;
		if (r == 1) {
			VecSink sk{&nb};
			c.outputs(*vol, cent[0], cent[1], cent[2], *nf, sk);
			if (c.ovf) { nb.clear(); return -1; }
		}
		return r;
	}
};

static inline int step_int(double a) { return a < 0 ? int(a) - 1 : int(a); }
static inline int step_mod(int a, int b) { return a >= 0 ? a % b : b - 1 - (b - 1 - a) % b; }
}  // namespace

extern "C" {

static int list_depth = 0;
static long long list_fallbacks = 0;
// G > 0: first try every cell through a candidate list of worklist depth G.
void hx_use_list(int G) { list_depth = G; list_fallbacks = 0; }
// on: wavefront cells continue where their list ended (default); off: they
// are computed again by the plain tiers.
void hx_list_resume(int on) { list_resume = on; }
long long hx_list_fallbacks(void) { return list_fallbacks; }

static int use_wave = 0;
static long long wave_done = 0, wave2_done = 0, wave4_done = 0;
// on: every cell is first tried by the wavefront-cell tiers (wcell_dev.h,
// 64 lanes emulated, up to 64 then 128 vertices); cells they cannot finish
// go on to the per-lane tiers.
void hx_use_wave(int on) { use_wave = on; wave_done = 0; wave2_done = 0; wave4_done = 0; }
long long hx_wave_done(void) { return wave_done; }
// cells the 128-vertex wavefront cell finished after the 64-vertex one overflowed
long long hx_wave2_done(void) { return wave2_done; }
// ... and the 254-vertex one after both overflowed
long long hx_wave4_done(void) { return wave4_done; }

static long long prd_passes = 0;
static int prd_replay = 1;
// Passes over the cells of the last periodic computation (1 + replays).
long long hx_image_passes(void) { return prd_passes; }
// 0: keep the first pass (image blocks in create_all_images order), to show
// what the replay of the first-visit order changes.
void hx_set_replay(int on) { prd_replay = on; }

static long long compute_prd(const hx_spec *s, int n, const int *ids, const double *xyz, const double *r,
			     hx_out *o, int *stats_out, long long *tier_counts);

// tier_counts[3]: cells finished by each tier.  Returns cells, or -1 on a
// cell no tier could finish.

long long hx_compute_neighbor(const hx_spec *s, int n, const int *ids, const double *xyz, const double *r,
			      hx_out *o, int *stats_out /* 4 ints per cell or null */, long long *tier_counts) {
	if (s->kind > 1) return compute_prd(s, n, ids, xyz, r, o, stats_out, tier_counts);
	const bool poly = s->kind == 1;
	RectView v = RectView();
	v.ax = s->geom[0]; v.bx = s->geom[1]; v.ay = s->geom[2]; v.by = s->geom[3]; v.az = s->geom[4]; v.bz = s->geom[5];
	v.nx = s->grid[0]; v.ny = s->grid[1]; v.nz = s->grid[2]; v.nxy = v.nx * v.ny;
	v.xper = s->periodic[0]; v.yper = s->periodic[1]; v.zper = s->periodic[2];
	v.boxx = (v.bx - v.ax) / v.nx; v.boxy = (v.by - v.ay) / v.ny; v.boxz = (v.bz - v.az) / v.nz;
	v.xsp = 1 / v.boxx; v.ysp = 1 / v.boxy; v.zsp = 1 / v.boxz;
	v.hx = v.xper ? 2 * v.nx + 1 : v.nx; v.hy = v.yper ? 2 * v.ny + 1 : v.ny; v.hz = v.zper ? 2 * v.nz + 1 : v.nz;
	v.bxsq = v.boxx * v.boxx + v.boxy * v.boxy + v.boxz * v.boxz;
	v.max_len_sq = (v.bx - v.ax) * (v.bx - v.ax) * (v.xper ? 0.25 : 1) + (v.by - v.ay) * (v.by - v.ay) * (v.yper ? 0.25 : 1)
		       + (v.bz - v.az) * (v.bz - v.az) * (v.zper ? 0.25 : 1);
	const int nxyz = v.nxy * v.nz;
	// binning: block index as put_remap computes it; stable by input order
	std::vector<int> blk(n, -1), cnt(nxyz + 1, 0);
	std::vector<dbl4> rem(n);
	double maxr = 0;
	for (int t = 0; t < n; t++) {
		double x = xyz[3 * t], y = xyz[3 * t + 1], z = xyz[3 * t + 2];
		int i = step_int((x - v.ax) * v.xsp), j, k, l;
		if (v.xper) { l = step_mod(i, v.nx); x += v.boxx * (l - i); i = l; } else if (i < 0 || i >= v.nx) continue;
		j = step_int((y - v.ay) * v.ysp);
		if (v.yper) { l = step_mod(j, v.ny); y += v.boxy * (l - j); j = l; } else if (j < 0 || j >= v.ny) continue;
		k = step_int((z - v.az) * v.zsp);
		if (v.zper) { l = step_mod(k, v.nz); z += v.boxz * (l - k); k = l; } else if (k < 0 || k >= v.nz) continue;
		blk[t] = i + v.nx * j + v.nxy * k;
		double rr = r ? r[t] : 0.5;
		rem[t] = dbl4{x, y, z, rr};
		if (poly && maxr < rr) maxr = rr;
		cnt[blk[t] + 1]++;
	}
	for (int b = 0; b < nxyz; b++) cnt[b + 1] += cnt[b];
	std::vector<int> fill(cnt.begin(), cnt.end() - 1);
	int m = cnt[nxyz];
	std::vector<dbl4> P(m);
	std::vector<int> pid(m);
	for (int t = 0; t < n; t++) if (blk[t] >= 0) { int d = fill[blk[t]]++; P[d] = rem[t]; pid[d] = ids[t]; }
	std::vector<double> mrad(4096), walls;
	worklist_radii(v.boxx, v.boxy, v.boxz, mrad.data());
	// device wall records (type, w_id, the constants the wall constructors derive)
	auto add = [&](int type, int id, const double *a) {
		double r[VGPU_WALL_STRIDE] = {0};
		r[0] = type; r[1] = id; r[2] = a[0]; r[3] = a[1]; r[4] = a[2];
		if (type <= 1) r[5] = a[3];
		else {
			r[5] = a[3]; r[6] = a[4]; r[7] = a[5];
			r[8] = 1 / (a[3] * a[3] + a[4] * a[4] + a[5] * a[5]);
			if (type == 2) r[9] = a[6];
			else { r[9] = std::sin(a[6]); r[10] = std::cos(a[6]); r[11] = std::tan(a[6]); }
		}
		walls.insert(walls.end(), r, r + VGPU_WALL_STRIDE);
	};
	for (int w = 0; w < s->n_spheres; w++) add(0, (int)s->spheres[5 * w + 4], s->spheres + 5 * w);
	for (int w = 0; w < s->n_walls; w++) add((int)s->walls[9 * w], (int)s->walls[9 * w + 1], s->walls + 9 * w + 2);
	v.gy = v.ny;
	v.P = P.data(); v.pid = pid.data(); v.bstart = cnt.data(); v.cstart = cnt.data(); v.mrad = mrad.data(); v.wl = vgpu_worklist_table;
	v.walls = walls.data(); v.nwalls = (int)(walls.size() / VGPU_WALL_STRIDE); v.max_radius = maxr;

	std::vector<int> cellblk(m);
	for (int b = 0; b < nxyz; b++) for (int q = cnt[b]; q < cnt[b + 1]; q++) cellblk[q] = b;
	ListRunner<false> lr0(v.max_len_sq);
	ListRunner<true> lr1(v.max_len_sq);
	ListRunner<false, WCell, WLayout::bytes> lw0(v.max_len_sq);
	ListRunner<true, WCell, WLayout::bytes> lw1(v.max_len_sq);

	Runner<WCell, WLayout::bytes> rw(v.max_len_sq);
	Runner<WCell2, WLayoutT<2>::bytes> rw2(v.max_len_sq);
	Runner<WCell4, WLayoutT<4>::bytes> rw4(v.max_len_sq);
	Runner<VCell3, Layout3<>::lane_bytes> r0(v.max_len_sq);
	Runner<VCell<TierWide, 1>, Layout<TierWide, 1>::bytes> r1(v.max_len_sq);
	Runner<VCell<TierHuge, 1>, Layout<TierHuge, 1>::bytes> *r2 = 0;
	std::vector<int> nb;
	long long nn = 0;
	int cell = 0;
	bool bad = false;
	for (int ijk = 0; ijk < nxyz; ijk++) {
		int ci = ijk % v.nx, cj = (ijk / v.nx) % v.ny, ck = ijk / v.nxy;
		for (int q = cnt[ijk]; q < cnt[ijk + 1]; q++, cell++) {
			double vol = 0, cent[3] = {0, 0, 0};
			int nf = 0, res;
			CellStats st;
			nb.clear();
			int tier = 0;
			res = -1;
			if (list_depth > 0) {
				if (use_wave)
					res = poly ? lw1.run(v, cellblk.data(), cell, list_depth, ijk, q - cnt[ijk], &vol, cent, &nf, nb, st, &list_fallbacks)
						   : lw0.run(v, cellblk.data(), cell, list_depth, ijk, q - cnt[ijk], &vol, cent, &nf, nb, st, &list_fallbacks);
				else
					res = poly ? lr1.run(v, cellblk.data(), cell, list_depth, ijk, q - cnt[ijk], &vol, cent, &nf, nb, st, &list_fallbacks)
						   : lr0.run(v, cellblk.data(), cell, list_depth, ijk, q - cnt[ijk], &vol, cent, &nf, nb, st, &list_fallbacks);
				if (res < 0) nb.clear();
			}
			if (res < 0 && use_wave) {
				res = poly ? rw.run<true>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st)
					   : rw.run<false>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st);
				if (res >= 0 && nf > (int)WLayout::NBCAP) res = -1;
				if (res < 0) {
					nb.clear();
					res = poly ? rw2.run<true>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st)
						   : rw2.run<false>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st);
					if (res >= 0 && nf > (int)WLayoutT<2>::NBCAP) res = -1;
					if (res < 0) {
						nb.clear();
						res = poly ? rw4.run<true>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st)
							   : rw4.run<false>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st);
						if (res >= 0 && nf > (int)WLayoutT<4>::NBCAP) res = -1;
						if (res < 0) nb.clear(); else wave4_done++;
					} else wave2_done++;
				} else wave_done++;
			}
			if (res < 0) res = poly ? r0.run<true>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st)
				   : r0.run<false>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st);
			if (res < 0) {
				tier = 1; nb.clear();
				res = poly ? r1.run<true>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st)
					   : r1.run<false>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st);
			}
			if (res < 0) {
				tier = 2; nb.clear();
				if (!r2) r2 = new Runner<VCell<TierHuge, 1>, Layout<TierHuge, 1>::bytes>(v.max_len_sq);
				res = poly ? r2->run<true>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st)
					   : r2->run<false>(v, ijk, q - cnt[ijk], ci, cj, ck, &vol, cent, &nf, nb, st);
			}
			if (res < 0) { bad = true; res = 0; }
			if (tier_counts) tier_counts[tier]++;
			if (stats_out) { stats_out[4 * cell] = st.n_planes; stats_out[4 * cell + 1] = st.wl_exit; stats_out[4 * cell + 2] = st.phase; stats_out[4 * cell + 3] = tier; }
			o->id[cell] = pid[q]; o->ijk[cell] = ijk; o->q[cell] = q - cnt[ijk]; o->exists[cell] = res;
			if (o->pos) { o->pos[3 * cell] = P[q].x; o->pos[3 * cell + 1] = P[q].y; o->pos[3 * cell + 2] = P[q].z; }
			o->vol[cell] = vol; o->cent[3 * cell] = cent[0]; o->cent[3 * cell + 1] = cent[1]; o->cent[3 * cell + 2] = cent[2];
			o->nfaces[cell] = res ? nf : 0;
			o->nbr_off[cell] = nn;
			if (res) for (size_t t = 0; t < nb.size(); t++) { if (nn < o->nbr_cap) o->nbr[nn] = nb[t]; nn++; }
		}
	}
	o->nbr_off[cell] = nn;
	delete r2;
	return bad ? -1 : cell;
}

}

// Periodic containers: the engine's algorithm (binning as put_locate_block,
// image blocks first in create_all_images order, then replayed in the
// recorded first-visit order until that order reproduces itself), one cell
// at a time.  Returns cells, -1 on a cell no tier could finish, -3 on a
// duplicate particle, -4 when the unit cell fails, -5 when the first-visit
// order does not settle.
static long long compute_prd(const hx_spec *s, int n, const int *ids, const double *xyz, const double *r,
			     hx_out *o, int *stats_out, long long *tier_counts) {
	const bool poly = s->kind == 3;
	PrdGeom g;
	if (!g.build(s->geom[0], s->geom[1], s->geom[2], s->geom[3], s->geom[4], s->geom[5], s->grid[0], s->grid[1],
		     s->grid[2]))
		return -4;
	const int nb = (int)g.nblocks();
	// binning (src/container_prd.cc:189-217), stable by input order
	std::vector<int> blk(n, -1), cnt(nb + 1, 0);
	std::vector<dbl4> rem(n);
	double maxr = 0;
	for (int t = 0; t < n; t++) {
		double x = xyz[3 * t], y = xyz[3 * t + 1], z = xyz[3 * t + 2];
		int k = step_int(z * g.zsp);
		if (k < 0 || k >= g.nz) { int ak = prd_step_div(k, g.nz); z -= ak * g.bz; y -= ak * g.byz; x -= ak * g.bxz; k -= ak * g.nz; }
		int j = step_int(y * g.ysp);
		if (j < 0 || j >= g.ny) { int aj = prd_step_div(j, g.ny); y -= aj * g.by; x -= aj * g.bxy; j -= aj * g.ny; }
		int i = step_int(x * g.xsp);
		if (i < 0 || i >= g.nx) { int ai = prd_step_div(i, g.nx); x -= ai * g.bx; i -= ai * g.nx; }
		j += g.ey; k += g.ez;
		blk[t] = i + g.nx * (j + g.oy * k);
		double rr = r ? r[t] : 0.5;
		rem[t] = dbl4{x, y, z, rr};
		if (poly && maxr < rr) maxr = rr;
		cnt[blk[t] + 1]++;
	}
	for (int b = 0; b < nb; b++) cnt[b + 1] += cnt[b];
	std::vector<int> fill(cnt.begin(), cnt.end() - 1);
	const int m = cnt[nb];
	std::vector<dbl4> P(std::max(m, 1));
	std::vector<int> pid(std::max(m, 1)), cellblk(std::max(m, 1));
	for (int t = 0; t < n; t++) {
		int d = fill[blk[t]]++;
		P[d] = rem[t]; pid[d] = ids[t]; cellblk[d] = blk[t];
	}
	for (int d = 0; d < m; d++)
		for (int u = cnt[cellblk[d]]; u < d; u++) {
			double dx = P[u].x - P[d].x, dy = P[u].y - P[d].y, dz = P[u].z - P[d].z;
			if (dx * dx + dy * dy + dz * dz < 1e-10) return -3;
		}
	std::vector<double> mrad(4096);
	worklist_radii(g.boxx, g.boxy, g.boxz, mrad.data());
	RectView v = RectView();
	v.ax = v.ay = v.az = 0; v.bx = g.bx; v.by = g.by; v.bz = g.bz;
	v.nx = g.nx; v.ny = g.ny; v.nz = g.nz; v.gy = g.oy; v.nxy = g.nx * g.oy;
	v.boxx = g.boxx; v.boxy = g.boxy; v.boxz = g.boxz; v.xsp = g.xsp; v.ysp = g.ysp; v.zsp = g.zsp;
	v.hx = 2 * g.nx + 1; v.hy = 2 * g.ey + 1; v.hz = 2 * g.ez + 1;
	v.bxsq = g.boxx * g.boxx + g.boxy * g.boxy + g.boxz * g.boxz;
	v.max_len_sq = g.max_len_sq; v.max_radius = maxr;
	v.cstart = cnt.data(); v.mrad = mrad.data(); v.wl = vgpu_worklist_table; v.nwalls = 0;
	v.bxy = g.bxy; v.bxz = g.bxz; v.byz = g.byz;
	v.ey = g.ey; v.ez = g.ez; v.wy = g.wy; v.wz = g.wz; v.oy = g.oy; v.oz = g.oz;
	v.unit = g.unit.data(); v.unit_p = (int)g.unit.size();
	std::vector<unsigned long long> fv(nb, ~0ull), prev;
	v.fv = fv.data();

	ListRunner<false, VCell3, Layout3<>::lane_bytes, true> lr0(v.max_len_sq);
	ListRunner<true, VCell3, Layout3<>::lane_bytes, true> lr1(v.max_len_sq);
	ListRunner<false, WCell, WLayout::bytes, true> lw0(v.max_len_sq);
	ListRunner<true, WCell, WLayout::bytes, true> lw1(v.max_len_sq);
	Runner<WCell, WLayout::bytes> rw(v.max_len_sq);
	Runner<WCell2, WLayoutT<2>::bytes> rw2(v.max_len_sq);
	Runner<WCell4, WLayoutT<4>::bytes> rw4(v.max_len_sq);
	Runner<VCell3, Layout3<>::lane_bytes> r0(v.max_len_sq);
	Runner<VCell<TierWide, 1>, Layout<TierWide, 1>::bytes> r1(v.max_len_sq);
	Runner<VCell<TierHuge, 1>, Layout<TierHuge, 1>::bytes> *r2 = 0;
	struct Res { double vol, cent[3]; int nf, exists, tier; std::vector<int> nb; CellStats st; };
	std::vector<Res> res(m);
	std::vector<unsigned char> touched(m, 0);
	bool bad = false;
	// list: the first pass tries the candidate list first (as the device's
	// learned tier does); the replays take the plain tiers
	auto one = [&](int cell, bool list) {
		const int ijk = cellblk[cell], q = cell - cnt[ijk];
		const int ci = ijk % g.nx, cj = (ijk / g.nx) % g.oy, ck = ijk / (g.nx * g.oy);
		Res &R = res[cell];
		R.vol = 0; R.cent[0] = R.cent[1] = R.cent[2] = 0; R.nf = 0; R.nb.clear(); R.tier = 0;
		const unsigned long long key = (unsigned long long)cell << 24;
		int rr = -1;
		if (list && list_depth > 0) {
			const int *cb = cellblk.data();
			if (use_wave)
				rr = poly ? lw1.run(v, cb, cell, list_depth, ijk, q, &R.vol, R.cent, &R.nf, R.nb, R.st, &list_fallbacks, key)
					  : lw0.run(v, cb, cell, list_depth, ijk, q, &R.vol, R.cent, &R.nf, R.nb, R.st, &list_fallbacks, key);
			else
				rr = poly ? lr1.run(v, cb, cell, list_depth, ijk, q, &R.vol, R.cent, &R.nf, R.nb, R.st, &list_fallbacks, key)
					  : lr0.run(v, cb, cell, list_depth, ijk, q, &R.vol, R.cent, &R.nf, R.nb, R.st, &list_fallbacks, key);
			if (rr < 0) R.nb.clear();
		}
		if (rr < 0 && use_wave) {
			rr = poly ? rw.run<true, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key)
				  : rw.run<false, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key);
			if (rr >= 0 && R.nf > (int)WLayout::NBCAP) rr = -1;
			if (rr < 0) {
				R.nb.clear();
				rr = poly ? rw2.run<true, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key)
					  : rw2.run<false, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key);
				if (rr >= 0 && R.nf > (int)WLayoutT<2>::NBCAP) rr = -1;
				if (rr < 0) {
					R.nb.clear();
					rr = poly ? rw4.run<true, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key)
						  : rw4.run<false, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key);
					if (rr >= 0 && R.nf > (int)WLayoutT<4>::NBCAP) rr = -1;
					if (rr < 0) R.nb.clear(); else wave4_done++;
				} else wave2_done++;
			} else wave_done++;
		}
		if (rr < 0)
			rr = poly ? r0.run<true, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key)
				  : r0.run<false, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key);
		if (rr < 0) {
			R.tier = 1; R.nb.clear();
			rr = poly ? r1.run<true, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key)
				  : r1.run<false, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key);
		}
		if (rr < 0) {
			R.tier = 2; R.nb.clear();
			if (!r2) r2 = new Runner<VCell<TierHuge, 1>, Layout<TierHuge, 1>::bytes>(v.max_len_sq);
			rr = poly ? r2->run<true, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key)
				  : r2->run<false, true>(v, ijk, q, ci, cj, ck, &R.vol, R.cent, &R.nf, R.nb, R.st, key);
		}
		if (rr < 0) { bad = true; rr = 0; }
		R.exists = rr;
		if (R.st.img) touched[cell] = 1;
	};
	std::vector<dbl4> Pa;
	std::vector<int> ia, ba;
	{
		ImageSet im(g, P.data(), pid.data(), cnt.data());
		for (int b : prd_margin_images(g, PRD_FIRST_MARGIN)) im.create_block(b);
		im.assemble(Pa, ia, ba);
		v.P = Pa.data(); v.pid = ia.data(); v.bstart = ba.data();
	}
	for (int c = 0; c < m; c++) one(c, true);
	prd_passes = 1;
	bool settled = false;
	if (!prd_replay) settled = true;
	for (int pass = 0; pass < 8 && !settled; pass++) {
		prev = fv;
		ImageSet im(g, P.data(), pid.data(), cnt.data());
		for (int b : prd_visit_order(g, prev.data())) im.create_block(b);
		if (im.late) { delete r2; return -5; }
		im.assemble(Pa, ia, ba);
		v.P = Pa.data(); v.pid = ia.data(); v.bstart = ba.data();
		std::fill(fv.begin(), fv.end(), ~0ull);
		for (int c = 0; c < m; c++) if (touched[c]) one(c, false);
		prd_passes = pass + 2;
		settled = fv == prev;
	}
	delete r2;
	if (!settled) return -5;
	long long nn = 0;
	for (int c = 0; c < m; c++) {
		const Res &R = res[c];
		const int ijk = cellblk[c];
		if (tier_counts) tier_counts[R.tier]++;
		if (stats_out) { stats_out[4 * c] = R.st.n_planes; stats_out[4 * c + 1] = R.st.wl_exit; stats_out[4 * c + 2] = R.st.phase; stats_out[4 * c + 3] = R.tier; }
		o->id[c] = pid[c]; o->ijk[c] = ijk; o->q[c] = c - cnt[ijk]; o->exists[c] = R.exists;
		if (o->pos) { o->pos[3 * c] = P[c].x; o->pos[3 * c + 1] = P[c].y; o->pos[3 * c + 2] = P[c].z; }
		o->vol[c] = R.vol; o->cent[3 * c] = R.cent[0]; o->cent[3 * c + 1] = R.cent[1]; o->cent[3 * c + 2] = R.cent[2];
		o->nfaces[c] = R.exists ? R.nf : 0;
		o->nbr_off[c] = nn;
		if (R.exists) for (size_t t = 0; t < R.nb.size(); t++) { if (nn < o->nbr_cap) o->nbr[nn] = R.nb[t]; nn++; }
	}
	o->nbr_off[m] = nn;
	return bad ? -1 : m;
}


// ---- Polyhedron statistics (geom_dev.h) from given cell tables ----
#include "../../voro_amd/csrc/geom_dev.h"

extern "C" long long hx_geometry_stats(int ncells, const int *exists, const long long *vert_off, const double *pts,
				       const int *orders, const long long *edge_off, const int *edges,
				       const int *edge_rel, const long long *face_off, const long long *fv_off,
				       double tol, double *vertices, double *face_area, double *face_perim,
				       double *normal, int *face_order, int *face_verts, double *surface,
				       double *edge_dist, int *n_edges, double *maxr2) {
	const long long nv = vert_off[ncells], ne = edge_off[ncells];
	std::vector<int> vstart(ncells), np(ncells), eoff(nv > 0 ? nv : 1), status(ncells), fo(ncells), vo(ncells),
		fvo(ncells), eo(ncells);
	std::vector<unsigned char> nu(nv > 0 ? nv : 1), rel(ne > 0 ? ne : 1);
	std::vector<short> ed(ne > 0 ? ne : 1);
	std::vector<double> xyz(3 * (nv > 0 ? nv : 1));
	for (int c = 0; c < ncells; c++) {
		vstart[c] = (int)vert_off[c];
		np[c] = (int)(vert_off[c + 1] - vert_off[c]);
		status[c] = exists[c] ? ST_EXISTS : 0;
		fo[c] = (int)face_off[c]; vo[c] = (int)vert_off[c]; fvo[c] = (int)fv_off[c]; eo[c] = (int)edge_off[c];
		long long e = edge_off[c];
		for (long long v = vert_off[c]; v < vert_off[c + 1]; v++) { eoff[v] = (int)e; e += orders[v]; }
	}
	for (long long v = 0; v < nv; v++) {
		nu[v] = (unsigned char)orders[v];
		for (int a = 0; a < 3; a++) xyz[3 * v + a] = pts[3 * v + a];
	}
	for (long long t = 0; t < ne; t++) { ed[t] = (short)edges[t]; rel[t] = (unsigned char)edge_rel[t]; }
	GeoSink g;
	g.vstart = vstart.data(); g.np = np.data(); g.xyz = xyz.data(); g.nu = nu.data(); g.eoff = eoff.data();
	g.ed = ed.data(); g.rel = rel.data();
	GeoStats o;
	o.status = status.data(); o.face_off = fo.data(); o.vert_off = vo.data(); o.edge_off = eo.data();
	o.fv_off = fvo.data(); o.tol = tol;
	o.vertices = vertices; o.face_area = face_area; o.face_perim = face_perim; o.normal = normal;
	o.face_order = face_order; o.face_verts = face_verts; o.surface = surface; o.edge_dist = edge_dist;
	o.n_edges = n_edges; o.maxr2 = maxr2;
	for (int c = 0; c < ncells; c++) cell_geometry_stats(g, o, c);
	// the tables must come back unchanged (edges restored after marking)
	for (long long t = 0; t < ne; t++) if (ed[t] != edges[t]) return -1;
	return ncells;
}
