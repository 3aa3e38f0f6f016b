// Per-particle cell computation over the binned block grid: start cell,
// walls, own block, worklist phases 1 and 2 with the security-radius stop,
// and the phase-3 breadth-first block search.
//
// Restates src/v_compute.cc:291-621 (compute_cell), 634-780 (region tests),
// 797-944 (compute_min_max_radius), 236-264 (queue/mask helpers), together
// with the container adapters src/container.hh:192-251 (start cell,
// frac_pos, region_index), src/rad_option.hh:24-158 (radical policy) and
// src/wall.cc:28-103 (sphere, plane, cylinder and cone walls).  The candidate order, the count_list
// schedule on which the maximum vertex radius is refreshed and every
// floating-point expression are kept as in the reference, so the sequence
// of real cuts, and hence the cell, is identical.
//
// Differences of representation: the container-wide block mask with its
// visit stamp becomes a per-lane window of (2W+1)^3 bits around the
// particle's block whose words carry their own stamp (no clearing between
// cells), and the growable circular block queue becomes a fixed ring; a cell
// that leaves the window or fills the ring is recomputed by a wider tier.
#pragma once
#include "cell_dev.h"
#include "cell3_dev.h"
#include "wcell_dev.h"

namespace vgpu {

struct dbl4 { double x, y, z, w; };

// Worklist table (generated; identical to the reference's table).
#include "worklist_table.h"

// Read-only view of a binned container.
//
// Rectangular containers: blocks nx*ny*nz.  Periodic containers
// (container_periodic_base, src/container_prd.hh:83-211): blocks nx*oy*oz,
// the primary domain at j in [ey,wy), k in [ez,wz), image blocks around it
// in y and z; ax=ay=az=0, (bx,by,bz) the diagonal of the lattice matrix and
// bxy, bxz, byz its off-diagonal terms.
struct RectView {
	double ax, bx, ay, by, az, bz;
	int nx, ny, nz, nxy;    // nxy = nx * gy
	int gy;                 // block rows in y: ny, or oy for a periodic container
	double boxx, boxy, boxz, xsp, ysp, zsp;
	int xper, yper, zper;
	int hx, hy, hz;
	double bxsq, max_len_sq, max_radius;
	const dbl4 *P;          // particles (x,y,z,r), block-major, every block
	const int *pid;         // their ids
	const int *bstart;      // block start offsets into P, nblocks+1
	const int *cstart;      // offsets of the cells (stored particles) per block, nblocks+1
	const double *mrad;     // 64x64 worklist radii for this block size
	const unsigned *wl;     // 64x64 worklist table
	const double *walls;    // walls, VGPU_WALL_STRIDE doubles each (wall_cut_plane)
	int nwalls;
	// periodic containers
	double bxy, bxz, byz;
	int ey, ez, wy, wz, oy, oz;
	const UnitVert *unit;   // the unit Voronoi cell
	int unit_p;
	unsigned long long *fv; // first visit of each block (null: not recorded)
	// ghost cells of a periodic container: the blocks they may read (null:
	// any); a search that reaches another block is given up (overflow)
	const unsigned char *img_ok;
	// batches of containers: this container's first block in the batch's
	// block arrays (bstart and cstart above already point at it)
	int boff;
};

// A wall record: type (0 sphere, 1 plane, 2 cylinder, 3 cone), w_id, then
// the wall's constants as the reference's constructors store them
// (src/wall.hh:27-115): sphere xc,yc,zc,rc; plane xc,yc,zc,ac; cylinder
// xc,yc,zc,xa,ya,za,asi,rc; cone xc,yc,zc,xa,ya,za,asi,sang,cang,gra.
enum { VGPU_WALL_STRIDE = 12 };

// The plane a wall cuts the cell of the particle at (x,y,z) with, as the
// wall's cut_cell_base computes it (src/wall.cc:28-36 sphere, 50-53 plane,
// 64-74 cylinder, 88-103 cone): c.nplane(px,py,pz,rs,pid).  False if the
// wall makes no cut (the particle within 1e-5 of the sphere centre or the
// axis).
VG_INL bool wall_cut_plane(const double *wp, double x, double y, double z, double &px, double &py, double &pz,
			   double &rs, int &pid) {
	const int type = (int)wp[0];
	pid = (int)wp[1];
	const double xc = wp[2], yc = wp[3], zc = wp[4];
	if (type == 1) {
		px = xc; py = yc; pz = zc;
		rs = 2 * (wp[5] - x * xc - y * yc - z * zc);
		return true;
	}
	if (type == 0) {
		const double xd = x - xc, yd = y - yc, zd = z - zc;
		double dq = xd * xd + yd * yd + zd * zd;
		if (dq > 1e-5) {
			dq = 2 * (sqrt(dq) * wp[5] - dq);
			px = xd; py = yd; pz = zd; rs = dq;
			return true;
		}
		return false;
	}
	const double xa = wp[5], ya = wp[6], za = wp[7], asi = wp[8];
	double xd = x - xc, yd = y - yc, zd = z - zc, pa = (xd * xa + yd * ya + zd * za) * asi;
	xd -= xa * pa; yd -= ya * pa; zd -= za * pa;
	pa = xd * xd + yd * yd + zd * zd;
	if (!(pa > 1e-5)) return false;
	if (type == 2) {
		pa = 2 * (sqrt(pa) * wp[9] - pa);
		px = xd; py = yd; pz = zd; rs = pa;
		return true;
	}
	const double sang = wp[9], cang = wp[10];
	pa = 1 / sqrt(pa);
	const double q = sqrt(asi);
	const double xf = -sang * q * xa + cang * pa * xd;
	const double yf = -sang * q * ya + cang * pa * yd;
	const double zf = -sang * q * za + cang * pa * zd;
	pa = 2 * (xf * (xc - x) + yf * (yc - y) + zf * (zc - z));
	px = xf; py = yf; pz = zf; rs = pa;
	return true;
}

// Per-cell statistics (labels for the candidate predictor).
struct CellStats {
	int n_planes;   // nplane calls
	int wl_exit;    // worklist index at which the radius test stopped (64 = phase 3)
	int phase;      // 1, 2 or 3
	int max_p;
	int img;        // periodic containers: an image block was read
};

// Worklist indices at which the reference refreshes mrs, src/v_compute.cc
// (count_list {7,11,15,19,26,35,45,59}), one byte each.
VG_INL int count_list(int i) { return (int)((0x3b2d231a130f0b07ull >> (8 * i)) & 0xffull); }

template<bool POLY>
struct RadPolicy {
	double r_rad, r_mul, r_val, max_radius;
	VG_INL void init(double r, double maxr) {
		max_radius = maxr;
		if (POLY) { r_rad = r * r; r_mul = (r - maxr) * (r + maxr); }
	}
	VG_INL void prime(double rv) { if (POLY) r_val = 1 + r_mul / rv; }
	VG_INL bool ctest(double crs, double mrs) const {
		if (!POLY) return crs > mrs;
		double cc = crs + r_mul;
		return cc > 0 && cc * cc > mrs * crs;
	}
	VG_INL double cutoff(double lrs) const { return POLY ? lrs * r_val : lrs; }
	VG_INL double scale(double rs, double rj) const { return POLY ? rs + r_rad - rj * rj : rs; }
	// the value scale_check leaves in rs
	VG_INL double scale_checked(double rs, double rj) const { return POLY ? rs + (r_rad - rj * rj) : rs; }
	VG_INL bool scale_check(double &rs, double mrs, double rj) const {
		if (!POLY) return rs < mrs;
		double trs = rs;
		rs += r_rad - rj * rj;
		return rs <= 0 || rs * rs < mrs * trs;
	}
};

template<class CELL>
struct BlockSearch {
	typedef typename CELL::L L;
	enum { W = L::W, WD = L::WD, MAXQ = L::MAXQ };
	CELL &c;
	unsigned stamp;
	int ci0, cj0, ck0;      // centre block (mask coordinates)
	int qh, qt;             // ring head / tail
	VG_INL BlockSearch(CELL &c_, unsigned st, int i, int j, int k)
		: c(c_), stamp(st), ci0(i), cj0(j), ck0(k), qh(0), qt(0) {}
	VG_INL auto BM(int w) { return c.template A<unsigned>(L::o_bm, w); }
	VG_INL auto BS(int w) { return c.template A<unsigned>(L::o_bms, w); }
	VG_INL auto QU(int t) { return c.template A<int>(L::o_qu, t); }
	// Window index of block (ei,ej,ek), or -1 (overflow) if outside.
	VG_INL int widx(int ei, int ej, int ek) {
		int a = ei - ci0 + W, b = ej - cj0 + W, d = ek - ck0 + W;
		if (a < 0 || a >= WD || b < 0 || b >= WD || d < 0 || d >= WD) { c.ovf = true; return -1; }
		return a + WD * (b + WD * d);
	}
	VG_INL bool marked(int ei, int ej, int ek) {
		int t = widx(ei, ej, ek);
		if (t < 0) return true;
		int w = t >> 5;
		return (unsigned)BS(w) == stamp && (((unsigned)BM(w) >> (t & 31)) & 1u);
	}
	VG_INL void mark(int ei, int ej, int ek) {
		int t = widx(ei, ej, ek);
		if (t < 0) return;
		int w = t >> 5;
		unsigned b = 1u << (t & 31);
		if ((unsigned)BS(w) != stamp) { BS(w) = stamp; BM(w) = b; }
		else BM(w) |= b;
	}
	VG_INL void push(int ei, int ej, int ek) {
		if (qt - qh >= MAXQ) { c.ovf = true; return; }
		QU(qt % MAXQ) = (ei - ci0 + 512) | ((ej - cj0 + 512) << 10) | ((ek - ck0 + 512) << 20);
		qt++;
	}
	VG_INL bool pop(int &ei, int &ej, int &ek) {
		if (qh == qt) return false;
		int v = QU(qh % MAXQ);
		qh++;
		ei = (v & 1023) - 512 + ci0; ej = ((v >> 10) & 1023) - 512 + cj0; ek = ((v >> 20) & 1023) - 512 + ck0;
		return true;
	}
	VG_INL void mark_push(int ei, int ej, int ek) { mark(ei, ej, ek); push(ei, ej, ek); }
};

// compute_min_max_radius, src/v_compute.cc:797-944.  The reference spells
// out all 26 neighbour directions; they reduce to seven forms (corner, three
// edges, three faces) once a negative direction is expressed through its
// sign: (-b)*x == -(b*x) and a+(-y) == a-y exactly in IEEE arithmetic, so
// every form evaluates the reference's expression for that direction
// operation for operation.
// The two radii of compute_min_max_radius for worklist offset (di,dj,dk):
// crs_near, the squared distance to the near corner/edge/face, which the
// reference tests against mrs first, and crs_far, which it goes on to
// compute from crs_near when the block is not skipped.
VG_FN void min_max_radius_pair(const RectView &v, int di, int dj, int dk, double fx, double fy, double fz, double gxs,
			       double gys, double gzs, double &crs_near, double &crs_far, int &form) {
	const double boxx = v.boxx, boxy = v.boxy, boxz = v.boxz;
	const double sx = di > 0 ? 1.0 : -1.0, sy = dj > 0 ? 1.0 : -1.0, sz = dk > 0 ? 1.0 : -1.0;
	const double xlo = (di > 0 ? di : di + 1) * boxx - fx;
	const double ylo = (dj > 0 ? dj : dj + 1) * boxy - fy;
	const double zlo = (dk > 0 ? dk : dk + 1) * boxz - fz;
	form = (di != 0 ? 4 : 0) | (dj != 0 ? 2 : 0) | (dk != 0 ? 1 : 0);
	// squared distance to the near corner/edge/face, axes in x,y,z order
	bool first = true;
	double crs = 0;
	if (di != 0) { crs = xlo * xlo; first = false; }
	if (dj != 0) { if (first) crs = ylo * ylo; else crs += ylo * ylo; first = false; }
	if (dk != 0) { if (first) crs = zlo * zlo; else crs += zlo * zlo; }
	crs_near = crs;
	const double ax = boxx * ((2 * sx) * xlo + boxx);
	const double ay = boxy * ((2 * sy) * ylo + boxy);
	const double az = boxz * ((2 * sz) * zlo + boxz);
	switch (form) {
	case 7: crs += v.bxsq + 2 * ((sx * boxx) * xlo + (sy * boxy) * ylo + (sz * boxz) * zlo); break;  // corner
	case 6: crs += ax + ay + gzs; break;                  // x,y: edge along z
	case 5: crs += az; crs += gys + ax; break;            // x,z: edge along y
	case 4: crs += gzs; crs += gys + ax; break;           // x: face
	case 3: crs += az; crs += ay; crs += gxs; break;      // y,z: edge along x
	case 2: crs += gzs; crs += ay; crs += gxs; break;     // y: face
	case 1: crs += az; crs += gys; crs += gxs; break;     // z: face
	default: crs = 0; crs += gys; crs += gxs; break;      // central block: not reached
	}
	crs_far = crs;
}

template<bool POLY>
VG_FN bool min_max_radius(const RectView &v, const RadPolicy<POLY> &rp, int di, int dj, int dk, double fx,
			  double fy, double fz, double gxs, double gys, double gzs, double &crs, double mrs) {
	double crs_far;
	int form;
	min_max_radius_pair(v, di, dj, dk, fx, fy, fz, gxs, gys, gzs, crs, crs_far, form);
	if (form != 0 && rp.ctest(crs, mrs)) return true;
	crs = crs_far;
	return false;
}

// Region tests, src/v_compute.cc:634-780 (corner_test, edge_[xyz]_test,
// face_[xyz]_test): returns true when no particle of the block can cut the
// cell.  Each test probes 6 (corner, edge) or 4 (face) points of the block
// with plane_intersects, the first with the _guess variant.  Per axis the
// point takes the near (bit 0) or far (bit 1) coordinate; an axis on which
// the block is level with the particle's block uses low (0) / high (1).  The
// plane offset is the sum over the non-level axes of near*coordinate, in x,
// y, z order, and r_prime gets the sum of the near squares -- exactly the
// arguments the reference passes.  The probe order per form is tabled here
// (3 bits per probe, x in bit 2).
template<class CELL, bool POLY>
VG_FN bool region_skip(CELL &c, RadPolicy<POLY> &rp, int ei, int ej, int ek, int i, int j, int k,
		       double xlo, double xhi, double ylo, double yhi, double zlo, double zhi) {
	const int sx = ei > i ? 1 : (ei < i ? -1 : 0);
	const int sy = ej > j ? 1 : (ej < j ? -1 : 0);
	const int sz = ek > k ? 1 : (ek < k ? -1 : 0);
	// near / far coordinate per axis
	const double xn = sx < 0 ? xhi : xlo, xf = sx < 0 ? xlo : xhi;
	const double yn = sy < 0 ? yhi : ylo, yf = sy < 0 ? ylo : yhi;
	const double zn = sz < 0 ? zhi : zlo, zf = sz < 0 ? zlo : zhi;
	const int form = (sx != 0 ? 4 : 0) | (sy != 0 ? 2 : 0) | (sz != 0 ? 1 : 0);
	// probe sequences, 3 bits each (first probe in the low bits)
	unsigned seq;
	int np = 6;
	switch (form) {
	case 7: seq = 4u | 6u << 3 | 2u << 6 | 3u << 9 | 1u << 12 | 5u << 15; break;  // corner_test
	case 3: seq = 1u | 5u << 3 | 4u << 6 | 0u << 9 | 2u << 12 | 6u << 15; break;  // edge_x_test
	case 5: seq = 1u | 3u << 3 | 2u << 6 | 0u << 9 | 4u << 12 | 6u << 15; break;  // edge_y_test
	case 6: seq = 2u | 3u << 3 | 1u << 6 | 0u << 9 | 4u << 12 | 5u << 15; break;  // edge_z_test
	case 4: seq = 0u | 1u << 3 | 3u << 6 | 2u << 9; np = 4; break;                  // face_x_test
	case 2: seq = 0u | 1u << 3 | 5u << 6 | 4u << 9; np = 4; break;                  // face_y_test
	case 1: seq = 0u | 2u << 3 | 6u << 6 | 4u << 9; np = 4; break;                  // face_z_test
	default: c.fatal = c.ovf = true; return true;                                    // central block
	}
	double pr;
	if (form == 7) pr = xn * xn + yn * yn + zn * zn;
	else if (form == 3) pr = yn * yn + zn * zn;
	else if (form == 5) pr = xn * xn + zn * zn;
	else if (form == 6) pr = xn * xn + yn * yn;
	else if (form == 4) pr = xn * xn;
	else if (form == 2) pr = yn * yn;
	else pr = zn * zn;
	rp.prime(pr);
	for (int t = 0; t < np; t++) {
		const unsigned b = (seq >> (3 * t)) & 7u;
		const double X = (b & 4u) ? xf : xn, Y = (b & 2u) ? yf : yn, Z = (b & 1u) ? zf : zn;
		double rsq;
		if (form == 7) rsq = xn * X + yn * Y + zn * Z;
		else if (form == 3) rsq = yn * Y + zn * Z;
		else if (form == 5) rsq = xn * X + zn * Z;
		else if (form == 6) rsq = xn * X + yn * Y;
		else if (form == 4) rsq = xn * X;
		else if (form == 2) rsq = yn * Y;
		else rsq = zn * Z;
		if (c.plane_probe(X, Y, Z, rp.cutoff(rsq), t == 0)) return false;
	}
	return true;
}

// container_base::region_index, src/container.hh:243-251
VG_INL int rect_region(const RectView &v, int ci, int cj, int ck, int ei, int ej, int ek, double &qx, double &qy,
		       double &qz, int disp) {
	if (v.xper) { if (ci + ei < v.nx) { ei += v.nx; qx = -(v.bx - v.ax); } else if (ci + ei >= (v.nx << 1)) { ei -= v.nx; qx = v.bx - v.ax; } else qx = 0; }
	if (v.yper) { if (cj + ej < v.ny) { ej += v.ny; qy = -(v.by - v.ay); } else if (cj + ej >= (v.ny << 1)) { ej -= v.ny; qy = v.by - v.ay; } else qy = 0; }
	if (v.zper) { if (ck + ek < v.nz) { ek += v.nz; qz = -(v.bz - v.az); } else if (ck + ek >= (v.nz << 1)) { ek -= v.nz; qz = v.bz - v.az; } else qz = 0; }
	return disp + ei + v.nx * (ej + v.ny * ek);
}

VG_INL int step_div_i(int a, int b) { return a >= 0 ? a / b : -1 + (a + 1) / b; }

// Records the first visit of block b: the smallest (cell << 24 | visit)
// key over all cells.
VG_INL void note_first_visit(unsigned long long *p, unsigned long long key) {
#if defined(__HIP_DEVICE_COMPILE__)
	if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > key) atomicMin(p, key);
#else
	if (*p > key) *p = key;
#endif
}

// container_periodic_base::region_index, src/container_prd.hh:163-168.
// The reference creates the image block here when it is first reached;
// the engine's image blocks are prepared beforehand (prd_host.h), and the
// visit is recorded so that their creation order can be replayed.
VG_INL int prd_region(const RectView &v, int ci, int cj, int ck, int ei, int ej, int ek, double &qx,
		      unsigned long long key, int &img) {
	int qi = ci + (ei - v.nx);
	const int qj = cj + (ej - v.ey), qk = ck + (ek - v.ez);
	const int iv = step_div_i(qi, v.nx);
	if (iv != 0) { qx = iv * v.bx; qi -= v.nx * iv; } else qx = 0;
	const int b = qi + v.nx * (qj + v.oy * qk);
	if (qj < v.ey || qj >= v.wy || qk < v.ez || qk >= v.wz) {
		img = 1;
		if (v.fv) note_first_visit(v.fv + b, key);
	}
	return b;
}

// Cell types that evaluate worklist steps one per lane (wavefront cells on
// the device).
template<class C, class = void> struct LaneSteps { static constexpr bool value = false; };
template<class C> struct LaneSteps<C, std::void_t<decltype(C::lane_parallel)>> {
	static constexpr bool value = C::lane_parallel;
};
#if VG_WDEV
VG_INL unsigned long long steps_ballot(bool b) { return __ballot(b); }
VG_INL int steps_lane() { return (int)threadIdx.x; }
// a value every lane holds alike, read back from memory into a scalar register
VG_INL int steps_uniform(int a) { return __builtin_amdgcn_readfirstlane(a); }
#else
VG_INL unsigned long long steps_ballot(bool) { return 0; }
VG_INL int steps_lane() { return 0; }
VG_INL int steps_uniform(int a) { return a; }
#endif
VG_INL unsigned long long steps_uniform64(unsigned long long a) {
	return (unsigned long long)(unsigned)steps_uniform((int)(unsigned)a)
	       | ((unsigned long long)(unsigned)steps_uniform((int)(unsigned)(a >> 32)) << 32);
}
VG_INL double steps_uniform_d(double a) {
	unsigned long long b;
	__builtin_memcpy(&b, &a, 8);
	b = steps_uniform64(b);
	__builtin_memcpy(&a, &b, 8);
	return a;
}

// Candidate data of compute_cell_rect: the worklist entries and radii, the
// particle range of a worklist block, and the particles themselves, each
// loaded where it is used.  (Staging them in the lanes of a wavefront cell,
// 64 at a time, measured slower: the staging registers cost more than the
// loads they save.)
template<class CELL, bool PRD>
struct CandFetch {
	VG_INL void wl_begin(const RectView &, const double *, const unsigned *, int, int, int, unsigned, unsigned, int,
			     int, int, int) {}
	VG_INL double rad(const double *radp, int g) const { return radp[g]; }
	VG_INL unsigned ent(const unsigned *e, int g) const { return e[g]; }
	VG_INL void range(const RectView &v, int, int rijk, int &bt, int &be) const {
		bt = v.bstart[rijk]; be = v.bstart[rijk + 1];
	}
	VG_INL void block(int, int) {}
	VG_INL void part(const RectView &v, int idx, dbl4 &o, int &pid) { o = v.P[idx]; pid = v.pid[idx]; }
};


// Where a learned-path list ended (compute_cell_list returning -1): the
// cell holds every cut of worklist steps 0 .. g-1, and the plain candidate
// loop continues at the top of step g, before its radius refresh, with the
// list loop's radius state (g < 0: nothing to continue).  acc: bit s+1 for
// each step s whose block was cut (not skipped by its distance), from which
// the phase-2 bookkeeping is rebuilt; img, key: the periodic image-read
// flag and the next first-visit key.  The list loop records only mh, the
// radius bound at the worklist start (mh[0]) and after each refresh
// (mh[1..9]); list_resume_fill derives the rest from the records.  A
// wavefront cell keeps the record in LDS, read where the worklist starts, so
// that it holds no registers across the cuts.
struct ListResume {
	int g, next_count, cpi, img;
	double mrs;
	unsigned long long acc, key;
	double mh[10];
};

// compute_cell (src/v_compute.cc:291-621) for a rectangular container or,
// with PRD, a periodic one (start cell, frac_pos and region_index of
// src/container_prd.hh:124-168).  (ijk,s) is the particle's block and
// index among the block's cells, (ci,cj,ck) the block coordinates, key0 the
// cell's loop position << 24 (first-visit keys).  Returns the reference's
// compute_cell result; c.ovf / c.fatal report capacity or internal trouble.
//
// The reference's nested loops (walls, own block, worklist phases 1-2, block
// search) are driven as one resumable candidate generator feeding a single
// nplane call site.  Block-level decisions (radius refresh, radius stop,
// block skip, region tests) are taken when the previous block's particles
// have all been cut, exactly when the reference takes them.  Phase 2's
// bookkeeping (mask marks and queue pushes) is only read by the block search,
// which under 1% of cells reach; it is therefore recorded as one bit per
// worklist entry and replayed, in the original order, when the block search
// starts.
//
// ghost (compute_ghost_cell, src/container.hh:488-498,710-721): the cell of a
// point that is not stored, as if put last in block ijk (s = the block's
// count) with the container's maximum radius ghost_maxr; a block search that
// comes back to block ijk (where the reference would also meet the point's
// own periodic image) ends as an overflow, and so does one that reaches a
// block v.img_ok excludes (periodic containers: an image block the reference
// would create during the call, copying the point into it).
template<class CELL, bool POLY, bool PRD = false>
VG_FN bool compute_cell_rect(CELL &c, const RectView &v, int ijk, int s, int ci, int cj, int ck,
			     unsigned stamp, CellStats &stt, unsigned long long key0 = 0, const dbl4 *ghost = nullptr,
			     double ghost_maxr = 0, const ListResume *resume = nullptr) {
	double x, y, z, qx = 0, qy = 0, qz = 0;
	int i, j, k, di, dj, dk, ei, ej, ek, f = 0, g = 0, disp = 0;
	double fx0 = 0, fy0 = 0, fz0 = 0, gxs0 = 0, gys0 = 0, gzs0 = 0;
	unsigned q, m1 = 0, m2 = 0;
	unsigned long long key = key0;
	const bool resumed = resume && steps_uniform(resume->g) >= 0;
	stt.n_planes = 0; stt.wl_exit = 0; stt.phase = 1; stt.img = resumed ? steps_uniform(resume->img) : 0;

	const int b0 = v.bstart[ijk], co0 = v.bstart[ijk + 1] - b0;
	const dbl4 me = ghost ? *ghost : v.P[b0 + s];
	x = me.x; y = me.y; z = me.z;
	if (PRD) {
		if (!resumed) {
			c.init_unit(v.unit, v.unit_p);
			if (c.ovf) return false;
		}
		i = v.nx; j = v.ey; k = v.ez;
	} else {
		double xa, xb, ya, yb, za, zb;
		if (v.xper) { xb = 0.5 * (v.bx - v.ax); xa = -xb; i = v.nx; } else { xa = v.ax - x; xb = v.bx - x; i = ci; }
		if (v.yper) { yb = 0.5 * (v.by - v.ay); ya = -yb; j = v.ny; } else { ya = v.ay - y; yb = v.by - y; j = cj; }
		if (v.zper) { zb = 0.5 * (v.bz - v.az); za = -zb; k = v.nz; } else { za = v.az - z; zb = v.bz - z; k = ck; }
		if (!resumed) c.init_box(xa, xb, ya, yb, za, zb);
		disp = ijk - i - v.nx * (j + v.ny * k);
	}
	RadPolicy<POLY> rp;
	rp.init(me.w, ghost ? ghost_maxr : v.max_radius);
	BlockSearch<CELL> bs(c, stamp, i, j, k);
	CandFetch<CELL, PRD> cf;
	cf.block(b0, b0 + co0);
	// Phase-2 bookkeeping is deferred: bit g of p2_cut says block g was
	// not skipped by the radius test, so its outside neighbours were due.
	unsigned long long p2_cut = 0;
	int g2_first = 0;

	double mrs = 0, crs = 0;
	int next_count = 3, cpi = 0;
	const double *radp = v.mrad;
	const unsigned *e = v.wl;
	// Generator state: 0 walls, 1 own block, 2 worklist, 4 replay of the
	// phase-2 bookkeeping, 3 block search.  A continued cell has its walls
	// and own block cut already.
	int stage = resumed ? 1 : 0, w = resumed ? v.nwalls : 0, t = resumed ? co0 : 0;
	bool marking = false;
	int bt = 0, be = 0;           // particle range of the block being cut
	bool bcheck = false;          // per-particle radius check for this block
	// the particle shifted by the current block's periodic image (the
	// particle itself until the worklist starts)
	double x2 = x, y2 = y, z2 = z;
	// The particle's own position is read again where it is needed after
	// the cuts have started (block changes, worklist selection) instead of
	// being held across them: wavefront cells run short of scalar registers.
	const int me_idx = b0 + s;
	auto reload = [&]() {
		if constexpr (LaneSteps<CELL>::value) {
			if (ghost) { x = ghost->x; y = ghost->y; z = ghost->z; }
			else {
				int q_ = me_idx;
				asm volatile("" : "+s"(q_));
				const dbl4 m_ = v.P[q_];
				x = m_.x; y = m_.y; z = m_.z;
			}
		}
	};

	// Upper bound on the cell's current max_radius_squared() (the cell only
	// shrinks), refreshed with the reference's mrs and after every 16
	// candidates; see the skip before the nplane call.
	double mrs_b = 1.7976931348623157e308;
	int since = 0;
	// Lane-parallel worklist steps (wavefront cells): the steps of one radius
	// interval, taken from step gbase, whose blocks are still to be cut
	// (pend) and need the per-particle check (bmask), bit l = step gbase + l;
	// gstop >= 0: the radius test stops the cell at that step once they are.
	unsigned long long pend = 0, bmask = 0;
	int gbase = 0, gstop = -1;
	// Lane-parallel particle checks: particles cbase + l of the current block
	// that are candidates (bit l of cm), up to 64 evaluated at a time.
	unsigned long long cm = 0;
	int cbase = 0;

	// frac_pos and the squared distances to the far block faces
	// (src/v_compute.cc:328-362), recomputed from the particle where the
	// wavefront cell needs them rather than held across the cuts.
	auto frac = [&](double &fx_, double &fy_, double &fz_, double &gx_, double &gy_, double &gz_) {
		reload();
		if (PRD) { fx_ = x - v.boxx * ci; fy_ = y - v.boxy * (cj - v.ey); fz_ = z - v.boxz * (ck - v.ez); }
		else { fx_ = x - v.ax - v.boxx * ci; fy_ = y - v.ay - v.boxy * cj; fz_ = z - v.az - v.boxz * ck; }
		gx_ = (m1 & 127u) ? fx_ : v.boxx - fx_;
		gy_ = (m1 & (127u << 7)) ? fy_ : v.boxy - fy_;
		gz_ = (m1 & (127u << 14)) ? fz_ : v.boxz - fz_;
		gx_ *= gx_; gy_ *= gy_; gz_ *= gz_;
	};
	constexpr bool refrac = LaneSteps<CELL>::value;
	// the worklist of the particle's sub-block (src/v_compute.cc:328-362)
	auto select_worklist = [&]() {
		double fx, fy, fz, gxs, gys, gzs;
		reload();
		if (PRD) { fx = x - v.boxx * ci; fy = y - v.boxy * (cj - v.ey); fz = z - v.boxz * (ck - v.ez); }
		else { fx = x - v.ax - v.boxx * ci; fy = y - v.ay - v.boxy * cj; fz = z - v.az - v.boxz * ck; }
		di = int(fx * v.xsp * VGPU_WL_FGRID); dj = int(fy * v.ysp * VGPU_WL_FGRID); dk = int(fz * v.zsp * VGPU_WL_FGRID);
		if (di >= VGPU_WL_HGRID) {
			gxs = fx;
			m1 = 127 + (3 << 21); m2 = 1 + (1 << 21); di = VGPU_WL_FGRID - 1 - di; if (di < 0) di = 0;
		} else { m1 = m2 = 0; gxs = v.boxx - fx; }
		if (dj >= VGPU_WL_HGRID) {
			gys = fy;
			m1 |= (127 << 7) + (3 << 24); m2 |= (1 << 7) + (1 << 24); dj = VGPU_WL_FGRID - 1 - dj; if (dj < 0) dj = 0;
		} else gys = v.boxy - fy;
		if (dk >= VGPU_WL_HGRID) {
			gzs = fz;
			m1 |= (127 << 14) + (3 << 27); m2 |= (1 << 14) + (1 << 27); dk = VGPU_WL_FGRID - 1 - dk; if (dk < 0) dk = 0;
		} else gzs = v.boxz - fz;
		gxs *= gxs; gys *= gys; gzs *= gzs;
		if constexpr (!refrac) { fx0 = fx; fy0 = fy; fz0 = fz; gxs0 = gxs; gys0 = gys; gzs0 = gzs; }
		const int wli = di + VGPU_WL_HGRID * (dj + VGPU_WL_HGRID * dk);
		radp = v.mrad + wli * 64;
		e = v.wl + wli * 64;
		cf.wl_begin(v, radp, e, i, j, k, m1, m2, ci, cj, ck, disp);
		f = cf.ent(e, 0); g = 0;
	};
	for (;;) {
		double px_ = 0, py_ = 0, pz_ = 0, rs = 0;
		double rs_raw = -1;           // squared distance of a particle candidate (walls: -1)
		int pid = 0;
		bool have = false;
		while (!have) {
			if constexpr (LaneSteps<CELL>::value) {
				if (cm) {
					// the next candidate of the block, in storage order
					const int idx = cbase + __builtin_ctzll(cm);
					cm &= cm - 1;
					dbl4 o;
					cf.part(v, idx, o, pid);
					px_ = o.x - x2; py_ = o.y - y2; pz_ = o.z - z2;
					rs_raw = px_ * px_ + py_ * py_ + pz_ * pz_;
					rs = bcheck ? rp.scale_checked(rs_raw, o.w) : rp.scale(rs_raw, o.w);
					have = true;
					continue;
				}
				if (bt < be) {
					// Up to 64 particles of the block at once, one per lane,
					// each with the per-particle test of src/v_compute.cc:402-428
					// (mrs is constant within a block).
					const int cnt = be - bt < 64 ? be - bt : 64;
					const int lane = steps_lane();
					bool pass = false;
					if (lane < cnt) {
						dbl4 o;
						int pl;
						cf.part(v, bt + lane, o, pl);
						const double ax = o.x - x2, ay = o.y - y2, az = o.z - z2;
						double r2 = ax * ax + ay * ay + az * az;
						pass = !bcheck || rp.scale_check(r2, mrs, o.w);
					}
					cm = steps_ballot(pass);
					cbase = bt;
					bt += cnt;
					continue;
				}
			}
			if (bt < be) {
				// Next particle of the current block (src/v_compute.cc:402-428).
				dbl4 o;
				cf.part(v, bt, o, pid);
				bt++;
				px_ = o.x - x2; py_ = o.y - y2; pz_ = o.z - z2;
				rs = px_ * px_ + py_ * py_ + pz_ * pz_;
				rs_raw = rs;
				if (!bcheck) { rs = rp.scale(rs, o.w); have = true; }
				else have = rp.scale_check(rs, mrs, o.w);
				continue;
			}
			if (stage == 0) {
				// Walls in the order they were added (src/container.hh:86-89).
				if (w < v.nwalls) {
					const double *wp = v.walls + VGPU_WALL_STRIDE * w;
					w++;
					have = wall_cut_plane(wp, x2, y2, z2, px_, py_, pz_, rs, pid);
					continue;
				}
				stage = 1;
				continue;
			}
			if (stage == 1) {
				// Own block in storage order, skipping the particle itself
				// (src/v_compute.cc:307-323).
				if (t < co0) {
					if (t == s) { t++; continue; }
					dbl4 o;
					cf.part(v, b0 + t, o, pid);
					t++;
					px_ = o.x - x2; py_ = o.y - y2; pz_ = o.z - z2;
					rs_raw = px_ * px_ + py_ * py_ + pz_ * pz_;
					rs = rp.scale(rs_raw, o.w);
					have = true;
					continue;
				}
				// Worklist selection (src/v_compute.cc:328-362).
				mrs = c.max_radius_squared();
				mrs_b = mrs; since = 0;
				select_worklist();
				if (resume) {
					const int gr = steps_uniform(resume->g);
					if (gr >= 0) {
						// continue a list that ended: the worklist starts
						// again at the top of step gr
						g = gr; mrs = steps_uniform_d(resume->mrs); mrs_b = mrs;
						next_count = steps_uniform(resume->next_count);
						cpi = steps_uniform(resume->cpi);
						key = steps_uniform64(resume->key);
						const int s1 = f > 1 ? f : 1;     // the step at whose top phase 2 starts
						if (g > s1) {
							marking = true; stt.phase = 2; g2_first = s1 + 1;
							p2_cut = steps_uniform64(resume->acc) & ~((2ull << s1) - 1);   // bits s1+1 ..
						}
					}
				}
				stage = 2;
				continue;
			}
			if constexpr (LaneSteps<CELL>::value) if (stage == 2) {
				if (pend) {
					// the next block of the interval, in worklist order
					const int l = __builtin_ctzll(pend);
					pend &= pend - 1;
					const int g1 = gbase + l + 1;
					q = cf.ent(e, g1); q ^= m1; q += m2;
					ei = (int)(q & 127) - 64 + i;
					ej = (int)((q >> 7) & 127) - 64 + j;
					ek = (int)((q >> 14) & 127) - 64 + k;
					const int rijk = PRD ? prd_region(v, ci, cj, ck, ei, ej, ek, qx, key++, stt.img)
							     : rect_region(v, ci, cj, ck, ei, ej, ek, qx, qy, qz, disp);
					if (ghost && (rijk == ijk || (PRD && v.img_ok && !v.img_ok[rijk]))) { c.ovf = true; stt.n_planes = c.n_cuts; return false; }
					cf.range(v, g1, rijk, bt, be);
					cf.block(bt, be);
					reload();
					x2 = x - qx; y2 = y - qy; z2 = z - qz;
					bcheck = ((bmask >> l) & 1ull) != 0;
					continue;
				}
				if (gstop >= 0) { stt.wl_exit = gstop; stt.n_planes = c.n_cuts; return true; }
				if (g >= 63) {
					if (rp.ctest(cf.rad(radp, g), mrs)) { stt.wl_exit = g; stt.n_planes = c.n_cuts; return true; }
					stt.wl_exit = 64;
					stt.phase = 3;
					stage = 4;
					continue;
				}
				// The top of step g (phase-2 start, radius refresh), then every
				// step up to the next refresh at once: mrs is constant there, so
				// each step's radius stop, block bounds and block skip
				// (src/v_compute.cc:371-431,443-526) are decided in its own lane
				// with the expressions of the sequential loop; the blocks are
				// then cut in worklist order.
				if (!marking && g >= f && g > 0) { marking = true; stt.phase = 2; g2_first = g + 1; }
				if (g == next_count) {
					mrs = c.max_radius_squared();
					mrs_b = mrs; since = 0;
					if (cpi != 8) next_count = count_list(cpi++);
				}
				const int gE = (next_count > g && next_count < 63) ? next_count : 63;
				const int n = gE - g;
				const int lane = steps_lane();
				double fx = fx0, fy = fy0, fz = fz0, gxs = gxs0, gys = gys0, gzs = gzs0;
				if constexpr (refrac) frac(fx, fy, fz, gxs, gys, gzs);
				const int gg = g + lane;
				bool stop = false, acc = false, bc = false, mk = false;
				if (lane < n) {
					stop = rp.ctest(radp[gg], mrs);
					unsigned qq = e[gg + 1];
					qq ^= m1; qq += m2;
					const int ddi = (int)(qq & 127) - 64, ddj = (int)((qq >> 7) & 127) - 64,
						  ddk = (int)((qq >> 14) & 127) - 64;
					const int eei = ddi + i, eej = ddj + j, eek = ddk + k;
					if (eei >= 0 && eei < v.hx && eej >= 0 && eej < v.hy && eek >= 0 && eek < v.hz) {
						double crs_l = 0;
						if (!min_max_radius<POLY>(v, rp, ddi, ddj, ddk, fx, fy, fz, gxs, gys, gzs, crs_l, mrs)) {
							acc = true;
							bc = rp.ctest(crs_l, mrs);
						}
					}
					mk = marking || (gg >= f && gg > 0);
				}
				const unsigned long long sm = steps_ballot(stop);
				const int s0 = sm ? __builtin_ctzll(sm) : n;
				const unsigned long long below = (1ull << s0) - 1;   // s0 <= 62
				pend = steps_ballot(acc) & below;
				bmask = steps_ballot(bc) & below;
				p2_cut |= (steps_ballot(acc && mk) & below) << (g + 1);
				// phase 2 starts at the first later step whose top is reached
				const int last = s0 < n ? g + s0 : g + n - 1;
				if (!marking) {
					const int gm = f > g + 1 ? f : g + 1;
					if (gm <= last) { marking = true; stt.phase = 2; g2_first = gm + 1; }
				}
				gbase = g;
				if (s0 < n) gstop = g + s0;
				g += s0 < n ? s0 : n;
				continue;
			}
			if (stage == 2) {
				if (g >= 63) {
					// Worklist exhausted (src/v_compute.cc:529).
					if (rp.ctest(cf.rad(radp, g), mrs)) { stt.wl_exit = g; stt.n_planes = c.n_cuts; return true; }
					stt.wl_exit = 64;
					stt.phase = 3;
					stage = 4;
					continue;
				}
				// Phase 1 is a do-while over g < f (src/v_compute.cc:366-431):
				// its first iteration always runs.  Phase 2 marks the mask
				// and queues outside neighbours (src/v_compute.cc:443-526).
				if (!marking && g >= f && g > 0) { marking = true; stt.phase = 2; g2_first = g + 1; }
				if (g == next_count) {
					mrs = c.max_radius_squared();
					mrs_b = mrs; since = 0;
					if (cpi != 8) {
						next_count = count_list(cpi++);
					}
				}
				if (rp.ctest(cf.rad(radp, g), mrs)) { stt.wl_exit = g; stt.n_planes = c.n_cuts; return true; }
				g++;
				q = cf.ent(e, g); q ^= m1; q += m2;
				di = q & 127; di -= 64;
				dj = (q >> 7) & 127; dj -= 64;
				dk = (q >> 14) & 127; dk -= 64;
				ei = di + i; if (ei < 0 || ei >= v.hx) continue;
				ej = dj + j; if (ej < 0 || ej >= v.hy) continue;
				ek = dk + k; if (ek < 0 || ek >= v.hz) continue;
				if (min_max_radius<POLY>(v, rp, di, dj, dk, fx0, fy0, fz0, gxs0, gys0, gzs0, crs, mrs)) continue;
				if (marking) p2_cut |= 1ull << g;
				{
					const int rijk = PRD ? prd_region(v, ci, cj, ck, ei, ej, ek, qx, key++, stt.img)
							     : rect_region(v, ci, cj, ck, ei, ej, ek, qx, qy, qz, disp);
					if (ghost && (rijk == ijk || (PRD && v.img_ok && !v.img_ok[rijk]))) { c.ovf = true; stt.n_planes = c.n_cuts; return false; }
					cf.range(v, g, rijk, bt, be);
					cf.block(bt, be);
					reload();
					x2 = x - qx; y2 = y - qy; z2 = z - qz;
					bcheck = rp.ctest(crs, mrs);
				}
				continue;
			}
			if (stage == 4) {
				// Entering the block search: replay phase 2's mask marks and
				// queue pushes (src/v_compute.cc:443-526) in their original
				// order -- each visited worklist block is marked, and a block
				// not skipped by the radius test queues its outside
				// neighbours (scan_bits_mask_add, src/v_compute.cc:249-264).
				const unsigned b1 = 1u << 21, b2 = 1u << 22, b3 = 1u << 24, b4 = 1u << 25, b5 = 1u << 27, b6 = 1u << 28;
				for (int gg = g2_first; gg <= 63 && g2_first > 0; gg++) {
					q = cf.ent(e, gg); q ^= m1; q += m2;
					ei = (int)(q & 127) - 64 + i; if (ei < 0 || ei >= v.hx) continue;
					ej = (int)((q >> 7) & 127) - 64 + j; if (ej < 0 || ej >= v.hy) continue;
					ek = (int)((q >> 14) & 127) - 64 + k; if (ek < 0 || ek >= v.hz) continue;
					bs.mark(ei, ej, ek);
					if (!((p2_cut >> gg) & 1ull)) continue;
					if ((q & b2) == b2) {
						if (ei > 0) bs.mark_push(ei - 1, ej, ek);
						if ((q & b1) == 0 && ei < v.hx - 1) bs.mark_push(ei + 1, ej, ek);
					} else if ((q & b1) == b1 && ei < v.hx - 1) bs.mark_push(ei + 1, ej, ek);
					if ((q & b4) == b4) {
						if (ej > 0) bs.mark_push(ei, ej - 1, ek);
						if ((q & b3) == 0 && ej < v.hy - 1) bs.mark_push(ei, ej + 1, ek);
					} else if ((q & b3) == b3 && ej < v.hy - 1) bs.mark_push(ei, ej + 1, ek);
					if ((q & b6) == b6) {
						if (ek > 0) bs.mark_push(ei, ej, ek - 1);
						if ((q & b5) == 0 && ek < v.hz - 1) bs.mark_push(ei, ej, ek + 1);
					} else if ((q & b5) == b5 && ek < v.hz - 1) bs.mark_push(ei, ej, ek + 1);
					if (c.ovf) { stt.n_planes = c.n_cuts; return false; }
				}
				stage = 3;
				continue;
			}
			// stage 3: breadth-first over queued blocks (src/v_compute.cc:534-618).
			if (!bs.pop(ei, ej, ek)) { stt.n_planes = c.n_cuts; return true; }
			{
				double fx = fx0, fy = fy0, fz = fz0, gxs = gxs0, gys = gys0, gzs = gzs0;
				if constexpr (refrac) frac(fx, fy, fz, gxs, gys, gzs);
				(void)gxs; (void)gys; (void)gzs;
				const double xlo = (ei - i) * v.boxx - fx, xhi = xlo + v.boxx;
				const double ylo = (ej - j) * v.boxy - fy, yhi = ylo + v.boxy;
				const double zlo = (ek - k) * v.boxz - fz, zhi = zlo + v.boxz;
				if (region_skip<CELL, POLY>(c, rp, ei, ej, ek, i, j, k, xlo, xhi, ylo, yhi, zlo, zhi)) {
					if (c.ovf) { stt.n_planes = c.n_cuts; return false; }
					continue;
				}
			}
			{
				const int rijk = PRD ? prd_region(v, ci, cj, ck, ei, ej, ek, qx, key++, stt.img)
						     : rect_region(v, ci, cj, ck, ei, ej, ek, qx, qy, qz, disp);
				if (ghost && (rijk == ijk || (PRD && v.img_ok && !v.img_ok[rijk]))) { c.ovf = true; stt.n_planes = c.n_cuts; return false; }
				bt = v.bstart[rijk]; be = v.bstart[rijk + 1];
				cf.block(bt, be);
				reload();
				x2 = x - qx; y2 = y - qy; z2 = z - qz;
				bcheck = false;
			}
			// add_to_mask, src/v_compute.cc:236-243
			if (ek > 0 && !bs.marked(ei, ej, ek - 1)) bs.mark_push(ei, ej, ek - 1);
			if (ej > 0 && !bs.marked(ei, ej - 1, ek)) bs.mark_push(ei, ej - 1, ek);
			if (ei > 0 && !bs.marked(ei - 1, ej, ek)) bs.mark_push(ei - 1, ej, ek);
			if (ei < v.hx - 1 && !bs.marked(ei + 1, ej, ek)) bs.mark_push(ei + 1, ej, ek);
			if (ej < v.hy - 1 && !bs.marked(ei, ej + 1, ek)) bs.mark_push(ei, ej + 1, ek);
			if (ek < v.hz - 1 && !bs.marked(ei, ej, ek + 1)) bs.mark_push(ei, ej, ek + 1);
			if (c.ovf) { stt.n_planes = c.n_cuts; return false; }
		}
		if (rs_raw >= 0) {
			// Every vertex V (doubled coordinates, |V|^2 <= mrs_b) has
			// x.V - rs <= sqrt(rs_raw mrs_b) - rs: when that is below
			// -2 tol the plane misses the cell with every vertex strictly
			// inside (rounding of both sides is far below tol), and the
			// reference's nplane returns without changing the cell
			// (src/cell.cc:719-757).  The call is skipped; it still counts.
			if (++since >= 16) { mrs_b = c.max_radius_squared(); since = 0; }
			if (rs - sqrt(rs_raw * mrs_b) > 2.0 * c.tol) { c.n_cuts++; continue; }
		}
		if (!c.nplane(px_, py_, pz_, rs, pid)) { stt.n_planes = c.n_cuts; return false; }
	}
}

// ---------------------------------------------------------------------------
// Candidate lists (the learned path).  A list holds one record per worklist
// step of one cell up to a depth G, in the reference's order
// (src/v_compute.cc:328-431), then an end record:
//   LK_BLOCK    step g -> g+1, block rijk in the grid: radp[g], crs_near,
//               crs_far, the image shift of the block
//   LK_NOBLOCK  step g -> g+1, block outside the grid: radp[g]
//   LK_END      the end of the list at depth G: radp[G]
// crs_near / crs_far are the two radii compute_min_max_radius derives for
// the block (src/v_compute.cc:797-944) before and after its mrs test.  A
// record is independent of the cell's shape, so the records are written by a
// separate gather pass; the particles are read from the block arrays when
// the block is cut, and the decisions that depend on the shape (count_list
// refresh, the radius stop, the block skip, the per-particle check) are taken
// here exactly as compute_cell_rect takes them.  info: the kind (bits 0-2),
// the block is a periodic container's image block (bit 3; its first visit is
// recorded here, in the order and with the key compute_cell_rect records it:
// the gather must not record visits, the cell may stop before the block),
// and the shift of the block's image per axis (bits 4-5 x, 6-7 y, 8-9 z:
// 0 none, 1 minus one period, 2 plus one period).
struct ListRec { double rad, cn, cf; int rijk, info; };
enum { LK_BLOCK = 2, LK_NOBLOCK = 3, LK_END = 4 };

// The record's image shift per axis code: the values rect_region and
// prd_region give.
VG_INL double list_shift(int code, double period) { return code == 1 ? -period : (code == 2 ? period : 0.0); }
VG_INL int list_code(double q) { return q < 0 ? 1 : (q > 0 ? 2 : 0); }

// compute_cell over a candidate list.  Returns 1 / 0 (the cell exists /
// was deleted), -1 when the list ends before the radius test stops the
// cell (the plain path must finish it), -2 on overflow (c.ovf), -3 for an
// empty list (no list was gathered: the plain path computes the cell).
// key0: the cell's loop position << 24 (periodic first-visit keys).  res:
// where a list that ended left off (-1), for compute_cell_rect to continue
// the cell from there instead of computing it again.
template<class CELL, bool POLY, bool PRD = false>
VG_FN int compute_cell_list(CELL &c, const RectView &v, int ijk, int s, const ListRec *__restrict__ L, int stride,
			    int nrec, CellStats &stt, unsigned long long key0 = 0, ListResume *res = nullptr) {
	stt.n_planes = 0; stt.wl_exit = 0; stt.phase = 1; stt.img = 0;
	if (nrec <= 0) return -3;
	unsigned long long key = key0;
	const int b0 = v.bstart[ijk], co0 = v.bstart[ijk + 1] - b0;
	const dbl4 me = v.P[b0 + s];
	double x = me.x, y = me.y, z = me.z;
	if (PRD) {
		c.init_unit(v.unit, v.unit_p);
		if (c.ovf) { stt.n_planes = 0; return -2; }
	} else {
		double xa, xb, ya, yb, za, zb;
		if (v.xper) { xb = 0.5 * (v.bx - v.ax); xa = -xb; } else { xa = v.ax - x; xb = v.bx - x; }
		if (v.yper) { yb = 0.5 * (v.by - v.ay); ya = -yb; } else { ya = v.ay - y; yb = v.by - y; }
		if (v.zper) { zb = 0.5 * (v.bz - v.az); za = -zb; } else { za = v.az - z; zb = v.bz - z; }
		c.init_box(xa, xb, ya, yb, za, zb);
	}
	RadPolicy<POLY> rp;
	rp.init(me.w, v.max_radius);
	double mrs = 0;
	int next_count = 3, cpi = 0, g = 0, w = 0, t = 0, k = 0;
	bool own_done = false, bcheck = false;
	int bt = 0, be = 0;              // particle range of the block being cut
	// the cell's particle, shifted by the block's image (the particle itself
	// before the first list block); the particle is read again at each block
	// change rather than held across the cuts (as in compute_cell_rect)
	double x2 = x, y2 = y, z2 = z;
	const int me_idx = b0 + s;
	auto reload = [&]() {
		if constexpr (LaneSteps<CELL>::value) {
			int q_ = me_idx;
			asm volatile("" : "+s"(q_));
			const dbl4 m_ = v.P[q_];
			x = m_.x; y = m_.y; z = m_.z;
		}
	};
	// upper bound on max_radius_squared(), as in compute_cell_rect
	double mrs_b = 1.7976931348623157e308;
	int since = 0;
	// lane-parallel particle checks (as in compute_cell_rect): particles
	// cbase + l of a checked block that are candidates
	unsigned long long cm = 0;
	int cbase = 0;
	for (;;) {
		double px_ = 0, py_ = 0, pz_ = 0, rs = 0;
		double rs_raw = -1;
		int pid = 0;
		bool have = false;
		while (!have) {
			if (w < v.nwalls) {
				// Walls in the order they were added (src/container.hh:86-89).
				const double *wp = v.walls + VGPU_WALL_STRIDE * w;
				w++;
				have = wall_cut_plane(wp, x2, y2, z2, px_, py_, pz_, rs, pid);
				continue;
			}
			if constexpr (LaneSteps<CELL>::value) {
				if (cm) {
					const int idx = cbase + __builtin_ctzll(cm);
					cm &= cm - 1;
					const dbl4 o = v.P[idx];
					pid = v.pid[idx];
					px_ = o.x - x2; py_ = o.y - y2; pz_ = o.z - z2;
					rs_raw = px_ * px_ + py_ * py_ + pz_ * pz_;
					rs = rp.scale_checked(rs_raw, o.w);
					have = true;
					continue;
				}
				if (bcheck && bt < be) {
					// up to 64 particles of a checked block at once, one per
					// lane (src/v_compute.cc:402-428; mrs is constant within a block)
					const int cnt = be - bt < 64 ? be - bt : 64;
					const int lane = steps_lane();
					bool pass = false;
					if (lane < cnt) {
						const dbl4 o = v.P[bt + lane];
						const double ax = o.x - x2, ay = o.y - y2, az = o.z - z2;
						double r2 = ax * ax + ay * ay + az * az;
						pass = rp.scale_check(r2, mrs, o.w);
					}
					cm = steps_ballot(pass);
					cbase = bt;
					bt += cnt;
					continue;
				}
			}
			if (bt < be) {
				// next particle of the block (src/v_compute.cc:402-428)
				const dbl4 o = v.P[bt];
				pid = v.pid[bt];
				bt++;
				px_ = o.x - x2; py_ = o.y - y2; pz_ = o.z - z2;
				rs = px_ * px_ + py_ * py_ + pz_ * pz_;
				rs_raw = rs;
				if (!bcheck) { rs = rp.scale(rs, o.w); have = true; }
				else have = rp.scale_check(rs, mrs, o.w);
				continue;
			}
			if (!own_done) {
				// own block in storage order, skipping the particle itself
				// (src/v_compute.cc:307-323)
				if (t < co0) {
					if (t == s) { t++; continue; }
					const dbl4 o = v.P[b0 + t];
					pid = v.pid[b0 + t];
					t++;
					px_ = o.x - x2; py_ = o.y - y2; pz_ = o.z - z2;
					rs_raw = px_ * px_ + py_ * py_ + pz_ * pz_;
					rs = rp.scale(rs_raw, o.w);
					have = true;
					continue;
				}
				// worklist selection (src/v_compute.cc:328)
				mrs = c.max_radius_squared();
				mrs_b = mrs; since = 0;
				own_done = true;
				if (res) res->mh[0] = mrs;
			}
			if (k >= nrec) { stt.n_planes = c.n_cuts; return -1; }
			const ListRec a = L[(size_t)k * stride];
			k++;
			if (g < 63 && g == next_count) {
				mrs = c.max_radius_squared();
				mrs_b = mrs; since = 0;
				if (res) res->mh[cpi + 1] = mrs;
				if (cpi != 8) next_count = count_list(cpi++);
			}
			if (rp.ctest(a.rad, mrs)) { stt.wl_exit = g; stt.n_planes = c.n_cuts; return 1; }
			const int kind = a.info & 7;
			if (kind == LK_END) {
				stt.wl_exit = g >= 63 ? 64 : g;
				stt.n_planes = c.n_cuts;
				return -1;
			}
			g++;
			if (kind == LK_NOBLOCK) continue;
			if (rp.ctest(a.cn, mrs)) continue;   // compute_min_max_radius skip
			if (PRD) {
				// region_index of the block (src/container_prd.hh:163-168):
				// every call takes a key, image blocks record the visit
				if ((a.info >> 3) & 1) {
					stt.img = 1;
					if (v.fv) note_first_visit(v.fv + a.rijk, key);
				}
				key++;
			}
			bcheck = rp.ctest(a.cf, mrs);
			{
				double qx, qy = 0, qz = 0;
				if (PRD) qx = list_shift((a.info >> 4) & 3, v.bx);
				else {
					qx = list_shift((a.info >> 4) & 3, v.bx - v.ax);
					qy = list_shift((a.info >> 6) & 3, v.by - v.ay);
					qz = list_shift((a.info >> 8) & 3, v.bz - v.az);
				}
				reload();
				x2 = x - qx; y2 = y - qy; z2 = z - qz;
			}
			bt = v.bstart[a.rijk];
			be = v.bstart[a.rijk + 1];
		}
		if (rs_raw >= 0) {
			// the exact skip of a plane that provably misses the cell (see
			// compute_cell_rect); the call still counts
			if (++since >= 16) { mrs_b = c.max_radius_squared(); since = 0; }
			if (rs - sqrt(rs_raw * mrs_b) > 2.0 * c.tol) { c.n_cuts++; continue; }
		}
		if (!c.nplane(px_, py_, pz_, rs, pid)) { stt.n_planes = c.n_cuts; return c.ovf ? -2 : 0; }
	}
}

// The rest of a ListResume after compute_cell_list returned -1 over the
// nrec records of L (mh filled by that call): the records are one per
// worklist step, then the end record if the list has one, so the list ended
// at the top of step g; the loop's refresh schedule, block skips and
// first-visit keys are replayed from the records with the recorded radius
// bounds.
template<bool POLY, bool PRD>
VG_FN void list_resume_fill(const RectView &v, int ijk, int s, const ListRec *__restrict__ L, int stride, int nrec,
			    unsigned long long key0, ListResume &res) {
	const int g = (L[(size_t)(nrec - 1) * stride].info & 7) == LK_END ? nrec - 1 : nrec;
	RadPolicy<POLY> rp;
	rp.init(v.P[v.bstart[ijk] + s].w, v.max_radius);
	int nc = 3, cp = 0, img = 0;
	double mrs = steps_uniform_d(res.mh[0]);
	unsigned long long acc = 0, key = key0;
	for (int t = 0; t < g; t++) {
		if (t < 63 && t == nc) {
			mrs = steps_uniform_d(res.mh[cp + 1]);
			if (cp != 8) nc = count_list(cp++);
		}
		const ListRec a = L[(size_t)t * stride];
		if ((a.info & 7) != LK_BLOCK || rp.ctest(a.cn, mrs)) continue;
		acc |= 1ull << (t + 1);
		if (PRD) {
			if ((a.info >> 3) & 1) img = 1;
			key++;
		}
	}
	res.g = g; res.next_count = nc; res.cpi = cp; res.img = img;
	res.mrs = mrs; res.acc = acc; res.key = key;
}

}  // namespace vgpu
