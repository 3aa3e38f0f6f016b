// Host side of the periodic containers (container_periodic and
// container_periodic_poly, src/container_prd.hh/.cc, src/unitcell.cc): the
// geometry, the unit Voronoi cell every particle's cell starts from, and the
// periodic image blocks.
//
// Image blocks.  The reference fills the blocks beyond the primary domain in
// y and z lazily: the first time a cell's computation reaches an image block
// (region_index, src/container_prd.hh:163-168), create_periodic_image copies
// the particles that belong to it -- and the part of its neighbours' share
// that comes from the same primary blocks -- out of the primary domain.  The
// order of the particles in an image block, which is candidate order and so
// decides vertex numbering and rounding, therefore depends on the order in
// which image blocks are first reached over the whole loop.  A block is
// complete before it is first read and never changes afterwards, so the
// reference's loop is the same as computing every cell against the block
// arrays that result from creating the image blocks in order of first visit.
// ImageSet restates the creation routines so that the engine can replay
// them in the order of first visits it records on the GPU.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "compute_dev.h"

namespace vgpu {

inline int prd_step_int(double a) { return a < 0 ? int(a) - 1 : int(a); }
inline int prd_step_div(int a, int b) { return a >= 0 ? a / b : -1 + (a + 1) / b; }

struct PrdGeom {
	double bx, bxy, by, bxz, byz, bz;
	int nx, ny, nz;
	double boxx, boxy, boxz, xsp, ysp, zsp;
	int ey = 0, ez = 0, wy = 0, wz = 0, oy = 0, oz = 0;
	double max_len_sq = 0;              // unit_voro.max_radius_squared()
	double max_uv_y = 0, max_uv_z = 0;
	std::vector<UnitVert> unit;         // unit_voro, vertex by vertex
	int unit_max_order = 0;

	// unitcell::unitcell (src/unitcell.cc:26-85): a box of +-10 periods cut
	// by shells of lattice images until a shell no longer cuts it; then the
	// container's derived sizes (src/container_prd.cc:30-37).  Returns false
	// where the reference stops with "Periodic cell computation failed" or
	// the cell does not fit the engine's start-cell record.
	bool build(double bx_, double bxy_, double by_, double bxz_, double byz_, double bz_, int nx_, int ny_,
		   int nz_) {
		bx = bx_; bxy = bxy_; by = by_; bxz = bxz_; byz = byz_; bz = bz_;
		nx = nx_; ny = ny_; nz = nz_;
		boxx = bx / nx; boxy = by / ny; boxz = bz / nz;
		xsp = 1 / boxx; ysp = 1 / boxy; zsp = 1 / boxz;
		const int shells = 10;   // max_unit_voro_shells, src/config.hh
		typedef VCell<TierWide, 1> UC;
		std::vector<char> slab(UC::L::bytes + 2 * 65536, 0);
		UC c;
		c.setup(slab.data() + 65536, 0, shells * shells * 4 * (bx * bx + by * by + bz * bz));
		const double ucx = shells * bx, ucy = shells * by, ucz = shells * bz;
		c.init_box(-ucx, ucx, -ucy, ucy, -ucz, ucz);
		auto test = [&](int i, int j, int k) {   // unit_voro_test, src/unitcell.cc:196-200
			double x = i * bx + j * bxy + k * bxz, y = j * by + k * byz, z = k * bz;
			double rsq = x * x + y * y + z * z;
			return c.plane_probe(x, y, z, rsq, false);
		};
		bool good = true;
		auto apply = [&](int i, int j, int k) {  // unit_voro_apply, src/unitcell.cc:149-154
			double x = i * bx + j * bxy + k * bxz, y = j * by + k * byz, z = k * bz;
			if (!good) return;
			c.nplane(x, y, z, x * x + y * y + z * z, 0);
			c.nplane(-x, -y, -z, (-x) * (-x) + (-y) * (-y) + (-z) * (-z), 0);
			if (c.ovf) good = false;
		};
		auto intersect = [&](int l) {            // unit_voro_intersect, src/unitcell.cc:174-190
			if (test(l, 0, 0)) return true;
			for (int i = 1; i < l; i++) { if (test(l, i, 0)) return true; if (test(-l, i, 0)) return true; }
			for (int i = -l; i <= l; i++) if (test(i, l, 0)) return true;
			for (int i = 1; i < l; i++) for (int j = -l + 1; j <= l; j++) {
				if (test(l, j, i)) return true;
				if (test(-j, l, i)) return true;
				if (test(-l, -j, i)) return true;
				if (test(j, -l, i)) return true;
			}
			for (int i = -l; i <= l; i++) for (int j = -l; j <= l; j++) if (test(i, j, l)) return true;
			return false;
		};
		for (int l = 1; l < 2 * shells; l++) {
			if (intersect(l)) {
				apply(l, 0, 0);
				for (int i = 1; i < l; i++) { apply(l, i, 0); apply(-l, i, 0); }
				for (int i = -l; i <= l; i++) apply(i, l, 0);
				for (int i = 1; i < l; i++) for (int j = -l + 1; j <= l; j++) {
					apply(l, j, i); apply(-j, l, i); apply(-l, -j, i); apply(j, -l, i);
				}
				for (int i = -l; i <= l; i++) for (int j = -l; j <= l; j++) apply(i, j, l);
				if (!good) return false;
				continue;
			}
			// reach of the cell in y and z (src/unitcell.cc:61-77)
			max_uv_y = max_uv_z = 0;
			for (int v = 0; v < c.p; v++) {
				double q = c.X(v), y = c.Y(v), z = c.Z(v);
				q = std::sqrt(q * q + y * y + z * z);
				if (y + q > max_uv_y) max_uv_y = y + q;
				if (z + q > max_uv_z) max_uv_z = z + q;
			}
			max_uv_z *= 0.5;
			max_uv_y *= 0.5;
			max_len_sq = c.max_radius_squared();
			ey = int(max_uv_y * ysp + 1);
			ez = int(max_uv_z * zsp + 1);
			wy = ny + ey; wz = nz + ez;
			oy = ny + 2 * ey; oz = nz + 2 * ez;
			unit.assign(c.p, UnitVert());
			unit_max_order = 0;
			for (int v = 0; v < c.p; v++) {
				UnitVert &u = unit[v];
				u.x = c.X(v); u.y = c.Y(v); u.z = c.Z(v);
				u.nu = c.NU(v);
				if (u.nu > unit_max_order) unit_max_order = u.nu;
				if (u.nu > 8) return false;
				for (int i = 0; i < u.nu; i++) { u.ed[i] = (short)c.E(v, i); u.rel[i] = (signed char)c.R(v, i); }
			}
			// a finished cell holds no order-1 or order-2 vertex
			return c.np1 == 0 && c.np2 == 0 && !c.fatal;
		}
		return false;
	}
	long long nblocks() const { return (long long)nx * oy * oz; }
	bool primary(int b) const {
		const int j = (b / nx) % oy, k = b / (nx * oy);
		return j >= ey && j < wy && k >= ez && k < wz;
	}
};

// The image blocks of a periodic container, built by the reference's
// creation routines (src/container_prd.cc:583-775, src/container_prd.hh:
// 188-194) from the binned primary blocks: block b's particles are
// P[cstart[b] .. cstart[b+1]) with ids pid[...] for a primary block.  Image
// particles are appended to one flat list with their block; assemble() sorts
// them by block, stably, so each block keeps its order of creation.
struct ImageSet {
	const PrdGeom &g;
	const dbl4 *P;
	const int *pid;
	const int *cstart;
	std::vector<int> reg;                   // block of each image particle
	std::vector<dbl4> ip;                   // image particles, in creation order
	std::vector<int> iid;
	std::vector<unsigned char> img, created;
	bool late = false;   // a particle reached a block that was already complete
	ImageSet(const PrdGeom &g_, const dbl4 *P_, const int *pid_, const int *cstart_)
		: g(g_), P(P_), pid(pid_), cstart(cstart_), img(g_.nblocks(), 0), created(g_.nblocks(), 0) {}
	// Starts again with no image block over new primary blocks, keeping the
	// lists' storage (a periodic compute builds several sets).
	void reset(const dbl4 *P_, const int *pid_, const int *cstart_) {
		P = P_; pid = pid_; cstart = cstart_;
		reg.clear(); ip.clear(); iid.clear();
		std::fill(img.begin(), img.end(), (unsigned char)0);
		std::fill(created.begin(), created.end(), (unsigned char)0);
		late = false;
	}

	int co(int f) const { return cstart[f + 1] - cstart[f]; }
	double px(int f, int l) const { return P[cstart[f] + l].x; }
	double py(int f, int l) const { return P[cstart[f] + l].y; }

	// put_image, src/container_prd.cc:766-774 (the source is a primary block)
	void put_image(int r, int fijk, int l, double dx, double dy, double dz) {
		const dbl4 s = P[cstart[fijk] + l];
		if (created[r]) late = true;
		reg.push_back(r);
		ip.push_back(dbl4{s.x + dx, s.y + dy, s.z + dz, s.w});
		iid.push_back(pid[cstart[fijk] + l]);
	}
	// create_periodic_image, src/container_prd.hh:188-194
	void create(int di, int dj, int dk) {
		if (dk >= g.ez && dk < g.wz) {
			if (dj < g.ey || dj >= g.wy) side(di, dj, dk);
		} else vertical(di, dj, dk);
		created[di + g.nx * (dj + g.oy * dk)] = 1;
	}
	void create_block(int b) { create(b % g.nx, (b / g.nx) % g.oy, b / (g.nx * g.oy)); }

	// create_side_image, src/container_prd.cc:583-626
	void side(int di, int dj, int dk) {
		const int nx = g.nx, oy = g.oy, ny = g.ny, ey = g.ey;
		const double bx = g.bx, bxy = g.bxy, by = g.by, boxx = g.boxx, xsp = g.xsp;
		int l, dijk = di + nx * (dj + oy * dk), odijk, ima = prd_step_div(dj - ey, ny);
		int qua = di + prd_step_int(-ima * bxy * xsp), quadiv = prd_step_div(qua, nx);
		int fi = qua - quadiv * nx, fijk = fi + nx * (dj - ima * ny + oy * dk);
		double dis = ima * bxy + quadiv * bx, switchx = di * boxx - ima * bxy - quadiv * bx, adis;
		if ((img[dijk] & 1) == 0) {
			if (di > 0) { odijk = dijk - 1; adis = dis; }
			else { odijk = dijk + nx - 1; adis = dis + bx; }
			img[odijk] |= 2;
			for (l = 0; l < co(fijk); l++) {
				if (px(fijk, l) > switchx) put_image(dijk, fijk, l, dis, by * ima, 0);
				else put_image(odijk, fijk, l, adis, by * ima, 0);
			}
		}
		if ((img[dijk] & 2) == 0) {
			if (fi == nx - 1) { fijk += 1 - nx; switchx += (1 - nx) * boxx; dis += bx; }
			else { fijk++; switchx += boxx; }
			if (di == nx - 1) { odijk = dijk - nx + 1; adis = dis - bx; }
			else { odijk = dijk + 1; adis = dis; }
			img[odijk] |= 1;
			for (l = 0; l < co(fijk); l++) {
				if (px(fijk, l) < switchx) put_image(dijk, fijk, l, dis, by * ima, 0);
				else put_image(odijk, fijk, l, adis, by * ima, 0);
			}
		}
		img[dijk] = 3;
	}

	// create_vertical_image, src/container_prd.cc:635-757
	void vertical(int di, int dj, int dk) {
		const int nx = g.nx, oy = g.oy, ny = g.ny, nz = g.nz, ey = g.ey, ez = g.ez, wy = g.wy;
		const double bx = g.bx, bxy = g.bxy, by = g.by, bxz = g.bxz, byz = g.byz, bz = g.bz;
		const double boxx = g.boxx, boxy = g.boxy, xsp = g.xsp, ysp = g.ysp;
		int l, dijk = di + nx * (dj + oy * dk), dijkl, dijkr, ima = prd_step_div(dk - ez, nz);
		int qj = dj + prd_step_int(-ima * byz * ysp), qjdiv = prd_step_div(qj - ey, ny);
		int qi = di + prd_step_int((-ima * bxz - qjdiv * bxy) * xsp), qidiv = prd_step_div(qi, nx);
		int fi = qi - qidiv * nx, fj = qj - qjdiv * ny, fijk = fi + nx * (fj + oy * (dk - ima * nz)), fijk2;
		double disy = ima * byz + qjdiv * by, switchy = (dj - ey) * boxy - ima * byz - qjdiv * by;
		double disx = ima * bxz + qjdiv * bxy + qidiv * bx, switchx = di * boxx - ima * bxz - qjdiv * bxy - qidiv * bx;
		double switchx2, disxl, disxr, disx2, disxr2;

		if (di == 0) { dijkl = dijk + nx - 1; disxl = disx + bx; }
		else { dijkl = dijk - 1; disxl = disx; }
		if (di == nx - 1) { dijkr = dijk - nx + 1; disxr = disx - bx; }
		else { dijkr = dijk + 1; disxr = disx; }

		// down-left
		bool y_exist = dj != 0;
		if ((img[dijk] & 1) == 0) {
			img[dijkl] |= 2;
			if (y_exist) { img[dijkl - nx] |= 8; img[dijk - nx] |= 4; }
			for (l = 0; l < co(fijk); l++) {
				if (py(fijk, l) > switchy) {
					if (px(fijk, l) > switchx) put_image(dijk, fijk, l, disx, disy, bz * ima);
					else put_image(dijkl, fijk, l, disxl, disy, bz * ima);
				} else {
					if (!y_exist) continue;
					if (px(fijk, l) > switchx) put_image(dijk - nx, fijk, l, disx, disy, bz * ima);
					else put_image(dijkl - nx, fijk, l, disxl, disy, bz * ima);
				}
			}
		}
		// down-right
		if ((img[dijk] & 2) == 0) {
			if (fi == nx - 1) { fijk2 = fijk + 1 - nx; switchx2 = switchx + (1 - nx) * boxx; disx2 = disx + bx; disxr2 = disxr + bx; }
			else { fijk2 = fijk + 1; switchx2 = switchx + boxx; disx2 = disx; disxr2 = disxr; }
			img[dijkr] |= 1;
			if (y_exist) { img[dijkr - nx] |= 4; img[dijk - nx] |= 8; }
			for (l = 0; l < co(fijk2); l++) {
				if (py(fijk2, l) > switchy) {
					if (px(fijk2, l) > switchx2) put_image(dijkr, fijk2, l, disxr2, disy, bz * ima);
					else put_image(dijk, fijk2, l, disx2, disy, bz * ima);
				} else {
					if (!y_exist) continue;
					if (px(fijk2, l) > switchx2) put_image(dijkr - nx, fijk2, l, disxr2, disy, bz * ima);
					else put_image(dijk - nx, fijk2, l, disx2, disy, bz * ima);
				}
			}
		}
		// the primary row above, wrapping in y with its shift in x
		if (fj == wy - 1) {
			fijk += nx * (1 - ny) - fi;
			switchy += (1 - ny) * boxy;
			disy += by;
			qi = di + prd_step_int(-(ima * bxz + (qjdiv + 1) * bxy) * xsp);
			int dqidiv = prd_step_div(qi, nx) - qidiv;
			qidiv += dqidiv;
			fi = qi - qidiv * nx;
			fijk += fi;
			disx += bxy + bx * dqidiv;
			disxl += bxy + bx * dqidiv;
			disxr += bxy + bx * dqidiv;
			switchx -= bxy + bx * dqidiv;
		} else {
			fijk += nx;
			switchy += boxy;
		}
		// up-left
		y_exist = dj != oy - 1;
		if ((img[dijk] & 4) == 0) {
			img[dijkl] |= 8;
			if (y_exist) { img[dijkl + nx] |= 2; img[dijk + nx] |= 1; }
			for (l = 0; l < co(fijk); l++) {
				if (py(fijk, l) > switchy) {
					if (!y_exist) continue;
					if (px(fijk, l) > switchx) put_image(dijk + nx, fijk, l, disx, disy, bz * ima);
					else put_image(dijkl + nx, fijk, l, disxl, disy, bz * ima);
				} else {
					if (px(fijk, l) > switchx) put_image(dijk, fijk, l, disx, disy, bz * ima);
					else put_image(dijkl, fijk, l, disxl, disy, bz * ima);
				}
			}
		}
		// up-right
		if ((img[dijk] & 8) == 0) {
			if (fi == nx - 1) { fijk2 = fijk + 1 - nx; switchx2 = switchx + (1 - nx) * boxx; disx2 = disx + bx; disxr2 = disxr + bx; }
			else { fijk2 = fijk + 1; switchx2 = switchx + boxx; disx2 = disx; disxr2 = disxr; }
			img[dijkr] |= 4;
			if (y_exist) { img[dijkr + nx] |= 1; img[dijk + nx] |= 2; }
			for (l = 0; l < co(fijk2); l++) {
				if (py(fijk2, l) > switchy) {
					if (!y_exist) continue;
					if (px(fijk2, l) > switchx2) put_image(dijkr + nx, fijk2, l, disxr2, disy, bz * ima);
					else put_image(dijk + nx, fijk2, l, disx2, disy, bz * ima);
				} else {
					if (px(fijk2, l) > switchx2) put_image(dijkr, fijk2, l, disxr2, disy, bz * ima);
					else put_image(dijk, fijk2, l, disx2, disy, bz * ima);
				}
			}
		}
		img[dijk] = 15;
	}

	// Block-major arrays over all nx*oy*oz blocks: primary blocks as binned,
	// image blocks as built.  bstart has nblocks+1 entries.
	void assemble(std::vector<dbl4> &Pall, std::vector<int> &pidall, std::vector<int> &bstart) const {
		const int nb = (int)g.nblocks();
		std::vector<int> icnt(nb, 0);
		for (int r : reg) icnt[r]++;
		bstart.assign(nb + 1, 0);
		long long tot = 0;
		for (int b = 0; b < nb; b++) {
			bstart[b] = (int)tot;
			tot += g.primary(b) ? co(b) : icnt[b];
		}
		bstart[nb] = (int)tot;
		Pall.resize(std::max<long long>(tot, 1));
		pidall.resize(std::max<long long>(tot, 1));
		for (int b = 0; b < nb; b++) {
			if (!g.primary(b)) { icnt[b] = bstart[b]; continue; }
			int d = bstart[b];
			for (int l = 0; l < co(b); l++, d++) { Pall[d] = P[cstart[b] + l]; pidall[d] = pid[cstart[b] + l]; }
		}
		for (size_t t = 0; t < reg.size(); t++) {
			const int d = icnt[reg[t]]++;
			Pall[d] = ip[t];
			pidall[d] = iid[t];
		}
	}
};

// Image blocks within `margin` block layers of the primary domain, in the
// order of create_all_images (src/container_prd.cc:551-554: k, then j, then
// i).  The margins the reference allocates (ey, ez layers) can be far wider
// than what the cells reach; blocks beyond `margin` that a cell does reach
// are created once their first visit has been recorded.
enum { PRD_FIRST_MARGIN = 6 };   // block layers created before any visit is known
inline std::vector<int> prd_margin_images(const PrdGeom &g, int margin) {
	std::vector<int> o;
	const int j0 = std::max(0, g.ey - margin), j1 = std::min(g.oy, g.wy + margin);
	const int k0 = std::max(0, g.ez - margin), k1 = std::min(g.oz, g.wz + margin);
	for (int k = k0; k < k1; k++)
		for (int j = j0; j < j1; j++)
			for (int i = 0; i < g.nx; i++) {
				const int b = i + g.nx * (j + g.oy * k);
				if (!g.primary(b)) o.push_back(b);
			}
	return o;
}

// Image blocks that were visited, in order of their first visit (fv: per
// block, the smallest (cell << 24 | visit) key, ~0 if never visited).
inline std::vector<int> prd_visit_order(const PrdGeom &g, const unsigned long long *fv) {
	std::vector<int> o;
	for (int b = 0; b < (int)g.nblocks(); b++) if (fv[b] != ~0ull && !g.primary(b)) o.push_back(b);
	std::sort(o.begin(), o.end(), [&](int a, int b) { return fv[a] < fv[b]; });
	return o;
}

}  // namespace vgpu
