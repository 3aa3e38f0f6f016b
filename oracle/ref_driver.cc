// Test-infrastructure driver around the unmodified Voro++ sources.
//
// This file is compiled together with the reference library's own source
// files (taken where they lie under the reference checkout, never copied into
// this repository) into oracle/_ref/libvororef.so.  It is a checker and a CPU
// baseline only: tests, the smoke check and the benchmark's CPU leg load it;
// the product library never does.
//
// It exposes a plain C interface so that Python (ctypes) can hand it seeded
// inputs and read back, in c_loop_all order, exactly what the GPU engine
// returns: particle id, block, existence flag, volume, centroid, face count
// and neighbour IDs.  The calls it makes are the reference's public API:
//   container / container_poly            src/container.hh:288-728
//   container_periodic / _poly            src/container_prd.hh:213-651
//   put()                                 src/container.cc:90-115
//   c_loop_all / c_loop_all_periodic      src/c_loops.hh:165-209,325-393
//   compute_cell(c,vl)                    src/container.hh:463-465
//   voronoicell_neighbor outputs          src/cell.cc:1584,1861,2168,2660
//   compute_all_cells()                   src/container.cc:438-443
#include <chrono>
#include <cstring>
#include <vector>

#include "voro++.hh"

using namespace voro;

extern "C" {

struct ref_spec {
	int kind;            // 0 container, 1 container_poly, 2 container_periodic, 3 container_periodic_poly
	double geom[6];      // rectangular: ax,bx,ay,by,az,bz ; periodic: bx,bxy,by,bxz,byz,bz
	int grid[3];         // nx,ny,nz
	int periodic[3];     // rectangular containers only
	int init_mem;
	int n_spheres;       // sphere walls (rectangular containers only)
	const double *spheres;  // n_spheres x {xc,yc,zc,rc,w_id}
};

struct ref_out {
	int *id;             // [n] particle id in loop order
	int *ijk;            // [n] block index
	int *q;              // [n] index in block
	int *exists;         // [n] compute_cell() return value
	double *vol;         // [n]
	double *cent;        // [3n] centroid relative to the particle
	int *nfaces;         // [n]
	long long *nbr_off;  // [n+1]
	int *nbr;            // [nbr_cap]
	long long nbr_cap;
	double *pos;         // [3n] stored (possibly remapped) position, or null
};

}

namespace {

template<class c_class, class c_loop>
long long run_loop(c_class &con, ref_out *o, double *secs) {
	voronoicell_neighbor c(con);
	c_loop vl(con);
	std::vector<int> v;
	long long m = 0, nn = 0;
	double sink = 0;
	auto t0 = std::chrono::steady_clock::now();
	if (vl.start()) do {
		bool ok = con.compute_cell(c, vl);
		if (!o && ok) {
			// Timing mode: produce the same outputs, keep only a checksum.
			double cx, cy, cz;
			c.centroid(cx, cy, cz);
			c.neighbors(v);
			sink += c.volume() + cx + c.number_of_faces() + (double)v.size();
		}
		if (o) {
			o->id[m] = vl.pid();
			o->ijk[m] = vl.ijk;
			o->q[m] = vl.q;
			o->exists[m] = ok ? 1 : 0;
			if (o->pos) {
				double x, y, z;
				vl.pos(x, y, z);
				o->pos[3 * m] = x; o->pos[3 * m + 1] = y; o->pos[3 * m + 2] = z;
			}
			o->nbr_off[m] = nn;
			if (ok) {
				o->vol[m] = c.volume();
				double cx, cy, cz;
				c.centroid(cx, cy, cz);
				o->cent[3 * m] = cx; o->cent[3 * m + 1] = cy; o->cent[3 * m + 2] = cz;
				o->nfaces[m] = c.number_of_faces();
				c.neighbors(v);
				for (size_t t = 0; t < v.size(); t++) {
					if (nn < o->nbr_cap) o->nbr[nn] = v[t];
					nn++;
				}
			} else {
				o->vol[m] = 0; o->cent[3 * m] = o->cent[3 * m + 1] = o->cent[3 * m + 2] = 0;
				o->nfaces[m] = 0;
			}
		}
		m++;
	} while (vl.inc());
	auto t1 = std::chrono::steady_clock::now();
	if (o) o->nbr_off[m] = nn;
	if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
	if (sink == -1.0) m = -m;  // keeps the timing-mode work observable
	return m;
}

template<class c_class>
void put_all(c_class &con, int n, const int *ids, const double *xyz, const double *) {
	for (int i = 0; i < n; i++) con.put(ids[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

template<class c_class>
void put_all_r(c_class &con, int n, const int *ids, const double *xyz, const double *r) {
	for (int i = 0; i < n; i++) con.put(ids[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], r[i]);
}

template<class c_class>
void add_spheres(c_class &con, const ref_spec *s, std::vector<wall_sphere*> &keep) {
	for (int w = 0; w < s->n_spheres; w++) {
		const double *q = s->spheres + 5 * w;
		wall_sphere *ws = new wall_sphere(q[0], q[1], q[2], q[3], (int)q[4]);
		keep.push_back(ws);
		con.add_wall(*ws);
	}
}

// mode 0: neighbour loop with outputs; mode 1: time compute_all_cells() only;
// mode 2: time the neighbour loop only (outputs discarded).
long long dispatch(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
		   ref_out *o, double *secs, int mode) {
	const double *g = s->geom;
	std::vector<wall_sphere*> keep;
	long long m = 0;
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		add_spheres(con, s, keep);
		put_all(con, n, ids, xyz, r);
		if (mode == 1) {
			auto t0 = std::chrono::steady_clock::now();
			con.compute_all_cells();
			auto t1 = std::chrono::steady_clock::now();
			*secs = std::chrono::duration<double>(t1 - t0).count();
			m = n;
		} else m = run_loop<container, c_loop_all>(con, mode == 0 ? o : 0, secs);
	} else if (s->kind == 1) {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		add_spheres(con, s, keep);
		put_all_r(con, n, ids, xyz, r);
		if (mode == 1) {
			auto t0 = std::chrono::steady_clock::now();
			con.compute_all_cells();
			auto t1 = std::chrono::steady_clock::now();
			*secs = std::chrono::duration<double>(t1 - t0).count();
			m = n;
		} else m = run_loop<container_poly, c_loop_all>(con, mode == 0 ? o : 0, secs);
	} else if (s->kind == 2) {
		container_periodic con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				       s->init_mem);
		put_all(con, n, ids, xyz, r);
		if (mode == 1) {
			auto t0 = std::chrono::steady_clock::now();
			con.compute_all_cells();
			auto t1 = std::chrono::steady_clock::now();
			*secs = std::chrono::duration<double>(t1 - t0).count();
			m = n;
		} else m = run_loop<container_periodic, c_loop_all_periodic>(con, mode == 0 ? o : 0, secs);
	} else {
		container_periodic_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1],
					    s->grid[2], s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		if (mode == 1) {
			auto t0 = std::chrono::steady_clock::now();
			con.compute_all_cells();
			auto t1 = std::chrono::steady_clock::now();
			*secs = std::chrono::duration<double>(t1 - t0).count();
			m = n;
		} else m = run_loop<container_periodic_poly, c_loop_all_periodic>(con, mode == 0 ? o : 0, secs);
	}
	for (size_t w = 0; w < keep.size(); w++) delete keep[w];
	return m;
}

}  // namespace

extern "C" {

// Runs the voronoicell_neighbor loop and fills `o` in c_loop_all order.
// Returns the number of loop entries (particles stored in the container).
long long ref_compute_neighbor(const ref_spec *s, int n, const int *ids, const double *xyz,
			       const double *r, ref_out *o, double *secs) {
	return dispatch(s, n, ids, xyz, r, o, secs, 0);
}

// Times container::compute_all_cells() (steady_clock around the call only, as
// examples/timing/timing_test.cc:51-57 does).  Returns seconds.
double ref_time_compute_all_cells(const ref_spec *s, int n, const int *ids, const double *xyz,
				  const double *r) {
	double secs = 0;
	dispatch(s, n, ids, xyz, r, 0, &secs, 1);
	return secs;
}

// container(_poly)::find_voronoi_cell (src/container.cc:237-297) for nq
// query points; found[i] is its return value.  With `secs`, the time of the
// query loop alone.  Rectangular kinds only (returns -1 otherwise).
long long ref_find_voronoi_cells(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
				 int nq, const double *q, int *found, int *pid, double *pos, double *secs) {
	const double *g = s->geom;
	if (s->kind > 1) return -1;
	long long nf = 0;
	std::chrono::steady_clock::time_point t0, t1;
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all(con, n, ids, xyz, r);
		t0 = std::chrono::steady_clock::now();
		for (int i = 0; i < nq; i++) {
			found[i] = con.find_voronoi_cell(q[3 * i], q[3 * i + 1], q[3 * i + 2], pos[3 * i], pos[3 * i + 1],
							 pos[3 * i + 2], pid[i]) ? 1 : 0;
			nf += found[i];
		}
		t1 = std::chrono::steady_clock::now();
	} else {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		t0 = std::chrono::steady_clock::now();
		for (int i = 0; i < nq; i++) {
			found[i] = con.find_voronoi_cell(q[3 * i], q[3 * i + 1], q[3 * i + 2], pos[3 * i], pos[3 * i + 1],
							 pos[3 * i + 2], pid[i]) ? 1 : 0;
			nf += found[i];
		}
		t1 = std::chrono::steady_clock::now();
	}
	if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
	return nf;
}

// container(_poly)::compute_ghost_cell (src/container.hh:488-498,710-721)
// for nq points: return value, volume, centroid, face count and up to
// max_nbr neighbour ids per point.  Rectangular kinds only (returns -1).
long long ref_compute_ghost_cells(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
				  int nq, const double *q, const double *qr, int *exists, double *vol, double *cent,
				  int *nfaces, int *nbr, int max_nbr) {
	const double *g = s->geom;
	if (s->kind > 1) return -1;
	std::vector<int> nb;
	long long m = 0;
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all(con, n, ids, xyz, r);
		voronoicell_neighbor c(con);
		for (int i = 0; i < nq; i++) {
			exists[i] = con.compute_ghost_cell(c, q[3 * i], q[3 * i + 1], q[3 * i + 2]) ? 1 : 0;
			vol[i] = 0; cent[3 * i] = cent[3 * i + 1] = cent[3 * i + 2] = 0; nfaces[i] = 0;
			if (!exists[i]) continue;
			m++;
			vol[i] = c.volume();
			c.centroid(cent[3 * i], cent[3 * i + 1], cent[3 * i + 2]);
			nfaces[i] = c.number_of_faces();
			c.neighbors(nb);
			for (int f = 0; f < (int)nb.size() && f < max_nbr; f++) nbr[(size_t)i * max_nbr + f] = nb[f];
		}
	} else {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		voronoicell_neighbor c(con);
		for (int i = 0; i < nq; i++) {
			exists[i] = con.compute_ghost_cell(c, q[3 * i], q[3 * i + 1], q[3 * i + 2], qr[i]) ? 1 : 0;
			vol[i] = 0; cent[3 * i] = cent[3 * i + 1] = cent[3 * i + 2] = 0; nfaces[i] = 0;
			if (!exists[i]) continue;
			m++;
			vol[i] = c.volume();
			c.centroid(cent[3 * i], cent[3 * i + 1], cent[3 * i + 2]);
			nfaces[i] = c.number_of_faces();
			c.neighbors(nb);
			for (int f = 0; f < (int)nb.size() && f < max_nbr; f++) nbr[(size_t)i * max_nbr + f] = nb[f];
		}
	}
	return m;
}

// container(_poly)::print_custom(format, filename) (src/container.cc:401-432).
// Rectangular kinds only (returns -1).
long long ref_print_custom(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
			   const char *format, const char *filename) {
	const double *g = s->geom;
	if (s->kind > 1) return -1;
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all(con, n, ids, xyz, r);
		con.print_custom(format, filename);
	} else {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		con.print_custom(format, filename);
	}
	return 0;
}

// Times the voronoicell_neighbor loop with volume/centroid/faces/neighbours.
double ref_time_neighbor_loop(const ref_spec *s, int n, const int *ids, const double *xyz,
			      const double *r) {
	double secs = 0;
	dispatch(s, n, ids, xyz, r, 0, &secs, 2);
	return secs;
}

}
