// voro_amd.hh -- header-only C++ drop-in for the reference's container
// interface, over the C ABI of voro_gpu.h (link with -lvoro_gpu).
//
// Mirrors the reference classes a program uses:
//   container / container_poly          src/container.hh:288-728
//   container_periodic / _poly          src/container_prd.hh:213-651
//   wall_sphere / _plane / _cylinder / _cone, add_wall   src/wall.hh:27-115
//   c_loop_all / c_loop_all_periodic    src/c_loops.hh:165-209, 325-393
//   c_loop_subset                       src/c_loops.hh:218-258
//   particle_order / c_loop_order       src/c_loops.hh:37-69, 268-317
//   pre_container / pre_container_poly  src/pre_container.hh
//   voronoicell_neighbor / voronoicell  src/cell.hh:130-549 (outputs and
//                                       polyhedron statistics)
//   print_custom, import, draw_*        src/container.cc:333-432, container.hh:325-450
// A reference program that does
//     container con(...); con.put(i,x,y,z); ...
//     c_loop_all vl(con); voronoicell_neighbor c(con);
//     if(vl.start()) do if(con.compute_cell(c,vl)) { c.volume(); c.neighbors(v); c.face_areas(f); ... } while(vl.inc());
// compiles unchanged against this header (namespace voro_amd instead of
// voro).  The first compute_cell (or compute_all_cells) computes every cell
// on the GPU; later compute_cell calls read the stored cell.  The first
// polyhedron statistic asked for (face_areas(), vertices(), ...) computes
// the container again with the cell tables kept (VGPU_F_GEOMETRY).  Errors
// the reference reports with exit() are thrown as voro_amd::error here.
// Walls are the four reference classes; a user-defined wall class (the
// reference's virtual wall interface) is not supported.
#ifndef VORO_AMD_HH
#define VORO_AMD_HH

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "voro_gpu.h"

namespace voro_amd {

struct error : std::runtime_error {
	int status;
	error(int s, const char *what)
		: std::runtime_error(std::string(what) + ": " + vgpu_status_string(s)), status(s) {}
};

inline void check(int s, const char *what) {
	if (s != VGPU_OK) throw error(s, what);
}

// ---- walls (src/wall.hh:27-115) ----
struct wall {
	int type;
	double a[7];
	int w_id;
    protected:
	wall(int t, int id, double a0, double a1, double a2, double a3, double a4 = 0, double a5 = 0, double a6 = 0)
		: type(t), a{a0, a1, a2, a3, a4, a5, a6}, w_id(id) {}
};
struct wall_sphere : wall {
	wall_sphere(double xc, double yc, double zc, double rc, int w_id_ = -99) : wall(0, w_id_, xc, yc, zc, rc) {}
};
struct wall_plane : wall {
	wall_plane(double xc, double yc, double zc, double ac, int w_id_ = -99) : wall(1, w_id_, xc, yc, zc, ac) {}
};
struct wall_cylinder : wall {
	wall_cylinder(double xc, double yc, double zc, double xa, double ya, double za, double rc, int w_id_ = -99)
		: wall(2, w_id_, xc, yc, zc, xa, ya, za, rc) {}
};
struct wall_cone : wall {
	wall_cone(double xc, double yc, double zc, double xa, double ya, double za, double ang, int w_id_ = -99)
		: wall(3, w_id_, xc, yc, zc, xa, ya, za, ang) {}
};

// ---- particle_order (src/c_loops.hh:37-69): the put calls it saw ----
class particle_order {
    public:
	std::vector<long long> seq;     // put-sequence numbers of the particles put through it
	const void *owner = nullptr;
};

class voronoicell_neighbor;
class c_loop_all;
class c_loop_subset;
class c_loop_order;

class container_base_gpu {
    public:
	~container_base_gpu() { if (h) vgpu_destroy(h); }
	container_base_gpu(const container_base_gpu &) = delete;
	container_base_gpu &operator=(const container_base_gpu &) = delete;

	// add_wall, src/container.hh:63-69 (walls cut cells in the order added)
	void add_wall(wall &w) {
		check(vgpu_add_wall(h, w.type, w.a, w.w_id), "add_wall");
		dirty = true;
	}
	// container_base::point_inside, src/container.cc:484-487
	bool point_inside(double x, double y, double z) {
		int in = 0;
		check(vgpu_point_inside(h, x, y, z, &in), "point_inside");
		return in != 0;
	}
	void clear() { check(vgpu_clear(h), "clear"); nput = 0; dirty = true; }
	// Computes every cell and keeps the results (src/container.cc:438-443).
	void compute_all_cells() { compute(false); }
	double sum_cell_volumes() {
		ensure();
		double v = 0;
		check(vgpu_sum_volumes(h, &v), "sum_cell_volumes");
		return v;
	}
	template<class v_cell, class c_loop> bool compute_cell(v_cell &c, c_loop &vl);
	vgpu_container *handle() { return h; }

	// text formats, byte for byte the reference's
	void print_custom(const char *format, const char *filename) {
		check(vgpu_print_custom(h, format, filename), "print_custom");
		fetch(); dirty = false;
	}
	void import(const char *filename) {
		long long n = 0;
		int st = vgpu_import(h, filename, radii(), &n);
		nput += n; dirty = true;
		check(st, "import");
	}
	void draw_particles(const char *filename) { check(vgpu_draw_particles(h, filename, 0), "draw_particles"); }
	void draw_particles_pov(const char *filename) { check(vgpu_draw_particles(h, filename, 1), "draw_particles_pov"); }
	void draw_cells_gnuplot(const char *filename) {
		check(vgpu_draw_cells(h, filename, 0), "draw_cells_gnuplot");
		fetch(); dirty = false;
	}
	void draw_cells_pov(const char *filename) {
		check(vgpu_draw_cells(h, filename, 1), "draw_cells_pov");
		fetch(); dirty = false;
	}

	// loop-order results
	std::vector<int> id, ijk, q, exists, nfaces, nbr;
	std::vector<double> pos, vol, cent, rad;
	std::vector<long long> nbr_off;
	// polyhedron statistics (filled on first use)
	bool have_geo = false;
	std::vector<long long> voff, eoff, fvoff;
	std::vector<double> gvert, farea, fperim, normal, surface, edist, maxr2;
	std::vector<int> gord, gedges, gerel, forder, fverts, nedges;

    protected:
	container_base_gpu(int kind, const double *geom, const int *grid, const int *per)
		: gnx(grid[0]), gny(grid[1]), rect(kind <= VGPU_CONTAINER_POLY) {
		check(vgpu_create(kind, geom, grid, per, 8, &h), "container");
	}
	int gnx, gny;
	bool rect;
	int radii() const {
		int k = 0;
		vgpu_kind(h, &k);
		return k == VGPU_CONTAINER_POLY || k == VGPU_CONTAINER_PERIODIC_POLY;
	}
	void compute(bool geometry) {
		check(vgpu_compute_all(h, VGPU_F_NEIGHBORS | VGPU_F_CSR | (geometry ? VGPU_F_GEOMETRY : 0u), nullptr),
		      "compute_all_cells");
		fetch();
		dirty = false;
		if (geometry) fetch_geometry();
	}
	void ensure() { if (dirty) compute_all_cells(); }
	void ensure_geometry() {
		if (dirty || !have_geo) compute(true);
	}
	void fetch() {
		long long n = 0;
		have_geo = false;
		check(vgpu_num_cells(h, &n), "num_cells");
		id.resize(n); ijk.resize(n); q.resize(n); exists.resize(n); nfaces.resize(n);
		pos.resize(3 * n); vol.resize(n); cent.resize(3 * n); nbr_off.resize(n + 1); rad.resize(n);
		long long tot = 0;
		check(vgpu_get_results(h, id.data(), ijk.data(), q.data(), exists.data(), pos.data(), vol.data(),
				       cent.data(), nfaces.data(), nbr_off.data(), nullptr, 0, &tot), "results");
		nbr.resize(tot > 0 ? tot : 1);
		check(vgpu_get_results(h, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
				       nullptr, nbr.data(), tot, &tot), "results");
		long long m = 0;
		check(vgpu_get_particles(h, &m, nullptr, nullptr, rad.data()), "particles");
	}
	void fetch_geometry() {
		long long nv, ne, nf, nfv, n = (long long)id.size();
		check(vgpu_geometry_sizes(h, &nv, &ne, &nf, &nfv), "geometry");
		voff.resize(n + 1); eoff.resize(n + 1); fvoff.resize(n + 1);
		gvert.resize(3 * nv + 3); gord.resize(nv + 1); gedges.resize(ne + 1); gerel.resize(ne + 1);
		farea.resize(nf + 1); fperim.resize(nf + 1); normal.resize(3 * nf + 3); forder.resize(nf + 1);
		fverts.resize(nfv + 1); surface.resize(n); edist.resize(n); nedges.resize(n); maxr2.resize(n);
		check(vgpu_get_geometry(h, voff.data(), gvert.data(), gord.data(), eoff.data(), gedges.data(), gerel.data(),
					farea.data(), fperim.data(), normal.data(), forder.data(), fvoff.data(), fverts.data(),
					surface.data(), edist.data(), nedges.data(), maxr2.data()), "geometry");
		have_geo = true;
	}
	void put_one(int n, double x, double y, double z, const double *r, particle_order *po) {
		double p[3] = {x, y, z};
		check(vgpu_put(h, 1, &n, p, r), "put");
		if (po) { po->owner = this; po->seq.push_back(nput); }
		nput++;
		dirty = true;
	}
	bool find(double x, double y, double z, double &rx, double &ry, double &rz, int &pid) {
		double xyz[3] = {x, y, z}, r[3];
		int found = 0, id_ = 0;
		check(vgpu_find_voronoi_cells(h, 1, xyz, &found, &id_, r, 0), "find_voronoi_cell");
		if (!found) return false;
		rx = r[0]; ry = r[1]; rz = r[2]; pid = id_;
		return true;
	}
	vgpu_container *h = nullptr;
	bool dirty = true;
	long long nput = 0;
	friend class c_loop_all;
	friend class c_loop_subset;
	friend class c_loop_order;
	friend class voronoicell_neighbor;
};

struct geom6 { double v[6]; geom6(double a, double b, double c, double d, double e, double f) : v{a, b, c, d, e, f} {} };
struct grid3 { int v[3]; grid3(int a, int b, int c) : v{a, b, c} {} };

// container(ax,bx,ay,by,az,bz,nx,ny,nz,xp,yp,zp,init_mem), src/container.hh:290-296
class container : public container_base_gpu {
    public:
	container(double ax, double bx, double ay, double by, double az, double bz, int nx, int ny, int nz, bool xp,
		  bool yp, bool zp, int init_mem = 8)
		: container_base_gpu(VGPU_CONTAINER, geom6(ax, bx, ay, by, az, bz).v, grid3(nx, ny, nz).v,
				     grid3(xp, yp, zp).v) { (void)init_mem; }
	void put(int n, double x, double y, double z) { put_one(n, x, y, z, nullptr, nullptr); }
	void put(particle_order &vo, int n, double x, double y, double z) { put_one(n, x, y, z, nullptr, &vo); }
	// find_voronoi_cell, src/container.cc:237-262 (batches: vgpu_find_voronoi_cells)
	bool find_voronoi_cell(double x, double y, double z, double &rx, double &ry, double &rz, int &pid) {
		return find(x, y, z, rx, ry, rz, pid);
	}
};

// container_poly, src/container.hh:510-520
class container_poly : public container_base_gpu {
    public:
	container_poly(double ax, double bx, double ay, double by, double az, double bz, int nx, int ny, int nz,
		       bool xp, bool yp, bool zp, int init_mem = 8)
		: container_base_gpu(VGPU_CONTAINER_POLY, geom6(ax, bx, ay, by, az, bz).v, grid3(nx, ny, nz).v,
				     grid3(xp, yp, zp).v) { (void)init_mem; }
	void put(int n, double x, double y, double z, double r) { put_one(n, x, y, z, &r, nullptr); }
	void put(particle_order &vo, int n, double x, double y, double z, double r) { put_one(n, x, y, z, &r, &vo); }
	bool find_voronoi_cell(double x, double y, double z, double &rx, double &ry, double &rz, int &pid) {
		return find(x, y, z, rx, ry, rz, pid);
	}
};

// container_periodic(bx,bxy,by,bxz,byz,bz,nx,ny,nz,init_mem), src/container_prd.hh:213-217
class container_periodic : public container_base_gpu {
    public:
	container_periodic(double bx, double bxy, double by, double bxz, double byz, double bz, int nx, int ny, int nz,
			   int init_mem = 8)
		: container_base_gpu(VGPU_CONTAINER_PERIODIC, geom6(bx, bxy, by, bxz, byz, bz).v, grid3(nx, ny, nz).v,
				     nullptr) { (void)init_mem; }
	void put(int n, double x, double y, double z) { put_one(n, x, y, z, nullptr, nullptr); }
	void put(particle_order &vo, int n, double x, double y, double z) { put_one(n, x, y, z, nullptr, &vo); }
	// find_voronoi_cell, src/container_prd.cc:303-326
	bool find_voronoi_cell(double x, double y, double z, double &rx, double &ry, double &rz, int &pid) {
		return find(x, y, z, rx, ry, rz, pid);
	}
};

// container_periodic_poly, src/container_prd.hh:434-441
class container_periodic_poly : public container_base_gpu {
    public:
	container_periodic_poly(double bx, double bxy, double by, double bxz, double byz, double bz, int nx, int ny,
				int nz, int init_mem = 8)
		: container_base_gpu(VGPU_CONTAINER_PERIODIC_POLY, geom6(bx, bxy, by, bxz, byz, bz).v, grid3(nx, ny, nz).v,
				     nullptr) { (void)init_mem; }
	void put(int n, double x, double y, double z, double r) { put_one(n, x, y, z, &r, nullptr); }
	void put(particle_order &vo, int n, double x, double y, double z, double r) { put_one(n, x, y, z, &r, &vo); }
	// find_voronoi_cell, src/container_prd.cc:338-360
	bool find_voronoi_cell(double x, double y, double z, double &rx, double &ry, double &rz, int &pid) {
		return find(x, y, z, rx, ry, rz, pid);
	}
};

// Shared part of the loops: the loop position m (an index into the
// container's loop-order results) and the reference's c_loop_base readers
// (src/c_loops.hh:130-158).
class c_loop_base_gpu {
    public:
	int ijk = 0, q = 0;
	int pid() { return con.id[m]; }
	void pos(double &x, double &y, double &z) { x = con.pos[3 * m]; y = con.pos[3 * m + 1]; z = con.pos[3 * m + 2]; }
	void pos(int &pid_, double &x, double &y, double &z, double &r) {
		pid_ = con.id[m]; pos(x, y, z); r = con.rad[m];
	}
	double x() { return con.pos[3 * m]; }
	double y() { return con.pos[3 * m + 1]; }
	double z() { return con.pos[3 * m + 2]; }
	long long index() const { return m; }

    protected:
	explicit c_loop_base_gpu(container_base_gpu &c) : con(c) {}
	void at(long long m_) { m = m_; ijk = con.ijk[m]; q = con.q[m]; }
	container_base_gpu &con;
	long long m = 0;
};

// c_loop_all, src/c_loops.hh:165-209: blocks by ascending index, then
// storage order.
class c_loop_all : public c_loop_base_gpu {
    public:
	int i = 0, j = 0, k = 0;
	template<class c_class>
	explicit c_loop_all(c_class &con_) : c_loop_base_gpu(con_) {}
	bool start() {
		con.ensure();
		if (con.id.empty()) return false;
		set(0);
		return true;
	}
	bool inc() {
		if (m + 1 >= (long long)con.id.size()) return false;
		set(m + 1);
		return true;
	}

    private:
	void set(long long m_) {
		at(m_);
		// block coordinates of block ijk (rectangular containers; a periodic
		// container's blocks are numbered with its image rows)
		if (con.rect) { i = ijk % con.gnx; j = (ijk / con.gnx) % con.gny; k = ijk / (con.gnx * con.gny); }
	}
};

// c_loop_all_periodic, src/c_loops.hh:325-393: the primary domain's blocks
// by ascending index, then storage order -- the order the engine stores a
// periodic container's cells in.
typedef c_loop_all c_loop_all_periodic;

// c_loop_subset, src/c_loops.hh:218-258: the reference's loop over the
// blocks (and, with bounds_test, the particles) of a sphere, a box or a
// block range, step for step (vgpu_loop_subset).  compute_cell at a subset
// position gives the cell c_loop_all computes there.
class c_loop_subset : public c_loop_base_gpu {
    public:
	int i = 0, j = 0, k = 0;
	template<class c_class>
	explicit c_loop_subset(c_class &con_) : c_loop_base_gpu(con_) {}
	void setup_sphere(double vx, double vy, double vz, double r, bool bounds_test = true) {
		double v[4] = {vx, vy, vz, r};
		run(0, v, nullptr, bounds_test);
	}
	void setup_box(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
		       bool bounds_test = true) {
		double v[6] = {xmin, xmax, ymin, ymax, zmin, zmax};
		run(1, v, nullptr, bounds_test);
	}
	void setup_intbox(int ai, int bi, int aj, int bj, int ak, int bk) {
		int iv[6] = {ai, bi, aj, bj, ak, bk};
		run(2, nullptr, iv, false);
	}
	bool start() {
		con.ensure();
		t = 0;
		if (idx.empty()) return false;
		set();
		return true;
	}
	bool inc() {
		if (t + 1 >= idx.size()) return false;
		t++;
		set();
		return true;
	}

    private:
	void run(int mode, const double *v, const int *iv, bool bt) {
		long long n = 0;
		check(vgpu_loop_subset(con.h, mode, v, iv, bt ? 1 : 0, nullptr, nullptr, 0, &n), "c_loop_subset");
		idx.resize(n > 0 ? n : 1); blk.resize(3 * (n > 0 ? n : 1));
		check(vgpu_loop_subset(con.h, mode, v, iv, bt ? 1 : 0, idx.data(), blk.data(), n, &n), "c_loop_subset");
		idx.resize(n); blk.resize(3 * n);
		t = 0;
	}
	void set() { at(idx[t]); i = blk[3 * t]; j = blk[3 * t + 1]; k = blk[3 * t + 2]; }
	std::vector<long long> idx;
	std::vector<int> blk;
	size_t t = 0;
};

// c_loop_order, src/c_loops.hh:268-317: the stored particles a
// particle_order recorded, in record order (particles the container dropped
// are not stored, so not visited).
class c_loop_order : public c_loop_base_gpu {
    public:
	template<class c_class>
	c_loop_order(c_class &con_, particle_order &vo_) : c_loop_base_gpu(con_), vo(vo_) {}
	bool start() {
		con.ensure();
		long long n = (long long)con.id.size();
		std::vector<int> put(n > 0 ? n : 1);
		check(vgpu_get_put_index(con.h, put.data()), "c_loop_order");
		long long mx = 0;
		for (long long t = 0; t < n; t++) mx = put[t] > mx ? put[t] : mx;
		std::vector<long long> where(mx + 1, -1);
		for (long long t = 0; t < n; t++) where[put[t]] = t;
		list.clear();
		for (long long s : vo.seq) if (s <= mx && where[s] >= 0) list.push_back(where[s]);
		t = 0;
		if (list.empty()) return false;
		at(list[0]);
		return true;
	}
	bool inc() {
		if (t + 1 >= list.size()) return false;
		at(list[++t]);
		return true;
	}

    private:
	particle_order &vo;
	std::vector<long long> list;
	size_t t = 0;
};

// voronoicell_neighbor: the outputs and polyhedron statistics of the cell at
// a loop position (src/cell.hh:130-549).  The statistics need the cell
// tables: the first one asked for computes the container again with them.
class voronoicell_neighbor {
    public:
	int p = 0;    // vertex count (after a statistic was asked for)
	voronoicell_neighbor() {}
	template<class c_class> explicit voronoicell_neighbor(c_class &) {}
	double volume() { return con->vol[m]; }
	void centroid(double &x, double &y, double &z) { x = con->cent[3 * m]; y = con->cent[3 * m + 1]; z = con->cent[3 * m + 2]; }
	int number_of_faces() { return con->nfaces[m]; }
	void neighbors(std::vector<int> &v) { v.assign(con->nbr.begin() + con->nbr_off[m], con->nbr.begin() + con->nbr_off[m + 1]); }
	// src/cell.cc:1791-2358
	void face_areas(std::vector<double> &v) { faces(con->farea, v); }
	void face_perimeters(std::vector<double> &v) { faces(con->fperim, v); }
	void face_orders(std::vector<int> &v) { faces(con->forder, v); }
	void normals(std::vector<double> &v) {
		geo();
		v.assign(con->normal.begin() + 3 * con->nbr_off[m], con->normal.begin() + 3 * con->nbr_off[m + 1]);
	}
	void face_vertices(std::vector<int> &v) {
		geo();
		v.assign(con->fverts.begin() + con->fvoff[m], con->fverts.begin() + con->fvoff[m + 1]);
	}
	void face_freq_table(std::vector<int> &v) {
		std::vector<int> o;
		face_orders(o);
		v.clear();
		for (int q : o) { if ((size_t)q >= v.size()) v.resize(q + 1, 0); v[q]++; }
	}
	void vertex_orders(std::vector<int> &v) {
		geo();
		v.assign(con->gord.begin() + con->voff[m], con->gord.begin() + con->voff[m + 1]);
	}
	void vertices(std::vector<double> &v) {
		geo();
		v.assign(con->gvert.begin() + 3 * con->voff[m], con->gvert.begin() + 3 * con->voff[m + 1]);
	}
	void vertices(double x, double y, double z, std::vector<double> &v) {
		vertices(v);
		for (size_t t = 0; t < v.size(); t += 3) { v[t] = x + v[t]; v[t + 1] = y + v[t + 1]; v[t + 2] = z + v[t + 2]; }
	}
	double surface_area() { geo(); return con->surface[m]; }
	double total_edge_distance() { geo(); return con->edist[m]; }
	int number_of_edges() { geo(); return con->nedges[m]; }
	double max_radius_squared() { geo(); return con->maxr2[m]; }

    private:
	void geo() {
		if (!con->have_geo) con->ensure_geometry();
		p = (int)(con->voff[m + 1] - con->voff[m]);
	}
	template<class T>
	void faces(const std::vector<T> &a, std::vector<T> &v) {
		geo();
		v.assign(a.begin() + con->nbr_off[m], a.begin() + con->nbr_off[m + 1]);
	}
	container_base_gpu *con = nullptr;
	long long m = 0;
	friend class container_base_gpu;
};
typedef voronoicell_neighbor voronoicell;

template<class v_cell, class c_loop>
bool container_base_gpu::compute_cell(v_cell &c, c_loop &vl) {
	ensure();
	c.con = this;
	c.m = vl.index();
	if (have_geo) c.p = (int)(voff[c.m + 1] - voff[c.m]);
	return exists[c.m] != 0;
}

// pre_container / pre_container_poly (src/pre_container.hh,
// src/pre_container.cc): particles kept before the container exists, the
// block grid guessed from them, then put in put order.
template<bool POLY>
class pre_container_t {
    public:
	pre_container_t(double ax_, double bx_, double ay_, double by_, double az_, double bz_, bool xp, bool yp, bool zp)
		: ax(ax_), bx(bx_), ay(ay_), by(by_), az(az_), bz(bz_), xperiodic(xp), yperiodic(yp), zperiodic(zp) {}
	// src/pre_container.cc:71-79 (particles outside a non-periodic range are not kept)
	void put(int n, double x, double y, double z, double r = 0.5) {
		if ((xperiodic || (x >= ax && x <= bx)) && (yperiodic || (y >= ay && y <= by)) &&
		    (zperiodic || (z >= az && z <= bz))) {
			ids.push_back(n); xyz.push_back(x); xyz.push_back(y); xyz.push_back(z); rr.push_back(r);
		}
	}
	int total_particles() const { return (int)ids.size(); }
	// src/pre_container.cc:57-63
	void guess_optimal(int &nx, int &ny, int &nz) {
		const double dx = bx - ax, dy = by - ay, dz = bz - az;
		const double ilscale = std::pow(total_particles() / (5.6 * dx * dy * dz), 1 / 3.0);
		nx = int(dx * ilscale + 1); ny = int(dy * ilscale + 1); nz = int(dz * ilscale + 1);
	}
	template<class c_class>
	void setup(c_class &con) {
		for (size_t t = 0; t < ids.size(); t++) put_into(con, t, nullptr);
	}
	template<class c_class>
	void setup(particle_order &vo, c_class &con) {
		for (size_t t = 0; t < ids.size(); t++) put_into(con, t, &vo);
	}

    private:
	template<class c_class>
	void put_into(c_class &con, size_t t, particle_order *vo) {
		if constexpr (POLY) {
			if (vo) con.put(*vo, ids[t], xyz[3 * t], xyz[3 * t + 1], xyz[3 * t + 2], rr[t]);
			else con.put(ids[t], xyz[3 * t], xyz[3 * t + 1], xyz[3 * t + 2], rr[t]);
		} else {
			if (vo) con.put(*vo, ids[t], xyz[3 * t], xyz[3 * t + 1], xyz[3 * t + 2]);
			else con.put(ids[t], xyz[3 * t], xyz[3 * t + 1], xyz[3 * t + 2]);
		}
	}
	double ax, bx, ay, by, az, bz;
	bool xperiodic, yperiodic, zperiodic;
	std::vector<int> ids;
	std::vector<double> xyz, rr;
};
typedef pre_container_t<false> pre_container;
typedef pre_container_t<true> pre_container_poly;

}  // namespace voro_amd

#endif
