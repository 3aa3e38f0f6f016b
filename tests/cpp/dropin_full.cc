// A program written against the reference's interface, compiled against the
// drop-in header: walls (sphere, plane, cylinder, cone) added through the
// wall base class, put through a particle_order, c_loop_all with the cell
// outputs and polyhedron statistics, c_loop_subset, c_loop_order,
// print_custom, draw_cells_gnuplot and a pre_container.  Prints one summary
// line per part; the GPU test compares it with the same quantities the
// reference library gives.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "voro_amd.hh"

using namespace voro_amd;

static uint64_t mix(uint64_t z) {
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

int main(int argc, char **argv) {
	const char *dir = argc > 1 ? argv[1] : ".";
	container con(0, 10, 0, 10, 0, 10, 6, 6, 6, false, false, false, 8);
	wall_sphere ws(5, 5, 5, 4.9, -5);
	wall_plane wp(0.2, 0.1, 1, 8.7, -6);
	wall_cylinder wc(5, 5, 5, 1, 0.3, 0.2, 4.4, -7);
	wall_cone wk(5, 5, -1, 0.1, 0.2, 1, 0.55, -8);
	con.add_wall(ws); con.add_wall(wp); con.add_wall(wc); con.add_wall(wk);
	particle_order po;
	uint64_t s = mix(777);
	for (int i = 0; i < 3000; i++) {
		double p[3];
		for (int a = 0; a < 3; a++) p[a] = 10.0 * ((mix(s += 0x9E3779B97F4A7C15ull) >> 11) * (1.0 / 9007199254740992.0));
		if (i % 3 == 0) con.put(po, i, p[0], p[1], p[2]);
		else con.put(i, p[0], p[1], p[2]);
	}
	// c_loop_all: outputs and statistics
	{
		c_loop_all vl(con);
		voronoicell_neighbor c(con);
		std::vector<int> nb, fo, fv, vo;
		std::vector<double> fa, nr, vt;
		double vsum = 0, area = 0, area_f = 0, edges = 0;
		long long cells = 0, faces = 0, fvsum = 0, vorders = 0;
		unsigned long long chk = 0;
		if (vl.start()) do {
			if (con.compute_cell(c, vl)) {
				vsum += c.volume();
				faces += c.number_of_faces();
				c.neighbors(nb);
				for (int t : nb) chk = chk * 31u + (unsigned long long)(long long)(t + 101);
				c.face_areas(fa);
				for (double a : fa) area_f += a;
				area += c.surface_area();
				c.face_orders(fo);
				c.face_vertices(fv);
				for (int t : fv) fvsum += t;
				c.vertex_orders(vo);
				for (int t : vo) vorders += t;
				c.normals(nr);
				c.vertices(vt);
				edges += c.total_edge_distance();
				cells++;
			}
		} while (vl.inc());
		printf("all %lld %.17g %lld %llu %.17g %.17g %lld %lld %.17g\n", cells, vsum, faces, chk, area, area_f, fvsum,
		       vorders, edges);
	}
	// c_loop_subset: a sphere with bounds test
	{
		c_loop_subset vl(con);
		vl.setup_sphere(4.0, 6.0, 5.0, 2.2, true);
		long long n = 0, idsum = 0, blk = 0;
		if (vl.start()) do { n++; idsum += vl.pid(); blk += vl.i + 7 * vl.j + 49 * vl.k; } while (vl.inc());
		printf("subset %lld %lld %lld\n", n, idsum, blk);
	}
	// c_loop_order: the particles put through po, in put order
	{
		c_loop_order vl(con, po);
		long long n = 0;
		unsigned long long chk = 0;
		if (vl.start()) do { n++; chk = chk * 131u + (unsigned long long)vl.pid(); } while (vl.inc());
		printf("order %lld %llu\n", n, chk);
	}
	std::string f1 = std::string(dir) + "/custom.txt", f2 = std::string(dir) + "/cells.gnu";
	con.print_custom("%i %q %v %s %n %f %t %l %P %A", f1.c_str());
	con.draw_cells_gnuplot(f2.c_str());
	// pre_container: guess the grid, then put
	{
		pre_container pc(0, 10, 0, 10, 0, 10, false, false, false);
		uint64_t s2 = mix(4242);
		for (int i = 0; i < 1000; i++) {
			double p[3];
			for (int a = 0; a < 3; a++) p[a] = 12.0 * ((mix(s2 += 0x9E3779B97F4A7C15ull) >> 11) * (1.0 / 9007199254740992.0)) - 1;
			pc.put(i, p[0], p[1], p[2]);
		}
		int nx, ny, nz;
		pc.guess_optimal(nx, ny, nz);
		container c2(0, 10, 0, 10, 0, 10, nx, ny, nz, false, false, false, 8);
		pc.setup(c2);
		printf("pre %d %d %d %d %.17g\n", pc.total_particles(), nx, ny, nz, c2.sum_cell_volumes());
	}
	return 0;
}
