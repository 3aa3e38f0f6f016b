// A program written against the reference's interface (container, put,
// c_loop_all, voronoicell_neighbor, compute_cell), compiled against the
// drop-in header.  Prints: cells, sum of volumes, total faces, and a
// checksum of neighbour ids, for 1,000 seeded particles in the unit cube.
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

int main() {
	container con(0, 1, 0, 1, 0, 1, 6, 6, 6, false, false, false, 8);
	uint64_t s = mix(12345);
	for (int i = 0; i < 1000; i++) {
		double p[3];
		for (int a = 0; a < 3; a++) p[a] = (mix(s += 0x9E3779B97F4A7C15ull) >> 11) * (1.0 / 9007199254740992.0);
		con.put(i, p[0], p[1], p[2]);
	}
	c_loop_all vl(con);
	voronoicell_neighbor c(con);
	std::vector<int> nb;
	double vsum = 0;
	long long faces = 0, cells = 0;
	unsigned long long chk = 0;
	if (vl.start()) do {
		if (con.compute_cell(c, vl)) {
			vsum += c.volume();
			faces += c.number_of_faces();
			c.neighbors(nb);
			for (size_t t = 0; t < nb.size(); t++) chk = chk * 31u + (unsigned long long)(long long)(nb[t] + 7);
			cells++;
		}
	} while (vl.inc());
	printf("%lld %.17g %lld %llu\n", cells, vsum, faces, chk);
	return 0;
}
