// Text formats of the containers, over the C ABI: print_custom (the
// per-cell line of voronoicell_base::output_custom), import, and the
// particle, cell and domain drawings.  Every number is printed with the
// reference's printf conversions, so the text is the reference's byte for
// byte (the engine's values being its values bit for bit).  Host code only:
// the cells come from the engine's stored results and cell tables.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/voro_gpu.h"

namespace {

// Everything one pass over the stored cells needs, in loop order.
struct Cells {
	long long n = 0;
	std::vector<int> id, exists, nfaces;
	std::vector<double> pos, vol, cent, rad;
	std::vector<long long> nbr_off;
	std::vector<int> nbr;
	// cell tables and statistics (with_geometry)
	bool geo = false;
	std::vector<long long> voff, eoff, fvoff;
	std::vector<double> vert, farea, fperim, normal, surface, edist, maxr2;
	std::vector<int> ord, edges, erel, forder, fverts, nedges;

	int load(vgpu_container *h, bool with_geometry) {
		int st;
		if ((st = vgpu_num_cells(h, &n))) return st;
		const size_t N = (size_t)n;
		id.resize(N); exists.resize(N); nfaces.resize(N); pos.resize(3 * N); vol.resize(N); cent.resize(3 * N);
		rad.resize(N); nbr_off.resize(N + 1);
		long long tot = 0;
		if ((st = vgpu_get_results(h, id.data(), nullptr, nullptr, exists.data(), pos.data(), vol.data(), cent.data(),
					   nfaces.data(), nbr_off.data(), nullptr, 0, &tot)))
			return st;
		nbr.resize((size_t)tot + 1);
		if ((st = vgpu_get_results(h, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
					   nbr.data(), tot, &tot)))
			return st;
		long long nb = 0;
		if ((st = vgpu_get_particles(h, &nb, nullptr, nullptr, rad.data()))) return st;
		geo = with_geometry;
		if (!geo) return VGPU_OK;
		long long nv, ne, nf, nfv;
		if ((st = vgpu_geometry_sizes(h, &nv, &ne, &nf, &nfv))) return st;
		voff.resize(N + 1); eoff.resize(N + 1); fvoff.resize(N + 1);
		vert.resize(3 * (size_t)nv + 3); ord.resize((size_t)nv + 1); edges.resize((size_t)ne + 1);
		erel.resize((size_t)ne + 1); farea.resize((size_t)nf + 1); fperim.resize((size_t)nf + 1);
		normal.resize(3 * (size_t)nf + 3); forder.resize((size_t)nf + 1); fverts.resize((size_t)nfv + 1);
		surface.resize(N); edist.resize(N); nedges.resize(N); maxr2.resize(N);
		return vgpu_get_geometry(h, voff.data(), vert.data(), ord.data(), eoff.data(), edges.data(), erel.data(),
					 farea.data(), fperim.data(), normal.data(), forder.data(), fvoff.data(), fverts.data(),
					 surface.data(), edist.data(), nedges.data(), maxr2.data());
	}
};

// voro_print_vector, src/common.cc:67-100
template<class T>
void print_vector(const T *v, int s, const char *f1, FILE *fp) {
	char f4[64], f3[64], f2[64], f4s[64];
	snprintf(f4s, sizeof f4s, "%s %s %s %s ", f1, f1, f1, f1);
	snprintf(f4, sizeof f4, "%s %s %s %s", f1, f1, f1, f1);
	snprintf(f3, sizeof f3, "%s %s %s", f1, f1, f1);
	snprintf(f2, sizeof f2, "%s %s", f1, f1);
	int k = 0;
	while (k + 4 < s) {
		fprintf(fp, f4s, v[k], v[k + 1], v[k + 2], v[k + 3]);
		k += 4;
	}
	if (k + 3 <= s) {
		if (k + 4 == s) fprintf(fp, f4, v[k], v[k + 1], v[k + 2], v[k + 3]);
		else fprintf(fp, f3, v[k], v[k + 1], v[k + 2]);
	} else if (s > 0) {
		if (k + 2 == s) fprintf(fp, f2, v[k], v[k + 1]);
		else fprintf(fp, f1, v[k]);
	}
}

// voro_print_face_vertices, src/common.cc:110-136
void print_face_vertices(const int *v, int s, FILE *fp) {
	int j, k = 0, l;
	bool first = true;
	while (k < s) {
		l = v[k++];
		const char *sp = first ? "" : " ";
		first = false;
		if (l <= 1) {
			if (l == 1) fprintf(fp, "%s(%d)", sp, v[k++]);
			else fprintf(fp, "%s()", sp);
		} else {
			j = k + l;
			fprintf(fp, "%s(%d", sp, v[k++]);
			while (k < j) fprintf(fp, ",%d", v[k++]);
			fputs(")", fp);
		}
	}
}

bool is_geometry_code(char c) { return c != 0 && strchr("wpPomgEeFAaftl", c) != nullptr; }

// The codes a format uses, read left to right as output_custom reads it
// (a '%' followed by anything is a code; "%%" is the code '%').
bool format_needs_geometry(const char *f) {
	for (; *f; f++) {
		if (*f != '%') continue;
		f++;
		if (*f == 0) break;
		if (is_geometry_code(*f)) return true;
	}
	return false;
}

// voronoicell_base::output_custom, src/cell.cc:2485-2558, for stored cell m.
void output_custom(const Cells &c, long long m, const char *format, FILE *fp) {
	const int i = c.id[m];
	const double x = c.pos[3 * m], y = c.pos[3 * m + 1], z = c.pos[3 * m + 2], r = c.rad[m];
	const long long f0 = c.nbr_off[m], nf = c.nbr_off[m + 1] - f0;
	const long long v0 = c.geo ? c.voff[m] : 0, nv = c.geo ? c.voff[m + 1] - v0 : 0;
	for (const char *fmp = format; *fmp; fmp++) {
		if (*fmp != '%') { putc(*fmp, fp); continue; }
		fmp++;
		switch (*fmp) {
		case 'i': fprintf(fp, "%d", i); break;
		case 'x': fprintf(fp, "%g", x); break;
		case 'y': fprintf(fp, "%g", y); break;
		case 'z': fprintf(fp, "%g", z); break;
		case 'q': fprintf(fp, "%g %g %g", x, y, z); break;
		case 'r': fprintf(fp, "%g", r); break;
		case 'w': fprintf(fp, "%d", (int)nv); break;
		case 'p':
			for (long long t = 0; t < nv; t++)
				fprintf(fp, t ? " (%g,%g,%g)" : "(%g,%g,%g)", c.vert[3 * (v0 + t)], c.vert[3 * (v0 + t) + 1],
					c.vert[3 * (v0 + t) + 2]);
			break;
		case 'P':
			for (long long t = 0; t < nv; t++)
				fprintf(fp, t ? " (%g,%g,%g)" : "(%g,%g,%g)", x + c.vert[3 * (v0 + t)], y + c.vert[3 * (v0 + t) + 1],
					z + c.vert[3 * (v0 + t) + 2]);
			break;
		case 'o':
			for (long long t = 0; t < nv; t++) fprintf(fp, t ? " %d" : "%d", c.ord[v0 + t]);
			break;
		case 'm': fprintf(fp, "%g", 0.25 * c.maxr2[m]); break;
		case 'g': fprintf(fp, "%d", c.nedges[m]); break;
		case 'E': fprintf(fp, "%g", c.edist[m]); break;
		case 'e': print_vector(c.fperim.data() + f0, (int)nf, "%g", fp); break;
		case 's': fprintf(fp, "%d", c.nfaces[m]); break;
		case 'F': fprintf(fp, "%g", c.surface[m]); break;
		case 'A': {
			// face_freq_table, src/cell.cc:2333-2354
			std::vector<int> v;
			for (long long t = 0; t < nf; t++) {
				const int q = c.forder[f0 + t];
				if ((size_t)q >= v.size()) v.resize(q + 1, 0);
				v[q]++;
			}
			print_vector(v.data(), (int)v.size(), "%d", fp);
		} break;
		case 'a': print_vector(c.forder.data() + f0, (int)nf, "%d", fp); break;
		case 'f': print_vector(c.farea.data() + f0, (int)nf, "%g", fp); break;
		case 't': print_face_vertices(c.fverts.data() + c.fvoff[m], (int)(c.fvoff[m + 1] - c.fvoff[m]), fp); break;
		case 'l':
			for (long long t = 0; t < nf; t++)
				fprintf(fp, t ? " (%g,%g,%g)" : "(%g,%g,%g)", c.normal[3 * (f0 + t)], c.normal[3 * (f0 + t) + 1],
					c.normal[3 * (f0 + t) + 2]);
			break;
		case 'n': print_vector(c.nbr.data() + f0, (int)nf, "%d", fp); break;
		case 'v': fprintf(fp, "%g", c.vol[m]); break;
		case 'c': fprintf(fp, "%g %g %g", c.cent[3 * m], c.cent[3 * m + 1], c.cent[3 * m + 2]); break;
		case 'C':
			fprintf(fp, "%g %g %g", x + c.cent[3 * m], y + c.cent[3 * m + 1], z + c.cent[3 * m + 2]);
			break;
		case 0: fmp--; break;
		default: putc('%', fp); putc(*fmp, fp);
		}
	}
	fputs("\n", fp);
}

// voronoicell_base::draw_gnuplot, src/cell.cc:1960-1979: every edge chain,
// marking edges as walked (on a copy of the cell's edge table).
void draw_gnuplot(const Cells &c, long long m, FILE *fp) {
	const double x = c.pos[3 * m], y = c.pos[3 * m + 1], z = c.pos[3 * m + 2];
	const long long v0 = c.voff[m], p = c.voff[m + 1] - v0, e0 = c.eoff[m];
	std::vector<int> eo((size_t)p + 1);
	for (long long t = 0; t < p; t++) eo[t + 1] = eo[t] + c.ord[v0 + t];
	std::vector<int> ed(c.edges.begin() + e0, c.edges.begin() + e0 + eo[p]);
	const int *rel = c.erel.data() + e0;
	const int *nu = c.ord.data() + v0;
	auto P = [&](int v, double &a, double &b, double &d) {
		a = x + c.vert[3 * (v0 + v)]; b = y + c.vert[3 * (v0 + v) + 1]; d = z + c.vert[3 * (v0 + v) + 2];
	};
	for (int i = 1; i < p; i++)
		for (int j = 0; j < nu[i]; j++) {
			int k = ed[eo[i] + j];
			if (k < 0) continue;
			double a, b, d;
			P(i, a, b, d);
			fprintf(fp, "%g %g %g\n", a, b, d);
			int l = i, mm = j;
			for (;;) {
				ed[eo[k] + rel[eo[l] + mm]] = -1 - l;
				ed[eo[l] + mm] = -1 - k;
				l = k;
				P(k, a, b, d);
				fprintf(fp, "%g %g %g\n", a, b, d);
				// search_edge, src/cell.cc:1981-1987
				bool found = false;
				for (mm = 0; mm < nu[l]; mm++) {
					k = ed[eo[l] + mm];
					if (k >= 0) { found = true; break; }
				}
				if (!found) break;
			}
			fputs("\n\n", fp);
		}
}

// voronoicell_base::draw_pov, src/cell.cc:1940-1955.
void draw_pov(const Cells &c, long long m, FILE *fp) {
	const double x = c.pos[3 * m], y = c.pos[3 * m + 1], z = c.pos[3 * m + 2];
	const long long v0 = c.voff[m], p = c.voff[m + 1] - v0, e0 = c.eoff[m];
	char b1[128], b2[128];
	long long e = e0;
	for (long long i = 0; i < p; i++) {
		snprintf(b1, sizeof b1, "%g,%g,%g", x + c.vert[3 * (v0 + i)], y + c.vert[3 * (v0 + i) + 1],
			 z + c.vert[3 * (v0 + i) + 2]);
		fprintf(fp, "sphere{<%s>,r}\n", b1);
		for (int j = 0; j < c.ord[v0 + i]; j++) {
			const int k = c.edges[e + j];
			if (k < i) {
				snprintf(b2, sizeof b2, "%g,%g,%g", x + c.vert[3 * (v0 + k)], y + c.vert[3 * (v0 + k) + 1],
					 z + c.vert[3 * (v0 + k) + 2]);
				if (strcmp(b1, b2) != 0) fprintf(fp, "cylinder{<%s>,<%s>,r}\n", b1, b2);
			}
		}
		e += c.ord[v0 + i];
	}
}

int compute(vgpu_container *h, bool geometry) {
	return vgpu_compute_all(h, VGPU_F_NEIGHBORS | VGPU_F_CSR | (geometry ? VGPU_F_GEOMETRY : 0u), nullptr);
}

}  // namespace

extern "C" {

int vgpu_print_custom(vgpu_container *h, const char *format, const char *filename) {
	if (!h || !format || !filename) return VGPU_INVALID_ARGUMENT;
	const bool geo = format_needs_geometry(format);
	int st;
	if ((st = compute(h, geo))) return st;
	Cells c;
	if ((st = c.load(h, geo))) return st;
	FILE *fp = fopen(filename, "w");
	if (!fp) return VGPU_FILE_ERROR;
	for (long long m = 0; m < c.n; m++)
		if (c.exists[m]) output_custom(c, m, format, fp);
	return fclose(fp) == 0 ? VGPU_OK : VGPU_FILE_ERROR;
}

int vgpu_import(vgpu_container *h, const char *filename, int radii, long long *n_read) {
	if (!h || !filename) return VGPU_INVALID_ARGUMENT;
	FILE *fp = fopen(filename, "r");
	if (!fp) return VGPU_FILE_ERROR;
	std::vector<int> ids;
	std::vector<double> xyz, r;
	int i, j;
	double x, y, z, rr;
	if (radii) {
		while ((j = fscanf(fp, "%d %lg %lg %lg %lg", &i, &x, &y, &z, &rr)) == 5) {
			ids.push_back(i); xyz.push_back(x); xyz.push_back(y); xyz.push_back(z); r.push_back(rr);
		}
	} else {
		while ((j = fscanf(fp, "%d %lg %lg %lg", &i, &x, &y, &z)) == 4) {
			ids.push_back(i); xyz.push_back(x); xyz.push_back(y); xyz.push_back(z);
		}
	}
	fclose(fp);
	if (n_read) *n_read = (long long)ids.size();
	int st = ids.empty() ? VGPU_OK
			      : vgpu_put(h, (int)ids.size(), ids.data(), xyz.data(), radii ? r.data() : nullptr);
	if (st) return st;
	// src/container.cc:337: anything but a clean end of file is an import error
	return j == EOF ? VGPU_OK : VGPU_FILE_ERROR;
}

int vgpu_draw_particles(vgpu_container *h, const char *filename, int pov) {
	if (!h || !filename) return VGPU_INVALID_ARGUMENT;
	int st;
	if ((st = vgpu_bin(h))) return st;
	long long n = 0;
	if ((st = vgpu_get_particles(h, &n, nullptr, nullptr, nullptr))) return st;
	std::vector<int> id((size_t)n + 1);
	std::vector<double> pos(3 * (size_t)n + 3), rad((size_t)n + 1);
	if ((st = vgpu_get_particles(h, &n, id.data(), pos.data(), rad.data()))) return st;
	int kind = 0;
	if ((st = vgpu_kind(h, &kind))) return st;
	const bool poly = kind == VGPU_CONTAINER_POLY || kind == VGPU_CONTAINER_PERIODIC_POLY;
	FILE *fp = fopen(filename, "w");
	if (!fp) return VGPU_FILE_ERROR;
	for (long long m = 0; m < n; m++) {
		const double x = pos[3 * m], y = pos[3 * m + 1], z = pos[3 * m + 2];
		// src/container.hh:329-335, 353-360 and container_poly's 551-557, 577-584
		if (pov) {
			if (poly) fprintf(fp, "// id %d\nsphere{<%g,%g,%g>,%g}\n", id[m], x, y, z, rad[m]);
			else fprintf(fp, "// id %d\nsphere{<%g,%g,%g>,s}\n", id[m], x, y, z);
		} else {
			if (poly) fprintf(fp, "%d %g %g %g %g\n", id[m], x, y, z, rad[m]);
			else fprintf(fp, "%d %g %g %g\n", id[m], x, y, z);
		}
	}
	return fclose(fp) == 0 ? VGPU_OK : VGPU_FILE_ERROR;
}

int vgpu_draw_cells(vgpu_container *h, const char *filename, int pov) {
	if (!h || !filename) return VGPU_INVALID_ARGUMENT;
	int st;
	if ((st = compute(h, true))) return st;
	Cells c;
	if ((st = c.load(h, true))) return st;
	FILE *fp = fopen(filename, "w");
	if (!fp) return VGPU_FILE_ERROR;
	// src/container.hh:379-385 and 405-412
	for (long long m = 0; m < c.n; m++) {
		if (!c.exists[m]) continue;
		if (pov) {
			fprintf(fp, "// cell %d\n", c.id[m]);
			draw_pov(c, m, fp);
		} else draw_gnuplot(c, m, fp);
	}
	return fclose(fp) == 0 ? VGPU_OK : VGPU_FILE_ERROR;
}

}  // extern "C"
