// tg_load.cpp -- the N-GenIC / Gadget-1 initial-condition reader (src/load.cpp:13-96): the
// 256-byte header and the type-1 positions and velocities, each block framed by Fortran record
// markers.  Host code.
//
// Positions are shifted by half the mean particle spacing, 0.5 N^(-1/3) (N-GenIC puts its grid at
// the cell corners), wrapped into the unit box and converted to fixed point; velocities are scaled
// by a^(3/2), a = 1 / (1 + z).  Every conversion keeps the reference's float / double steps.
#include "tg_load.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace tg {
void set_capi_message(const char* msg);  // tg_capi.cpp
int32_t double_to_pos(double x);         // tg_tree.cpp
}  // namespace tg

static_assert(sizeof(tg_ngenic_header) == 256, "N-GenIC header is 256 bytes");

namespace {

struct File {
	FILE* fp = nullptr;
	explicit File(const char* path) : fp(std::fopen(path, "rb")) {}
	~File() {
		if (fp) std::fclose(fp);
	}
};

bool read_marked(FILE* fp, void* dst, size_t bytes) {
	int32_t marker = 0;
	if (std::fread(&marker, 4, 1, fp) != 1) return false;
	if (bytes && std::fread(dst, 1, bytes, fp) != bytes) return false;
	return std::fread(&marker, 4, 1, fp) == 1;
}

int fail(const std::string& msg) {
	tg::set_capi_message(msg.c_str());
	return TG_EINVAL;
}

}  // namespace

extern "C" {

int tg_ngenic_read_header(const char* path, tg_ngenic_header* h) {
	if (!path || !h) return fail("tg_ngenic_read_header: null argument");
	File f(path);
	if (!f.fp) return fail(std::string("unable to open ") + path);
	if (!read_marked(f.fp, h, sizeof(*h))) return fail(std::string("truncated header in ") + path);
	return TG_OK;
}

int tg_ngenic_load(const char* path, int64_t capacity, int32_t* x, float* v, int64_t* n) {
	if (!path || !n) return fail("tg_ngenic_load: null argument");
	File f(path);
	if (!f.fp) return fail(std::string("unable to open ") + path);
	tg_ngenic_header h;
	if (!read_marked(f.fp, &h, sizeof(h))) return fail(std::string("truncated header in ") + path);
	const int64_t np = int64_t(h.npart[1]);
	*n = np;
	if (!x && !v) return TG_OK;  // size query
	if (np > capacity || !x || !v) return fail("tg_ngenic_load: arrays too small for the file's particles");
	int32_t marker = 0;
	if (std::fread(&marker, 4, 1, f.fp) != 1) return fail("truncated position block");
	// src/load.cpp:55-76: x += sep in double, back to float; wrapped while above 1
	const double sep = 0.5 * std::pow(double(h.npart[1]), -1.0 / 3.0);
	for (int64_t i = 0; i < np; i++) {
		float p[3];
		if (std::fread(p, 4, 3, f.fp) != 3) return fail("truncated position block");
		for (int d = 0; d < 3; d++) {
			float c = p[d];
			c += sep;
			while (c > 1.0) c -= 1.0;
			x[3 * i + d] = tg::double_to_pos(c);
		}
	}
	if (std::fread(&marker, 4, 1, f.fp) != 1 || std::fread(&marker, 4, 1, f.fp) != 1) return fail("truncated velocity block");
	// src/load.cpp:80-88: v * (1 / (1 + z))^1.5 in double, stored as float
	const double c0 = 1.0 / (1.0 + h.redshift);
	const double s = std::pow(c0, 1.5);
	for (int64_t i = 0; i < np; i++) {
		float w[3];
		if (std::fread(w, 4, 3, f.fp) != 3) return fail("truncated velocity block");
		for (int d = 0; d < 3; d++) v[3 * i + d] = float(w[d] * s);
	}
	return TG_OK;
}

}  // extern "C"
