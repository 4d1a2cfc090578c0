// gravity_hip.cpp -- drop-in replacement for tigergrav's src/gravity.cpp.
//
// A tigergrav build that compiles this file instead of src/gravity.cpp (and links
// libtigergrav_amd.so) runs every interaction kernel on the MI355X: the eight functions keep the
// signatures of tigergrav/gravity.hpp:39-51 and forward to the C ABI of include/tg_gravity.h.
// The reference types are layout-compatible with the C structs (force = {double phi; vect<double>
// g}, multi_src = {multipole<float> m; vect<pos_type> x}, expansion<double> = 35 doubles,
// vect<pos_type> = 3 x int32), so the vectors are passed without copies.
//
// Kernel parameters are read from options::get() on every call (the reference freezes them in
// function-local statics at the first call, src/gravity.cpp:176-187; reading them each time
// gives the same values in a normal run).  Errors abort, as the reference's own error paths do.
#include <tigergrav/gravity.hpp>
#include <tigergrav/options.hpp>

#include <cstdio>
#include <cstdlib>

#include "tg_gravity.h"

static_assert(sizeof(force) == sizeof(tg_force), "force layout");
static_assert(sizeof(multi_src) == sizeof(tg_multi_src), "multi_src layout");
static_assert(sizeof(vect<pos_type>) == 3 * sizeof(int32_t), "position layout");
static_assert(sizeof(expansion<double>) == 35 * sizeof(double), "expansion layout");

namespace {
tg_params params() {
	const auto& o = options::get();
	tg_params p;
	p.ewald = o.ewald ? 1 : 0;
	p.soft_len = o.soft_len;
	p.m = o.m_tot / o.problem_size;
	p.device = 0;
	return p;
}
inline const int32_t* pos(const std::vector<vect<pos_type>>& v) { return reinterpret_cast<const int32_t*>(v.data()); }
inline const int32_t* pos(const vect<pos_type>& v) { return reinterpret_cast<const int32_t*>(&v); }
inline std::uint64_t check(int rc, std::uint64_t flop, const char* what) {
	if (rc != TG_OK) {
		std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, tg_last_error());
		std::abort();
	}
	return flop;
}
}  // namespace

std::uint64_t gravity_PP_direct(std::vector<force>& f, const std::vector<vect<pos_type>>& x, std::vector<vect<pos_type>> y,
		bool do_phi) {
	const tg_params p = params();
	std::uint64_t flop = 0;
	int rc = tg_gravity_PP_direct(&p, reinterpret_cast<tg_force*>(f.data()), pos(x), int64_t(x.size()), pos(y),
			int64_t(y.size()), do_phi ? 1 : 0, &flop);
	return check(rc, flop, "tg_gravity_PP_direct");
}

std::uint64_t gravity_PC_direct(std::vector<force>& f, const std::vector<vect<pos_type>>& x, std::vector<multi_src>& y) {
	const tg_params p = params();
	std::uint64_t flop = 0;
	int rc = tg_gravity_PC_direct(&p, reinterpret_cast<tg_force*>(f.data()), pos(x), int64_t(x.size()),
			reinterpret_cast<const tg_multi_src*>(y.data()), int64_t(y.size()), &flop);
	return check(rc, flop, "tg_gravity_PC_direct");
}

std::uint64_t gravity_CC_direct(expansion<double>& L, const vect<pos_type>& x, std::vector<multi_src>& y) {
	const tg_params p = params();
	std::uint64_t flop = 0;
	int rc = tg_gravity_CC_direct(&p, L.data(), pos(x), reinterpret_cast<const tg_multi_src*>(y.data()), int64_t(y.size()),
			&flop);
	return check(rc, flop, "tg_gravity_CC_direct");
}

std::uint64_t gravity_CP_direct(expansion<double>& L, const vect<pos_type>& x, std::vector<vect<pos_type>> y) {
	const tg_params p = params();
	std::uint64_t flop = 0;
	int rc = tg_gravity_CP_direct(&p, L.data(), pos(x), pos(y), int64_t(y.size()), &flop);
	return check(rc, flop, "tg_gravity_CP_direct");
}

std::uint64_t gravity_PP_ewald(std::vector<force>& f, const std::vector<vect<pos_type>>& x, std::vector<vect<pos_type>> y) {
	const tg_params p = params();
	std::uint64_t flop = 0;
	int rc = tg_gravity_PP_ewald(&p, reinterpret_cast<tg_force*>(f.data()), pos(x), int64_t(x.size()), pos(y),
			int64_t(y.size()), &flop);
	return check(rc, flop, "tg_gravity_PP_ewald");
}

std::uint64_t gravity_PC_ewald(std::vector<force>& f, const std::vector<vect<pos_type>>& x, std::vector<multi_src>& y) {
	const tg_params p = params();
	std::uint64_t flop = 0;
	int rc = tg_gravity_PC_ewald(&p, reinterpret_cast<tg_force*>(f.data()), pos(x), int64_t(x.size()),
			reinterpret_cast<const tg_multi_src*>(y.data()), int64_t(y.size()), &flop);
	return check(rc, flop, "tg_gravity_PC_ewald");
}

std::uint64_t gravity_CC_ewald(expansion<double>& L, const vect<pos_type>& x, std::vector<multi_src>& y) {
	const tg_params p = params();
	std::uint64_t flop = 0;
	int rc = tg_gravity_CC_ewald(&p, L.data(), pos(x), reinterpret_cast<const tg_multi_src*>(y.data()), int64_t(y.size()),
			&flop);
	return check(rc, flop, "tg_gravity_CC_ewald");
}

std::uint64_t gravity_CP_ewald(expansion<double>& L, const vect<pos_type>& x, std::vector<vect<pos_type>> y) {
	const tg_params p = params();
	std::uint64_t flop = 0;
	int rc = tg_gravity_CP_ewald(&p, L.data(), pos(x), pos(y), int64_t(y.size()), &flop);
	return check(rc, flop, "tg_gravity_CP_ewald");
}

double ewald_near_separation(const vect<double> x) {
	const double d[3] = {x[0], x[1], x[2]};
	return tg_ewald_near_separation(d);
}

double ewald_far_separation(const vect<double> x, double r) {
	const double d[3] = {x[0], x[1], x[2]};
	return tg_ewald_far_separation(d, r);
}

void init_ewald() {}
