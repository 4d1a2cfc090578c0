// ref_driver.cpp -- TEST INFRASTRUCTURE (oracle).  Not product code.
//
// Links the reference's own compiled kernels (src/gravity.cpp, built from where it lies under
// the reference tree by oracle/Makefile) and exposes them through a C ABI so the test suite can
// (a) generate golden vectors and (b) time the reference CPU path on the same interaction lists
// as the GPU path.
//
// The reference reads its kernel parameters from the global `options` singleton
// (tigergrav/options.hpp:217-219, defined in src/options.cpp:18-26 next to the Boost/HPX command
// line parser).  This driver defines that singleton and its two accessors itself, then sets the
// four fields the kernels read (ewald, soft_len, m_tot, problem_size -- src/gravity.cpp:176-187)
// once before the first kernel call, as the reference's main() does.
#include <tigergrav/gravity.hpp>
#include <tigergrav/options.hpp>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

options options::global;
options& options::get() { return global; }
void options::set(options o) { global = o; }

namespace {
struct cforce {
	double phi;
	double g[3];
};
struct cmulti {
	float m[17];
	int32_t x[3];
};
static_assert(sizeof(force) == sizeof(cforce), "force layout");
static_assert(sizeof(multi_src) == sizeof(cmulti), "multi_src layout");

std::vector<vect<pos_type>> to_pos(const int32_t* p, int64_t n) {
	std::vector<vect<pos_type>> v(n);
	for (int64_t i = 0; i < n; i++)
		for (int d = 0; d < 3; d++) v[i][d] = p[3 * i + d];
	return v;
}
std::vector<force> to_force(const cforce* f, int64_t n) {
	std::vector<force> v(n);
	for (int64_t i = 0; i < n; i++) {
		v[i].phi = f[i].phi;
		for (int d = 0; d < 3; d++) v[i].g[d] = f[i].g[d];
	}
	return v;
}
void from_force(const std::vector<force>& v, cforce* f) {
	for (size_t i = 0; i < v.size(); i++) {
		f[i].phi = v[i].phi;
		for (int d = 0; d < 3; d++) f[i].g[d] = v[i].g[d];
	}
}
std::vector<multi_src> to_multi(const cmulti* m, int64_t n) {
	std::vector<multi_src> v(n);
	for (int64_t i = 0; i < n; i++) {
		for (int q = 0; q < 17; q++) v[i].m[q] = m[i].m[q];
		for (int d = 0; d < 3; d++) v[i].x[d] = m[i].x[d];
	}
	return v;
}
vect<pos_type> to_x(const int32_t* p) {
	vect<pos_type> x;
	for (int d = 0; d < 3; d++) x[d] = p[d];
	return x;
}
}  // namespace

extern "C" {

int tgref_set_options(int ewald, double soft_len, double m_tot, uint64_t problem_size, double theta) {
	options o;
	o.ewald = ewald;
	o.soft_len = soft_len;
	o.m_tot = m_tot;
	o.problem_size = problem_size;
	o.theta = theta;
	o.parts_per_node = 64;
	options::set(o);
	return 0;
}

uint64_t tgref_PP_direct(cforce* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj, int do_phi) {
	auto F = to_force(f, ns);
	auto rc = gravity_PP_direct(F, to_pos(x, ns), to_pos(y, nj), do_phi != 0);
	from_force(F, f);
	return rc;
}
uint64_t tgref_PC_direct(cforce* f, const int32_t* x, int64_t ns, const cmulti* y, int64_t nc) {
	auto F = to_force(f, ns);
	auto Y = to_multi(y, nc);
	auto rc = gravity_PC_direct(F, to_pos(x, ns), Y);
	from_force(F, f);
	return rc;
}
uint64_t tgref_CC_direct(double* L, const int32_t* x, const cmulti* y, int64_t nc) {
	expansion<double> E;
	for (int i = 0; i < LP; i++) E[i] = L[i];
	auto Y = to_multi(y, nc);
	auto rc = gravity_CC_direct(E, to_x(x), Y);
	for (int i = 0; i < LP; i++) L[i] = E[i];
	return rc;
}
uint64_t tgref_CP_direct(double* L, const int32_t* x, const int32_t* y, int64_t nj) {
	expansion<double> E;
	for (int i = 0; i < LP; i++) E[i] = L[i];
	auto rc = gravity_CP_direct(E, to_x(x), to_pos(y, nj));
	for (int i = 0; i < LP; i++) L[i] = E[i];
	return rc;
}
uint64_t tgref_PP_ewald(cforce* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj) {
	auto F = to_force(f, ns);
	auto rc = gravity_PP_ewald(F, to_pos(x, ns), to_pos(y, nj));
	from_force(F, f);
	return rc;
}
uint64_t tgref_PC_ewald(cforce* f, const int32_t* x, int64_t ns, const cmulti* y, int64_t nc) {
	auto F = to_force(f, ns);
	auto Y = to_multi(y, nc);
	auto rc = gravity_PC_ewald(F, to_pos(x, ns), Y);
	from_force(F, f);
	return rc;
}
uint64_t tgref_CC_ewald(double* L, const int32_t* x, const cmulti* y, int64_t nc) {
	expansion<double> E;
	for (int i = 0; i < LP; i++) E[i] = L[i];
	auto Y = to_multi(y, nc);
	auto rc = gravity_CC_ewald(E, to_x(x), Y);
	for (int i = 0; i < LP; i++) L[i] = E[i];
	return rc;
}
uint64_t tgref_CP_ewald(double* L, const int32_t* x, const int32_t* y, int64_t nj) {
	expansion<double> E;
	for (int i = 0; i < LP; i++) E[i] = L[i];
	auto rc = gravity_CP_ewald(E, to_x(x), to_pos(y, nj));
	for (int i = 0; i < LP; i++) L[i] = E[i];
	return rc;
}

}  // extern "C"
