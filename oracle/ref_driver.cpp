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

// One force solve over given interaction lists with the reference's own code: its L2L
// (expansion<double>::operator<<, tigergrav/expansion.hpp:155-233), L2P (translate_L2,
// :236-269) and the eight gravity_* kernels, called in kick_fmm's order (src/tree.cpp:400-401,
// 481-491, 603-626).  Nodes are ordered with parents first; with stride > 1 only every stride-th
// node's / leaf's kernel work is done (a timing sample).  Returns the reference's flop count.
uint64_t tgref_solve(int64_t N, const int32_t* pos, int64_t nn, const int32_t* node_range, const int32_t* node_child,
		const int32_t* node_parent, const int32_t* node_x, const float* node_m, const int64_t* const* ptr,
		const int32_t* const* lst, int stride, cforce* out) {
	const bool ewald = options::get().ewald;
	std::vector<expansion<double>> L(nn);
	uint64_t flop = 0;
	auto gather_pos = [&](int k, int64_t b, int64_t e) {
		std::vector<vect<pos_type>> y;
		for (int64_t q = b; q < e; q += 2)
			for (int32_t p = lst[k][q]; p < lst[k][q + 1]; p++) y.push_back(to_x(pos + 3 * (int64_t) p));
		return y;
	};
	auto gather_multi = [&](int k, int64_t b, int64_t e) {
		std::vector<multi_src> y(e - b);
		for (int64_t q = b; q < e; q++) {
			const int32_t id = lst[k][q];
			for (int c = 0; c < 17; c++) y[q - b].m[c] = node_m[17 * (int64_t) id + c];
			y[q - b].x = to_x(node_x + 3 * (int64_t) id);
		}
		return y;
	};
	int64_t start = 0;
	while (start < nn) {
		int64_t end = start + 1;
		while (end < nn && node_parent[end] < start) end++;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : flop)
		for (int64_t n = start; n < end; n++) {
			if (node_range[2 * n] == node_range[2 * n + 1]) continue;
			const auto X = pos_to_double(to_x(node_x + 3 * n));
			expansion<double> Ln;
			if (node_parent[n] >= 0) {
				const auto Xp = pos_to_double(to_x(node_x + 3 * (int64_t) node_parent[n]));
				Ln = L[node_parent[n]] << (X - Xp);
			} else {
				Ln = 0.0;
			}
			if (stride <= 1 || n % stride == 0) {
				auto y0 = gather_multi(0, ptr[0][n], ptr[0][n + 1]);
				flop += gravity_CC_direct(Ln, to_x(node_x + 3 * n), y0);
				flop += gravity_CP_direct(Ln, to_x(node_x + 3 * n), gather_pos(1, ptr[1][n], ptr[1][n + 1]));
				if (ewald) {
					auto y2 = gather_multi(2, ptr[2][n], ptr[2][n + 1]);
					flop += gravity_CC_ewald(Ln, to_x(node_x + 3 * n), y2);
					flop += gravity_CP_ewald(Ln, to_x(node_x + 3 * n), gather_pos(3, ptr[3][n], ptr[3][n + 1]));
				}
			}
			L[n] = Ln;
		}
		start = end;
	}
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : flop)
	for (int64_t n = 0; n < nn; n++) {
		if (node_child[2 * n] >= 0) continue;
		const int32_t pb = node_range[2 * n], pe = node_range[2 * n + 1];
		if (pb == pe || (stride > 1 && n % stride)) continue;
		const auto X = pos_to_double(to_x(node_x + 3 * n));
		std::vector<vect<pos_type>> x = to_pos(pos + 3 * (int64_t) pb, pe - pb);
		std::vector<force> f(x.size());
		for (size_t j = 0; j < x.size(); j++) {
			force t = L[n].translate_L2(vect<float>(pos_to_double(x[j]) - X));
			f[j].phi = t.phi;
			f[j].g = t.g;
		}
		auto y4 = gather_multi(4, ptr[4][n], ptr[4][n + 1]);
		flop += gravity_PC_direct(f, x, y4);
		flop += gravity_PP_direct(f, x, gather_pos(5, ptr[5][n], ptr[5][n + 1]), true);
		if (ewald) {
			auto y6 = gather_multi(6, ptr[6][n], ptr[6][n + 1]);
			flop += gravity_PC_ewald(f, x, y6);
			flop += gravity_PP_ewald(f, x, gather_pos(7, ptr[7][n], ptr[7][n + 1]));
		}
		from_force(f, out + pb);
	}
	return flop;
}

}  // extern "C"
