// tg_groups.cpp -- the statistics of friends-of-friends groups (src/groups.cpp:129-228): centre,
// mean velocity, velocity dispersion, largest and median radius, from two passes over the
// particles in tree order with the reference's float / double steps.  Host code, built without
// contraction.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "tg_fof.hpp"

namespace tg {

namespace {

struct Acc {
	int N = 0;
	float rmax = 0.0f, rc = 0.0f;
	double x[3] = {0.0, 0.0, 0.0};
	float v[3] = {0.0f, 0.0f, 0.0f};
	float dv[3] = {0.0f, 0.0f, 0.0f};
	std::vector<float> radii;
};

}  // namespace

int group_catalog(const Tree& T, const std::vector<uint64_t>& group_tree, const float* v_input, double a, int32_t min_n, tg_group* out,
		int64_t capacity, int64_t* n_out) {
	const int64_t n = int64_t(T.order.size());
	const double ainv = 1.0 / a;
	const float ainvf = float(ainv);
	std::map<uint64_t, Acc> groups;
	auto vel = [&](int64_t i, int d) { return v_input[3 * T.order[size_t(i)] + d]; };
	// groups_add_particle1 (src/groups.cpp:129-157)
	for (int64_t i = 0; i < n; i++) {
		const uint64_t id = group_tree[size_t(i)];
		if (id == TG_FOF_NONE) continue;
		Acc& g = groups[id];
		if (g.N == 0) {
			for (int d = 0; d < 3; d++) g.x[d] = pos_to_double(T.pos[3 * size_t(i) + d]);
			g.N++;
		} else {
			double dx[3];
			for (int d = 0; d < 3; d++) dx[d] = pos_to_double(T.pos[3 * size_t(i) + d]) - g.x[d] / double(g.N);
			for (int d = 0; d < 3; d++) {
				const double absdx = std::abs(dx[d]);
				dx[d] = std::copysign(std::min(absdx, 1.0 - absdx), dx[d] * (0.5 - absdx));
			}
			double step[3];
			for (int d = 0; d < 3; d++) step[d] = dx[d] + g.x[d] / double(g.N);
			for (int d = 0; d < 3; d++) g.x[d] += step[d];
			g.N++;
			for (int d = 0; d < 3; d++) {
				while (g.x[d] < 0.0) g.x[d] += double(g.N);
				while (g.x[d] > double(g.N)) g.x[d] -= double(g.N);
			}
		}
		for (int d = 0; d < 3; d++) g.v[d] += vel(i, d) * ainvf;
	}
	// groups_finish1 (src/groups.cpp:159-177)
	for (auto& e : groups) {
		Acc& g = e.second;
		for (int d = 0; d < 3; d++) {
			g.x[d] /= double(g.N);
			g.v[d] /= float(g.N);
		}
	}
	// groups_add_particle2 (src/groups.cpp:179-198)
	for (int64_t i = 0; i < n; i++) {
		const uint64_t id = group_tree[size_t(i)];
		if (id == TG_FOF_NONE) continue;
		Acc& g = groups[id];
		for (int d = 0; d < 3; d++) {
			const float dv = vel(i, d) * ainvf - g.v[d];
			g.dv[d] += dv * dv;
		}
		double dx[3];
		for (int d = 0; d < 3; d++) {
			dx[d] = pos_to_double(T.pos[3 * size_t(i) + d]) - g.x[d];
			dx[d] = std::min(std::abs(dx[d]), 1.0 - std::abs(dx[d]));
		}
		double dot = dx[0] * dx[0];
		dot += dx[1] * dx[1];
		dot += dx[2] * dx[2];
		const double r = a * std::sqrt(dot);
		g.rmax = std::max(g.rmax, float(r));
		g.radii.push_back(float(r));
	}
	// groups_finish2 (src/groups.cpp:200-228)
	int64_t k = 0;
	for (auto& e : groups) {
		Acc& g = e.second;
		for (int d = 0; d < 3; d++) {
			g.dv[d] /= float(g.N);
			g.dv[d] = std::sqrt(g.dv[d]);
		}
		std::sort(g.radii.begin(), g.radii.end());
		g.rc = g.radii[g.radii.size() / 2];
		if (g.N < min_n) continue;
		if (out && k < capacity) {
			tg_group& o = out[k];
			o.id = e.first;
			o.N = g.N;
			o.rmax = g.rmax;
			o.rc = g.rc;
			for (int d = 0; d < 3; d++) {
				o.x[d] = g.x[d];
				o.v[d] = g.v[d];
				o.dv[d] = g.dv[d];
			}
		}
		k++;
	}
	*n_out = k;
	return out && k > capacity ? TG_EINVAL : TG_OK;
}

}  // namespace tg
