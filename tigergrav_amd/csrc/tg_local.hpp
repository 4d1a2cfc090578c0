// tg_local.hpp -- local-expansion translation (L2L) and evaluation (L2P) in FP64.
//
// The exact 4th-order Taylor re-expansion of tigergrav/expansion.hpp:185-233 and evaluation of
// tigergrav/expansion.hpp:236-269.  Slots carry the symmetric-tensor weights mult(s)/|s|!
// (the reference's expansion_factors), so with W_s(d) = mult(s)/|s|! d^s:
//   L2L:  L'_t = sum_{|s| <= 4-|t|} L_{t u s} W_s(d)
//   L2P:  phi  = sum_s L_s W_s(d),   g_a = -sum_{|s| <= 3} L_{s u a} W_s(d)
#pragma once
#include "tg_math.hpp"

namespace tg {

struct Slot {
	int ord;
	int i[4];
};
constexpr Slot SL[35] = {{0, {0, 0, 0, 0}}, {1, {0}}, {1, {1}}, {1, {2}}, {2, {0, 0}}, {2, {0, 1}}, {2, {0, 2}},
		{2, {1, 1}}, {2, {1, 2}}, {2, {2, 2}}, {3, {0, 0, 0}}, {3, {0, 0, 1}}, {3, {0, 0, 2}}, {3, {0, 1, 1}},
		{3, {0, 1, 2}}, {3, {0, 2, 2}}, {3, {1, 1, 1}}, {3, {1, 1, 2}}, {3, {1, 2, 2}}, {3, {2, 2, 2}},
		{4, {0, 0, 0, 0}}, {4, {0, 0, 0, 1}}, {4, {0, 0, 0, 2}}, {4, {0, 0, 1, 1}}, {4, {0, 0, 1, 2}},
		{4, {0, 0, 2, 2}}, {4, {0, 1, 1, 1}}, {4, {0, 1, 1, 2}}, {4, {0, 1, 2, 2}}, {4, {0, 2, 2, 2}},
		{4, {1, 1, 1, 1}}, {4, {1, 1, 1, 2}}, {4, {1, 1, 2, 2}}, {4, {1, 2, 2, 2}}, {4, {2, 2, 2, 2}}};

// slot of the multiset union of slots t and s (requires ord(t) + ord(s) <= 4)
__host__ __device__ constexpr int count_idx(const Slot& a, int v) {
	return (a.ord > 0 && a.i[0] == v) + (a.ord > 1 && a.i[1] == v) + (a.ord > 2 && a.i[2] == v) + (a.ord > 3 && a.i[3] == v);
}
__host__ __device__ constexpr int slot_of_counts(int n0, int n1, int n2) {
	// sorted tuple with n0 zeros, n1 ones, n2 twos
	return (n0 + n1 + n2) == 0 ? 0
	     : (n0 + n1 + n2) == 1 ? 1 + (n0 ? 0 : (n1 ? 1 : 2))
	     : (n0 + n1 + n2) == 2 ? 4 + s2(n0 ? 0 : (n1 ? 1 : 2), n2 ? 2 : (n1 ? 1 : 0))
	     : (n0 + n1 + n2) == 3 ? 10 + s3sorted(n0 ? 0 : (n1 ? 1 : 2), n0 >= 2 ? 0 : (n0 + n1 >= 2 ? 1 : 2), n2 ? 2 : (n1 ? 1 : 0))
	     : 20 + s4sorted(n0 ? 0 : (n1 ? 1 : 2), n0 >= 2 ? 0 : (n0 + n1 >= 2 ? 1 : 2), n0 >= 3 ? 0 : (n0 + n1 >= 3 ? 1 : 2),
	                     n2 ? 2 : (n1 ? 1 : 0));
}
__host__ __device__ constexpr int merge_slot(int t, int s) {
	return slot_of_counts(count_idx(SL[t], 0) + count_idx(SL[s], 0), count_idx(SL[t], 1) + count_idx(SL[s], 1),
			count_idx(SL[t], 2) + count_idx(SL[s], 2));
}
__host__ __device__ constexpr double fact(int n) { return n <= 1 ? 1.0 : n * fact(n - 1); }
__host__ __device__ constexpr double slot_weight(int s) {
	// mult(s)/|s|! = 1 / prod(count_v!)
	return 1.0 / (fact(count_idx(SL[s], 0)) * fact(count_idx(SL[s], 1)) * fact(count_idx(SL[s], 2)));
}

__device__ __forceinline__ void taylor_weights(double* W, const double* d) {
#pragma unroll
	for (int s = 0; s < 35; s++) {
		double w = slot_weight(s);
#pragma unroll
		for (int k = 0; k < 4; k++)
			if (k < SL[s].ord) w *= d[SL[s].i[k]];
		W[s] = w;
	}
}

// L (35, FP64) re-expanded about a centre displaced by d: out = L2L(L, d)
__device__ __forceinline__ void l2l(double* out, const double* L, const double* d) {
	double W[35];
	taylor_weights(W, d);
#pragma unroll
	for (int t = 0; t < 35; t++) {
		double acc = 0.0;
#pragma unroll
		for (int s = 0; s < 35; s++)
			if (SL[t].ord + SL[s].ord <= 4) acc = fma(L[merge_slot(t, s)], W[s], acc);
		out[t] = acc;
	}
}

__device__ __forceinline__ void l2p(double& phi, double* g, const double* L, const double* d) {
	double W[35];
	taylor_weights(W, d);
	double p = 0.0;
#pragma unroll
	for (int s = 0; s < 35; s++) p = fma(L[s], W[s], p);
	phi = p;
#pragma unroll
	for (int a = 0; a < 3; a++) {
		double acc = 0.0;
#pragma unroll
		for (int s = 0; s < 20; s++) acc = fma(L[merge_slot(s, 1 + a)], W[s], acc);
		g[a] = -acc;
	}
}

}  // namespace tg
