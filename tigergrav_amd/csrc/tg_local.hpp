// tg_local.hpp -- local-expansion translation (L2L) and evaluation (L2P) in FP64.
//
// The exact 4th-order Taylor re-expansion of tigergrav/expansion.hpp:185-233 and evaluation of
// tigergrav/expansion.hpp:236-269.  Slots carry the symmetric-tensor weights mult(s)/|s|!
// (the reference's expansion_factors), so with W_s(d) = mult(s)/|s|! d^s:
//   L2L:  L'_t = sum_{|s| <= 4-|t|} L_{t u s} W_s(d)
//   L2P:  phi  = sum_s L_s W_s(d),   g_a = -sum_{|s| <= 3} L_{s u a} W_s(d)
#pragma once
#include <utility>

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
__host__ __device__ constexpr double fact4(int n) {  // n! for n <= 4
	return n == 4 ? 24.0 : (n == 3 ? 6.0 : (n == 2 ? 2.0 : 1.0));
}
__host__ __device__ constexpr double slot_weight(int s) {
	// mult(s)/|s|! = 1 / prod(count_v!)
	return 1.0 / (fact4(count_idx(SL[s], 0)) * fact4(count_idx(SL[s], 1)) * fact4(count_idx(SL[s], 2)));
}

// Compile-time slot arithmetic: every index below is a template argument, so the expansions
// are straight-line code over registers.
template <int T, int S>
struct Merge {
	static constexpr bool valid = SL[T].ord + SL[S].ord <= 4;
	static constexpr int value = valid ? merge_slot(T, S) : 0;
};

template <int S>
__device__ __forceinline__ double taylor_weight(const double* d) {
	constexpr double w0 = slot_weight(S);
	constexpr int o = SL[S].ord;
	double w = w0;
	if constexpr (o > 0) w *= d[SL[S].i[0]];
	if constexpr (o > 1) w *= d[SL[S].i[1]];
	if constexpr (o > 2) w *= d[SL[S].i[2]];
	if constexpr (o > 3) w *= d[SL[S].i[3]];
	return w;
}

template <int... S>
__device__ __forceinline__ void taylor_weights_seq(double* W, const double* d, std::integer_sequence<int, S...>) {
	((W[S] = taylor_weight<S>(d)), ...);
}

__device__ __forceinline__ void taylor_weights(double* W, const double* d) {
	taylor_weights_seq(W, d, std::make_integer_sequence<int, 35>{});
}

// sum over the slots s with ord(t) + ord(s) <= 4 of L[t u s] W[s], in slot order
template <int T, int... S>
__device__ __forceinline__ double shift_row(const double* L, const double* W, std::integer_sequence<int, S...>) {
	double acc = 0.0;
	((acc = Merge<T, S>::valid ? fma(L[Merge<T, S>::value], W[S], acc) : acc), ...);
	return acc;
}

template <int... T>
__device__ __forceinline__ void shift_all(double* out, const double* L, const double* W, std::integer_sequence<int, T...>) {
	((out[T] = shift_row<T>(L, W, std::make_integer_sequence<int, 35>{})), ...);
}

// L (35, FP64) re-expanded about a centre displaced by d: out = L2L(L, d)
__device__ __forceinline__ void l2l(double* out, const double* L, const double* d) {
	double W[35];
	taylor_weights(W, d);
	shift_all(out, L, W, std::make_integer_sequence<int, 35>{});
}

// potential and acceleration at offset d: phi = sum_s L_s W_s, g_a = -sum_{|s| <= 3} L_{s u a} W_s
template <int A, int... S>
__device__ __forceinline__ double grad_row(const double* L, const double* W, std::integer_sequence<int, S...>) {
	double acc = 0.0;
	((acc = fma(L[Merge<S, 1 + A>::value], W[S], acc)), ...);
	return -acc;
}

template <int... S>
__device__ __forceinline__ double dot_seq(const double* L, const double* W, std::integer_sequence<int, S...>) {
	double acc = 0.0;
	((acc = fma(L[S], W[S], acc)), ...);
	return acc;
}

__device__ __forceinline__ void l2p(double& phi, double* g, const double* L, const double* d) {
	double W[35];
	taylor_weights(W, d);
	phi = dot_seq(L, W, std::make_integer_sequence<int, 35>{});
	g[0] = grad_row<0>(L, W, std::make_integer_sequence<int, 20>{});
	g[1] = grad_row<1>(L, W, std::make_integer_sequence<int, 20>{});
	g[2] = grad_row<2>(L, W, std::make_integer_sequence<int, 20>{});
}

}  // namespace tg
