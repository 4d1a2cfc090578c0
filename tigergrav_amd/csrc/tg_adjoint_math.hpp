// tg_adjoint_math.hpp -- device math of the adjoint: derivative tensors to fifth order.
//
// The adjoint of the interactions needs one derivative more than the forward: the acceleration of
// an octupole source (M2P, fourth derivatives of the Green's function) differentiated with respect
// to the sink position, and the local expansion of a point mass (P2L, up to fourth derivatives)
// differentiated with respect to the source position, need fifth derivatives.
//
// Symmetric tensors up to order 5 are stored in 56 slots: the 35 of the forward's expansion layout
// ([0], [1+a], [4+s2], [10+s3], [20+s4]; tigergrav/expansion.hpp:77-137) followed by the 21
// sorted index quintuples in lexicographic order.  Slot k holds the index multiset with counts
// (n0, n1, n2) of the values 0, 1, 2; every loop over slots is unrolled with compile-time indices.
// The periodic correction's derivatives come from the forward's table (green_ewald_tab, tg_math.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "tg_math.hpp"

namespace tg {
namespace adj {

constexpr int NS5 = 56;  // slots of orders 0..5

struct Cnt {
	int n0, n1, n2;
};
struct SlotTable {
	Cnt c[NS5];
};
// orders 0..5, each in lexicographic order of the sorted tuple 0^n0 1^n1 2^n2: more leading zeros
// first, then more ones
constexpr SlotTable make_slots() {
	SlotTable t{};
	int k = 0;
	for (int n = 0; n <= 5; n++)
		for (int n0 = n; n0 >= 0; n0--)
			for (int n1 = n - n0; n1 >= 0; n1--) {
				t.c[k].n0 = n0;
				t.c[k].n1 = n1;
				t.c[k].n2 = n - n0 - n1;
				k++;
			}
	return t;
}
constexpr SlotTable ST = make_slots();

constexpr int order_of(int k) { return ST.c[k].n0 + ST.c[k].n1 + ST.c[k].n2; }
constexpr int slot_of(int n0, int n1, int n2) {
	for (int k = 0; k < NS5; k++)
		if (ST.c[k].n0 == n0 && ST.c[k].n1 == n1 && ST.c[k].n2 == n2) return k;
	return -1;
}
// slot of the multiset union of slots k and j
constexpr int merge(int k, int j) {
	return slot_of(ST.c[k].n0 + ST.c[j].n0, ST.c[k].n1 + ST.c[j].n1, ST.c[k].n2 + ST.c[j].n2);
}
// slot k with one more index a
constexpr int add(int k, int a) { return merge(k, 1 + a); }
// the forward's layout agrees on orders <= 4
static_assert(slot_of(2, 0, 0) == 4 && slot_of(0, 1, 1) == 8 && slot_of(1, 1, 1) == 14 && slot_of(0, 0, 3) == 19 &&
				slot_of(2, 1, 1) == 24 && slot_of(0, 0, 4) == 34 && slot_of(5, 0, 0) == 35 && slot_of(0, 0, 5) == 55,
		"slot layout");

// number of ways to pair p couples among n equal indices: n! / (2^p p! (n-2p)!)
constexpr float pairings(int n, int p) {
	float num = 1.0f;
	for (int i = 0; i < 2 * p; i++) num *= float(n - i);
	float den = 1.0f;
	for (int i = 1; i <= p; i++) den *= 2.0f * float(i);
	return num / den;
}

// component k of the derivative tensor of a radial function with d_j = (r^-1 d/dr)^j f:
//   D_k = sum over p0, p1, p2 couples of equal indices of d_{|k| - p} prod_i pairings(n_i, p_i) x_i^(n_i - 2 p_i)
// px/py/pz hold the powers 0..5 of the coordinates; all bounds are compile-time, so the loops unroll
template <int K>
__device__ __forceinline__ float radial_comp(const float* d, const float* px, const float* py, const float* pz) {
	constexpr int n0 = ST.c[K].n0, n1 = ST.c[K].n1, n2 = ST.c[K].n2, n = n0 + n1 + n2;
	float acc = 0.0f;
#pragma unroll
	for (int p0 = 0; 2 * p0 <= n0; p0++)
#pragma unroll
		for (int p1 = 0; 2 * p1 <= n1; p1++)
#pragma unroll
			for (int p2 = 0; 2 * p2 <= n2; p2++) {
				const float c = pairings(n0, p0) * pairings(n1, p1) * pairings(n2, p2);
				acc = fmaf(c * px[n0 - 2 * p0] * py[n1 - 2 * p1] * pz[n2 - 2 * p2], d[n - p0 - p1 - p2], acc);
			}
	return acc;
}

template <int... K>
__device__ __forceinline__ void radial_add_seq(float* D, const float* d, const float* px, const float* py, const float* pz,
		std::integer_sequence<int, K...>) {
	((D[K] += radial_comp<K>(d, px, py, pz)), ...);
}
__device__ __forceinline__ void powers5(float* p, float x) {
	p[0] = 1.0f;
#pragma unroll
	for (int e = 1; e <= 5; e++) p[e] = p[e - 1] * x;
}
// D += the first NS slots (orders 0..5: 1, 4, 10, 20, 35, 56) of the derivative tensor of the radial
// function with factors d[0..5] at x
template <int NS>
__device__ __forceinline__ void radial_add(float* D, const float* d, const float* x) {
	float px[6], py[6], pz[6];
	powers5(px, x[0]);
	powers5(py, x[1]);
	powers5(pz, x[2]);
	radial_add_seq(D, d, px, py, pz, std::make_integer_sequence<int, NS>{});
}

// derivatives of -1/r (first NS slots), zero below r = 1e-9 like the forward (tigergrav/green.hpp:167-186)
template <int NS>
__device__ __forceinline__ void green_direct_n(float* D, const float* x) {
	const float r2 = fmaf(x[2], x[2], fmaf(x[1], x[1], x[0] * x[0]));
	const float ri = (r2 > 1e-18f) ? __builtin_amdgcn_rsqf(r2) : 0.0f;
	const float ri2 = ri * ri;
	float d[6];
	d[0] = -ri;
#pragma unroll
	for (int k = 0; k < 5; k++) d[k + 1] = -float(2 * k + 1) * d[k] * ri2;
#pragma unroll
	for (int k = 0; k < NS; k++) D[k] = 0.0f;
	radial_add<NS>(D, d, x);
}

}  // namespace adj
}  // namespace tg
