// tg_math.hpp -- device math for the gfx950 gravity kernels.
//
// Symmetric-tensor storage follows the reference layouts so results can be handed back
// unchanged: expansion slots [0], [1+a], [4+s2], [10+s3], [20+s4] (tigergrav/expansion.hpp:77-137),
// multipole slots [0], [1+s2], [7+s3] (tigergrav/multipole.hpp:56-82), where s2/s3/s4 rank the
// sorted index tuple lexicographically.  Every loop over tensor slots is fully unrolled with
// compile-time indices, so tensors live in VGPRs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "tg_ewald_table.hpp"

namespace tg {

// ---- compile-time slot tables ----------------------------------------------------------------
struct Tup {
	int i[4];
};
constexpr Tup T2[6] = {{{0, 0}}, {{0, 1}}, {{0, 2}}, {{1, 1}}, {{1, 2}}, {{2, 2}}};
constexpr Tup T3[10] = {{{0, 0, 0}}, {{0, 0, 1}}, {{0, 0, 2}}, {{0, 1, 1}}, {{0, 1, 2}},
		{{0, 2, 2}}, {{1, 1, 1}}, {{1, 1, 2}}, {{1, 2, 2}}, {{2, 2, 2}}};
constexpr Tup T4[15] = {{{0, 0, 0, 0}}, {{0, 0, 0, 1}}, {{0, 0, 0, 2}}, {{0, 0, 1, 1}}, {{0, 0, 1, 2}},
		{{0, 0, 2, 2}}, {{0, 1, 1, 1}}, {{0, 1, 1, 2}}, {{0, 1, 2, 2}}, {{0, 2, 2, 2}}, {{1, 1, 1, 1}},
		{{1, 1, 1, 2}}, {{1, 1, 2, 2}}, {{1, 2, 2, 2}}, {{2, 2, 2, 2}}};

__host__ __device__ constexpr int s2(int a, int b) {
	return a > b ? s2(b, a) : (a == 0 ? b : (a == 1 ? 2 + b : 5));
}
__host__ __device__ constexpr int s3sorted(int a, int b, int c) {
	return a == 0 ? s2(b, c) : (a == 1 ? (b == 1 ? (c == 1 ? 6 : 7) : 8) : 9);
}
__host__ __device__ constexpr int s3(int a, int b, int c) {
	// sort three values
	return (a <= b) ? ((b <= c) ? s3sorted(a, b, c) : ((a <= c) ? s3sorted(a, c, b) : s3sorted(c, a, b)))
	                : ((a <= c) ? s3sorted(b, a, c) : ((b <= c) ? s3sorted(b, c, a) : s3sorted(c, b, a)));
}
__host__ __device__ constexpr int s4sorted(int a, int b, int c, int d) {
	return a == 0 ? s3sorted(b, c, d) : (a == 1 ? (b == 1 ? (c == 1 ? (d == 1 ? 10 : 11) : 12) : 13) : 14);
}
// a - b in position units (2^32 per box) between a particle and an expansion centre: the int32
// wrap-around difference when the box is periodic, else the exact difference.  The two agree for a
// particle inside its node; a particle moved across the box edge with the tree held fixed keeps a
// continuous offset only through the first.
__host__ __device__ inline int64_t centre_offset(int32_t a, int32_t b, bool wrap) {
	return wrap ? int64_t(int32_t(uint32_t(a) - uint32_t(b))) : int64_t(a) - int64_t(b);
}
__host__ __device__ constexpr int min2(int a, int b) { return a < b ? a : b; }
__host__ __device__ constexpr int max2(int a, int b) { return a < b ? b : a; }
__host__ __device__ constexpr int s4(int a, int b, int c, int d) {
	// sorting network for four values
	return s4sorted(min2(min2(a, b), min2(c, d)),
			min2(max2(min2(a, b), min2(c, d)), min2(max2(a, b), max2(c, d))),
			max2(max2(min2(a, b), min2(c, d)), min2(max2(a, b), max2(c, d))),
			max2(max2(a, b), max2(c, d)));
}
// expansion slots
__host__ __device__ constexpr int E1(int a) { return 1 + a; }
__host__ __device__ constexpr int E2(int a, int b) { return 4 + s2(a, b); }
__host__ __device__ constexpr int E3(int a, int b, int c) { return 10 + s3(a, b, c); }
__host__ __device__ constexpr int E4(int a, int b, int c, int d) { return 20 + s4(a, b, c, d); }
// multipole slots
__host__ __device__ constexpr int M2(int a, int b) { return 1 + s2(a, b); }
__host__ __device__ constexpr int M3(int a, int b, int c) { return 7 + s3(a, b, c); }

// number of distinct index values of a sorted tuple, and multiplicity (permutations) of it
__host__ __device__ constexpr int ndistinct(const Tup& t, int n) {
	return n == 2 ? 1 + (t.i[0] != t.i[1])
	              : (n == 3 ? 1 + (t.i[0] != t.i[1]) + (t.i[1] != t.i[2]) : 0);
}
__host__ __device__ constexpr int mult2(const Tup& t) { return t.i[0] == t.i[1] ? 1 : 2; }
__host__ __device__ constexpr int mult3(const Tup& t) {
	return (t.i[0] == t.i[1] && t.i[1] == t.i[2]) ? 1 : ((t.i[0] == t.i[1] || t.i[1] == t.i[2]) ? 3 : 6);
}
// expansion_factors (tigergrav/expansion.hpp:163-182): multiplicity / k!
__host__ __device__ constexpr float ef2(int s) { return mult2(T2[s]) * 0.5f; }
__host__ __device__ constexpr float ef3(int s) { return mult3(T3[s]) * (1.0f / 6.0f); }

constexpr double POS_INV = 1.0 / 4294967296.0;
constexpr float POS_INVF = 1.0f / 4294967296.0f;

// calls f(std::integral_constant<int, i>) for i = 0..N-1, so slot indices are compile-time
template <int... I, class F>
__device__ __forceinline__ void static_for_seq(std::integer_sequence<int, I...>, F&& f) {
	(f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
	static_for_seq(std::make_integer_sequence<int, N>{}, f);
}

// ---- derivative tensor from radial factors ----------------------------------------------------
// For a radial function f(r), with d_k = (r^-1 d/dr)^k f, the Cartesian derivatives are
//   D_a    = x_a d1
//   D_ab   = x_a x_b d2 + delta_ab d1
//   D_abc  = x_a x_b x_c d3 + (delta_ab x_c + delta_ac x_b + delta_bc x_a) d2
//   D_abcd = x_a x_b x_c x_d d4 + (6 delta x x terms) d3 + (3 delta delta terms) d2
// (the reference builds the same tensor by in-place slot updates, tigergrav/green.hpp:62-117).
// Each index pattern takes its delta terms factored out, with per-axis factors shared between
// patterns (a, b, c distinct):
//   D_aaa  = x_a (x_a^2 d3 + 3 d2)              D_aab = x_b w_a,   w_a = x_a^2 d3 + d2
//   D_abc  = x_a x_b x_c d3
//   D_aaaa = x_a^2 (x_a^2 d4 + 6 d3) + 3 d2     D_aaab = x_a x_b u_a,   u_a = x_a^2 d4 + 3 d3
//   D_aabb = x_a^2 v_b + w_b                    D_aabc = x_b x_c v_a,   v_a = x_a^2 d4 + d3
// about 60 operations for the 35 slots where the expanded sums take about 140.
__host__ __device__ constexpr int count_of(const Tup& t, int n, int v) {
	return (n > 0 && t.i[0] == v) + (n > 1 && t.i[1] == v) + (n > 2 && t.i[2] == v) + (n > 3 && t.i[3] == v);
}
// the axis that occurs exactly k times in the first n indices of t (-1 if none)
__host__ __device__ constexpr int axis_with(const Tup& t, int n, int k) {
	return count_of(t, n, 0) == k ? 0 : count_of(t, n, 1) == k ? 1 : count_of(t, n, 2) == k ? 2 : -1;
}
// the axis that does not occur (-1 if all do)
__host__ __device__ constexpr int axis_missing(const Tup& t, int n) { return axis_with(t, n, 0); }

__device__ __forceinline__ void radial_tensor(float* D, float d0, float d1, float d2, float d3, float d4, const float* x) {
	D[0] = d0;
#pragma unroll
	for (int a = 0; a < 3; a++) D[1 + a] = x[a] * d1;
	float x2[6];
#pragma unroll
	for (int s = 0; s < 6; s++) {
		const int a = T2[s].i[0], b = T2[s].i[1];
		x2[s] = x[a] * x[b];
		D[4 + s] = (a == b) ? fmaf(x2[s], d2, d1) : x2[s] * d2;
	}
	const float d2x3 = 3.0f * d2, d3x3 = 3.0f * d3, d3x6 = 6.0f * d3;
	float w[3], u[3], v[3];
	static_for<3>([&](auto a_) {
		constexpr int a = decltype(a_)::value;
		constexpr int aa = s2(a, a);
		const float xaa = x2[aa];
		w[a] = fmaf(xaa, d3, d2);
		u[a] = fmaf(xaa, d4, d3x3);
		v[a] = fmaf(xaa, d4, d3);
	});
	static_for<10>([&](auto s_) {
		constexpr int s = decltype(s_)::value;
		constexpr Tup t = T3[s];
		constexpr int a3 = axis_with(t, 3, 3), a2 = axis_with(t, 3, 2), a1 = axis_with(t, 3, 1);
		if constexpr (a3 >= 0) {
			constexpr int aa = s2(a3, a3);
			D[10 + s] = x[a3] * fmaf(x2[aa], d3, d2x3);
		} else if constexpr (a2 >= 0) {
			D[10 + s] = x[a1] * w[a2];
		} else {
			D[10 + s] = x2[1] * (x[2] * d3);  // x2[1] = x_0 x_1
		}
	});
	static_for<15>([&](auto s_) {
		constexpr int s = decltype(s_)::value;
		constexpr Tup t = T4[s];
		constexpr int a4 = axis_with(t, 4, 4), a3 = axis_with(t, 4, 3), a2 = axis_with(t, 4, 2), a1 = axis_with(t, 4, 1);
		if constexpr (a4 >= 0) {
			constexpr int aa = s2(a4, a4);
			const float xaa = x2[aa];
			D[20 + s] = fmaf(xaa, fmaf(xaa, d4, d3x6), d2x3);
		} else if constexpr (a3 >= 0) {
			constexpr int ab = s2(a3, a1);
			D[20 + s] = x2[ab] * u[a3];
		} else if constexpr (a2 >= 0 && a1 < 0) {
			// two axes twice each: a2 is the first of them, b the one neither missing nor a2
			constexpr int b = 3 - a2 - axis_missing(t, 4);
			constexpr int aa = s2(a2, a2);
			D[20 + s] = fmaf(x2[aa], v[b], w[b]);
		} else {
			// one axis twice (a2), the other two once
			constexpr int b = a2 == 0 ? 1 : 0, c = a2 == 2 ? 1 : 2;
			constexpr int bc = s2(b, c);
			D[20 + s] = x2[bc] * v[a2];
		}
	});
}

// derivatives of -1/r (tigergrav/green.hpp:167-186): zero below r = 1e-9 like the reference
__device__ __forceinline__ void green_direct(float* D, const float* x) {
	const float r2 = fmaf(x[2], x[2], fmaf(x[1], x[1], x[0] * x[0]));
	const float ri = (r2 > 1e-18f) ? __builtin_amdgcn_rsqf(r2) : 0.0f;
	const float ri2 = ri * ri;
	const float d0 = -ri;
	const float d1 = ri * ri2;
	const float d2 = -3.0f * d1 * ri2;
	const float d3 = -5.0f * d2 * ri2;
	const float d4 = -7.0f * d3 * ri2;
	radial_tensor(D, d0, d1, d2, d3, d4, x);
}

// ---- contractions (src/gravity.cpp:26-169) ----------------------------------------------------
// M2L: L += contraction of the multipole M with D (reference gravity.cpp:26-90)
// Every term is added straight into L, in the reference's order of terms (as its loop adds each
// term into its L).  HARM: D is traceless in any two of its indices (derivatives of 1/r away from
// the origin), so D_{k,zz} = -D_{k,xx} - D_{k,yy} and a multipole's zz-holding components fold
// into the others: the quadrupole contracts in 5 terms and the octupole in 7 (each L component
// whose terms reach D through them: L0, L_a, L_ab), with the 8 folded weights formed once per
// entry.
template <bool HARM>
__device__ __forceinline__ void m2l(float* L, const float* M, const float* D) {
	float q[6], o[10];
#pragma unroll
	for (int s = 0; s < 6; s++) q[s] = M[1 + s] * ef2(s);
#pragma unroll
	for (int s = 0; s < 10; s++) o[s] = -M[7 + s] * ef3(s);
	if (HARM) {
		// quadrupole slots xx 0, xy 1, xz 2, yy 3, yz 4, zz 5; octupole xxx 0, xxy 1, xxz 2, xyy 3,
		// xyz 4, xzz 5, yyy 6, yyz 7, yzz 8, zzz 9
		q[0] -= q[5];
		q[3] -= q[5];
		o[0] -= o[5];
		o[3] -= o[5];
		o[1] -= o[8];
		o[6] -= o[8];
		o[2] -= o[9];
		o[7] -= o[9];
	}
	auto quad = [&](auto slot_of, float acc) {
#pragma unroll
		for (int s = 0; s < 6; s++)
			if (!HARM || s != 5) acc = fmaf(q[s], D[slot_of(s)], acc);
		return acc;
	};
	auto oct = [&](auto slot_of, float acc) {
#pragma unroll
		for (int s = 0; s < 10; s++)
			if (!HARM || (s != 5 && s != 8 && s != 9)) acc = fmaf(o[s], D[slot_of(s)], acc);
		return acc;
	};
	float l0 = fmaf(M[0], D[0], L[0]);
	l0 = quad([](int s) { return 4 + s; }, l0);
	l0 = oct([](int s) { return 10 + s; }, l0);
	L[0] = l0;
#pragma unroll
	for (int a = 0; a < 3; a++) {
		float la = fmaf(M[0], D[1 + a], L[1 + a]);
		la = quad([a](int s) { return E3(a, T2[s].i[0], T2[s].i[1]); }, la);
		la = oct([a](int s) { return E4(a, T3[s].i[0], T3[s].i[1], T3[s].i[2]); }, la);
		L[1 + a] = la;
	}
#pragma unroll
	for (int t = 0; t < 6; t++) {
		const int a = T2[t].i[0], b = T2[t].i[1];
		float lab = fmaf(M[0], D[4 + t], L[4 + t]);
		lab = quad([a, b](int s) { return E4(a, b, T2[s].i[0], T2[s].i[1]); }, lab);
		L[4 + t] = lab;
	}
#pragma unroll
	for (int s = 10; s < 35; s++) L[s] = fmaf(M[0], D[s], L[s]);
}

// P2L: L += m D (reference gravity.cpp:92-121)
__device__ __forceinline__ void p2l(float* L, float m, const float* D) {
#pragma unroll
	for (int s = 0; s < 35; s++) L[s] = fmaf(m, D[s], L[s]);
}

// M2P: potential and acceleration at the sink (reference gravity.cpp:123-169).  The potential's
// octupole term reproduces the reference's loop (a over all values, b <= c): each sorted triple
// is weighted by its number of distinct index values.
__host__ __device__ constexpr float wq3(int s) { return ndistinct(T3[s], 3) * ef3(s); }

__device__ __forceinline__ void m2p(float& phi, float* g, const float* M, const float* D) {
	float p = M[0] * D[0];
#pragma unroll
	for (int s = 0; s < 6; s++) p = fmaf(M[1 + s] * ef2(s), D[4 + s], p);
#pragma unroll
	for (int s = 0; s < 10; s++) p = fmaf(-M[7 + s] * wq3(s), D[10 + s], p);
	phi += p;
#pragma unroll
	for (int a = 0; a < 3; a++) {
		float ga = -M[0] * D[1 + a];
#pragma unroll
		for (int s = 0; s < 6; s++) ga = fmaf(-M[1 + s] * ef2(s), D[E3(a, T2[s].i[0], T2[s].i[1])], ga);
#pragma unroll
		for (int s = 0; s < 10; s++) ga = fmaf(M[7 + s] * ef3(s), D[E4(a, T3[s].i[0], T3[s].i[1], T3[s].i[2])], ga);
		g[a] += ga;
	}
}

// ---- Ewald periodic correction ----------------------------------------------------------------
// The tabulated correction (tg_ewald_table.hpp): c_ewt points at the per-device table (defined
// in tg_kernels.hip; another translation unit keeps its own copy of the pointer).
extern __constant__ const float* c_ewt;

// slots read for the first nout outputs: every slot up to three orders above the highest output
// order, rounded up to whole float4 loads
constexpr int ewt_nload(int nout) {
	const int o = nout <= 1 ? 0 : nout <= 4 ? 1 : nout <= 10 ? 2 : nout <= 20 ? 3 : 4;
	const int s = (o + 4) * (o + 5) * (o + 6) / 6;
	return (s + 3) / 4 * 4;
}

template <int V>
constexpr int ewt_cx = V;  // forces a constant slot index to be folded at compile time

template <int... S, class F>
__device__ __forceinline__ void ewt_for_seq(std::integer_sequence<int, S...>, F&& f) {
	(f(std::integral_constant<int, S>{}), ...);
}

// exchanges components k and swap_ab(k, a, b) of D[0..NOUT) where cond holds: the tensor with
// coordinates a and b exchanged
template <int NOUT, int A, int B>
__device__ __forceinline__ void ewt_transpose_if(float* D, bool cond) {
	ewt_for_seq(std::make_integer_sequence<int, NOUT>{}, [&](auto k_) {
		constexpr int k = decltype(k_)::value;
		constexpr int p = ewt::swap_ab(k, A, B);
		if constexpr (p > k && p < NOUT) {
			const float a = D[k], b = D[p];
			D[k] = cond ? b : a;
			D[p] = cond ? a : b;
		}
	});
}

// The first NOUT components (all of orders 0..1, 0..2, 0..4 or 0..5: NOUT = 4, 10, 35, 56) of
// the periodic-correction tensor (tigergrav/green.hpp:189-286) at a wrapped separation x
// (|x_i| <= 1/2).  |x| is sorted into the wedge u0 >= u1 >= u2 by three compare-exchanges; the
// components there come from a third-order Taylor expansion about the nearest grid point of
// `table` (FP64 values of the correction, rounded to float, on a 1/64 grid of the wedge); the
// exchanges are then undone on the components in reverse order, and the sign of each component is
// flipped by the coordinates that are negative and occur an odd number of times.  Interpolation
// error below 5e-6 of each order's largest value (1.7e-6 RMS over orders 0..4, 1.1e-7 at order 1).
// Order-5 components (the adjoint's) take a second-order expansion, the table ending at order 7:
// error about 1e-4 of that order's largest value.  At |x| <= 1e-6 the reference's convention
// D = (2.8372975, 0, ..., 0) applies (tigergrav/green.hpp:272): the table's last row, read with a
// zero offset.
template <int NOUT>
__device__ __forceinline__ void green_ewald_tab(float* D, const float* x, const float* table) {
	constexpr int NLOAD = NOUT == 56 ? TG_EWT_S : ewt_nload(NOUT);
	static_assert(NOUT == 56 || NOUT == 35 || NOUT == 10 || NOUT == 4, "orders 0..5, 0..4, 0..2 or 0..1");
	static_assert(NLOAD <= TG_EWT_S, "slots read");
	const float r2 = fmaf(x[2], x[2], fmaf(x[1], x[1], x[0] * x[0]));
	float u[3], sg[3];
#pragma unroll
	for (int a = 0; a < 3; a++) {
		u[a] = fabsf(x[a]);
		sg[a] = x[a] < 0.0f ? -1.0f : 1.0f;
	}
	// sort |x| descending: exchange (0,1), (1,2), (0,1) where the later coordinate is larger
	auto cx = [](float& p, float& q, bool c) {
		const float a = p, b = q;
		p = c ? b : a;
		q = c ? a : b;
	};
	const bool s1 = u[1] > u[0];
	cx(u[0], u[1], s1);
	const bool s2 = u[2] > u[1];
	cx(u[1], u[2], s2);
	const bool s3 = u[1] > u[0];
	cx(u[0], u[1], s3);
	// a coincident pair reads the table's last row, (2.8372975, 0, ..., 0), with a zero offset
	const bool z = !(r2 > 1e-12f);
	float d[3];
	int gi[3];
#pragma unroll
	for (int a = 0; a < 3; a++) {
		gi[a] = min(int(fmaf(u[a], float(TG_EWT_INV), 0.5f)), TG_EWT_N - 1);
		d[a] = z ? 0.0f : u[a] - float(gi[a]) * (1.0f / TG_EWT_INV);
	}
	const int row = z ? TG_EWT_P : ewt::wedge_index(gi[0], gi[1], gi[2]);
	const float4* tp = reinterpret_cast<const float4*>(table + size_t(row) * TG_EWT_STRIDE);
	// the row holds the slots with the third axis at most once; the others follow from the trace
	// relation (a constant Laplacian, -4 pi; 0 for the coincident row), each from two slots before it
	constexpr int MAXO = ewt::ST.c[NLOAD - 1].n0 + ewt::ST.c[NLOAD - 1].n1 + ewt::ST.c[NLOAD - 1].n2;  // highest order read
	constexpr int NSTORED = ewt::stored_count(MAXO);
	float c[(NSTORED + 3) / 4 * 4];
#pragma unroll
	for (int q = 0; q < (NSTORED + 3) / 4; q++) {
		const float4 v = tp[q];
		c[4 * q] = v.x;
		c[4 * q + 1] = v.y;
		c[4 * q + 2] = v.z;
		c[4 * q + 3] = v.w;
	}
	const float lap = z ? 0.0f : -12.566370614359172f;
	float t[NLOAD];
	ewt_for_seq(std::make_integer_sequence<int, NLOAD>{}, [&](auto k_) {
		constexpr int k = decltype(k_)::value;
		constexpr int n0 = ewt::ST.c[k].n0, n1 = ewt::ST.c[k].n1, n2 = ewt::ST.c[k].n2;
		if constexpr (n2 <= 1) {
			constexpr int si = ewt::stored_index(k);
			static_assert(si >= 0 && si < (NSTORED + 3) / 4 * 4, "stored slot read");
			t[k] = c[si];
		} else {
			constexpr int ka = ewt::slot_of(n0 + 2, n1, n2 - 2), kb = ewt::slot_of(n0, n1 + 2, n2 - 2);
			static_assert(ka >= 0 && ka < k && kb >= 0 && kb < k, "trace terms come first");
			t[k] = ((n0 == 0 && n1 == 0 && n2 == 2) ? lap : 0.0f) - t[ka] - t[kb];
		}
	});
	// the Taylor monomials d^j / j! of orders 1..3 (slots 1..19)
	float w[20];
	w[0] = 1.0f;
	ewt_for_seq(std::make_integer_sequence<int, 19>{}, [&](auto j_) {
		constexpr int j = decltype(j_)::value + 1;
		constexpr int n0 = ewt::ST.c[j].n0, n1 = ewt::ST.c[j].n1, n2 = ewt::ST.c[j].n2;
		float m = ewt::taylor_coef(j);
#pragma unroll
		for (int e = 0; e < n0; e++) m *= d[0];
#pragma unroll
		for (int e = 0; e < n1; e++) m *= d[1];
#pragma unroll
		for (int e = 0; e < n2; e++) m *= d[2];
		w[j] = m;
	});
	// The correction has a constant Laplacian in the cell (-4 pi: the periodic potential's uniform
	// background), so every derivative tensor's trace is fixed by lower slots: a component holding
	// the wedge's third axis twice or more is C - D[k - 2z + 2x] - D[k - 2z + 2y], C = -4 pi for
	// D_zz and 0 otherwise.  Those components (10 of the 35 to order 4) come from that relation,
	// in slot order (their two terms come earlier), the others from the Taylor step.
	ewt_for_seq(std::make_integer_sequence<int, NOUT>{}, [&](auto k_) {
		constexpr int k = decltype(k_)::value;
		constexpr int n0 = ewt::ST.c[k].n0, n1 = ewt::ST.c[k].n1, n2 = ewt::ST.c[k].n2;
		if constexpr (n2 >= 2) {
			constexpr int ka = ewt::slot_of(n0 + 2, n1, n2 - 2), kb = ewt::slot_of(n0, n1 + 2, n2 - 2);
			static_assert(ka >= 0 && ka < k && kb >= 0 && kb < k, "trace terms come first");
			const float c = (n0 == 0 && n1 == 0 && n2 == 2) ? lap : 0.0f;
			D[k] = c - D[ka] - D[kb];
		} else {
			// monomials of orders 1..3 (slots 1..19), or 1..2 (slots 1..9) where order 8 would be needed
			constexpr int jn = n0 + n1 + n2 + 3 <= TG_EWT_ORDER ? 19 : 9;
			float v = t[k];
			ewt_for_seq(std::make_integer_sequence<int, jn>{}, [&](auto j_) {
				constexpr int j = decltype(j_)::value + 1;
				static_assert(ewt::merge(k, j) >= 0 && ewt::merge(k, j) < NLOAD, "slot read");
				v = fmaf(t[ewt_cx<ewt::merge(k, j)>], w[j], v);
			});
			D[k] = v;
		}
	});
	// back from the wedge to the separation's own axes
	ewt_transpose_if<NOUT, 0, 1>(D, s3);
	ewt_transpose_if<NOUT, 1, 2>(D, s2);
	ewt_transpose_if<NOUT, 0, 1>(D, s1);
	// the sign of a component is the product of the signs of the coordinates it holds an odd
	// number of times: one multiply per component from the products of the three signs
	const float sgp[8] = {1.0f, sg[0], sg[1], sg[0] * sg[1], sg[2], sg[0] * sg[2], sg[1] * sg[2], sg[0] * sg[1] * sg[2]};
	ewt_for_seq(std::make_integer_sequence<int, NOUT>{}, [&](auto k_) {
		constexpr int k = decltype(k_)::value;
		constexpr int o = (ewt::odd_in(k, 0) ? 1 : 0) | (ewt::odd_in(k, 1) ? 2 : 0) | (ewt::odd_in(k, 2) ? 4 : 0);
		if constexpr (o != 0) D[k] *= sgp[o];
	});
}

__device__ __forceinline__ void green_ewald(float* D, const float* x) { green_ewald_tab<35>(D, x, c_ewt); }

}  // namespace tg
