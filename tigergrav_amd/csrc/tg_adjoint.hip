// tg_adjoint.hip -- the adjoint of the batched force solve.
//
// For a loss L = sum_i w_i . g_i of the accelerations of one solve, the gradients dL/dx_k and
// dL/dm_k, with the tree, the interaction lists and the expansion centres of that solve held
// fixed (the function tg_solver_update_particles evaluates).  The solve is linear in every
// expansion and moment, so the adjoint runs the solve's operators transposed, recomputing every
// interaction from the stored lists (nothing of the forward is kept but its expansions dL):
//   1. sinks, over the forward's leaf work items:
//      k_adj_l2p       g = -grad T_L(x - X):   dL/dx += -w^T Hess T_L,  lambda_leaf = dg/dL^T w;
//      k_adj_up        Lambda_a = sum over leaves B below a of S(X_B - X_a)^T lambda_B (L2L transposed),
//                      level by level from the children's
//      k_adj_pp(_ewald) tidal tensor J(d) = dF/dd per pair; the lists are symmetric (leaf B is in
//                      A's PP list iff A is in B's), so each particle's source-side terms are gathered
//                      from its own list: dL/dx_k += J(d_kj) (m_j w_k - m_k w_j), dL/dm_k += -w_j . F(d_kj)
//      k_adj_pc_sink   dL/dx_i += w . dg/dx_i (fifth derivatives); k_adj_pc_t (by source node c)
//                      mu_c += dg/dM_c^T w
//   2. node lists:
//      k_adj_cc_t  mu_c += M2L(., X_a - X_c)^T Lambda_a, over the CC entries sorted by source c
//      k_adj_cp_t  CP, by source leaf: dL/dm_j += Lambda_a . D(X_a - x_j), dL/dx_j += -m_j Lambda_a . grad D
//                  (fifth derivatives), no atomics
//   3. k_adj_down, level by level from the root: mu~_c = mu_c + M2M^T mu~_parent; then k_adj_moments,
//      per leaf: the transposed P2M onto the leaf's particles.
// Ewald kinds use the periodic correction's derivatives in the same places.  Per-term arithmetic is
// FP32 like the forward; sums are FP64.  The multipole adjoints mu are accumulated with FP64
// atomics (one per source node and item), so their summation order varies between runs; Lambda and
// the particle gradients are not.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tg_adjoint_math.hpp"
#include "tg_gravity.h"
#include "tg_local.hpp"
#include "tg_solver.hpp"

namespace tg {

__constant__ const float* c_adj_ewt;  // the device's tabulated correction (tg_kernels.hip)

namespace {

using adj::add;
using adj::merge;

hipError_t init_adjoint_tables(int device) {
	const float* table = ewald_table_on(device);
	if (!table) return hipErrorNotInitialized;
	return hipMemcpyToSymbol(HIP_SYMBOL(c_adj_ewt), &table, sizeof(table));
}

// ---- compile-time contraction helpers ---------------------------------------------------------
// multipole slot of expansion slot K of order 2 or 3 (the multipole has no order 1)
constexpr int mslot(int K) { return K < 10 ? 1 + (K - 4) : 7 + (K - 10); }
constexpr float mfac(int K) { return K < 10 ? ef2(K - 4) : -ef3(K - 10); }  // M2P/M2L sign and expansion factor

// forces a constant index to be folded at compile time (the slot searches are loops)
template <int V>
constexpr int cx = V;

template <int... S, class F>
__device__ __forceinline__ void for_seq(std::integer_sequence<int, S...>, F&& f) {
	(f(std::integral_constant<int, S>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void for_n(F&& f) {
	for_seq(std::make_integer_sequence<int, N>{}, f);
}

// M2P acceleration at offset with derivative tensor D (the forward's m2p):
//   g_a = -sum_K fac_K M_K D_{K u a}, K over {0} u order-2 u order-3 slots, fac_0 = 1
// its derivative along b adds index b; its transpose in M contracts with w.
// returns dg/dx_b contracted with w (3 values) -- needs D to order 5
__device__ __forceinline__ void m2p_sink_adjoint(float* gx, const float* M, const float* D, const float* w) {
	for_n<3>([&](auto b_) {
		constexpr int b = decltype(b_)::value;
		float acc = 0.0f;
		for_n<3>([&](auto a_) {
			constexpr int a = decltype(a_)::value;
			float t = -M[0] * D[cx<add(add(0, a), b)>];
			for_n<16>([&](auto q_) {
				constexpr int K = 4 + decltype(q_)::value;  // order-2 and order-3 slots 4..19
				constexpr float f = -mfac(K);
				t = fmaf(f * M[cx<mslot(K)>], D[cx<add(add(K, a), b)>], t);
			});
			acc = fmaf(w[a], t, acc);
		});
		gx[b] += acc;
	});
}
// mu_K += dg/dM_K^T w
__device__ __forceinline__ void m2p_source_adjoint(float* mu, const float* D, const float* w) {
	float t0 = 0.0f;
	for_n<3>([&](auto a_) {
		constexpr int a = decltype(a_)::value;
		t0 = fmaf(-w[a], D[cx<add(0, a)>], t0);
	});
	mu[0] += t0;
	for_n<16>([&](auto q_) {
		constexpr int K = 4 + decltype(q_)::value;
		float t = 0.0f;
		for_n<3>([&](auto a_) {
			constexpr int a = decltype(a_)::value;
			t = fmaf(w[a], D[cx<add(K, a)>], t);
		});
		constexpr float f = -mfac(K);
		mu[cx<mslot(K)>] += f * t;
	});
}
// M2L transposed: mu_K += sum_t Lambda_t dL_t/dM_K for the forward's m2l (L_t += sum_K fac_K M_K D_{t u K}
// over |t| + |K| <= 4, fac_0 = 1)
__device__ __forceinline__ void m2l_adjoint(float* mu, const float* Lam, const float* D) {
	float t0 = 0.0f;
	for_n<35>([&](auto t_) {
		constexpr int t = decltype(t_)::value;
		t0 = fmaf(Lam[t], D[t], t0);
	});
	mu[0] += t0;
	for_n<16>([&](auto q_) {
		constexpr int K = 4 + decltype(q_)::value;
		float acc = 0.0f;
		for_n<35>([&](auto t_) {
			constexpr int t = decltype(t_)::value;
			if constexpr (adj::order_of(t) + adj::order_of(K) <= 4) acc = fmaf(Lam[t], D[cx<merge(t, K)>], acc);
		});
		constexpr float f = mfac(K);
		mu[cx<mslot(K)>] += f * acc;
	});
}

// spline-softened pair (src/gravity.cpp:231-287) in box units: F(d) = -d f(r), and f'(r)/r
__device__ __forceinline__ void spline_f(float r2, float h, float& f, float& fr) {
	if (r2 > h * h) {
		const float ri = __builtin_amdgcn_rsqf(r2);
		const float ri2 = ri * ri;
		f = ri2 * ri;
		fr = -3.0f * f * ri2;
		return;
	}
	const float r = sqrtf(r2), hi = 1.0f / h, q = r * hi;
	if (q < 0.5f) {
		const float hi3 = hi * hi * hi;
		f = fmaf(q * q, fmaf(32.0f, q, -192.0f / 5.0f), 32.0f / 3.0f) * hi3;
		fr = fmaf(96.0f, q, -384.0f / 5.0f) * hi3 * hi * hi;
	} else {
		const float q3 = q * q * q;
		const float P = fmaf(fmaf(fmaf(fmaf(-32.0f / 3.0f, q, 192.0f / 5.0f), q, -48.0f), q, 64.0f / 3.0f), q3, -1.0f / 15.0f);
		const float dP = q * q * fmaf(fmaf(fmaf(-64.0f, q, 192.0f), q, -192.0f), q, 64.0f);
		const float ri = 1.0f / r, ri3 = ri * ri * ri;
		f = P * ri3;
		fr = (dP * hi - 3.0f * P * ri) * ri3 * ri;
	}
}

struct AdjArgs {
	const int4* parts;      // tree order (x, y, z, mass bits)
	const float4* w;        // loss weights, tree order
	const MultiSrc* multis;
	const int4* nodepos;
	const double* dL;       // the forward's per-node expansions
	const int4* leaves;     // (node, pbegin, pend, path begin)
	const int* path;
	const int* level;
	double* grad;           // 4 per particle, tree order: dL/dx, dL/dy, dL/dz, dL/dm
	double* Lam;            // 35 per node
	double* mu;             // 17 per node
	float h;                // softening, box units
	int wrap;
	int maxdepth;
};

constexpr float PINV = 1.0f / 4294967296.0f;
__device__ __forceinline__ float sepf(int a, int b, bool wrap) {
	const unsigned d = unsigned(a) - unsigned(b);
	if (wrap) return float(int(d)) * PINV;
	return float(double(a) - double(b)) * PINV;
}

// sum of row[0..n) with four partial sums (short dependency chains)
__device__ __forceinline__ double sum4(const double* row, int n) {
	double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
	int k = 0;
	for (; k + 4 <= n; k += 4) {
		a0 += row[k];
		a1 += row[k + 1];
		a2 += row[k + 2];
		a3 += row[k + 3];
	}
	for (; k < n; k++) a0 += row[k];
	return (a0 + a1) + (a2 + a3);
}

// ---- 1. leaves -------------------------------------------------------------------------------
// L2P and the shifts: per leaf, writes each sink's gradient (the first term; later kernels add)
__global__ __launch_bounds__(64) void k_adj_l2p(AdjArgs A) {
	const int4 lf = A.leaves[blockIdx.x];
	const int t = threadIdx.x;
	const int depth = A.level[lf.x] + 1;
	const int ns = lf.z - lf.y;
	extern __shared__ double lds[];  // [35][64] per-sink lambda, [35] L then lambda of the leaf, [35][maxdepth]
	double* lam64 = lds;
	double* Ls = lds + 35 * 64;
	double* sh = Ls + 35;
	const int4 c = A.nodepos[lf.x];
	// the leaf's local expansion, as the forward forms it (sum over the path of shifted dL)
	double off[3] = {0.0, 0.0, 0.0};
	if (t < depth) {
		const int a = A.path[lf.w + t];
		const int4 X = A.nodepos[a];
		off[0] = double(int64_t(c.x) - int64_t(X.x)) * POS_INV;
		off[1] = double(int64_t(c.y) - int64_t(X.y)) * POS_INV;
		off[2] = double(int64_t(c.z) - int64_t(X.z)) * POS_INV;
		double La[35], Lt[35];
#pragma unroll
		for (int q = 0; q < 35; q++) La[q] = A.dL[35 * int64_t(a) + q];
		l2l(Lt, La, off);
#pragma unroll
		for (int q = 0; q < 35; q++) sh[q * A.maxdepth + t] = Lt[q];
	}
	__syncthreads();
	if (t < 35) Ls[t] = sum4(sh + t * A.maxdepth, depth);
	__syncthreads();
	const bool sink = t < ns;
	const int pi = lf.y + min(t, ns - 1);
	const int4 me = A.parts[pi];
	const float4 wv = A.w[pi];
	const float wi[3] = {sink ? wv.x : 0.0f, sink ? wv.y : 0.0f, sink ? wv.z : 0.0f};
	double gx[3] = {0.0, 0.0, 0.0};
	{
		const double d[3] = {double(float(double(centre_offset(me.x, c.x, A.wrap)) * POS_INV)),
				double(float(double(centre_offset(me.y, c.y, A.wrap)) * POS_INV)),
				double(float(double(centre_offset(me.z, c.z, A.wrap)) * POS_INV))};
		double W[35];
		taylor_weights(W, d);
		// g_a = -sum_{|s|<=3} L_{s u a} W_s  ->  dg_a/dd_b = -sum_{|s|<=2} L_{s u a u b} W_s
		for_n<3>([&](auto a_) {
			constexpr int a = decltype(a_)::value;
			for_n<3>([&](auto b_) {
				constexpr int b = decltype(b_)::value;
				double hab = 0.0;
				for_n<10>([&](auto s_) {
					constexpr int s = decltype(s_)::value;
					hab = fma(Ls[Merge<Merge<s, 1 + a>::value, 1 + b>::value], W[s], hab);
				});
				gx[b] -= double(wi[a]) * hab;
			});
		});
		// lambda_t = dg/dL_t^T w = -sum_a w_a W_{t minus a}
		double lam[35];
#pragma unroll
		for (int q = 0; q < 35; q++) lam[q] = 0.0;
		for_n<20>([&](auto s_) {
			constexpr int s = decltype(s_)::value;
			for_n<3>([&](auto a_) {
				constexpr int a = decltype(a_)::value;
				lam[Merge<s, 1 + a>::value] -= double(wi[a]) * W[s];
			});
		});
#pragma unroll
		for (int q = 0; q < 35; q++) lam64[q * 64 + t] = lam[q];
	}
	__syncthreads();
	// lambda of the leaf; k_adj_up carries it to the ancestors
	if (t < 35) A.Lam[35 * int64_t(lf.x) + t] = sum4(lam64 + t * 64, ns);
	if (sink) {
		double* o = A.grad + 4 * int64_t(pi);
		o[0] = gx[0];
		o[1] = gx[1];
		o[2] = gx[2];
		o[3] = 0.0;
	}
}

// leaf work items of one kind (the forward's: first sink, count [| upper bits], list begin, end)
struct AdjItems {
	const int4* items;
	const int* list;
	int nitems;
};

// PP direct: one item per group, lane = sink, sources in LDS tiles of 64 with their weights.
// The lists are symmetric, so particle k's own list carries its source-side terms as well:
//   dL/dx_k += sum_j J(d_kj) (m_j w_k - m_k w_j),  J v = -f v - (f'/r)(d.v) d,
//   dL/dm_k += sum_j f(d_kj) (d_kj . w_j)
__global__ __launch_bounds__(64) void k_adj_pp(AdjArgs A, AdjItems I) {
	const int4 it = I.items[blockIdx.x];
	const int t = threadIdx.x;
	const int ns = it.y & PP_COUNT_MASK;
	// sinks padded to a power of two nsp; the P = 64 / nsp lane groups (phases) split each tile's
	// sources, so leaves of 32 or fewer particles keep every lane busy
	const int lg = ns <= 1 ? 0 : 32 - __clz(ns - 1);
	const int nsp = 1 << lg;
	const int P = 64 >> lg;
	const int si = t & (nsp - 1);
	const int ph = t >> lg;
	const int near_len = int(unsigned(it.y) >> PP_NEAR_SHIFT);
	const int near_end = near_len == PP_NEAR_ALL ? INT_MAX : it.z + near_len;
	const int pi = it.x + min(si, ns - 1);
	const int4 me = A.parts[pi];
	const float4 wv = A.w[pi];
	const float wi[3] = {wv.x, wv.y, wv.z};
	const float mi = __int_as_float(me.w);
	__shared__ int4 tp[64];
	__shared__ float4 tw[64];
	double gx[3] = {0.0, 0.0, 0.0}, gm = 0.0;
	for (int base = it.z; base < it.w; base += 64) {
		const int n = min(64, it.w - base);
		__syncthreads();
		if (t < n) {
			const int j = I.list[base + t];
			tp[t] = A.parts[j];
			tw[t] = A.w[j];
		}
		__syncthreads();
		float fx[3] = {0.f, 0.f, 0.f}, fm = 0.f;
		// tiles inside the list's near prefix may hold pairs closer than h (spline); later tiles
		// hold only pairs at least h apart (the forward's split)
		const bool near = base < near_end;
		for (int k = ph; k < n; k += P) {
			const int4 s = tp[k];
			const float4 ws = tw[k];
			const float mj = __int_as_float(s.w);
			const float d[3] = {sepf(me.x, s.x, A.wrap), sepf(me.y, s.y, A.wrap), sepf(me.z, s.z, A.wrap)};
			const float v[3] = {mj * wi[0] - mi * ws.x, mj * wi[1] - mi * ws.y, mj * wi[2] - mi * ws.z};
			const float r2 = fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
			float f, fr;
			if (near) {
				spline_f(r2, A.h, f, fr);
			} else {
				const float ri = __builtin_amdgcn_rsqf(r2);
				const float ri2 = ri * ri;
				f = ri2 * ri;
				fr = -3.0f * f * ri2;
			}
			const float dv = fmaf(d[2], v[2], fmaf(d[1], v[1], d[0] * v[0]));
			const float dw = fmaf(d[2], ws.z, fmaf(d[1], ws.y, d[0] * ws.x));
			const float frdv = fr * dv;
#pragma unroll
			for (int b = 0; b < 3; b++) fx[b] = fmaf(-f, v[b], fmaf(-frdv, d[b], fx[b]));
			fm = fmaf(f, dw, fm);
		}
#pragma unroll
		for (int b = 0; b < 3; b++) gx[b] += fx[b];
		gm += fm;
	}
	for (int off = nsp; off < 64; off <<= 1) {
#pragma unroll
		for (int b = 0; b < 3; b++) gx[b] += __shfl_xor(gx[b], off, 64);
		gm += __shfl_xor(gm, off, 64);
	}
	if (ph == 0 && si < ns) {
		double* o = A.grad + 4 * int64_t(pi);
		o[0] += gx[0];
		o[1] += gx[1];
		o[2] += gx[2];
		o[3] += gm;
	}
}

// PP Ewald: the correction's F = -D1, J = -D2, per pair from the tabulated correction (its
// real-space images, Fourier terms and n = 0 term together, as the forward's k_pp_ewald reads them):
//   dL/dx_k += sum_j J(d_kj) (m_j w_k - m_k w_j),  dL/dm_k += sum_j w_j . D1(d_kj)
__global__ __launch_bounds__(64) void k_adj_pp_ewald(AdjArgs A, AdjItems I) {
	const int4 it = I.items[blockIdx.x];
	const int t = threadIdx.x;
	const int ns = it.y & PP_COUNT_MASK;
	const bool sink = t < ns;
	const int pi = it.x + min(t, ns - 1);
	const int4 me = A.parts[pi];
	const float4 wv = A.w[pi];
	const float wi[3] = {wv.x, wv.y, wv.z};
	const float mi = __int_as_float(me.w);
	__shared__ int4 tp[64];
	__shared__ float4 tw[64];
	double gx[3] = {0.0, 0.0, 0.0}, gm = 0.0;
	for (int base = it.z; base < it.w; base += 64) {
		const int n = min(64, it.w - base);
		__syncthreads();
		if (t < n) {
			const int j = I.list[base + t];
			tp[t] = A.parts[j];
			tw[t] = A.w[j];
		}
		__syncthreads();
		float fx[3] = {0.f, 0.f, 0.f}, fm = 0.f;
		for (int k = 0; k < n; k++) {
			const int4 s = tp[k];
			const float4 ws = tw[k];
			const float mj = __int_as_float(s.w);
			const float x[3] = {sepf(me.x, s.x, true), sepf(me.y, s.y, true), sepf(me.z, s.z, true)};
			const float r2 = fmaf(x[2], x[2], fmaf(x[1], x[1], x[0] * x[0]));
			if (!(r2 > 1e-12f)) continue;  // the reference's constant below r = 1e-6: no derivatives
			float D[10];
			green_ewald_tab<10>(D, x, c_adj_ewt);
			const float v[3] = {mj * wi[0] - mi * ws.x, mj * wi[1] - mi * ws.y, mj * wi[2] - mi * ws.z};
			for_n<3>([&](auto b_) {
				constexpr int b = decltype(b_)::value;
				float jv = 0.0f;
				for_n<3>([&](auto a_) {
					constexpr int a = decltype(a_)::value;
					jv = fmaf(-D[cx<add(add(0, a), b)>], v[a], jv);
				});
				fx[b] += jv;
			});
			fm = fmaf(ws.x, D[1], fmaf(ws.y, D[2], fmaf(ws.z, D[3], fm)));
		}
#pragma unroll
		for (int b = 0; b < 3; b++) gx[b] += fx[b];
		gm += fm;
	}
	if (!sink) return;
	double* o = A.grad + 4 * int64_t(pi);
	o[0] += gx[0];
	o[1] += gx[1];
	o[2] += gx[2];
	o[3] += gm;
}

// PC, sink side: lane = sink, over the item's multipoles, dL/dx_i += w . dg/dx_i (fifth derivatives)
template <bool EWALD>
__global__ __launch_bounds__(64) void k_adj_pc_sink(AdjArgs A, AdjItems I) {
	const int4 it = I.items[blockIdx.x];
	const int t = threadIdx.x;
	const int ns = it.y & PP_COUNT_MASK;
	if (t >= ns) return;
	const int pi = it.x + t;
	const int4 me = A.parts[pi];
	const float4 wv = A.w[pi];
	const float wi[3] = {wv.x, wv.y, wv.z};
	float gx[3] = {0.f, 0.f, 0.f};
	const bool wrapx = EWALD || A.wrap;
	for (int e = it.z; e < it.w; e++) {
		const MultiSrc& ms = A.multis[I.list[e]];
		float M[17];
#pragma unroll
		for (int q = 0; q < 17; q++) M[q] = ms.m[q];
		const float d[3] = {sepf(me.x, ms.x[0], wrapx), sepf(me.y, ms.x[1], wrapx), sepf(me.z, ms.x[2], wrapx)};
		float D[adj::NS5];
		if (EWALD)
			green_ewald_tab<adj::NS5>(D, d, c_adj_ewt);
		else
			adj::green_direct_n<adj::NS5>(D, d);
		m2p_sink_adjoint(gx, M, D, wi);
	}
	double* o = A.grad + 4 * int64_t(pi);
	o[0] += double(gx[0]);
	o[1] += double(gx[1]);
	o[2] += double(gx[2]);
}

// PC, source side: lane = multipole of the item's list, over the item's sinks (staged in LDS):
// mu_c += dg/dM_c^T w (fourth derivatives), one FP64 atomic per slot and entry
// PC source side by source node: mu_c += sum over the sinks i of every leaf whose PC list holds c of
// dg_i/dM_c^T w_i.  One work group per item (node c, a run of at most 16 sink leaves); the lanes
// cover 64 / L leaves at a time, L = 2^lgL lanes per leaf (L >= the largest leaf), one sink each;
// the 17 sums are reduced over the wave once per item and added with one atomic each.
struct AdjPCArgs {
	const int4* items;   // (node c, first sink leaf, end, 0)
	const int2* leaves;  // sink leaves: (first particle, count)
	const int4* parts;
	const float4* w;
	const MultiSrc* multis;
	double* mu;
	int lgL;
	int wrap;
};

template <bool EWALD>
__global__ __launch_bounds__(64) void k_adj_pc_t(AdjPCArgs A) {
	const int4 it = A.items[blockIdx.x];
	const int t = threadIdx.x;
	const int Q = 64 >> A.lgL;         // leaves per pass
	const int qq = t >> A.lgL;         // this lane's leaf within the pass
	const int o = t & ((1 << A.lgL) - 1);  // and its sink within the leaf
	const MultiSrc& ms = A.multis[it.x];
	const int cx0 = ms.x[0], cx1 = ms.x[1], cx2 = ms.x[2];
	const bool wrapx = EWALD || A.wrap;
	float mu[17];
#pragma unroll
	for (int q = 0; q < 17; q++) mu[q] = 0.0f;
	for (int k = it.y + qq; k - qq < it.z; k += Q) {
		if (k >= it.z) continue;
		const int2 lf = A.leaves[k];
		if (o >= lf.y) continue;
		const int pi = lf.x + o;
		const int4 me = A.parts[pi];
		const float4 wv = A.w[pi];
		const float wi[3] = {wv.x, wv.y, wv.z};
		const float d[3] = {sepf(me.x, cx0, wrapx), sepf(me.y, cx1, wrapx), sepf(me.z, cx2, wrapx)};
		float D[35];
		if (EWALD)
			green_ewald_tab<35>(D, d, c_adj_ewt);
		else
			adj::green_direct_n<35>(D, d);
		m2p_source_adjoint(mu, D, wi);
	}
	__shared__ float red[17][65];
#pragma unroll
	for (int q = 0; q < 17; q++) red[q][t] = mu[q];
	__syncthreads();
	if (t < 17) {
		double acc = 0.0;
		for (int k = 0; k < 64; k++) acc += double(red[t][k]);
		atomicAdd(A.mu + 17 * int64_t(it.x) + t, acc);
	}
}

// ---- 2. node lists -------------------------------------------------------------------------------
// CP by source leaf: the forward adds m_j D(X_a - x_j) to the expansion of every node a whose CP
// list holds leaf B's range, so  dL/dm_j = sum_a Lambda_a . D(X_a - x_j)  and
// dL/dx_j = -m_j sum_a Lambda_a . grad D(X_a - x_j)  (fifth derivatives).  One work group per source
// leaf (item: first particle, count, sink-node list range); lanes are the leaf's particles, and the
// P = 64 / nsp lane groups split the sink nodes, staged 64 at a time in LDS.  Every particle belongs
// to one item, so the gradient is added without atomics.
struct AdjCPArgs {
	const int4* items;
	const int* sinks;
	const int4* nodepos;
	const int4* parts;
	const double* Lam;
	double* grad;
	int wrap;
};

template <bool EWALD>
__global__ __launch_bounds__(64) void k_adj_cp_t(AdjCPArgs A) {
	const int4 it = A.items[blockIdx.x];
	const int t = threadIdx.x;
	const int ns = it.y;
	const int lg = ns <= 1 ? 0 : 32 - __clz(ns - 1);
	const int nsp = 1 << lg;
	const int P = 64 >> lg;
	const int si = t & (nsp - 1);
	const int ph = t >> lg;
	const int pi = it.x + min(si, ns - 1);
	const int4 me = A.parts[pi];
	const bool wrapx = EWALD || A.wrap;
	__shared__ float lam[64][36];
	__shared__ int4 xs[64];
	double gm = 0.0, g3[3] = {0.0, 0.0, 0.0};
	for (int base = it.z; base < it.w; base += 64) {
		const int n = min(64, it.w - base);
		__syncthreads();
		if (t < n) {
			const int a = A.sinks[base + t];
			xs[t] = A.nodepos[a];
#pragma unroll
			for (int q = 0; q < 35; q++) lam[t][q] = float(A.Lam[35 * int64_t(a) + q]);
		}
		__syncthreads();
		float fm = 0.0f, f3[3] = {0.f, 0.f, 0.f};
		for (int k = ph; k < n; k += P) {
			const int4 X = xs[k];
			const float d[3] = {sepf(X.x, me.x, wrapx), sepf(X.y, me.y, wrapx), sepf(X.z, me.z, wrapx)};
			float D[adj::NS5];
			if (EWALD)
				green_ewald_tab<adj::NS5>(D, d, c_adj_ewt);
			else
				adj::green_direct_n<adj::NS5>(D, d);
			for_n<35>([&](auto t_) {
				constexpr int tt = decltype(t_)::value;
				const float l = lam[k][tt];
				fm = fmaf(l, D[tt], fm);
				for_n<3>([&](auto b_) {
					constexpr int bb = decltype(b_)::value;
					f3[bb] = fmaf(l, D[cx<add(tt, bb)>], f3[bb]);
				});
			});
		}
		gm += fm;
#pragma unroll
		for (int q = 0; q < 3; q++) g3[q] += f3[q];
	}
	for (int off = nsp; off < 64; off <<= 1) {
		gm += __shfl_xor(gm, off, 64);
#pragma unroll
		for (int q = 0; q < 3; q++) g3[q] += __shfl_xor(g3[q], off, 64);
	}
	if (ph == 0 && si < ns) {
		const double mj = double(__int_as_float(me.w));
		double* o = A.grad + 4 * int64_t(pi);
		o[0] -= mj * g3[0];
		o[1] -= mj * g3[1];
		o[2] -= mj * g3[2];
		o[3] += gm;
	}
}

// CC by source: the entries (source c, sink node a) sorted by source and cut into waves of <= 64; each
// lane forms M2L(., X_a - X_c)^T Lambda_a, then 17 lanes add the wave's entries source by source and
// issue one atomic per source segment and slot
struct AdjCCArgs {
	const int2* waves;
	const int* src;
	const int* snk;
	const int4* nodepos;
	const MultiSrc* multis;
	const double* Lam;
	double* mu;
	int wrap;
};

template <bool EWALD>
__global__ __launch_bounds__(64) void k_adj_cc_t(AdjCCArgs A) {
	const int2 wv = A.waves[blockIdx.x];
	const int t = threadIdx.x;
	const int j = wv.x + t;
	const int cnt = wv.y - wv.x;
	__shared__ float red[17][65];
	__shared__ int sc[64];
	float mu[17];
#pragma unroll
	for (int q = 0; q < 17; q++) mu[q] = 0.0f;
	int c = -1;
	if (j < wv.y) {
		c = A.src[j];
		const int a = A.snk[j];
		const int4 X = A.nodepos[a];
		const MultiSrc& ms = A.multis[c];
		float Lam[35];
#pragma unroll
		for (int q = 0; q < 35; q++) Lam[q] = float(A.Lam[35 * int64_t(a) + q]);
		const bool wrapx = EWALD || A.wrap;
		const float d[3] = {sepf(X.x, ms.x[0], wrapx), sepf(X.y, ms.x[1], wrapx), sepf(X.z, ms.x[2], wrapx)};
		float D[35];
		if (EWALD)
			green_ewald_tab<35>(D, d, c_adj_ewt);
		else
			adj::green_direct_n<35>(D, d);
		m2l_adjoint(mu, Lam, D);
	}
#pragma unroll
	for (int q = 0; q < 17; q++) red[q][t] = mu[q];
	sc[t] = c;
	__syncthreads();
	// segments of one source from a ballot of their first entries; lanes 0..16 add one slot each
	const bool first = t < cnt && (t == 0 || sc[t] != sc[t - 1]);
	unsigned long long starts = __ballot(first);
	if (t < 17) {
		while (starts) {
			const int b = __ffsll(starts) - 1;
			starts &= starts - 1;
			const int e = starts ? __ffsll(starts) - 1 : cnt;
			double a0 = 0.0, a1 = 0.0;
			int k = b;
			for (; k + 2 <= e; k += 2) {
				a0 += double(red[t][k]);
				a1 += double(red[t][k + 1]);
			}
			if (k < e) a0 += double(red[t][k]);
			atomicAdd(A.mu + 17 * int64_t(sc[b]) + t, a0 + a1);
		}
	}
}

// ---- up-sweep of Lambda ------------------------------------------------------------------------------
// The forward gives leaf B the expansion sum_a S(X_B - X_a) dL_a over its path, so dL/d(dL_a) =
// sum over leaves B below a of S(X_B - X_a)^T lambda_B.  Shifts compose exactly
// (S(X_B - X_a) = S(X_B - X_c) S(X_c - X_a) for a child c of a on the way), hence level by level,
// deepest first:  Lambda_a = sum_children S(X_c - X_a)^T Lambda_c,  with Lambda_B = lambda_B at the
// leaves.  One lane per (node, child); the two children of a node sit on adjacent lanes and are
// added with one exchange.  No atomics: every Lambda is written once.
__global__ __launch_bounds__(64) void k_adj_up(const int2* __restrict__ pairs, int npairs, const int4* __restrict__ nodepos,
		double* __restrict__ Lam) {
	const int j = blockIdx.x * 64 + threadIdx.x;
	const int2 pc = j < npairs ? pairs[j] : make_int2(-1, -1);
	double out[35];
#pragma unroll
	for (int q = 0; q < 35; q++) out[q] = 0.0;
	if (pc.y >= 0) {
		const int4 Xa = nodepos[pc.x], Xc = nodepos[pc.y];
		// (S^T l)_u = sum_{t u s = u} l_t W_s(X_c - X_a), W the Taylor weights of the offset
		const double off[3] = {double(int64_t(Xc.x) - int64_t(Xa.x)) * POS_INV, double(int64_t(Xc.y) - int64_t(Xa.y)) * POS_INV,
				double(int64_t(Xc.z) - int64_t(Xa.z)) * POS_INV};
		double W[35];
		taylor_weights(W, off);
		const double* lc = Lam + 35 * int64_t(pc.y);
		for_n<35>([&](auto t_) {
			constexpr int tt = decltype(t_)::value;
			const double lt = lc[tt];
			for_n<35>([&](auto s_) {
				constexpr int s = decltype(s_)::value;
				if constexpr (Merge<tt, s>::valid) out[Merge<tt, s>::value] = fma(lt, W[s], out[Merge<tt, s>::value]);
			});
		});
	}
#pragma unroll
	for (int q = 0; q < 35; q++) out[q] += __shfl_xor(out[q], 1, 64);
	if ((j & 1) == 0 && j < npairs) {
		double* dst = Lam + 35 * int64_t(pc.x);
#pragma unroll
		for (int q = 0; q < 35; q++) dst[q] = out[q];
	}
}

// ---- 3. moments ----------------------------------------------------------------------------------
// multipole<float>::operator>>= transposed (tigergrav/multipole.hpp:414-433, restated in the host
// tree): M3'_pql = M3 + M0 Y_p Y_q Y_l + M2_pq Y_l + M2_ql Y_p + M2_lp Y_q, M2'_pq = M2 + M0 Y_p Y_q
__device__ __forceinline__ void m2m_adjoint(float* mu, const float* mup, const float* Y) {
	mu[0] = mup[0];
#pragma unroll
	for (int s = 0; s < 6; s++) mu[1 + s] = mup[1 + s];
#pragma unroll
	for (int s = 0; s < 10; s++) mu[7 + s] = mup[7 + s];
#pragma unroll
	for (int s = 0; s < 6; s++) mu[0] = fmaf(mup[1 + s] * Y[T2[s].i[0]], Y[T2[s].i[1]], mu[0]);
#pragma unroll
	for (int s = 0; s < 10; s++) {
		const int p = T3[s].i[0], q = T3[s].i[1], l = T3[s].i[2];
		const float u = mup[7 + s];
		mu[0] = fmaf(u * Y[p] * Y[q], Y[l], mu[0]);
		mu[1 + s2(p, q)] = fmaf(u, Y[l], mu[1 + s2(p, q)]);
		mu[1 + s2(q, l)] = fmaf(u, Y[p], mu[1 + s2(q, l)]);
		mu[1 + s2(l, p)] = fmaf(u, Y[q], mu[1 + s2(l, p)]);
	}
}

// The forward forms a parent's moments from its children's (M2M), so the multipole adjoint of a
// node is its own mu plus the M2M transpose of its parent's:  mu~_c = mu_c + M2M(X_c - X_a)^T mu~_a,
// level by level from the root (the up-sweep's (node, child) pairs in reverse level order); in place.
__global__ __launch_bounds__(64) void k_adj_down(const int2* __restrict__ pairs, int npairs, const int4* __restrict__ nodepos,
		double* __restrict__ mu) {
	const int j = blockIdx.x * 64 + threadIdx.x;
	if (j >= npairs) return;
	const int2 pc = pairs[j];
	if (pc.y < 0) return;
	const int4 Xa = nodepos[pc.x], Xc = nodepos[pc.y];
	const float Y[3] = {float(double(int64_t(Xc.x) - int64_t(Xa.x)) * POS_INV), float(double(int64_t(Xc.y) - int64_t(Xa.y)) * POS_INV),
			float(double(int64_t(Xc.z) - int64_t(Xa.z)) * POS_INV)};
	float m[17], n[17];
#pragma unroll
	for (int q = 0; q < 17; q++) m[q] = float(mu[17 * int64_t(pc.x) + q]);
	m2m_adjoint(n, m, Y);
	double* dst = mu + 17 * int64_t(pc.y);
#pragma unroll
	for (int q = 0; q < 17; q++) dst[q] += double(n[q]);
}

// P2M transposed onto a leaf's particles, from the leaf's swept multipole adjoint
__global__ __launch_bounds__(64) void k_adj_moments(AdjArgs A) {
	const int4 lf = A.leaves[blockIdx.x];
	const int t = threadIdx.x;
	__shared__ float mus[17];
	if (t < 17) mus[t] = float(A.mu[17 * int64_t(lf.x) + t]);
	__syncthreads();
	const int ns = lf.z - lf.y;
	if (t >= ns) return;
	const int pi = lf.y + t;
	const int4 me = A.parts[pi];
	const int4 c = A.nodepos[lf.x];
	const float mj = __int_as_float(me.w);
	// P2M (part_vect_multipole_info): M2_pq += m d_p d_q, M3_pql += m d_p d_q d_l, d = float(x - X)
	const float d[3] = {float(double(centre_offset(me.x, c.x, A.wrap)) * POS_INV), float(double(centre_offset(me.y, c.y, A.wrap)) * POS_INV),
			float(double(centre_offset(me.z, c.z, A.wrap)) * POS_INV)};
	float gm = mus[0], g3[3] = {0.f, 0.f, 0.f};
#pragma unroll
	for (int s = 0; s < 6; s++) {
		const int p = T2[s].i[0], q = T2[s].i[1];
		const float u = mus[1 + s];
		gm = fmaf(u * d[p], d[q], gm);
		g3[p] = fmaf(u, d[q], g3[p]);
		g3[q] = fmaf(u, d[p], g3[q]);
	}
#pragma unroll
	for (int s = 0; s < 10; s++) {
		const int p = T3[s].i[0], q = T3[s].i[1], l = T3[s].i[2];
		const float u = mus[7 + s];
		gm = fmaf(u * d[p] * d[q], d[l], gm);
		g3[p] = fmaf(u * d[q], d[l], g3[p]);
		g3[q] = fmaf(u * d[p], d[l], g3[q]);
		g3[l] = fmaf(u * d[p], d[q], g3[l]);
	}
	double* o = A.grad + 4 * int64_t(pi);
	o[0] += double(mj * g3[0]);
	o[1] += double(mj * g3[1]);
	o[2] += double(mj * g3[2]);
	o[3] += double(gm);
}

__global__ void k_adj_zero(double* p, int64_t n) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0.0;
}

}  // namespace

#define ACHK(call)                                                   \
	do {                                                             \
		hipError_t e_ = (call);                                      \
		if (e_ != hipSuccess) {                                      \
			err = std::string(#call) + ": " + hipGetErrorString(e_); \
			return e_ == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP;  \
		}                                                            \
	} while (0)

int Solver::adjoint(const float* w_in, double* grad_out, float* ms) {
	if (device < 0 || n == 0 || tree.nodes.empty()) {
		err = "adjoint needs a device solver with particles";
		return TG_EINVAL;
	}
	if (parts_per_node > 64) {
		err = "the adjoint handles leaves of at most 64 particles";
		return TG_EINVAL;
	}
	for (int32_t lf : tree.leaves)
		if (tree.nodes[size_t(lf)].pend - tree.nodes[size_t(lf)].pbegin > 64) {
			// only more coincident particles than a leaf holds make one (the tree's depth limit)
			err = "the adjoint handles leaves of at most 64 particles; a leaf at the tree's depth limit holds more";
			return TG_EINVAL;
		}
	if (!adj_pp_symmetric) {
		err = "the adjoint needs symmetric PP lists";
		return TG_EINVAL;
	}
	ACHK(hipSetDevice(device));
	if (!adj_lists_ready) {
		{
			const int rc = host_lists();  // the by-source transposes read the CC list values
			if (rc != TG_OK) return rc;
		}
		// CC entries by source: (source, sink node) pairs of the direct and Ewald lists, sorted by source
		auto put_v = [&](DeviceArray& d, const void* h, size_t bytes) -> hipError_t {
			hipError_t r = d.alloc(std::max<size_t>(bytes, 16));
			if (r != hipSuccess || bytes == 0) return r;
			return hipMemcpy(d.p, h, bytes, hipMemcpyHostToDevice);
		};
		for (int e = 0; e < 2; e++) {
			const hvec<int64_t>& ptr = e ? tree.lists.cc_e_ptr : tree.lists.cc_d_ptr;
			const hvec<int32_t>& v = e ? tree.lists.cc_e : tree.lists.cc_d;
			std::vector<std::pair<int32_t, int32_t>> pr;
			for (size_t a = 0; a + 1 < ptr.size(); a++)
				for (int64_t q = ptr[a]; q < ptr[a + 1]; q++) pr.emplace_back(v[q], int32_t(a));
			std::sort(pr.begin(), pr.end());
			std::vector<int> src(pr.size()), snk(pr.size());
			std::vector<int2> waves;
			for (size_t q = 0; q < pr.size(); q++) {
				src[q] = pr[q].first;
				snk[q] = pr[q].second;
			}
			for (size_t q = 0; q < pr.size(); q += 64) waves.push_back(make_int2(int(q), int(std::min(pr.size(), q + 64))));
			ACHK(put_v(d_adj_cct_src[e], src.data(), src.size() * sizeof(int)));
			ACHK(put_v(d_adj_cct_snk[e], snk.data(), snk.size() * sizeof(int)));
			ACHK(put_v(d_adj_cct_waves[e], waves.data(), waves.size() * sizeof(int2)));
			adj_cct_nwaves[e] = int(waves.size());
		}
		// CP entries by source leaf (direct, Ewald): every CP range is a whole leaf's particles
		std::vector<int32_t> leaf_at(size_t(n) + 1, -1);
		for (int32_t lf : tree.leaves)
			if (!tree.nodes[lf].empty()) leaf_at[size_t(tree.nodes[lf].pbegin)] = lf;
		for (int e = 0; e < 2; e++) {
			const hvec<int64_t>& ptr = e ? tree.lists.cp_e_ptr : tree.lists.cp_d_ptr;
			const hvec<int32_t>& v = e ? tree.lists.cp_e : tree.lists.cp_d;
			std::vector<std::pair<int32_t, int32_t>> pr;  // (source leaf first particle, sink node)
			for (size_t a = 0; a + 1 < ptr.size(); a++)
				for (int64_t q = ptr[a]; q < ptr[a + 1]; q += 2) {
					const int32_t b = v[q], en = v[q + 1];
					if (en <= b) continue;
					const int32_t lf = leaf_at[size_t(b)];
					if (lf < 0 || tree.nodes[lf].pend != en || en - b > 64) {
						err = "the adjoint needs CP sources that are whole leaves of at most 64 particles";
						return TG_EINVAL;
					}
					pr.emplace_back(b, int32_t(a));
				}
			std::sort(pr.begin(), pr.end());
			std::vector<int> sinks(pr.size());
			std::vector<int4> items;
			for (size_t q = 0; q < pr.size(); q++) {
				sinks[q] = pr[q].second;
				if (q == 0 || pr[q].first != pr[q - 1].first) {
					const int32_t b = pr[q].first, lf = leaf_at[size_t(b)];
					items.push_back(make_int4(b, tree.nodes[lf].pend - b, int(q), int(q)));
				}
				items.back().w = int(q + 1);
			}
			ACHK(put_v(d_adj_cpt_sinks[e], sinks.data(), sinks.size() * sizeof(int)));
			ACHK(put_v(d_adj_cpt_items[e], items.data(), items.size() * sizeof(int4)));
			adj_cpt_nitems[e] = int(items.size());
		}
		// PC entries by source node (direct, Ewald): runs of at most 16 sink leaves per item
		int maxleaf = 1;
		for (int32_t lf : tree.leaves) maxleaf = std::max(maxleaf, int(tree.nodes[lf].pend - tree.nodes[lf].pbegin));
		adj_pc_lgL = 0;
		while ((1 << adj_pc_lgL) < maxleaf) adj_pc_lgL++;
		if (adj_pc_lgL > 6) {
			err = "the adjoint handles leaves of at most 64 particles";
			return TG_EINVAL;
		}
		for (int e = 0; e < 2; e++) {
			const hvec<int64_t>& ptr = e ? tree.lists.pc_e_ptr : tree.lists.pc_d_ptr;
			const hvec<int32_t>& v = e ? tree.lists.pc_e : tree.lists.pc_d;
			std::vector<std::pair<int32_t, int32_t>> pr;  // (source node, sink leaf)
			for (int32_t lf : tree.leaves) {
				if (tree.nodes[lf].empty() || size_t(lf) + 1 >= ptr.size()) continue;
				for (int64_t q = ptr[lf]; q < ptr[lf + 1]; q++) pr.emplace_back(v[q], lf);
			}
			std::sort(pr.begin(), pr.end());
			std::vector<int2> sl(pr.size());
			std::vector<int4> items;
			for (size_t q = 0; q < pr.size(); q++) {
				const Node& nd = tree.nodes[pr[q].second];
				sl[q] = make_int2(nd.pbegin, nd.pend - nd.pbegin);
				if (q == 0 || pr[q].first != pr[q - 1].first || items.back().z - items.back().y >= 16)
					items.push_back(make_int4(pr[q].first, int(q), int(q), 0));
				items.back().z = int(q + 1);
			}
			ACHK(put_v(d_adj_pct_leaves[e], sl.data(), sl.size() * sizeof(int2)));
			ACHK(put_v(d_adj_pct_items[e], items.data(), items.size() * sizeof(int4)));
			adj_pct_nitems[e] = int(items.size());
		}
		// (node, child) pairs of the internal nodes, deepest level first; an empty child gets -1
		std::vector<std::vector<int2>> by_level(size_t(tree.max_level) + 1);
		for (size_t i = 0; i < tree.nodes.size(); i++) {
			const Node& nd = tree.nodes[i];
			if (nd.leaf()) continue;
			for (int c = 0; c < 2; c++) {
				const int ch = nd.child[c];
				by_level[size_t(nd.level)].push_back(make_int2(int(i), ch >= 0 && !tree.nodes[size_t(ch)].empty() ? ch : -1));
			}
		}
		std::vector<int2> pairs;
		adj_up_ptr.assign(1, 0);
		for (int l = tree.max_level; l >= 0; l--) {
			pairs.insert(pairs.end(), by_level[size_t(l)].begin(), by_level[size_t(l)].end());
			if (int(pairs.size()) > adj_up_ptr.back()) adj_up_ptr.push_back(int(pairs.size()));
		}
		ACHK(d_adj_up.alloc(std::max<size_t>(pairs.size() * sizeof(int2), 16)));
		if (!pairs.empty()) ACHK(hipMemcpy(d_adj_up.p, pairs.data(), pairs.size() * sizeof(int2), hipMemcpyHostToDevice));
		adj_lists_ready = true;
	}
	static thread_local int tables_ready = -1;
	if (tables_ready != device) {
		ACHK(init_adjoint_tables(device));
		tables_ready = device;
	}
	const int64_t nn = int64_t(tree.nodes.size());
	std::vector<float4> w(n);
	for (int64_t i = 0; i < n; i++) {
		const int64_t o = tree.order[i];
		w[i] = make_float4(w_in[3 * o], w_in[3 * o + 1], w_in[3 * o + 2], 0.0f);
	}
	ACHK(d_adj_w.alloc(std::max<size_t>(n * sizeof(float4), 16)));
	ACHK(d_adj_grad.alloc(std::max<size_t>(n * 4 * sizeof(double), 16)));
	ACHK(d_adj_lam.alloc(std::max<size_t>(nn * 35 * sizeof(double), 16)));
	ACHK(d_adj_mu.alloc(std::max<size_t>(nn * 17 * sizeof(double), 16)));
	hipStream_t s = stream;
	ACHK(hipMemcpyAsync(d_adj_w.p, w.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
	if (!adj_ev[0]) {
		ACHK(hipEventCreate(&adj_ev[0]));
		ACHK(hipEventCreate(&adj_ev[1]));
	}
	hipEvent_t e0 = adj_ev[0], e1 = adj_ev[1];
	if (ms) ACHK(hipEventRecord(e0, s));
	hipLaunchKernelGGL(k_adj_zero, dim3(unsigned((nn * 35 + 255) / 256)), dim3(256), 0, s, d_adj_lam.as<double>(), nn * 35);
	hipLaunchKernelGGL(k_adj_zero, dim3(unsigned((nn * 17 + 255) / 256)), dim3(256), 0, s, d_adj_mu.as<double>(), nn * 17);
	ACHK(hipGetLastError());
	AdjArgs a{};
	a.parts = d_parts.as<int4>();
	a.w = d_adj_w.as<float4>();
	a.multis = d_multis.as<MultiSrc>();
	a.nodepos = d_nodepos.as<int4>();
	a.dL = d_dL.as<double>();
	a.leaves = d_leaf_items.as<int4>();
	a.path = d_path.as<int>();
	a.level = d_level.as<int>();
	a.grad = d_adj_grad.as<double>();
	a.Lam = d_adj_lam.as<double>();
	a.mu = d_adj_mu.as<double>();
	a.h = float(soft_len);
	a.wrap = ewald;
	a.maxdepth = tree.max_level + 1;
	const size_t lds = size_t(35 * 64 + 35 + 35 * a.maxdepth) * sizeof(double);
	if (n_leaf_items > 0) {
		hipLaunchKernelGGL(k_adj_l2p, dim3(n_leaf_items), dim3(64), lds, s, a);
		ACHK(hipGetLastError());
	}
	for (size_t l = 0; l + 1 < adj_up_ptr.size(); l++) {
		const int b = adj_up_ptr[l], np = adj_up_ptr[l + 1] - b;
		hipLaunchKernelGGL(k_adj_up, dim3(unsigned((np + 63) / 64)), dim3(64), 0, s, d_adj_up.as<int2>() + b, np, d_nodepos.as<int4>(),
				d_adj_lam.as<double>());
		ACHK(hipGetLastError());
	}
	for (int k : {PP_DIRECT, PP_EWALD, PC_DIRECT, PC_EWALD}) {
		if (work[k].nitems <= 0) continue;
		AdjItems I{work[k].items.as<int4>(), work[k].list.as<int>(), work[k].nitems};
		const dim3 g(work[k].nitems), b(64);
		if (k == PP_DIRECT)
			hipLaunchKernelGGL(k_adj_pp, g, b, 0, s, a, I);
		else if (k == PP_EWALD)
			hipLaunchKernelGGL(k_adj_pp_ewald, g, b, 0, s, a, I);
		else if (k == PC_DIRECT)
			hipLaunchKernelGGL(k_adj_pc_sink<false>, g, b, 0, s, a, I);
		else
			hipLaunchKernelGGL(k_adj_pc_sink<true>, g, b, 0, s, a, I);
		ACHK(hipGetLastError());
	}
	for (int e = 0; e < 2; e++) {
		if (adj_cct_nwaves[e] <= 0) continue;
		AdjCCArgs ca{d_adj_cct_waves[e].as<int2>(), d_adj_cct_src[e].as<int>(), d_adj_cct_snk[e].as<int>(), d_nodepos.as<int4>(),
				d_multis.as<MultiSrc>(), d_adj_lam.as<double>(), d_adj_mu.as<double>(), ewald};
		if (e)
			hipLaunchKernelGGL(k_adj_cc_t<true>, dim3(adj_cct_nwaves[e]), dim3(64), 0, s, ca);
		else
			hipLaunchKernelGGL(k_adj_cc_t<false>, dim3(adj_cct_nwaves[e]), dim3(64), 0, s, ca);
		ACHK(hipGetLastError());
	}
	for (int e = 0; e < 2; e++) {
		if (adj_pct_nitems[e] <= 0) continue;
		AdjPCArgs pa{d_adj_pct_items[e].as<int4>(), d_adj_pct_leaves[e].as<int2>(), d_parts.as<int4>(), d_adj_w.as<float4>(),
				d_multis.as<MultiSrc>(), d_adj_mu.as<double>(), adj_pc_lgL, ewald};
		if (e)
			hipLaunchKernelGGL(k_adj_pc_t<true>, dim3(adj_pct_nitems[e]), dim3(64), 0, s, pa);
		else
			hipLaunchKernelGGL(k_adj_pc_t<false>, dim3(adj_pct_nitems[e]), dim3(64), 0, s, pa);
		ACHK(hipGetLastError());
	}
	for (int e = 0; e < 2; e++) {
		if (adj_cpt_nitems[e] <= 0) continue;
		AdjCPArgs pa{d_adj_cpt_items[e].as<int4>(), d_adj_cpt_sinks[e].as<int>(), d_nodepos.as<int4>(), d_parts.as<int4>(),
				d_adj_lam.as<double>(), d_adj_grad.as<double>(), ewald};
		if (e)
			hipLaunchKernelGGL(k_adj_cp_t<true>, dim3(adj_cpt_nitems[e]), dim3(64), 0, s, pa);
		else
			hipLaunchKernelGGL(k_adj_cp_t<false>, dim3(adj_cpt_nitems[e]), dim3(64), 0, s, pa);
		ACHK(hipGetLastError());
	}
	for (size_t l = adj_up_ptr.size(); l-- > 1;) {
		const int b = adj_up_ptr[l - 1], np = adj_up_ptr[l] - b;
		hipLaunchKernelGGL(k_adj_down, dim3(unsigned((np + 63) / 64)), dim3(64), 0, s, d_adj_up.as<int2>() + b, np, d_nodepos.as<int4>(),
				d_adj_mu.as<double>());
		ACHK(hipGetLastError());
	}
	if (n_leaf_items > 0) {
		hipLaunchKernelGGL(k_adj_moments, dim3(n_leaf_items), dim3(64), 0, s, a);
		ACHK(hipGetLastError());
	}
	if (ms) ACHK(hipEventRecord(e1, s));
	std::vector<double> g(n * 4);
	ACHK(hipMemcpyAsync(g.data(), d_adj_grad.p, n * 4 * sizeof(double), hipMemcpyDeviceToHost, s));
	ACHK(hipStreamSynchronize(s));
	if (ms) ACHK(hipEventElapsedTime(ms, e0, e1));
	for (int64_t i = 0; i < n; i++) {
		const int64_t o = tree.order[i];
		for (int c = 0; c < 4; c++) grad_out[4 * o + c] = g[4 * i + c];
	}
	return TG_OK;
}

}  // namespace tg
