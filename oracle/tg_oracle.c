/*
 * tg_oracle.c -- TEST INFRASTRUCTURE ONLY (see tg_oracle.h).
 *
 * Build with -ffp-contract=off and without -ffast-math: every float operation below is meant
 * to round exactly where the reference's does.  `fmaf` appears only where the reference calls
 * its fused `fmadd` (tigergrav/simd.hpp:694-700).
 */
#define _GNU_SOURCE
#include "tg_oracle.h"
#include "ewald_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define POS_INV (1.0 / 4294967296.0) /* tigergrav/position.hpp:29-30 */
#define NL 16                        /* lanes of the reference's simd_float (AVX2: 2 x 8) */

/* ------------------------------------------------------------------------------------------ */
/* symmetric-tensor slot maps (tigergrav/expansion.hpp:95-137, tigergrav/multipole.hpp:356-380) */
static int S2[3][3], S3[3][3][3], S4[3][3][3][3];
static float EF[35];    /* expansion_factors<float> (tigergrav/expansion.hpp:163-182) */
static double EF64[35]; /* expansion_factors<double> */
static float FOUR_H_OK = 0;

static int rank_sorted(int order, int a, int b, int c, int d) {
	int s[4] = {a, b, c, d};
	for (int i = 0; i < order; i++)
		for (int j = i + 1; j < order; j++)
			if (s[j] < s[i]) {
				int t = s[i];
				s[i] = s[j];
				s[j] = t;
			}
	int t[4] = {0, 0, 0, 0}, rank = 0;
	for (;;) {
		int eq = 1;
		for (int i = 0; i < order; i++) eq &= (t[i] == s[i]);
		if (eq) return rank;
		rank++;
		int p = order - 1;
		while (p >= 0 && t[p] == 2) p--;
		t[p]++;
		for (int q = p + 1; q < order; q++) t[q] = t[p];
	}
}

__attribute__((constructor)) static void tgo_init(void) {
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++) {
			S2[a][b] = rank_sorted(2, a, b, 0, 0);
			for (int c = 0; c < 3; c++) {
				S3[a][b][c] = rank_sorted(3, a, b, c, 0);
				for (int d = 0; d < 3; d++) S4[a][b][c][d] = rank_sorted(4, a, b, c, d);
			}
		}
	/* the reference builds the factors by repeated float/double += over all ordered tuples */
	for (int i = 0; i < 35; i++) EF[i] = 0.0f, EF64[i] = 0.0;
	EF[0] += 1.0f;
	EF64[0] += 1.0;
	for (int a = 0; a < 3; a++) {
		EF[1 + a] += 1.0f;
		EF64[1 + a] += 1.0;
		for (int b = 0; b < 3; b++) {
			EF[4 + S2[a][b]] += 0.5f;
			EF64[4 + S2[a][b]] += 0.5;
			for (int c = 0; c < 3; c++) {
				EF[10 + S3[a][b][c]] += (float) (1.0 / 6.0);
				EF64[10 + S3[a][b][c]] += 1.0 / 6.0;
				for (int d = 0; d < 3; d++) {
					EF[20 + S4[a][b][c][d]] += (float) (1.0 / 24.0);
					EF64[20 + S4[a][b][c][d]] += 1.0 / 24.0;
				}
			}
		}
	}
	FOUR_H_OK = 1;
}

#define E1(a) (1 + (a))
#define E2(a, b) (4 + S2[a][b])
#define E3(a, b, c) (10 + S3[a][b][c])
#define E4(a, b, c, d) (20 + S4[a][b][c][d])
#define M2(a, b) (1 + S2[a][b])
#define M3(a, b, c) (7 + S3[a][b][c])

static const int LANE_SUM_ORDER[NL] = {0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15};

/* simd_double::sum() (tigergrav/simd.hpp:533-542): the 4x4 register layout visits lanes
 * 0,4,8,12,1,5,... */
static double lane_sum(const double* v) {
	double s = 0.0;
	for (int i = 0; i < NL; i++) s += v[LANE_SUM_ORDER[i]];
	return s;
}

/* _mm256_max_ps / _mm256_min_ps semantics: second operand wins ties and NaNs */
static inline float vmax(float a, float b) { return a > b ? a : b; }
static inline float vmin(float a, float b) { return a < b ? a : b; }
static inline float neg(float a) { return 0.0f - a; } /* simd_float unary minus */
static inline float dot3(const float* v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

/* ------------------------------------------------------------------------------------------ */
/* the reference's float math (tigergrav/simd.hpp:556-692) */

/* 2^r: round-to-nearest-even split + degree-8 Taylor polynomial in ln2 + exponent bits
 * (simd.hpp:556-620) */
static float two_pow(float r) {
	const float c1 = (float) log(2.0);
	const float c2 = (float) (0.5 * pow(log(2.0), 2));
	const float c3 = (float) ((1.0 / 6.0) * pow(log(2.0), 3));
	const float c4 = (float) ((1.0 / 24.0) * pow(log(2.0), 4));
	const float c5 = (float) ((1.0 / 120.0) * pow(log(2.0), 5));
	const float c6 = (float) ((1.0 / 720.0) * pow(log(2.0), 6));
	const float c7 = (float) ((1.0 / 5040.0) * pow(log(2.0), 7));
	const float c8 = (float) ((1.0 / 40320.0) * pow(log(2.0), 8));
	const float r0 = rintf(r);
	const int32_t n = (int32_t) r0;
	const float x = r - (float) n;
	float y = c8;
	y = fmaf(y, x, c7);
	y = fmaf(y, x, c6);
	y = fmaf(y, x, c5);
	y = fmaf(y, x, c4);
	y = fmaf(y, x, c3);
	y = fmaf(y, x, c2);
	y = fmaf(y, x, c1);
	y = fmaf(y, x, 1.0f);
	uint32_t bits = ((uint32_t) (n + 0x7f)) << 23;
	float scale;
	memcpy(&scale, &bits, 4);
	return y * scale;
}

float tgo_exp(float a) { /* simd.hpp:667-674 */
	const float c0 = (float) (1.0 / log(2.0));
	a = vmin(a, 88.0f);
	a = vmax(a, -88.0f);
	return two_pow(a * c0);
}

/* Abramowitz & Stegun 7.1.26 erfc with exp(-x^2) returned through e (simd.hpp:676-692) */
float tgo_erfcexp(float x, float* e) {
	const float p = (float) 0.3275911, a1 = (float) 0.254829592, a2 = (float) -0.284496736;
	const float a3 = (float) 1.421413741, a4 = (float) -1.453152027, a5 = (float) 1.061405429;
	const float t1 = 1.0f / (1.0f + p * x);
	const float t2 = t1 * t1;
	const float t3 = t2 * t1;
	const float t4 = t2 * t2;
	const float t5 = t2 * t3;
	*e = tgo_exp(neg(x) * x);
	return (a1 * t1 + a2 * t2 + a3 * t3 + a4 * t4 + a5 * t5) * *e;
}

/* range reduction by 2pi + odd polynomial with split pi (simd.hpp:635-652) */
float tgo_sin(float x0) {
	const float pi_major = (float) 3.1415927, pi_minor = (float) -0.00000008742278;
	float x = x0;
	x = x - rintf(x * (float) (1.0 / (2.0 * M_PI))) * (float) (2.0 * M_PI);
	const float x2 = x * x;
	float p = (float) 0.00000000013291342;
	p = fmaf(p, x2, (float) -0.000000023317787);
	p = fmaf(p, x2, (float) 0.0000025222919);
	p = fmaf(p, x2, (float) -0.00017350505);
	p = fmaf(p, x2, (float) 0.0066208798);
	p = fmaf(p, x2, (float) -0.10132118);
	const float x1 = (x - pi_major) - pi_minor;
	const float x3 = (x + pi_major) + pi_minor;
	return x1 * x3 * p * x;
}

static inline float ref_cos(float x) { return tgo_sin(x + (float) (M_PI / 2.0)); } /* simd.hpp:654-656 */

/* ------------------------------------------------------------------------------------------ */
/* separation of sink X and source Y (src/gravity.cpp:221-230) */
static inline void sep(int ewald, const int32_t* X, const int32_t* Y, float* dx) {
	for (int d = 0; d < 3; d++) {
		if (ewald) {
			const int32_t w = (int32_t) ((uint32_t) X[d] - (uint32_t) Y[d]);
			dx[d] = (float) ((double) w * POS_INV);
		} else {
			dx[d] = (float) ((double) X[d] * POS_INV - (double) Y[d] * POS_INV);
		}
	}
}

/* green_deriv_direct (tigergrav/green.hpp:62-117): order of slot updates kept, since slots
 * alias when indices repeat */
static void green_deriv_direct(float* D, float d0, float d1, float d2, float d3, float d4, const float* dx) {
	D[0] = d0;
	for (int a = 0; a < 3; a++) {
		D[E1(a)] = dx[a] * d1;
		for (int b = 0; b <= a; b++) {
			const float dxadxb = dx[a] * dx[b];
			D[E2(a, b)] = dxadxb * d2;
			for (int c = 0; c <= b; c++) {
				const float dxadxbdxc = dxadxb * dx[c];
				D[E3(a, b, c)] = dxadxbdxc * d3;
				for (int d = 0; d <= c; d++) D[E4(a, b, c, d)] = dxadxbdxc * dx[d] * d4;
			}
		}
	}
	for (int a = 0; a < 3; a++) {
		D[E2(a, a)] += d1;
		D[E3(a, a, a)] = fmaf(dx[a], d2, D[E3(a, a, a)]);
		D[E4(a, a, a, a)] = fmaf(dx[a] * dx[a], d3, D[E4(a, a, a, a)]);
		D[E4(a, a, a, a)] = fmaf(2.0f, d2, D[E4(a, a, a, a)]);
		for (int b = 0; b <= a; b++) {
			const float dxadxb = dx[a] * dx[b];
			D[E3(a, a, b)] = fmaf(dx[b], d2, D[E3(a, a, b)]);
			D[E3(a, b, b)] = fmaf(dx[a], d2, D[E3(a, b, b)]);
			D[E4(a, a, a, b)] = fmaf(dxadxb, d3, D[E4(a, a, a, b)]);
			D[E4(a, b, b, b)] = fmaf(dxadxb, d3, D[E4(a, b, b, b)]);
			D[E4(a, a, b, b)] += d2;
			for (int c = 0; c <= b; c++) {
				D[E4(a, a, b, c)] = fmaf(dx[b], dx[c] * d3, D[E4(a, a, b, c)]);
				D[E4(a, b, c, c)] = fmaf(dxadxb, d3, D[E4(a, b, c, c)]);
				D[E4(a, b, b, c)] = fmaf(dx[a], dx[c] * d3, D[E4(a, b, b, c)]);
			}
		}
	}
}

/* green_direct (tigergrav/green.hpp:167-186): derivatives of -1/r, zero at r <= 1e-9 */
void tgo_green_direct(const float* dX, float* D) {
	const float r0 = (float) 1.0e-9;
	const float r2 = dot3(dX);
	const float r = sqrtf(r2);
	const float rinv = (r > r0 ? 1.0f : 0.0f) / vmax(r, r0);
	const float r2inv = rinv * rinv;
	const float d0 = neg(rinv);
	const float d1 = neg(d0) * r2inv;
	const float d2 = -3.0f * d1 * r2inv;
	const float d3 = -5.0f * d2 * r2inv;
	const float d4 = -7.0f * d3 * r2inv;
	green_deriv_direct(D, d0, d1, d2, d3, d4, dX);
}

/* green_deriv_ewald (tigergrav/green.hpp:119-165): FP64 += of float terms */
static void green_deriv_ewald(double* D, float d0, float d1, float d2, float d3, float d4, const float* dx) {
	D[0] += d0;
	for (int a = 0; a < 3; a++) {
		D[E1(a)] += dx[a] * d1;
		D[E2(a, a)] += d1;
		D[E3(a, a, a)] += dx[a] * d2;
		D[E4(a, a, a, a)] += dx[a] * dx[a] * d3;
		D[E4(a, a, a, a)] += 2.0f * d2;
		for (int b = 0; b <= a; b++) {
			const float dxadxb = dx[a] * dx[b];
			D[E2(a, b)] += dxadxb * d2;
			D[E3(a, a, b)] += dx[b] * d2;
			D[E3(a, b, b)] += dx[a] * d2;
			D[E4(a, a, a, b)] += dxadxb * d3;
			D[E4(a, b, b, b)] += dxadxb * d3;
			D[E4(a, a, b, b)] += d2;
			for (int c = 0; c <= b; c++) {
				const float dxadxbdxc = dxadxb * dx[c];
				D[E3(a, b, c)] += dxadxbdxc * d3;
				D[E4(a, a, b, c)] += dx[b] * dx[c] * d3;
				D[E4(a, b, c, c)] += dxadxb * d3;
				D[E4(a, b, b, c)] += dx[a] * dx[c] * d3;
				for (int d = 0; d <= c; d++) D[E4(a, b, c, d)] += dxadxbdxc * dx[d] * d4;
			}
		}
	}
}

/* green_ewald (tigergrav/green.hpp:189-286): periodic correction derivative tensor */
void tgo_green_ewald(const float* X, float* rcD) {
	const float rcut = (float) 1.0e-6;
	const float fouroversqrtpi = (float) (4.0 / sqrt(M_PI));
	double D[35];
	for (int i = 0; i < 35; i++) D[i] = 0.0;
	const float r = sqrtf(dot3(X));
	const float zmask = r > rcut ? 1.0f : 0.0f;
	for (int i = 0; i < TGO_NREAL; i++) {
		const float* n = tgo_real_idx[i];
		float dx[3] = {X[0] - n[0], X[1] - n[1], X[2] - n[2]};
		const float r2 = dot3(dx);
		const float r4 = r2 * r2;
		const float r = sqrtf(r2);
		const float mask = (r < 3.6f ? 1.0f : 0.0f) * zmask;
		const float rinv = mask / vmax(r, rcut);
		const float r2inv = rinv * rinv;
		const float r3inv = r2inv * rinv;
		const float r5inv = r2inv * r3inv;
		const float r7inv = r2inv * r5inv;
		const float r9inv = r2inv * r7inv;
		float expfac;
		const float erfc = tgo_erfcexp(2.0f * r, &expfac);
		const float expfactor = fouroversqrtpi * r * expfac;
		const float d0 = neg(erfc) * rinv;
		const float d1 = (expfactor + erfc) * r3inv;
		const float d2 = neg(fmaf(expfactor, fmaf(8.0f, r2, 3.0f), 3.0f * erfc)) * r5inv;
		const float d3 = fmaf(expfactor, (15.0f + fmaf(40.0f, r2, 64.0f * r4)), 15.0f * erfc) * r7inv;
		const float d4 = neg(fmaf(expfactor, fmaf(8.0f * r2, (35.0f + fmaf(56.0f, r2, 64.0f * r4)), 105.0f), 105.0f * erfc)) * r9inv;
		green_deriv_ewald(D, d0, d1, d2, d3, d4, dx);
	}
	const float twopi = (float) (2.0 * M_PI);
	for (int i = 0; i < TGO_NFOUR; i++) {
		const float* h = tgo_four_idx[i];
		const float* H = tgo_four_H[i];
		float hdotdx = h[0] * X[0];
		hdotdx += X[1] * h[1];
		hdotdx += X[2] * h[2];
		const float omega = twopi * hdotdx;
		float si = tgo_sin(omega), co = ref_cos(omega);
		si *= zmask;
		co *= zmask;
		D[0] += H[0] * co;
		for (int a = 0; a < 3; a++) {
			D[E1(a)] += H[E1(a)] * si;
			for (int b = 0; b <= a; b++) {
				D[E2(a, b)] += H[E2(a, b)] * co;
				for (int c = 0; c <= b; c++) {
					D[E3(a, b, c)] += H[E3(a, b, c)] * si;
					for (int d = 0; d <= c; d++) D[E4(a, b, c, d)] += H[E4(a, b, c, d)] * co;
				}
			}
		}
	}
	for (int i = 0; i < 35; i++) rcD[i] = (float) D[i];
	float D1[35];
	tgo_green_direct(X, D1);
	const float rinv = neg(zmask) * D1[0];
	rcD[0] = (float) (M_PI / 4.0) + rcD[0] + rinv;
	rcD[0] = fmaf((float) 2.8372975, (1.0f - zmask), rcD[0] * zmask);
	for (int a = 0; a < 3; a++) {
		rcD[E1(a)] = zmask * (rcD[E1(a)] - D1[E1(a)]);
		for (int b = 0; b <= a; b++) {
			rcD[E2(a, b)] = zmask * (rcD[E2(a, b)] - D1[E2(a, b)]);
			for (int c = 0; c <= b; c++) {
				rcD[E3(a, b, c)] = zmask * (rcD[E3(a, b, c)] - D1[E3(a, b, c)]);
				for (int d = 0; d <= c; d++) rcD[E4(a, b, c, d)] = zmask * (rcD[E4(a, b, c, d)] - D1[E4(a, b, c, d)]);
			}
		}
	}
}

/* ------------------------------------------------------------------------------------------ */
/* contractions (src/gravity.cpp:26-169), float products accumulated into FP64 */

/* M2L: multipole -> local expansion (gravity.cpp:26-90) */
static void m2l_acc(double* L, const float* M, const float* D) {
	L[0] += M[0] * D[0];
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++) {
			L[0] += M[M2(a, b)] * D[E2(a, b)] * EF[E2(a, b)];
			for (int c = b; c < 3; c++) L[0] -= M[M3(a, b, c)] * D[E3(a, b, c)] * EF[E3(a, b, c)];
		}
	for (int a = 0; a < 3; a++) {
		L[E1(a)] += M[0] * D[E1(a)];
		for (int b = 0; b < 3; b++)
			for (int c = b; c < 3; c++) {
				L[E1(a)] += M[M2(c, b)] * D[E3(a, b, c)] * EF[E2(c, b)];
				for (int d = c; d < 3; d++) L[E1(a)] -= M[M3(b, c, d)] * D[E4(a, b, c, d)] * EF[E3(b, c, d)];
			}
	}
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++) {
			L[E2(a, b)] += M[0] * D[E2(a, b)];
			for (int c = 0; c < 3; c++)
				for (int d = c; d < 3; d++) L[E2(a, b)] += M[M2(c, d)] * D[E4(a, b, c, d)] * EF[E2(c, d)];
		}
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++)
			for (int c = b; c < 3; c++) L[E3(a, b, c)] += M[0] * D[E3(a, b, c)];
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++)
			for (int c = b; c < 3; c++)
				for (int d = c; d < 3; d++) L[E4(a, b, c, d)] += M[0] * D[E4(a, b, c, d)];
}

/* P2L: point mass -> local expansion (gravity.cpp:92-121) */
static void p2l_acc(double* L, float M, const float* D) {
	L[0] += M * D[0];
	for (int a = 0; a < 3; a++) {
		L[E1(a)] += M * D[E1(a)];
		for (int b = a; b < 3; b++) {
			L[E2(a, b)] += M * D[E2(a, b)];
			for (int c = b; c < 3; c++) {
				L[E3(a, b, c)] += M * D[E3(a, b, c)];
				for (int d = c; d < 3; d++) L[E4(a, b, c, d)] += M * D[E4(a, b, c, d)];
			}
		}
	}
}

/* M2P: multipole -> potential and acceleration (gravity.cpp:123-169).  The octupole term of
 * the potential loops b over all three values (not b >= a), so sorted triples are weighted by
 * their number of distinct indices -- kept as the reference computes it. */
static void m2p(double* phi, double* g, const float* M, const float* D) {
	double f0 = M[0] * D[0];
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++) f0 += M[M2(a, b)] * D[E2(a, b)] * EF[E2(a, b)];
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++)
			for (int c = b; c < 3; c++) f0 -= M[M3(a, b, c)] * D[E3(a, b, c)] * EF[E3(a, b, c)];
	double f1[3] = {0.0, 0.0, 0.0};
	for (int a = 0; a < 3; a++) f1[a] -= M[0] * D[E1(a)];
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++)
			for (int c = b; c < 3; c++) f1[a] -= M[M2(c, b)] * D[E3(a, b, c)] * EF[E2(c, b)];
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++)
			for (int c = b; c < 3; c++)
				for (int d = c; d < 3; d++) f1[a] += M[M3(c, b, d)] * D[E4(a, b, c, d)] * EF[E3(b, c, d)];
	*phi = f0;
	g[0] = f1[0];
	g[1] = f1[1];
	g[2] = f1[2];
}

/* ------------------------------------------------------------------------------------------ */
/* kernels */

static int64_t pad16(int64_t n) { return ((n + NL - 1) / NL) * NL; }

/* gravity_PP_direct (src/gravity.cpp:171-300) */
uint64_t tgo_PP_direct(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj, int do_phi) {
	if (ns == 0) return 0;
	const double h = P->soft_len, m = P->m;
	const float H = (float) h, halfH = (float) (0.5 * h), Hinv = (float) (1.0 / h), H3inv = (float) (1.0 / h / h / h);
	const double phi_self = (-315.0 / 128.0) * m / h;
	double* G = calloc((size_t) ns * 3 * NL, sizeof(double));
	double* Phi = calloc((size_t) ns * NL, sizeof(double));
	const int64_t cnt2 = pad16(nj);
	for (int64_t J = 0; J < cnt2; J += NL) {
		for (int64_t i = 0; i < ns; i++) {
			for (int k = 0; k < NL; k++) {
				const int64_t j = J + k;
				const int32_t zero[3] = {0, 0, 0};
				const int32_t* Y = j < nj ? y + 3 * j : zero;
				const float Mk = j < nj ? (float) m : 0.0f;
				float dX[3];
				sep(P->ewald, x + 3 * i, Y, dX);
				const float r2 = dot3(dX);
				const float r = sqrtf(r2);
				const float rinv = 1.0f / vmax(r, halfH);
				const float rinv3 = rinv * rinv * rinv;
				const float sw1 = r > H ? 1.0f : 0.0f;
				const float sw3 = r < halfH ? 1.0f : 0.0f;
				const float sw2 = 1.0f - sw1 - sw3;
				const float roh = vmin(r * Hinv, 1.0f);
				const float roh2 = roh * roh;
				const float roh3 = roh2 * roh;
				const float f1 = rinv3;
				float f2 = (float) (-32.0 / 3.0);
				f2 = fmaf(f2, roh, (float) (+192.0 / 5.0));
				f2 = fmaf(f2, roh, -48.0f);
				f2 = fmaf(f2, roh, (float) (+64.0 / 3.0));
				f2 = fmaf(f2, roh3, (float) (-1.0 / 15.0));
				f2 *= rinv3;
				float f3 = 32.0f;
				f3 = fmaf(f3, roh, (float) (-192.0 / 5.0));
				f3 = fmaf(f3, roh2, (float) (+32.0 / 3.0));
				f3 *= H3inv;
				const float fr = sw1 * f1 + sw2 * f2 + sw3 * f3;
				for (int d = 0; d < 3; d++) G[(i * 3 + d) * NL + k] -= (double) (dX[d] * Mk * fr);
				if (do_phi) {
					const float p1 = rinv;
					float p2 = (float) (+32.0 / 15.0);
					p2 = fmaf(p2, roh, (float) (-48.0 / 5.0));
					p2 = fmaf(p2, roh, 16.0f);
					p2 = fmaf(p2, roh, (float) (-32.0 / 3.0));
					p2 = fmaf(p2, roh2, (float) (+16.0 / 5.0));
					p2 = fmaf(p2, roh, (float) (-1.0 / 15.0));
					p2 *= rinv;
					float p3 = (float) (-32.0 / 5.0);
					p3 = fmaf(p3, roh, (float) (+48.0 / 5.0));
					p3 = fmaf(p3, roh2, (float) (-16.0 / 3.0));
					p3 = fmaf(p3, roh2, (float) (+14.0 / 5.0));
					p3 *= Hinv;
					Phi[i * NL + k] -= (double) ((sw1 * p1 + sw2 * p2 + sw3 * p3) * Mk);
				}
			}
		}
	}
	for (int64_t i = 0; i < ns; i++) {
		for (int d = 0; d < 3; d++) f[i].g[d] += lane_sum(G + (i * 3 + d) * NL);
		f[i].phi += lane_sum(Phi + i * NL);
		f[i].phi -= phi_self;
	}
	free(G);
	free(Phi);
	return ((74 + (uint64_t) do_phi * 15) * (uint64_t) nj + NL * 4 + 5) * (uint64_t) ns;
}

/* shared driver for the multipole->particle kernels (PC direct/ewald, gravity.cpp:302-366,616-678) */
static void pc_common(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc, int ewald_kernel) {
	double* G = calloc((size_t) ns * 3 * NL, sizeof(double));
	double* Phi = calloc((size_t) ns * NL, sizeof(double));
	const int64_t cnt2 = pad16(nc);
	for (int64_t J = 0; J < cnt2; J += NL) {
		for (int64_t i = 0; i < ns; i++) {
			for (int k = 0; k < NL; k++) {
				const int64_t j = J + k;
				float M[17];
				const int32_t* Y;
				if (j < nc) {
					memcpy(M, y[j].m, sizeof(M));
					Y = y[j].x;
				} else {
					for (int q = 0; q < 17; q++) M[q] = 0.0f;
					Y = y[nc - 1].x;
				}
				float dX[3], D[35];
				sep(ewald_kernel ? 1 : P->ewald, x + 3 * i, Y, dX);
				if (ewald_kernel)
					tgo_green_ewald(dX, D);
				else
					tgo_green_direct(dX, D);
				double phi, g[3];
				m2p(&phi, g, M, D);
				for (int d = 0; d < 3; d++) G[(i * 3 + d) * NL + k] += g[d];
				Phi[i * NL + k] += phi;
			}
		}
	}
	for (int64_t i = 0; i < ns; i++) {
		for (int d = 0; d < 3; d++) f[i].g[d] += lane_sum(G + (i * 3 + d) * NL);
		f[i].phi += lane_sum(Phi + i * NL);
	}
	free(G);
	free(Phi);
}

uint64_t tgo_PC_direct(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc) {
	if (ns == 0) return 0;
	if (nc > 0) pc_common(P, f, x, ns, y, nc, 0);
	return 581 * (uint64_t) nc * (uint64_t) ns;
}

uint64_t tgo_PC_ewald(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc) {
	if (ns == 0) return 0;
	if (nc > 0) pc_common(P, f, x, ns, y, nc, 1);
	return 251531 * (uint64_t) nc * (uint64_t) ns;
}

/* multipole->local (CC direct/ewald, gravity.cpp:368-426,680-731) */
static void cc_common(const tgo_params* P, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc, int ewald_kernel) {
	double* Lacc = calloc(35 * NL, sizeof(double));
	double Ll[35];
	const int64_t cnt2 = pad16(nc);
	for (int64_t J = 0; J < cnt2; J += NL) {
		for (int k = 0; k < NL; k++) {
			const int64_t j = J + k;
			float M[17];
			const int32_t* Y;
			if (j < nc) {
				memcpy(M, y[j].m, sizeof(M));
				Y = y[j].x;
			} else {
				for (int q = 0; q < 17; q++) M[q] = 0.0f;
				Y = y[nc - 1].x;
			}
			float dX[3], D[35];
			sep(ewald_kernel ? 1 : P->ewald, x, Y, dX);
			if (ewald_kernel)
				tgo_green_ewald(dX, D);
			else
				tgo_green_direct(dX, D);
			for (int q = 0; q < 35; q++) Ll[q] = Lacc[q * NL + k];
			m2l_acc(Ll, M, D);
			for (int q = 0; q < 35; q++) Lacc[q * NL + k] = Ll[q];
		}
	}
	for (int q = 0; q < 35; q++) L[q] += lane_sum(Lacc + q * NL);
	free(Lacc);
}

uint64_t tgo_CC_direct(const tgo_params* P, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc) {
	if (nc == 0) return 0;
	cc_common(P, L, x, y, nc, 0);
	return 1025 * (uint64_t) nc;
}

uint64_t tgo_CC_ewald(const tgo_params* P, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc) {
	if (nc == 0) return 0;
	cc_common(P, L, x, y, nc, 1);
	return 251975 * (uint64_t) nc;
}

/* particle->local (CP direct/ewald, gravity.cpp:428-486,733-785) */
static void cp_common(const tgo_params* P, double* L, const int32_t* x, const int32_t* y, int64_t nj, int ewald_kernel) {
	double* Lacc = calloc(35 * NL, sizeof(double));
	double Ll[35];
	const int64_t cnt2 = pad16(nj);
	for (int64_t J = 0; J < cnt2; J += NL) {
		for (int k = 0; k < NL; k++) {
			const int64_t j = J + k;
			const int32_t* Y = j < nj ? y + 3 * j : y + 3 * (nj - 1);
			const float M = j < nj ? (float) P->m : 0.0f;
			float dX[3], D[35];
			sep(ewald_kernel ? 1 : P->ewald, x, Y, dX);
			if (ewald_kernel)
				tgo_green_ewald(dX, D);
			else
				tgo_green_direct(dX, D);
			for (int q = 0; q < 35; q++) Ll[q] = Lacc[q * NL + k];
			p2l_acc(Ll, M, D);
			for (int q = 0; q < 35; q++) Lacc[q * NL + k] = Ll[q];
		}
	}
	for (int q = 0; q < 35; q++) L[q] += lane_sum(Lacc + q * NL);
	free(Lacc);
}

uint64_t tgo_CP_direct(const tgo_params* P, double* L, const int32_t* x, const int32_t* y, int64_t nj) {
	if (nj == 0) return 0;
	cp_common(P, L, x, y, nj, 0);
	return 455 * (uint64_t) nj;
}

uint64_t tgo_CP_ewald(const tgo_params* P, double* L, const int32_t* x, const int32_t* y, int64_t nj) {
	if (nj == 0) return 0;
	cp_common(P, L, x, y, nj, 1);
	return 251233 * (uint64_t) nj;
}

/* gravity_PP_ewald (src/gravity.cpp:491-614) */
uint64_t tgo_PP_ewald(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj) {
	if (ns == 0) return 0;
	const double M = P->m;
	const float rcut = (float) 1.0e-6, twopi = (float) (2 * M_PI), pioverfour = (float) (M_PI / 4.0);
	const float fouroversqrtpi = (float) (4.0 / sqrt(M_PI));
	const double phi0 = 2.8372975;
	double* G = calloc((size_t) ns * 3 * NL, sizeof(double));
	double* Phi = calloc((size_t) ns * NL, sizeof(double));
	const int64_t cnt2 = pad16(nj);
	for (int64_t J = 0; J < cnt2; J += NL) {
		for (int64_t I = 0; I < ns; I++) {
			for (int k = 0; k < NL; k++) {
				const int64_t j = J + k;
				const int32_t zero[3] = {0, 0, 0};
				const int32_t* Y = j < nj ? y + 3 * j : zero;
				const double mask = j < nj ? 1.0 : 0.0;
				float dX0[3];
				sep(1, x + 3 * I, Y, dX0);
				double phi = 0.0, g[3] = {0.0, 0.0, 0.0};
				const float r = sqrtf(dot3(dX0));
				const float cut_mask = r > rcut ? 1.0f : 0.0f;
				for (int i = 0; i < TGO_NREAL; i++) {
					const float* n = tgo_real_idx[i];
					const float dx[3] = {dX0[0] - n[0], dX0[1] - n[1], dX0[2] - n[2]};
					const float r2 = dot3(dx);
					const float r = sqrtf(r2);
					const float mask = cut_mask * (r < 3.6f ? 1.0f : 0.0f);
					const float rinv = mask * 1.0f / vmax(r, rcut);
					const float r2inv = rinv * rinv;
					const float r3inv = r2inv * rinv;
					float expfac;
					const float erfc = tgo_erfcexp(2.0f * r, &expfac);
					const float d0 = neg(erfc) * rinv;
					const float expfactor = fouroversqrtpi * r * expfac;
					const float d1 = (expfactor + erfc) * r3inv;
					phi += d0;
					for (int a = 0; a < 3; a++) g[a] -= (dX0[a] - n[a]) * d1;
				}
				for (int i = 0; i < TGO_NFOUR; i++) {
					const float* H = tgo_four_H[i];
					const float* h = tgo_four_idx[i];
					float hdotdx = h[0] * dX0[0];
					for (int d = 1; d < 3; d++) hdotdx += dX0[d] * h[d];
					const float omega = twopi * hdotdx;
					const float s = tgo_sin(omega), c = ref_cos(omega);
					phi += H[0] * c * cut_mask;
					for (int d = 0; d < 3; d++) g[d] -= H[E1(d)] * s * cut_mask;
				}
				const float rinv = 1.0f / vmax(r, rcut);
				phi += pioverfour + rinv;
				const double sw = cut_mask;
				phi = phi0 * (double) (1.0f - (float) sw) + phi * sw;
				const float rinv3 = cut_mask * rinv * rinv * rinv;
				for (int d = 0; d < 3; d++) g[d] += dX0[d] * rinv3;
				Phi[I * NL + k] += M * phi * mask;
				for (int d = 0; d < 3; d++) G[(I * 3 + d) * NL + k] += g[d] * M * mask;
			}
		}
	}
	for (int64_t i = 0; i < ns; i++) {
		for (int d = 0; d < 3; d++) f[i].g[d] += lane_sum(G + (i * 3 + d) * NL);
		f[i].phi += lane_sum(Phi + i * NL);
	}
	free(G);
	free(Phi);
	return 56245 * (uint64_t) nj * (uint64_t) ns;
}

/* ========================================================================================== */
/* FP64 restatement of the same formulas (libm math, same lattice truncation, same quirks)   */

static void green_deriv64(double* D, double d0, double d1, double d2, double d3, double d4, const double* dx, int accumulate) {
	/* closed form of the derivative tensor of f(r): D_{i..} built from radial factors d_k, the
	 * same algebra as green_deriv_direct/green_deriv_ewald */
	double T[35];
	T[0] = d0;
	for (int a = 0; a < 3; a++) {
		T[E1(a)] = dx[a] * d1;
		for (int b = 0; b <= a; b++) {
			T[E2(a, b)] = dx[a] * dx[b] * d2 + (a == b ? d1 : 0.0);
			for (int c = 0; c <= b; c++) {
				double t = dx[a] * dx[b] * dx[c] * d3;
				t += ((a == b) * dx[c] + (a == c) * dx[b] + (b == c) * dx[a]) * d2;
				T[E3(a, b, c)] = t;
				for (int d = 0; d <= c; d++) {
					double u = dx[a] * dx[b] * dx[c] * dx[d] * d4;
					u += ((a == b) * dx[c] * dx[d] + (a == c) * dx[b] * dx[d] + (a == d) * dx[b] * dx[c] + (b == c) * dx[a] * dx[d] +
							(b == d) * dx[a] * dx[c] + (c == d) * dx[a] * dx[b]) *
						d3;
					u += ((a == b) * (c == d) + (a == c) * (b == d) + (a == d) * (b == c)) * d2;
					T[E4(a, b, c, d)] = u;
				}
			}
		}
	}
	for (int i = 0; i < 35; i++) D[i] = accumulate ? D[i] + T[i] : T[i];
}

void tgo64_green_direct(const double* dX, double* D) {
	const double r = sqrt(dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2]);
	const double rinv = r > 1e-9 ? 1.0 / r : 0.0;
	const double r2inv = rinv * rinv;
	const double d0 = -rinv, d1 = rinv * r2inv, d2 = -3.0 * d1 * r2inv, d3 = -5.0 * d2 * r2inv, d4 = -7.0 * d3 * r2inv;
	green_deriv64(D, d0, d1, d2, d3, d4, dX, 0);
}

/* radial factors of -erfc(2r)/r and its derivatives (green.hpp:228-234 in FP64) */
static void ewald_real_factors(double r, double* d) {
	const double r2 = r * r, r4 = r2 * r2;
	const double rinv = 1.0 / r, r2inv = rinv * rinv, r3inv = r2inv * rinv, r5inv = r2inv * r3inv, r7inv = r2inv * r5inv,
		     r9inv = r2inv * r7inv;
	const double erfc_ = erfc(2.0 * r);
	const double expfactor = 4.0 / sqrt(M_PI) * r * exp(-4.0 * r2);
	d[0] = -erfc_ * rinv;
	d[1] = (expfactor + erfc_) * r3inv;
	d[2] = -(expfactor * (8.0 * r2 + 3.0) + 3.0 * erfc_) * r5inv;
	d[3] = (expfactor * (15.0 + 40.0 * r2 + 64.0 * r4) + 15.0 * erfc_) * r7inv;
	d[4] = -(expfactor * (8.0 * r2 * (35.0 + 56.0 * r2 + 64.0 * r4) + 105.0) + 105.0 * erfc_) * r9inv;
}

/* the reference's truncated periodic correction, evaluated in FP64.  Near r = 0 the n = 0
 * real-space term and the subtracted Newtonian term are combined analytically as
 * d^k[erf(2r)/r] so the result stays accurate where the float formula cancels. */
static void green_ewald64_trunc(const double* X, double* D, int nreal2, int nfour2) {
	for (int i = 0; i < 35; i++) D[i] = 0.0;
	const double r0 = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
	const int zmask = r0 > 1.0e-6;
	if (!zmask) {
		D[0] = 2.8372975; /* the reference's r -> 0 value (green.hpp:272) */
		return;
	}
	const int nm = (int) sqrt((double) nreal2) + 1;
	for (int i = -nm; i <= nm; i++)
		for (int j = -nm; j <= nm; j++)
			for (int k = -nm; k <= nm; k++) {
				if (i * i + j * j + k * k > nreal2) continue;
				const double dx[3] = {X[0] - i, X[1] - j, X[2] - k};
				const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
				if (!(r < 3.6)) continue;
				double d[5];
				if (i == 0 && j == 0 && k == 0) {
					/* -erfc(2r)/r + 1/r = erf(2r)/r; radial factors d_k = (1/r d/dr)^k of it */
					double e[5], n[5];
					ewald_real_factors(r, e);
					const double rinv = 1.0 / r, r2inv = rinv * rinv;
					n[0] = -rinv;
					n[1] = rinv * r2inv;
					n[2] = -3.0 * n[1] * r2inv;
					n[3] = -5.0 * n[2] * r2inv;
					n[4] = -7.0 * n[3] * r2inv;
					if (r < 0.05) {
						/* series of erf(2r)/r = (2/sqrt(pi)) sum_k (-1)^k (2r)^{2k+1}/(k!(2k+1)) / r */
						double c[12];
						const double s = 2.0 / sqrt(M_PI);
						double fact = 1.0;
						for (int q = 0; q < 12; q++) {
							if (q > 0) fact *= q;
							c[q] = s * ((q & 1) ? -1.0 : 1.0) * pow(2.0, 2 * q + 1) / (fact * (2 * q + 1));
						}
						/* f(r) = sum_q c_q r^{2q};  (1/r d/dr)^k f = sum_q c_q (2q)(2q-2)...(2q-2k+2) r^{2q-2k} */
						for (int kk = 0; kk < 5; kk++) {
							double acc = 0.0;
							for (int q = kk; q < 12; q++) {
								double coef = c[q];
								for (int t = 0; t < kk; t++) coef *= (2 * q - 2 * t);
								acc += coef * pow(r, 2 * (q - kk));
							}
							d[kk] = acc;
						}
					} else {
						for (int kk = 0; kk < 5; kk++) d[kk] = e[kk] - n[kk];
					}
				} else {
					ewald_real_factors(r, d);
				}
				green_deriv64(D, d[0], d[1], d[2], d[3], d[4], dx, 1);
			}
	const int hm = (int) sqrt((double) nfour2) + 1;
	for (int i = -hm; i <= hm; i++)
		for (int j = -hm; j <= hm; j++)
			for (int k = -hm; k <= hm; k++) {
				const int h2 = i * i + j * j + k * k;
				if (h2 > nfour2 || h2 == 0) continue;
				const double h[3] = {i, j, k};
				const double c0 = exp(-M_PI * M_PI * h2 / 4.0) / h2;
				const double om = 2.0 * M_PI * (h[0] * X[0] + h[1] * X[1] + h[2] * X[2]);
				const double co = cos(om), si = sin(om);
				D[0] += -(1.0 / M_PI) * c0 * co;
				for (int a = 0; a < 3; a++) {
					D[E1(a)] += 2.0 * h[a] * c0 * si;
					for (int b = 0; b <= a; b++) {
						D[E2(a, b)] += 4.0 * M_PI * h[a] * h[b] * c0 * co;
						for (int c = 0; c <= b; c++) {
							D[E3(a, b, c)] += -8.0 * M_PI * M_PI * h[a] * h[b] * h[c] * c0 * si;
							for (int d = 0; d <= c; d++) D[E4(a, b, c, d)] += -16.0 * M_PI * M_PI * M_PI * h[a] * h[b] * h[c] * h[d] * c0 * co;
						}
					}
				}
			}
	D[0] += M_PI / 4.0;
}

void tgo64_green_ewald(const double* X, double* D) { green_ewald64_trunc(X, D, 17, 9); }

static inline void sep64(int ewald, const int32_t* X, const int32_t* Y, double* dx) {
	for (int d = 0; d < 3; d++) {
		if (ewald)
			dx[d] = (double) (int32_t) ((uint32_t) X[d] - (uint32_t) Y[d]) * POS_INV;
		else
			dx[d] = ((double) X[d] - (double) Y[d]) * POS_INV;
	}
}

static void m2l64d(double* L, const double* M, const double* D);
static void m2l64(double* L, const float* Mf, const double* D) {
	double M[17];
	for (int i = 0; i < 17; i++) M[i] = Mf[i];
	m2l64d(L, M, D);
}

static void m2l64d(double* L, const double* M, const double* D) {
	L[0] += M[0] * D[0];
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++) {
			L[0] += M[M2(a, b)] * D[E2(a, b)] * EF64[E2(a, b)];
			for (int c = b; c < 3; c++) L[0] -= M[M3(a, b, c)] * D[E3(a, b, c)] * EF64[E3(a, b, c)];
		}
	for (int a = 0; a < 3; a++) {
		L[E1(a)] += M[0] * D[E1(a)];
		for (int b = 0; b < 3; b++)
			for (int c = b; c < 3; c++) {
				L[E1(a)] += M[M2(c, b)] * D[E3(a, b, c)] * EF64[E2(c, b)];
				for (int d = c; d < 3; d++) L[E1(a)] -= M[M3(b, c, d)] * D[E4(a, b, c, d)] * EF64[E3(b, c, d)];
			}
	}
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++) {
			L[E2(a, b)] += M[0] * D[E2(a, b)];
			for (int c = 0; c < 3; c++)
				for (int d = c; d < 3; d++) L[E2(a, b)] += M[M2(c, d)] * D[E4(a, b, c, d)] * EF64[E2(c, d)];
		}
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++)
			for (int c = b; c < 3; c++) L[E3(a, b, c)] += M[0] * D[E3(a, b, c)];
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++)
			for (int c = b; c < 3; c++)
				for (int d = c; d < 3; d++) L[E4(a, b, c, d)] += M[0] * D[E4(a, b, c, d)];
}

static void m2p64d(double* phi, double* g, const double* M, const double* D);
static void m2p64(double* phi, double* g, const float* Mf, const double* D) {
	double M[17];
	for (int i = 0; i < 17; i++) M[i] = Mf[i];
	m2p64d(phi, g, M, D);
}

static void m2p64d(double* phi, double* g, const double* M, const double* D) {
	double f0 = M[0] * D[0];
	for (int a = 0; a < 3; a++)
		for (int b = a; b < 3; b++) f0 += M[M2(a, b)] * D[E2(a, b)] * EF64[E2(a, b)];
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++)
			for (int c = b; c < 3; c++) f0 -= M[M3(a, b, c)] * D[E3(a, b, c)] * EF64[E3(a, b, c)];
	double f1[3] = {0, 0, 0};
	for (int a = 0; a < 3; a++) {
		f1[a] -= M[0] * D[E1(a)];
		for (int b = 0; b < 3; b++)
			for (int c = b; c < 3; c++) {
				f1[a] -= M[M2(c, b)] * D[E3(a, b, c)] * EF64[E2(c, b)];
				for (int d = c; d < 3; d++) f1[a] += M[M3(c, b, d)] * D[E4(a, b, c, d)] * EF64[E3(b, c, d)];
			}
	}
	*phi += f0;
	for (int a = 0; a < 3; a++) g[a] += f1[a];
}

void tgo64_PP_direct(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj, int do_phi) {
	if (ns == 0) return;
	const double h = P->soft_len, m = P->m;
	for (int64_t i = 0; i < ns; i++) {
		double g[3] = {0, 0, 0}, phi = 0.0;
		for (int64_t j = 0; j < nj; j++) {
			double dx[3];
			sep64(P->ewald, x + 3 * i, y + 3 * j, dx);
			const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
			double fr, pr;
			if (r > h) {
				fr = 1.0 / (r * r * r);
				pr = 1.0 / r;
			} else {
				const double q = r / h;
				if (r < 0.5 * h) {
					fr = (32.0 / 3.0 - 192.0 / 5.0 * q * q + 32.0 * q * q * q) / (h * h * h);
					pr = (14.0 / 5.0 - 16.0 / 3.0 * q * q + 48.0 / 5.0 * q * q * q * q - 32.0 / 5.0 * q * q * q * q * q) / h;
				} else {
					const double q3 = q * q * q;
					fr = (64.0 / 3.0 * q3 - 48.0 * q3 * q + 192.0 / 5.0 * q3 * q * q - 32.0 / 3.0 * q3 * q3 - 1.0 / 15.0) / (r * r * r);
					pr = (-1.0 / 15.0 + 16.0 / 5.0 * q - 32.0 / 3.0 * q3 + 16.0 * q3 * q - 48.0 / 5.0 * q3 * q * q + 32.0 / 15.0 * q3 * q3) / r;
				}
			}
			for (int d = 0; d < 3; d++) g[d] -= dx[d] * m * fr;
			phi -= m * pr;
		}
		for (int d = 0; d < 3; d++) f[i].g[d] += g[d];
		if (do_phi) f[i].phi += phi;
		f[i].phi -= (-315.0 / 128.0) * m / h;
	}
}

static void pc64(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc, int ewald_kernel) {
	for (int64_t i = 0; i < ns; i++)
		for (int64_t j = 0; j < nc; j++) {
			double dx[3], D[35];
			sep64(ewald_kernel ? 1 : P->ewald, x + 3 * i, y[j].x, dx);
			if (ewald_kernel)
				tgo64_green_ewald(dx, D);
			else
				tgo64_green_direct(dx, D);
			m2p64(&f[i].phi, f[i].g, y[j].m, D);
		}
}

void tgo64_PC_direct(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc) { pc64(P, f, x, ns, y, nc, 0); }
void tgo64_PC_ewald(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc) { pc64(P, f, x, ns, y, nc, 1); }

static void cc64(const tgo_params* P, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc, int ewald_kernel) {
	for (int64_t j = 0; j < nc; j++) {
		double dx[3], D[35];
		sep64(ewald_kernel ? 1 : P->ewald, x, y[j].x, dx);
		if (ewald_kernel)
			tgo64_green_ewald(dx, D);
		else
			tgo64_green_direct(dx, D);
		m2l64(L, y[j].m, D);
	}
}

void tgo64_CC_direct(const tgo_params* P, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc) { cc64(P, L, x, y, nc, 0); }
void tgo64_CC_ewald(const tgo_params* P, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc) { cc64(P, L, x, y, nc, 1); }

static void cp64(const tgo_params* P, double* L, const int32_t* x, const int32_t* y, int64_t nj, int ewald_kernel) {
	for (int64_t j = 0; j < nj; j++) {
		double dx[3], D[35];
		sep64(ewald_kernel ? 1 : P->ewald, x, y + 3 * j, dx);
		if (ewald_kernel)
			tgo64_green_ewald(dx, D);
		else
			tgo64_green_direct(dx, D);
		for (int q = 0; q < 35; q++) L[q] += P->m * D[q];
	}
}

void tgo64_CP_direct(const tgo_params* P, double* L, const int32_t* x, const int32_t* y, int64_t nj) { cp64(P, L, x, y, nj, 0); }
void tgo64_CP_ewald(const tgo_params* P, double* L, const int32_t* x, const int32_t* y, int64_t nj) { cp64(P, L, x, y, nj, 1); }

void tgo64_PP_ewald(const tgo_params* P, tgo_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj) {
	for (int64_t i = 0; i < ns; i++)
		for (int64_t j = 0; j < nj; j++) {
			double dx[3], D[35];
			sep64(1, x + 3 * i, y + 3 * j, dx);
			tgo64_green_ewald(dx, D);
			f[i].phi += P->m * D[0];
			for (int d = 0; d < 3; d++) f[i].g[d] -= P->m * D[E1(d)];
		}
}

/* ========================================================================================== */
/* FP64 physics oracle                                                                        */

void tgo_ewald_corr64(const double* dx, double* phi, double* g) {
	double D[35];
	green_ewald64_trunc(dx, D, 25, 16);
	/* D holds derivatives of the correction to -1/r; the potential of a unit mass is D[0] and the
	 * acceleration is -grad */
	*phi = D[0];
	for (int d = 0; d < 3; d++) g[d] = -D[E1(d)];
}

void tgo_direct64(int periodic, double h, const int32_t* xs, int64_t ns, const int32_t* ys, const double* m, int64_t nj, double* phi,
		double* g) {
#pragma omp parallel for schedule(dynamic, 4)
	for (int64_t i = 0; i < ns; i++) {
		double gi[3] = {0, 0, 0}, pi = 0.0;
		for (int64_t j = 0; j < nj; j++) {
			double dx[3];
			sep64(periodic, xs + 3 * i, ys + 3 * j, dx);
			const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
			double fr, pr;
			if (r >= h) {
				fr = 1.0 / (r * r * r);
				pr = 1.0 / r;
			} else {
				const double q = r / h, q2 = q * q, q3 = q2 * q;
				if (r < 0.5 * h) {
					fr = (32.0 / 3.0 - 192.0 / 5.0 * q2 + 32.0 * q3) / (h * h * h);
					pr = (14.0 / 5.0 - 16.0 / 3.0 * q2 + 48.0 / 5.0 * q2 * q2 - 32.0 / 5.0 * q3 * q2) / h;
				} else {
					fr = (64.0 / 3.0 * q3 - 48.0 * q3 * q + 192.0 / 5.0 * q3 * q2 - 32.0 / 3.0 * q3 * q3 - 1.0 / 15.0) / (r * r * r);
					pr = (-1.0 / 15.0 + 16.0 / 5.0 * q - 32.0 / 3.0 * q3 + 16.0 * q3 * q - 48.0 / 5.0 * q3 * q2 + 32.0 / 15.0 * q3 * q3) / r;
				}
			}
			for (int d = 0; d < 3; d++) gi[d] -= m[j] * dx[d] * fr;
			pi -= m[j] * pr;
			if (periodic) {
				double cp, cg[3];
				tgo_ewald_corr64(dx, &cp, cg);
				pi += m[j] * cp;
				for (int d = 0; d < 3; d++) gi[d] += m[j] * cg[d];
			}
		}
		phi[i] = pi;
		for (int d = 0; d < 3; d++) g[3 * i + d] = gi[d];
	}
}

/* ========================================================================================== */
/* One force solve over given interaction lists, the reference's way: node by node (parents     */
/* first) L = L_parent << dX; L += CC_d, CP_d, CC_e, CP_e; at each leaf f = L2P(L) then         */
/* += PC_d, PP_d, PC_e, PP_e (src/tree.cpp:400-401,481-491,603-626).                            */

/* expansion<double>::operator<<= (tigergrav/expansion.hpp:185-233), in place */
static void l2l_ref(double* me, const double* dX) {
	for (int a = 0; a < 3; a++) {
		me[0] += me[E1(a)] * dX[a];
		for (int b = 0; b <= a; b++) {
			me[0] += me[E2(a, b)] * dX[a] * dX[b] * EF64[E2(a, b)];
			for (int c = 0; c <= b; c++) {
				me[0] += me[E3(a, b, c)] * dX[a] * dX[b] * dX[c] * EF64[E3(a, b, c)];
				for (int d = 0; d <= c; d++) me[0] += me[E4(a, b, c, d)] * dX[a] * dX[b] * dX[c] * dX[d] * EF64[E4(a, b, c, d)];
			}
		}
	}
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++) {
			me[E1(a)] += me[E2(a, b)] * dX[b];
			for (int c = 0; c <= b; c++) {
				me[E1(a)] += me[E3(a, b, c)] * dX[b] * dX[c] * EF64[E2(b, c)];
				for (int d = 0; d <= c; d++) me[E1(a)] += me[E4(a, b, c, d)] * dX[b] * dX[c] * dX[d] * EF64[E3(b, c, d)];
			}
		}
	for (int a = 0; a < 3; a++)
		for (int b = 0; b <= a; b++)
			for (int c = 0; c < 3; c++) {
				me[E2(a, b)] += me[E3(a, b, c)] * dX[c];
				for (int d = 0; d <= c; d++) me[E2(a, b)] += me[E4(a, b, c, d)] * dX[c] * dX[d] * EF64[E2(c, d)];
			}
	for (int a = 0; a < 3; a++)
		for (int b = 0; b <= a; b++)
			for (int c = 0; c <= b; c++)
				for (int d = 0; d < 3; d++) me[E3(a, b, c)] += me[E4(a, b, c, d)] * dX[d];
}

/* expansion<double>::translate_L2 (tigergrav/expansion.hpp:236-269) */
static void l2p_ref(const double* L, const double* dX, tgo_force* f) {
	double phi = L[0];
	for (int a = 0; a < 3; a++) {
		phi += L[E1(a)] * dX[a];
		for (int b = a; b < 3; b++) {
			phi += L[E2(a, b)] * dX[a] * dX[b] * EF64[E2(a, b)];
			for (int c = b; c < 3; c++) {
				phi += L[E3(a, b, c)] * dX[a] * dX[b] * dX[c] * EF64[E3(a, b, c)];
				for (int d = c; d < 3; d++) phi += L[E4(a, b, c, d)] * dX[a] * dX[b] * dX[c] * dX[d] * EF64[E4(a, b, c, d)];
			}
		}
	}
	f->phi = phi;
	for (int a = 0; a < 3; a++) {
		double g = -L[E1(a)];
		for (int b = 0; b < 3; b++) {
			g -= L[E2(a, b)] * dX[b];
			for (int c = b; c < 3; c++) {
				g -= L[E3(a, b, c)] * dX[b] * dX[c] * EF64[E2(b, c)];
				for (int d = c; d < 3; d++) g -= L[E4(a, b, c, d)] * dX[b] * dX[c] * dX[d] * EF64[E3(b, c, d)];
			}
		}
		f->g[a] = g;
	}
}

void tgo_l2l(double* L, const double* dX) { l2l_ref(L, dX); }
void tgo_l2p(const double* L, const double* dX, tgo_force* f) { l2p_ref(L, dX, f); }

static int32_t* gather_pos(const int32_t* pos, const int32_t* ranges, int64_t b, int64_t e, int64_t* n) {
	int64_t tot = 0;
	for (int64_t k = b; k < e; k += 2) tot += ranges[k + 1] - ranges[k];
	int32_t* out = malloc((size_t) (tot > 0 ? tot : 1) * 3 * sizeof(int32_t));
	int64_t o = 0;
	for (int64_t k = b; k < e; k += 2)
		for (int32_t p = ranges[k]; p < ranges[k + 1]; p++, o++) memcpy(out + 3 * o, pos + 3 * (int64_t) p, 12);
	*n = tot;
	return out;
}

static tgo_multi_src* gather_multi(const int32_t* node_x, const float* node_m, const int32_t* ids, int64_t b, int64_t e) {
	tgo_multi_src* out = malloc((size_t) (e > b ? e - b : 1) * sizeof(tgo_multi_src));
	for (int64_t k = b; k < e; k++) {
		memcpy(out[k - b].m, node_m + 17 * (int64_t) ids[k], 17 * sizeof(float));
		memcpy(out[k - b].x, node_x + 3 * (int64_t) ids[k], 12);
	}
	return out;
}

/* kinds: 0 cc_d, 1 cp_d, 2 cc_e, 3 cp_e, 4 pc_d, 5 pp_d, 6 pc_e, 7 pp_e.  Nodes are ordered with
 * parents before children.  With stride > 1 only every stride-th node's node-kind work and every
 * stride-th leaf's leaf-kind work is done (a timing sample; results are then partial).  With want
 * != NULL only nodes with want[n] != 0 are evaluated (their L2L, node kinds and, for leaves, leaf
 * kinds); marking a set of leaves and all their ancestors gives those leaves' exact results.
 * Returns the reference-weighted flop count. */
static uint64_t solve_impl(const tgo_params* P, int64_t N, const int32_t* pos, int64_t nn, const int32_t* node_range,
		const int32_t* node_child, const int32_t* node_parent, const int32_t* node_x, const float* node_m,
		const int64_t* const* ptr, const int32_t* const* lst, int stride, const uint8_t* want, tgo_force* out);

uint64_t tgo_solve(const tgo_params* P, int64_t N, const int32_t* pos, int64_t nn, const int32_t* node_range,
		const int32_t* node_child, const int32_t* node_parent, const int32_t* node_x, const float* node_m,
		const int64_t* const* ptr, const int32_t* const* lst, int stride, tgo_force* out, uint64_t* counts) {
	(void) counts;
	return solve_impl(P, N, pos, nn, node_range, node_child, node_parent, node_x, node_m, ptr, lst, stride, NULL, out);
}

uint64_t tgo_solve_want(const tgo_params* P, int64_t N, const int32_t* pos, int64_t nn, const int32_t* node_range,
		const int32_t* node_child, const int32_t* node_parent, const int32_t* node_x, const float* node_m,
		const int64_t* const* ptr, const int32_t* const* lst, const uint8_t* want, tgo_force* out) {
	return solve_impl(P, N, pos, nn, node_range, node_child, node_parent, node_x, node_m, ptr, lst, 1, want, out);
}

static uint64_t solve_impl(const tgo_params* P, int64_t N, const int32_t* pos, int64_t nn, const int32_t* node_range,
		const int32_t* node_child, const int32_t* node_parent, const int32_t* node_x, const float* node_m,
		const int64_t* const* ptr, const int32_t* const* lst, int stride, const uint8_t* want, tgo_force* out) {
	(void) N;
	double* L = calloc((size_t) nn * 35, sizeof(double));
	uint64_t flop = 0, cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	/* node pass: levels are contiguous runs of nodes whose parents come earlier */
	int64_t start = 0;
	while (start < nn) {
		int64_t end = start + 1;
		/* nodes in [start, end) have parents < start */
		while (end < nn && node_parent[end] < start) end++;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : flop)
		for (int64_t n = start; n < end; n++) {
			if (node_range[2 * n] == node_range[2 * n + 1]) continue;
			if (want && !want[n]) continue;
			double* Ln = L + 35 * n;
			const int32_t p = node_parent[n];
			if (p >= 0) {
				memcpy(Ln, L + 35 * (int64_t) p, 35 * sizeof(double));
				double dX[3];
				for (int d = 0; d < 3; d++)
					dX[d] = ((double) node_x[3 * n + d] * POS_INV + 0.5) - ((double) node_x[3 * (int64_t) p + d] * POS_INV + 0.5);
				l2l_ref(Ln, dX);
			}
			if (stride > 1 && n % stride) continue;
			for (int k = 0; k < 4; k++) {
				const int64_t b = ptr[k][n], e = ptr[k][n + 1];
				if (e == b) continue;
				if (k == 0 || k == 2) {
					tgo_multi_src* y = gather_multi(node_x, node_m, lst[k], b, e);
					flop += (k == 0 ? tgo_CC_direct : tgo_CC_ewald)(P, Ln, node_x + 3 * n, y, e - b);
					free(y);
				} else {
					int64_t nj;
					int32_t* y = gather_pos(pos, lst[k], b, e, &nj);
					flop += (k == 1 ? tgo_CP_direct : tgo_CP_ewald)(P, Ln, node_x + 3 * n, y, nj);
					free(y);
				}
			}
		}
		start = end;
	}
	/* leaf pass */
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : flop)
	for (int64_t n = 0; n < nn; n++) {
		if (node_child[2 * n] >= 0) continue;
		const int32_t pb = node_range[2 * n], pe = node_range[2 * n + 1];
		if (pb == pe) continue;
		if (stride > 1 && n % stride) continue;
		if (want && !want[n]) continue;
		const int64_t ns = pe - pb;
		tgo_force* f = out + pb;
		const int32_t* x = pos + 3 * (int64_t) pb;
		for (int64_t i = 0; i < ns; i++) {
			double dX[3];
			for (int d = 0; d < 3; d++)
				dX[d] = (double) (float) (((double) x[3 * i + d] * POS_INV + 0.5) - ((double) node_x[3 * n + d] * POS_INV + 0.5));
			l2p_ref(L + 35 * n, dX, &f[i]);
		}
		for (int k = 4; k < 8; k++) {
			const int64_t b = ptr[k][n], e = ptr[k][n + 1];
			if (k == 6 && e == b) continue;
			if (k == 7 && e == b) continue;
			if (k == 4 || k == 6) {
				tgo_multi_src* y = gather_multi(node_x, node_m, lst[k], b, e);
				flop += (k == 4 ? tgo_PC_direct : tgo_PC_ewald)(P, f, x, ns, y, e - b);
				free(y);
			} else {
				int64_t nj;
				int32_t* y = gather_pos(pos, lst[k], b, e, &nj);
				if (k == 5)
					flop += tgo_PP_direct(P, f, x, ns, y, nj, 1);
				else
					flop += tgo_PP_ewald(P, f, x, ns, y, nj);
				free(y);
			}
		}
	}
	(void) cnt;
	free(L);
	return flop;
}

/* ========================================================================================== */
/* FP64 frozen-tree solve: the function the batched solve evaluates on a kept tree (the adjoint's  */
/* function), every step in FP64.  Moments about the kept centres node_x (no dipole slot, as the   */
/* reference's multipole has none): leaves by P2M of their particles' offsets (wrap-around in      */
/* periodic mode), other nodes by the raw-moment shift of their children (tigergrav/multipole.hpp: */
/* 117-135); then the reference's solve sequence (src/tree.cpp:400-401,481-491,603-626) with the   */
/* tgo64_ formulas, FP64 moments and unrounded L2P offsets.  Smooth in the positions, so central   */
/* differences of it are the derivative to FP64 precision.                                         */
static void shift_moments64(double* M, const double* Y) {
	for (int p = 0; p < 3; p++)
		for (int q = p; q < 3; q++)
			for (int l = q; l < 3; l++)
				M[M3(p, q, l)] += M[0] * Y[p] * Y[q] * Y[l] + M[M2(p, q)] * Y[l] + M[M2(q, l)] * Y[p] + M[M2(l, p)] * Y[q];
	for (int p = 0; p < 3; p++)
		for (int q = p; q < 3; q++) M[M2(p, q)] += M[0] * Y[p] * Y[q];
}

/* the spline-softened pair (src/gravity.cpp:231-287) in FP64: g += -m dx f(r), phi += -m p(r) */
static void spline_pair64(double h, double m, const double* dx, tgo_force* f) {
	const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
	double fr, pr;
	if (r > h) {
		fr = 1.0 / (r * r * r);
		pr = 1.0 / r;
	} else {
		const double q = r / h;
		if (r < 0.5 * h) {
			fr = (32.0 / 3.0 - 192.0 / 5.0 * q * q + 32.0 * q * q * q) / (h * h * h);
			pr = (14.0 / 5.0 - 16.0 / 3.0 * q * q + 48.0 / 5.0 * q * q * q * q - 32.0 / 5.0 * q * q * q * q * q) / h;
		} else {
			const double q3 = q * q * q;
			fr = (64.0 / 3.0 * q3 - 48.0 * q3 * q + 192.0 / 5.0 * q3 * q * q - 32.0 / 3.0 * q3 * q3 - 1.0 / 15.0) / (r * r * r);
			pr = (-1.0 / 15.0 + 16.0 / 5.0 * q - 32.0 / 3.0 * q3 + 16.0 * q3 * q - 48.0 / 5.0 * q3 * q * q + 32.0 / 15.0 * q3 * q3) / r;
		}
	}
	for (int d = 0; d < 3; d++) f->g[d] -= dx[d] * m * fr;
	f->phi -= m * pr;
}

void tgo64_frozen_solve(const tgo_params* P, int64_t N, const int32_t* pos, const double* mass, int64_t nn,
		const int32_t* node_range, const int32_t* node_child, const int32_t* node_parent, const int32_t* node_x,
		const int64_t* const* ptr, const int32_t* const* lst, tgo_force* out) {
	double* M = calloc((size_t) nn * 17, sizeof(double));
	double* L = calloc((size_t) nn * 35, sizeof(double));
#define MASS(i) (mass ? mass[i] : P->m)
	/* moments, children before parents (nodes are numbered parents first) */
	for (int64_t n = nn - 1; n >= 0; n--) {
		double* Mn = M + 17 * n;
		const int32_t pb = node_range[2 * n], pe = node_range[2 * n + 1];
		if (pb == pe) continue;
		if (node_child[2 * n] < 0) {
			for (int32_t i = pb; i < pe; i++) {
				double dx[3];
				sep64(P->ewald, pos + 3 * (int64_t) i, node_x + 3 * n, dx);
				const double mi = MASS(i);
				Mn[0] += mi;
				for (int a = 0; a < 3; a++)
					for (int b = a; b < 3; b++) {
						Mn[M2(a, b)] += mi * dx[a] * dx[b];
						for (int c = b; c < 3; c++) Mn[M3(a, b, c)] += mi * dx[a] * dx[b] * dx[c];
					}
			}
		} else {
			for (int k = 0; k < 2; k++) {
				const int32_t c = node_child[2 * n + k];
				double Mc[17], Y[3];
				memcpy(Mc, M + 17 * (int64_t) c, sizeof(Mc));
				for (int d = 0; d < 3; d++) Y[d] = ((double) node_x[3 * (int64_t) c + d] - (double) node_x[3 * n + d]) * POS_INV;
				shift_moments64(Mc, Y);
				for (int q = 0; q < 17; q++) Mn[q] += Mc[q];
			}
		}
	}
	/* node pass, level by level */
	int64_t start = 0;
	while (start < nn) {
		int64_t end = start + 1;
		while (end < nn && node_parent[end] < start) end++;
#pragma omp parallel for schedule(dynamic, 1)
		for (int64_t n = start; n < end; n++) {
			if (node_range[2 * n] == node_range[2 * n + 1]) continue;
			double* Ln = L + 35 * n;
			const int32_t p = node_parent[n];
			if (p >= 0) {
				memcpy(Ln, L + 35 * (int64_t) p, 35 * sizeof(double));
				double dX[3];
				for (int d = 0; d < 3; d++) dX[d] = ((double) node_x[3 * n + d] - (double) node_x[3 * (int64_t) p + d]) * POS_INV;
				l2l_ref(Ln, dX);
			}
			for (int k = 0; k < 4; k++) {
				const int ew = k >= 2;
				for (int64_t q = ptr[k][n]; q < ptr[k][n + 1]; q += (k & 1) ? 2 : 1) {
					if (k & 1) { /* CP: particle ranges */
						for (int32_t j = lst[k][q]; j < lst[k][q + 1]; j++) {
							double dx[3], D[35];
							sep64(ew ? 1 : P->ewald, node_x + 3 * n, pos + 3 * (int64_t) j, dx);
							if (ew) tgo64_green_ewald(dx, D); else tgo64_green_direct(dx, D);
							const double mj = MASS(j);
							for (int r = 0; r < 35; r++) Ln[r] += mj * D[r];
						}
					} else { /* CC: source nodes */
						const int32_t c = lst[k][q];
						double dx[3], D[35];
						sep64(ew ? 1 : P->ewald, node_x + 3 * n, node_x + 3 * (int64_t) c, dx);
						if (ew) tgo64_green_ewald(dx, D); else tgo64_green_direct(dx, D);
						m2l64d(Ln, M + 17 * (int64_t) c, D);
					}
				}
			}
		}
		start = end;
	}
	/* leaf pass */
#pragma omp parallel for schedule(dynamic, 1)
	for (int64_t n = 0; n < nn; n++) {
		if (node_child[2 * n] >= 0) continue;
		const int32_t pb = node_range[2 * n], pe = node_range[2 * n + 1];
		for (int32_t i = pb; i < pe; i++) {
			double dX[3];
			sep64(P->ewald, pos + 3 * (int64_t) i, node_x + 3 * n, dX);
			l2p_ref(L + 35 * n, dX, &out[i]);
			out[i].phi -= (-315.0 / 128.0) * P->m / P->soft_len;
			for (int k = 4; k < 8; k++) {
				const int ew = k >= 6;
				for (int64_t q = ptr[k][n]; q < ptr[k][n + 1]; q += (k & 1) ? 2 : 1) {
					if (k & 1) { /* PP: particle ranges */
						for (int32_t j = lst[k][q]; j < lst[k][q + 1]; j++) {
							double dx[3];
							sep64(ew ? 1 : P->ewald, pos + 3 * (int64_t) i, pos + 3 * (int64_t) j, dx);
							if (ew) {
								double D[35];
								tgo64_green_ewald(dx, D);
								out[i].phi += MASS(j) * D[0];
								for (int d = 0; d < 3; d++) out[i].g[d] -= MASS(j) * D[E1(d)];
							} else {
								spline_pair64(P->soft_len, MASS(j), dx, &out[i]);
							}
						}
					} else { /* PC: source nodes */
						const int32_t c = lst[k][q];
						double dx[3], D[35];
						sep64(ew ? 1 : P->ewald, pos + 3 * (int64_t) i, node_x + 3 * (int64_t) c, dx);
						if (ew) tgo64_green_ewald(dx, D); else tgo64_green_direct(dx, D);
						m2p64d(&out[i].phi, out[i].g, M + 17 * (int64_t) c, D);
					}
				}
			}
		}
	}
#undef MASS
	(void) N;
	free(M);
	free(L);
}

/* ------------------------------------------------------------------------------------------ */
/* time step: part_vect_kick (src/part_vect.cpp:341-436) and part_vect_drift (src/part_vect.cpp:
 * 493-540) over one particle range, in particle order, with the time grid of tigergrav/time.hpp.
 * Every particle's force is given (f[i]); the reference compacts them to the active ones, which
 * reads the same values.  float here is the reference's vect<float>, double its vect<double>. */
static double tgo_rung_to_dt(double t_max, int rung) { return t_max / (double) (1 << rung); } /* time.hpp:36-39 */

static int tgo_dt_to_rung(double t_max, double dt) { /* time.hpp:41-48; -1 past rung 30 */
	int rung = 0;
	while (tgo_rung_to_dt(t_max, rung) > dt) {
		rung++;
		if (rung > 30) return -1;
	}
	return rung;
}

int tgo_kick(const tgo_step_params* P, int64_t n, const tgo_force* f, float* v, int8_t* rung, int min_rung, int do_out,
		double* stats /* g[3], p[3], pot, kin, max_rung */) {
	const double eps = 10.0 * 2.2250738585072014e-308; /* 10 numeric_limits<double>::min() */
	const double scale = P->scale;
	const double a3inv = 1.0 / (scale * scale * scale);
	const double m = P->m;
	const double sgn = P->glass ? -1.0 : 1.0;
	const double glass_drag = 1.0;
	double g[3] = {0, 0, 0}, p[3] = {0, 0, 0}, pot = 0.0, kin = 0.0;
	int rc_rung = 0, bad = 0;
	for (int64_t i = 0; i < n; i++) {
		float* vi = v + 3 * i;
		/* p = p + v / problem_size / (scale scale): vect<float> / float, / float, then to double */
		for (int d = 0; d < 3; d++) p[d] = p[d] + (double) ((vi[d] / (float) P->problem_size) / (float) (scale * scale));
		float dot = vi[0] * vi[0];
		dot += vi[1] * vi[1];
		dot += vi[2] * vi[2];
		kin += 0.5 * m * (double) dot / (scale * scale);
		const int r = rung[i];
		if (r >= min_rung || do_out) {
			if (r >= min_rung) {
				if (r != 0) {
					const double dt = tgo_rung_to_dt(P->t_max, r);
					const double dt1 = P->cosmic ? P->kick_dt1[r] : 0.5 * dt;
					for (int d = 0; d < 3; d++) vi[d] = vi[d] + (float) (f[i].g[d] * (dt1 * sgn * P->G));
					if (P->glass)
						for (int d = 0; d < 3; d++) vi[d] = vi[d] / (float) (1.0 + glass_drag * dt1);
				}
				double gG[3];
				for (int d = 0; d < 3; d++) gG[d] = f[i].g[d] * P->G;
				double gg = gG[0] * gG[0];
				gg += gG[1] * gG[1];
				gg += gG[2] * gG[2];
				const double a = sqrt(gg) * a3inv;
				double dt = P->eta * sqrt(P->soft_len * (1.0 / 2.8) / (a + eps));
				if (P->dt_max < dt) dt = P->dt_max; /* std::min(dt_max, dt) */
				int nr = tgo_dt_to_rung(P->t_max, dt);
				if (nr < 0) {
					bad = 1;
					nr = 30;
				}
				if (nr < min_rung) nr = min_rung;
				if (nr > rc_rung) rc_rung = nr;
				dt = tgo_rung_to_dt(P->t_max, nr);
				int st = nr > r - 1 ? nr : r - 1;
				if (st < 1) st = 1;
				rung[i] = (int8_t) st;
				const double dt2 = P->ewald ? P->kick_dt2[st] : 0.5 * dt;
				for (int d = 0; d < 3; d++) vi[d] = vi[d] + (float) (f[i].g[d] * (dt2 * sgn * P->G));
				if (P->glass)
					for (int d = 0; d < 3; d++) vi[d] = vi[d] / (float) (1.0 + glass_drag * dt2);
			}
			if (do_out) {
				for (int d = 0; d < 3; d++) g[d] = g[d] + f[i].g[d] * P->G / P->problem_size / (scale * scale * scale);
				pot += 0.5 * m * f[i].phi * P->G / scale;
			}
		}
	}
	for (int d = 0; d < 3; d++) {
		stats[d] = g[d];
		stats[3 + d] = p[d];
	}
	stats[6] = pot;
	stats[7] = kin;
	stats[8] = rc_rung;
	return bad ? -1 : 0;
}

static int32_t tgo_double_to_pos(double x) { /* tigergrav/position.hpp:30-37 */
	const double POS_MAX = 4294967296.0;
	const double y = x - 0.5;
	long z = lround(y * POS_MAX);
	if (z >= POS_MAX / 2.0) z -= (long) POS_MAX;
	return (int32_t) z;
}

double tgo_drift(const tgo_step_params* P, int64_t n, int32_t* x, const float* v, int max_rung, double drift_dt, double a1) {
	const double m = P->m;
	const double a1inv2 = 1.0 / (a1 * a1);
	const double dt = P->cosmic ? drift_dt : tgo_rung_to_dt(P->t_max, max_rung);
	double ekin = 0.0;
	for (int64_t j = 0; j < n; j++) {
		const float* vj = v + 3 * j;
		float dot = vj[0] * vj[0];
		dot += vj[1] * vj[1];
		dot += vj[2] * vj[2];
		ekin += 0.5 * m * (double) dot * a1inv2;
		for (int d = 0; d < 3; d++) {
			const double dx = (double) (vj[d] * (float) dt); /* vect<float> * float */
			double xd = (double) x[3 * j + d] * POS_INV + 0.5;
			xd += dx;
			while (xd >= 1.0) xd -= 1.0;
			while (xd < 0.0) xd += 1.0;
			x[3 * j + d] = tgo_double_to_pos(xd);
		}
	}
	return ekin;
}

/* ------------------------------------------------------------------------------------------ */
/* friends-of-friends: part_vect_find_groups (src/part_vect.cpp:84-120) over all pairs, repeated
 * until no particle changes (the reference's groups_execute_finders loop, src/main.cpp:41-45).
 * pos is tree order; group[i] starts at DEFAULT_GROUP; a first link gives a particle its own index,
 * then each link takes the smaller of the two ids.  O(n^2) per sweep: small n only.  A pair whose
 * coordinate difference wraps (across the box edge) is not linked: the reference searches
 * unwrapped ranges there (src/range.cpp:80-116), which leaves such pairs to the kd boxes' reach. */
void tgo_fof(int64_t n, const int32_t* pos, uint64_t L2, uint64_t* group) {
	const uint64_t NONE = 0xFFFFFFFFFFFFFFull;
	for (int64_t i = 0; i < n; i++) group[i] = NONE;
	int changed = 1;
	while (changed) {
		changed = 0;
		for (int64_t i = 0; i < n; i++) {
			for (int64_t j = 0; j < n; j++) {
				if (j == i) continue;
				uint64_t d2 = 0;
				int wraps = 0;
				for (int d = 0; d < 3; d++) {
					const int32_t dx = (int32_t) ((uint32_t) pos[3 * i + d] - (uint32_t) pos[3 * j + d]);
					const uint64_t u = (uint64_t) (int64_t) dx;
					d2 += u * u;
					wraps |= (int64_t) pos[3 * i + d] - (int64_t) pos[3 * j + d] != (int64_t) dx;
				}
				if (d2 < L2 && d2 != 0 && !wraps) {
					uint64_t id = group[i];
					if (id == NONE) {
						id = (uint64_t) i;
						group[i] = id;
						changed = 1;
					}
					const uint64_t other = group[j];
					if (id != other) {
						const uint64_t g = id < other ? id : other;
						if (group[i] != g) changed = 1;
						group[i] = g;
					}
				}
			}
		}
	}
}
