// tg_ewald_table.cpp -- FP64 evaluation of the periodic correction and its derivatives to order
// TG_EWT_ORDER, and the table the device kernels interpolate (tg_ewald_table.hpp).  Host code.
#include "tg_ewald_table.hpp"

#include <cmath>
#include <utility>

#include "tg_gravity.h"

namespace tg {

namespace {

// ways to pair p couples among n equal indices: n! / (2^p p! (n - 2p)!)
double pairings(int n, int p) {
	double num = 1.0, den = 1.0;
	for (int i = 0; i < 2 * p; i++) num *= double(n - i);
	for (int i = 1; i <= p; i++) den *= 2.0 * double(i);
	return num / den;
}

// The derivative tensor of a radial function f(|y|) with reduced derivatives g_j = (r^-1 d/dr)^j f:
// component (n0, n1, n2) is the sum over p_i couples of equal indices of
// g_(n - p0 - p1 - p2) prod_i pairings(n_i, p_i) y_i^(n_i - 2 p_i).  Its terms, per slot:
struct RadialTerm {
	double c;
	int e0, e1, e2, g;
};
struct RadialTerms {
	std::vector<RadialTerm> t;
	int begin[TG_EWT_S + 1];
	RadialTerms() {
		for (int k = 0; k < TG_EWT_S; k++) {
			begin[k] = int(t.size());
			const int n0 = ewt::ST.c[k].n0, n1 = ewt::ST.c[k].n1, n2 = ewt::ST.c[k].n2, n = n0 + n1 + n2;
			for (int p0 = 0; 2 * p0 <= n0; p0++)
				for (int p1 = 0; 2 * p1 <= n1; p1++)
					for (int p2 = 0; 2 * p2 <= n2; p2++)
						t.push_back({pairings(n0, p0) * pairings(n1, p1) * pairings(n2, p2), n0 - 2 * p0, n1 - 2 * p1, n2 - 2 * p2,
								n - p0 - p1 - p2});
		}
		begin[TG_EWT_S] = int(t.size());
	}
};
const RadialTerms& radial_terms() {
	static const RadialTerms R;
	return R;
}

void radial_add(double* D, const double* g, const double* y) {
	const RadialTerms& R = radial_terms();
	double pw[3][TG_EWT_ORDER + 1];
	for (int a = 0; a < 3; a++) {
		pw[a][0] = 1.0;
		for (int e = 1; e <= TG_EWT_ORDER; e++) pw[a][e] = pw[a][e - 1] * y[a];
	}
	for (int k = 0; k < TG_EWT_S; k++) {
		double acc = 0.0;
		for (int q = R.begin[k]; q < R.begin[k + 1]; q++) {
			const RadialTerm& T = R.t[size_t(q)];
			acc += T.c * pw[0][T.e0] * pw[1][T.e1] * pw[2][T.e2] * g[T.g];
		}
		D[k] += acc;
	}
}

// Fourier vectors h != 0, |h|^2 <= 9 (tigergrav/green.hpp:30-59): the term -(c0/pi) cos(2 pi h.x),
// c0 = exp(-pi^2 h^2 / 4) / h^2, has derivative (n0, n1, n2) = -(c0/pi) (2 pi)^n h^(n0,n1,n2) cos^(n)
struct FourierTerms {
	std::vector<double> h;     // 3 per vector
	std::vector<double> coef;  // TG_EWT_S per vector
	FourierTerms() {
		for (int i = -3; i <= 3; i++)
			for (int j = -3; j <= 3; j++)
				for (int l = -3; l <= 3; l++) {
					const int h2 = i * i + j * j + l * l;
					if (h2 == 0 || h2 > 9) continue;
					const double c0 = std::exp(-M_PI * M_PI * h2 / 4.0) / h2;
					const double hv[3] = {double(i), double(j), double(l)};
					h.insert(h.end(), hv, hv + 3);
					for (int k = 0; k < TG_EWT_S; k++) {
						const int n0 = ewt::ST.c[k].n0, n1 = ewt::ST.c[k].n1, n2 = ewt::ST.c[k].n2, n = n0 + n1 + n2;
						coef.push_back(-(c0 / M_PI) * std::pow(2.0 * M_PI, n) * std::pow(hv[0], n0) * std::pow(hv[1], n1) *
								std::pow(hv[2], n2));
					}
				}
	}
};
const FourierTerms& fourier_terms() {
	static const FourierTerms F;
	return F;
}

}  // namespace

void ewald_green64(const double* x, double* D) {
	for (int k = 0; k < TG_EWT_S; k++) D[k] = 0.0;
	const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
	double g[TG_EWT_ORDER + 1];
	{
		// the n = 0 image with the Newtonian term it corrects: erf(2r)/r = sum_q c_q u^q, u = r^2,
		// c_q = (2/sqrt(pi)) (-1)^q 2^(2q+1) / (q! (2q+1)); (r^-1 d/dr)^k = sum_q c_q (2q)(2q-2)...(2q-2k+2) u^(q-k)
		const int NQ = 60;
		for (int k = 0; k <= TG_EWT_ORDER; k++) {
			double acc = 0.0, upow = 1.0, fact = 1.0;
			for (int q = 0; q < NQ; q++) {
				if (q > 0) fact *= double(q);
				if (q >= k) {
					double c = 2.0 / std::sqrt(M_PI) * ((q & 1) ? -1.0 : 1.0) * std::ldexp(1.0, 2 * q + 1) / fact / double(2 * q + 1);
					for (int t = 0; t < k; t++) c *= double(2 * q - 2 * t);
					acc += c * upow;
					upow *= r2;
				}
			}
			g[k] = acc;
		}
		radial_add(D, g, x);
	}
	// real-space images n != 0, |n|^2 <= 17, |x - n| < 3.6 (tigergrav/green.hpp:3-26,197-251):
	// -erfc(2r)/r, whose reduced derivatives are (-1)^(k+1) B_k with B_0 = erfc(2r)/r and
	// B_k = ((2k - 1) B_(k-1) + 8^k / (2 sqrt(pi)) exp(-4 r^2)) / r^2
	for (int i = -4; i <= 4; i++)
		for (int j = -4; j <= 4; j++)
			for (int l = -4; l <= 4; l++) {
				const int n2 = i * i + j * j + l * l;
				if (n2 == 0 || n2 > 17) continue;
				const double y[3] = {x[0] - i, x[1] - j, x[2] - l};
				const double q2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
				const double q = std::sqrt(q2);
				if (!(q < 3.6)) continue;
				const double ex = std::exp(-4.0 * q2) / (2.0 * std::sqrt(M_PI));
				double B = std::erfc(2.0 * q) / q, p8 = 1.0;
				g[0] = -B;
				for (int k = 1; k <= TG_EWT_ORDER; k++) {
					p8 *= 8.0;
					B = (double(2 * k - 1) * B + p8 * ex) / q2;
					g[k] = (k & 1) ? B : -B;
				}
				radial_add(D, g, y);
			}
	// Fourier terms: d^n/dt^n cos t = cos, -sin, -cos, sin in turn
	const FourierTerms& F = fourier_terms();
	for (size_t v = 0; v < F.h.size() / 3; v++) {
		const double w = 2.0 * M_PI * (F.h[3 * v] * x[0] + F.h[3 * v + 1] * x[1] + F.h[3 * v + 2] * x[2]);
		const double cs = std::cos(w), sn = std::sin(w);
		const double trig[4] = {cs, -sn, -cs, sn};
		const double* c = F.coef.data() + v * TG_EWT_S;
		for (int k = 0; k < TG_EWT_S; k++) {
			const int n = ewt::ST.c[k].n0 + ewt::ST.c[k].n1 + ewt::ST.c[k].n2;
			D[k] += c[k] * trig[n & 3];
		}
	}
	D[0] += M_PI / 4.0;
}

void ewald_table_build(std::vector<float>& T) {
	T.assign(size_t(TG_EWT_ROWS) * TG_EWT_STRIDE, 0.0f);
	T[size_t(TG_EWT_P) * TG_EWT_STRIDE] = 2.8372975f;  // the coincident-pair row (tigergrav/green.hpp:272)
	radial_terms();  // the term tables, before the threads read them
	fourier_terms();
	// the points with i >= j >= l; the kernels map every separation onto them by symmetry
#pragma omp parallel for schedule(dynamic, 1)
	for (int i = 0; i < TG_EWT_N; i++)
		for (int j = 0; j <= i; j++)
			for (int l = 0; l <= j; l++) {
				const double x[3] = {double(i) / TG_EWT_INV, double(j) / TG_EWT_INV, double(l) / TG_EWT_INV};
				double D[TG_EWT_S];
				ewald_green64(x, D);
				float* row = T.data() + size_t(ewt::wedge_index(i, j, l)) * TG_EWT_STRIDE;
				for (int k = 0; k < TG_EWT_S; k++)
					if (ewt::stored_index(k) >= 0) row[ewt::stored_index(k)] = float(D[k]);
			}
}

}  // namespace tg

extern "C" int tg_ewald_green64(const double* x, int64_t n, double* D) {
	if (n < 0 || (n > 0 && (!x || !D))) return TG_EINVAL;
	for (int64_t i = 0; i < n; i++) tg::ewald_green64(x + 3 * i, D + TG_EWT_S * i);
	return TG_OK;
}
