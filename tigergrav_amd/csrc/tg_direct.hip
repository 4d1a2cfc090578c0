// tg_direct.hip -- FP64 direct summation: the accuracy reference for a force solve.
//
// The reference checks a solve by comparing it with a direct sum over all particles
// (--solver_test, src/main.cpp:137-164, error statistics src/output.cpp:17-51).  This is that
// direct sum, in FP64, for a sample of sink particles against all N sources:
//   open:      spline-softened Newton (src/gravity.cpp:231-287, no self-potential offset);
//   periodic:  spline-softened Newton on the nearest image plus the converged Ewald correction
//              of the periodic Green's function minus 1/r (the function tigergrav's Ewald kernels
//              truncate, tigergrav/green.hpp:189-286), with its +pi/alpha^2 constant.
// The Ewald sum is split at alpha = 6: the real-space part runs over the 27 images with
// |n_i| <= 1 (every omitted image is >= 1.5 away, erfc(9) ~ 1e-37) with a cutoff at r = 1
// (erfc(6) ~ 2e-17); the Fourier part over 0 < |h|^2 <= 144 (exp(-pi^2 145 / 36) ~ 5e-18),
// evaluated through structure factors C(h) = sum_j m_j cos(2 pi h.x_j), S(h) = sum_j m_j sin(...),
// one half-space of h with a factor 2.  Every sum has a fixed order: results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "tg_direct.h"
#include "tg_gravity.h"

namespace {

constexpr double kAlpha = 6.0;
constexpr int kHMax2 = 144;
constexpr int kSFChunks = 128;   // particle chunks of the structure-factor pass
constexpr int kSrcChunks = 64;   // source chunks of the pair pass
constexpr double kPosInv = 1.0 / 4294967296.0;

thread_local std::string g_direct_err;

__device__ __forceinline__ double sep(int ewald, int a, int b) {
	return ewald ? double(int(unsigned(a) - unsigned(b))) * kPosInv : (double(a) - double(b)) * kPosInv;
}

// spline-softened Newton: f (g = -m dx f) and p (phi = -m p), as the reference's three branches
__device__ __forceinline__ void spline64(double r, double h, double& f, double& p) {
	if (r >= h) {
		f = 1.0 / (r * r * r);
		p = 1.0 / r;
		return;
	}
	const double q = r / h, q2 = q * q, q3 = q2 * q;
	if (r < 0.5 * h) {
		f = (32.0 / 3.0 - 192.0 / 5.0 * q2 + 32.0 * q3) / (h * h * h);
		p = (14.0 / 5.0 - 16.0 / 3.0 * q2 + 48.0 / 5.0 * q2 * q2 - 32.0 / 5.0 * q3 * q2) / h;
	} else {
		f = (64.0 / 3.0 * q3 - 48.0 * q3 * q + 192.0 / 5.0 * q3 * q2 - 32.0 / 3.0 * q3 * q3 - 1.0 / 15.0) / (r * r * r);
		p = (-1.0 / 15.0 + 16.0 / 5.0 * q - 32.0 / 3.0 * q3 + 16.0 * q3 * q - 48.0 / 5.0 * q3 * q2 + 32.0 / 15.0 * q3 * q3) / r;
	}
}

// the real-space part of the correction for one pair at nearest-image offset x:
//   U(r0) - sum_{n != 0, r_n < 1} erfc(a r_n)/r_n,  U(r) = erf(a r)/r  (= erfc(a r)/r's n = 0 term minus -1/r)
// returned as the potential term u and d = (1/r) du/dr per image folded into the gradient.
__device__ void real_space(const double* x, double& u, double* du) {
	const double a = kAlpha, two_over_sqrtpi = 1.1283791670955126;
	u = 0.0;
	du[0] = du[1] = du[2] = 0.0;
	for (int i = -1; i <= 1; i++)
		for (int j = -1; j <= 1; j++)
			for (int k = -1; k <= 1; k++) {
				const double y[3] = {x[0] + i, x[1] + j, x[2] + k};
				const double r2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
				double val, dr;  // potential term and (1/r) d/dr of it
				if (i == 0 && j == 0 && k == 0) {
					const double z = a * sqrt(r2), z2 = z * z;
					if (z < 0.1) {
						// erf(z)/r and its (1/r) d/dr as series in z^2
						const double c = two_over_sqrtpi * a;
						val = c * (1.0 + z2 * (-1.0 / 3.0 + z2 * (1.0 / 10.0 + z2 * (-1.0 / 42.0 + z2 * (1.0 / 216.0 + z2 * (-1.0 / 1320.0 + z2 / 9360.0))))));
						dr = c * a * a *
						     (-2.0 / 3.0 + z2 * (4.0 / 10.0 + z2 * (-6.0 / 42.0 + z2 * (8.0 / 216.0 + z2 * (-10.0 / 1320.0 + z2 * (12.0 / 9360.0 - z2 * 14.0 / 75600.0))))));
					} else {
						const double r = sqrt(r2);
						const double e = erf(z);
						val = e / r;
						dr = (-e / r + two_over_sqrtpi * a * exp(-z2)) / r2;
					}
				} else {
					if (r2 >= 1.0) continue;
					const double r = sqrt(r2), z = a * r;
					const double ec = erfc(z);
					val = -ec / r;
					dr = (ec / r + two_over_sqrtpi * a * exp(-z * z)) / r2;
				}
				u += val;
				du[0] += y[0] * dr;
				du[1] += y[1] * dr;
				du[2] += y[2] * dr;
			}
}

// the real-space correction's gradient and Hessian for one pair at nearest-image offset x:
// du = grad u, hu = the 6 distinct entries of grad grad u (xx, xy, xz, yy, yz, zz); for a radial
// psi with d_k = (r^-1 d/dr)^k psi, grad psi = d1 y, grad grad psi = d2 y y^T + d1 I
//   n = 0:  psi = erf(a r)/r,   d_k = (-1)^k (2k-1)!! r^-(2k+1) [erf - g S_k]   (series for a r < 0.1)
//   n != 0: psi = -erfc(a r)/r, d_k = -(-1)^k (2k-1)!! r^-(2k+1) [erfc + g S_k]
// with g = (2a/sqrt(pi)) exp(-a^2 r^2), S_1 = r, S_2 = r + 2 a^2 r^3 / 3
__device__ void real_space2(const double* x, double* du, double* hu) {
	const double a = kAlpha, c = 1.1283791670955126 * kAlpha;
	for (int q = 0; q < 3; q++) du[q] = 0.0;
	for (int q = 0; q < 6; q++) hu[q] = 0.0;
	for (int i = -1; i <= 1; i++)
		for (int j = -1; j <= 1; j++)
			for (int k = -1; k <= 1; k++) {
				const double y[3] = {x[0] + i, x[1] + j, x[2] + k};
				const double r2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
				double d1, d2;
				if (i == 0 && j == 0 && k == 0) {
					const double z2 = a * a * r2;
					if (z2 < 0.01) {
						// erf(z)/r = c sum_k (-1)^k z^2k / (k! (2k+1)): differentiate term by term in u = r^2
						double s1 = 0.0, s2 = 0.0, zp = 1.0, zq = 1.0, fact = 1.0;  // zp = z^2(q-1), zq = z^2(q-2)
						for (int q = 1; q <= 14; q++) {
							fact *= q;
							const double sg = (q & 1) ? -1.0 : 1.0;
							s1 += sg * 2.0 * q / (fact * (2 * q + 1)) * zp;
							if (q >= 2) {
								s2 += sg * 4.0 * q * (q - 1) / (fact * (2 * q + 1)) * zq;
								zq *= z2;
							}
							zp *= z2;
						}
						d1 = c * a * a * s1;
						d2 = c * a * a * a * a * s2;
					} else {
						const double r = sqrt(r2), e = erf(a * r), g = c * exp(-z2);
						d1 = -(e - g * r) / (r2 * r);
						d2 = 3.0 * (e - g * (r + 2.0 * a * a * r2 * r / 3.0)) / (r2 * r2 * r);
					}
				} else {
					if (r2 >= 1.0) continue;
					const double r = sqrt(r2), ec = erfc(a * r), g = c * exp(-a * a * r2);
					d1 = (ec + g * r) / (r2 * r);
					d2 = -3.0 * (ec + g * (r + 2.0 * a * a * r2 * r / 3.0)) / (r2 * r2 * r);
				}
				du[0] += y[0] * d1;
				du[1] += y[1] * d1;
				du[2] += y[2] * d1;
				hu[0] += d2 * y[0] * y[0] + d1;
				hu[1] += d2 * y[0] * y[1];
				hu[2] += d2 * y[0] * y[2];
				hu[3] += d2 * y[1] * y[1] + d1;
				hu[4] += d2 * y[1] * y[2];
				hu[5] += d2 * y[2] * y[2] + d1;
			}
}

// spline pair in FP64: f (F = -d f) and fr = f'(r)/r
__device__ __forceinline__ void spline64_grad(double r, double h, double& f, double& fr) {
	if (r >= h) {
		f = 1.0 / (r * r * r);
		fr = -3.0 * f / (r * r);
		return;
	}
	const double q = r / h, q2 = q * q, q3 = q2 * q;
	if (r < 0.5 * h) {
		f = (32.0 / 3.0 - 192.0 / 5.0 * q2 + 32.0 * q3) / (h * h * h);
		fr = (-384.0 / 5.0 + 96.0 * q) / (h * h * h * h * h);
	} else {
		const double P = 64.0 / 3.0 * q3 - 48.0 * q3 * q + 192.0 / 5.0 * q3 * q2 - 32.0 / 3.0 * q3 * q3 - 1.0 / 15.0;
		const double dP = 64.0 * q2 - 192.0 * q3 + 192.0 * q3 * q - 64.0 * q3 * q2;
		f = P / (r * r * r);
		fr = (dP / h - 3.0 * P / r) / (r * r * r * r);
	}
}

// structure factors of one particle chunk: thread = half-space vector h, block = 64 vectors
__global__ __launch_bounds__(64) void k_structure(const int4* __restrict__ hv, int nh, const int* __restrict__ x,
		const float* __restrict__ mass, double m, int64_t n, double* __restrict__ part) {
	const int t = blockIdx.x * 64 + threadIdx.x;
	const int64_t per = (n + kSFChunks - 1) / kSFChunks;
	const int64_t b = int64_t(blockIdx.y) * per, e = min(n, b + per);
	__shared__ double sx[256][4];
	const int4 h = t < nh ? hv[t] : make_int4(0, 0, 0, 0);
	double C = 0.0, S = 0.0;
	for (int64_t base = b; base < e; base += 256) {
		__syncthreads();
		for (int q = threadIdx.x; q < 256; q += 64) {
			const int64_t j = base + q;
			if (j < e) {
				sx[q][0] = double(x[3 * j]) * kPosInv + 0.5;
				sx[q][1] = double(x[3 * j + 1]) * kPosInv + 0.5;
				sx[q][2] = double(x[3 * j + 2]) * kPosInv + 0.5;
				sx[q][3] = mass ? double(mass[j]) : m;
			} else {
				sx[q][0] = sx[q][1] = sx[q][2] = sx[q][3] = 0.0;
			}
		}
		__syncthreads();
		const int cnt = int(min<int64_t>(256, e - base));
		for (int q = 0; q < cnt; q++) {
			// 2 h.x is exact in FP64 (x carries 32 fraction bits), so sincospi reduces exactly
			const double arg = 2.0 * (h.x * sx[q][0] + h.y * sx[q][1] + h.z * sx[q][2]);
			double s, c;
			sincospi(arg, &s, &c);
			C = fma(sx[q][3], c, C);
			S = fma(sx[q][3], s, S);
		}
	}
	if (t < nh) {
		part[2 * (int64_t(blockIdx.y) * nh + t)] = C;
		part[2 * (int64_t(blockIdx.y) * nh + t) + 1] = S;
	}
}

__global__ __launch_bounds__(256) void k_structure_sum(const double* __restrict__ part, int nh, double* __restrict__ cs) {
	const int t = blockIdx.x * 256 + threadIdx.x;
	if (t >= nh) return;
	double C = 0.0, S = 0.0;
	for (int c = 0; c < kSFChunks; c++) {
		C += part[2 * (int64_t(c) * nh + t)];
		S += part[2 * (int64_t(c) * nh + t) + 1];
	}
	cs[2 * t] = C;
	cs[2 * t + 1] = S;
}

// pair pass: thread = sink, block = 64 sinks x one source chunk
__global__ __launch_bounds__(64) void k_pairs(const int* __restrict__ x, const float* __restrict__ mass, double m, int64_t n,
		const int64_t* __restrict__ sinks, int64_t ns, int ewald, double h, double* __restrict__ part) {
	const int64_t si = int64_t(blockIdx.x) * 64 + threadIdx.x;
	const int64_t i = sinks[min(si, ns - 1)];
	const int xi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
	const int64_t per = (n + kSrcChunks - 1) / kSrcChunks;
	const int64_t b = int64_t(blockIdx.y) * per, e = min(n, b + per);
	__shared__ int sx[256][3];
	__shared__ double sm[256];
	double phi = 0.0, g[3] = {0.0, 0.0, 0.0};
	for (int64_t base = b; base < e; base += 256) {
		__syncthreads();
		for (int q = threadIdx.x; q < 256; q += 64) {
			const int64_t j = base + q;
			const bool in = j < e;
			sx[q][0] = in ? x[3 * j] : 0;
			sx[q][1] = in ? x[3 * j + 1] : 0;
			sx[q][2] = in ? x[3 * j + 2] : 0;
			sm[q] = in ? (mass ? double(mass[j]) : m) : 0.0;
		}
		__syncthreads();
		const int cnt = int(min<int64_t>(256, e - base));
		for (int q = 0; q < cnt; q++) {
			const double dx[3] = {sep(ewald, xi[0], sx[q][0]), sep(ewald, xi[1], sx[q][1]), sep(ewald, xi[2], sx[q][2])};
			const double mj = sm[q];
			const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
			double f, p;
			spline64(r, h, f, p);
			phi -= mj * p;
			g[0] -= mj * dx[0] * f;
			g[1] -= mj * dx[1] * f;
			g[2] -= mj * dx[2] * f;
			if (ewald) {
				double u, du[3];
				real_space(dx, u, du);
				phi += mj * u;
				g[0] -= mj * du[0];
				g[1] -= mj * du[1];
				g[2] -= mj * du[2];
			}
		}
	}
	if (si < ns) {
		double* o = part + 4 * (si * kSrcChunks + blockIdx.y);
		o[0] = phi;
		o[1] = g[0];
		o[2] = g[1];
		o[3] = g[2];
	}
}

// per sink: pair partials in chunk order, then the Fourier part and the constant
__global__ __launch_bounds__(64) void k_finish(const int* __restrict__ x, const int64_t* __restrict__ sinks, int64_t ns,
		const double* __restrict__ part, int ewald, const int4* __restrict__ hv, const double* __restrict__ cs, int nh,
		double mtot, double* __restrict__ out) {
	const int64_t si = int64_t(blockIdx.x) * 64 + threadIdx.x;
	if (si >= ns) return;
	double phi = 0.0, g[3] = {0.0, 0.0, 0.0};
	for (int c = 0; c < kSrcChunks; c++) {
		const double* o = part + 4 * (si * kSrcChunks + c);
		phi += o[0];
		g[0] += o[1];
		g[1] += o[2];
		g[2] += o[3];
	}
	if (ewald) {
		const int64_t i = sinks[si];
		const double xi[3] = {double(x[3 * i]) * kPosInv + 0.5, double(x[3 * i + 1]) * kPosInv + 0.5,
				double(x[3 * i + 2]) * kPosInv + 0.5};
		const double pi = 3.141592653589793, a2 = kAlpha * kAlpha;
		double fp = 0.0, fg[3] = {0.0, 0.0, 0.0};
		for (int t = 0; t < nh; t++) {
			const int4 h = hv[t];
			const double h2 = double(h.x * h.x + h.y * h.y + h.z * h.z);
			const double c0 = 2.0 * exp(-pi * pi * h2 / a2) / h2;  // both h and -h
			double s, c;
			sincospi(2.0 * (h.x * xi[0] + h.y * xi[1] + h.z * xi[2]), &s, &c);
			const double C = cs[2 * t], S = cs[2 * t + 1];
			fp += c0 * (c * C + s * S);
			const double gg = 2.0 * c0 * (s * C - c * S);
			fg[0] += h.x * gg;
			fg[1] += h.y * gg;
			fg[2] += h.z * gg;
		}
		phi += -fp / pi + mtot * pi / a2;
		g[0] -= fg[0];
		g[1] -= fg[1];
		g[2] -= fg[2];
	}
	double* o = out + 4 * si;
	o[0] = phi;
	o[1] = g[0];
	o[2] = g[1];
	o[3] = g[2];
}

// gradient: structure factors C, S (mass weighted) and Cw, Sw (weight weighted), 8 per h
__global__ __launch_bounds__(64) void k_structure8(const int4* __restrict__ hv, int nh, const int* __restrict__ x,
		const float* __restrict__ mass, double m, const float* __restrict__ w, int64_t n, double* __restrict__ part) {
	const int t = blockIdx.x * 64 + threadIdx.x;
	const int64_t per = (n + kSFChunks - 1) / kSFChunks;
	const int64_t b = int64_t(blockIdx.y) * per, e = min(n, b + per);
	__shared__ double sx[128][7];
	const int4 h = t < nh ? hv[t] : make_int4(0, 0, 0, 0);
	double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	for (int64_t base = b; base < e; base += 128) {
		__syncthreads();
		for (int q = threadIdx.x; q < 128; q += 64) {
			const int64_t j = base + q;
			const bool in = j < e;
			sx[q][0] = in ? double(x[3 * j]) * kPosInv + 0.5 : 0.0;
			sx[q][1] = in ? double(x[3 * j + 1]) * kPosInv + 0.5 : 0.0;
			sx[q][2] = in ? double(x[3 * j + 2]) * kPosInv + 0.5 : 0.0;
			sx[q][3] = in ? (mass ? double(mass[j]) : m) : 0.0;
			sx[q][4] = in ? double(w[3 * j]) : 0.0;
			sx[q][5] = in ? double(w[3 * j + 1]) : 0.0;
			sx[q][6] = in ? double(w[3 * j + 2]) : 0.0;
		}
		__syncthreads();
		const int cnt = int(min<int64_t>(128, e - base));
		for (int q = 0; q < cnt; q++) {
			double sn, cs;
			sincospi(2.0 * (h.x * sx[q][0] + h.y * sx[q][1] + h.z * sx[q][2]), &sn, &cs);
			acc[0] = fma(sx[q][3], cs, acc[0]);
			acc[1] = fma(sx[q][3], sn, acc[1]);
			for (int d = 0; d < 3; d++) {
				acc[2 + d] = fma(sx[q][4 + d], cs, acc[2 + d]);
				acc[5 + d] = fma(sx[q][4 + d], sn, acc[5 + d]);
			}
		}
	}
	if (t < nh)
		for (int q = 0; q < 8; q++) part[8 * (int64_t(blockIdx.y) * nh + t) + q] = acc[q];
}

__global__ __launch_bounds__(256) void k_structure8_sum(const double* __restrict__ part, int nh, double* __restrict__ cs) {
	const int t = blockIdx.x * 256 + threadIdx.x;
	if (t >= nh) return;
	double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	for (int c = 0; c < kSFChunks; c++)
		for (int q = 0; q < 8; q++) acc[q] += part[8 * (int64_t(c) * nh + t) + q];
	for (int q = 0; q < 8; q++) cs[8 * t + q] = acc[q];
}

// gradient pair pass: d/dx_k += J(d_kj) (m_j w_k - m_k w_j),  d/dm_k += -w_j . F(d_kj)
__global__ __launch_bounds__(64) void k_pairs_grad(const int* __restrict__ x, const float* __restrict__ mass, double m,
		const float* __restrict__ w, int64_t n, const int64_t* __restrict__ sinks, int64_t ns, int ewald, double h,
		double* __restrict__ part) {
	const int64_t si = int64_t(blockIdx.x) * 64 + threadIdx.x;
	const int64_t i = sinks[min(si, ns - 1)];
	const int xi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
	const double mi = mass ? double(mass[i]) : m;
	const double wi[3] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
	const int64_t per = (n + kSrcChunks - 1) / kSrcChunks;
	const int64_t b = int64_t(blockIdx.y) * per, e = min(n, b + per);
	__shared__ int sx[256][3];
	__shared__ double sm[256], sw[256][3];
	double g[3] = {0.0, 0.0, 0.0}, gm = 0.0;
	for (int64_t base = b; base < e; base += 256) {
		__syncthreads();
		for (int q = threadIdx.x; q < 256; q += 64) {
			const int64_t j = base + q;
			const bool in = j < e;
			sx[q][0] = in ? x[3 * j] : 0;
			sx[q][1] = in ? x[3 * j + 1] : 0;
			sx[q][2] = in ? x[3 * j + 2] : 0;
			sm[q] = in ? (mass ? double(mass[j]) : m) : 0.0;
			for (int d = 0; d < 3; d++) sw[q][d] = in ? double(w[3 * j + d]) : 0.0;
		}
		__syncthreads();
		const int cnt = int(min<int64_t>(256, e - base));
		for (int q = 0; q < cnt; q++) {
			const double dx[3] = {sep(ewald, xi[0], sx[q][0]), sep(ewald, xi[1], sx[q][1]), sep(ewald, xi[2], sx[q][2])};
			const double mj = sm[q];
			const double v[3] = {mj * wi[0] - mi * sw[q][0], mj * wi[1] - mi * sw[q][1], mj * wi[2] - mi * sw[q][2]};
			const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
			double f, fr;
			spline64_grad(r, h, f, fr);
			const double dv = dx[0] * v[0] + dx[1] * v[1] + dx[2] * v[2];
			const double dw = dx[0] * sw[q][0] + dx[1] * sw[q][1] + dx[2] * sw[q][2];
			for (int d = 0; d < 3; d++) g[d] += -f * v[d] - fr * dv * dx[d];
			gm += f * dw;
			if (ewald) {
				// correction: F = -grad u, J = -grad grad u
				double du[3], hu[6];
				real_space2(dx, du, hu);
				g[0] -= hu[0] * v[0] + hu[1] * v[1] + hu[2] * v[2];
				g[1] -= hu[1] * v[0] + hu[3] * v[1] + hu[4] * v[2];
				g[2] -= hu[2] * v[0] + hu[4] * v[1] + hu[5] * v[2];
				gm += du[0] * sw[q][0] + du[1] * sw[q][1] + du[2] * sw[q][2];
			}
		}
	}
	if (si < ns) {
		double* o = part + 4 * (si * kSrcChunks + blockIdx.y);
		o[0] = g[0];
		o[1] = g[1];
		o[2] = g[2];
		o[3] = gm;
	}
}

// per sink: pair partials, then the Fourier part: with c0 = 2 exp(-pi^2 h^2/a^2)/h^2 per half-space h
//   d/dx_k += -4 pi c0 h [(h.w_k)(cos_k C + sin_k S) - m_k (cos_k h.Cw + sin_k h.Sw)]
//   d/dm_k += 2 c0 [sin_k h.Cw - cos_k h.Sw]
__global__ __launch_bounds__(64) void k_finish_grad(const int* __restrict__ x, const float* __restrict__ mass, double m,
		const float* __restrict__ w, const int64_t* __restrict__ sinks, int64_t ns, const double* __restrict__ part, int ewald,
		const int4* __restrict__ hv, const double* __restrict__ cs8, int nh, double* __restrict__ out) {
	const int64_t si = int64_t(blockIdx.x) * 64 + threadIdx.x;
	if (si >= ns) return;
	double g[4] = {0.0, 0.0, 0.0, 0.0};
	for (int c = 0; c < kSrcChunks; c++)
		for (int q = 0; q < 4; q++) g[q] += part[4 * (si * kSrcChunks + c) + q];
	if (ewald) {
		const int64_t i = sinks[si];
		const double xi[3] = {double(x[3 * i]) * kPosInv + 0.5, double(x[3 * i + 1]) * kPosInv + 0.5,
				double(x[3 * i + 2]) * kPosInv + 0.5};
		const double mi = mass ? double(mass[i]) : m;
		const double wi[3] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
		const double pi = 3.141592653589793, a2 = kAlpha * kAlpha;
		for (int t = 0; t < nh; t++) {
			const int4 h = hv[t];
			const double hd[3] = {double(h.x), double(h.y), double(h.z)};
			const double h2 = hd[0] * hd[0] + hd[1] * hd[1] + hd[2] * hd[2];
			const double c0 = 2.0 * exp(-pi * pi * h2 / a2) / h2;
			double sn, cs;
			sincospi(2.0 * (hd[0] * xi[0] + hd[1] * xi[1] + hd[2] * xi[2]), &sn, &cs);
			const double* S8 = cs8 + 8 * t;
			const double hw = hd[0] * wi[0] + hd[1] * wi[1] + hd[2] * wi[2];
			const double hCw = hd[0] * S8[2] + hd[1] * S8[3] + hd[2] * S8[4];
			const double hSw = hd[0] * S8[5] + hd[1] * S8[6] + hd[2] * S8[7];
			// F_F(d) = -2 sum_half c0 h sin(2 pi h.d), J_F(d) = -4 pi sum_half c0 h h^T cos(2 pi h.d)
			const double coef = -4.0 * pi * c0 * (hw * (cs * S8[0] + sn * S8[1]) - mi * (cs * hCw + sn * hSw));
			for (int d = 0; d < 3; d++) g[d] += coef * hd[d];
			g[3] += 2.0 * c0 * (sn * hCw - cs * hSw);
		}
	}
	double* o = out + 4 * si;
	for (int q = 0; q < 4; q++) o[q] = g[q];
}

struct Buf {
	void* p = nullptr;
	~Buf() {
		if (p) (void) hipFree(p);
	}
};

}  // namespace

extern "C" int tg_direct_forces(const tg_params* prm, const int32_t* x, const float* mass, int64_t n, const int64_t* sinks,
		int64_t ns, tg_force* out) {
	if (!prm || !x || !sinks || !out || n <= 0 || ns <= 0 || !(prm->soft_len > 0.0)) return TG_EINVAL;
	for (int64_t k = 0; k < ns; k++)
		if (sinks[k] < 0 || sinks[k] >= n) {
			g_direct_err = "sink index out of range";
			return TG_EINVAL;
		}
	const int ewald = prm->ewald ? 1 : 0;
	double mtot = 0.0;
	if (mass)
		for (int64_t j = 0; j < n; j++) mtot += mass[j];
	else
		mtot = prm->m * double(n);
	// half-space of 0 < |h|^2 <= kHMax2
	std::vector<int4> hv;
	for (int i = -12; i <= 12; i++)
		for (int j = -12; j <= 12; j++)
			for (int k = -12; k <= 12; k++) {
				const int h2 = i * i + j * j + k * k;
				if (h2 == 0 || h2 > kHMax2) continue;
				if (i < 0 || (i == 0 && (j < 0 || (j == 0 && k < 0)))) continue;
				hv.push_back(make_int4(i, j, k, 0));
			}
	const int nh = int(hv.size());
	Buf dx, dm, ds, dh, dsf, dcs, dpart, dout;
	hipStream_t st = nullptr;
	int rc = TG_OK;
#define DCHK(call)                                                   \
	do {                                                             \
		hipError_t e_ = (call);                                      \
		if (e_ != hipSuccess) {                                      \
			g_direct_err = std::string(#call) + ": " + hipGetErrorString(e_); \
			rc = e_ == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP;    \
			goto done;                                               \
		}                                                            \
	} while (0)
	{
		DCHK(hipSetDevice(prm->device));
		DCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
		DCHK(hipMalloc(&dx.p, size_t(n) * 12));
		DCHK(hipMemcpyAsync(dx.p, x, size_t(n) * 12, hipMemcpyHostToDevice, st));
		if (mass) {
			DCHK(hipMalloc(&dm.p, size_t(n) * 4));
			DCHK(hipMemcpyAsync(dm.p, mass, size_t(n) * 4, hipMemcpyHostToDevice, st));
		}
		DCHK(hipMalloc(&ds.p, size_t(ns) * 8));
		DCHK(hipMemcpyAsync(ds.p, sinks, size_t(ns) * 8, hipMemcpyHostToDevice, st));
		DCHK(hipMalloc(&dh.p, hv.size() * sizeof(int4)));
		DCHK(hipMemcpyAsync(dh.p, hv.data(), hv.size() * sizeof(int4), hipMemcpyHostToDevice, st));
		DCHK(hipMalloc(&dcs.p, size_t(nh) * 16));
		if (ewald) {
			DCHK(hipMalloc(&dsf.p, size_t(kSFChunks) * nh * 16));
			hipLaunchKernelGGL(k_structure, dim3((nh + 63) / 64, kSFChunks), dim3(64), 0, st, static_cast<const int4*>(dh.p), nh,
					static_cast<const int*>(dx.p), static_cast<const float*>(dm.p), prm->m, n, static_cast<double*>(dsf.p));
			DCHK(hipGetLastError());
			hipLaunchKernelGGL(k_structure_sum, dim3((nh + 255) / 256), dim3(256), 0, st, static_cast<const double*>(dsf.p), nh,
					static_cast<double*>(dcs.p));
			DCHK(hipGetLastError());
		}
		DCHK(hipMalloc(&dpart.p, size_t(ns) * kSrcChunks * 32));
		const unsigned sblocks = unsigned((ns + 63) / 64);
		hipLaunchKernelGGL(k_pairs, dim3(sblocks, kSrcChunks), dim3(64), 0, st, static_cast<const int*>(dx.p),
				static_cast<const float*>(dm.p), prm->m, n, static_cast<const int64_t*>(ds.p), ns, ewald, prm->soft_len,
				static_cast<double*>(dpart.p));
		DCHK(hipGetLastError());
		DCHK(hipMalloc(&dout.p, size_t(ns) * 32));
		hipLaunchKernelGGL(k_finish, dim3(sblocks), dim3(64), 0, st, static_cast<const int*>(dx.p),
				static_cast<const int64_t*>(ds.p), ns, static_cast<const double*>(dpart.p), ewald, static_cast<const int4*>(dh.p),
				static_cast<const double*>(dcs.p), nh, mtot, static_cast<double*>(dout.p));
		DCHK(hipGetLastError());
		DCHK(hipMemcpyAsync(out, dout.p, size_t(ns) * 32, hipMemcpyDeviceToHost, st));
		DCHK(hipStreamSynchronize(st));
	}
done:
#undef DCHK
	if (st) (void) hipStreamDestroy(st);
	return rc;
}

extern "C" int tg_direct_gradient(const tg_params* prm, const int32_t* x, const float* mass, int64_t n, const float* w,
		const int64_t* sinks, int64_t ns, double* out) {
	if (!prm || !x || !w || !sinks || !out || n <= 0 || ns <= 0 || !(prm->soft_len > 0.0)) return TG_EINVAL;
	for (int64_t k = 0; k < ns; k++)
		if (sinks[k] < 0 || sinks[k] >= n) {
			g_direct_err = "sink index out of range";
			return TG_EINVAL;
		}
	const int ewald = prm->ewald ? 1 : 0;
	std::vector<int4> hv;
	for (int i = -12; i <= 12; i++)
		for (int j = -12; j <= 12; j++)
			for (int k = -12; k <= 12; k++) {
				const int h2 = i * i + j * j + k * k;
				if (h2 == 0 || h2 > kHMax2) continue;
				if (i < 0 || (i == 0 && (j < 0 || (j == 0 && k < 0)))) continue;
				hv.push_back(make_int4(i, j, k, 0));
			}
	const int nh = int(hv.size());
	Buf dx, dm, dw, ds, dh, dsf, dcs, dpart, dout;
	hipStream_t st = nullptr;
	int rc = TG_OK;
#define DCHK(call)                                                   \
	do {                                                             \
		hipError_t e_ = (call);                                      \
		if (e_ != hipSuccess) {                                      \
			g_direct_err = std::string(#call) + ": " + hipGetErrorString(e_); \
			rc = e_ == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP;    \
			goto done;                                               \
		}                                                            \
	} while (0)
	{
		DCHK(hipSetDevice(prm->device));
		DCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
		DCHK(hipMalloc(&dx.p, size_t(n) * 12));
		DCHK(hipMemcpyAsync(dx.p, x, size_t(n) * 12, hipMemcpyHostToDevice, st));
		DCHK(hipMalloc(&dw.p, size_t(n) * 12));
		DCHK(hipMemcpyAsync(dw.p, w, size_t(n) * 12, hipMemcpyHostToDevice, st));
		if (mass) {
			DCHK(hipMalloc(&dm.p, size_t(n) * 4));
			DCHK(hipMemcpyAsync(dm.p, mass, size_t(n) * 4, hipMemcpyHostToDevice, st));
		}
		DCHK(hipMalloc(&ds.p, size_t(ns) * 8));
		DCHK(hipMemcpyAsync(ds.p, sinks, size_t(ns) * 8, hipMemcpyHostToDevice, st));
		DCHK(hipMalloc(&dh.p, hv.size() * sizeof(int4)));
		DCHK(hipMemcpyAsync(dh.p, hv.data(), hv.size() * sizeof(int4), hipMemcpyHostToDevice, st));
		DCHK(hipMalloc(&dcs.p, size_t(nh) * 64));
		if (ewald) {
			DCHK(hipMalloc(&dsf.p, size_t(kSFChunks) * nh * 64));
			hipLaunchKernelGGL(k_structure8, dim3((nh + 63) / 64, kSFChunks), dim3(64), 0, st, static_cast<const int4*>(dh.p), nh,
					static_cast<const int*>(dx.p), static_cast<const float*>(dm.p), prm->m, static_cast<const float*>(dw.p), n,
					static_cast<double*>(dsf.p));
			DCHK(hipGetLastError());
			hipLaunchKernelGGL(k_structure8_sum, dim3((nh + 255) / 256), dim3(256), 0, st, static_cast<const double*>(dsf.p), nh,
					static_cast<double*>(dcs.p));
			DCHK(hipGetLastError());
		}
		DCHK(hipMalloc(&dpart.p, size_t(ns) * kSrcChunks * 32));
		const unsigned sblocks = unsigned((ns + 63) / 64);
		hipLaunchKernelGGL(k_pairs_grad, dim3(sblocks, kSrcChunks), dim3(64), 0, st, static_cast<const int*>(dx.p),
				static_cast<const float*>(dm.p), prm->m, static_cast<const float*>(dw.p), n, static_cast<const int64_t*>(ds.p), ns, ewald,
				prm->soft_len, static_cast<double*>(dpart.p));
		DCHK(hipGetLastError());
		DCHK(hipMalloc(&dout.p, size_t(ns) * 32));
		hipLaunchKernelGGL(k_finish_grad, dim3(sblocks), dim3(64), 0, st, static_cast<const int*>(dx.p),
				static_cast<const float*>(dm.p), prm->m, static_cast<const float*>(dw.p), static_cast<const int64_t*>(ds.p), ns,
				static_cast<const double*>(dpart.p), ewald, static_cast<const int4*>(dh.p), static_cast<const double*>(dcs.p), nh,
				static_cast<double*>(dout.p));
		DCHK(hipGetLastError());
		DCHK(hipMemcpyAsync(out, dout.p, size_t(ns) * 32, hipMemcpyDeviceToHost, st));
		DCHK(hipStreamSynchronize(st));
	}
done:
#undef DCHK
	if (st) (void) hipStreamDestroy(st);
	return rc;
}

extern "C" const char* tg_direct_last_error(void) { return g_direct_err.c_str(); }
