// tg_step.hip -- kick and drift with hierarchical rungs on device-resident particles
// (include/tg_step.h): part_vect_kick (src/part_vect.cpp:341-436), part_vect_drift
// (src/part_vect.cpp:493-540) and the time grid of tigergrav/time.hpp.
//
// One thread per particle.  Every per-particle expression keeps the reference's types and operand
// order -- float velocities, FP64 forces, the float conversions its vect<float> operators make --
// with no contraction, so velocities, rungs and positions come out as the reference's.  The
// statistics are summed per work group (tree reduction in LDS) and then over the groups in a fixed
// order by one more work group: deterministic, though not the reference's particle order.
#include "tg_step.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

namespace tg {
void set_capi_error(const char* where, hipError_t e);  // tg_capi.cpp
void set_capi_message(const char* msg);
}  // namespace tg

namespace {

constexpr int BLOCK = 256;
constexpr int NSTAT = 8;  // g[3], p[3], pot, kin

__device__ __forceinline__ double rung_dt(double t_max, int rung) {
	return t_max / double(1 << rung);  // rung_to_dt (tigergrav/time.hpp:36-39)
}

__device__ __forceinline__ double d_pos_to_double(int32_t x) {
#pragma clang fp contract(off)
	return double(x) * (1.0 / 4294967296.0) + 0.5;
}

__device__ __forceinline__ int32_t d_double_to_pos(double x) {
#pragma clang fp contract(off)
	const double POS_MAX = 4294967296.0;
	const double y = x - 0.5;
	long long z = llround(y * POS_MAX);
	if (double(z) >= POS_MAX / 2.0) z -= (long long) POS_MAX;
	return int32_t(z);
}

// float dot product in the reference's order (tigergrav/vect.hpp:222-230)
__device__ __forceinline__ float dotf(const float* v) {
#pragma clang fp contract(off)
	float r = v[0] * v[0];
	r += v[1] * v[1];
	r += v[2] * v[2];
	return r;
}

// sums of NSTAT doubles over the work group (values of lanes past the end are zero); lane 0 writes
// the group's partial
__device__ void group_sum(double* s, double* out, int* rmax_lds, int rung_max, int* rung_out) {
	__shared__ double red[NSTAT][BLOCK];
	const int t = threadIdx.x;
	for (int k = 0; k < NSTAT; k++) red[k][t] = s[k];
	if (rung_out) atomicMax(rmax_lds, rung_max);
	__syncthreads();
	for (int w = BLOCK / 2; w > 0; w >>= 1) {
		if (t < w)
			for (int k = 0; k < NSTAT; k++) red[k][t] += red[k][t + w];
		__syncthreads();
	}
	if (t == 0) {
		for (int k = 0; k < NSTAT; k++) out[NSTAT * blockIdx.x + k] = red[k][0];
		if (rung_out) rung_out[blockIdx.x] = *rmax_lds;
	}
}

struct KickArgs {
	double G, eta, soft_len, dt_max, t_max, m, N, scale, sgn;
	int glass, cosmic, ewald, min_rung, do_out;
	double dt1[TG_RUNG_MAX + 1], dt2[TG_RUNG_MAX + 1];
};

__global__ __launch_bounds__(BLOCK) void k_kick(KickArgs A, int64_t n, const tg_force* __restrict__ f, float* __restrict__ v,
		int8_t* __restrict__ rung, double* __restrict__ partial, int* __restrict__ rung_part, int* __restrict__ bad) {
#pragma clang fp contract(off)
	__shared__ int rmax;
	if (threadIdx.x == 0) rmax = 0;
	__syncthreads();
	const int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
	double s[NSTAT] = {0, 0, 0, 0, 0, 0, 0, 0};
	int my_rung = 0;
	if (i < n) {
		const double eps = 10.0 * DBL_MIN;
		const double a3inv = 1.0 / (A.scale * A.scale * A.scale);
		const double glass_drag = 1.0;
		float vv[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
		// statistics of every particle: momentum and kinetic energy before the kick
		const float Nf = float(A.N), s2f = float(A.scale * A.scale);
		for (int d = 0; d < 3; d++) s[3 + d] = double((vv[d] / Nf) / s2f);
		s[7] = 0.5 * A.m * double(dotf(vv)) / (A.scale * A.scale);
		int r = rung[i];
		if (r < 0 || r > TG_RUNG_MAX) {
			atomicOr(bad, 2);  // not a rung of the time grid: left as it is
		} else if (r >= A.min_rung || A.do_out) {
			const tg_force F = f[i];
			if (r >= A.min_rung) {
				if (r != 0) {
					const double dt = rung_dt(A.t_max, r);
					const double dt1 = A.cosmic ? A.dt1[r] : 0.5 * dt;
					const double c = dt1 * A.sgn * A.G;
					for (int d = 0; d < 3; d++) vv[d] = vv[d] + float(F.g[d] * c);
					if (A.glass) {
						const float q = float(1.0 + glass_drag * dt1);
						for (int d = 0; d < 3; d++) vv[d] = vv[d] / q;
					}
				}
				double gG[3];
				for (int d = 0; d < 3; d++) gG[d] = F.g[d] * A.G;
				double gg = gG[0] * gG[0];
				gg += gG[1] * gG[1];
				gg += gG[2] * gG[2];
				const double a = sqrt(gg) * a3inv;
				const double x = A.eta * sqrt(A.soft_len * (1.0 / 2.8) / (a + eps));
				double dt = x < A.dt_max ? x : A.dt_max;
				// dt_to_rung (tigergrav/time.hpp:41-48): the first rung whose step is not above dt
				int nr = 0;
				while (nr <= TG_RUNG_MAX && rung_dt(A.t_max, nr) > dt) nr++;
				if (nr > TG_RUNG_MAX) {
					atomicOr(bad, 1);  // beyond the time grid (the reference's shift overflows there)
					nr = TG_RUNG_MAX;
				}
				nr = nr > A.min_rung ? nr : A.min_rung;
				my_rung = nr;
				dt = rung_dt(A.t_max, nr);
				int keep = r - 1;
				int stored = nr > keep ? nr : keep;
				stored = stored > 1 ? stored : 1;
				rung[i] = int8_t(stored);
				const double dt2 = A.ewald ? A.dt2[stored] : 0.5 * dt;
				const double c2 = dt2 * A.sgn * A.G;
				for (int d = 0; d < 3; d++) vv[d] = vv[d] + float(F.g[d] * c2);
				if (A.glass) {
					const float q = float(1.0 + glass_drag * dt2);
					for (int d = 0; d < 3; d++) vv[d] = vv[d] / q;
				}
				for (int d = 0; d < 3; d++) v[3 * i + d] = vv[d];
			}
			if (A.do_out) {
				for (int d = 0; d < 3; d++) s[d] = F.g[d] * A.G / A.N / (A.scale * A.scale * A.scale);
				s[6] = 0.5 * A.m * F.phi * A.G / A.scale;
			}
		}
	}
	group_sum(s, partial, &rmax, my_rung, rung_part);
}

__global__ __launch_bounds__(BLOCK) void k_drift(int64_t n, int32_t* __restrict__ x, const float* __restrict__ v, double drift_dt,
		double m, double a1inv2, double* __restrict__ partial) {
#pragma clang fp contract(off)
	const int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
	double s[NSTAT] = {0, 0, 0, 0, 0, 0, 0, 0};
	if (i < n) {
		const float vv[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
		s[7] = 0.5 * m * double(dotf(vv)) * a1inv2;
		const float dtf = float(drift_dt);
		for (int d = 0; d < 3; d++) {
			const double dx = double(vv[d] * dtf);
			double p = d_pos_to_double(x[3 * i + d]);
			p += dx;
			while (p >= 1.0) p -= 1.0;
			while (p < 0.0) p += 1.0;
			x[3 * i + d] = d_double_to_pos(p);
		}
	}
	group_sum(s, partial, nullptr, 0, nullptr);
}

// the groups' partials: 256 threads each add a contiguous run of groups in group order, then
// thread 0 adds the 256 run sums in order (a fixed order, so the totals are deterministic)
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, const int* __restrict__ rung_part, int ng,
		double* __restrict__ out) {
	__shared__ double run[NSTAT][256];
	__shared__ int rrun[256];
	const int t = threadIdx.x;
	const int per = (ng + 255) / 256;
	const int g0 = t * per, g1 = g0 + per < ng ? g0 + per : ng;
	double a[NSTAT] = {0, 0, 0, 0, 0, 0, 0, 0};
	int r = 0;
	for (int g = g0; g < g1; g++) {
		for (int k = 0; k < NSTAT; k++) a[k] += partial[NSTAT * g + k];
		if (rung_part) r = rung_part[g] > r ? rung_part[g] : r;
	}
	for (int k = 0; k < NSTAT; k++) run[k][t] = a[k];
	rrun[t] = r;
	__syncthreads();
	if (t < NSTAT) {
		double s = 0.0;
		for (int q = 0; q < 256; q++) s += run[t][q];
		out[t] = s;
	}
	if (t == NSTAT) {
		int m = 0;
		for (int q = 0; q < 256; q++) m = rrun[q] > m ? rrun[q] : m;
		out[NSTAT] = double(m);
	}
}

#define STCHK(call)                                           \
	do {                                                      \
		hipError_t e_ = (call);                               \
		if (e_ != hipSuccess) {                               \
			tg::set_capi_error(#call, e_);                    \
			rc = e_ == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP; \
			goto done;                                        \
		}                                                     \
	} while (0)

// Per-device scratch of the calls (group partials, the rung maxima, the error flag, the totals),
// kept between calls and grown when a larger particle count needs it: a call allocates nothing in
// the steady state of a time step.  Calls are serialised by the mutex.
struct Scratch {
	double* partial = nullptr;
	int* rung_part = nullptr;
	int* bad = nullptr;
	double* sum = nullptr;
	int cap = 0;  // groups
};
std::mutex g_scratch_mtx;
std::vector<Scratch> g_scratch;

hipError_t scratch_for(int device, int ng, Scratch** out) {
	if (size_t(device) >= g_scratch.size()) g_scratch.resize(size_t(device) + 1);
	Scratch& S = g_scratch[size_t(device)];
	if (ng > S.cap || !S.bad) {
		if (S.partial) (void) hipFree(S.partial);
		if (S.rung_part) (void) hipFree(S.rung_part);
		if (S.bad) (void) hipFree(S.bad);
		if (S.sum) (void) hipFree(S.sum);
		S = Scratch();
		const int cap = ng > 1024 ? ng : 1024;
		hipError_t e = hipMalloc(&S.partial, size_t(cap) * NSTAT * sizeof(double));
		if (e == hipSuccess) e = hipMalloc(&S.rung_part, size_t(cap) * sizeof(int));
		if (e == hipSuccess) e = hipMalloc(&S.bad, sizeof(int));
		if (e == hipSuccess) e = hipMalloc(&S.sum, (NSTAT + 1) * sizeof(double));
		if (e != hipSuccess) return e;
		S.cap = cap;
	}
	*out = &S;
	return hipSuccess;
}

__global__ void k_clear_flag(int* f) {
	if (threadIdx.x == 0) *f = 0;
}

bool bad_params(const tg_step_params* p) {
	if (!p) return true;
	int nd = 0;
	if (hipGetDeviceCount(&nd) != hipSuccess || p->device < 0 || p->device >= nd) return true;
	return !(p->t_max > 0.0) || !(p->scale > 0.0);
}

}  // namespace

extern "C" {

int tg_kick(const tg_step_params* p, int64_t n, const tg_force* f, float* v, int8_t* rung, int min_rung, int do_out,
		tg_kick_stats* stats) {
	if (bad_params(p) || n < 0 || (n > 0 && (!f || !v || !rung)) || min_rung < 0 || min_rung > TG_RUNG_MAX) {
		tg::set_capi_message("tg_kick: bad argument");
		return TG_EINVAL;
	}
	int rc = TG_OK;
	const int ng = int((n + BLOCK - 1) / BLOCK) > 0 ? int((n + BLOCK - 1) / BLOCK) : 1;
	double* partial = nullptr;
	int* rung_part = nullptr;
	int* bad = nullptr;
	double* sum = nullptr;
	Scratch* S = nullptr;
	std::lock_guard<std::mutex> lock(g_scratch_mtx);
	double hs[NSTAT + 1] = {};
	int hbad = 0;
	KickArgs A{};
	A.G = p->G;
	A.eta = p->eta;
	A.soft_len = p->soft_len;
	A.dt_max = p->dt_max;
	A.t_max = p->t_max;
	A.m = p->m;
	A.N = p->problem_size;
	A.scale = p->scale;
	A.sgn = p->glass ? -1.0 : 1.0;
	A.glass = p->glass;
	A.cosmic = p->cosmic;
	A.ewald = p->ewald;
	A.min_rung = min_rung;
	A.do_out = do_out;
	for (int r = 0; r <= TG_RUNG_MAX; r++) {
		A.dt1[r] = p->kick_dt1[r];
		A.dt2[r] = p->kick_dt2[r];
	}
	STCHK(hipSetDevice(p->device));
	STCHK(scratch_for(p->device, ng, &S));
	partial = S->partial;
	rung_part = S->rung_part;
	bad = S->bad;
	sum = S->sum;
	hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, 0, bad);
	hipLaunchKernelGGL(k_kick, dim3(ng), dim3(BLOCK), 0, 0, A, n, f, v, rung, partial, rung_part, bad);
	STCHK(hipGetLastError());
	hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, 0, partial, rung_part, ng, sum);
	STCHK(hipGetLastError());
	STCHK(hipMemcpy(hs, sum, sizeof(hs), hipMemcpyDeviceToHost));
	STCHK(hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost));
	if (stats) {
		for (int d = 0; d < 3; d++) {
			stats->g[d] = hs[d];
			stats->p[d] = hs[3 + d];
		}
		stats->pot = hs[6];
		stats->kin = hs[7];
		stats->max_rung = int32_t(hs[8]);
	}
	if (hbad) {
		tg::set_capi_message(hbad & 2 ? "tg_kick: a particle's rung is outside 0..30"
		                              : "tg_kick: a particle's time step falls below the finest rung (30)");
		rc = TG_EINVAL;
	}
done:
	return rc;
}

int tg_drift(const tg_step_params* p, int64_t n, int32_t* x, const float* v, int max_rung, double drift_dt, double a1,
		double* ekin) {
	if (bad_params(p) || n < 0 || (n > 0 && (!x || !v)) || max_rung < 0 || max_rung > TG_RUNG_MAX || !(a1 > 0.0)) {
		tg::set_capi_message("tg_drift: bad argument");
		return TG_EINVAL;
	}
	int rc = TG_OK;
	const int ng = int((n + BLOCK - 1) / BLOCK) > 0 ? int((n + BLOCK - 1) / BLOCK) : 1;
	double* partial = nullptr;
	double* sum = nullptr;
	Scratch* S = nullptr;
	std::lock_guard<std::mutex> lock(g_scratch_mtx);
	double hs[NSTAT + 1] = {};
	// the drift step: the finest active rung's, or drift_dt with cosmology (src/part_vect.cpp:505-506)
	const double dt = p->cosmic ? drift_dt : p->t_max / double(1 << max_rung);
	STCHK(hipSetDevice(p->device));
	STCHK(scratch_for(p->device, ng, &S));
	partial = S->partial;
	sum = S->sum;
	hipLaunchKernelGGL(k_drift, dim3(ng), dim3(BLOCK), 0, 0, n, x, v, dt, p->m, 1.0 / (a1 * a1), partial);
	STCHK(hipGetLastError());
	hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, 0, partial, nullptr, ng, sum);
	STCHK(hipGetLastError());
	STCHK(hipMemcpy(hs, sum, sizeof(hs), hipMemcpyDeviceToHost));
	if (ekin) *ekin = hs[7];
done:
	return rc;
}

double tg_rung_to_dt(double t_max, int rung) { return rung < 0 || rung > TG_RUNG_MAX ? -1.0 : t_max / double(1 << rung); }

int tg_dt_to_rung(double t_max, double dt) {
	int rung = 0;
	while (tg_rung_to_dt(t_max, rung) > dt) {
		rung++;
		if (rung > TG_RUNG_MAX) return -1;
	}
	return rung;
}

int tg_min_rung(uint32_t t) {
	int r = 31;
	while ((t & 1u) == 0 && r != 0) {
		r--;
		t >>= 1;
	}
	return r;
}

uint32_t tg_time_inc(uint32_t t, int max_rung) { return t + (uint32_t(1) << uint32_t(31 - max_rung)); }

double tg_time_to_double(double t_max, uint32_t t) { return (double(t) / double(uint32_t(1) << 31)) * t_max; }

}  // extern "C"
