// tg_calib.hip -- the VALU ceilings the interaction kernels are measured against.
//
// Two register-only loops, timed with HIP events on the calling device:
//   * independent FP32 FMA chains: the sustained vector FP32 rate of the part as clocked under
//     load (the published 157.3 TFLOP/s counts packed FP32, which this library does not use);
//   * the PP far-field pair body (3 integer differences, 3 conversions, r^2, rsqrt, r^-3, 3 FMA for
//     g, 1 add for phi) on synthetic register data: the best rate the PP kernel's instruction mix
//     can reach with no memory, LDS or loop overhead.
// Each lane stores one float at the end, inside the buffer the host allocates for the grid.
#include <hip/hip_runtime.h>

#include <string>

#include "tg_calib.h"

namespace {

constexpr int kBlocks = 256 * 8 * 4, kThreads = 256, kIters = 4096;

__global__ __launch_bounds__(256) void k_calib_fma(float* out, int iters, float a, float b) {
	float v[8];
#pragma unroll
	for (int k = 0; k < 8; k++) v[k] = threadIdx.x * 1e-3f + k;
	for (int i = 0; i < iters; i++) {
#pragma unroll
		for (int k = 0; k < 8; k++) v[k] = fmaf(v[k], a, b);
	}
	float s = 0.f;
#pragma unroll
	for (int k = 0; k < 8; k++) s += v[k];
	out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_calib_pair(float* out, int iters, int sx, int sy, int sz) {
	const int xi = threadIdx.x * 977, yi = threadIdx.x * 131, zi = threadIdx.x * 7919;
	float ax = 0.f, ay = 0.f, az = 0.f, pp = 0.f;
	for (int i = 0; i < iters; i++) {
		const int xj = sx + i * 3, yj = sy + i * 5, zj = sz + i * 7;
		const float fx = float(xi - xj), fy = float(yi - yj), fz = float(zi - zj);
		const float r2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
		const float ri = __builtin_amdgcn_rsqf(r2);
		const float ri3 = ri * ri * ri;
		ax = fmaf(fx, ri3, ax);
		ay = fmaf(fy, ri3, ay);
		az = fmaf(fz, ri3, az);
		pp += ri;
	}
	out[blockIdx.x * blockDim.x + threadIdx.x] = ax + ay + az + pp;
}

thread_local std::string g_calib_err;

}  // namespace

extern "C" int tg_calibrate(int device, int repeats, double* fma_tflops, double* pp_pairs_per_s) {
	if (!fma_tflops || !pp_pairs_per_s || repeats < 1) return 1;
	*fma_tflops = *pp_pairs_per_s = 0.0;
	float* out = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	hipStream_t s = nullptr;
	int rc = 0;
#define CCHK(x)                                          \
	do {                                                 \
		hipError_t e_ = (x);                             \
		if (e_ != hipSuccess) {                          \
			g_calib_err = std::string(#x) + ": " + hipGetErrorString(e_); \
			rc = 2;                                      \
			goto done;                                   \
		}                                                \
	} while (0)
	CCHK(hipSetDevice(device));
	CCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
	CCHK(hipMalloc(&out, size_t(kBlocks) * kThreads * sizeof(float)));
	CCHK(hipEventCreate(&e0));
	CCHK(hipEventCreate(&e1));
	for (int r = 0; r < repeats + 1; r++) {  // the first launch of each loop is a warmup
		float ms = 0.f;
		CCHK(hipEventRecord(e0, s));
		hipLaunchKernelGGL(k_calib_fma, dim3(kBlocks), dim3(kThreads), 0, s, out, kIters, 0.999f, 1e-3f);
		CCHK(hipGetLastError());
		CCHK(hipEventRecord(e1, s));
		CCHK(hipEventSynchronize(e1));
		CCHK(hipEventElapsedTime(&ms, e0, e1));
		const double tf = double(kBlocks) * kThreads * kIters * 8 * 2 / (ms * 1e-3) / 1e12;
		if (r > 0 && tf > *fma_tflops) *fma_tflops = tf;
		CCHK(hipEventRecord(e0, s));
		hipLaunchKernelGGL(k_calib_pair, dim3(kBlocks), dim3(kThreads), 0, s, out, kIters, 12345, 67890, 13579);
		CCHK(hipGetLastError());
		CCHK(hipEventRecord(e1, s));
		CCHK(hipEventSynchronize(e1));
		CCHK(hipEventElapsedTime(&ms, e0, e1));
		const double pr = double(kBlocks) * kThreads * kIters / (ms * 1e-3);
		if (r > 0 && pr > *pp_pairs_per_s) *pp_pairs_per_s = pr;
	}
done:
#undef CCHK
	if (e0) (void) hipEventDestroy(e0);
	if (e1) (void) hipEventDestroy(e1);
	if (out) (void) hipFree(out);
	if (s) (void) hipStreamDestroy(s);
	return rc;
}

extern "C" const char* tg_calibrate_last_error(void) { return g_calib_err.c_str(); }
