// tg_noalloc.hip -- tg_gravity_PP_direct_noalloc: the per-call particle-particle kernel of
// tg_gravity_PP_direct with no allocation at all.  Its staging (the sink and source records, the
// work items, the caller's forces and the kernel's sums) lives in `scratch`, a file-scope device
// array of fixed size: every device holds its own copy from the library's load until the process
// ends, and every call on one device reuses it.  The host side reaches it only through the address
// hipGetSymbolAddress gives on each device (the array's name is a host-side stand-in there).
// Calls that overlap on one device race on it; calls on different devices do not.
//
// The steps give the same bits as tg_gravity_PP_direct: the same kernel on the same records, its
// sums added to the caller's forces one addition per component, then phi_self subtracted
// (src/gravity.cpp:187,290-296).  Buffers are cleared by kernels (no memset), so the sequence is
// also fit for stream capture.
#include "tg_gravity.h"
#include "tg_kernels.hpp"

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

namespace tg {
void set_capi_error(const char* where, hipError_t e);  // tg_capi.cpp: the calling thread's tg_last_error()
void set_capi_message(const char* msg);
}  // namespace tg

namespace {

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

constexpr int64_t MAX_SINKS = TG_NOALLOC_MAX_SINKS;
constexpr int64_t MAX_SOURCES = TG_NOALLOC_MAX_SOURCES;
constexpr int64_t MAX_ITEMS = (MAX_SINKS + 63) / 64;
// byte offsets of the staging areas in the array
constexpr size_t OFF_SINKS = 0;                                     // int4 per sink
constexpr size_t OFF_SUMS = align256(OFF_SINKS + 16 * MAX_SINKS);   // 4 doubles per sink: the kernel's sums
constexpr size_t OFF_F = align256(OFF_SUMS + 32 * MAX_SINKS);       // 4 doubles per sink: the caller's forces
constexpr size_t OFF_X = align256(OFF_F + 32 * MAX_SINKS);          // 3 int32 per sink, as passed
constexpr size_t OFF_ITEMS = align256(OFF_X + 12 * MAX_SINKS);      // int4 per 64 sinks
constexpr size_t OFF_SRCS = align256(OFF_ITEMS + 16 * MAX_ITEMS);   // int4 per source
constexpr size_t OFF_Y = align256(OFF_SRCS + 16 * MAX_SOURCES);     // 3 int32 per source, as passed
constexpr size_t SCRATCH_BYTES = align256(OFF_Y + 12 * MAX_SOURCES);
static_assert(SCRATCH_BYTES <= (size_t(9) << 20), "scratch stays below 9 MiB");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex scratch_mtx;
std::vector<unsigned char*> scratch_at;  // the array's device address, per device (null: not set up)

// the array's address on `device` (current), looked up once per device
hipError_t scratch_on(int device, unsigned char** out) {
	std::lock_guard<std::mutex> lock(scratch_mtx);
	if ((int) scratch_at.size() <= device) scratch_at.resize(size_t(device) + 1, nullptr);
	if (!scratch_at[size_t(device)]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess) return e;
		scratch_at[size_t(device)] = static_cast<unsigned char*>(p);
	}
	*out = scratch_at[size_t(device)];
	return hipSuccess;
}

// sink and source records (position, mass bits), the work items (64 sinks each, the whole source
// list, every tile softened), and zeroed sums
__global__ void k_noalloc_stage(const int32_t* __restrict__ x, int64_t ns, const int32_t* __restrict__ y, int64_t nj, float m,
		int4* __restrict__ sinks, int4* __restrict__ srcs, int4* __restrict__ items, double* __restrict__ sums) {
	const int wbits = __float_as_int(m);
	const int64_t ngroups = (ns + 63) / 64;
	const int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nj || i < 4 * ns; i += stride) {
		if (i < ns) sinks[i] = make_int4(x[3 * i], x[3 * i + 1], x[3 * i + 2], wbits);
		if (i < nj) srcs[i] = make_int4(y[3 * i], y[3 * i + 1], y[3 * i + 2], wbits);
		if (i < ngroups) {
			const int64_t b = 64 * i;
			const int cnt = int(ns - b < 64 ? ns - b : 64);
			items[i] = make_int4(int(b), tg::pp_item_y(cnt, tg::PP_NEAR_ALL), 0, int(nj));
		}
		if (i < 4 * ns) sums[i] = 0.0;
	}
}

// the caller's forces plus the kernel's sums (when there were sources), then phi_self off the
// potential: the additions tg_gravity_PP_direct makes on the host, in the same order
__global__ void k_noalloc_finish(double* __restrict__ f, const double* __restrict__ sums, int64_t ns, int add, double phi_self) {
	const int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < 4 * ns; i += stride) {
		double v = f[i];
		if (add) v += sums[i];
		if ((i & 3) == 0) v -= phi_self;
		f[i] = v;
	}
}

unsigned grid_for(int64_t n) {
	const int64_t g = (n + 255) / 256;
	return unsigned(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

#define NACHK(call)                                       \
	do {                                                  \
		hipError_t e_ = (call);                           \
		if (e_ != hipSuccess) {                           \
			tg::set_capi_error(#call, e_);                \
			return e_ == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP; \
		}                                                 \
	} while (0)

}  // namespace

extern "C" int tg_gravity_PP_direct_noalloc(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const int32_t* y,
		int64_t nj, int do_phi, uint64_t* flop) {
	if (!p || ns < 0 || nj < 0 || (ns > 0 && (!f || !x)) || (nj > 0 && !y)) {
		tg::set_capi_message("tg_gravity_PP_direct_noalloc: bad argument");
		return TG_EINVAL;
	}
	if (ns > MAX_SINKS || nj > MAX_SOURCES) {
		tg::set_capi_message(ns > MAX_SINKS ? "tg_gravity_PP_direct_noalloc: more sinks than TG_NOALLOC_MAX_SINKS"
		                                    : "tg_gravity_PP_direct_noalloc: more sources than TG_NOALLOC_MAX_SOURCES");
		return TG_EINVAL;
	}
	if (flop) *flop = 0;
	if (ns == 0) return TG_OK;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
		tg::set_capi_message("no HIP device");
		return TG_ENODEV;
	}
	if (p->device < 0 || p->device >= ndev) {
		tg::set_capi_message("device ordinal out of range");
		return TG_EINVAL;
	}
	NACHK(hipSetDevice(p->device));
	unsigned char* base = nullptr;
	NACHK(scratch_on(p->device, &base));
	int4* sinks = reinterpret_cast<int4*>(base + OFF_SINKS);
	double* sums = reinterpret_cast<double*>(base + OFF_SUMS);
	double* fd = reinterpret_cast<double*>(base + OFF_F);
	int32_t* xd = reinterpret_cast<int32_t*>(base + OFF_X);
	int4* items = reinterpret_cast<int4*>(base + OFF_ITEMS);
	int4* srcs = reinterpret_cast<int4*>(base + OFF_SRCS);
	int32_t* yd = reinterpret_cast<int32_t*>(base + OFF_Y);
	const hipStream_t s = hipStreamPerThread;
	static_assert(sizeof(tg_force) == 32, "tg_force is 4 doubles");
	NACHK(hipMemcpyAsync(fd, f, size_t(ns) * sizeof(tg_force), hipMemcpyHostToDevice, s));
	if (nj > 0) {
		NACHK(hipMemcpyAsync(xd, x, size_t(ns) * 12, hipMemcpyHostToDevice, s));
		NACHK(hipMemcpyAsync(yd, y, size_t(nj) * 12, hipMemcpyHostToDevice, s));
		const float m = float(p->m);
		hipLaunchKernelGGL(k_noalloc_stage, dim3(grid_for(nj > 4 * ns ? nj : 4 * ns)), dim3(256), 0, s, xd, ns, yd, nj, m, sinks, srcs,
				items, sums);
		NACHK(hipGetLastError());
		tg::LeafArgs a{};
		a.sinks = sinks;
		a.srcs = srcs;
		a.multis = nullptr;
		a.items = items;
		a.list = nullptr;
		a.out = sums;
		a.nitems = int((ns + 63) / 64);
		a.wrap = p->ewald ? 1 : 0;
		a.uniform = 1;
		a.m = m;
		const double hu = p->soft_len * 4294967296.0;
		a.h2u = float(hu * hu);
		a.hinvu = float(1.0 / hu);
		a.hinv3u = float(1.0 / (hu * hu * hu));
		a.queue = nullptr;
		NACHK(tg::launch_pp_direct(a, do_phi != 0, s));
	}
	const double phi_self = (-315.0 / 128.0) * p->m / p->soft_len;
	hipLaunchKernelGGL(k_noalloc_finish, dim3(grid_for(4 * ns)), dim3(256), 0, s, fd, sums, ns, int(nj > 0), phi_self);
	NACHK(hipGetLastError());
	NACHK(hipMemcpyAsync(f, fd, size_t(ns) * sizeof(tg_force), hipMemcpyDeviceToHost, s));
	NACHK(hipStreamSynchronize(s));
	if (flop) *flop = ((74 + uint64_t(do_phi != 0) * 15) * uint64_t(nj) + 16 * 4 + 5) * uint64_t(ns);
	return TG_OK;
}
