// tg_capi.cpp -- the per-call C ABI (include/tg_gravity.h): the eight reference kernel entry
// points, each a synchronous device round trip on the calling thread's own stream.
//
// Semantics follow the reference functions they replace (src/gravity.cpp:171-785):
//   * results are added to the caller's f / L;
//   * PP/PC with no sinks and CC/CP with no sources do nothing and report 0 flops;
//   * PP_direct subtracts phi_self = -315/128 m/h from every sink's potential on every call,
//     also when do_phi is false or there are no sources (src/gravity.cpp:187,290-296);
//   * the reference's weighted flop count is returned through *flop.
#include "tg_gravity.h"
#include "tg_kernels.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int set_hip_error(hipError_t e, const char* where) {
	g_err = std::string(where) + ": " + hipGetErrorString(e);
	return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? TG_ENOMEM : TG_EHIP;
}

#define HIPCHK(call)                                          \
	do {                                                      \
		hipError_t e_ = (call);                               \
		if (e_ != hipSuccess) return set_hip_error(e_, #call); \
	} while (0)

// growable device buffer (freed by its owner, on the device it was allocated on)
struct DevBuf {
	void* p = nullptr;
	size_t cap = 0;
	void release() {
		if (p) hipFree(p);
		p = nullptr;
		cap = 0;
	}
	hipError_t ensure(size_t bytes) {
		if (bytes <= cap) return hipSuccess;
		if (p) hipFree(p);
		p = nullptr;
		cap = 0;
		size_t want = bytes < 4096 ? 4096 : bytes + bytes / 2;
		hipError_t e = hipMalloc(&p, want);
		if (e == hipSuccess) cap = want;
		return e;
	}
};

// per host thread: a stream and staging buffers on one device
struct ThreadCtx {
	int device = -1;
	hipStream_t stream = nullptr;
	DevBuf sinks, srcs, multis, items, out;
	std::vector<int4> hsinks, hsrcs, hitems;
	std::vector<int2> hwaves;
	std::vector<double> hout;
	~ThreadCtx() {
		// device memory is left to process teardown: a thread may outlive the HIP runtime
	}
};

thread_local ThreadCtx t_ctx;

int get_ctx(const tg_params* p, ThreadCtx** out) {
	if (!p) {
		g_err = "null tg_params";
		return TG_EINVAL;
	}
	ThreadCtx& T = t_ctx;
	if (T.device != p->device) {
		int n = 0;
		if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
			g_err = "no HIP device";
			return TG_ENODEV;
		}
		if (p->device < 0 || p->device >= n) {
			g_err = "device ordinal out of range";
			return TG_EINVAL;
		}
		if (T.device >= 0 && hipSetDevice(T.device) == hipSuccess) {
			// the previous device's stream and staging buffers, released on that device
			if (T.stream) hipStreamDestroy(T.stream);
			for (DevBuf* b : {&T.sinks, &T.srcs, &T.multis, &T.items, &T.out}) b->release();
		}
		T = ThreadCtx();  // device -1 until the new device is set up
		HIPCHK(hipSetDevice(p->device));
		HIPCHK(tg::init_device_tables(p->device));
		HIPCHK(hipStreamCreateWithFlags(&T.stream, hipStreamNonBlocking));
		T.device = p->device;
	} else {
		HIPCHK(hipSetDevice(p->device));
	}
	*out = &T;
	return TG_OK;
}

inline int4 pack(const int32_t* x, float w) {
	int4 v;
	v.x = x[0];
	v.y = x[1];
	v.z = x[2];
	std::memcpy(&v.w, &w, 4);
	return v;
}

// particle sinks x (ns) against particle (or multipole) sources: shared by PP and PC
int leaf_call(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const void* y, int64_t nj, int kind,
		int do_phi) {
	ThreadCtx* T;
	int rc = get_ctx(p, &T);
	if (rc) return rc;
	const float m = float(p->m);
	T->hsinks.resize(ns);
	for (int64_t i = 0; i < ns; i++) T->hsinks[i] = pack(x + 3 * i, m);
	const bool multi = (kind == tg::PC_DIRECT || kind == tg::PC_EWALD);
	const int64_t ngroups = (ns + 63) / 64;
	T->hitems.resize(ngroups);
	for (int64_t g = 0; g < ngroups; g++) {
		const int64_t b = g * 64;
		T->hitems[g] = make_int4(int(b), tg::pp_item_y(int(std::min<int64_t>(64, ns - b)), tg::PP_NEAR_ALL), 0, int(nj));
	}
	HIPCHK(T->sinks.ensure(ns * sizeof(int4)));
	HIPCHK(T->items.ensure(ngroups * sizeof(int4)));
	HIPCHK(T->out.ensure(ns * 4 * sizeof(double)));
	hipStream_t s = T->stream;
	HIPCHK(hipMemcpyAsync(T->sinks.p, T->hsinks.data(), ns * sizeof(int4), hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(T->items.p, T->hitems.data(), ngroups * sizeof(int4), hipMemcpyHostToDevice, s));
	HIPCHK(hipMemsetAsync(T->out.p, 0, ns * 4 * sizeof(double), s));
	if (multi) {
		HIPCHK(T->multis.ensure(nj * sizeof(tg_multi_src)));
		HIPCHK(hipMemcpyAsync(T->multis.p, y, nj * sizeof(tg_multi_src), hipMemcpyHostToDevice, s));
	} else {
		T->hsrcs.resize(nj);
		const int32_t* yp = static_cast<const int32_t*>(y);
		for (int64_t j = 0; j < nj; j++) T->hsrcs[j] = pack(yp + 3 * j, m);
		HIPCHK(T->srcs.ensure(nj * sizeof(int4)));
		HIPCHK(hipMemcpyAsync(T->srcs.p, T->hsrcs.data(), nj * sizeof(int4), hipMemcpyHostToDevice, s));
	}
	tg::LeafArgs a{};
	a.sinks = static_cast<const int4*>(T->sinks.p);
	a.srcs = static_cast<const int4*>(T->srcs.p);
	a.multis = static_cast<const tg::MultiSrc*>(T->multis.p);
	a.items = static_cast<const int4*>(T->items.p);
	a.list = nullptr;
	a.out = static_cast<double*>(T->out.p);
	a.nitems = int(ngroups);
	a.wrap = p->ewald ? 1 : 0;
	a.uniform = 1;
	a.m = m;
	const double hu = p->soft_len * 4294967296.0;
	a.h2u = float(hu * hu);
	a.hinvu = float(1.0 / hu);
	a.hinv3u = float(1.0 / (hu * hu * hu));
	switch (kind) {
		case tg::PP_DIRECT: HIPCHK(tg::launch_pp_direct(a, do_phi != 0, s)); break;
		case tg::PC_DIRECT: HIPCHK(tg::launch_pc(a, false, s)); break;
		case tg::PC_EWALD: HIPCHK(tg::launch_pc(a, true, s)); break;
		case tg::PP_EWALD: HIPCHK(tg::launch_pp_ewald(a, s)); break;
	}
	T->hout.resize(ns * 4);
	HIPCHK(hipMemcpyAsync(T->hout.data(), T->out.p, ns * 4 * sizeof(double), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	for (int64_t i = 0; i < ns; i++) {
		f[i].phi += T->hout[4 * i];
		f[i].g[0] += T->hout[4 * i + 1];
		f[i].g[1] += T->hout[4 * i + 2];
		f[i].g[2] += T->hout[4 * i + 3];
	}
	return TG_OK;
}

// one expansion centre x against nj sources: shared by CC and CP
int node_call(const tg_params* p, double* L, const int32_t* x, const void* y, int64_t nj, int kind) {
	ThreadCtx* T;
	int rc = get_ctx(p, &T);
	if (rc) return rc;
	const float m = float(p->m);
	const bool multi = (kind == tg::CC_DIRECT || kind == tg::CC_EWALD);
	// one wave (and one partial slot) per 64 sources
	const int64_t chunk = 64;
	const int64_t nitems = (nj + chunk - 1) / chunk;
	T->hwaves.resize(nitems);
	for (int64_t i = 0; i < nitems; i++) T->hwaves[i] = make_int2(int(i * chunk), int(std::min<int64_t>(nj, (i + 1) * chunk)));
	T->hsinks.assign(1, pack(x, 0.0f));
	HIPCHK(T->sinks.ensure(sizeof(int4)));
	HIPCHK(T->items.ensure(nitems * sizeof(int2)));
	HIPCHK(T->out.ensure(nitems * 35 * sizeof(double)));
	hipStream_t s = T->stream;
	HIPCHK(hipMemcpyAsync(T->sinks.p, T->hsinks.data(), sizeof(int4), hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(T->items.p, T->hwaves.data(), nitems * sizeof(int2), hipMemcpyHostToDevice, s));
	if (multi) {
		HIPCHK(T->multis.ensure(nj * sizeof(tg_multi_src)));
		HIPCHK(hipMemcpyAsync(T->multis.p, y, nj * sizeof(tg_multi_src), hipMemcpyHostToDevice, s));
	} else {
		T->hsrcs.resize(nj);
		const int32_t* yp = static_cast<const int32_t*>(y);
		for (int64_t j = 0; j < nj; j++) T->hsrcs[j] = pack(yp + 3 * j, m);
		HIPCHK(T->srcs.ensure(nj * sizeof(int4)));
		HIPCHK(hipMemcpyAsync(T->srcs.p, T->hsrcs.data(), nj * sizeof(int4), hipMemcpyHostToDevice, s));
	}
	tg::NodeArgs a{};
	a.sinkpos = static_cast<const int4*>(T->sinks.p);
	a.srcs = static_cast<const int4*>(T->srcs.p);
	a.multis = static_cast<const tg::MultiSrc*>(T->multis.p);
	a.waves = static_cast<const int2*>(T->items.p);
	a.list = nullptr;
	a.owner = nullptr;
	a.slot_node = nullptr;
	a.partial = static_cast<double*>(T->out.p);
	a.nwaves = int(nitems);
	a.wrap = p->ewald ? 1 : 0;
	switch (kind) {
		case tg::CC_DIRECT: HIPCHK(tg::launch_cc(a, false, s)); break;
		case tg::CC_EWALD: HIPCHK(tg::launch_cc(a, true, s)); break;
		case tg::CP_DIRECT: HIPCHK(tg::launch_cp(a, false, s)); break;
		case tg::CP_EWALD: HIPCHK(tg::launch_cp(a, true, s)); break;
	}
	T->hout.resize(nitems * 35);
	HIPCHK(hipMemcpyAsync(T->hout.data(), T->out.p, nitems * 35 * sizeof(double), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	for (int c = 0; c < 35; c++) {
		double acc = 0.0;
		for (int64_t i = 0; i < nitems; i++) acc += T->hout[35 * i + c];
		L[c] += acc;
	}
	return TG_OK;
}

bool bad_sizes(const void* f, const void* x, int64_t ns, const void* y, int64_t nj) {
	return ns < 0 || nj < 0 || (ns > 0 && (!f || !x)) || (nj > 0 && !y) || ns > (int64_t(1) << 30) || nj > (int64_t(1) << 30);
}

}  // namespace

namespace tg {
// the calling thread's tg_last_error() message, set from other translation units
void set_capi_error(const char* where, hipError_t e) { set_hip_error(e, where); }
void set_capi_message(const char* msg) { g_err = msg; }
}  // namespace tg

extern "C" {

int tg_gravity_PP_direct(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj,
		int do_phi, uint64_t* flop) {
	if (!p || bad_sizes(f, x, ns, y, nj)) return TG_EINVAL;
	if (flop) *flop = 0;
	if (ns == 0) return TG_OK;
	if (nj > 0) {
		int rc = leaf_call(p, f, x, ns, y, nj, tg::PP_DIRECT, do_phi);
		if (rc) return rc;
	}
	const double phi_self = (-315.0 / 128.0) * p->m / p->soft_len;
	for (int64_t i = 0; i < ns; i++) f[i].phi -= phi_self;
	if (flop) *flop = ((74 + uint64_t(do_phi != 0) * 15) * uint64_t(nj) + 16 * 4 + 5) * uint64_t(ns);
	return TG_OK;
}

int tg_gravity_PC_direct(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const tg_multi_src* y, int64_t nc,
		uint64_t* flop) {
	if (!p || bad_sizes(f, x, ns, y, nc)) return TG_EINVAL;
	if (flop) *flop = 0;
	if (ns == 0 || nc == 0) return TG_OK;
	int rc = leaf_call(p, f, x, ns, y, nc, tg::PC_DIRECT, 1);
	if (rc == TG_OK && flop) *flop = 581 * uint64_t(nc) * uint64_t(ns);
	return rc;
}

int tg_gravity_PP_ewald(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj,
		uint64_t* flop) {
	if (!p || bad_sizes(f, x, ns, y, nj)) return TG_EINVAL;
	if (flop) *flop = 0;
	if (ns == 0 || nj == 0) return TG_OK;
	int rc = leaf_call(p, f, x, ns, y, nj, tg::PP_EWALD, 1);
	if (rc == TG_OK && flop) *flop = 56245 * uint64_t(nj) * uint64_t(ns);
	return rc;
}

int tg_gravity_PC_ewald(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const tg_multi_src* y, int64_t nc,
		uint64_t* flop) {
	if (!p || bad_sizes(f, x, ns, y, nc)) return TG_EINVAL;
	if (flop) *flop = 0;
	if (ns == 0 || nc == 0) return TG_OK;
	int rc = leaf_call(p, f, x, ns, y, nc, tg::PC_EWALD, 1);
	if (rc == TG_OK && flop) *flop = 251531 * uint64_t(nc) * uint64_t(ns);
	return rc;
}

int tg_gravity_CC_direct(const tg_params* p, double* L, const int32_t* x, const tg_multi_src* y, int64_t nc, uint64_t* flop) {
	if (!p || nc < 0 || (nc > 0 && (!L || !x || !y))) return TG_EINVAL;
	if (flop) *flop = 0;
	if (nc == 0) return TG_OK;
	int rc = node_call(p, L, x, y, nc, tg::CC_DIRECT);
	if (rc == TG_OK && flop) *flop = 1025 * uint64_t(nc);
	return rc;
}

int tg_gravity_CC_ewald(const tg_params* p, double* L, const int32_t* x, const tg_multi_src* y, int64_t nc, uint64_t* flop) {
	if (!p || nc < 0 || (nc > 0 && (!L || !x || !y))) return TG_EINVAL;
	if (flop) *flop = 0;
	if (nc == 0) return TG_OK;
	int rc = node_call(p, L, x, y, nc, tg::CC_EWALD);
	if (rc == TG_OK && flop) *flop = 251975 * uint64_t(nc);
	return rc;
}

int tg_gravity_CP_direct(const tg_params* p, double* L, const int32_t* x, const int32_t* y, int64_t nj, uint64_t* flop) {
	if (!p || nj < 0 || (nj > 0 && (!L || !x || !y))) return TG_EINVAL;
	if (flop) *flop = 0;
	if (nj == 0) return TG_OK;
	int rc = node_call(p, L, x, y, nj, tg::CP_DIRECT);
	if (rc == TG_OK && flop) *flop = 455 * uint64_t(nj);
	return rc;
}

int tg_gravity_CP_ewald(const tg_params* p, double* L, const int32_t* x, const int32_t* y, int64_t nj, uint64_t* flop) {
	if (!p || nj < 0 || (nj > 0 && (!L || !x || !y))) return TG_EINVAL;
	if (flop) *flop = 0;
	if (nj == 0) return TG_OK;
	int rc = node_call(p, L, x, y, nj, tg::CP_EWALD);
	if (rc == TG_OK && flop) *flop = 251233 * uint64_t(nj);
	return rc;
}

double tg_ewald_near_separation(const double x[3]) {
	double d = 0.0;
	for (int dim = 0; dim < 3; dim++) {
		const double ax = std::abs(x[dim]);
		const double t = std::min(ax, 1.0 - ax);
		d += t * t;
	}
	return std::sqrt(d);
}

double tg_ewald_far_separation(const double x[3], double /*r*/) { return std::max(0.25, tg_ewald_near_separation(x)); }

const char* tg_last_error(void) { return g_err.c_str(); }

int tg_device_count(void) {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) return 0;
	return n;
}

int tg_version(void) { return 1; }

}  // extern "C"
