// tg_kernels.hip -- gfx950 kernels for tigergrav's eight interaction kinds.
//
// Mapping to CDNA4:
//  * leaf kinds (PP, PC): one 64-lane wave per item; lane = (sink, phase).  A leaf of n <= 64
//    sinks is padded to the next power of two nsp and the wave runs 64/nsp phases, each phase
//    taking every (64/nsp)-th source of the tile, so small leaves still fill the wave.  Sources
//    are staged 64 at a time through LDS and read back as broadcasts.  Per-tile partial sums are
//    float, the running sum FP64 (the reference accumulates float terms in FP64 lanes,
//    src/gravity.cpp:259-261), and phases are merged with cross-lane FP64 shuffles.
//  * node kinds (CC, CP): lane = source; a wave packs the short lists of several expansion
//    centres (<= 64 entries), 35 float accumulators per lane, merged per centre in FP64 through
//    LDS by a segmented scan.
//  * Ewald kinds use tg::green_ewald (tg_math.hpp) and, for PP, green_ewald1_real below.
#include "tg_kernels.hpp"
#include "tg_math.hpp"

#include <climits>
#include <cmath>
#include <mutex>
#include <vector>

namespace tg {

__constant__ const float* c_ewt;

// ---- the tabulated correction -------------------------------------------------------------
static std::mutex table_mtx;                     // guards device_table and the per-device set-up
static std::vector<const float*> device_table;  // the tabulated correction on each device

const float* ewald_table_on(int device) {
	std::lock_guard<std::mutex> lock(table_mtx);
	return device >= 0 && device < (int) device_table.size() ? device_table[size_t(device)] : nullptr;
}

hipError_t init_device_tables(int device) {
	static std::vector<int> done;
	std::lock_guard<std::mutex> lock(table_mtx);
	if ((int) done.size() <= device) done.resize(device + 1, 0);
	if (done[device]) return hipSuccess;
	hipError_t e;
	if ((e = hipSetDevice(device)) != hipSuccess) return e;
	{
		// the tabulated correction: built once per process on the host, copied to every device, and
		// kept until the process ends
		static std::vector<float> table;
		if (table.empty()) ewald_table_build(table);
		float* dt = nullptr;
		if ((e = hipMalloc(&dt, table.size() * sizeof(float))) != hipSuccess) return e;
		if ((e = hipMemcpy(dt, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) return e;
		const float* cdt = dt;
		if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_ewt), &cdt, sizeof(cdt))) != hipSuccess) return e;
		if ((int) device_table.size() <= device) device_table.resize(size_t(device) + 1, nullptr);
		device_table[size_t(device)] = cdt;
	}
	done[device] = 1;
	return hipSuccess;
}

// ---- separations -----------------------------------------------------------------------------
// position difference a - b in position units (2^32 per box), as float:
//   WRAP: the int32 wrap-around difference (periodic minimum image, src/gravity.cpp:221-225);
//   else: the exact 33-bit difference, rounded once (src/gravity.cpp:227-229).
template <bool WRAP>
__device__ __forceinline__ float sepu(int a, int b) {
	const unsigned d = unsigned(a) - unsigned(b);
	if (WRAP) return float(int(d));
	const bool neg = a < b;
	const float f = float(neg ? 0u - d : d);
	return neg ? -f : f;
}

__device__ __forceinline__ int ceil_log2_64(int n) { return n <= 1 ? 0 : 32 - __clz(n - 1); }

__device__ __forceinline__ double wave_xor(double v, int off) { return __shfl_xor(v, off, 64); }

// ---- PP direct -------------------------------------------------------------------------------
// Spline-softened (GADGET-2 cubic spline, support h) potential and acceleration
// (src/gravity.cpp:231-287).  Beyond h it is Newton's 1/r.  A leaf item's source list starts with
// the sources that may come within h of its sinks (the near prefix, PP_NEAR_LEN of the item);
// tiles that reach into it run the per-lane softened branch, every later tile the straight-line
// far field.  The 64-source tile is stored phase-major in LDS, so every lane walks consecutive
// entries with a wave-uniform trip count; a partial last tile is padded with zero-weight entries.
// With equal masses (UNIFORM) the mass multiplies the sums once at the end.
template <bool USE_W, bool PHI>
__device__ __forceinline__ void pp_pair(const int4& s, float h2u, float hinvu, float hinv3u, float& ax, float& ay, float& az,
		float& pp, float fx, float fy, float fz) {
	const float r2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
	const float w = __int_as_float(s.w);
	if (r2 > h2u) {
		const float ri = __builtin_amdgcn_rsqf(r2);
		const float wri = USE_W ? w * ri : ri;
		const float wri3 = wri * ri * ri;
		ax = fmaf(fx, wri3, ax);
		ay = fmaf(fy, wri3, ay);
		az = fmaf(fz, wri3, az);
		if (PHI) pp += wri;
	} else {
		const float ru = sqrtf(r2);
		const float q = ru * hinvu;
		float f, p;
		if (q < 0.5f) {
			const float q2 = q * q;
			f = fmaf(q2, fmaf(32.0f, q, -192.0f / 5.0f), 32.0f / 3.0f) * hinv3u;
			p = fmaf(q2, fmaf(q2, fmaf(-32.0f / 5.0f, q, 48.0f / 5.0f), -16.0f / 3.0f), 14.0f / 5.0f) * hinvu;
		} else {
			const float ri = 1.0f / ru;
			const float q3 = q * q * q;
			f = fmaf(fmaf(fmaf(fmaf(-32.0f / 3.0f, q, 192.0f / 5.0f), q, -48.0f), q, 64.0f / 3.0f), q3, -1.0f / 15.0f) *
			    (ri * ri * ri);
			p = fmaf(fmaf(fmaf(fmaf(fmaf(32.0f / 15.0f, q, -48.0f / 5.0f), q, 16.0f), q, -32.0f / 3.0f), q * q, 16.0f / 5.0f), q,
			         -1.0f / 15.0f) * ri;
		}
		const float wf = w * f;
		ax = fmaf(fx, wf, ax);
		ay = fmaf(fy, wf, ay);
		az = fmaf(fz, wf, az);
		if (PHI) pp = fmaf(w, p, pp);
	}
}

// a pair known to be at least h apart, from its exact separation (position units).  The
// separation of two distinct positions is at least one unit, so the floor of r2 changes no pair of
// the list and keeps a zero-weight pad that lands on a sink finite.
template <bool USE_W, bool PHI>
__device__ __forceinline__ void pp_far(const int4& s, float& ax, float& ay, float& az, float& pp, float fx, float fy, float fz) {
	const float r2 = fmaxf(fmaf(fz, fz, fmaf(fy, fy, fx * fx)), 1.0f);
	const float ri = __builtin_amdgcn_rsqf(r2);
	const float wri = USE_W ? __int_as_float(s.w) * ri : ri;
	const float wri3 = wri * ri * ri;
	ax = fmaf(fx, wri3, ax);
	ay = fmaf(fy, wri3, ay);
	az = fmaf(fz, wri3, az);
	if (PHI) pp += wri;
}

// the same far pair from offsets to a common centre: sink offset (ox, oy, oz) minus the source's
// offset (s.x, s.y, s.z), source weight s.w
template <bool USE_W, bool PHI>
__device__ __forceinline__ void pp_far_off(const float4& s, float ox, float oy, float oz, float& ax, float& ay, float& az,
		float& pp) {
	const float fx = ox - s.x, fy = oy - s.y, fz = oz - s.z;
	const float r2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
	const float ri = __builtin_amdgcn_rsqf(r2);
	const float wri = USE_W ? s.w * ri : ri;
	const float wri3 = wri * ri * ri;
	ax = fmaf(fx, wri3, ax);
	ay = fmaf(fy, wri3, ay);
	az = fmaf(fz, wri3, az);
	if (PHI) pp += wri;
}

// offsets of at least 2^30 position units (a quarter box) from the item's centre: the difference
// of two offsets below it is the wrapped separation itself
__device__ __forceinline__ bool off_big(float ox, float oy, float oz) {
	return fabsf(ox) >= 1073741824.0f || fabsf(oy) >= 1073741824.0f || fabsf(oz) >= 1073741824.0f;
}

// one PP item (<= 64 sinks of a leaf against its source list), by one 64-lane work group.
// Lane = (sink si, phase ph) with P = 64 / ns phases (lanes from P ns on idle), so a leaf of ns
// sinks keeps P ns of the 64 lanes busy.  A 64-source tile is staged phase-major in LDS: phase ph
// takes sources ph, ph + P, ... from K = ceil(64 / P) consecutive slots (slots no source fills hold
// zero-weight pads), so every lane walks its slots with a wave-uniform trip count.  The phases
// start an odd number of 16-byte slots apart, so the P addresses one read instruction broadcasts
// lie in different LDS banks.  Per-tile partial
// sums are float, the running sums FP64; the phases of a sink are added in phase order through LDS.
// Tiles of the near prefix take separations as exact integer differences rounded once
// (src/gravity.cpp:221-229) and the softened pair.  Far tiles take every separation as the
// difference of two float offsets from the item's first sink: a source's offset is converted
// once per tile when it is staged, so the pair costs one subtract per axis instead of an integer
// subtract and a conversion.  The offsets are integer differences rounded once, so a separation
// carries at most the rounding of the larger offset (2^-24 of it) beside its own.  Periodic tiles
// with an offset of a quarter box or more (possible only when leaves span that much) keep the
// exact form, so the wrap-around always picks the nearest image.
struct PPShared {
	int4 tile[128];
	float4 ftile[128];
	double red[4][64];
};

template <bool WRAP, bool PHI, bool UNIFORM>
__device__ __forceinline__ void pp_item(const LeafArgs& A, const int4 it, PPShared& S) {
	const int lane = threadIdx.x;
	const int ns = max(it.y & PP_COUNT_MASK, 1);
	const int near_len = int(unsigned(it.y) >> PP_NEAR_SHIFT);
	const int near_end = near_len == PP_NEAR_ALL ? INT_MAX : it.z + near_len;
	const int P = 64 / ns;             // phases
	const int K = (64 + P - 1) / P;    // slots per phase
	// phase stride in the tile: odd, so the phases' broadcast reads in one instruction fall on
	// different 4-bank groups of the LDS (conflict-free for up to 8 phases)
	const int KS = K | 1;
	const int ph = min(lane / ns, P - 1);
	const int si = lane - (lane / ns) * ns;
	const int4 me = A.sinks[it.x + si];
	const int4 c0 = A.sinks[it.x];     // the item's centre for the far tiles
	const float ox = sepu<WRAP>(me.x, c0.x), oy = sepu<WRAP>(me.y, c0.y), oz = sepu<WRAP>(me.z, c0.z);
	const bool item_off = !WRAP || __ballot(off_big(ox, oy, oz)) == 0ull;
	// source l of a tile goes to phase l % P, slot l / P; slots past the tile's sources are pads
	const int dst = (lane % P) * KS + lane / P;
	const int4 pad = make_int4(INT_MIN, INT_MIN, INT_MIN, 0);
	const float4 fpad = make_float4(3.0e30f, 3.0e30f, 3.0e30f, 0.0f);
	S.tile[lane] = pad;
	S.tile[lane + 64] = pad;
	S.ftile[lane] = fpad;
	S.ftile[lane + 64] = fpad;
	const int4* mine = S.tile + ph * KS;
	const float4* fmine = S.ftile + ph * KS;
	float ax = 0.f, ay = 0.f, az = 0.f, pp = 0.f;
	double dax = 0.0, day = 0.0, daz = 0.0, dpp = 0.0;
	// software pipeline: the next tile's source record (and the list index after it) are loaded
	// while the current tile is evaluated
	auto index_at = [&](int b) { return b + lane < it.w ? (A.list ? A.list[b + lane] : b + lane) : -1; };
	auto source_at = [&](int j) {
		int4 v = pad;  // zero weight
		if (j >= 0) {
			v = A.srcs[j];
			if (UNIFORM) v.w = __float_as_int(1.0f);
		}
		return v;
	};
	int next_idx0 = index_at(it.z);
	int4 next = source_at(next_idx0);
	bool next_pad = next_idx0 < 0;
	int next_idx = index_at(it.z + 64);
	for (int base = it.z; base < it.w; base += 64) {
		const int n = min(64, it.w - base);
		const bool near = base < near_end;
		// the staged source as an offset from the centre; a pad sits out of reach (r2 overflows to
		// infinity, its reciprocal root and weight are zero)
		float4 f = fpad;
		if (!next_pad) f = make_float4(sepu<WRAP>(next.x, c0.x), sepu<WRAP>(next.y, c0.y), sepu<WRAP>(next.z, c0.z),
				__int_as_float(next.w));
		const unsigned long long bigs = WRAP ? __ballot(!next_pad && off_big(f.x, f.y, f.z)) : 0ull;
		const bool use_off = item_off && !near && bigs == 0ull;
		__syncthreads();
		S.tile[dst] = next;
		S.ftile[dst] = f;
		__syncthreads();
		next_pad = next_idx < 0;
		next = source_at(next_idx);
		next_idx = index_at(base + 128);
		const int trip = n == 64 ? K : (n + P - 1) / P;
		if (near) {
			for (int j = 0; j < trip; j++) {
				const int4 s = mine[j];
				pp_pair<true, PHI>(s, A.h2u, A.hinvu, A.hinv3u, ax, ay, az, pp, sepu<WRAP>(me.x, s.x), sepu<WRAP>(me.y, s.y),
						sepu<WRAP>(me.z, s.z));
			}
		} else if (use_off) {
			// pads add exact zeros with or without their weight
#pragma unroll 8
			for (int j = 0; j < trip; j++) pp_far_off<!UNIFORM, PHI>(fmine[j], ox, oy, oz, ax, ay, az, pp);
		} else {
			for (int j = 0; j < trip; j++) {
				const int4 s = mine[j];
				pp_far<true, PHI>(s, ax, ay, az, pp, sepu<WRAP>(me.x, s.x), sepu<WRAP>(me.y, s.y), sepu<WRAP>(me.z, s.z));
			}
		}
		dax += ax;
		day += ay;
		daz += az;
		dpp += pp;
		ax = ay = az = pp = 0.f;
	}
	// the phases of each sink, added in phase order
	S.red[0][lane] = dax;
	S.red[1][lane] = day;
	S.red[2][lane] = daz;
	S.red[3][lane] = dpp;
	__syncthreads();
	if (lane < ns) {
		for (int q = 1; q < P; q++) {
			dax += S.red[0][q * ns + lane];
			day += S.red[1][q * ns + lane];
			daz += S.red[2][q * ns + lane];
			dpp += S.red[3][q * ns + lane];
		}
		// position units: g = -sum w dx r^-3 = -2^64 sum (w du ru^-3), phi = -2^32 sum (w / ru)
		const double ms = UNIFORM ? double(A.m) : 1.0;
		double* o = A.out + 4 * (size_t) (it.x + lane);
		if (PHI) o[0] += -4294967296.0 * ms * dpp;
		o[1] += -18446744073709551616.0 * ms * dax;
		o[2] += -18446744073709551616.0 * ms * day;
		o[3] += -18446744073709551616.0 * ms * daz;
	}
}

// one work group per item; the dispatcher hands items out in array order (heaviest first)
template <bool WRAP, bool PHI, bool UNIFORM>
__global__ __launch_bounds__(64) void k_pp_direct(LeafArgs A) {
	__shared__ PPShared S;
	pp_item<WRAP, PHI, UNIFORM>(A, A.items[blockIdx.x], S);
}

// persistent form: as many work groups as the machine keeps resident, each taking the next
// item from a queue counter (A.queue, zero at launch) until the items run out
template <bool WRAP, bool PHI, bool UNIFORM>
__global__ __launch_bounds__(64) void k_pp_direct_queue(LeafArgs A) {
	__shared__ PPShared S;
	__shared__ int next;
	for (;;) {
		__syncthreads();
		if (threadIdx.x == 0) next = atomicAdd(A.queue, 1);
		__syncthreads();
		const int k = next;
		if (k >= A.nitems) break;
		pp_item<WRAP, PHI, UNIFORM>(A, A.items[k], S);
	}
}

__global__ void k_zero_counter(int* c) {
	if (threadIdx.x == 0) c[0] = 0;
}

hipError_t launch_pp_direct(const LeafArgs& a, bool do_phi, hipStream_t s) {
	if (a.nitems <= 0) return hipSuccess;
	if (a.queue) {
		// the counter is cleared by a kernel (not a memset) so the launch sequence stays graph-safe
		hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, a.queue);
		const int groups = a.nitems < a.resident ? a.nitems : a.resident;
#define TG_PPQ(W, PH, U) hipLaunchKernelGGL((k_pp_direct_queue<W, PH, U>), dim3(groups), dim3(64), 0, s, a)
		const bool u = a.uniform != 0;
		if (a.wrap) {
			if (do_phi) { if (u) TG_PPQ(true, true, true); else TG_PPQ(true, true, false); }
			else { if (u) TG_PPQ(true, false, true); else TG_PPQ(true, false, false); }
		} else {
			if (do_phi) { if (u) TG_PPQ(false, true, true); else TG_PPQ(false, true, false); }
			else { if (u) TG_PPQ(false, false, true); else TG_PPQ(false, false, false); }
		}
#undef TG_PPQ
		return hipGetLastError();
	}
#define TG_PP(W, PH, U) hipLaunchKernelGGL((k_pp_direct<W, PH, U>), dim3(a.nitems), dim3(64), 0, s, a)
	const bool u = a.uniform != 0;
	if (a.wrap) {
		if (do_phi) { if (u) TG_PP(true, true, true); else TG_PP(true, true, false); }
		else { if (u) TG_PP(true, false, true); else TG_PP(true, false, false); }
	} else {
		if (do_phi) { if (u) TG_PP(false, true, true); else TG_PP(false, true, false); }
		else { if (u) TG_PP(false, false, true); else TG_PP(false, false, false); }
	}
#undef TG_PP
	return hipGetLastError();
}

// ---- PC (multipole -> particle) ----------------------------------------------------------------
template <bool EWALD, bool WRAP>
__global__ __launch_bounds__(64) void k_pc(LeafArgs A) {
	const int4 it = A.items[blockIdx.x];
	const int lane = threadIdx.x;
	const int ns = it.y & PP_COUNT_MASK;
	const int lg = ceil_log2_64(ns);
	const int nsp = 1 << lg;
	const int P = 64 >> lg;
	const int si = lane & (nsp - 1);
	const int ph = lane >> lg;
	const int4 me = A.sinks[it.x + min(si, ns - 1)];
	__shared__ float tile[64][20];
	float phi = 0.f, g[3] = {0.f, 0.f, 0.f};
	double dphi = 0.0, dg[3] = {0.0, 0.0, 0.0};
	for (int base = it.z; base < it.w; base += 64) {
		const int n = min(64, it.w - base);
		__syncthreads();
		if (lane < n) {
			const float4* src = reinterpret_cast<const float4*>(A.multis + (A.list ? A.list[base + lane] : base + lane));
#pragma unroll
			for (int q = 0; q < 5; q++) {
				const float4 v = src[q];
				tile[lane][4 * q] = v.x;
				tile[lane][4 * q + 1] = v.y;
				tile[lane][4 * q + 2] = v.z;
				tile[lane][4 * q + 3] = v.w;
			}
		}
		__syncthreads();
		for (int k = ph; k < n; k += P) {
			float M[17];
#pragma unroll
			for (int q = 0; q < 17; q++) M[q] = tile[k][q];
			const float x[3] = {sepu<EWALD || WRAP>(me.x, __float_as_int(tile[k][17])) * POS_INVF,
					sepu<EWALD || WRAP>(me.y, __float_as_int(tile[k][18])) * POS_INVF,
					sepu<EWALD || WRAP>(me.z, __float_as_int(tile[k][19])) * POS_INVF};
			float D[35];
			if (EWALD)
				green_ewald(D, x);
			else
				green_direct(D, x);
			m2p(phi, g, M, D);
		}
		dphi += phi;
		dg[0] += g[0];
		dg[1] += g[1];
		dg[2] += g[2];
		phi = g[0] = g[1] = g[2] = 0.f;
	}
	for (int off = nsp; off < 64; off <<= 1) {
		dphi += wave_xor(dphi, off);
		dg[0] += wave_xor(dg[0], off);
		dg[1] += wave_xor(dg[1], off);
		dg[2] += wave_xor(dg[2], off);
	}
	if (ph == 0 && si < ns) {
		double* o = A.out + 4 * (size_t) (it.x + si);
		o[0] += dphi;
		o[1] += dg[0];
		o[2] += dg[1];
		o[3] += dg[2];
	}
}

hipError_t launch_pc(const LeafArgs& a, bool ewald, hipStream_t s) {
	if (a.nitems <= 0) return hipSuccess;
	if (ewald)
		hipLaunchKernelGGL((k_pc<true, true>), dim3(a.nitems), dim3(64), 0, s, a);
	else if (a.wrap)
		hipLaunchKernelGGL((k_pc<false, true>), dim3(a.nitems), dim3(64), 0, s, a);
	else
		hipLaunchKernelGGL((k_pc<false, false>), dim3(a.nitems), dim3(64), 0, s, a);
	return hipGetLastError();
}

// ---- PP Ewald correction -------------------------------------------------------------------------
// The periodic correction between particle pairs (src/gravity.cpp:491-614): per pair the potential
// and gradient of the tabulated correction (green_ewald_tab<4>: the reference's real-space images,
// Fourier terms, n = 0 term and pi/4 in one table), weighted by the source mass; pairs closer than
// 1e-6 take the reference's constant 2.8372975 and zero force (its zmask).  Sinks and sources are
// laid out as in k_pc: 2^lg lanes per sink, the other lane bits split the tile's sources.
__global__ __launch_bounds__(64) void k_pp_ewald(LeafArgs A) {
	const int4 it = A.items[blockIdx.x];
	const int lane = threadIdx.x;
	const int ns = it.y & PP_COUNT_MASK;
	const int lg = ceil_log2_64(ns);
	const int nsp = 1 << lg;
	const int P = 64 >> lg;
	const int si = lane & (nsp - 1);
	const int ph = lane >> lg;
	const int4 me = A.sinks[it.x + min(si, ns - 1)];
	__shared__ int4 tile[64];
	double dphi = 0.0, dg[3] = {0.0, 0.0, 0.0};
	for (int base = it.z; base < it.w; base += 64) {
		const int n = min(64, it.w - base);
		__syncthreads();
		if (lane < n) tile[lane] = A.srcs[A.list ? A.list[base + lane] : base + lane];
		__syncthreads();
		float phi = 0.f, g[3] = {0.f, 0.f, 0.f};
		for (int k = ph; k < n; k += P) {
			const int4 s = tile[k];
			const float x[3] = {sepu<true>(me.x, s.x) * POS_INVF, sepu<true>(me.y, s.y) * POS_INVF,
					sepu<true>(me.z, s.z) * POS_INVF};
			float D[4];
			green_ewald_tab<4>(D, x, c_ewt);
			const float w = __int_as_float(s.w);
			phi = fmaf(w, D[0], phi);
			g[0] = fmaf(-w, D[1], g[0]);
			g[1] = fmaf(-w, D[2], g[1]);
			g[2] = fmaf(-w, D[3], g[2]);
		}
		dphi += phi;
		dg[0] += g[0];
		dg[1] += g[1];
		dg[2] += g[2];
	}
	for (int off = nsp; off < 64; off <<= 1) {
		dphi += wave_xor(dphi, off);
		dg[0] += wave_xor(dg[0], off);
		dg[1] += wave_xor(dg[1], off);
		dg[2] += wave_xor(dg[2], off);
	}
	if (ph == 0 && si < ns) {
		double* o = A.out + 4 * (size_t) (it.x + si);
		o[0] += dphi;
		o[1] += dg[0];
		o[2] += dg[1];
		o[3] += dg[2];
	}
}

hipError_t launch_pp_ewald(const LeafArgs& a, hipStream_t s) {
	if (a.nitems <= 0) return hipSuccess;
	hipLaunchKernelGGL(k_pp_ewald, dim3(a.nitems), dim3(64), 0, s, a);
	return hipGetLastError();
}

// ---- node kinds ----------------------------------------------------------------------------------
// Segmented FP64 reduction of the lanes' 35 float accumulators by owner slot, through LDS.  The
// entries of one slot are consecutive, so lanes 0..34 (one coefficient each) scan the wave once.
__device__ __forceinline__ void node_segment_store(const float* L, int slot, int cnt, const int* slot_node_unused,
		double* partial) {
	__shared__ float red[35][65];
	__shared__ int sl[64];
	const int lane = threadIdx.x;
#pragma unroll
	for (int c = 0; c < 35; c++) red[c][lane] = L[c];
	sl[lane] = slot;
	__syncthreads();
	// the wave's segments (runs of one slot) from a ballot of their first entries; lanes 0..34 then
	// add one coefficient each over every segment, four FP64 partial sums per segment
	const bool first = lane < cnt && (lane == 0 || sl[lane] != sl[lane - 1]);
	unsigned long long starts = __ballot(first);
	if (lane < 35) {
		while (starts) {
			const int b = __ffsll(starts) - 1;
			starts &= starts - 1;
			const int e = starts ? __ffsll(starts) - 1 : cnt;
			double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
			int j = b;
			for (; j + 4 <= e; j += 4) {
				a0 += double(red[lane][j]);
				a1 += double(red[lane][j + 1]);
				a2 += double(red[lane][j + 2]);
				a3 += double(red[lane][j + 3]);
			}
			for (; j < e; j++) a0 += double(red[lane][j]);
			partial[35 * (size_t) sl[b] + lane] = (a0 + a1) + (a2 + a3);
		}
	}
}

// A wave takes a run of list entries: up to 64 entries of one or more partial slots (one entry per
// lane), or the entries of one slot beyond 64 (a chunk longer than a wave), which the lanes take
// 64 apart, each adding its entries' terms in order before the store.
template <bool EWALD, bool WRAP>
__global__ __launch_bounds__(64) void k_cc(NodeArgs A) {
	const int2 wv = A.waves[blockIdx.x];
	const int lane = threadIdx.x;
	float L[35];
#pragma unroll
	for (int c = 0; c < 35; c++) L[c] = 0.f;
	int slot = -1;
	int j = wv.x + lane;
	if (j < wv.y) {
		slot = A.owner ? A.owner[j] : int(blockIdx.x);  // the plain drop-in: one slot per wave, one sink
		const int4 X = A.sinkpos[A.slot_node ? A.slot_node[slot] : 0];
		// software pipeline (direct kinds): the next entry's multipole and the list index after it
		// are loaded while the current entry is evaluated, into the other of two buffers (the loop
		// is unrolled twice, so no copy between them).  The Ewald kinds load in order: their table
		// row already holds most registers, and the 20 more cost a wave per SIMD (measured: no
		// gain)
		auto index_at = [&](int jj) { return jj < wv.y ? (A.list ? A.list[jj] : jj) : -1; };
		auto load_multi = [&](int idx, float4* m) {
			if (idx >= 0) {
				const float4* src = reinterpret_cast<const float4*>(A.multis + idx);
#pragma unroll
				for (int q = 0; q < 5; q++) m[q] = src[q];
			}
		};
		auto entry = [&](const float4* m) {
			float v[20];
#pragma unroll
			for (int q = 0; q < 5; q++) {
				v[4 * q] = m[q].x;
				v[4 * q + 1] = m[q].y;
				v[4 * q + 2] = m[q].z;
				v[4 * q + 3] = m[q].w;
			}
			const float x[3] = {sepu<EWALD || WRAP>(X.x, __float_as_int(v[17])) * POS_INVF,
					sepu<EWALD || WRAP>(X.y, __float_as_int(v[18])) * POS_INVF,
					sepu<EWALD || WRAP>(X.z, __float_as_int(v[19])) * POS_INVF};
			float D[35];
			if (EWALD)
				green_ewald(D, x);
			else
				green_direct(D, x);
			m2l<!EWALD>(L, v, D);
		};
		float4 ma[5];
		if (EWALD) {
			for (; j < wv.y; j += 64) {
				const float4* src = reinterpret_cast<const float4*>(A.multis + (A.list ? A.list[j] : j));
#pragma unroll
				for (int q = 0; q < 5; q++) ma[q] = src[q];
				entry(ma);
			}
		} else {
			float4 mb[5];
			load_multi(index_at(j), ma);
			int next_idx = index_at(j + 64);
			for (;;) {
				load_multi(next_idx, mb);
				next_idx = index_at(j + 128);
				entry(ma);
				j += 64;
				if (j >= wv.y) break;
				load_multi(next_idx, ma);
				next_idx = index_at(j + 128);
				entry(mb);
				j += 64;
				if (j >= wv.y) break;
			}
		}
	}
	node_segment_store(L, slot, min(wv.y - wv.x, 64), A.slot_node, A.partial);
}

template <bool EWALD, bool WRAP>
__global__ __launch_bounds__(64) void k_cp(NodeArgs A) {
	const int2 wv = A.waves[blockIdx.x];
	const int lane = threadIdx.x;
	float L[35];
#pragma unroll
	for (int c = 0; c < 35; c++) L[c] = 0.f;
	int slot = -1;
	int j = wv.x + lane;
	if (j < wv.y) {
		slot = A.owner ? A.owner[j] : int(blockIdx.x);  // the plain drop-in: one slot per wave, one sink
		const int4 X = A.sinkpos[A.slot_node ? A.slot_node[slot] : 0];
		for (; j < wv.y; j += 64) {
			const int4 s = A.srcs[A.list ? A.list[j] : j];
			const float x[3] = {sepu<EWALD || WRAP>(X.x, s.x) * POS_INVF, sepu<EWALD || WRAP>(X.y, s.y) * POS_INVF,
					sepu<EWALD || WRAP>(X.z, s.z) * POS_INVF};
			float D[35];
			if (EWALD)
				green_ewald(D, x);
			else
				green_direct(D, x);
			p2l(L, __int_as_float(s.w), D);
		}
	}
	node_segment_store(L, slot, min(wv.y - wv.x, 64), A.slot_node, A.partial);
}

hipError_t launch_cc(const NodeArgs& a, bool ewald, hipStream_t s) {
	if (a.nwaves <= 0) return hipSuccess;
	if (ewald)
		hipLaunchKernelGGL((k_cc<true, true>), dim3(a.nwaves), dim3(64), 0, s, a);
	else if (a.wrap)
		hipLaunchKernelGGL((k_cc<false, true>), dim3(a.nwaves), dim3(64), 0, s, a);
	else
		hipLaunchKernelGGL((k_cc<false, false>), dim3(a.nwaves), dim3(64), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_cp(const NodeArgs& a, bool ewald, hipStream_t s) {
	if (a.nwaves <= 0) return hipSuccess;
	if (ewald)
		hipLaunchKernelGGL((k_cp<true, true>), dim3(a.nwaves), dim3(64), 0, s, a);
	else if (a.wrap)
		hipLaunchKernelGGL((k_cp<false, true>), dim3(a.nwaves), dim3(64), 0, s, a);
	else
		hipLaunchKernelGGL((k_cp<false, false>), dim3(a.nwaves), dim3(64), 0, s, a);
	return hipGetLastError();
}

}  // namespace tg
