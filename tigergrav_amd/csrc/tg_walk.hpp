// tg_walk.hpp -- the dual-tree walk (interaction lists) on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "tg_tree.hpp"

namespace tg {

// a node as the device walk reads it
struct WalkNode {
	int32_t pbegin, pend, c0, c1;
	int32_t x[3];
	float r;
	float rmax;  // span of the node's particle box (the sink rmax of the Ewald self test)
};

// fills out[i] for every node of T (host, OpenMP; tg_upload.cpp)
void walk_nodes(const Tree& T, WalkNode* out);

// The walk's lists left in device memory: node-order CSR per kind, kinds in the order cc_d, cp_d,
// cc_e, cp_e, pc_d, pp_d, pc_e, pp_e (node kinds hold node ids or [begin, end) particle ranges as
// the host Lists do).  The table of nodes the walk read stays too.  Every array is owned here.
struct DeviceLists {
	WalkNode* nodes = nullptr;   // nn nodes
	int64_t* ptr_all = nullptr;  // 8 x (nn + 1) pointers, kind-major
	int32_t* val[8] = {};        // values per kind
	int64_t len[8] = {};         // values per kind (ints)
	int64_t nn = 0;
	const int64_t* ptr(int k) const { return ptr_all + int64_t(k) * (nn + 1); }
	void release();
};

// Tree::walk on the device: fills T.lists and T.census with exactly the lists the host walk
// builds (same entries, same order).  Returns hipSuccess or the first HIP error; err names it.
// t, if given, receives the seconds spent in set-up, the walk steps, list assembly with the copy
// to the host, and the census.  With keep, nothing is copied to the host: the lists stay on the
// device in *keep (released first), T.lists is left empty and T.census zero.
// With nodes (a device table of nn nodes, which device_walk then owns and frees or hands to keep),
// T supplies only its parameters: the tree itself may be on the device only.
// With hint, the walk's working buffers start at the sizes *hint holds (those the caller's last
// walk reached, which this walk writes back), so a rebuild of similar particles grows none of them:
// a grown output pool copies what it holds, and a grown per-step buffer waits for the stream.
struct WalkHint {
	int64_t out[8] = {};    // entries of each output kind
	int64_t lists = 0;      // the longest checklist set of a step
	int64_t counts = 0;     // per-step count buffer
	int64_t codes = 0;      // per-step category bytes
};
hipError_t device_walk(Tree& T, int device, hipStream_t s, std::string& err, double* t = nullptr, DeviceLists* keep = nullptr,
		WalkNode* nodes = nullptr, int64_t nn = 0, WalkHint* hint = nullptr);

struct DeviceTree;
// the walk's node table of a device-built tree, made on the device (*out: nn nodes, device memory
// the caller frees): the centres, radii and children, and each node's sink span from a bottom-up
// union of its non-empty leaves' particle boxes, as Tree::sink_spans forms it
hipError_t walk_nodes_device(const DeviceTree& D, WalkNode** out, hipStream_t s, std::string& err);

}  // namespace tg
