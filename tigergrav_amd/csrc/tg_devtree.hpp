// tg_devtree.hpp -- the kd-tree and its multipoles built on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "tg_tree.hpp"

namespace tg {

// a node's split record and its multipole data as the device build keeps them
struct DNode {
	int32_t pbegin, pend, c0, c1, parent, level;
	uint64_t boxid;
};

struct DMulti {
	float m[17];
	float r;
	int32_t x[3];
	int32_t pad;
	uint64_t N;
	double rmin[3], rmax[3];
};

// A device-built tree left in device memory (tree order), owned here: what device_build_tree
// would copy into a host Tree.
struct DeviceTree {
	int32_t* pos = nullptr;   // 3 per particle
	float* mass = nullptr;
	int64_t* order = nullptr;  // input index of each tree-order particle
	DNode* nodes = nullptr;
	DMulti* multi = nullptr;
	int64_t n = 0, nn = 0;
	std::vector<int> level_begin;  // first node of each level, then nn
	void release();
};

// Tree::build followed by Tree::compute_multipoles, on the device: the same particle order, the
// same nodes in the same numbering, the same centres, moments, radii and particle boxes, bit for
// bit (src/tree.cpp:122-296, src/part_vect.cpp:542-635,895-968).  The result is written into T
// (host) as the host build leaves it or, with keep, left on the device in *keep (released first;
// T then gets only its parameters, particle mass and an empty node array).  t, if given, receives
// the seconds of the split passes (with the upload of the particles) and of the multipole passes
// (with the copy back to the host).  x and masses are device pointers on `device` when on_device
// is set, host pointers otherwise.
hipError_t device_build_tree(Tree& T, const int32_t* x, const float* masses, int64_t n, const TreeParams& p,
		double particle_mass, int device, hipStream_t s, std::string& err, double* t = nullptr, bool on_device = false,
		DeviceTree* keep = nullptr);

// the host tree (positions, masses, order, nodes, leaves, depth) of a kept device tree, as
// device_build_tree without keep writes it
hipError_t device_tree_to_host(const DeviceTree& D, Tree& T, hipStream_t s, std::string& err);

}  // namespace tg
