// tg_devtree.hpp -- the kd-tree and its multipoles built on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "tg_tree.hpp"

namespace tg {

// Tree::build followed by Tree::compute_multipoles, on the device: the same particle order, the
// same nodes in the same numbering, the same centres, moments, radii and particle boxes, bit for
// bit (src/tree.cpp:122-296, src/part_vect.cpp:542-635,895-968).  The result is written into T
// (host) as the host build leaves it.  t, if given, receives the seconds of the split passes
// (with the upload of the particles) and of the multipole passes (with the copy back to the host).
// x and masses are device pointers on `device` when on_device is set, host pointers otherwise.
hipError_t device_build_tree(Tree& T, const int32_t* x, const float* masses, int64_t n, const TreeParams& p,
		double particle_mass, int device, hipStream_t s, std::string& err, double* t = nullptr, bool on_device = false);

}  // namespace tg
