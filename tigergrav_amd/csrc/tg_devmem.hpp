// tg_devmem.hpp -- a caching device allocator for the rebuild path.
//
// A time step rebuilds the tree, the lists and the work arrays with nearly the same sizes every
// time.  hipMalloc and hipFree of those buffers (the free implicitly waits for the device) would
// cost more than several of the build's kernels, so freed blocks are kept per device, by size
// class (four classes per power of two, so a block is at most 25% larger than asked), and handed
// out again.  A block is only reused by work enqueued after its free on the same stream, or after
// the caller has waited for the stream that last used it (the solver waits for both of its
// streams before a rebuild).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace tg {

// a block of at least `bytes` on the current device (hipMalloc's contract otherwise)
hipError_t dev_malloc(void** p, size_t bytes);
template <class T>
hipError_t dev_malloc(T** p, size_t bytes) {
	return dev_malloc(reinterpret_cast<void**>(p), bytes);
}
// back to the cache (null: nothing)
void dev_free(void* p);
// frees every cached block of `device` (blocks in use stay); returns the bytes freed
size_t dev_trim(int device);

}  // namespace tg
