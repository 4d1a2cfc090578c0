// tg_devmem.cpp -- the caching device allocator of the rebuild path (tg_devmem.hpp).  Host code.
#include "tg_devmem.hpp"

#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace tg {
namespace {

struct Block {
	size_t bytes;
	int device;
};

std::mutex g_mtx;
std::unordered_map<void*, Block> g_live;           // handed out
std::vector<std::multimap<size_t, void*>> g_free;  // per device: size class -> cached block
std::vector<size_t> g_cached;                      // per device: bytes cached
constexpr size_t CACHE_LIMIT = size_t(16) << 30;   // per device; beyond it frees go to hipFree

// the size class of a request: 256-byte granules below 4 KiB, else four classes per power of two
size_t size_class(size_t b) {
	if (b <= 4096) return (b + 255) & ~size_t(255);
	int e = 63 - __builtin_clzll(b);  // 2^e <= b < 2^(e+1)
	const size_t step = size_t(1) << (e - 2);
	return (b + step - 1) / step * step;
}

}  // namespace

hipError_t dev_malloc(void** p, size_t bytes) {
	*p = nullptr;
	int device = 0;
	hipError_t e = hipGetDevice(&device);
	if (e != hipSuccess) return e;
	const size_t c = size_class(bytes == 0 ? 1 : bytes);
	{
		std::lock_guard<std::mutex> lock(g_mtx);
		if (size_t(device) < g_free.size()) {
			auto& fl = g_free[size_t(device)];
			auto it = fl.find(c);
			if (it != fl.end()) {
				*p = it->second;
				fl.erase(it);
				g_cached[size_t(device)] -= c;
				g_live[*p] = Block{c, device};
				return hipSuccess;
			}
		}
	}
	e = hipMalloc(p, c);
	if (e != hipSuccess) {
		// the cache may hold what the device lacks: give it back and try once more
		dev_trim(device);
		e = hipMalloc(p, c);
		if (e != hipSuccess) {
			*p = nullptr;
			return e;
		}
	}
	std::lock_guard<std::mutex> lock(g_mtx);
	g_live[*p] = Block{c, device};
	return hipSuccess;
}

void dev_free(void* p) {
	if (!p) return;
	std::lock_guard<std::mutex> lock(g_mtx);
	auto it = g_live.find(p);
	if (it == g_live.end()) {
		(void) hipFree(p);  // not from this allocator
		return;
	}
	const Block b = it->second;
	g_live.erase(it);
	if (size_t(b.device) >= g_free.size()) {
		g_free.resize(size_t(b.device) + 1);
		g_cached.resize(size_t(b.device) + 1, 0);
	}
	if (g_cached[size_t(b.device)] + b.bytes > CACHE_LIMIT) {
		(void) hipFree(p);
		return;
	}
	g_free[size_t(b.device)].emplace(b.bytes, p);
	g_cached[size_t(b.device)] += b.bytes;
}

size_t dev_trim(int device) {
	std::vector<void*> out;
	size_t freed = 0;
	{
		std::lock_guard<std::mutex> lock(g_mtx);
		if (size_t(device) >= g_free.size()) return 0;
		for (auto& kv : g_free[size_t(device)]) out.push_back(kv.second);
		g_free[size_t(device)].clear();
		freed = g_cached[size_t(device)];
		g_cached[size_t(device)] = 0;
	}
	int cur = 0;
	(void) hipGetDevice(&cur);
	if (cur != device) (void) hipSetDevice(device);
	for (void* p : out) (void) hipFree(p);
	if (cur != device) (void) hipSetDevice(cur);
	return freed;
}

}  // namespace tg
