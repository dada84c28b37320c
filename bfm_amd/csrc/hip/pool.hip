// Cache of device blocks for the multigrid set-up. The hierarchy's buffers
// and the topology build's temporaries come from per-device free lists and
// go back to them, instead of hipMalloc / hipFree: a hipMalloc of a block of
// several MB costs about 0.1 ms, and hipFree waits for the whole device, so a
// fresh system's set-up that allocates and frees a few hundred blocks would
// spend several milliseconds there. Sizes round up to classes a quarter of a
// power of two apart (at most 25% more than asked). A block is handed back
// only once every stream that used it has drained (pool_free's contract), so
// the next owner may use it on any stream. Cached memory beyond kKeep bytes
// per device goes back to HIP; when hipMalloc runs out of memory, the
// device's cached blocks are released and the allocation tried once more.

#include <algorithm>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace bfmx {

namespace {

constexpr size_t kKeep = size_t(8) << 30;   // cached bytes kept per device
constexpr size_t kMinClass = 256;

struct Block {
	int dev;
	size_t cls;
};

struct Pool {
	std::mutex m;
	std::unordered_map<void*, Block> live;                   // every block the pool handed out
	std::unordered_map<int, std::map<size_t, std::vector<void*>>> free;   // per device, per class
	std::unordered_map<int, size_t> cached;                  // per device, bytes in the free lists
};

// Never destroyed: blocks may still be returned while the process exits,
// after the HIP runtime is gone.
Pool& pool() {
	static Pool* p = new Pool;
	return *p;
}

size_t size_class(size_t b) {
	if (b <= kMinClass) {
		return kMinClass;
	}
	int e = 63 - __builtin_clzll(b - 1);   // 2^e < b <= 2^(e + 1)
	size_t const q = size_t(1) << (e - 1);  // a quarter of 2^(e + 1)
	return (b + q - 1) / q * q;
}

// Releases the cached blocks of device dev (lock held; dev current).
void release_cached(Pool& P, int dev) {
	for (auto& [cls, v] : P.free[dev]) {
		for (void* q : v) {
			(void) hipFree(q);
			P.live.erase(q);
		}
		v.clear();
	}
	P.cached[dev] = 0;
}

}  // namespace

int pool_alloc_bytes(void** out, size_t bytes) {
	*out = nullptr;
	int dev = 0;
	BFMX_HIP(hipGetDevice(&dev));
	size_t const cls = size_class(bytes);
	Pool& P = pool();
	std::lock_guard<std::mutex> lock(P.m);
	auto& lists = P.free[dev];
	auto it = lists.find(cls);
	if (it != lists.end() && !it->second.empty()) {
		*out = it->second.back();
		it->second.pop_back();
		P.cached[dev] -= cls;
		return BFMX_OK;
	}
	void* q = nullptr;
	hipError_t e = hipMalloc(&q, cls);
	if (e == hipErrorOutOfMemory && P.cached[dev] > 0) {
		(void) hipGetLastError();
		release_cached(P, dev);
		e = hipMalloc(&q, cls);
	}
	if (e != hipSuccess) {
		set_error("hipMalloc of %zu bytes failed: %s", cls, hipGetErrorString(e));
		return e == hipErrorOutOfMemory ? BFMX_E_NOMEM : BFMX_E_HIP;
	}
	P.live[q] = Block{dev, cls};
	*out = q;
	return BFMX_OK;
}

void pool_free_bytes(void* q) {
	if (!q) {
		return;
	}
	Pool& P = pool();
	std::lock_guard<std::mutex> lock(P.m);
	auto it = P.live.find(q);
	if (it == P.live.end()) {
		(void) hipFree(q);   // not one of the pool's blocks
		return;
	}
	Block const b = it->second;
	if (P.cached[b.dev] + b.cls > kKeep) {
		P.live.erase(it);
		(void) hipFree(q);
		return;
	}
	P.free[b.dev][b.cls].push_back(q);
	P.cached[b.dev] += b.cls;
}

// Small pinned host blocks (a system's scalar mirror) cached per size: a
// hipHostMalloc costs tens of microseconds, a fresh system's first solve
// needs one.
namespace {
struct PinnedCache {
	std::mutex m;
	std::map<size_t, std::vector<void*>> free;
};
PinnedCache& pinned_cache() {
	static PinnedCache* c = new PinnedCache;   // never destroyed, as the pool
	return *c;
}
}  // namespace

int pinned_small_alloc(void** out, size_t bytes) {
	*out = nullptr;
	{
		PinnedCache& C = pinned_cache();
		std::lock_guard<std::mutex> lock(C.m);
		auto it = C.free.find(bytes);
		if (it != C.free.end() && !it->second.empty()) {
			*out = it->second.back();
			it->second.pop_back();
			return BFMX_OK;
		}
	}
	hipError_t const e = hipHostMalloc(out, bytes, hipHostMallocDefault);
	if (e != hipSuccess) {
		*out = nullptr;
		set_error("hipHostMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
		return BFMX_E_NOMEM;
	}
	return BFMX_OK;
}

void pinned_small_free(void* p, size_t bytes) {
	if (!p) {
		return;
	}
	PinnedCache& C = pinned_cache();
	std::lock_guard<std::mutex> lock(C.m);
	C.free[bytes].push_back(p);
}

int pinned_staging(void** out, size_t bytes, int slot) {
	struct Staging {
		void* p = nullptr;
		size_t cap = 0;
	};
	// not freed at thread exit: the HIP runtime may already be gone then
	thread_local Staging* slots = new Staging[kStagingSlots];
	if (slot < 0 || slot >= kStagingSlots) {
		set_error("pinned staging slot %d out of range", slot);
		return BFMX_E_ARG;
	}
	Staging* const st = &slots[slot];
	if (st->cap < bytes) {
		if (st->p) {
			(void) hipHostFree(st->p);
			st->p = nullptr;
			st->cap = 0;
		}
		size_t const cap = std::max(bytes, size_t(4) << 10);
		hipError_t const e = hipHostMalloc(&st->p, cap, hipHostMallocDefault);
		if (e != hipSuccess) {
			st->p = nullptr;
			set_error("hipHostMalloc of %zu bytes failed: %s", cap, hipGetErrorString(e));
			return BFMX_E_NOMEM;
		}
		st->cap = cap;
	}
	*out = st->p;
	return BFMX_OK;
}

}  // namespace bfmx
