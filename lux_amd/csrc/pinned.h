// Pinned host staging for the served forms (server.hip, dloop.hip), pooled by size: hipHostMalloc is slow, so
// requests of one size class reuse blocks.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

namespace lux {

constexpr size_t kPinnedPoolCap = 4ull << 30;  // pinned staging kept for reuse

struct PinnedPool {
    unsigned flags = hipHostMallocDefault;  // hipHostMallocCoherent where the device reads and writes them itself
    std::mutex mu;
    std::multimap<size_t, void*> free_blocks;
    size_t held = 0;

    void* get(size_t n, size_t* got)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = free_blocks.lower_bound(n);
            if (it != free_blocks.end() && it->first <= 2 * n + (1 << 20)) {
                void* p = it->second;
                *got = it->first;
                held -= it->first;
                free_blocks.erase(it);
                return p;
            }
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, n, flags) != hipSuccess) return nullptr;
        *got = n;
        return p;
    }
    void put(void* p, size_t n)
    {
        if (!p) return;
        std::unique_lock<std::mutex> lk(mu);
        if (held + n > kPinnedPoolCap) {
            lk.unlock();
            (void)hipHostFree(p);
            return;
        }
        free_blocks.emplace(n, p);
        held += n;
    }
    void clear()
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& kv : free_blocks) (void)hipHostFree(kv.second);
        free_blocks.clear();
        held = 0;
    }
};

}  // namespace lux
