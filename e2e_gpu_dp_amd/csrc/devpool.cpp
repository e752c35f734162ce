// devpool.cpp — reuse of large device allocations across models.
//
// A build and a solve allocate a few large buffers (the row staging, the CSR entries,
// the compressed stream, the value buffers) and free them when the model closes. Mapping
// and unmapping gigabytes costs milliseconds per call, so blocks of at least kMinPooled
// bytes that are freed go to a per-device cache instead and serve later requests of a
// similar size. A block is cached only after the device is idle (as hipFree would wait),
// so no kernel still in flight can see its next owner's data.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace e2e {

namespace {

constexpr size_t kMinPooled = size_t(1) << 20;   // smaller blocks go straight to hipMalloc
constexpr size_t kGranule = size_t(2) << 20;     // pooled sizes are rounded up to 2 MiB

struct Block {
    void* p;
    size_t bytes;
    int dev;
};

struct Pool {
    std::mutex mu;
    std::unordered_map<void*, Block> out;  // pooled blocks in use
    std::vector<Block> cached;             // idle blocks
    size_t cached_bytes = 0;

    static size_t limit() {  // E2E_POOL_BYTES: most bytes kept idle (0 turns the pool off)
        static const size_t cap = [] {
            const char* e = std::getenv("E2E_POOL_BYTES");
            return e ? (size_t)std::strtoull(e, nullptr, 10) : (size_t(32) << 30);
        }();
        return cap;
    }
    // frees idle blocks of device dev (all devices if dev < 0) until at most keep bytes stay
    // (oldest first)
    void trim(int dev, size_t keep) {
        for (size_t i = 0; i < cached.size() && cached_bytes > keep;) {
            if (dev >= 0 && cached[i].dev != dev) {
                i++;
                continue;
            }
            int cur = 0;
            (void)hipGetDevice(&cur);
            if (cached[i].dev != cur) (void)hipSetDevice(cached[i].dev);
            (void)hipFree(cached[i].p);
            if (cached[i].dev != cur) (void)hipSetDevice(cur);
            cached_bytes -= cached[i].bytes;
            cached.erase(cached.begin() + (std::ptrdiff_t)i);
        }
    }
};

Pool& pool() {
    static Pool* P = new Pool();  // never destroyed: frees at exit would race the runtime's teardown
    return *P;
}

}  // namespace

hipError_t pool_malloc(void** p, size_t bytes) {
    *p = nullptr;
    Pool& P = pool();
    if (bytes < kMinPooled || P.limit() == 0) {  // not pooled; on failure the idle blocks go back first
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess && P.limit() != 0) {
            (void)hipGetLastError();
            int dev = 0;
            if (hipGetDevice(&dev) == hipSuccess) {
                std::lock_guard<std::mutex> lk(P.mu);
                P.trim(dev, 0);
            }
            e = hipMalloc(p, bytes);
        }
        if (e != hipSuccess) (void)hipGetLastError();
        return e;
    }
    const size_t want = (bytes + kGranule - 1) / kGranule * kGranule;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(P.mu);
        // the smallest idle block of this device that holds want, wasting at most an eighth
        size_t best = P.cached.size();
        for (size_t i = 0; i < P.cached.size(); i++) {
            const Block& b = P.cached[i];
            if (b.dev == dev && b.bytes >= want && b.bytes - want <= want / 8 &&
                (best == P.cached.size() || b.bytes < P.cached[best].bytes))
                best = i;
        }
        if (best < P.cached.size()) {
            const Block b = P.cached[best];
            P.cached.erase(P.cached.begin() + (std::ptrdiff_t)best);
            P.cached_bytes -= b.bytes;
            P.out[b.p] = b;
            *p = b.p;
            return hipSuccess;
        }
    }
    e = hipMalloc(p, want);
    if (e != hipSuccess) {  // idle blocks of this device go back to the runtime, then once more
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(P.mu);
            P.trim(dev, 0);
        }
        e = hipMalloc(p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return e;
        }
    }
    std::lock_guard<std::mutex> lk(P.mu);
    P.out[*p] = Block{*p, want, dev};
    return hipSuccess;
}

void pool_free(void* p) {
    if (!p) return;
    Pool& P = pool();
    Block b{};
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.out.find(p);
        if (it == P.out.end()) {  // not a pooled block
            (void)hipFree(p);
            return;
        }
        b = it->second;
        P.out.erase(it);
    }
    // the block's last kernels may still run on any stream of its device: wait for the
    // device, as hipFree does, before the block can reach another owner
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (b.dev != cur) (void)hipSetDevice(b.dev);
    (void)hipDeviceSynchronize();
    if (b.dev != cur) (void)hipSetDevice(cur);
    std::lock_guard<std::mutex> lk(P.mu);
    if (b.bytes > P.limit()) {
        (void)hipFree(b.p);
        return;
    }
    P.cached.push_back(b);
    P.cached_bytes += b.bytes;
    P.trim(-1, P.limit());  // the oldest blocks leave first when the idle bytes pass the limit
}

void pool_release_all() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    P.trim(-1, 0);
}

size_t pool_cached_bytes() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    return P.cached_bytes;
}

}  // namespace e2e
