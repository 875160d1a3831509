// sdfx_runtime.cpp — error state, per-thread streams, caching allocator.
#include "sdfx_runtime.hpp"

#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace sdfx {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }
const std::string& last_error() { return g_last_error; }

int current_device() {
    int dev = 0;
    SDFX_HIP(hipGetDevice(&dev));
    return dev;
}

namespace {

struct ThreadStreams {
    std::unordered_map<int, hipStream_t> by_device;
    ~ThreadStreams() {
        for (auto& kv : by_device) (void)hipStreamDestroy(kv.second);
    }
};
thread_local ThreadStreams t_streams;

// Size classes: powers of two from 256 B, exact sizes above 1 GiB.
size_t size_class(size_t bytes) {
    if (bytes > (size_t(1) << 30)) return (bytes + 0xFFFFF) & ~size_t(0xFFFFF);
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
}

struct Pool {
    std::mutex mu;
    // device -> size class -> free blocks
    std::map<int, std::map<size_t, std::vector<void*>>> free_blocks;
    // pointer -> (device, size class)
    std::unordered_map<void*, std::pair<int, size_t>> live;
};

Pool& pool() {
    static Pool* p = new Pool();  // intentionally leaked: blocks live until process exit
    return *p;
}

}  // namespace

hipStream_t thread_stream() {
    int dev = current_device();
    auto it = t_streams.by_device.find(dev);
    if (it != t_streams.by_device.end()) return it->second;
    hipStream_t s;
    SDFX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    t_streams.by_device[dev] = s;
    return s;
}

void* dev_alloc(size_t bytes) {
    int dev = current_device();
    size_t cls = size_class(bytes ? bytes : 1);
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto& fl = P.free_blocks[dev][cls];
        if (!fl.empty()) {
            void* p = fl.back();
            fl.pop_back();
            P.live[p] = {dev, cls};
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
        // release cached blocks of this device and retry once
        std::vector<void*> drop;
        {
            std::lock_guard<std::mutex> lk(P.mu);
            for (auto& kv : P.free_blocks[dev])
                for (void* q : kv.second) drop.push_back(q);
            P.free_blocks[dev].clear();
        }
        for (void* q : drop) (void)hipFree(q);
        (void)hipGetLastError();
        e = hipMalloc(&p, cls);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(SDFX_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
    }
    std::lock_guard<std::mutex> lk(P.mu);
    P.live[p] = {dev, cls};
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(p);
    if (it == P.live.end()) return;
    P.free_blocks[it->second.first][it->second.second].push_back(p);
    P.live.erase(it);
}

}  // namespace sdfx
