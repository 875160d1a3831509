// sdfx_runtime.cpp — error state, per-thread streams, caching allocator.
#include "sdfx_runtime.hpp"

#include <cstring>
#include <charconv>
#include <cmath>
#include <string>

#include <cstdio>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace sdfx {

std::string rust_debug_f32(float x) {
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
    const bool neg = std::signbit(x);
    const float a = std::fabs(x);
    if (a == 0.0f) return neg ? "-0.0" : "0.0";
    // shortest round-trip digits d1 d2 ... dk and decimal exponent e: a = d1.d2...dk * 10^e
    char buf[48];
    const auto r = std::to_chars(buf, buf + sizeof buf, a, std::chars_format::scientific);
    const std::string sci(buf, r.ptr);
    const size_t epos = sci.find('e');
    std::string digits;
    for (size_t i = 0; i < epos; ++i)
        if (sci[i] != '.') digits += sci[i];
    const int e = std::stoi(sci.substr(epos + 1));
    std::string out = neg ? "-" : "";
    if (a < 1e-4f || a >= 1e16f) {  // float_to_exponential_common_shortest
        out += digits[0];
        if (digits.size() > 1) out += "." + digits.substr(1);
        return out + "e" + std::to_string(e);
    }
    if (e >= 0) {  // float_to_decimal_common_shortest with at least one fractional digit
        std::string ip = digits.substr(0, std::min<size_t>(digits.size(), (size_t)e + 1));
        ip.append((size_t)e + 1 > digits.size() ? (size_t)e + 1 - digits.size() : 0, '0');
        std::string fp = (size_t)e + 1 < digits.size() ? digits.substr((size_t)e + 1) : "0";
        return out + ip + "." + fp;
    }
    return out + "0." + std::string((size_t)(-e - 1), '0') + digits;
}

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

namespace {
struct PinnedPool {
    std::mutex mu;
    std::map<size_t, std::vector<void*>> free_blocks;  // size class -> blocks
    std::unordered_map<void*, size_t> live;
};
PinnedPool& pinned_pool() {
    static PinnedPool* p = new PinnedPool();  // intentionally leaked: blocks live until process exit
    return *p;
}
}  // namespace

void* pinned_alloc(size_t bytes) {
    const size_t cls = size_class(bytes ? bytes : 1);
    PinnedPool& P = pinned_pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto& fl = P.free_blocks[cls];
        if (!fl.empty()) {
            void* p = fl.back();
            fl.pop_back();
            P.live[p] = cls;
            return p;
        }
    }
    void* p = nullptr;
    SDFX_HIP(hipHostMalloc(&p, cls, hipHostMallocDefault));
    std::lock_guard<std::mutex> lk(P.mu);
    P.live[p] = cls;
    return p;
}

void pinned_free(void* p) {
    if (!p) return;
    PinnedPool& P = pinned_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(p);
    if (it == P.live.end()) return;
    P.free_blocks[it->second].push_back(p);
    P.live.erase(it);
}

void Arena::upload(void* d_dst, const void* h_src, size_t bytes) {
    if (!bytes) return;
    if (region) throw Error(SDFX_ERR_INVALID_ARGUMENT, "host uploads need pinned staging memory, which a call on a fixed region cannot allocate");
    void* h = pinned_alloc(bytes);
    staged.push_back(h);
    std::memcpy(h, h_src, bytes);
    SDFX_HIP(hipMemcpyAsync(d_dst, h, bytes, hipMemcpyHostToDevice, stream));
}

// ---------------------------------------------------------------- tuning

}  // namespace sdfx
#include "sdfx_internal.hpp"
namespace sdfx {

Tuning& tuning() {
    // 2^26 pairs: about 10k queries against 6k triangles, where the exhaustive
    // kernel finishes well before an LBVH build and traversal would.
    // 1,536 batched propagation blocks: one resident generation of the
    // item-major kernel (6 blocks per CU, bound by its 24.8 KB of LDS)
    static Tuning t{int64_t(1) << 26, 4096, 0, 1536, 1 << 20};
    return t;
}

// ---------------------------------------------------------------- profiling

namespace {
struct ProfRecord {
    std::string name;
    hipEvent_t a, b;
};
struct ProfState {
    std::mutex mu;
    bool enabled = false;
    bool counters_on = true;  // work counters (distance evaluations, node visits)
    std::string select;       // only kernels whose name starts with this get events ("" = all)
    std::vector<ProfRecord> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, std::pair<double, uint64_t>> totals;  // name -> (ms, launches)
    std::map<std::string, uint64_t> counters;
};
ProfState& prof() {
    static ProfState* p = new ProfState();
    return *p;
}
hipEvent_t take_event(ProfState& P) {
    if (!P.free_events.empty()) {
        hipEvent_t e = P.free_events.back();
        P.free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    // timing-only events: no system-scope fence (cache writeback and
    // invalidation) when recorded, so the kernels between them run undisturbed
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}
void resolve(ProfState& P) {
    for (auto& r : P.pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto& t = P.totals[r.name];
            t.first += ms;
            t.second += 1;
        }
        P.free_events.push_back(r.a);
        P.free_events.push_back(r.b);
    }
    P.pending.clear();
}
}  // namespace

bool profile_enabled() {
    ProfState& P = prof();
    return P.enabled && P.counters_on;
}

void profile_count(const char* name, uint64_t value) {
    ProfState& P = prof();
    if (!P.enabled) return;
    std::lock_guard<std::mutex> lk(P.mu);
    P.counters[name] += value;
}

ProfPair prof_begin(const char* name) {
    ProfPair p;
    ProfState& P = prof();
    if (!P.enabled) return p;
    std::lock_guard<std::mutex> lk(P.mu);
    if (!P.select.empty() && std::strncmp(name, P.select.c_str(), P.select.size()) != 0) return p;
    p.a = take_event(P);
    p.b = take_event(P);
    if (!p.a || !p.b) {
        if (p.a) P.free_events.push_back(p.a);
        if (p.b) P.free_events.push_back(p.b);
        p.a = p.b = nullptr;
        return p;
    }
    p.name = name;
    return p;
}

void prof_end(const ProfPair& p) {
    if (!p.a) return;
    ProfState& P = prof();
    std::lock_guard<std::mutex> lk(P.mu);
    P.pending.push_back({p.name, p.a, p.b});
}

}  // namespace sdfx

extern "C" {

void sdfx_profile_enable(int on) {
    auto& P = sdfx::prof();
    std::lock_guard<std::mutex> lk(P.mu);
    sdfx::resolve(P);
    P.totals.clear();
    P.counters.clear();
    P.enabled = on != 0;
    P.counters_on = true;
    P.select.clear();
}

void sdfx_profile_select(const char* kernel_prefix, int counters) {
    auto& P = sdfx::prof();
    std::lock_guard<std::mutex> lk(P.mu);
    P.select = kernel_prefix ? kernel_prefix : "";
    P.counters_on = counters != 0;
}

namespace {
int64_t* tuning_slot(const char* key) {
    if (!key) return nullptr;
    const std::string k(key);
    if (k == "exhaustive_max_pairs") return &sdfx::tuning().exhaustive_max_pairs;
    if (k == "grid_prop_blocks") return &sdfx::tuning().grid_prop_blocks;
    if (k == "lbvh_visit_budget") return &sdfx::tuning().lbvh_visit_budget;
    if (k == "grid_batch_prop_blocks") return &sdfx::tuning().grid_batch_prop_blocks;
    if (k == "grid_item_major_min") return &sdfx::tuning().grid_item_major_min;
    return nullptr;
}
}  // namespace

int sdfx_set_tuning(const char* key, int64_t value) {
    int64_t* slot = tuning_slot(key);
    if (!slot) return -1;
    if ((std::string(key) == "grid_prop_blocks" || std::string(key) == "grid_batch_prop_blocks") &&
        (value < 8 || value > 65536))
        return -1;
    if (std::string(key) == "lbvh_visit_budget" && value != 0 && (value < 16 || value > INT32_MAX)) return -1;
    if (std::string(key) == "grid_item_major_min" && value < 0) return -1;
    *slot = value;
    return 0;
}

int64_t sdfx_get_tuning(const char* key) {
    int64_t* slot = tuning_slot(key);
    return slot ? *slot : -1;
}

uint64_t sdfx_profile_counter(const char* name) {
    auto& P = sdfx::prof();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.counters.find(name ? name : "");
    return it == P.counters.end() ? 0 : it->second;
}

int sdfx_profile_read(int max_entries, char* names, double* total_ms, uint64_t* launches) {
    auto& P = sdfx::prof();
    std::lock_guard<std::mutex> lk(P.mu);
    sdfx::resolve(P);
    int i = 0;
    for (auto& kv : P.totals) {
        if (i >= max_entries) break;
        std::snprintf(names + 64 * i, 64, "%s", kv.first.c_str());
        total_ms[i] = kv.second.first;
        launches[i] = kv.second.second;
        ++i;
    }
    return (int)P.totals.size();
}

}  // extern "C"
