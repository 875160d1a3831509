// sdfx_runtime.hpp — host-side runtime shared by the libsdfx translation units:
// status/error plumbing, per-thread streams and a caching device allocator.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sdfx.h"

namespace sdfx {

// Thrown inside the library, converted to a status at the C boundary.
struct Error : std::runtime_error {
    sdfx_status status;
    Error(sdfx_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

// Rust's `{:?}` of an f32 (core::fmt::float, float_to_general_debug): the
// shortest digits that round-trip, "1.0"-style decimals for 1e-4 <= |x| < 1e16
// and zero, "1.5e-7"-style exponentials otherwise, "NaN", "inf", "-inf", "-0.0".
std::string rust_debug_f32(float x);

void set_last_error(const std::string& msg);
void clear_last_error();
const std::string& last_error();

#define SDFX_HIP(expr)                                                                                 \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::sdfx::Error(e_ == hipErrorOutOfMemory ? SDFX_ERR_OUT_OF_MEMORY : SDFX_ERR_DEVICE,  \
                                std::string(#expr) + ": " + hipGetErrorString(e_));                    \
    } while (0)

// Stream used by host-pointer entry points: one per calling thread, created
// lazily with the default priority.
hipStream_t thread_stream();

// Pinned host blocks from a process-wide cache (hipHostMalloc, kept until exit).
void* pinned_alloc(size_t bytes);
void pinned_free(void* p);

// Caching device allocator. Blocks are returned to a per-device free list and
// reused by later calls; memory is released only at process exit.
void* dev_alloc(size_t bytes);
void dev_free(void* p);

// RAII owner of one device block.
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { reset(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n) {
        o.ptr = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            n = o.n;
            o.ptr = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void reset(size_t count) {
        release();
        n = count;
        ptr = count ? static_cast<T*>(dev_alloc(count * sizeof(T))) : nullptr;
    }
    void release() {
        if (ptr) dev_free(ptr);
        ptr = nullptr;
        n = 0;
    }
    T* get() const { return ptr; }
    size_t bytes() const { return n * sizeof(T); }
};

// Temporary device memory of one call. Blocks go back to the pool only after
// the stream is synchronised (in the destructor), so an exception thrown while
// kernels are in flight cannot hand their memory to another call. Declare the
// Arena after every DevBuf a call owns, so it is destroyed (and syncs) first.
//
// An Arena built over a fixed region (the scratch of a no-allocation entry
// point) hands out consecutive 256-byte-aligned pieces of it instead, never
// calls an allocator, and fails a request beyond its end; upload(), which
// needs pinned staging memory, is unavailable there.
struct Arena {
    hipStream_t stream;
    std::vector<void*> blocks;
    std::vector<void*> staged;  // pinned host blocks of upload()
    char* region = nullptr;     // fixed region, or null
    size_t region_bytes = 0, region_used = 0;
    explicit Arena(hipStream_t s) : stream(s) {}
    Arena(hipStream_t s, void* base, size_t bytes) : stream(s), region(static_cast<char*>(base)), region_bytes(bytes) {}
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    template <typename T>
    T* alloc(size_t count) {
        const size_t bytes = (count ? count : 1) * sizeof(T);
        if (region) {
            const size_t off = (region_used + 255) & ~size_t(255);
            if (off > region_bytes || bytes > region_bytes - off)
                throw Error(SDFX_ERR_INVALID_ARGUMENT, "the call needs more scratch than the fixed region holds");
            region_used = off + bytes;
            return reinterpret_cast<T*>(region + off);
        }
        void* p = dev_alloc(bytes);
        blocks.push_back(p);
        return static_cast<T*>(p);
    }
    // Copies `bytes` of host memory to d_dst on the stream without blocking the
    // host: the data goes through a pinned block the Arena keeps until it has
    // synchronised.
    void upload(void* d_dst, const void* h_src, size_t bytes);
    // Waits for the stream before returning its blocks. An Arena over a fixed
    // region owns nothing to return, so it does not wait: its caller decides
    // when the region's work is done (and under stream capture no wait is
    // allowed).
    ~Arena() {
        if (region && blocks.empty() && staged.empty()) return;
        (void)hipStreamSynchronize(stream);
        for (void* p : blocks) dev_free(p);
        for (void* p : staged) pinned_free(p);
    }
};

// Device-side error flags of one call, read back once at the end.
struct DevFlags {
    uint32_t non_finite_query;
    uint32_t index_out_of_range;
    uint32_t nan_distance;
    uint32_t internal;  // a device-side capacity or consistency check failed
};

int current_device();

// Kernel timing for measurement tools (off by default). While enabled, each
// launch made through SDFX_LAUNCH goes through hipExtLaunchKernelGGL with a
// start/stop event pair that the runtime ties to the kernel's own dispatch, so
// the pair brackets exactly that kernel's execution on its stream;
// sdfx_profile_read() resolves the pairs into per-name totals.
bool profile_enabled();
// Adds to a named work counter (e.g. distance evaluations) while profiling.
void profile_count(const char* name, uint64_t value);
struct ProfPair {
    const char* name = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
};
ProfPair prof_begin(const char* name);  // {nullptr events} when profiling is off
void prof_end(const ProfPair& p);

// hipLaunchKernelGGL, or its timed form while kernel timing is enabled.
#define SDFX_LAUNCH(kernel, grid, block, shmem, stream, ...)                                                 \
    do {                                                                                                     \
        const ::sdfx::ProfPair sdfx_pp_ = ::sdfx::prof_begin(#kernel);                                       \
        if (sdfx_pp_.a)                                                                                      \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, sdfx_pp_.a, sdfx_pp_.b, 0, __VA_ARGS__); \
        else                                                                                                 \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                             \
        ::sdfx::prof_end(sdfx_pp_);                                                                          \
    } while (0)

// Launch helper: blocks for n items with `threads` per block.
inline unsigned blocks_for(uint64_t n, unsigned threads) {
    uint64_t b = (n + threads - 1) / threads;
    return (unsigned)(b ? b : 1);
}

}  // namespace sdfx
