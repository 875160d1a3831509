// sdfx_mesh.hip — topology decoding, triangle gathering, query validation and
// a u64 prefix scan used by the grid pipeline.
#include "sdfx_internal.hpp"

namespace sdfx {

// lib.rs:193-215 get_triangles: TriangleList asserts len % 3 == 0 (with
// indices) or vertices.len() % 3 == 0 (without); TriangleStrip yields
// len - 2 sliding windows (none for len < 3).
sdfx_status topology_triangle_count(uint64_t nv, bool has_idx, uint64_t ni, sdfx_topology topo, uint64_t* nt) {
    *nt = 0;
    if (topo == SDFX_TRIANGLE_LIST) {
        if (has_idx) {
            if (ni % 3 != 0) return SDFX_ERR_INDEX_COUNT;
            *nt = ni / 3;
        } else {
            if (nv % 3 != 0) return SDFX_ERR_VERTEX_COUNT;
            *nt = nv / 3;
        }
        return SDFX_OK;
    }
    if (topo == SDFX_TRIANGLE_STRIP) {
        uint64_t n = has_idx ? ni : nv;
        *nt = n >= 3 ? n - 2 : 0;
        return SDFX_OK;
    }
    return SDFX_ERR_INVALID_ARGUMENT;
}

// The triangle t of one mesh: its vertex indices (clamped, with the
// out-of-range flag raised, so no read leaves the vertex buffer) and its
// TriRec.
__device__ __forceinline__ void gather_one(const float* __restrict__ v, uint64_t nv, const uint32_t* __restrict__ idx,
                                           int topo, uint64_t t, TriRec* __restrict__ out, DevFlags* flags) {
    uint64_t base = topo == SDFX_TRIANGLE_LIST ? 3 * t : t;
    uint64_t vi[3];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint64_t x = idx ? (uint64_t)idx[base + k] : base + k;
        if (x >= nv) {
            bad = true;
            x = nv ? nv - 1 : 0;
        }
        vi[k] = x;
    }
    if (bad) atomicOr(&flags->index_out_of_range, 1u);
    if (nv == 0) return;  // only reachable with bad indices; the call fails
    const float* a = v + 3 * vi[0];
    const float* b = v + 3 * vi[1];
    const float* c = v + 3 * vi[2];
    TriRec r;
    make_trirec(a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], (uint32_t)t, r);
    out[t] = r;
}

__global__ void __launch_bounds__(256) k_gather(const float* __restrict__ v, uint64_t nv,
                                                const uint32_t* __restrict__ idx, int topo, uint64_t nt,
                                                TriRec* __restrict__ out, DevFlags* flags) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt) gather_one(v, nv, idx, topo, t, out, flags);
}

// Many meshes in one launch: blockIdx.y is the mesh.
__global__ void __launch_bounds__(256) k_gather_batch(const GatherJob* __restrict__ jobs) {
    const GatherJob& J = jobs[blockIdx.y];
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < J.nt) gather_one(J.v, J.nv, J.idx, (int)J.topo, t, J.out, J.flags);
}

void gather_triangles(const float* d_v, uint64_t nv, const uint32_t* d_idx, sdfx_topology topo, uint64_t nt,
                      TriRec* d_tris, DevFlags* d_flags, hipStream_t s) {
    if (!nt) return;
    SDFX_LAUNCH(k_gather, dim3(blocks_for(nt, 256)), dim3(256), 0, s, d_v, nv, d_idx, (int)topo, nt, d_tris,
                       d_flags);
    SDFX_HIP(hipGetLastError());
}

void gather_triangles_batch(const std::vector<GatherJob>& jobs, Arena& A) {
    std::vector<GatherJob> live;
    uint64_t max_nt = 0;
    for (const GatherJob& J : jobs)
        if (J.nt) {
            live.push_back(J);
            max_nt = std::max(max_nt, J.nt);
        }
    for (size_t c0 = 0; c0 < live.size(); c0 += 65535) {  // grid.y limit
        const size_t c1 = std::min(live.size(), c0 + 65535);
        GatherJob* d_jobs = A.alloc<GatherJob>(c1 - c0);
        SDFX_HIP(hipMemcpyAsync(d_jobs, live.data() + c0, (c1 - c0) * sizeof(GatherJob), hipMemcpyHostToDevice,
                                A.stream));
        SDFX_LAUNCH(k_gather_batch, dim3(blocks_for(max_nt, 256), (unsigned)(c1 - c0)), dim3(256), 0, A.stream,
                    d_jobs);
        SDFX_HIP(hipGetLastError());
    }
}

__global__ void __launch_bounds__(256) k_check_finite(const float* __restrict__ q, uint64_t n, DevFlags* flags) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float x = q[i];
        bad |= !(fabsf(x) <= FLT_MAX);
    }
    if (bad) atomicOr(&flags->non_finite_query, 1u);
}

void check_queries_finite(const float* d_q, uint64_t nq, DevFlags* d_flags, hipStream_t s) {
    if (!nq) return;
    uint64_t n = 3 * nq;
    unsigned blocks = blocks_for(n, 256);
    if (blocks > 4096) blocks = 4096;
    SDFX_LAUNCH(k_check_finite, dim3(blocks), dim3(256), 0, s, d_q, n, d_flags);
    SDFX_HIP(hipGetLastError());
}

__global__ void k_fill(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_u32(uint32_t* d_ptr, uint64_t n, uint32_t v, hipStream_t s) {
    if (!n) return;
    unsigned blocks = blocks_for(n, 256);
    if (blocks > 8192) blocks = 8192;
    SDFX_LAUNCH(k_fill, dim3(blocks), dim3(256), 0, s, d_ptr, n, v);
    SDFX_HIP(hipGetLastError());
}

uint64_t sum_block_counters(const unsigned long long* d_slots, uint64_t n, hipStream_t s) {
    std::vector<unsigned long long> h(n);
    SDFX_HIP(hipMemcpyAsync(h.data(), d_slots, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    SDFX_HIP(hipStreamSynchronize(s));
    uint64_t t = 0;
    for (auto x : h) t += x;
    return t;
}

// ---------------------------------------------------------------- scans

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Tile-local exclusive scan (4 items per lane, wave shuffles, LDS across waves);
// the tile total goes to tile_sums[blockIdx.x].
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tile(const T* __restrict__ in, T* __restrict__ out,
                                                            uint64_t n, T* __restrict__ tile_sums) {
    __shared__ T wave_sums[SCAN_THREADS / 64];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T local = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = base + k < n ? in[base + k] : T(0);
        local += v[k];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wave_sums[wave] = x;
    __syncthreads();
    T wave_off = 0;
    for (int w = 0; w < wave; ++w) wave_off += wave_sums[w];
    T run = wave_off + x - local;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        T total = 0;
        for (int w = 0; w < SCAN_THREADS / 64; ++w) total += wave_sums[w];
        tile_sums[blockIdx.x] = total;
    }
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_add(T* __restrict__ out, uint64_t n,
                                                           const T* __restrict__ tile_offsets) {
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    T off = tile_offsets[blockIdx.x];
    for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_THREADS)
        if (base + k < n) out[base + k] += off;
}

template <typename T>
__global__ void k_copy_one(const T* src, T* dst) {
    *dst = *src;
}

template <typename T>
static void scan_rec(const T* d_in, T* d_out, uint64_t n, T* d_total, Arena& A) {
    uint64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    T* sums = A.alloc<T>(tiles);
    SDFX_LAUNCH(k_scan_tile<T>, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, A.stream, d_in, d_out, n, sums);
    SDFX_HIP(hipGetLastError());
    if (tiles == 1) {
        if (d_total) SDFX_LAUNCH(k_copy_one<T>, dim3(1), dim3(1), 0, A.stream, sums, d_total);
        return;
    }
    T* offs = A.alloc<T>(tiles);
    scan_rec<T>(sums, offs, tiles, d_total, A);
    SDFX_LAUNCH(k_scan_add<T>, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, A.stream, d_out, n, offs);
    SDFX_HIP(hipGetLastError());
}

void exclusive_scan_u64(const uint64_t* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total, Arena& A) {
    if (!n) return;
    scan_rec<uint64_t>(d_in, d_out, n, d_total, A);
}

void exclusive_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, uint32_t* d_total, Arena& A) {
    if (!n) return;
    scan_rec<uint32_t>(d_in, d_out, n, d_total, A);
}

}  // namespace sdfx
