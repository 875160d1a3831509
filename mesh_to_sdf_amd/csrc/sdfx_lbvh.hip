// sdfx_lbvh.hip — device LBVH over the padded triangle boxes (geo.rs:4-22) and
// the point-query kernels of the Bvh, Rtree and RtreeBvh paths.
//
// Build: centroid bounds -> 30-bit Morton codes -> stable LSD radix sort (8-bit
// digits, LDS histograms, wave-ballot multisplit ranking) -> Karras 2012
// hierarchy over (code, position) keys -> bottom-up refit with one atomic
// visit counter per node. Each internal node stores both children's boxes, so
// one 64-byte fetch decides both descents.
//
// Queries. The reference's R-tree (rstar 0.12) and BVH (bvh 0.10) searches
// are both conservative, so their results are fixed by the triangles
// themselves (generic/bvh.rs:131-176, generic/rtree.rs:147-154,
// generic/rtree_bvh.rs:153-198):
//  * nearest triangle = minimum FP32 d^2 (lowest topology index on exact ties);
//  * each of the +X/+Y/+Z ray counts = number of triangles the aligned test
//    hits, since a hit point lies inside the triangle's padded box.
// Nodes are pruned only when their box distance is clearly above the best d^2
// (2^-16 relative slack) and ray descent keeps every box the ray meets, so the
// device results equal the exhaustive ones.
#include "sdfx_internal.hpp"

namespace sdfx {

// ------------------------------------------------------------------ bounds

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = fbits(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return bitsf((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ void tri_centroid(const TriRec& t, float* c) {
    Box b;
    tri_bbox(t.ax, t.ay, t.az, t.bx, t.by, t.bz, t.cx, t.cy, t.cz, b);
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = 0.5f * (b.lo[k] + b.hi[k]);
}

// One mesh of a (batched) LBVH build. Kernels run on 2-D grids: blockIdx.y is
// the mesh, blockIdx.x * blockDim.x + threadIdx.x the triangle (or internal
// node) within it; blocks past a small mesh's end exit at once.
struct BuildSeg {
    const TriRec* tris;  // the mesh's triangles in topology order
    LeafRec* leaves;     // out: n triangles in Morton order
    BvhNode* nodes;      // out: n - 1 internal nodes
    uint64_t tri_off;    // first triangle (and sorted position) of the mesh in the batch
    uint64_t node_off;   // first internal node of the mesh in the batch
    uint64_t n;
};

// Per mesh six bound words (min x, y, z as ordered ints, then max), each 64
// u32 (256 bytes) apart: device-scope atomics on one cache line serialise.
constexpr uint32_t BOUND_STRIDE = 64;
constexpr uint32_t BOUND_WORDS = 6 * BOUND_STRIDE;
// Centroid-bound blocks each reduce this many triangles of one mesh.
constexpr uint32_t BOUND_CHUNK = 256 * 16;

__global__ void k_init_bounds(uint32_t* ob, uint32_t nseg) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)nseg * 6) return;
    ob[(i / 6) * BOUND_WORDS + (i % 6) * BOUND_STRIDE] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
}

__global__ void __launch_bounds__(256) k_centroid_bounds(const BuildSeg* __restrict__ segs, uint32_t* __restrict__ ob) {
    const BuildSeg& S = segs[blockIdx.y];
    const uint64_t c0 = (uint64_t)blockIdx.x * BOUND_CHUNK;
    if (c0 >= S.n) return;  // block-uniform
    const uint64_t c1 = min(S.n, c0 + BOUND_CHUNK);
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    for (uint64_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
        float c[3];
        tri_centroid(S.tris[i], c);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uint32_t o = f2ord(c[k]);
            mn[k] = min(mn[k], o);
            mx[k] = max(mx[k], o);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], off, 64));
            mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], off, 64));
        }
    }
    // block-wide, then one check-before-atomic per block and word
    __shared__ uint32_t wmn[4][3], wmx[4][3];
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            wmn[wave][k] = mn[k];
            wmx[wave][k] = mx[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x % 3;
        uint32_t* w = &ob[(uint64_t)blockIdx.y * BOUND_WORDS + threadIdx.x * BOUND_STRIDE];
        if (threadIdx.x < 3) {
            const uint32_t m = min(min(wmn[0][k], wmn[1][k]), min(wmn[2][k], wmn[3][k]));
            if (m < __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(w, m);
        } else {
            const uint32_t m = max(max(wmx[0][k], wmx[1][k]), max(wmx[2][k], wmx[3][k]));
            if (m > __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(w, m);
        }
    }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Sort key of each triangle: its mesh in the top seg_bits bits, then the top
// morton_bits bits of its 30-bit Morton code (centroid in the mesh's centroid
// box, 10 bits per axis). With one mesh the key is the full Morton code.
__global__ void __launch_bounds__(256) k_morton(const BuildSeg* __restrict__ segs, const uint32_t* __restrict__ ob,
                                                int morton_bits, uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
    const BuildSeg& S = segs[blockIdx.y];
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    const uint32_t* b = ob + (uint64_t)blockIdx.y * BOUND_WORDS;
    float c[3];
    tri_centroid(S.tris[i], c);
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lo = ord2f(b[k * BOUND_STRIDE]), hi = ord2f(b[(3 + k) * BOUND_STRIDE]);
        float ext = hi - lo;
        float u = ext > 0.0f ? (c[k] - lo) / ext : 0.0f;
        float sc = u * 1024.0f;
        q[k] = sc <= 0.0f ? 0u : (sc >= 1023.0f ? 1023u : (uint32_t)sc);  // NaN -> 0
    }
    const uint32_t code = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
    const uint32_t seg_part = morton_bits < 32 ? (uint32_t)blockIdx.y << morton_bits : 0u;
    keys[S.tri_off + i] = seg_part | (code >> (30 - morton_bits));
    vals[S.tri_off + i] = (uint32_t)(S.tri_off + i);
}

// ------------------------------------------------------------------ radix sort

constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;
constexpr int RS_WAVES = RS_THREADS / 64;

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint32_t* __restrict__ keys, uint64_t n, int shift,
                                                        uint32_t* __restrict__ hist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        uint64_t i = base + (uint64_t)r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: items are ranked in position order. Within one round of 256
// items, a lane's rank among equal digits of its wave comes from 8 ballots
// (multisplit); earlier waves' counts come from LDS; earlier rounds' counts
// from the running per-digit offsets.
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint32_t* __restrict__ keys_in,
                                                           const uint32_t* __restrict__ vals_in, uint64_t n,
                                                           int shift, const uint32_t* __restrict__ offsets,
                                                           uint32_t nblocks, uint32_t* __restrict__ keys_out,
                                                           uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wave_cnt[RS_WAVES][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    run[tid] = offsets[(uint64_t)tid * nblocks + blockIdx.x];
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < RS_ROUNDS; ++r) {
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) wave_cnt[w][tid] = 0;
        __syncthreads();
        uint64_t i = base + (uint64_t)r * RS_THREADS + tid;
        bool valid = i < n;
        uint32_t key = valid ? keys_in[i] : 0u;
        uint32_t val = valid ? vals_in[i] : 0u;
        uint32_t digit = (key >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (digit >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rank_in_wave = (uint32_t)__popcll(peers & lt_mask);
        // the highest lane of each peer group publishes the group's count
        if (valid && (peers >> lane) == 1ull) wave_cnt[wave][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t before = 0;
            for (int w = 0; w < wave; ++w) before += wave_cnt[w][digit];
            uint32_t dst = run[digit] + before + rank_in_wave;
            if (dst < n) {
                keys_out[dst] = key;
                vals_out[dst] = val;
            }
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) tot += wave_cnt[w][tid];
        run[tid] += tot;
        __syncthreads();
    }
}

// Sorts (keys, vals) by the low `bits` bits of keys; stable. Result ends in keys/vals.
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint64_t n, int bits, Arena& A) {
    if (n <= 1) return;
    uint32_t nblocks = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
    uint32_t* k2 = A.alloc<uint32_t>(n);
    uint32_t* v2 = A.alloc<uint32_t>(n);
    uint32_t* hist = A.alloc<uint32_t>((uint64_t)256 * nblocks);
    uint32_t *ki = keys, *vi = vals, *ko = k2, *vo = v2;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        SDFX_LAUNCH(k_rs_hist, dim3(nblocks), dim3(RS_THREADS), 0, A.stream, ki, n, shift, hist, nblocks);
        exclusive_scan_u32(hist, hist, (uint64_t)256 * nblocks, nullptr, A);
        SDFX_LAUNCH(k_rs_scatter, dim3(nblocks), dim3(RS_THREADS), 0, A.stream, ki, vi, n, shift, hist,
                           nblocks, ko, vo);
        SDFX_HIP(hipGetLastError());
        uint32_t* t;
        t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    if (ki != keys) {
        SDFX_HIP(hipMemcpyAsync(keys, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, A.stream));
        SDFX_HIP(hipMemcpyAsync(vals, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, A.stream));
    }
}

// ------------------------------------------------------------------ Karras build

// Common-prefix length of sorted keys i and j; equal codes fall back to the
// positions so every key is distinct (Karras 2012, section 4).
__device__ __forceinline__ int delta(const uint32_t* __restrict__ keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = keys[i], b = keys[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)(i ^ j));
}

// Karras 2012 on one mesh's sorted keys (keys[0, n) of that mesh); the
// outputs are the mesh's own arrays, indices local to the mesh.
__global__ void __launch_bounds__(256) k_karras(const BuildSeg* __restrict__ segs, const uint32_t* __restrict__ skeys,
                                                int32_t* __restrict__ g_left, int32_t* __restrict__ g_right,
                                                int32_t* __restrict__ g_parent_int,
                                                int32_t* __restrict__ g_parent_leaf, int32_t* __restrict__ g_last) {
    const BuildSeg& S = segs[blockIdx.y];
    const int64_t n = (int64_t)S.n;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const uint32_t* keys = skeys + S.tri_off;
    int32_t* left = g_left + S.node_off;
    int32_t* right = g_right + S.node_off;
    int32_t* parent_int = g_parent_int + S.node_off;
    int32_t* parent_leaf = g_parent_leaf + S.tri_off;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int64_t lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    int32_t lc = lo == gamma ? ~(int32_t)gamma : (int32_t)gamma;
    int32_t rc = hi == gamma + 1 ? ~(int32_t)(gamma + 1) : (int32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    g_last[S.node_off + i] = (int32_t)hi;
    if (lc < 0) parent_leaf[~lc] = (int32_t)i;
    else parent_int[lc] = (int32_t)i;
    if (rc < 0) parent_leaf[~rc] = (int32_t)i;
    else parent_int[rc] = (int32_t)i;
    if (i == 0) parent_int[0] = -1;
}

__global__ void __launch_bounds__(256) k_leaves(const BuildSeg* __restrict__ segs, const uint32_t* __restrict__ order,
                                                Box* __restrict__ g_leaf_box) {
    const BuildSeg& S = segs[blockIdx.y];
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    const TriRec t = S.tris[order[S.tri_off + i] - S.tri_off];
    S.leaves[i] = make_leafrec(t);
    Box b;
    tri_bbox(t.ax, t.ay, t.az, t.bx, t.by, t.bz, t.cx, t.cy, t.cz, b);
    g_leaf_box[S.tri_off + i] = b;
}

// Boxes handed from one thread to another (possibly on another XCD) are
// stored and read with agent-scope atomics, which go to the device's coherence
// point instead of an XCD's own L2. Only the counter that says "both children
// are done" then has to be ordered after them, which a workgroup-scope fence
// (a wait for the stores to complete) does without writing back or
// invalidating the L2 as a device-scope fence would on every tree level.
__device__ __forceinline__ void store_box_agent(Box* b, const Box& v) {
    float* f = reinterpret_cast<float*>(b);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        __hip_atomic_store(&f[k], v.lo[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&f[3 + k], v.hi[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
__device__ __forceinline__ Box load_box_agent(const Box* b) {
    const float* f = reinterpret_cast<const float*>(b);
    Box r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = __hip_atomic_load(&f[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        r.hi[k] = __hip_atomic_load(&f[3 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return r;
}

// Bottom-up refit: one thread per leaf climbs; at each internal node the first
// of the two children to arrive stops, the second (which finds the sibling's
// box published) writes the node and goes on with the union.
__global__ void __launch_bounds__(256) k_refit(const BuildSeg* __restrict__ segs,
                                               const int32_t* __restrict__ g_parent_leaf,
                                               const int32_t* __restrict__ g_parent_int,
                                               const int32_t* __restrict__ g_left, const int32_t* __restrict__ g_right,
                                               const Box* __restrict__ g_leaf_box, Box* g_node_box, uint32_t* g_visit,
                                               const int32_t* __restrict__ g_last) {
    const BuildSeg& S = segs[blockIdx.y];
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (S.n < 2 || i >= S.n) return;
    const int32_t* parent_int = g_parent_int + S.node_off;
    const int32_t* left = g_left + S.node_off;
    const int32_t* right = g_right + S.node_off;
    const Box* leaf_box = g_leaf_box + S.tri_off;
    Box* node_box = g_node_box + S.node_off;
    uint32_t* visit = g_visit + S.node_off;
    const int32_t* last = g_last + S.node_off;
    BvhNode* nodes = S.nodes;
    int32_t p = g_parent_leaf[S.tri_off + i];
    for (int level = 0; p >= 0 && level < 128; ++level) {  // depth <= 64 in a Karras tree
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // this thread's box stores are complete
        if (__hip_atomic_fetch_add(&visit[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
            return;  // the sibling subtree is not finished yet
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int32_t l = left[p], r = right[p];
        // leaf boxes come from an earlier kernel; internal ones were published above
        const Box lb = l < 0 ? leaf_box[~l] : load_box_agent(&node_box[l]);
        const Box rb = r < 0 ? leaf_box[~r] : load_box_agent(&node_box[r]);
        Box u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u.lo[k] = fminf(lb.lo[k], rb.lo[k]);
            u.hi[k] = fmaxf(lb.hi[k], rb.hi[k]);
        }
        store_box_agent(&node_box[p], u);
        BvhNode nd;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            nd.lmin[k] = lb.lo[k];
            nd.lmax[k] = lb.hi[k];
            nd.rmin[k] = rb.lo[k];
            nd.rmax[k] = rb.hi[k];
        }
        nd.left = l;
        nd.right = r;
        nd.last = last[p];
        nd.parent = parent_int[p];
        nodes[p] = nd;
        p = nd.parent;
    }
}

void lbvh_build_batch(const std::vector<LbvhBuildJob>& jobs, Arena& A) {
    hipStream_t s = A.stream;
    std::vector<BuildSeg> segs;
    uint64_t total = 0, total_nodes = 0, max_n = 0;
    for (const LbvhBuildJob& J : jobs) {
        Lbvh& bvh = *J.out;
        bvh.n = J.n;
        if (J.n == 0) continue;
        // ray-traversal stack entries pack a node index with 3 axis bits
        if (J.n >= (1ull << 29)) throw Error(SDFX_ERR_INVALID_ARGUMENT, "too many triangles for one LBVH");
        bvh.leaves.reset(J.n);
        bvh.nodes.reset(J.n > 1 ? J.n - 1 : 1);
        bvh.root = J.n > 1 ? 0 : ~0;
        segs.push_back(BuildSeg{J.tris, bvh.leaves.get(), bvh.nodes.get(), total, total_nodes, J.n});
        total += J.n;
        total_nodes += J.n - 1;
        max_n = std::max(max_n, J.n);
    }
    if (segs.empty()) return;
    if (total >= (1ull << 32) || segs.size() > 65535)
        throw Error(SDFX_ERR_INVALID_ARGUMENT, "an LBVH build batch holds fewer than 2^32 triangles and 65,536 meshes");
    const uint32_t nseg = (uint32_t)segs.size();
    int seg_bits = 0;  // the mesh field of the sort key
    while ((1u << seg_bits) < nseg) ++seg_bits;
    const int morton_bits = std::min(30, 32 - seg_bits);
    BuildSeg* d_segs = A.alloc<BuildSeg>(nseg);
    SDFX_HIP(hipMemcpyAsync(d_segs, segs.data(), nseg * sizeof(BuildSeg), hipMemcpyHostToDevice, s));
    uint32_t* bounds = A.alloc<uint32_t>((uint64_t)nseg * BOUND_WORDS);
    uint32_t* keys = A.alloc<uint32_t>(total);
    uint32_t* vals = A.alloc<uint32_t>(total);
    Box* leaf_box = A.alloc<Box>(total);
    const unsigned per_mesh = blocks_for(max_n, 256);
    SDFX_LAUNCH(k_init_bounds, dim3(blocks_for(6ull * nseg, 256)), dim3(256), 0, s, bounds, nseg);
    SDFX_LAUNCH(k_centroid_bounds, dim3(blocks_for(max_n, BOUND_CHUNK), nseg), dim3(256), 0, s, d_segs, bounds);
    SDFX_LAUNCH(k_morton, dim3(per_mesh, nseg), dim3(256), 0, s, d_segs, bounds, morton_bits, keys, vals);
    SDFX_HIP(hipGetLastError());
    // sorted by (mesh, Morton code): mesh m's triangles land on [tri_off, tri_off + n)
    radix_sort_pairs(keys, vals, total, seg_bits + morton_bits, A);
    SDFX_LAUNCH(k_leaves, dim3(per_mesh, nseg), dim3(256), 0, s, d_segs, vals, leaf_box);
    SDFX_HIP(hipGetLastError());
    if (total_nodes == 0) return;  // single-triangle meshes only
    int32_t* left = A.alloc<int32_t>(total_nodes);
    int32_t* right = A.alloc<int32_t>(total_nodes);
    int32_t* parent_int = A.alloc<int32_t>(total_nodes);
    int32_t* parent_leaf = A.alloc<int32_t>(total);
    Box* node_box = A.alloc<Box>(total_nodes);
    uint32_t* visit = A.alloc<uint32_t>(total_nodes);
    int32_t* last = A.alloc<int32_t>(total_nodes);
    fill_u32(visit, total_nodes, 0u, s);
    SDFX_LAUNCH(k_karras, dim3(per_mesh, nseg), dim3(256), 0, s, d_segs, keys, left, right, parent_int, parent_leaf,
                last);
    SDFX_LAUNCH(k_refit, dim3(per_mesh, nseg), dim3(256), 0, s, d_segs, parent_leaf, parent_int, left, right,
                leaf_box, node_box, visit, last);
    SDFX_HIP(hipGetLastError());
}

void lbvh_build(const TriRec* d_tris, uint64_t n, Lbvh& bvh, Arena& A) {
    lbvh_build_batch({LbvhBuildJob{d_tris, n, &bvh}}, A);
}

// ------------------------------------------------------------------ traversal

// Traversal stack of one lane: a ring of STACK_RING entries in LDS (the lane's
// own column, so lanes never share words and need no barrier), two u32 words
// per entry. A push onto a full ring overwrites the oldest entry and the
// stack remembers that it dropped one; every traversal below recovers what
// was dropped without other storage, so no traversal touches scratch memory.
// The traversal kernels run 256-thread blocks: 8 entries x 8 bytes x 256
// lanes = 16 KB of LDS per block keeps 8 waves per SIMD resident.
// Blocks per stretch of the XCD mapping of the Morton-ordered query kernels
// (xcd_chunk_block). Stretches of 1 to 16 blocks measured alike; one
// contiguous eighth of the queries per XCD left the XCDs unevenly loaded
// (served burst 69.6 -> 60.0 ms, noise-sphere Bvh(Raycast) 24.2 -> 23.3 ms,
// Rtree 20.9 -> 21.3 ms).
constexpr uint32_t QUERY_XCD_CHUNK = 8;
constexpr int STACK_RING = 8;
constexpr int STACK_LANES = 256;

__device__ __forceinline__ uint32_t* stack_lds() {
    __shared__ uint32_t words[2 * STACK_RING * STACK_LANES];
    return words;
}

struct RingStack {
    uint32_t* col;  // this lane's column: entry k at col[2 * k * STACK_LANES] / col[(2 * k + 1) * STACK_LANES]
    int head = 0, count = 0;
    bool dropped = false;  // an entry was overwritten since the flag was last cleared
    __device__ RingStack() : col(stack_lds() + threadIdx.x) {}
    __device__ __forceinline__ bool empty() const { return count == 0; }
    // Returns true when the ring was full: the oldest entry was overwritten
    // and its second word is in dropped_b.
    __device__ __forceinline__ bool push(uint32_t a, uint32_t b, uint32_t& dropped_b) {
        const int h = head & (STACK_RING - 1);  // with a full ring, the oldest entry's slot
        const bool full = count == STACK_RING;
        if (full) {
            dropped_b = col[(2 * h + 1) * STACK_LANES];
            dropped = true;
            --count;
        }
        col[2 * h * STACK_LANES] = a;
        col[(2 * h + 1) * STACK_LANES] = b;
        ++head;
        ++count;
        return full;
    }
    __device__ __forceinline__ void pop(uint32_t& a, uint32_t& b) {  // requires !empty()
        --head;
        --count;
        const int h = head & (STACK_RING - 1);
        a = col[2 * h * STACK_LANES];
        b = col[(2 * h + 1) * STACK_LANES];
    }
};

// Stack of the first pass's nearest search: a ring of NEAR_RING one-word
// entries in the LDS words of the two-word ring. A pushed node is always
// internal (leaf children are visited at once), so its index fits in the high
// `nb` bits, where nb covers the tree's internal nodes, and the low 32 - nb
// bits hold the top bits of its box's d^2: non-negative float bits order as
// integers, so the dropped low bits make the stored value a lower bound, and
// popping with it keeps every box the exact value keeps (the node's own
// children are then tested with exact boxes).
//
// A push onto a full ring overwrites the oldest entry, which the ring hands
// back; nearest_from keeps the newest dropped one and recovers the others by
// climbing parent links (no scratch memory: half of the uniform queries of a
// noise sphere hold more than eight pending boxes within their final bound).
constexpr int NEAR_RING = 2 * STACK_RING;

// The packing of one tree: `cap` bounds its node count (visit_cap, at least
// twice the internal nodes), so internal indices stay below cap / 2.
struct NearCode {
    uint32_t dbits, dshift;  // d^2 bits kept and the shift that drops the rest
    __device__ explicit NearCode(uint32_t cap) {
        const uint32_t nb = 32u - (uint32_t)__clz((int)max(cap >> 1, 1u));  // 1..31
        dbits = 32u - nb;
        dshift = nb - 1u;  // 31 significant bits of a non-negative float, dbits kept
    }
    __device__ __forceinline__ uint32_t pack(int32_t node, float d2) const {
        return ((uint32_t)node << dbits) | ((fbits(d2) & 0x7FFFFFFFu) >> dshift);
    }
    __device__ __forceinline__ int32_t node(uint32_t e) const { return (int32_t)(e >> dbits); }
    __device__ __forceinline__ float d2_floor(uint32_t e) const {
        return bitsf((e & ((1u << dbits) - 1u)) << dshift);
    }
};

struct NearRing {
    uint32_t* col;  // this lane's ring column: entry k at col[k * STACK_LANES]
    int head = 0, count = 0;
    __device__ NearRing() : col(stack_lds() + threadIdx.x) {}
    __device__ __forceinline__ bool empty() const { return count == 0; }
    // Returns true when the ring was full: its oldest entry, now overwritten, is in `dropped`.
    __device__ __forceinline__ bool push(uint32_t e, uint32_t& dropped) {
        uint32_t* slot = col + (head & (NEAR_RING - 1)) * STACK_LANES;  // with a full ring, the oldest entry
        const bool full = count == NEAR_RING;
        if (full) dropped = *slot;
        else ++count;
        *slot = e;
        ++head;
        return full;
    }
    __device__ __forceinline__ uint32_t pop() {  // requires !empty()
        --head;
        --count;
        return col[(head & (NEAR_RING - 1)) * STACK_LANES];
    }
};

struct Nearest {
    float d2;
    uint32_t index;  // topology index (tie-break)
    int32_t leaf;    // Morton position, -1 if none
};

// Traversal statistics gathered while kernel timing is on.
struct TravStats {
    uint32_t nodes = 0, leaves = 0, vote_nodes = 0, vote_leaves = 0;
    uint32_t drops = 0;      // nearest-search ring entries dropped (recovered by climbing)
    uint32_t restarts = 0;   // restarts of exact traversals after the ring dropped an entry
};

// Visit caps: a tree visits each node once per pass, plus at most one descent
// of at most 64 nodes per restart, and a restart follows a drop, which
// follows a push of a visited node; more visits mean a malformed tree.
__device__ __forceinline__ uint32_t exact_cap(uint32_t cap) { return cap >= (1u << 25) ? 0xFFFFFFFFu : cap * 65u; }

__device__ __forceinline__ void visit_leaf_nearest(const LeafRec* __restrict__ leaves, int32_t leaf, float px,
                                                   float py, float pz, Nearest& best, TravStats& ts) {
    ts.leaves++;
    const TriRec t = expand_leaf(leaves[leaf]);
    float d2 = tri_dist2(px, py, pz, t);
    if (d2 < best.d2 || (d2 == best.d2 && t.index < best.index)) {
        best.d2 = d2;
        best.index = t.index;
        best.leaf = leaf;
    }
}

// Left-first depth-first traversal of the subtree at `start` (an internal
// node), exact whatever the ring drops. With Karras's node numbering (see
// BvhNode) the leaves are reached in increasing position, and `done` is a
// position up to which every leaf has been visited or ruled out (start's first
// position minus one on entry). When the ring drained after dropping an entry,
// the traversal restarts at `start` and skips every left child that ends at or
// before `done`: no leaf is visited twice and none is missed, and since each
// drain finished at least one more subtree, `done` grows and the traversal
// ends.
//
// P, the traversal's policy, decides what a child box needs and what a leaf
// does: child(lo, hi, pay, child_pay) whether to enter a child box (pay is
// what the path carries: the vote's live ray axes, or the box's distance for
// pruning), keep(pay) whether a stacked entry is still wanted when popped (the
// bound may have tightened since the push), leaf(position, pay), and
// root_pay. Returns false when `budget` visits ran out below `cap` (the
// search is incomplete); past exact_cap(cap) visits the tree is malformed.
template <class P>
__device__ bool dfs_exact(const BvhNode* __restrict__ nodes, int32_t start, int32_t done, P& pol,
                          uint32_t& visits, uint32_t budget, uint32_t cap, bool& overflow, TravStats& ts) {
    RingStack st;
    int32_t node = start;
    uint32_t pay = pol.root_pay;
    const uint32_t limit = exact_cap(cap);
    for (;;) {
        // node: an internal node whose range extends beyond `done`
        ++visits;
        if (visits > budget && budget < cap) return false;
        if (visits > limit) {
            overflow = true;
            return true;
        }
        pol.count_node();
        const BvhNode nd = nodes[node];
        const int32_t s = nd.left < 0 ? ~nd.left : nd.left;
        uint32_t pl = 0, pr = 0;
        bool gl = s > done && pol.child(nd.lmin, nd.lmax, pay, pl);
        const bool gr = pol.child(nd.rmin, nd.rmax, pay, pr);
        if (gl && nd.left < 0) {
            pol.leaf(~nd.left, pl);
            gl = false;
        }
        if (gl) {  // the right child, leaf or node, waits for the left subtree
            uint32_t db;
            if (gr) st.push((uint32_t)nd.right, pr, db);
            node = nd.left;
            pay = pl;
            continue;
        }
        done = max(done, s);  // the left child is finished
        if (gr) {
            if (nd.right >= 0) {
                node = nd.right;
                pay = pr;
                continue;
            }
            pol.leaf(~nd.right, pr);
        }
        done = max(done, nd.last);
        // next: the newest stacked entry still wanted (a right child: its index
        // is its first position), else a restart after a drop, else the end
        bool found = false;
        while (!st.empty()) {
            uint32_t ea, eb;
            st.pop(ea, eb);
            const int32_t e = (int32_t)ea;
            const int32_t first = e < 0 ? ~e : e;
            done = max(done, first - 1);
            if (!pol.keep(eb)) continue;  // the next entry or a restart moves `done` past it
            if (e < 0) {
                pol.leaf(first, eb);
                done = first;
                continue;
            }
            node = e;
            pay = eb;
            found = true;
            break;
        }
        if (found) continue;
        if (!st.dropped) return true;
        st.dropped = false;
        ts.restarts++;
        node = start;
        pay = pol.root_pay;
    }
}

// Nearest search of one lane from the root, closest child first, folding
// into the lane's running best. Returns false when more than `budget` nodes
// were entered with nodes left (the search is incomplete; best holds what was
// found so far) if budget is below `cap`; more than `cap` visits means a
// malformed tree (overflow).
//
// Pruning keeps a box when box_d2 * PRUNE_SLACK <= bound, and bound is always
// the d^2 of a triangle already seen, so a box is dropped only if every
// triangle in it is strictly farther than one already seen. The answer is
// therefore the minimum of (d^2, topology index) over all triangles whatever
// the visiting order, start nodes, repeated visits or initial bound:
// k_lbvh_coop, which splits one query's tree among the lanes of a wave,
// returns the same answer as one lane.
//
// Pending boxes are the far children of the current path's ancestors, newest
// (deepest) on top of the ring. A full ring drops its oldest entry, so the
// dropped entries are the far children of the path's shallowest pushing
// ancestors, and the newest dropped entry E is the deepest of them. Once the
// ring drains, the search takes E up (if still within the bound) and, when
// E's subtree is done, climbs E's parent links to the root: at each ancestor
// it recomputes both child distances, which give the same near / far choice
// as on the way down, and enters the far child if it came up from the near
// one and the far box is still within the bound. Every ancestor above E's
// parent pushed its far child before E (so it was dropped too) or never
// pushed it (a leaf visited at once, or a box then outside the bound), so the
// climb finds exactly the boxes the ring lost. A drop inside a subtree taken
// up this way replaces E: its climb passes through that subtree's root.
__device__ bool nearest_from(const BvhNode* __restrict__ nodes, const LeafRec* __restrict__ leaves, int32_t root,
                             float px, float py, float pz, uint32_t budget, uint32_t cap, Nearest& best,
                             bool& overflow, TravStats& ts) {
    auto leaf = [&](int32_t lf) { visit_leaf_nearest(leaves, lf, px, py, pz, best, ts); };
    if (root < 0) {
        leaf(~root);
        return true;
    }
    NearRing st;
    const NearCode code(cap);
    int32_t node = root;
    uint32_t visits = 0;
    // the newest dropped entry not yet taken up, or NO_DROP (its node field
    // 2^nb - 1 is no internal node's index: those stay below cap / 2)
    constexpr uint32_t NO_DROP = 0xFFFFFFFFu;
    uint32_t drop = NO_DROP;
    int32_t resume = -1;  // once the ring drains, climb from this node (its subtree is then done)
    for (;;) {
        ++visits;
        if (visits > budget) {
            if (budget < cap) return false;
            overflow = true;  // a tree enters each node once; more means a malformed tree
            return true;
        }
        ts.nodes++;
        const BvhNode nd = nodes[node];
        float dl = box_dist2(px, py, pz, nd.lmin, nd.lmax);
        float dr = box_dist2(px, py, pz, nd.rmin, nd.rmax);
        float b = best.d2;
        bool gl = dl * PRUNE_SLACK <= b;
        bool gr = dr * PRUNE_SLACK <= b;
        // leaf children, left first; one call site, so lanes visiting the
        // left leaf and lanes visiting the right one share the evaluation
        const int32_t l0 = gl && nd.left < 0 ? ~nd.left : -1, l1 = gr && nd.right < 0 ? ~nd.right : -1;
        if (l0 >= 0) gl = false;
        if (l1 >= 0) gr = false;
        int32_t pending = l0 >= 0 ? l0 : l1, then = l0 >= 0 ? l1 : -1;
#pragma unroll 1
        while (pending >= 0) {
            leaf(pending);
            pending = then >= 0 && dr * PRUNE_SLACK <= best.d2 ? then : -1;  // the right leaf, pruned anew
            then = -1;
        }
        if (gl && gr) {
            const bool left_first = dl <= dr;
            uint32_t old;
            if (st.push(code.pack(left_first ? nd.right : nd.left, left_first ? dr : dl), old)) {
                drop = old;
                ts.drops++;
            }
            node = left_first ? nd.left : nd.right;
            continue;
        }
        if (gl) {
            node = nd.left;
            continue;
        }
        if (gr) {
            node = nd.right;
            continue;
        }
        // next: the newest ring entry still within the bound, else the newest
        // dropped entry, else a far child found by climbing
        bool found = false;
        while (!st.empty()) {
            const uint32_t e = st.pop();
            if (code.d2_floor(e) * PRUNE_SLACK <= best.d2) {
                node = code.node(e);
                found = true;
                break;
            }
        }
        if (found) continue;
        if (drop != NO_DROP) {
            resume = code.node(drop);
            const bool keep = code.d2_floor(drop) * PRUNE_SLACK <= best.d2;
            drop = NO_DROP;
            if (keep) {
                node = resume;
                continue;
            }
        }
        if (resume < 0) break;
        int32_t x = resume, p = nodes[resume].parent;
        found = false;
        for (int hops = 0; p >= 0; ++hops) {
            if (hops >= 64) {
                overflow = true;  // a Karras tree is at most 64 deep
                return true;
            }
            ts.nodes++;
            const BvhNode pn = nodes[p];
            const float cl = box_dist2(px, py, pz, pn.lmin, pn.lmax);
            const float cr = box_dist2(px, py, pz, pn.rmin, pn.rmax);
            const bool left_first = cl <= cr;
            const int32_t near = left_first ? pn.left : pn.right, far = left_first ? pn.right : pn.left;
            if (x == near && far >= 0 && (left_first ? cr : cl) * PRUNE_SLACK <= best.d2) {
                node = far;
                resume = far;
                found = true;
                break;
            }
            x = p;
            p = pn.parent;
        }
        if (!found) break;
    }
    return true;
}

// Minimum d^2 over all triangles, closest child first.
__device__ Nearest bvh_nearest(const BvhNode* __restrict__ nodes, const LeafRec* __restrict__ leaves, int32_t root,
                               float px, float py, float pz, uint32_t max_visits, bool& overflow, TravStats& ts) {
    Nearest best{INFINITY, 0xFFFFFFFFu, -1};
    nearest_from(nodes, leaves, root, px, py, pz, max_visits, max_visits, best, overflow, ts);
    return best;
}

// Bvh(Normal) (generic/bvh.rs:117-129): min_dist starts at f32::MAX and takes
// each candidate's signed distance when compare_distances(min_dist, d) is
// Greater. compare_distances is a strict weak order (a positive value comes
// before the negative values it approximately equals, below a threshold that
// grows with them), and candidates it calls Equal have the same sign and |d|,
// i.e. the same value, so the fold's result does not depend on the order the
// candidates come in, nor on a candidate coming twice: the traversal is the
// left-first exact one. Nodes are skipped only when their box is farther than
// any value that could tie with (ulps 2 / eps 1e-6) or beat |min_dist|.
struct FoldPolicy {
    const LeafRec* __restrict__ leaves;
    float px, py, pz;
    float acc = FLT_MAX;
    bool nan_seen = false;
    static constexpr uint32_t root_pay = 0u;
    __device__ float bound2() const {
        float b = fabsf(acc) * (1.0f + 1.0f / 1048576.0f) + 4e-6f;
        return b * b * (1.0f + 1.0f / 1048576.0f);
    }
    __device__ void count_node() {}
    __device__ bool child(const float* lo, const float* hi, uint32_t, uint32_t& cp) {
        const float d2 = box_dist2(px, py, pz, lo, hi);
        cp = fbits(d2);
        return d2 <= bound2();
    }
    __device__ bool keep(uint32_t b) { return bitsf(b) <= bound2(); }
    __device__ void leaf(int32_t pos, uint32_t) {
        const float d = tri_signed_dist(px, py, pz, expand_leaf(leaves[pos]));
        const int c = compare_distances(acc, d);
        nan_seen |= c == 2;
        if (c == 1) acc = d;
    }
};

__device__ float bvh_normal_fold(const BvhNode* __restrict__ nodes, const LeafRec* __restrict__ leaves,
                                 int32_t root, float px, float py, float pz, uint32_t max_visits, bool& nan_seen,
                                 bool& overflow, TravStats& ts) {
    FoldPolicy pol{leaves, px, py, pz};
    if (root < 0) {
        pol.leaf(~root, 0u);
    } else {
        uint32_t visits = 0;
        dfs_exact(nodes, root, -1, pol, visits, max_visits, max_visits, overflow, ts);
    }
    nan_seen = pol.nan_seen;
    return pol.acc;
}

__device__ __forceinline__ void ray_leaf(const LeafRec* __restrict__ leaves, int32_t leaf, uint32_t axes, float px,
                                         float py, float pz, uint32_t* cnt, TravStats& ts) {
    ts.vote_leaves++;
    const TriRec t = expand_leaf(leaves[leaf]);
    if ((axes & 1u) && ray_hit<0>(px, py, pz, t, nullptr)) cnt[0]++;
    if ((axes & 2u) && ray_hit<1>(px, py, pz, t, nullptr)) cnt[1]++;
    if ((axes & 4u) && ray_hit<2>(px, py, pz, t, nullptr)) cnt[2]++;
}

// Policy of dfs_exact for the 3-ray vote: the rays of root_pay's axes share
// one traversal and each path carries its live axes (a triangle must be
// counted exactly once per ray, which the exact traversal guarantees).
struct VotePolicy {
    const LeafRec* __restrict__ leaves;
    float px, py, pz;
    TravStats& ts;
    uint32_t cnt[3] = {0u, 0u, 0u};
    uint32_t root_pay = 7u;
    __device__ void count_node() { ts.vote_nodes++; }
    __device__ bool child(const float* lo, const float* hi, uint32_t axes, uint32_t& cp) {
        cp = axes & ray_axes(lo, hi, px, py, pz);
        return cp != 0u;
    }
    __device__ bool keep(uint32_t) { return true; }
    __device__ void leaf(int32_t pos, uint32_t axes) { ray_leaf(leaves, pos, axes, px, py, pz, cnt, ts); }
};

// The 3-ray vote of generic/bvh.rs:141-176 / generic/rtree_bvh.rs:163-198:
// inside when at least two of the +x, +y, +z rays cross the surface an odd
// number of times. The +x and +y rays go first; when their parities agree
// they decide the vote whatever the +z ray counts, so the +z ray is traced
// only where they disagree.
__device__ bool bvh_inside_vote(const BvhNode* __restrict__ nodes, const LeafRec* __restrict__ leaves, int32_t root,
                                float px, float py, float pz, uint32_t max_visits, bool& overflow, TravStats& ts) {
    VotePolicy pol{leaves, px, py, pz, ts};
    if (root < 0) {
        pol.leaf(~root, 7u);
    } else {
        uint32_t visits = 0;
        pol.root_pay = 3u;
        dfs_exact(nodes, root, -1, pol, visits, max_visits, max_visits, overflow, ts);
        if ((pol.cnt[0] ^ pol.cnt[1]) & 1u) {
            pol.root_pay = 4u;
            visits = 0;
            dfs_exact(nodes, root, -1, pol, visits, max_visits, max_visits, overflow, ts);
        }
    }
    const int insides = (int)(pol.cnt[0] & 1u) + (int)(pol.cnt[1] & 1u) + (int)(pol.cnt[2] & 1u);
    return insides > 1;
}

// generic/rtree_bvh.rs:217-276: closest_point.sub(point) of the nearest
// triangle (NaN without one).
__device__ __forceinline__ void closest_vector(const Nearest& nb, float px, float py, float pz,
                                               const LeafRec* __restrict__ leaves, float* __restrict__ v3) {
    float vx = NAN, vy = NAN, vz = NAN;
    if (nb.leaf >= 0) {
        float dx, dy, dz, cx, cy, cz;
        closest_diff(px, py, pz, expand_leaf(leaves[nb.leaf]), dx, dy, dz, cx, cy, cz);
        vx = cx - px;
        vy = cy - py;
        vz = cz - pz;
    }
    v3[0] = vx;
    v3[1] = vy;
    v3[2] = vz;
}

// The answer of a nearest-search mode from the nearest triangle (d^2, leaf)
// and the 3-ray vote.
template <int MODE>
__device__ __forceinline__ float nearest_answer(const Nearest& nb, bool inside, float px, float py, float pz,
                                                const LeafRec* __restrict__ leaves) {
    if (MODE == QM_NEAREST_SIGNED)
        return nb.leaf >= 0 ? tri_signed_dist(px, py, pz, expand_leaf(leaves[nb.leaf])) : sqrtf(nb.d2);
    float d = sqrtf(nb.d2);
    if (MODE == QM_RAYCAST_VOTE) d = fminf(FLT_MAX, d);  // generic/bvh.rs:131-139 folds from f32::MAX
    return inside ? -d : d;
}

// A query whose nearest search passed the visit budget: its index, the best
// triangle found so far (a bound for the rest of the search) and its vote.
struct DeferRec {
    uint64_t q;    // query index within its request
    uint32_t d2;   // bits of the best d^2 so far
    uint32_t index;
    int32_t leaf;
    uint32_t inside;
    uint32_t req;  // request of a batched query (0 otherwise)
    uint32_t pad;
};

// The answer of one query point in mode MODE (the per-query work of every
// query kernel below). With `defer`, a nearest search that passes `budget`
// visits stops and fills *defer instead of out[i] (returns true); the vote
// still runs here.
// One query of the point paths at (px, py, pz). Returns true when the nearest
// search ran out of its budget and the query goes to the per-wave pass (defer
// holds its state; the caller sets defer.q); otherwise `ans` is the answer.
template <int MODE>
__device__ __forceinline__ bool query_point(const BvhNode* __restrict__ nodes, const LeafRec* __restrict__ leaves,
                                            int32_t root, uint32_t max_visits, float px, float py, float pz,
                                            DevFlags* flags, TravStats& ts, uint32_t budget, bool can_defer,
                                            DeferRec& defer, float& ans) {
    bool overflow = false, deferred = false;
    if (MODE == QM_NORMAL_FOLD) {
        bool nan_seen = false;
        ans = bvh_normal_fold(nodes, leaves, root, px, py, pz, max_visits, nan_seen, overflow, ts);
        if (nan_seen) atomicOr(&flags->nan_distance, 1u);
    } else {
        Nearest nb{INFINITY, 0xFFFFFFFFu, -1};
        const bool done = nearest_from(nodes, leaves, root, px, py, pz, can_defer ? min(budget, max_visits) : max_visits,
                                       max_visits, nb, overflow, ts);
        const bool inside = MODE != QM_NEAREST_SIGNED &&
                            bvh_inside_vote(nodes, leaves, root, px, py, pz, max_visits, overflow, ts);
        if (done) {
            ans = nearest_answer<MODE>(nb, inside, px, py, pz, leaves);
        } else {
            defer = DeferRec{0ull, fbits(nb.d2), nb.index, nb.leaf, inside ? 1u : 0u, 0u, 0u};
            deferred = true;
        }
    }
    if (overflow) atomicOr(&flags->internal, 1u);
    return deferred;
}

template <int MODE>
// 8 waves per SIMD (at most 64 VGPRs; 65 and 7 waves otherwise).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_lbvh_query(const BvhNode* __restrict__ nodes,
                                                    const LeafRec* __restrict__ leaves, int32_t root,
                                                    uint32_t max_visits, const float* __restrict__ q, uint64_t nq,
                                                    const uint32_t* __restrict__ order, float* __restrict__ out,
                                                    DevFlags* flags, unsigned long long* stats, uint32_t budget,
                                                    DeferRec* __restrict__ deferred, uint32_t* n_deferred) {
    // with a Morton order, the XCDs take stretches of QUERY_XCD_CHUNK blocks in turn
    uint64_t i = (uint64_t)(order ? xcd_chunk_block(QUERY_XCD_CHUNK) : blockIdx.x) * blockDim.x + threadIdx.x;
    const bool active = i < nq;  // no early return: the statistics below use block barriers
    TravStats ts;
    bool dfr = false;
    DeferRec rec;
    if (active) {
        if (order) i = order[i];  // lanes of a wave take Morton-adjacent queries
        float ans;
        dfr = query_point<MODE>(nodes, leaves, root, max_visits, q[3 * i], q[3 * i + 1], q[3 * i + 2], flags, ts,
                                budget, deferred != nullptr, rec, ans);
        if (dfr) rec.q = i;
        else out[i] = ans;
    }
    if (deferred) {  // kernel argument: uniform; one atomic per wave appends its deferred queries
        const uint64_t m = __ballot(dfr);
        if (m) {
            const uint32_t lane = threadIdx.x & 63u;
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(n_deferred, (uint32_t)__popcll(m));
            base = __shfl(base, leader, 64);
            if (dfr) deferred[base + __popcll(m & ((1ull << lane) - 1ull))] = rec;
        }
    }
    if (stats) {  // kernel argument: uniform, every lane of the block gets here
        const uint64_t nb = gridDim.x;
        block_count(stats, ts.nodes);
        block_count(stats + nb, ts.leaves);
        block_count(stats + 2 * nb, ts.vote_nodes);
        block_count(stats + 3 * nb, ts.vote_leaves);
        // lane slots the wave spends: 64 x its slowest lane's node visits (a
        // wave steps until every lane is done)
        uint32_t mx = ts.nodes, mv = ts.vote_nodes;
        for (int off = 32; off > 0; off >>= 1) {
            mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
            mv = max(mv, (uint32_t)__shfl_xor((int)mv, off, 64));
        }
        block_count(stats + 4 * nb, (threadIdx.x & 63u) == 0 ? 64u * mx : 0u);
        block_count(stats + 5 * nb, (threadIdx.x & 63u) == 0 ? 64u * mv : 0u);
        block_count(stats + 6 * nb, ts.drops);
        block_count(stats + 7 * nb, ts.restarts);
    }
}

// One request of a batched point query (one tree, its queries and outputs).
struct PointReq {
    const BvhNode* nodes;
    const LeafRec* leaves;
    const float* q;
    float* out;
    uint64_t q_off;  // first query of the request in the batch numbering
    int32_t root;
    uint32_t max_visits;
    uint32_t flag;
    uint32_t budget;  // nearest-search visits before the per-wave pass (batched queries)
    uint64_t p_off;   // first launch position of the request: a multiple of 64 (k_lbvh_query_batch)
    uint32_t nq;      // queries of the request
    uint32_t pad;
};

// Entries per wave in the work pool of k_lbvh_coop (node, box d^2 bits).
constexpr uint32_t COOP_POOL = 128;

// The deferred nearest searches, one wave per query, wave-synchronous: each
// step every busy lane visits one node, idle lanes take subtrees that busy
// lanes donate to a pool in LDS while the pool holds fewer than 64 entries,
// and nodes are pruned against the wave's minimum d^2. A lane walks its
// subtree left child first, keeping the right children it does not donate on
// its ring; after the ring dropped an entry, the lane restarts at its subtree
// once the ring drains, skipping the left children that end at or before the
// last position it finished (dfs_exact explains the numbering; donated right
// children are all right children, so a subtree's index is its first
// position). The answer is the wave's minimum of (d^2, topology index),
// starting from the best triangle the first pass found (nearest_from explains
// why the order, and visiting a subtree twice, do not matter). Persistent
// waves take queries from *next until all *n_deferred are done.
template <int MODE>
// 6 waves per SIMD (at most 80 VGPRs, 20-24 bytes of spills per lane): left to
// itself the compiler takes 87-89 VGPRs and 5 waves, measured 1-2% slower on
// the noise-sphere workload.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_lbvh_coop(const PointReq* __restrict__ reqs, PointReq single,
                                                   const DeferRec* __restrict__ recs, const uint32_t* n_deferred,
                                                   uint32_t* next, DevFlags* flags, unsigned long long* stats) {
    __shared__ uint32_t pool_node[4][COOP_POOL], pool_d2[4][COOP_POOL];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t count = *n_deferred;
    bool overflow = false;
    TravStats ts;
    for (;;) {
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(next, 1u);
        j = __shfl(j, 0, 64);
        if (j >= count) break;
        const DeferRec r = recs[j];
        const PointReq R = reqs ? reqs[r.req] : single;
        const BvhNode* __restrict__ nodes = R.nodes;
        const LeafRec* __restrict__ leaves = R.leaves;
        const uint32_t limit = exact_cap(R.max_visits);
        const float px = R.q[3 * r.q], py = R.q[3 * r.q + 1], pz = R.q[3 * r.q + 2];
        Nearest best = lane == 0 ? Nearest{bitsf(r.d2), r.index, r.leaf} : Nearest{INFINITY, 0xFFFFFFFFu, -1};
        RingStack st;
        int32_t cur = -1;   // the lane's node, -1 when idle (its ring is then empty)
        int32_t sub = -1;   // the lane's subtree (restart point)
        int32_t done = -1;  // leaf position up to which the lane's subtree is finished
        uint32_t pn = 1;    // pool entries: the same value in every lane
        if (lane == 0) {
            pool_node[w][0] = (uint32_t)R.root;
            pool_d2[w][0] = 0u;
        }
        uint32_t visits = 0;
        for (;;) {
            float bound = best.d2;
            for (int off = 32; off > 0; off >>= 1) bound = fminf(bound, __shfl_xor(bound, off, 64));
            // idle lanes take the newest pool entries
            const uint64_t idle = __ballot(cur < 0);
            if (cur < 0) {
                const uint32_t rk = (uint32_t)__popcll(idle & below);
                if (rk < pn) {
                    const uint32_t slot = pn - 1u - rk;
                    if (bitsf(pool_d2[w][slot]) * PRUNE_SLACK <= bound) {
                        cur = sub = (int32_t)pool_node[w][slot];
                        done = sub - 1;
                        st.count = 0;  // empty already unless a malformed tree cut the last subtree short
                        st.dropped = false;
                    }
                }
            }
            pn -= min(pn, (uint32_t)__popcll(idle));
            if (__ballot(cur >= 0) == 0 && pn == 0) break;
            bool has_far = false;
            uint32_t far_node = 0, far_d2 = 0;
            if (cur >= 0) {
                if (++visits > limit) {  // see exact_cap: a malformed tree
                    overflow = true;
                    cur = -1;
                } else {
                    ts.nodes++;
                    const float bnd = fminf(bound, best.d2);
                    const BvhNode nd = nodes[cur];
                    const int32_t s = nd.left < 0 ? ~nd.left : nd.left;
                    const float dl = box_dist2(px, py, pz, nd.lmin, nd.lmax);
                    const float dr = box_dist2(px, py, pz, nd.rmin, nd.rmax);
                    bool gl = s > done && dl * PRUNE_SLACK <= bnd;
                    bool gr = dr * PRUNE_SLACK <= bnd;
                    const int32_t l0 = gl && nd.left < 0 ? ~nd.left : -1, l1 = gr && nd.right < 0 ? ~nd.right : -1;
                    if (l0 >= 0) gl = false;
                    if (l1 >= 0) gr = false;
                    int32_t pending = l0 >= 0 ? l0 : l1, then = l0 >= 0 ? l1 : -1;
#pragma unroll 1
                    while (pending >= 0) {  // one call site for left and right leaves
                        visit_leaf_nearest(leaves, pending, px, py, pz, best, ts);
                        pending = then >= 0 && dr * PRUNE_SLACK <= fminf(bound, best.d2) ? then : -1;
                        then = -1;
                    }
                    if (gl) {
                        cur = nd.left;
                        if (gr) {
                            far_node = (uint32_t)nd.right;
                            far_d2 = fbits(dr);
                            has_far = true;
                        }
                    } else if (gr) {
                        done = max(done, s);
                        cur = nd.right;
                    } else {
                        done = max(done, nd.last);
                        cur = -1;
                        while (!st.empty()) {
                            uint32_t a, bb;
                            st.pop(a, bb);
                            done = max(done, (int32_t)a - 1);  // a right child: its index is its first position
                            if (bitsf(bb) * PRUNE_SLACK <= fminf(bound, best.d2)) {
                                cur = (int32_t)a;
                                break;
                            }
                        }
                        if (cur < 0 && st.dropped) {  // restart the subtree, skipping what is finished
                            st.dropped = false;
                            ts.restarts++;
                            cur = sub;
                        }
                    }
                }
            }
            // right children go to the pool while it runs low (at most
            // 63 + 64 < COOP_POOL entries), else to the lane's own ring
            const uint64_t donors = __ballot(has_far);
            if (has_far) {
                if (pn < 64u) {
                    const uint32_t slot = pn + (uint32_t)__popcll(donors & below);
                    pool_node[w][slot] = far_node;
                    pool_d2[w][slot] = far_d2;
                } else {
                    uint32_t db;
                    st.push(far_node, far_d2, db);
                }
            }
            if (pn < 64u) pn += (uint32_t)__popcll(donors);
        }
        // lexicographic minimum of (d^2, index) over the wave, and its leaf
        const uint64_t key = ((uint64_t)fbits(best.d2) << 32) | best.index;
        uint64_t mk = key;
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = __shfl_xor(mk, off, 64);
            mk = o < mk ? o : mk;
        }
        const int wl = __ffsll((unsigned long long)__ballot(key == mk)) - 1;
        Nearest nb{bitsf((uint32_t)(mk >> 32)), (uint32_t)mk, __shfl(best.leaf, wl, 64)};
        if (lane == 0) {
            if constexpr (MODE == QM_CLOSEST_VECTOR) closest_vector(nb, px, py, pz, leaves, R.out + 3 * r.q);
            else R.out[r.q] = nearest_answer<MODE>(nb, r.inside != 0u, px, py, pz, leaves);
        }
        if (__ballot(overflow) && lane == 0) atomicOr(&flags[R.flag].internal, 1u);
        overflow = false;
    }
    if (stats) {
        block_count(stats, ts.nodes);
        block_count(stats + gridDim.x, ts.leaves);
        block_count(stats + 2 * gridDim.x, ts.restarts);
    }
}

static uint32_t visit_cap(const Lbvh& bvh) { return (uint32_t)(2 * bvh.n + 64); }

// Node visits after which a nearest search moves to the per-wave pass: the
// tuning knob, or with 0 (the default) 15 x cbrt(triangles), which measured
// best with Bvh(Raycast) on noise spheres of 5,120 to 327,680 triangles
// (budgets 256 to 1024; half near-surface, half uniform queries).
static uint32_t visit_budget(const Lbvh& bvh) {
    const int64_t k = tuning().lbvh_visit_budget;
    if (k > 0) return (uint32_t)k;
    return (uint32_t)std::max(64.0, 15.0 * std::cbrt((double)bvh.n));
}

// Morton code of each query inside the tree's root box (clamped outside it).
__global__ void __launch_bounds__(256) k_query_morton(const BvhNode* __restrict__ nodes, const float* __restrict__ q,
                                                      uint64_t nq, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const BvhNode r = nodes[0];
    uint32_t c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lo = fminf(r.lmin[k], r.rmin[k]), hi = fmaxf(r.lmax[k], r.rmax[k]);
        float ext = hi - lo;
        float u = ext > 0.0f ? (q[3 * i + k] - lo) / ext : 0.0f;
        float sc = u * 1024.0f;
        c[k] = sc <= 0.0f ? 0u : (sc >= 1023.0f ? 1023u : (uint32_t)sc);  // NaN -> 0
    }
    keys[i] = (expand10(c[0]) << 2) | (expand10(c[1]) << 1) | expand10(c[2]);
    vals[i] = (uint32_t)i;
}

// Query order for coherent traversal: the permutation sorting the queries by
// Morton code, so a wave's lanes walk nearby paths (null when not worth it).
static const uint32_t* query_order(const Lbvh& bvh, const float* d_q, uint64_t nq, Arena& A) {
    if (nq < 4096 || bvh.root < 0 || nq >= (1ull << 32)) return nullptr;
    uint32_t* keys = A.alloc<uint32_t>(nq);
    uint32_t* vals = A.alloc<uint32_t>(nq);
    SDFX_LAUNCH(k_query_morton, dim3(blocks_for(nq, 256)), dim3(256), 0, A.stream, bvh.nodes.get(), d_q, nq, keys,
                vals);
    radix_sort_pairs(keys, vals, nq, 30, A);
    return vals;
}

template <int MODE>
static void launch_query(const Lbvh& bvh, const float* d_q, uint64_t nq, const uint32_t* order, unsigned blocks,
                         float* d_out, DevFlags* d_flags, unsigned long long* stats, uint32_t budget,
                         DeferRec* recs, uint32_t* n_deferred, hipStream_t s) {
    SDFX_LAUNCH(k_lbvh_query<MODE>, dim3(blocks), dim3(256), 0, s, bvh.nodes.get(), bvh.leaves.get(), bvh.root,
                visit_cap(bvh), d_q, nq, order, d_out, d_flags, stats, budget, recs, n_deferred);
}

// The per-wave pass over the deferred queries: of one tree (reqs null, the
// tree and buffers in `single`) or of a batch's requests. words: n_deferred, next.
static void launch_coop(QueryMode mode, const PointReq* reqs, const PointReq& single, uint32_t* words,
                        const DeferRec* recs, unsigned blocks, DevFlags* d_flags, unsigned long long* stats,
                        hipStream_t s) {
    switch (mode) {
        case QM_RAYCAST_VOTE:
            SDFX_LAUNCH(k_lbvh_coop<QM_RAYCAST_VOTE>, dim3(blocks), dim3(256), 0, s, reqs, single, recs, words,
                        words + 1, d_flags, stats);
            break;
        case QM_NEAREST_SIGNED:
            SDFX_LAUNCH(k_lbvh_coop<QM_NEAREST_SIGNED>, dim3(blocks), dim3(256), 0, s, reqs, single, recs, words,
                        words + 1, d_flags, stats);
            break;
        case QM_NEAREST_VOTE:
            SDFX_LAUNCH(k_lbvh_coop<QM_NEAREST_VOTE>, dim3(blocks), dim3(256), 0, s, reqs, single, recs, words,
                        words + 1, d_flags, stats);
            break;
        case QM_CLOSEST_VECTOR:
            SDFX_LAUNCH(k_lbvh_coop<QM_CLOSEST_VECTOR>, dim3(blocks), dim3(256), 0, s, reqs, single, recs, words,
                        words + 1, d_flags, stats);
            break;
        default:
            break;
    }
}

void lbvh_query(const Lbvh& bvh, const float* d_q, uint64_t nq, QueryMode mode, float* d_out, DevFlags* d_flags,
                Arena& A) {
    if (!nq || !bvh.n) return;
    hipStream_t s = A.stream;
    const uint32_t* order = query_order(bvh, d_q, nq, A);
    const unsigned blocks = order ? xcd_chunk_grid(blocks_for(nq, 256), QUERY_XCD_CHUNK) : blocks_for(nq, 256);
    // Nearest searches longer than the visit budget finish in a second pass
    // that gives each of them a whole wave: a few interior queries can need
    // tens of thousands of visits, and one lane walking them alone would set
    // the pass's duration.
    const uint64_t budget = visit_budget(bvh);
    const bool split = mode != QM_NORMAL_FOLD && bvh.root >= 0 && budget < visit_cap(bvh);
    DeferRec* recs = nullptr;
    uint32_t* words = nullptr;  // n_deferred, next
    if (split) {
        recs = A.alloc<DeferRec>(nq);
        words = A.alloc<uint32_t>(2);
        fill_u32(words, 2, 0u, s);
    }
    unsigned long long* stats = nullptr;
    const unsigned coop_blocks = (unsigned)std::min<uint64_t>(2048, (nq + 3) / 4);
    if (profile_enabled() && mode != QM_NORMAL_FOLD) {
        stats = A.alloc<unsigned long long>(8ull * blocks + 3ull * coop_blocks);
        fill_u32(reinterpret_cast<uint32_t*>(stats), 16ull * blocks + 6ull * coop_blocks, 0u, s);
    }
    const uint32_t b32 = split ? (uint32_t)budget : 0xFFFFFFFFu;
    switch (mode) {
        case QM_RAYCAST_VOTE:
            launch_query<QM_RAYCAST_VOTE>(bvh, d_q, nq, order, blocks, d_out, d_flags, stats, b32, recs, words, s);
            break;
        case QM_NORMAL_FOLD:
            launch_query<QM_NORMAL_FOLD>(bvh, d_q, nq, order, blocks, d_out, d_flags, stats, b32, nullptr, nullptr, s);
            break;
        case QM_NEAREST_SIGNED:
            launch_query<QM_NEAREST_SIGNED>(bvh, d_q, nq, order, blocks, d_out, d_flags, stats, b32, recs, words, s);
            break;
        case QM_NEAREST_VOTE:
            launch_query<QM_NEAREST_VOTE>(bvh, d_q, nq, order, blocks, d_out, d_flags, stats, b32, recs, words, s);
            break;
        case QM_CLOSEST_VECTOR:  // vectors have their own entry point, lbvh_closest_vectors
            throw Error(SDFX_ERR_INVALID_ARGUMENT, "closest-point vectors are not an SDF query mode");
    }
    unsigned long long* cstats = stats ? stats + 8ull * blocks : nullptr;
    if (split) {
        const PointReq single{bvh.nodes.get(), bvh.leaves.get(), d_q, d_out, 0, bvh.root, visit_cap(bvh), 0u, 0u};
        launch_coop(mode, nullptr, single, words, recs, coop_blocks, d_flags, cstats, s);
    }
    SDFX_HIP(hipGetLastError());
    if (stats) {
        profile_count("lbvh.nearest_nodes", sum_block_counters(stats, blocks, s));
        profile_count("lbvh.nearest_leaves", sum_block_counters(stats + blocks, blocks, s));
        profile_count("lbvh.vote_nodes", sum_block_counters(stats + 2ull * blocks, blocks, s));
        profile_count("lbvh.vote_leaves", sum_block_counters(stats + 3ull * blocks, blocks, s));
        profile_count("lbvh.nearest_wave_slots", sum_block_counters(stats + 4ull * blocks, blocks, s));
        profile_count("lbvh.vote_wave_slots", sum_block_counters(stats + 5ull * blocks, blocks, s));
        profile_count("lbvh.nearest_drops", sum_block_counters(stats + 6ull * blocks, blocks, s));
        profile_count("lbvh.restarts", sum_block_counters(stats + 7ull * blocks, blocks, s));
        profile_count("lbvh.queries", nq);
        if (split) {
            profile_count("lbvh.deferred_nodes", sum_block_counters(cstats, coop_blocks, s));
            profile_count("lbvh.deferred_leaves", sum_block_counters(cstats + coop_blocks, coop_blocks, s));
            profile_count("lbvh.deferred_restarts", sum_block_counters(cstats + 2ull * coop_blocks, coop_blocks, s));
            uint32_t nd = 0;
            SDFX_HIP(hipMemcpyAsync(&nd, words, sizeof(nd), hipMemcpyDeviceToHost, s));
            SDFX_HIP(hipStreamSynchronize(s));
            profile_count("lbvh.deferred_queries", nd);
        }
    }
}

// The request holding batch query g: the last one whose q_off is <= g.
__device__ __forceinline__ uint32_t point_req_of(const PointReq* __restrict__ reqs, uint32_t nreq, uint64_t g) {
    uint32_t lo = 0, hi = nreq - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (reqs[mid].q_off <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Sort key of batch query g: its request in the top 8 bits, then a 24-bit
// Morton code inside its mesh's root box, so a sort groups each request's
// queries in spatial order.
__global__ void __launch_bounds__(256) k_point_morton_batch(const PointReq* __restrict__ reqs, uint32_t nreq,
                                                            uint64_t total, uint32_t* __restrict__ keys,
                                                            uint32_t* __restrict__ vals) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const uint32_t r = point_req_of(reqs, nreq, g);
    const PointReq& R = reqs[r];
    const uint64_t i = g - R.q_off;
    uint32_t code = 0;
    if (R.root >= 0) {
        const BvhNode n0 = R.nodes[0];
        uint32_t c[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float lo = fminf(n0.lmin[k], n0.rmin[k]), hi = fmaxf(n0.lmax[k], n0.rmax[k]);
            const float ext = hi - lo;
            const float u = ext > 0.0f ? (R.q[3 * i + k] - lo) / ext : 0.0f;
            const float sc = u * 256.0f;
            c[k] = sc <= 0.0f ? 0u : (sc >= 255.0f ? 255u : (uint32_t)sc);  // NaN -> 0
        }
        code = (expand10(c[0]) << 2) | (expand10(c[1]) << 1) | expand10(c[2]);
    }
    keys[g] = (r << 24) | code;
    vals[g] = (uint32_t)g;
}

// The request holding launch position i of k_lbvh_query_batch: the last one
// whose p_off is <= i. i is wave-uniform, so the search runs on scalar loads.
__device__ __forceinline__ uint32_t point_req_at(const PointReq* __restrict__ reqs, uint32_t nreq, uint64_t i) {
    uint32_t lo = 0, hi = nreq - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (wave_uniform_load(reqs, mid).p_off <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The queries of many requests in one launch. Launch positions run request by
// request, each request's first position a multiple of 64 (p_off), so every
// wave serves one request: its fields, the tree pointers among them, sit in
// scalar registers and the traversal is the single-request one of
// k_lbvh_query. Position p_off + j takes the request's j-th query in sorted
// order (batch query q_off + j when unsorted).
template <int MODE>
// 8 waves per SIMD (at most 64 VGPRs), as k_lbvh_query.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_lbvh_query_batch(const PointReq* __restrict__ reqs, uint32_t nreq,
                                                          uint64_t positions, const uint32_t* __restrict__ order,
                                                          DevFlags* flags, DeferRec* __restrict__ deferred,
                                                          uint32_t* n_deferred) {
    // the wave's first position (with a Morton order, the XCDs take
    // stretches of QUERY_XCD_CHUNK blocks in turn)
    const uint64_t w0 = (uint64_t)(order ? xcd_chunk_block(QUERY_XCD_CHUNK) : blockIdx.x) * blockDim.x +
                        __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    bool dfr = false;
    DeferRec rec;
    if (w0 < positions) {  // no early return: the wave appends its deferred queries together
        const uint32_t r = point_req_at(reqs, nreq, w0);
        const PointReq R = wave_uniform_load(reqs, r);
        const uint32_t j = (uint32_t)(w0 - R.p_off) + (threadIdx.x & 63u);
        if (j < R.nq) {
            const uint64_t s = R.q_off + j;
            const uint32_t k = (uint32_t)((order ? order[s] : s) - R.q_off);  // a batch holds fewer than 2^32 queries
            TravStats ts;
            float ans;
            dfr = query_point<MODE>(R.nodes, R.leaves, R.root, R.max_visits, R.q[3 * k], R.q[3 * k + 1],
                                    R.q[3 * k + 2], flags + R.flag, ts, R.budget, deferred != nullptr, rec, ans);
            if (dfr) {
                rec.q = k;
                rec.req = r;
            } else {
                R.out[k] = ans;
            }
        }
    }
    if (deferred) {  // kernel argument: uniform
        const uint64_t m = __ballot(dfr);
        if (m) {
            const uint32_t lane = threadIdx.x & 63u;
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(n_deferred, (uint32_t)__popcll(m));
            base = __shfl(base, leader, 64);
            if (dfr) deferred[base + __popcll(m & ((1ull << lane) - 1ull))] = rec;
        }
    }
}

void lbvh_query_batch(const std::vector<PointBatchItem>& items, QueryMode mode, DevFlags* d_flags, Arena& A) {
    hipStream_t s = A.stream;
    constexpr size_t CHUNK = 256;  // requests per launch: the sort key holds the request in 8 bits
    for (size_t c0 = 0; c0 < items.size(); c0 += CHUNK) {
        const size_t c1 = std::min(items.size(), c0 + CHUNK);
        std::vector<PointReq> h;
        uint64_t total = 0, positions = 0;
        for (size_t k = c0; k < c1; ++k) {
            const PointBatchItem& it = items[k];
            if (!it.nq || !it.bvh || !it.bvh->n) continue;
            if (it.nq >= (1ull << 32)) throw Error(SDFX_ERR_INVALID_ARGUMENT, "a point batch holds fewer than 2^32 queries");
            h.push_back(PointReq{it.bvh->nodes.get(), it.bvh->leaves.get(), it.d_q, it.d_out, total, it.bvh->root,
                                 visit_cap(*it.bvh), it.flag, visit_budget(*it.bvh), positions, (uint32_t)it.nq, 0u});
            total += it.nq;
            positions += (it.nq + 63) / 64 * 64;  // every request starts on a wave
        }
        if (h.empty()) continue;
        if (total >= (1ull << 32)) throw Error(SDFX_ERR_INVALID_ARGUMENT, "a point batch holds fewer than 2^32 queries");
        PointReq* d_reqs = A.alloc<PointReq>(h.size());
        A.upload(d_reqs, h.data(), h.size() * sizeof(PointReq));
        const uint32_t nreq = (uint32_t)h.size();
        const uint32_t* order = nullptr;
        if (total >= 4096) {
            uint32_t* keys = A.alloc<uint32_t>(total);
            uint32_t* vals = A.alloc<uint32_t>(total);
            SDFX_LAUNCH(k_point_morton_batch, dim3(blocks_for(total, 256)), dim3(256), 0, s, d_reqs, nreq, total, keys,
                        vals);
            int rbits = 0;  // the request field's width
            while ((1u << rbits) < nreq) ++rbits;
            radix_sort_pairs(keys, vals, total, 24 + rbits, A);
            order = vals;
        }
        const unsigned blocks = order ? xcd_chunk_grid(blocks_for(positions, 256), QUERY_XCD_CHUNK) : blocks_for(positions, 256);
        // long nearest searches finish one per wave, as in lbvh_query
        const bool split = mode != QM_NORMAL_FOLD;
        DeferRec* recs = nullptr;
        uint32_t* words = nullptr;  // n_deferred, next
        if (split) {
            recs = A.alloc<DeferRec>(total);
            words = A.alloc<uint32_t>(2);
            fill_u32(words, 2, 0u, s);
        }
        switch (mode) {
            case QM_RAYCAST_VOTE:
                SDFX_LAUNCH(k_lbvh_query_batch<QM_RAYCAST_VOTE>, dim3(blocks), dim3(256), 0, s, d_reqs, nreq,
                            positions, order, d_flags, recs, words);
                break;
            case QM_NORMAL_FOLD:
                SDFX_LAUNCH(k_lbvh_query_batch<QM_NORMAL_FOLD>, dim3(blocks), dim3(256), 0, s, d_reqs, nreq,
                            positions, order, d_flags, recs, words);
                break;
            case QM_NEAREST_SIGNED:
                SDFX_LAUNCH(k_lbvh_query_batch<QM_NEAREST_SIGNED>, dim3(blocks), dim3(256), 0, s, d_reqs, nreq,
                            positions, order, d_flags, recs, words);
                break;
            case QM_NEAREST_VOTE:
                SDFX_LAUNCH(k_lbvh_query_batch<QM_NEAREST_VOTE>, dim3(blocks), dim3(256), 0, s, d_reqs, nreq,
                            positions, order, d_flags, recs, words);
                break;
            case QM_CLOSEST_VECTOR:
                throw Error(SDFX_ERR_INVALID_ARGUMENT, "closest-point vectors are not an SDF query mode");
        }
        if (split)
            launch_coop(mode, d_reqs, PointReq{}, words, recs, (unsigned)std::min<uint64_t>(2048, (total + 3) / 4),
                        d_flags, nullptr, s);
        SDFX_HIP(hipGetLastError());
    }
}

// query_vector_to_closest_point: the vector to the nearest triangle's closest
// point and the 3-ray vote per query; nearest searches past `budget` visits
// are deferred to k_lbvh_coop (which writes their vectors) as in lbvh_query.
__global__ void __launch_bounds__(256) k_closest_vectors(const BvhNode* __restrict__ nodes,
                                                         const LeafRec* __restrict__ leaves, int32_t root,
                                                         uint32_t max_visits, const float* __restrict__ q, uint64_t nq,
                                                         const uint32_t* __restrict__ order, float* __restrict__ vec,
                                                         uint8_t* __restrict__ inside, DevFlags* flags,
                                                         uint32_t budget, DeferRec* __restrict__ deferred,
                                                         uint32_t* n_deferred) {
    uint64_t i = (uint64_t)(order ? xcd_chunk_block(QUERY_XCD_CHUNK) : blockIdx.x) * blockDim.x + threadIdx.x;
    bool dfr = false;
    DeferRec rec;
    if (i < nq) {  // no early return: the wave appends its deferred queries together
        if (order) i = order[i];
        const float px = q[3 * i], py = q[3 * i + 1], pz = q[3 * i + 2];
        bool overflow = false;
        TravStats ts;
        Nearest nb{INFINITY, 0xFFFFFFFFu, -1};
        const bool done = nearest_from(nodes, leaves, root, px, py, pz, deferred ? min(budget, max_visits) : max_visits,
                                       max_visits, nb, overflow, ts);
        if (done) closest_vector(nb, px, py, pz, leaves, vec + 3 * i);
        else rec = DeferRec{i, fbits(nb.d2), nb.index, nb.leaf, 0u, 0u, 0u};
        dfr = !done;
        inside[i] = bvh_inside_vote(nodes, leaves, root, px, py, pz, max_visits, overflow, ts) ? 1 : 0;
        if (overflow) atomicOr(&flags->internal, 1u);
    }
    if (deferred) {  // kernel argument: uniform; one atomic per wave appends its deferred queries
        const uint64_t m = __ballot(dfr);
        if (m) {
            const uint32_t lane = threadIdx.x & 63u;
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(n_deferred, (uint32_t)__popcll(m));
            base = __shfl(base, leader, 64);
            if (dfr) deferred[base + __popcll(m & ((1ull << lane) - 1ull))] = rec;
        }
    }
}

void lbvh_closest_vectors(const Lbvh& bvh, const float* d_q, uint64_t nq, float* d_vec, uint8_t* d_inside,
                          DevFlags* d_flags, Arena& A) {
    if (!nq || !bvh.n) return;
    hipStream_t s = A.stream;
    const uint32_t* order = query_order(bvh, d_q, nq, A);
    const unsigned blocks = order ? xcd_chunk_grid(blocks_for(nq, 256), QUERY_XCD_CHUNK) : blocks_for(nq, 256);
    const uint32_t budget = visit_budget(bvh);
    const bool split = bvh.root >= 0 && budget < visit_cap(bvh);
    DeferRec* recs = nullptr;
    uint32_t* words = nullptr;  // n_deferred, next
    if (split) {
        recs = A.alloc<DeferRec>(nq);
        words = A.alloc<uint32_t>(2);
        fill_u32(words, 2, 0u, s);
    }
    SDFX_LAUNCH(k_closest_vectors, dim3(blocks), dim3(256), 0, s, bvh.nodes.get(), bvh.leaves.get(), bvh.root,
                visit_cap(bvh), d_q, nq, order, d_vec, d_inside, d_flags, budget, recs, words);
    if (split) {
        const PointReq single{bvh.nodes.get(), bvh.leaves.get(), d_q, d_vec, 0, bvh.root, visit_cap(bvh), 0u, 0u};
        launch_coop(QM_CLOSEST_VECTOR, nullptr, single, words, recs, (unsigned)std::min<uint64_t>(2048, (nq + 3) / 4),
                    d_flags, nullptr, s);
    }
    SDFX_HIP(hipGetLastError());
}

}  // namespace sdfx
