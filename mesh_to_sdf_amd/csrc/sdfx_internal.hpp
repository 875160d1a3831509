// sdfx_internal.hpp — interfaces between libsdfx's translation units.
#pragma once

#include <string>
#include <vector>

#include "sdfx_geo.hpp"
#include "sdfx_runtime.hpp"

namespace sdfx {

// ---------------------------------------------------------------- mesh prep

// Triangle count of a topology (lib.rs:193-215), or the count error the
// reference's asserts raise.
sdfx_status topology_triangle_count(uint64_t n_vertices, bool has_indices, uint64_t n_indices,
                                    sdfx_topology topology, uint64_t* n_tris);

// Decodes the topology on the device and writes one TriRec per triangle in
// topology order. Indices >= n_vertices raise flags->index_out_of_range (the
// triangle then reads a clamped vertex so no access leaves the buffer).
void gather_triangles(const float* d_vertices, uint64_t n_vertices, const uint32_t* d_indices,
                      sdfx_topology topology, uint64_t n_tris, TriRec* d_tris, DevFlags* d_flags,
                      hipStream_t stream);

// gather_triangles for many meshes in one launch.
struct GatherJob {
    const float* v;
    uint64_t nv;
    const uint32_t* idx;
    sdfx_topology topo;
    uint64_t nt;
    TriRec* out;
    DevFlags* flags;
};
void gather_triangles_batch(const std::vector<GatherJob>& jobs, Arena& arena);

// Raises flags->non_finite_query if any coordinate is NaN or infinite (lib.rs:374-376).
void check_queries_finite(const float* d_queries, uint64_t n_queries, DevFlags* d_flags, hipStream_t stream);

// ---------------------------------------------------------------- None path

// generate/generic/default.rs:22-85: every query against every triangle, in
// triangle order. d_out receives the signed distances.
void none_query(const TriRec* d_tris, uint64_t n_tris, const float* d_queries, uint64_t n_queries,
                sdfx_sign_method sign, float* d_out, DevFlags* d_flags, Arena& arena);

// ---------------------------------------------------------------- LBVH

// Internal node: both children's boxes plus child links. A link >= 0 is an
// internal node, a link < 0 is the leaf ~link (position in Morton order).
// Node indices are Karras's (2012): node i covers the leaf positions
// [first, last] with i at one end; its left child (link or leaf position) is
// the split s, so the left child covers [first, s] and the right one
// [s + 1, last], and a right child's index is its first position.
struct __attribute__((aligned(16))) BvhNode {
    float lmin[3];
    int32_t left;
    float lmax[3];
    int32_t right;
    float rmin[3];
    int32_t last;    // last leaf position the node covers
    float rmax[3];
    int32_t parent;  // internal parent node, -1 at the root
};
static_assert(sizeof(BvhNode) == 64, "BvhNode layout");

struct Lbvh {
    uint64_t n = 0;          // leaves (triangles)
    int32_t root = 0;        // root link (leaf ~0 when n == 1)
    DevBuf<BvhNode> nodes;   // n - 1 internal nodes
    DevBuf<LeafRec> leaves;  // triangles in Morton order (LeafRec::index keeps topology order)
};

// Builds the LBVH: centroid bounds, 30-bit Morton codes, stable LDS radix
// sort, Karras hierarchy, bottom-up refit of the padded triangle boxes.
void lbvh_build(const TriRec* d_tris, uint64_t n_tris, Lbvh& bvh, Arena& arena);

// Builds the LBVHs of many meshes in shared launches (one sort over all of
// them, keyed by mesh then Morton code; with more than 4 meshes the Morton
// code keeps its top 32 - ceil(log2 meshes) bits). Each mesh gets its own node
// and leaf buffers, so it can be freed on its own.
struct LbvhBuildJob {
    const TriRec* tris;  // topology order
    uint64_t n;
    Lbvh* out;
};
void lbvh_build_batch(const std::vector<LbvhBuildJob>& jobs, Arena& arena);

enum QueryMode : int {
    QM_RAYCAST_VOTE = 0,     // min distance (clamped to f32::MAX) + 3-ray vote: Bvh(Raycast)
    QM_NORMAL_FOLD = 1,      // compare_distances fold: Bvh(Normal)
    QM_NEAREST_SIGNED = 2,   // nearest by d^2, signed distance to it: Rtree
    QM_NEAREST_VOTE = 3,     // nearest by d^2, distance to it + 3-ray vote: RtreeBvh
    QM_CLOSEST_VECTOR = 4,   // nearest by d^2, vector to its closest point (query_vector_to_closest_point)
};

void lbvh_query(const Lbvh& bvh, const float* d_queries, uint64_t n_queries, QueryMode mode, float* d_out,
                DevFlags* d_flags, Arena& arena);

// One request of a batched point query: queries d_q[0, nq) against the mesh
// whose LBVH is `bvh`, answers to d_out, flags d_flags[flag].
struct PointBatchItem {
    const Lbvh* bvh;
    const float* d_q;
    uint64_t nq;
    float* d_out;
    uint32_t flag;
};
// lbvh_query for many (mesh, queries) requests of one mode in shared launches
// (the queries of all requests sorted by request, then Morton code); each
// answer is the one lbvh_query gives.
void lbvh_query_batch(const std::vector<PointBatchItem>& items, QueryMode mode, DevFlags* d_flags, Arena& arena);

// generic/rtree_bvh.rs:217-276: closest - query and the 3-ray inside vote.
void lbvh_closest_vectors(const Lbvh& bvh, const float* d_queries, uint64_t n_queries, float* d_vectors,
                          uint8_t* d_inside, DevFlags* d_flags, Arena& arena);

// Exhaustive forms of the LBVH query modes (bit-identical results), used when
// queries x triangles is small enough that no acceleration structure pays.
void exhaustive_query(const TriRec* d_tris, uint64_t n_tris, const float* d_queries, uint64_t n_queries,
                      QueryMode mode, float* d_out, DevFlags* d_flags, Arena& arena);

// ---------------------------------------------------------------- mesh check

// generate/mesh_check.rs:24-81 on host arrays: kind 0 (Ok), 1 (face count),
// 2 (duplicate vertex), 3 (duplicate face), 4 (duplicate edge), 5 (edge without
// reverse); `where` is the vertex, face or edge (3 * face + k) index; message
// is the reference's Err string.
struct MeshCheck {
    int kind = 0;
    uint64_t where = 0;
    std::string message;
};
MeshCheck check_mesh_triangle_list(const float* vertices, uint64_t n_vertices, const uint32_t* faces,
                                   uint64_t n_face_indices, hipStream_t stream);

// Stable LSD radix sort of (keys, vals) by the low `bits` bits of keys; the
// result ends in keys / vals.
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint64_t n, int bits, Arena& A);

// ---------------------------------------------------------------- tuning

// Process-wide knobs (sdfx_set_tuning); none changes a result.
// exhaustive_max_pairs: queries x triangles at or below which the accelerated
// modes run exhaustively. grid_prop_blocks: blocks of each grid propagation
// launch (rounded down to a multiple of the 8 XCDs).
struct Tuning {
    int64_t exhaustive_max_pairs;
    int64_t grid_prop_blocks;
    int64_t lbvh_visit_budget;  // node visits after which a nearest search moves to the per-wave pass (0: by mesh size)
    int64_t grid_batch_prop_blocks;  // grid_prop_blocks for launches that serve several requests
    // rounds after one that took at least this many cells propagate item-major
    // (k_frontier_propagate_items, measured faster on batches whose rounds
    // reach millions of items, slower on a 128^3 grid); 0: every round
    int64_t grid_item_major_min;
};
Tuning& tuning();

// ---------------------------------------------------------------- grid

struct GridParams {
    float first[3];
    float size[3];
    uint64_t count[3];
};

// generate/grid.rs:265-378 on the device: narrow-band seeding, frontier
// propagation to a fixpoint, grid-aligned raycast sign (Raycast only).
void grid_sdf(const TriRec* d_tris, uint64_t n_tris, const GridParams& g, sdfx_sign_method sign, float* d_out,
              DevFlags* d_flags, Arena& arena);

// One request of a batched generate_grid_sdf: its n_tris triangles (in its
// topology order) are d_tris[0, n_tris), numbered [tri_off, tri_off + n_tris)
// in the batch (the requests' ranges are disjoint and increasing); its flags are
// d_flags[flag].
struct GridBatchItem {
    const TriRec* d_tris;
    uint64_t tri_off, n_tris;
    GridParams gp;
    float* d_out;
    uint32_t flag;
    // grid_sdf_exact: each cell's distance from the point queries (|d| for
    // Raycast, the signed value for Normal), which stands in for seeding and
    // propagation; the ray stage and the output are as for any request
    const float* exact = nullptr;
};
// All requests (one sign method) through shared launches, in runs of at most
// about 2^28 cells each; each result is bit-identical to grid_sdf on that
// request alone. With run_errors, a run that fails records its error for its
// requests (run_errors[i] for items[i]) and the others still run; without,
// the error is thrown.
// The schedule-free grid: each cell's distance is the FP32 minimum over all
// triangles, found by the LBVH nearest search at the cell centre (Normal: the
// compare_distances fold of Bvh(Normal)); Raycast signs come from the same
// grid raycasts as grid_sdf's. No counterpart in the reference, whose flood
// fill only approaches these distances, by an amount that depends on its
// thread count.
void grid_sdf_exact(const TriRec* d_tris, uint64_t nt, const GridParams& gp, sdfx_sign_method sign, float* d_out,
                    DevFlags* d_flags, Arena& arena);
void grid_sdf_batch(const std::vector<GridBatchItem>& items, sdfx_sign_method sign, DevFlags* d_flags,
                    Arena& arena, std::vector<std::pair<sdfx_status, std::string>>* run_errors);

// ---------------------------------------------------------------- utilities

// Adds v to out[blockIdx.x] once per block (block-wide sum in LDS). Each block
// owns its slot, so the counters of consecutive launches on one stream need
// no atomics; the host sums the slots.
// Element i of a table that the running kernel does not write, with i the same
// in every active lane: read as constant memory, so the loads go through the
// scalar cache (s_load) instead of one vector load per word and lane. Words
// the caller does not use are dropped by the compiler.
template <class T>
__device__ __forceinline__ T wave_uniform_load(const T* table, uint32_t i) {
    using ConstWord = const __attribute__((address_space(4))) uint32_t;
    static_assert(sizeof(T) % 4 == 0, "read as 32-bit words");
    ConstWord* w = (ConstWord*)(table + i);
    T r;
    uint32_t* d = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) d[k] = w[k];
    return r;
}

__device__ __forceinline__ void block_count(unsigned long long* out, uint32_t v) {
    __shared__ unsigned long long sdfx_block_sum;
    if (threadIdx.x == 0) sdfx_block_sum = 0;
    __syncthreads();
    unsigned long long w = v;
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
    if ((threadIdx.x & 63u) == 0 && w) atomicAdd(&sdfx_block_sum, w);
    __syncthreads();
    if (threadIdx.x == 0 && sdfx_block_sum) out[blockIdx.x] += sdfx_block_sum;
}
// Folds m into *out with atomicMin unless *out already holds a value <= m. The
// words only decrease, so a stale read can only cost an unneeded atomic; the
// read keeps thousands of blocks from queueing atomics on one address.
__device__ __forceinline__ void atomic_min_if_lower(uint32_t* out, uint32_t m) {
    if (m < __hip_atomic_load(out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(out, m);
}
// MI355X dispatches the blocks of a launch to its 8 XCDs round-robin (block b
// runs on XCD b % 8), and each XCD has its own L2. xcd_chunk_block() renumbers
// the blocks in stretches of `chunk` blocks dealt to the XCDs in turn (XCD x
// takes stretches x, x + 8, ...): each stretch of work laid out in spatial
// order shares one L2, and work that is heavier in one part of the range
// spreads over every XCD. The launch must have xcd_chunk_grid(n, chunk) blocks
// (a multiple of 8 * chunk); results >= n are idle.
constexpr uint32_t XCDS = 8;
__device__ __forceinline__ uint32_t xcd_chunk_block(uint32_t chunk) {
    const uint32_t x = blockIdx.x % XCDS, i = blockIdx.x / XCDS;
    return ((i / chunk) * XCDS + x) * chunk + i % chunk;
}
inline unsigned xcd_chunk_grid(uint64_t blocks, uint32_t chunk) {
    const uint64_t step = (uint64_t)XCDS * chunk;
    return (unsigned)((blocks + step - 1) / step * step);
}
// Block-wide minimum of v folded into *out (nothing when every lane holds
// UINT32_MAX). All lanes of the block must call it.
__device__ __forceinline__ void block_min_u32(uint32_t* out, uint32_t v) {
    __shared__ uint32_t sdfx_wave_min[16];
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
    if ((threadIdx.x & 63u) == 0) sdfx_wave_min[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = UINT32_MAX;
        for (uint32_t w = 0; w < (blockDim.x + 63u) / 64u; ++w) m = min(m, sdfx_wave_min[w]);
        if (m != UINT32_MAX) atomic_min_if_lower(out, m);
    }
    __syncthreads();
}
// Sums `n` per-block counter slots on the host (synchronises the stream).
uint64_t sum_block_counters(const unsigned long long* d_slots, uint64_t n, hipStream_t stream);

// Fills n u32 words with v using a kernel (no memset node, so the call is safe
// inside captured graphs).
void fill_u32(uint32_t* d_ptr, uint64_t n, uint32_t v, hipStream_t stream);

// ---------------------------------------------------------------- scans

// Exclusive prefix sum of n values on the device (in place allowed); the total
// goes to *d_total when it is not null. No host synchronisation.
void exclusive_scan_u64(const uint64_t* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total, Arena& arena);
void exclusive_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, uint32_t* d_total, Arena& arena);

}  // namespace sdfx
