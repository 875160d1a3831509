// sdfx_grid.hip — generate_grid_sdf (generate/grid.rs:265-378) on the device,
// for one request or a batch of independent (mesh, grid) requests run through
// the same launches.
//
// Cells hold a 64-bit key:
//   Raycast: bits(d) << 32 | triangle        (d >= 0, so the bits order as floats)
//   Normal:  bits(|d|) << 32 | negative << 31 | triangle
// and keep the minimum under the order in which the reference replaces a
// stored value (compare_distances, lib.rs:315-332), the lower triangle (the
// index within the request's mesh) first among equal distances. For Raycast
// that is the keys' integer order (one atomicMin). In Normal mode it is the
// integer order except that a positive value goes before a negative one it
// approximately equals (key_before); a compare-and-swap loop keeps it. That
// order is a strict total order, so the minimum does not depend on the order
// in which candidates arrive.
//
// 1. Seeding (generate/grid.rs:383-457): every (triangle, cell) pair of each
//    triangle's snapped, widened padded box, one lane per pair (the pair list
//    is a prefix sum over the box volumes, so large triangles cannot starve a
//    lane).
// 2. Propagation (generate/grid.rs:495-558): the reference's relaxation rule
//    (a cell's triangle is tried on its neighbours and kept on strict
//    improvement) run as rounds. Round r takes the flagged cells of each
//    request that lie in that request's distance band, snapshots (cell,
//    triangle) and tries each on its 26 neighbours; the snapshot makes each
//    round, and so the fixpoint, independent of execution order.
// 3. Sign (generate/grid.rs:568-684): the rays start at the cells with index 0
//    along +X/+Y/+Z. Instead of walking a BVH per ray, every triangle visits
//    the rays whose origin lies in its padded box (the rays the reference's
//    BVH traversal would return it for), runs the aligned test and flips bit
//    k = min(floor(t / size), n - 1) of that ray's row. A suffix XOR along each
//    row gives the parity of hits at or beyond each cell, i.e. of the
//    reference's `+1 on cells 0..=k` counters. The 2-of-3 vote is fused into
//    the output write.
//
// Batches: requests share the launches but never each other's state. Each
// request's cells start on a tile boundary, every per-round decision (its band
// minimum, its previous round's size) reads only that request's words, and the
// keys hold mesh-local triangle indices, so every result is bit-identical to
// the same request run alone.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "sdfx_grid_plan.hpp"
#include "sdfx_internal.hpp"

namespace sdfx {

constexpr uint64_t KEY_EMPTY = ~0ull;

struct GridDev {
    float first[3];
    float size[3];
    uint32_t n[3];
    uint32_t words[3];  // u32 words per ray row along each axis
};

// One request of a batch as the kernels see it.
struct GridReq {
    GridDev g;
    const TriRec* tris;     // the request's triangles, in its topology order
    LeafRec* ltris;         // the same as 48-byte records (written by k_seed_boxes, read by propagation)
    uint64_t cell_off;      // first cell in the batch's key array (a multiple of COMPACT_TILE)
    uint64_t tri_off;       // first triangle in the batch's triangle numbering
    uint64_t ntris;
    uint64_t row_off[3];    // first ray-parity word of each axis
    uint64_t row_first[3];  // first ray row of each axis in the batch's row numbering
    float* out;
    uint32_t ncells;
    float band;        // width of the distance band a round takes
    float far;         // a thin round with the band minimum beyond this makes every later round take all
    uint32_t release;  // after a round that took fewer cells, a round takes every flagged cell
    uint32_t flag;     // index of the request's DevFlags
    uint32_t reach;    // bit k: axis k's cells are thin enough for two-step neighbours (grid_reach)
};

// Request lookup. A batch reads its table; a single request travels by value
// in the kernel arguments (scalar registers), which saves a dependent load on
// every access. Kernels take MODE = sign | batch << 1.
template <bool BATCH>
struct ReqView {
    const GridReq* table;
    GridReq one;
    __device__ __forceinline__ const GridReq& operator[](uint32_t q) const {
        if constexpr (BATCH) return table[q];
        else return one;
    }
};
// A batch request that every active lane of the wave shares: the table is
// written before a launch and never during one.
__device__ __forceinline__ GridReq wave_req(const GridReq* table, uint32_t q0) {
    return wave_uniform_load(table, q0);
}
template <int MODE>
constexpr int mode_sign = MODE & 1;
template <int MODE>
constexpr bool mode_batch = (MODE & 2) != 0;

// Calls f(R, q) with request q of the lane. In a batch whose active lanes all
// hold the same request (the usual case: work is laid out request by request)
// the wave reads it once with wave_req; otherwise each lane reads its own.
template <int MODE, class F>
__device__ __forceinline__ void with_req(const ReqView<mode_batch<MODE>>& reqs, uint32_t q, F&& f) {
    if constexpr (mode_batch<MODE>) {
        const uint32_t q0 = __builtin_amdgcn_readfirstlane(q);
        if (__ballot(q != q0) == 0) {
            const GridReq R = wave_req(reqs.table, q0);
            f(R, q0);
        } else {
            f(reqs.table[q], q);
        }
    } else {
        f(reqs.one, 0u);
    }
}

__device__ __forceinline__ float cell_coord(const GridDev& g, int k, uint32_t i) {
    return g.first[k] + (float)i * g.size[k];  // grid.rs:289-295
}

__device__ __forceinline__ uint32_t cell_index(const GridDev& g, uint32_t x, uint32_t y, uint32_t z) {
    return (x * g.n[1] + y) * g.n[2] + z;  // grid.rs:276-278 (below 2^32: checked on the host)
}

// `f32 as isize` (saturating, NaN -> 0) after floor, clamped to [0, n - 1]
// (grid.rs:299-324).
__device__ __forceinline__ uint32_t snap_axis(float c, uint32_t n) {
    float f = floorf(c);
    int64_t i;
    if (f != f) i = 0;
    else if (f >= 9.2233720e18f) i = INT64_MAX;
    else if (f <= -9.2233720e18f) i = INT64_MIN;
    else i = (int64_t)f;
    int64_t hi = (int64_t)n - 1;
    return (uint32_t)(i < 0 ? 0 : (i > hi ? hi : i));
}

// compare_distances(d, f32::MAX) is Less: |d| below f32::MAX (not NaN), and
// for a negative d not within 2 ulps of it (those approximately equal the
// positive f32::MAX and come after it).
__device__ __forceinline__ bool key_valid(float d) {
    const float a = fabsf(d);
    return a < FLT_MAX && ((int32_t)fbits(d) >= 0 || fbits(a) < 0x7f7ffffdu);
}

template <int SIGN>
__device__ __forceinline__ uint64_t make_key(float d, uint32_t tri) {
    if (SIGN == SDFX_SIGN_RAYCAST) return ((uint64_t)fbits(d) << 32) | tri;
    return ((uint64_t)fbits(fabsf(d)) << 32) | ((uint64_t)(fbits(d) >> 31) << 31) | tri;
}

// Whether key a goes strictly before key b (see the top of the file). Valid
// keys all go before KEY_EMPTY (an empty cell holds f32::MAX).
template <int SIGN>
__device__ __forceinline__ bool key_before(uint64_t a, uint64_t b) {
    if (SIGN == SDFX_SIGN_NORMAL && ((a ^ b) >> 31 & 1ull) && a != KEY_EMPTY && b != KEY_EMPTY &&
        approx_eq(bitsf((uint32_t)(a >> 32)), bitsf((uint32_t)(b >> 32))))
        return ((b >> 31) & 1ull) != 0;  // an approximate tie of opposite signs: the positive one first
    return a < b;
}

// Lowers *slot to key when key goes before its value (seen: a recent read of
// it); returns whether it did. Values only ever move earlier in the order, so
// a stale `seen` costs at most a failed compare-and-swap.
template <int SIGN>
__device__ __forceinline__ bool key_lower(uint64_t* slot, uint64_t key, uint64_t seen) {
    if (!key_before<SIGN>(key, seen)) return false;
    if (SIGN == SDFX_SIGN_RAYCAST) return key < atomicMin((unsigned long long*)slot, (unsigned long long)key);
    uint64_t old = seen;
    for (;;) {
        const uint64_t prev = atomicCAS((unsigned long long*)slot, (unsigned long long)old, (unsigned long long)key);
        if (prev == old) return true;
        old = prev;
        if (!key_before<SIGN>(key, old)) return false;
    }
}

// key_lower inside one launch, where `seen` was read in the same launch:
// whether this launch lowered the slot. A Raycast key's atomicMin is issued
// without waiting for its old value. The answer is key < seen, which can
// differ from the atomic's own outcome only when another lane lowered the
// slot below key after seen was read, i.e. in this launch; that lane's
// key < seen holds too, so "some lane of this launch answered true" is
// exactly "this launch lowered the slot". Callers only use the answer to flag
// the slot for the next round (the same set either way), to store a direction
// word (any word is valid) and to fold the key into a running minimum over
// the slots' new values (each lowered slot's final key is answered true by
// the lane that wrote it, and every key answered true is at least its slot's
// final key, so the minimum is the same). Normal keys keep the
// compare-and-swap loop, which needs the old value.
template <int SIGN>
__device__ __forceinline__ bool key_lower_async(uint64_t* slot, uint64_t key, uint64_t seen) {
    if (SIGN != SDFX_SIGN_RAYCAST) return key_lower<SIGN>(slot, key, seen);
    if (!(key < seen)) return false;
    __hip_atomic_fetch_min((unsigned long long*)slot, (unsigned long long)key, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);  // old value unused: the no-return form
    return true;
}

__device__ __forceinline__ uint32_t key_tri(uint64_t k, int sign) {
    return sign == SDFX_SIGN_NORMAL ? (uint32_t)(k & 0x7fffffffu) : (uint32_t)k;
}

__device__ __forceinline__ float key_value(uint64_t k, int sign) {
    if (k == KEY_EMPTY) return FLT_MAX;
    float a = bitsf((uint32_t)(k >> 32));
    return (sign == SDFX_SIGN_NORMAL && ((k >> 31) & 1ull)) ? -a : a;
}

template <int SIGN>
__device__ __forceinline__ float cell_dist(const GridDev& g, uint32_t x, uint32_t y, uint32_t z, const TriRec& t) {
    float px = cell_coord(g, 0, x), py = cell_coord(g, 1, y), pz = cell_coord(g, 2, z);
    return SIGN == SDFX_SIGN_RAYCAST ? tri_dist(px, py, pz, t) : tri_signed_dist(px, py, pz, t);
}

// cell_dist, also returning the vector from the closest point to the cell
// centre (the same FP32 operations as tri_dist / tri_signed_dist).
template <int SIGN>
__device__ __forceinline__ float cell_dist_dir(const GridDev& g, uint32_t x, uint32_t y, uint32_t z, const TriRec& t,
                                               float& dx, float& dy, float& dz) {
    const float px = cell_coord(g, 0, x), py = cell_coord(g, 1, y), pz = cell_coord(g, 2, z);
    float cx, cy, cz;
    closest_diff_flat(px, py, pz, t, dx, dy, dz, cx, cy, cz);
    const float d = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
    if (SIGN == SDFX_SIGN_RAYCAST) return d;
    return dot3(dx, dy, dz, t.nx, t.ny, t.nz) > 0.0f ? d : -d;
}

// ------------------------------------------------ supporting-plane skip

// Direction word of a cell: the vector from its closest point to its centre as
// its last improvement computed it, over its largest component, in three
// signed 10-bit fields (0 when the vector is zero or not finite). It only
// steers how much the bound below skips: any value gives a valid bound, so the
// words need no clearing and racing improvements may leave any of theirs.
__device__ __forceinline__ uint32_t dir_pack(float dx, float dy, float dz) {
    const float m = fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fabsf(dz));
    if (!(m > 0.0f) || !(m <= 3.4e38f)) return 0u;
    const float sc = 511.0f / m;
    const int qx = (int)rintf(dx * sc), qy = (int)rintf(dy * sc), qz = (int)rintf(dz * sc);
    return ((uint32_t)qx & 1023u) | (((uint32_t)qy & 1023u) << 10) | (((uint32_t)qz & 1023u) << 20);
}

// Lower bound of a triangle's distance around an item's cell. For any unit
// vector u the triangle T lies in the half-space u.x <= h, h = max over its
// vertices of u.v, so every point n has d(n, T) >= u.n - h: with n = c + o
// (c the item's cell centre), d(n, T) >= g + u.o, g = u.c - h. g is lowered by
// a margin that covers the FP32 rounding of g, of u.o and of the reference's
// distance, whose closest point lies within a few ulps of T: 2^-16 of the
// largest coordinate involved plus 2^-12 of T's longest edge. A sliver
// (|ab x ac|^2 below 2^-8 |ab|^2 |ac|^2), whose closest point can stray
// further, gets no bound (g = -inf), nor does a zero direction word. A pair
// whose bound exceeds the neighbour's key distance cannot improve it, so
// skipping it changes no result. tests/test_grid_fixpoint.py checks the bound
// on the CPU, with exact and with random directions, on meshes, degenerate
// soups and slivers far from the origin.
struct PlaneBound {
    float ux, uy, uz, g;
};
__device__ __forceinline__ PlaneBound plane_bound(uint32_t w, float cx, float cy, float cz, const TriRec& t,
                                                  float max_size) {
    PlaneBound b{0.0f, 0.0f, 0.0f, -INFINITY};
    float ux = (float)(((int)(w << 22)) >> 22), uy = (float)(((int)(w << 12)) >> 22),
          uz = (float)(((int)(w << 2)) >> 22);
    const float l2 = dot3(ux, uy, uz, ux, uy, uz);
    const float ab2 = dot3(t.abx, t.aby, t.abz, t.abx, t.aby, t.abz);
    const float ac2 = dot3(t.acx, t.acy, t.acz, t.acx, t.acy, t.acz);
    const float bc2 = dot3(t.bcx, t.bcy, t.bcz, t.bcx, t.bcy, t.bcz);
    const float n2 = dot3(t.nx, t.ny, t.nz, t.nx, t.ny, t.nz);
    if (!(l2 > 0.0f) || (t.degen == 0u && !(n2 >= (ab2 * ac2) * (1.0f / 256.0f)))) return b;
    const float inv = 1.0f / sqrtf(l2);
    ux *= inv;
    uy *= inv;
    uz *= inv;
    const float h = fmaxf(dot3(ux, uy, uz, t.ax, t.ay, t.az),
                          fmaxf(dot3(ux, uy, uz, t.bx, t.by, t.bz), dot3(ux, uy, uz, t.cx, t.cy, t.cz)));
    float sc = fmaxf(fmaxf(fabsf(cx), fabsf(cy)), fabsf(cz));
    sc = fmaxf(sc, fmaxf(fmaxf(fabsf(t.ax), fabsf(t.ay)), fabsf(t.az)));
    sc = fmaxf(sc, fmaxf(fmaxf(fabsf(t.bx), fabsf(t.by)), fabsf(t.bz)));
    sc = fmaxf(sc, fmaxf(fmaxf(fabsf(t.cx), fabsf(t.cy)), fabsf(t.cz)));
    sc = sc + 2.0f * max_size;  // the farthest neighbour is at most two cells away
    const float e = sqrtf(fmaxf(ab2, fmaxf(ac2, bc2)));
    const float g = (dot3(ux, uy, uz, cx, cy, cz) - h) - (sc * (1.0f / 65536.0f) + e * (1.0f / 4096.0f));
    b.ux = ux;
    b.uy = uy;
    b.uz = uz;
    b.g = g == g ? g : -INFINITY;
    return b;
}

__global__ void k_fill_u32g(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// One launch that sets a grid run's state: every key empty, then the u32 words
// in runs of (count, value) (the round counters and band words, the ray-parity
// bits).
struct WordRuns {
    uint64_t end[6];  // run k covers words [end[k-1], end[k])
    uint32_t value[6];
};
__global__ void __launch_bounds__(256) k_grid_init(uint64_t* __restrict__ keys, uint64_t ncells,
                                                   uint32_t* __restrict__ words, WordRuns runs) {
    const uint64_t nwords = runs.end[5];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncells + nwords;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < ncells) {
            keys[i] = KEY_EMPTY;
        } else {
            const uint64_t w = i - ncells;
            int k = 0;
            while (w >= runs.end[k]) ++k;  // w < end[5]: k stays below 6
            words[w] = runs.value[k];
        }
    }
}

// ---------------------------------------------------------------- round bookkeeping

constexpr int COMPACT_PER_LANE = 16;
constexpr int COMPACT_TILE = 256 * COMPACT_PER_LANE;
// Width of the distance band a round takes, in units of the largest |cell size|.
constexpr float GRID_BAND_CELLS = 3.0f;
// A thin round (below) whose band minimum lies beyond GRID_FAR_CELLS largest
// |cell size|s makes every later round of its request take every flagged cell:
// in the far field the front carries mostly final values and the band only
// adds rounds.
constexpr float GRID_FAR_CELLS = 8.0f;
// A frontier item offers its triangle to the reference's 26 neighbours and,
// along each axis whose cells are at least 1.5 times thinner than the largest,
// to the two cells two steps away on that axis, so that a front moves about as
// far along a thin axis per round as along the others (more candidates only
// move values closer to the exact distance; the end state still satisfies the
// reference's fixpoint over the 26).
inline bool grid_reach(float max_size, float abs_size) { return max_size / abs_size >= 1.5f; }

// Offset of neighbour k of an item (k < 26: the 26 around it, the centre
// skipped; then two steps both ways along each axis whose bit is set in
// reach); false when the request has fewer neighbours.
__device__ __forceinline__ bool neighbour_offset(uint32_t k, uint32_t reach, int& ox, int& oy, int& oz) {
    if (k < 26u) {
        k += k >= 13u ? 1u : 0u;
        ox = (int)(k / 9u) - 1;
        oy = (int)((k / 3u) % 3u) - 1;
        oz = (int)(k % 3u) - 1;
        return true;
    }
    uint32_t e = k - 26u;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!((reach >> a) & 1u)) continue;
        if (e < 2u) {
            const int d = e ? -2 : 2;
            ox = a == 0 ? d : 0;
            oy = a == 1 ? d : 0;
            oz = a == 2 ? d : 0;
            return true;
        }
        e -= 2u;
    }
    return false;
}
// A round after one that took fewer than max(GRID_RELEASE_MIN, cells /
// GRID_RELEASE_DIV) cells takes every flagged cell: once the front is thin,
// waiting for the band only adds rounds.
constexpr uint64_t GRID_RELEASE_MIN = 1024;
constexpr uint64_t GRID_RELEASE_DIV = 128;
// Device-scope atomics on one address serialise (about 11 ns each on MI355X),
// so shared words are spread: the item counters over NCLASS classes (tile t
// belongs to class t * NCLASS / ntiles, a run of neighbouring tiles, so the
// item buffer stays in spatial order), each request's band minimum and round
// size over req_spread words (64 for a single request, which every block of a
// launch folds into; 16 per request of a batch, whose blocks spread over the
// requests), all WORD_STRIDE u32 (256 bytes) apart.
constexpr uint32_t NCLASS = 64;
template <int MODE>
constexpr uint32_t req_spread = mode_batch<MODE> ? 16u : 64u;
constexpr uint32_t WORD_STRIDE = 64;
// Frontier item coordinates are packed as x | y << 21 | z << 42.
constexpr uint32_t AXIS_XY_LIMIT = 1u << 21;
constexpr uint32_t AXIS_Z_LIMIT = 1u << 22;

// One frontier entry: a cell taken in this round and the mesh-local triangle
// its key held when the round started. A single request stores the cell's
// coordinates as they are, {tri, x, y, z}; a batch adds the request and packs
// the coordinates, {tri, req, x | y << 21 | z << 42 (two words)}.
struct FrontierItem {
    uint32_t w[4];
};

template <int MODE>
__device__ __forceinline__ FrontierItem make_item(uint32_t tri, uint32_t q, uint32_t x, uint32_t y, uint32_t z) {
    if constexpr (mode_batch<MODE>) {
        const uint64_t p = (uint64_t)x | ((uint64_t)y << 21) | ((uint64_t)z << 42);
        return FrontierItem{{tri, q, (uint32_t)p, (uint32_t)(p >> 32)}};
    } else {
        return FrontierItem{{tri, x, y, z}};
    }
}

template <int MODE>
__device__ __forceinline__ void read_item(const FrontierItem& it, uint32_t& tri, uint32_t& q, uint32_t& x, uint32_t& y,
                                          uint32_t& z) {
    tri = it.w[0];
    if constexpr (mode_batch<MODE>) {
        const uint64_t p = ((uint64_t)it.w[3] << 32) | it.w[2];
        q = it.w[1];
        x = (uint32_t)(p & (AXIS_XY_LIMIT - 1));
        y = (uint32_t)((p >> 21) & (AXIS_XY_LIMIT - 1));
        z = (uint32_t)(p >> 42);
    } else {
        q = 0;
        x = it.w[1];
        y = it.w[2];
        z = it.w[3];
    }
}

// Word s (taken modulo the spread) of request q in a per-request spread array.
template <int MODE>
__device__ __forceinline__ uint32_t* req_word(uint32_t* base, uint32_t q, uint32_t s) {
    return base + ((uint64_t)q * req_spread<MODE> + s % req_spread<MODE>) * WORD_STRIDE;
}

// Minimum and sum over the spread words of request q, computed by every wave
// on its own (lane l < req_spread holds word l).
template <int MODE>
__device__ __forceinline__ uint32_t req_min(const uint32_t* base, uint32_t q) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t m = lane < req_spread<MODE> ? base[((uint64_t)q * req_spread<MODE> + lane) * WORD_STRIDE] : UINT32_MAX;
    for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off, 64));
    return m;
}
template <int MODE>
__device__ __forceinline__ uint32_t req_sum(const uint32_t* base, uint32_t q) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t m = lane < req_spread<MODE> ? base[((uint64_t)q * req_spread<MODE> + lane) * WORD_STRIDE] : 0u;
    for (int off = 32; off > 0; off >>= 1) m += (uint32_t)__shfl_xor((int)m, off, 64);
    return m;
}

// Folds each lane's (request q, value m) into that request's minimum words;
// UINT32_MAX means "nothing". A wave whose lanes share one request makes one
// atomic (the usual case: work is laid out request by request). All lanes of
// the wave must call it.
template <int MODE>
__device__ __forceinline__ void fold_req_min(uint32_t* base, uint32_t q, uint32_t m, uint32_t slot) {
    const uint64_t has = __ballot(m != UINT32_MAX);
    if (!has) return;
    const uint32_t q0 = (uint32_t)__shfl((int)q, __ffsll((long long)has) - 1, 64);
    if (m == UINT32_MAX) q = q0;
    if (__all(q == q0)) {
        for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off, 64));
        if ((threadIdx.x & 63u) == 0) atomic_min_if_lower(req_word<MODE>(base, q0, slot), m);
    } else if (m != UINT32_MAX) {
        atomic_min_if_lower(req_word<MODE>(base, q, slot), m);
    }
}

// ---------------------------------------------------------------- seeding

struct SeedBox {
    uint32_t mn[3];
    uint32_t ext[3];  // mx - mn + 1
};

// generate/grid.rs:400-426: padded box, snapped corners, widened by one cell
// where the snapped corner's centre lies beyond the box.
template <int MODE>
__global__ void __launch_bounds__(256) k_seed_boxes(uint64_t nt, ReqView<mode_batch<MODE>> reqs,
                                                    const uint32_t* __restrict__ tri_req,
                                                    SeedBox* __restrict__ boxes, uint64_t* __restrict__ counts) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t q = mode_batch<MODE> ? tri_req[t] : 0u;
    if (q == UINT32_MAX) {  // a triangle of no request in this run
        counts[t] = 0;
        return;
    }
    GridDev g;
    TriRec tr;
    with_req<MODE>(reqs, q, [&](const GridReq& R, uint32_t) {
        g = R.g;
        tr = R.tris[t - R.tri_off];
        R.ltris[t - R.tri_off] = make_leafrec(tr);
    });
    Box b;
    tri_bbox(tr.ax, tr.ay, tr.az, tr.bx, tr.by, tr.bz, tr.cx, tr.cy, tr.cz, b);
    uint32_t mn[3], mx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float bmin = g.first[k] - g.size[k] * 0.5f;  // grid.rs:264-265
        mn[k] = snap_axis((b.lo[k] - bmin) / g.size[k], g.n[k]);
        mx[k] = snap_axis((b.hi[k] - bmin) / g.size[k], g.n[k]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (mn[k] > 0 && cell_coord(g, k, mn[k]) > b.lo[k]) mn[k] -= 1;
        if (mx[k] < g.n[k] - 1 && cell_coord(g, k, mx[k]) < b.hi[k]) mx[k] += 1;
    }
    SeedBox sb;
    uint64_t cnt = 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        sb.mn[k] = mn[k];
        sb.ext[k] = mx[k] >= mn[k] ? mx[k] - mn[k] + 1 : 0;
        cnt *= sb.ext[k];
    }
    boxes[t] = sb;
    counts[t] = cnt;
}

// The last t in [lo, hi] with offsets[t] <= target (offsets[lo] <= target),
// searched by a whole wave: each step tests 64 evenly spaced positions, so a
// range of n entries takes log64(n) dependent loads. All lanes of the wave
// must call it with the same arguments; every lane gets the result.
__device__ __forceinline__ uint64_t wave_search(const uint64_t* __restrict__ offsets, uint64_t lo, uint64_t hi,
                                                uint64_t target) {
    const uint32_t lane = threadIdx.x & 63u;
    while (hi > lo) {
        const uint64_t step = (hi - lo + 64) / 64;  // ceil((hi - lo + 1) / 64)
        const uint64_t q = lo + lane * step;
        const bool pred = q <= hi && offsets[q] <= target;  // true for lane 0
        const uint64_t m = __ballot(pred);
        const uint32_t last = 63u - (uint32_t)__clzll((long long)m);
        const uint64_t nlo = lo + last * step;
        hi = min(hi, nlo + step - 1);
        lo = nlo;
    }
    return lo;
}

// Seeding works in chunks of SEED_CHUNK consecutive (triangle, cell) pairs: a
// wave finds the chunk's first and last triangles, the block loads their pair
// offsets into LDS and each lane finds its pair's triangle there (a global
// search only when zero-pair triangles make the range longer than the chunk).
constexpr uint32_t SEED_CHUNK = 2048;

template <int MODE>
__global__ void __launch_bounds__(256) k_seed_pairs(uint64_t nt, ReqView<mode_batch<MODE>> reqs,
                                                    const uint32_t* __restrict__ tri_req,
                                                    const SeedBox* __restrict__ boxes,
                                                    const uint64_t* __restrict__ offsets,
                                                    const uint64_t* __restrict__ total_ptr, uint64_t* keys,
                                                    uint32_t* __restrict__ dirs, DevFlags* flags,
                                                    unsigned long long* evals) {
    constexpr int SIGN = mode_sign<MODE>;
    __shared__ uint64_t s_off[SEED_CHUNK];
    __shared__ uint64_t s_t0, s_t1;
    const uint64_t total = *total_ptr;
    uint32_t n_evals = 0;
    for (uint64_t c0 = (uint64_t)blockIdx.x * SEED_CHUNK; c0 < total; c0 += (uint64_t)gridDim.x * SEED_CHUNK) {
        const uint64_t c1 = min(total, c0 + SEED_CHUNK);
        if (threadIdx.x < 64u) {
            const uint64_t t0 = wave_search(offsets, 0, nt - 1, c0);
            const uint64_t t1 = wave_search(offsets, t0, nt - 1, c1 - 1);
            if (threadIdx.x == 0) {
                s_t0 = t0;
                s_t1 = t1;
            }
        }
        __syncthreads();
        const uint64_t t0 = s_t0, span = s_t1 - t0 + 1;
        const bool in_lds = span <= SEED_CHUNK;  // block-uniform
        if (in_lds)
            for (uint64_t k = threadIdx.x; k < span; k += blockDim.x) s_off[k] = offsets[t0 + k];
        __syncthreads();
        for (uint64_t p = c0 + threadIdx.x; p < c1; p += blockDim.x) {
            // last triangle whose offset is <= p (the one whose pairs hold p)
            uint64_t lo = 0, hi = span - 1;
            while (lo < hi) {
                const uint64_t mid = (lo + hi + 1) >> 1;
                if ((in_lds ? s_off[mid] : offsets[t0 + mid]) <= p) lo = mid;
                else hi = mid - 1;
            }
            const uint64_t t = t0 + lo;
            const uint64_t off_t = in_lds ? s_off[lo] : offsets[t];
            const SeedBox sb = boxes[t];
            with_req<MODE>(reqs, mode_batch<MODE> ? tri_req[t] : 0u, [&](const GridReq& R, uint32_t) {
                uint64_t local = p - off_t;
                uint64_t yz = (uint64_t)sb.ext[1] * sb.ext[2];
                if (yz == 0 || local >= yz * sb.ext[0]) {  // cannot happen with a consistent prefix sum
                    atomicOr(&flags[R.flag].internal, 1u);
                    return;
                }
                uint32_t x = sb.mn[0] + (uint32_t)(local / yz);
                uint64_t r = local % yz;
                uint32_t y = sb.mn[1] + (uint32_t)(r / sb.ext[2]);
                uint32_t z = sb.mn[2] + (uint32_t)(r % sb.ext[2]);
                // the 48-byte copy k_seed_boxes wrote: three loads instead of six, the same values
                const TriRec tr = expand_leaf(R.ltris[t - R.tri_off]);
                float dx, dy, dz;
                float d = cell_dist_dir<SIGN>(R.g, x, y, z, tr, dx, dy, dz);
                n_evals++;
                if (d != d) atomicOr(&flags[R.flag].nan_distance, 1u);
                if (key_valid(d)) {
                    uint64_t key = make_key<SIGN>(d, (uint32_t)(t - R.tri_off));
                    const uint64_t ci = R.cell_off + cell_index(R.g, x, y, z);
                    if (key_lower_async<SIGN>(&keys[ci], key, keys[ci])) dirs[ci] = dir_pack(dx, dy, dz);
                }
            });
        }
        __syncthreads();  // s_off and s_t0 / s_t1 are reused by the next chunk
    }
    if (evals) block_count(evals, n_evals);
}

// ---------------------------------------------------------------- exact values

// The cell centres of one grid as query points, cell i = (x * ny + y) * nz + z.
__global__ void k_cell_centres(GridDev g, uint64_t ncells, float* __restrict__ q) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncells;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t z = (uint32_t)(i % g.n[2]);
        const uint64_t r = i / g.n[2];
        q[3 * i] = cell_coord(g, 0, (uint32_t)(r / g.n[1]));
        q[3 * i + 1] = cell_coord(g, 1, (uint32_t)(r % g.n[1]));
        q[3 * i + 2] = cell_coord(g, 2, z);
    }
}

// Keys of a single request from per-cell values (the triangle field is not
// read by the finish).
template <int MODE>
__global__ void k_keys_from_values(const float* __restrict__ values, uint64_t ncells, uint64_t* __restrict__ keys,
                                   DevFlags* flags) {
    constexpr int SIGN = mode_sign<MODE>;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncells;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float d = values[i];
        if (SIGN == SDFX_SIGN_RAYCAST) d = fabsf(d);
        if (d != d) atomicOr(&flags->nan_distance, 1u);
        keys[i] = key_valid(d) ? make_key<SIGN>(d, 0u) : KEY_EMPTY;
    }
}

// ---------------------------------------------------------------- propagation

// Flags every seeded cell (round 0 takes the seeds) and folds each request's
// smallest seeded distance into its minimum words.
template <int MODE>
__global__ void __launch_bounds__(256) k_mark_seeded(const uint64_t* __restrict__ keys, uint64_t ncells,
                                                     const uint32_t* __restrict__ tile_req,
                                                     uint8_t* __restrict__ changed, uint8_t* __restrict__ tile_changed,
                                                     uint32_t* __restrict__ min_out) {
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < ncells; base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = base + threadIdx.x;  // the loop bounds are the same for every lane
        uint32_t m = UINT32_MAX, q = 0;
        if (i < ncells) {
            const uint64_t k = keys[i];
            const bool seeded = k != KEY_EMPTY;
            changed[i] = seeded ? 1 : 0;
            if (mode_batch<MODE>) q = tile_req[i / COMPACT_TILE];
            if (seeded) m = (uint32_t)(k >> 32);
            if (i % COMPACT_TILE == 0) tile_changed[i / COMPACT_TILE] = 1;  // every tile is scanned once
        }
        if (mode_batch<MODE>) fold_req_min<MODE>(min_out, q, m, blockIdx.x);
        else block_min_u32(req_word<MODE>(min_out, 0, blockIdx.x), m);
    }
}

// Snapshot of one round's frontier. Tile t belongs to request q; the round
// takes every flagged cell of q whose distance (the high word of its key,
// non-negative float bits) is at most min_q + band_q, where min_q is the
// smallest distance among q's flagged cells (the minimum of q's words in
// min_cur, complete when the round starts), or every flagged cell of q when
// q's previous round took fewer than release_q cells (the sum of q's words in
// taken_prev), and in every round of q once a thin round (one after a round
// that took fewer than release_q cells) has happened beyond far_q or for the
// second time (released[q]; thin_round[q] holds round_tag, this round's
// number plus one, of q's first thin round). Each taken cell becomes an item
// carrying the triangle of its current key and its flag is cleared. Flagged cells beyond the band keep
// their flags; their smallest distance is folded into min_next and the number
// of taken cells into taken_cur.
//
// Each block takes a contiguous tile of COMPACT_TILE cells (16 per lane, flags
// read as one 16-byte vector), ranks the taken cells with a block-wide prefix
// sum and reserves its output range in its class's region with a single
// atomic, so items stay in cell order within a tile. Also resets the counters
// the next compaction appends to.
template <int MODE>
__global__ void __launch_bounds__(256) k_frontier_compact(const uint64_t* __restrict__ keys,
                                                          ReqView<mode_batch<MODE>> reqs,
                                                          const uint32_t* __restrict__ tile_req, uint32_t ntiles,
                                                          uint8_t* __restrict__ changed,
                                                          uint8_t* __restrict__ tile_changed,
                                                          FrontierItem* __restrict__ items, uint32_t region,
                                                          uint32_t* __restrict__ ctr_cur,
                                                          uint32_t* __restrict__ ctr_next,
                                                          const uint32_t* __restrict__ min_cur,
                                                          uint32_t* __restrict__ min_next,
                                                          const uint32_t* __restrict__ taken_prev,
                                                          uint32_t* __restrict__ taken_cur,
                                                          uint32_t* __restrict__ released,
                                                          uint32_t* __restrict__ thin_round, uint32_t round_tag,
                                                          int sign) {
    __shared__ uint32_t wave_tot[4], wave_min[4];
    __shared__ uint32_t block_base, band_min, prev_taken, far_released;
    if (blockIdx.x == 0 && threadIdx.x < NCLASS) ctr_next[threadIdx.x * WORD_STRIDE] = 0u;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // Grids of up to 4,096 tiles read each lane's 16 cell flags together with
    // the tile's flag instead of after it: one dependent trip less per round,
    // for a few MB of flag reads that stay in L2. Larger ones read the flags of
    // flagged tiles only.
    const bool early_flags = ntiles <= 4096u;
    for (uint32_t tile_id = blockIdx.x; tile_id < ntiles; tile_id += gridDim.x) {
        // this lane's 16 flags: cells tile_id * COMPACT_TILE + 16 * threadIdx.x onward
        const uint64_t g0 = (uint64_t)tile_id * COMPACT_TILE + threadIdx.x * COMPACT_PER_LANE;
        const uint8_t tile_flag = tile_changed[tile_id];
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (early_flags) v = *reinterpret_cast<const uint4*>(changed + g0);
        // a single request's band words are known before the tile flag too
        // (released[0] may be set by another block this round: only in a thin
        // round, which takes every flagged cell either way)
        const bool early_words = early_flags && !mode_batch<MODE>;
        uint32_t m0 = 0, p0 = 0, r0 = 0;
        if (early_words && wave == 0) {
            m0 = req_min<MODE>(min_cur, 0u);
            p0 = req_sum<MODE>(taken_prev, 0u);
            r0 = released[0];
        }
        if (!tile_flag) continue;  // block-uniform: no flagged cell in the tile
        const uint32_t q = mode_batch<MODE> ? tile_req[tile_id] : 0u;
        // the tile's request, the same in every lane: read through the scalar cache
        const GridReq R = [&] {
            if constexpr (mode_batch<MODE>) return wave_req(reqs.table, __builtin_amdgcn_readfirstlane(q));
            else return reqs.one;
        }();
        const uint32_t local0 = (uint32_t)((uint64_t)tile_id * COMPACT_TILE - R.cell_off);
        const uint32_t c0 = local0 + threadIdx.x * COMPACT_PER_LANE;  // first cell of this lane, request-local
        // this lane's 16 cell flags (whole tiles are allocated), unless read above
        if (!early_flags) v = *reinterpret_cast<const uint4*>(changed + g0);
        // the request's band words, shared through LDS (already loaded above for
        // a small single grid: then these are LDS writes only)
        if (wave == 0) {
            const uint32_t m = early_words ? m0 : req_min<MODE>(min_cur, q);
            const uint32_t p = early_words ? p0 : req_sum<MODE>(taken_prev, q);
            if (lane == 0) {
                band_min = m;
                prev_taken = p;
                far_released = early_words ? r0 : released[q];
            }
        }
        uint32_t flagged = 0;  // bit k: cell c0 + k is flagged
        {
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < COMPACT_PER_LANE; ++k) flagged |= ((w[k >> 2] >> (8 * (k & 3))) & 1u) << k;
            if (c0 >= R.ncells) flagged = 0;
            else if (R.ncells - c0 < (uint32_t)COMPACT_PER_LANE) flagged &= (1u << (R.ncells - c0)) - 1u;
        }
        // every flagged cell's key, issued together and before the barrier
        // below, so they are in flight while the band words are shared
        uint64_t kv[COMPACT_PER_LANE];
#pragma unroll
        for (int k = 0; k < COMPACT_PER_LANE; ++k) kv[k] = ((flagged >> k) & 1u) ? keys[g0 + k] : KEY_EMPTY;
        __syncthreads();
        const float mf = bitsf(band_min);
        const bool thin = prev_taken < R.release;
        // Every block taking a tile of q this round decides the same; the flag
        // is read by the later rounds' compactions.
        if (thin && !far_released && threadIdx.x == 0) {
            if (mf > R.far) released[q] = 1u;
            // The first thin round records itself (every block of it writes
            // the same tag; a block may read the tag another block of this
            // round wrote, which is no later round); a later one releases q.
            const uint32_t t = thin_round[q];
            if (t == 0u) thin_round[q] = round_tag;
            else if (t != round_tag) released[q] = 1u;
        }
        float limit = mf + R.band;
        // a NaN band (or no flagged cell), a thin previous front, or the far field: every flagged cell
        if (!(limit >= mf) || thin || far_released) limit = INFINITY;
        const uint32_t lbits = fbits(limit);  // non-negative floats order as their bits
        uint32_t take = 0, dmin = UINT32_MAX;
#pragma unroll
        for (int k = 0; k < COMPACT_PER_LANE; ++k) {
            if (!((flagged >> k) & 1u)) continue;
            const uint32_t hi = (uint32_t)(kv[k] >> 32);
            if (hi <= lbits) take |= 1u << k;
            else dmin = min(dmin, hi);
        }
        const uint32_t deferred = flagged & ~take;
        const uint32_t mine = __popc(take);
        // block-wide exclusive prefix of `mine`, block minimum of `dmin`
        uint32_t x = mine;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(x, off, 64);
            if ((int)lane >= off) x += y;
        }
        uint32_t wm = dmin;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) wm = min(wm, (uint32_t)__shfl_xor((int)wm, off, 64));
        if (lane == 63) wave_tot[wave] = x;
        if (lane == 0) wave_min[wave] = wm;
        __syncthreads();
        uint32_t before = 0, total = 0, bmin = UINT32_MAX;
#pragma unroll
        for (int qw = 0; qw < 4; ++qw) {
            before += (uint32_t)qw < wave ? wave_tot[qw] : 0u;
            total += wave_tot[qw];
            bmin = min(bmin, wave_min[qw]);
        }
        const uint32_t cls = (uint32_t)((uint64_t)tile_id * NCLASS / ntiles);
        if (threadIdx.x == 0) block_base = total ? atomicAdd(ctr_cur + cls * WORD_STRIDE, total) : 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            tile_changed[tile_id] = bmin != UINT32_MAX ? 1 : 0;  // deferred cells keep the tile flagged
            // nothing waits for these
            if (bmin != UINT32_MAX) atomicMin(req_word<MODE>(min_next, q, tile_id), bmin);
            if (total) atomicAdd(req_word<MODE>(taken_cur, q, tile_id), total);
        }
        if (take) {
            uint64_t slot = (uint64_t)cls * region + block_base + before + x - mine;
#pragma unroll
            for (int k = 0; k < COMPACT_PER_LANE; ++k) {
                if (!((take >> k) & 1u)) continue;
                const uint32_t i = c0 + k;
                const uint32_t z = i % R.g.n[2];
                const uint32_t r = i / R.g.n[2];
                items[slot++] = make_item<MODE>(key_tri(kv[k], sign), q, r / R.g.n[1], r % R.g.n[1], z);
            }
            uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int k = 0; k < COMPACT_PER_LANE; ++k) w[k >> 2] |= ((deferred >> k) & 1u) << (8 * (k & 3));
            *reinterpret_cast<uint4*>(changed + g0) = make_uint4(w[0], w[1], w[2], w[3]);
        }
        __syncthreads();  // wave_tot / wave_min / block_base / band_min are reused by the next tile
    }
}

// Per-wave queue of (neighbour, triangle) pairs that passed the cheap tests: a
// wave fills it 64 candidates at a time and evaluates 64 queued pairs whenever
// it holds that many, so the distance evaluations run with every lane busy
// instead of idling the lanes whose candidate was skipped.
constexpr uint32_t PROP_QUEUE = 128;  // a fill adds at most 64 to at most 63 left over
// Frontier items one propagation launch takes: kmax neighbour slots each plus
// one grid stride stay below 2^32.
__host__ __device__ constexpr uint32_t prop_chunk_items(uint32_t kmax) {
    return (uint32_t)(((1ull << 32) - (1ull << 24)) / kmax);
}
constexpr uint32_t PROP_WAVES = 4;    // 256-thread blocks

template <bool BATCH>
struct PropQueue {
    uint32_t x[PROP_WAVES][PROP_QUEUE], y[PROP_WAVES][PROP_QUEUE], z[PROP_WAVES][PROP_QUEUE];
    uint32_t tri[PROP_WAVES][PROP_QUEUE], req[PROP_WAVES][BATCH ? PROP_QUEUE : 1];
    uint32_t cur_lo[PROP_WAVES][PROP_QUEUE], cur_hi[PROP_WAVES][PROP_QUEUE];
};

// The queue of k_frontier_propagate_items: a pair is its item's lane in the
// block (the item, its triangle and bound sit in LDS) and its neighbour slot,
// src << 8 | k, with the neighbour's key read by the candidate stage.
struct PropQueueRef {
    uint32_t sk[PROP_WAVES][PROP_QUEUE], cur_lo[PROP_WAVES][PROP_QUEUE], cur_hi[PROP_WAVES][PROP_QUEUE];
};

// LDS traffic between lanes of one wave: keep the compiler from moving the
// accesses across this point (the wave's LDS operations execute in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-lane running minimum of improved distances for one request at a time.
struct PropAcc {
    uint32_t n_evals = 0, imin = UINT32_MAX, req = 0;
};

// Evaluates one queued pair: the reference's distance of cell (x, y, z) of
// request q to its mesh-local triangle tri, kept if its key is below the
// neighbour's current key.
template <int MODE>
__device__ __forceinline__ void prop_eval_tri(const GridReq& R, const TriRec& tr, uint64_t* keys,
                                              uint32_t* __restrict__ dirs, uint8_t* __restrict__ changed,
                                              uint8_t* __restrict__ tile_changed, uint32_t* __restrict__ min_next,
                                              DevFlags* flags, uint32_t x, uint32_t y, uint32_t z, uint32_t tri,
                                              uint32_t q, uint64_t cur, PropAcc& acc) {
    constexpr int SIGN = mode_sign<MODE>;
    const float px = cell_coord(R.g, 0, x), py = cell_coord(R.g, 1, y), pz = cell_coord(R.g, 2, z);
    float dx, dy, dz, cx, cy, cz;
    closest_diff_flat(px, py, pz, tr, dx, dy, dz, cx, cy, cz);
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    acc.n_evals++;
    if (cur != KEY_EMPTY) {
        // The key holds |d| = c. sqrt is monotone and correctly rounded, so
        // d2 >= (next float above c)^2 gives |d| > c: no improvement, and the
        // square root and sign can be skipped. (c is kept in the range where
        // these squares are normal floats.) In Normal mode a positive value
        // still goes before a negative c it approximately equals: within
        // 1e-6 (plus the rounding of the difference) or 2 ulps above it, so
        // the bound is raised past both.
        uint32_t cb = (uint32_t)(cur >> 32);
        if (SIGN == SDFX_SIGN_NORMAL && ((cur >> 31) & 1ull))
            cb = max(fbits(bitsf(cb) + 2e-6f), cb + 2u);
        if (cb >= 0x26800000u && cb < 0x5f000000u) {  // 2^-50 <= c < 2^63
            const float cp = bitsf(cb + 1u);
            const float lim = bitsf(fbits(cp * cp) + 1u);  // >= the exact square of cp
            if (d2 >= lim) return;
        }
    }
    const float dist = sqrtf(d2);
    float d = dist;
    if (SIGN == SDFX_SIGN_NORMAL) d = dot3(dx, dy, dz, tr.nx, tr.ny, tr.nz) > 0.0f ? dist : -dist;
    if (d != d) atomicOr(&flags[R.flag].nan_distance, 1u);
    if (!key_valid(d)) return;
    const uint64_t key = make_key<SIGN>(d, tri);
    const uint64_t ci = R.cell_off + cell_index(R.g, x, y, z);
    if (key_lower_async<SIGN>(&keys[ci], key, cur)) {
        dirs[ci] = dir_pack(dx, dy, dz);
        changed[ci] = 1;
        tile_changed[ci / COMPACT_TILE] = 1;
        if (mode_batch<MODE> && q != acc.req && acc.imin != UINT32_MAX) {  // the lane moved on to another request
            atomic_min_if_lower(req_word<MODE>(min_next, acc.req, blockIdx.x), acc.imin);
            acc.imin = UINT32_MAX;
        }
        if (mode_batch<MODE>) acc.req = q;
        acc.imin = min(acc.imin, (uint32_t)(key >> 32));
    }
}

// prop_eval_tri with the triangle read from the request's 48-byte copies.
template <int MODE>
__device__ __forceinline__ void prop_eval_req(const GridReq& R, uint64_t* keys, uint32_t* __restrict__ dirs,
                                              uint8_t* __restrict__ changed,
                                              uint8_t* __restrict__ tile_changed, uint32_t* __restrict__ min_next,
                                              DevFlags* flags, uint32_t x, uint32_t y, uint32_t z, uint32_t tri,
                                              uint32_t q, uint64_t cur, PropAcc& acc) {
    const TriRec tr = expand_leaf(R.ltris[tri]);  // three 16-byte loads instead of six, the same values
    prop_eval_tri<MODE>(R, tr, keys, dirs, changed, tile_changed, min_next, flags, x, y, z, tri, q, cur, acc);
}

// prop_eval_req for lane's request q. In a batch the wave's lanes nearly always
// share one request (items lie request by request); the wave then reads it
// once through the scalar cache.
template <int MODE>
__device__ __forceinline__ void prop_eval(const ReqView<mode_batch<MODE>>& reqs,
                                          uint64_t* keys, uint32_t* __restrict__ dirs, uint8_t* __restrict__ changed,
                                          uint8_t* __restrict__ tile_changed, uint32_t* __restrict__ min_next,
                                          DevFlags* flags, uint32_t x, uint32_t y, uint32_t z, uint32_t tri,
                                          uint32_t q, uint64_t cur, PropAcc& acc) {
    if constexpr (mode_batch<MODE>) {
        const uint32_t q0 = __builtin_amdgcn_readfirstlane(q);
        if (__ballot(q != q0) == 0) {
            const GridReq R = wave_req(reqs.table, q0);
            prop_eval_req<MODE>(R, keys, dirs, changed, tile_changed, min_next, flags, x, y, z, tri, q0, cur, acc);
        } else {
            prop_eval_req<MODE>(reqs.table[q], keys, dirs, changed, tile_changed, min_next, flags, x, y, z, tri, q,
                                cur, acc);
        }
    } else {
        prop_eval_req<MODE>(reqs.one, keys, dirs, changed, tile_changed, min_next, flags, x, y, z, tri, 0u, cur, acc);
    }
}

// The start of a propagation launch: prefix[c] = items in the classes before
// c; min_cur and taken_next (read by this round's compaction, written by the
// next one) are reset; block 0 writes the frontier size to *total_out and,
// when host_probe is not null, to host_probe[0] in pinned host memory,
// followed (after a system-scope fence) by host_seq in host_probe[1]: the host
// reads the size once it sees the sequence number. Ends with a block barrier.
template <int MODE>
__device__ __forceinline__ void prop_prologue(const uint32_t* __restrict__ ctr_cur, uint32_t* prefix, uint32_t nreq,
                                              uint32_t* __restrict__ min_cur, uint32_t* __restrict__ taken_next,
                                              uint32_t* __restrict__ total_out, uint32_t* __restrict__ host_probe,
                                              uint32_t host_seq) {
    if (threadIdx.x < 64u) {
        const uint32_t c = ctr_cur[threadIdx.x * WORD_STRIDE];
        uint32_t x = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(x, off, 64);
            if ((int)threadIdx.x >= off) x += y;
        }
        prefix[threadIdx.x + 1] = x;
        if (threadIdx.x == 0) prefix[0] = 0;
        if (blockIdx.x == 0 && threadIdx.x == 63) {
            *total_out = x;
            if (host_probe) {
                __hip_atomic_store(&host_probe[0], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __threadfence_system();
                __hip_atomic_store(&host_probe[1], host_seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < (uint64_t)nreq * req_spread<MODE>;
         w += (uint64_t)gridDim.x * blockDim.x) {
        min_cur[w * WORD_STRIDE] = UINT32_MAX;
        taken_next[w * WORD_STRIDE] = 0u;
    }
    __syncthreads();
}

// Tries each item's triangle on the item's 26 neighbours: one lane per pair
// tests bounds and the neighbour's current key, the pairs left go through the
// wave's queue to prop_eval. A strictly smaller key is stored with atomicMin
// and flags the cell for a later round. The reads before the atomic are plain
// (cacheable) loads: keys only decrease, so a stale value can only cost an
// unneeded evaluation or atomic. Improved distances are folded into each
// request's min_next words; min_cur and taken_next (read by this round's
// compaction, written by the next one) are reset. The frontier size goes to
// *total_out and, when host_probe is not null, to host_probe[0] in pinned host
// memory, followed (after a system-scope fence) by host_seq in host_probe[1]:
// the host reads the size once it sees the sequence number.
template <int MODE>
__global__ void __launch_bounds__(256, 8) k_frontier_propagate(const FrontierItem* __restrict__ items, uint32_t region,
                                                            const uint32_t* __restrict__ ctr_cur,
                                                            ReqView<mode_batch<MODE>> reqs, uint32_t nreq,
                                                            uint64_t* keys, uint32_t* __restrict__ dirs,
                                                            uint8_t* __restrict__ changed,
                                                            uint8_t* __restrict__ tile_changed,
                                                            uint32_t* __restrict__ min_next,
                                                            uint32_t* __restrict__ min_cur,
                                                            uint32_t* __restrict__ taken_next,
                                                            uint32_t* __restrict__ total_out,
                                                            uint32_t* __restrict__ host_probe, uint32_t host_seq,
                                                            uint32_t item_base, uint32_t kmax,
                                                            DevFlags* flags,
                                                            unsigned long long* evals) {
    constexpr int SIGN = mode_sign<MODE>;
    // prefix[c]: items in the classes before c
    __shared__ uint32_t prefix[NCLASS + 1];
    __shared__ PropQueue<mode_batch<MODE>> Q;
    prop_prologue<MODE>(ctr_cur, prefix, nreq, min_cur, taken_next, total_out, host_probe, host_seq);
    // this launch's items: [item_base, item_base + prop_chunk_items(kmax)) of
    // the frontier (one launch covers the frontier of grids below 130M cells;
    // larger ones run several launches per round), so the 32-bit work index
    // (kmax neighbour slots per item) never overflows
    const uint32_t nfront = prefix[NCLASS];
    const uint32_t work = nfront > item_base ? min(nfront - item_base, prop_chunk_items(kmax)) * kmax : 0u;
    // A plain grid-stride loop: consecutive blocks, on different XCDs, take
    // consecutive stretches of pairs. Giving each XCD one contiguous eighth
    // of the work instead, for L2 locality, left the XCDs unevenly loaded
    // (128^3 grid 1.715 -> 1.686 ms, batched grids 35.8 -> 34.4 ms without it).
    const uint32_t j1 = work;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    PropAcc acc;
    uint32_t qn = 0;  // queued pairs (the same value in every lane of the wave)
    for (uint32_t base = blockIdx.x * blockDim.x + wave * 64u; base < j1; base += stride) {
        const uint32_t j = base + lane;
        bool keep = false;
        uint32_t x = 0, y = 0, z = 0, tri = 0, q = 0, k = 0, ix = 0, iy = 0, iz = 0;
        uint64_t cur = 0;
        if (j < j1) {
            const uint32_t f = item_base + j / kmax;
            k = j - (f - item_base) * kmax;
            uint32_t cls = 0;  // the last class whose prefix is <= f (a non-empty one)
#pragma unroll
            for (uint32_t step = NCLASS / 2; step; step >>= 1)
                if (prefix[cls + step] <= f) cls += step;
            read_item<MODE>(items[(uint64_t)cls * region + (f - prefix[cls])], tri, q, ix, iy, iz);
        }
        // A neighbour already holding this triangle holds exactly the key the
        // triangle would give it, and keys only decrease.
        auto candidate = [&](const GridReq& R) {
            int ox, oy, oz;
            if (j < j1 && neighbour_offset(k, R.reach, ox, oy, oz)) {
                x = ix + (uint32_t)ox;  // below 0 wraps past the axis length
                y = iy + (uint32_t)oy;
                z = iz + (uint32_t)oz;
                if (x < R.g.n[0] && y < R.g.n[1] && z < R.g.n[2]) {
                    cur = keys[R.cell_off + cell_index(R.g, x, y, z)];
                    keep = cur == KEY_EMPTY || key_tri(cur, SIGN) != tri;
                }
            }
        };
        if constexpr (mode_batch<MODE>) {
            // lane 0 is in range whenever the loop runs
            const uint32_t q0 = __builtin_amdgcn_readfirstlane(q);
            if (__ballot(j < j1 && q != q0) == 0) {
                const GridReq R = wave_req(reqs.table, q0);
                candidate(R);
            } else {
                candidate(reqs.table[q]);
            }
        } else {
            candidate(reqs.one);
        }
        const uint64_t m = __ballot(keep);
        if (keep) {
            const uint32_t pos = qn + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            Q.x[wave][pos] = x;
            Q.y[wave][pos] = y;
            Q.z[wave][pos] = z;
            Q.tri[wave][pos] = tri;
            if constexpr (mode_batch<MODE>) Q.req[wave][pos] = q;
            Q.cur_lo[wave][pos] = (uint32_t)cur;
            Q.cur_hi[wave][pos] = (uint32_t)(cur >> 32);
        }
        qn += (uint32_t)__popcll(m);
        wave_lds_sync();
        if (qn >= 64u) {  // evaluate the newest 64 queued pairs
            qn -= 64u;
            const uint32_t e = qn + lane;
            const uint64_t c = ((uint64_t)Q.cur_hi[wave][e] << 32) | Q.cur_lo[wave][e];
            prop_eval<MODE>(reqs, keys, dirs, changed, tile_changed, min_next, flags, Q.x[wave][e], Q.y[wave][e],
                            Q.z[wave][e], Q.tri[wave][e], mode_batch<MODE> ? Q.req[wave][e] : 0u, c, acc);
            wave_lds_sync();
        }
    }
    if (lane < qn) {
        const uint64_t c = ((uint64_t)Q.cur_hi[wave][lane] << 32) | Q.cur_lo[wave][lane];
        prop_eval<MODE>(reqs, keys, dirs, changed, tile_changed, min_next, flags, Q.x[wave][lane],
                        Q.y[wave][lane], Q.z[wave][lane], Q.tri[wave][lane],
                        mode_batch<MODE> ? Q.req[wave][lane] : 0u, c, acc);
    }
    if (mode_batch<MODE>) fold_req_min<MODE>(min_next, acc.req, acc.imin, blockIdx.x);
    else block_min_u32(req_word<MODE>(min_next, 0, blockIdx.x), acc.imin);
    if (evals) block_count(evals, acc.n_evals);
}

// k_frontier_propagate for large rounds, item-major: one lane per frontier
// item walks the item's 26 neighbours, so the supporting-plane bound
// (plane_bound) is computed once per item and skips most of the neighbours
// that cannot improve before their distance is evaluated; the item is read
// once instead of once per neighbour. A skipped pair cannot improve its
// neighbour, so the keys end exactly as k_frontier_propagate leaves them; the
// host picks this kernel for rounds of at least grid_item_major_min items,
// where 26 dependent steps per lane are hidden by many waves. Same arguments
// and round bookkeeping as k_frontier_propagate.
template <int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_frontier_propagate_items(const FrontierItem* __restrict__ items,
                                                                  uint32_t region, const uint32_t* __restrict__ ctr_cur,
                                                                  ReqView<mode_batch<MODE>> reqs, uint32_t nreq,
                                                                  uint64_t* keys, uint32_t* __restrict__ dirs,
                                                                  uint8_t* __restrict__ changed,
                                                                  uint8_t* __restrict__ tile_changed,
                                                                  uint32_t* __restrict__ min_next,
                                                                  uint32_t* __restrict__ min_cur,
                                                                  uint32_t* __restrict__ taken_next,
                                                                  uint32_t* __restrict__ total_out,
                                                                  uint32_t* __restrict__ host_probe, uint32_t host_seq,
                                                                  uint32_t item_base, uint32_t kmax, DevFlags* flags,
                                                                  unsigned long long* evals) {
    constexpr int SIGN = mode_sign<MODE>;
    __shared__ uint32_t prefix[NCLASS + 1];
    __shared__ PropQueueRef Q;  // pairs of one request at a time
    // each lane's item, bound and triangle, read back at every neighbour so
    // that they need not stay in registers across the queue's evaluations;
    // the queued pairs take their cell and triangle from here
    __shared__ uint4 s_item[256];
    __shared__ float4 s_bound[256];
    __shared__ float s_tv[10][256];  // triangle a, b, c and the degeneracy code (as bits) per lane
    prop_prologue<MODE>(ctr_cur, prefix, nreq, min_cur, taken_next, total_out, host_probe, host_seq);
    const uint32_t nfront = prefix[NCLASS];
    const uint32_t work = nfront > item_base ? min(nfront - item_base, prop_chunk_items(kmax)) : 0u;
    // a plain grid-stride loop over the items, as in k_frontier_propagate
    const uint32_t j1 = work;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    PropAcc acc;
    uint32_t qn = 0;  // queued pairs (the same value in every lane of the wave)
    // The queue holds pairs of one request, qreq (wave-uniform): a walk for
    // another request first evaluates what is queued. Every evaluation thus
    // reads its request from scalar registers.
    uint32_t qreq = 0;
    // evaluates queued pair e (its item's cell, triangle and the neighbour's key from LDS)
    auto eval_pair = [&](const GridReq& R, uint32_t q0, uint32_t e) {
        const uint32_t sk = Q.sk[wave][e], src = sk >> 8;
        const uint64_t cv = ((uint64_t)Q.cur_hi[wave][e] << 32) | Q.cur_lo[wave][e];
        int ox = 0, oy = 0, oz = 0;
        neighbour_offset(sk & 0xFFu, R.reach, ox, oy, oz);
        const uint4 it = s_item[src];
        // the same values expand_leaf gives (the topology index is not used by the grid)
        TriRec tr;
        make_trirec(s_tv[0][src], s_tv[1][src], s_tv[2][src], s_tv[3][src], s_tv[4][src], s_tv[5][src],
                    s_tv[6][src], s_tv[7][src], s_tv[8][src], 0u, tr);
        tr.degen = fbits(s_tv[9][src]);
        prop_eval_tri<MODE>(R, tr, keys, dirs, changed, tile_changed, min_next, flags, it.x + (uint32_t)ox,
                            it.y + (uint32_t)oy, it.z + (uint32_t)oz, it.w, q0, cv, acc);
    };
    auto drain = [&](const GridReq& R, uint32_t q0) {  // evaluates the qn < 64 queued pairs
        if (lane < qn) eval_pair(R, q0, lane);
        qn = 0;
        wave_lds_sync();
    };
    for (uint32_t base = blockIdx.x * blockDim.x + wave * 64u; base < j1; base += stride) {
        const uint32_t j = base + lane;
        const bool live = j < j1;
        uint32_t tri = 0, q = 0, ix = 0, iy = 0, iz = 0;
        if (live) {
            const uint32_t f = item_base + j;
            uint32_t cls = 0;  // the last class whose prefix is <= f (a non-empty one)
#pragma unroll
            for (uint32_t step = NCLASS / 2; step; step >>= 1)
                if (prefix[cls + step] <= f) cls += step;
            read_item<MODE>(items[(uint64_t)cls * region + (f - prefix[cls])], tri, q, ix, iy, iz);
        }
        auto walk = [&](const GridReq& R, uint32_t q0, bool mine) {
            if (mine) {
                const float max_size = fmaxf(fmaxf(fabsf(R.g.size[0]), fabsf(R.g.size[1])), fabsf(R.g.size[2]));
                const LeafRec lf = R.ltris[tri];
                s_tv[0][threadIdx.x] = lf.ax;
                s_tv[1][threadIdx.x] = lf.ay;
                s_tv[2][threadIdx.x] = lf.az;
                s_tv[3][threadIdx.x] = lf.bx;
                s_tv[4][threadIdx.x] = lf.by;
                s_tv[5][threadIdx.x] = lf.bz;
                s_tv[6][threadIdx.x] = lf.cx;
                s_tv[7][threadIdx.x] = lf.cy;
                s_tv[8][threadIdx.x] = lf.cz;
                s_tv[9][threadIdx.x] = bitsf(lf.degen);
                const PlaneBound pb = plane_bound(dirs[R.cell_off + cell_index(R.g, ix, iy, iz)],
                                                  cell_coord(R.g, 0, ix), cell_coord(R.g, 1, iy),
                                                  cell_coord(R.g, 2, iz), expand_leaf(lf), max_size);
                s_bound[threadIdx.x] = make_float4(pb.ux, pb.uy, pb.uz, pb.g);  // only this lane reads it
                s_item[threadIdx.x] = make_uint4(ix, iy, iz, tri);
            }
            for (uint32_t k = 0; k < kmax; ++k) {
                bool keep = false;
                uint32_t x = 0, y = 0, z = 0, t = 0;
                uint64_t cur = 0;
                int ox = 0, oy = 0, oz = 0;
                const bool has = neighbour_offset(k, R.reach, ox, oy, oz);
                if (mine && has) {
                    const uint4 it = s_item[threadIdx.x];
                    t = it.w;
                    x = it.x + (uint32_t)ox;  // -1 wraps past the axis length
                    y = it.y + (uint32_t)oy;
                    z = it.z + (uint32_t)oz;
                    if (x < R.g.n[0] && y < R.g.n[1] && z < R.g.n[2]) {
                        cur = keys[R.cell_off + cell_index(R.g, x, y, z)];
                        // a neighbour already holding this triangle holds exactly the key it would give
                        keep = cur == KEY_EMPTY || key_tri(cur, SIGN) != t;
                        if (keep && cur != KEY_EMPTY) {
                            // |d| must exceed the key's distance to come first (in Normal
                            // mode, past a negative value it approximately equals, as in
                            // prop_eval_req)
                            uint32_t cb = (uint32_t)(cur >> 32);
                            if (SIGN == SDFX_SIGN_NORMAL && ((cur >> 31) & 1ull))
                                cb = max(fbits(bitsf(cb) + 2e-6f), cb + 2u);
                            const float4 pb = s_bound[threadIdx.x];
                            const float lb = pb.w + (pb.x * ((float)ox * R.g.size[0]) + pb.y * ((float)oy * R.g.size[1]) +
                                                     pb.z * ((float)oz * R.g.size[2]));
                            if (lb > bitsf(cb)) keep = false;
                        }
                    }
                }
                const uint64_t m = __ballot(keep);
                if (keep) {
                    const uint32_t pos = qn + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                    Q.sk[wave][pos] = threadIdx.x << 8 | k;
                    Q.cur_lo[wave][pos] = (uint32_t)cur;
                    Q.cur_hi[wave][pos] = (uint32_t)(cur >> 32);
                }
                qn += (uint32_t)__popcll(m);
                wave_lds_sync();
                if (qn >= 64u) {  // evaluate the newest 64 queued pairs
                    qn -= 64u;
                    eval_pair(R, q0, qn + lane);
                    wave_lds_sync();
                }
            }
        };
        if constexpr (mode_batch<MODE>) {
            // one walk per request among the wave's items (nearly always one:
            // items lie request by request), each with every lane taking part
            // and the request read once through the scalar cache
            uint64_t pending = __ballot(live);
            while (pending) {
                const uint32_t q0 = __builtin_amdgcn_readlane(q, __ffsll((unsigned long long)pending) - 1);
                const bool mine = live && q == q0;
                if (q0 != qreq && qn) drain(wave_req(reqs.table, qreq), qreq);
                qreq = q0;
                walk(wave_req(reqs.table, q0), q0, mine);
                pending &= ~__ballot(mine);
            }
        } else {
            walk(reqs.one, 0u, live);
        }
        // the next items replace these in LDS: evaluate what is queued
        if (qn) {
            if constexpr (mode_batch<MODE>) drain(wave_req(reqs.table, qreq), qreq);
            else drain(reqs.one, 0u);
        }
    }
    if (mode_batch<MODE>) fold_req_min<MODE>(min_next, acc.req, acc.imin, blockIdx.x);
    else block_min_u32(req_word<MODE>(min_next, 0, blockIdx.x), acc.imin);
    if (evals) block_count(evals, acc.n_evals);
}

// ---------------------------------------------------------------- raycast sign

// Rows of axis a whose origin coordinate along perpendicular axis k lies in
// [lo, hi]: a conservative index range, checked exactly per row afterwards.
__device__ __forceinline__ void row_range(const GridDev& g, int k, float lo, float hi, uint32_t& i0, uint32_t& i1) {
    float a = (lo - g.first[k]) / g.size[k];
    float b = (hi - g.first[k]) / g.size[k];
    float fmn = fminf(a, b), fmx = fmaxf(a, b);
    float n1 = (float)(g.n[k] - 1);
    // one extra row on each side absorbs the rounding of the division
    float s = floorf(fmn) - 1.0f, e = ceilf(fmx) + 1.0f;
    if (!(s >= 0.0f)) s = 0.0f;  // also NaN
    if (!(e <= n1)) e = n1;
    if (s > n1) s = n1;
    if (e < 0.0f) e = 0.0f;
    i0 = (uint32_t)s;
    i1 = (uint32_t)e;
}

// The ray rows of triangle t along axis a (work item 3 t + a): the rows whose
// origin lies in the triangle's padded box, a0..a0+na-1 along the first
// perpendicular axis and c0..c0+nc-1 along the second (none when the box lies
// entirely behind the rows' origins).
struct RayRows {
    uint32_t a0, c0, na, nc;
};

template <int AXIS>
__device__ __forceinline__ RayRows ray_rows(const Box& b, const GridDev& g) {
    constexpr int K1 = AXIS == 0 ? 1 : 0;  // the row index is (K1, K2) with K2 fastest
    constexpr int K2 = AXIS == 2 ? 1 : 2;
    RayRows r{0u, 0u, 0u, 0u};
    if (!(b.hi[AXIS] >= cell_coord(g, AXIS, 0u))) return r;  // box entirely behind the origins
    uint32_t a1, c1;
    row_range(g, K1, b.lo[K1], b.hi[K1], r.a0, a1);
    row_range(g, K2, b.lo[K2], b.hi[K2], r.c0, c1);
    if (a1 >= r.a0 && c1 >= r.c0) {
        r.na = a1 - r.a0 + 1;
        r.nc = c1 - r.c0 + 1;
    }
    return r;
}

// One ray row (i, j) of triangle tr along AXIS: the exact per-row box test,
// the aligned ray test and the parity flip of the hit's cell.
template <int AXIS>
__device__ __forceinline__ void ray_row(const TriRec& tr, const Box& b, const GridDev& g, uint32_t i, uint32_t j,
                                        uint32_t* __restrict__ bits) {
    constexpr int K1 = AXIS == 0 ? 1 : 0;
    constexpr int K2 = AXIS == 2 ? 1 : 2;
    const float u = cell_coord(g, K1, i), v = cell_coord(g, K2, j);
    if (!(b.lo[K1] <= u && u <= b.hi[K1]) || !(b.lo[K2] <= v && v <= b.hi[K2])) return;
    float o[3];
    o[AXIS] = cell_coord(g, AXIS, 0u);
    o[K1] = u;
    o[K2] = v;
    float t;
    if (!ray_hit<AXIS>(o[0], o[1], o[2], tr, &t)) return;
    // generate/grid.rs:605-607: (t / size).floor() as usize, capped at n - 1
    const float q = floorf(t / g.size[AXIS]);
    uint32_t kcell;
    if (!(q > 0.0f)) kcell = 0;  // NaN and negatives saturate to 0
    else if (q >= (float)(g.n[AXIS] - 1)) kcell = g.n[AXIS] - 1;
    else kcell = (uint32_t)q;
    const uint64_t row = (uint64_t)i * g.n[K2] + j;
    atomicXor(&bits[row * g.words[AXIS] + (kcell >> 5)], 1u << (kcell & 31u));
}

// A (triangle, axis) with at most RAY_SMALL rows (a triangle spanning a few
// cells has up to about 36: row_range adds a row on each side) walks them on
// its own lane;
// a larger one goes to a list whose entries each get a whole block, so a few
// large triangles on a fine grid spread over many lanes instead of walking
// all their rows on one.
constexpr uint32_t RAY_SMALL = 64;

struct RayBig {
    uint64_t w;  // 3 t + axis
    RayRows r;
};

template <int MODE>
__device__ __forceinline__ void ray_rows_walk(const GridReq& R, uint64_t t, int axis, const RayRows& r, uint64_t k0,
                                              uint64_t k1, uint64_t kstep, uint32_t* __restrict__ bx,
                                              uint32_t* __restrict__ by, uint32_t* __restrict__ bz) {
    const TriRec tr = expand_leaf(R.ltris[t - R.tri_off]);
    Box b;
    tri_bbox(tr.ax, tr.ay, tr.az, tr.bx, tr.by, tr.bz, tr.cx, tr.cy, tr.cz, b);
    for (uint64_t k = k0; k < k1; k += kstep) {
        const uint32_t i = r.a0 + (uint32_t)(k / r.nc), j = r.c0 + (uint32_t)(k % r.nc);
        if (axis == 0) ray_row<0>(tr, b, R.g, i, j, bx + R.row_off[0]);
        else if (axis == 1) ray_row<1>(tr, b, R.g, i, j, by + R.row_off[1]);
        else ray_row<2>(tr, b, R.g, i, j, bz + R.row_off[2]);
    }
}

template <int AXIS>
__device__ __forceinline__ void ray_rows_small(const TriRec& tr, const Box& b, const GridDev& g, const RayRows& r,
                                               uint32_t* __restrict__ bits) {
    for (uint32_t i = 0; i < r.na; ++i)
        for (uint32_t j = 0; j < r.nc; ++j) ray_row<AXIS>(tr, b, g, r.a0 + i, r.c0 + j, bits);
}

// One lane per (triangle, axis): walks a small box's rows, lists a large one.
template <int MODE>
__global__ void __launch_bounds__(256) k_ray_small(uint64_t nt, ReqView<mode_batch<MODE>> reqs,
                                                   const uint32_t* __restrict__ tri_req, RayBig* __restrict__ big,
                                                   uint32_t* __restrict__ n_big, uint32_t* __restrict__ bx,
                                                   uint32_t* __restrict__ by, uint32_t* __restrict__ bz) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool listed = false;
    RayRows r{0u, 0u, 0u, 0u};
    if (id < 3 * nt) {
        const uint64_t t = id / 3;
        const int axis = (int)(id % 3);
        const uint32_t q = mode_batch<MODE> ? tri_req[t] : 0u;
        if (q != UINT32_MAX) {  // else a triangle of no request in this run
            with_req<MODE>(reqs, q, [&](const GridReq& R, uint32_t) {
                const TriRec tr = expand_leaf(R.ltris[t - R.tri_off]);
                Box b;
                tri_bbox(tr.ax, tr.ay, tr.az, tr.bx, tr.by, tr.bz, tr.cx, tr.cy, tr.cz, b);
                r = axis == 0 ? ray_rows<0>(b, R.g) : (axis == 1 ? ray_rows<1>(b, R.g) : ray_rows<2>(b, R.g));
                if ((uint64_t)r.na * r.nc > RAY_SMALL) {
                    listed = true;
                } else if (axis == 0) {
                    ray_rows_small<0>(tr, b, R.g, r, bx + R.row_off[0]);
                } else if (axis == 1) {
                    ray_rows_small<1>(tr, b, R.g, r, by + R.row_off[1]);
                } else {
                    ray_rows_small<2>(tr, b, R.g, r, bz + R.row_off[2]);
                }
            });
        }
    }
    // one atomic per wave appends its listed entries
    const uint64_t m = __ballot(listed);
    if (m) {
        const uint32_t lane = threadIdx.x & 63u;
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(n_big, (uint32_t)__popcll(m));
        base = __shfl(base, leader, 64);
        if (listed) big[base + __popcll(m & ((1ull << lane) - 1ull))] = RayBig{id, r};
    }
}

// The listed (triangle, axis) entries, one block each at a time.
template <int MODE>
__global__ void __launch_bounds__(256) k_ray_big(ReqView<mode_batch<MODE>> reqs, const uint32_t* __restrict__ tri_req,
                                                 const RayBig* __restrict__ big, const uint32_t* __restrict__ n_big,
                                                 uint32_t* __restrict__ bx, uint32_t* __restrict__ by,
                                                 uint32_t* __restrict__ bz) {
    const uint32_t count = *n_big;
    for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
        const RayBig E = big[e];
        const uint64_t t = E.w / 3;
        const int axis = (int)(E.w % 3);
        const uint64_t nrows = (uint64_t)E.r.na * E.r.nc;
        with_req<MODE>(reqs, mode_batch<MODE> ? tri_req[t] : 0u, [&](const GridReq& R, uint32_t) {
            ray_rows_walk<MODE>(R, t, axis, E.r, threadIdx.x, nrows, blockDim.x, bx, by, bz);
        });
    }
}

// Per row: bit i becomes the XOR of flip bits i..n-1 (the parity of hits whose
// last counted cell is at or beyond i). Rows of all requests along one axis are
// numbered consecutively; a row finds its request by binary search.
template <int MODE>
__global__ void __launch_bounds__(256) k_ray_suffix(uint32_t* __restrict__ bx, uint32_t* __restrict__ by,
                                                    uint32_t* __restrict__ bz, uint64_t rows_x, uint64_t rows_y,
                                                    uint64_t rows_z, ReqView<mode_batch<MODE>> reqs, uint32_t nreq) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // the rows of the three axes, one after the other
    int axis = 0;
    uint32_t* bits = bx;
    if (r >= rows_x) {
        r -= rows_x;
        axis = 1;
        bits = by;
        if (r >= rows_y) {
            r -= rows_y;
            axis = 2;
            bits = bz;
            if (r >= rows_z) return;
        }
    }
    uint32_t lo = 0, hi = nreq - 1;  // the last request whose first row is <= r
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (reqs[mid].row_first[axis] <= r) lo = mid;
        else hi = mid - 1;
    }
    const GridReq& R = reqs[lo];
    const uint32_t words = R.g.words[axis];
    uint32_t* row = bits + R.row_off[axis] + (r - R.row_first[axis]) * words;
    uint32_t carry = 0;
    for (int w = (int)words - 1; w >= 0; --w) {
        uint32_t x = row[w];
        x ^= x >> 1;
        x ^= x >> 2;
        x ^= x >> 4;
        x ^= x >> 8;
        x ^= x >> 16;
        if (carry) x = ~x;
        row[w] = x;
        carry = x & 1u;
    }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_grid_finish(const uint64_t* __restrict__ keys, uint64_t ncells,
                                                     ReqView<mode_batch<MODE>> reqs,
                                                     const uint32_t* __restrict__ tile_req,
                                                     const uint32_t* __restrict__ px,
                                                     const uint32_t* __restrict__ py,
                                                     const uint32_t* __restrict__ pz) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncells;
         i += (uint64_t)gridDim.x * blockDim.x) {
        with_req<MODE>(reqs, mode_batch<MODE> ? tile_req[i / COMPACT_TILE] : 0u, [&](const GridReq& R, uint32_t) {
            const uint64_t l = i - R.cell_off;
            if (l >= R.ncells) return;  // tile padding after the request's last cell
            const GridDev& g = R.g;
            constexpr int SIGN = mode_sign<MODE>;
            float v = key_value(keys[i], SIGN);
            if (SIGN == SDFX_SIGN_RAYCAST) {
                uint32_t z = (uint32_t)(l % g.n[2]);
                uint32_t y = (uint32_t)((l / g.n[2]) % g.n[1]);
                uint32_t x = (uint32_t)(l / ((uint64_t)g.n[1] * g.n[2]));
                uint32_t ox = (px[R.row_off[0] + ((uint64_t)y * g.n[2] + z) * g.words[0] + (x >> 5)] >> (x & 31u)) & 1u;
                uint32_t oy = (py[R.row_off[1] + ((uint64_t)x * g.n[2] + z) * g.words[1] + (y >> 5)] >> (y & 31u)) & 1u;
                uint32_t oz = (pz[R.row_off[2] + ((uint64_t)x * g.n[1] + y) * g.words[2] + (z >> 5)] >> (z & 31u)) & 1u;
                if (ox + oy + oz >= 2u) v = -v;  // generate/grid.rs:633-638
            }
            R.out[l] = v;
        });
    }
}

// ---------------------------------------------------------------- driver

// Owner request of every tile and every triangle of a batch (UINT32_MAX for a
// triangle of no request in the run), by binary search over the table.
__global__ void __launch_bounds__(256) k_batch_owners(const GridReq* __restrict__ reqs, uint32_t nreq,
                                                      uint32_t* __restrict__ tile_req, uint64_t ntiles,
                                                      uint32_t* __restrict__ tri_req, uint64_t nt) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntiles) {
        uint32_t lo = 0, hi = nreq - 1;  // the last request whose first tile is <= i
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (reqs[mid].cell_off / COMPACT_TILE <= i) lo = mid;
            else hi = mid - 1;
        }
        tile_req[i] = lo;
    }
    if (i < nt) {
        uint32_t lo = 0, hi = nreq - 1;  // the last request whose first triangle is <= i
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (reqs[mid].tri_off <= i) lo = mid;
            else hi = mid - 1;
        }
        tri_req[i] = (reqs[lo].tri_off <= i && i < reqs[lo].tri_off + reqs[lo].ntris) ? lo : UINT32_MAX;
    }
}

// A per-thread pinned buffer for small host-to-device uploads, so they are
// asynchronous copies. A call reuses it only after its previous call has
// synchronised the stream.
static void* pinned_stage(size_t bytes) {
    struct Stage {
        void* p = nullptr;
        size_t cap = 0;
        ~Stage() {
            if (p) (void)hipHostFree(p);
        }
    };
    static thread_local Stage st;
    if (bytes > st.cap) {
        if (st.p) SDFX_HIP(hipHostFree(st.p));
        st.p = nullptr;
        st.cap = 0;
        SDFX_HIP(hipHostMalloc(&st.p, bytes, hipHostMallocDefault));
        st.cap = bytes;
    }
    return st.p;
}

// Two pinned probes per thread: the last propagation of a batch of rounds
// stores its frontier size and then the batch's sequence number straight into
// one of them (no copy kernel, no event in the stream: an event record costs a
// few microseconds of idle GPU between batches), and the host polls for the
// sequence number while the next batch of rounds runs.
struct BatchProbe {
    static constexpr uint32_t STRIDE = 16;  // u32 per probe: one 64-byte line each
    struct Slots {
        uint32_t* words = nullptr;    // host address
        uint32_t* d_words = nullptr;  // the same words as the device addresses them
        uint32_t seq = 0;             // last sequence number handed out on this thread
        uint32_t want[2] = {0, 0};
        ~Slots() {
            if (words) (void)hipHostFree(words);
        }
    };
    Slots& slots() {
        static thread_local Slots sl;
        if (!sl.words) {
            SDFX_HIP(hipHostMalloc((void**)&sl.words, 2 * STRIDE * sizeof(uint32_t),
                                   hipHostMallocMapped | hipHostMallocCoherent));
            std::memset(sl.words, 0, 2 * STRIDE * sizeof(uint32_t));
            SDFX_HIP(hipHostGetDevicePointer((void**)&sl.d_words, sl.words, 0));
        }
        return sl;
    }
    uint32_t* probe(int slot) { return slots().d_words + slot * STRIDE; }
    // the sequence number the next batch in `slot` reports
    uint32_t arm(int slot) {
        Slots& sl = slots();
        sl.want[slot] = ++sl.seq ? sl.seq : ++sl.seq;  // never 0, the words' initial value
        return sl.want[slot];
    }
    // The frontier size reported by the batch armed in `slot`. The batch's
    // launches are already in stream s: if s goes idle without the report, or
    // a minute passes, the run failed and an error is raised instead of waiting.
    uint32_t wait(int slot, hipStream_t s) {
        Slots& sl = slots();
        const uint32_t* w = sl.words + slot * STRIDE;
        const auto t0 = std::chrono::steady_clock::now();
        for (uint64_t spin = 1;; ++spin) {
            if (__atomic_load_n(&w[1], __ATOMIC_ACQUIRE) == sl.want[slot]) break;
            if ((spin & 4095u) == 0) {
                const hipError_t q = hipStreamQuery(s);
                if (q == hipSuccess) {
                    if (__atomic_load_n(&w[1], __ATOMIC_ACQUIRE) == sl.want[slot]) break;
                    throw Error(SDFX_ERR_DEVICE, "grid round batch ended without reporting its frontier");
                }
                if (q != hipErrorNotReady) SDFX_HIP(q);
                if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60))
                    throw Error(SDFX_ERR_DEVICE, "grid round batch did not report its frontier within 60 s");
            }
            __builtin_ia32_pause();
        }
        return __atomic_load_n(&w[0], __ATOMIC_ACQUIRE);
    }
};

static unsigned stride_blocks(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b ? b : 1);
}

template <int MODE>
static void grid_run(uint64_t nt, const std::vector<GridBatchItem>& jobs, DevFlags* d_flags, Arena& A) {
    constexpr int SIGN = mode_sign<MODE>;
    hipStream_t s = A.stream;
    const uint32_t nreq = (uint32_t)jobs.size();
    // request table: cells padded to whole tiles, rows numbered per axis
    std::vector<GridReq> reqs(nreq);
    LeafRec* ltris = A.alloc<LeafRec>(nt);
    uint64_t ncells = 0, rows_total[3] = {0, 0, 0}, words_total[3] = {0, 0, 0}, max_extent = 0;
    uint32_t kmax = 26;  // neighbours per item of the request with the most (neighbour_offset)
    for (uint32_t q = 0; q < nreq; ++q) {
        const GridBatchItem& J = jobs[q];
        GridReq& R = reqs[q];
        for (int k = 0; k < 3; ++k) {
            R.g.first[k] = J.gp.first[k];
            R.g.size[k] = J.gp.size[k];
            R.g.n[k] = (uint32_t)J.gp.count[k];
            R.g.words[k] = (uint32_t)((J.gp.count[k] + 31) / 32);
        }
        if (J.gp.count[0] > AXIS_XY_LIMIT || J.gp.count[1] > AXIS_XY_LIMIT || J.gp.count[2] > AXIS_Z_LIMIT)
            throw Error(SDFX_ERR_INVALID_ARGUMENT, "grid axes are limited to 2^21 (x, y) and 2^22 (z) cells");
        const uint64_t nc = J.gp.count[0] * J.gp.count[1] * J.gp.count[2];
        R.cell_off = ncells;
        R.ncells = (uint32_t)nc;
        R.tris = J.d_tris;
        R.ltris = ltris + J.tri_off;
        R.tri_off = J.tri_off;
        R.ntris = J.n_tris;
        R.out = J.d_out;
        R.flag = J.flag;
        const float max_size = std::fmax(std::fmax(std::fabs(R.g.size[0]), std::fabs(R.g.size[1])),
                                         std::fabs(R.g.size[2]));
        R.band = GRID_BAND_CELLS * max_size;
        R.far = GRID_FAR_CELLS * max_size;
        R.release = (uint32_t)std::max<uint64_t>(GRID_RELEASE_MIN, nc / GRID_RELEASE_DIV);
        R.reach = 0;
        for (int k = 0; k < 3; ++k)
            if (grid_reach(max_size, std::fabs(R.g.size[k]))) R.reach |= 1u << k;
        kmax = std::max<uint32_t>(kmax, 26u + 2u * (uint32_t)__builtin_popcount(R.reach));
        const uint64_t rows[3] = {J.gp.count[1] * J.gp.count[2], J.gp.count[0] * J.gp.count[2],
                                  J.gp.count[0] * J.gp.count[1]};
        for (int a = 0; a < 3; ++a) {
            R.row_first[a] = rows_total[a];
            R.row_off[a] = words_total[a];
            rows_total[a] += rows[a];
            words_total[a] += rows[a] * R.g.words[a];
        }
        const uint64_t tiles = (nc + COMPACT_TILE - 1) / COMPACT_TILE;
        ncells += tiles * COMPACT_TILE;
        max_extent = std::max<uint64_t>(max_extent, J.gp.count[0] + J.gp.count[1] + J.gp.count[2]);
    }
    // frontier counts and cell indices within a request are 32-bit; a run
    // holds fewer than 2^32 tile-padded cells (grid_sdf_batch splits batches)
    if (ncells > (1ull << 32)) throw Error(SDFX_ERR_INVALID_ARGUMENT, "a grid run holds at most 2^32 cells");
    const uint32_t ntiles = (uint32_t)(ncells / COMPACT_TILE);
    // a single request travels in the kernel arguments; a batch gets its table,
    // and the owner request of every tile and triangle, on the device
    GridReq* d_reqs = nullptr;
    uint32_t *tile_req = nullptr, *tri_req = nullptr;
    if (mode_batch<MODE>) {
        d_reqs = A.alloc<GridReq>(nreq);
        tile_req = A.alloc<uint32_t>(ntiles);
        tri_req = A.alloc<uint32_t>(nt);
        void* stage = pinned_stage(nreq * sizeof(GridReq));
        std::memcpy(stage, reqs.data(), nreq * sizeof(GridReq));
        SDFX_HIP(hipMemcpyAsync(d_reqs, stage, nreq * sizeof(GridReq), hipMemcpyHostToDevice, s));
        SDFX_LAUNCH(k_batch_owners, dim3(blocks_for(std::max<uint64_t>(ntiles, nt), 256)), dim3(256), 0, s, d_reqs,
                    nreq, tile_req, (uint64_t)ntiles, tri_req, nt);
    }
    ReqView<mode_batch<MODE>> rv{d_reqs, reqs[0]};

    // state of the run: keys; per parity of the round NCLASS item counters,
    // then per request its band minima and its taken-cell counts; the two
    // frontier totals and the per-request far-field flags; the ray-parity bits
    uint64_t* keys = A.alloc<uint64_t>(ncells);
    uint32_t* dirs = A.alloc<uint32_t>(ncells);  // direction words: any value is valid, left as they are
    const uint64_t cls_words = (uint64_t)NCLASS * WORD_STRIDE;
    const uint64_t req_words = (uint64_t)nreq * req_spread<MODE> * WORD_STRIDE;
    const uint64_t bit_words = SIGN == SDFX_SIGN_RAYCAST ? words_total[0] + words_total[1] + words_total[2] : 0;
    WordRuns runs;
    runs.end[0] = 2 * cls_words;                   // item counters: 0
    runs.end[1] = runs.end[0] + 2 * req_words;     // band minima: none yet
    runs.end[2] = runs.end[1] + req_words;         // taken counts of even rounds: 0
    runs.end[3] = runs.end[2] + req_words;         // taken counts read by round 0: large, so the band applies
    runs.end[4] = runs.end[3] + 3 + 2 * nreq;      // totals, release flags, the ray stage's list count, first thin rounds: 0
    runs.end[5] = runs.end[4] + bit_words;         // ray-parity bits: 0
    runs.value[0] = 0u;
    runs.value[1] = UINT32_MAX;
    runs.value[2] = 0u;
    // 2^24 per word: the sum over a request's words, at least 2^28, exceeds
    // every release threshold and stays below 2^32
    runs.value[3] = 1u << 24;
    runs.value[4] = 0u;
    runs.value[5] = 0u;
    uint32_t* words = A.alloc<uint32_t>(runs.end[5]);
    SDFX_LAUNCH(k_grid_init, dim3(stride_blocks(ncells + runs.end[5])), dim3(256), 0, s, keys, ncells, words, runs);

    // 1. seeding
    SeedBox* boxes = A.alloc<SeedBox>(nt);
    uint64_t* counts = A.alloc<uint64_t>(nt);
    uint64_t* total = A.alloc<uint64_t>(1);
    SDFX_LAUNCH(k_seed_boxes<MODE>, dim3(blocks_for(nt, 256)), dim3(256), 0, s, nt, rv, tri_req, boxes, counts);
    // a schedule-free request brings its values: its keys come from them and
    // the seeding pairs and frontier rounds are skipped (the 48-byte triangle
    // copies k_seed_boxes wrote serve the ray stage)
    const bool exact = nreq == 1 && jobs[0].exact != nullptr;
    // a batch's rounds are large: one resident generation of blocks does best there
    const int64_t blocks_knob = mode_batch<MODE> ? tuning().grid_batch_prop_blocks : tuning().grid_prop_blocks;
    const unsigned prop_blocks = (unsigned)std::max<int64_t>(XCDS, blocks_knob) / XCDS * XCDS;
    // distance-evaluation counters, only while kernel timing is on: one slot
    // per block of k_seed_pairs (8192 blocks), then one per propagation block
    unsigned long long* evals = nullptr;
    const uint64_t seed_slots = 8192, prop_slots = prop_blocks;
    if (profile_enabled() && !exact) {
        evals = A.alloc<unsigned long long>(seed_slots + prop_slots);
        fill_u32(reinterpret_cast<uint32_t*>(evals), 2 * (seed_slots + prop_slots), 0u, s);
    }
    uint32_t* totals = words + 2 * cls_words + 4 * req_words;
    if (exact) {
        SDFX_LAUNCH(k_keys_from_values<MODE>, dim3(stride_blocks(reqs[0].ncells)), dim3(256), 0, s, jobs[0].exact,
                    (uint64_t)reqs[0].ncells, keys, d_flags + reqs[0].flag);
    } else {
    exclusive_scan_u64(counts, counts, nt, total, A);
    SDFX_LAUNCH(k_seed_pairs<MODE>, dim3(8192), dim3(256), 0, s, nt, rv, tri_req, boxes, counts, total, keys, dirs,
                d_flags, evals);
    SDFX_HIP(hipGetLastError());

    // 2. frontier rounds
    uint32_t* ctr[2] = {words, words + cls_words};
    uint32_t* minw[2] = {words + 2 * cls_words, words + 2 * cls_words + req_words};
    uint32_t* taken[2] = {words + 2 * cls_words + 2 * req_words, words + 2 * cls_words + 3 * req_words};
    uint32_t* released = totals + 2;  // per request: every later round takes every flagged cell
    uint32_t* thin_round = totals + 3 + nreq;  // per request: 1 + the round number of its first thin round, or 0
    const uint32_t region = (uint32_t)((ntiles + NCLASS - 1) / NCLASS * COMPACT_TILE);  // items per class
    uint8_t* changed = A.alloc<uint8_t>(ncells);  // whole tiles: 16-byte vector reads stay inside
    uint8_t* tile_changed = A.alloc<uint8_t>(ntiles);
    FrontierItem* items = A.alloc<FrontierItem>((uint64_t)region * NCLASS);
    SDFX_LAUNCH(k_mark_seeded<MODE>, dim3(stride_blocks(ncells)), dim3(256), 0, s, keys, ncells, tile_req, changed,
                tile_changed, minw[0]);
    SDFX_HIP(hipGetLastError());
    // one block per flagged-tile scan, up to 16,384 blocks (a block walks the tiles
    // beyond that one after another; batched grids: 4,096 -> 16,384 cut compaction 9%)
    const unsigned compact_blocks = std::min<uint32_t>(ntiles, 16384u);
    int batch = 8;
    // Every round takes at least the nearest flagged cell of each request with
    // flagged cells and every improvement strictly lowers a key, so the loop
    // ends; the cap only turns an inconsistency into an error instead of a hang.
    const uint64_t max_rounds = 256ull * max_extent + 4096;
    // Batches of rounds are enqueued one ahead of the one being checked: while
    // the host waits for batch k's frontier size, batch k+1 is already on the
    // GPU. Once the front thins out the batches shrink, so fewer empty rounds
    // run after the last one that took a cell.
    BatchProbe probe;
    int cur = 0;
    uint64_t round = 0;
    // Once the front is thin (the last reported round took fewer than 2,048
    // cells: at most 53,248 item-neighbour pairs), 256 blocks cover every pair
    // in one pass and the other blocks would only cost dispatch time. The
    // propagation's loops are grid-strided, so a round that grows again is
    // still complete.
    uint32_t last_taken = UINT32_MAX;
    // a round's frontier holds at most ncells items; each propagation launch
    // takes prop_chunk_items(kmax) of them (one launch below 130M cells)
    const uint32_t chunk_items = prop_chunk_items(kmax);
    const uint32_t prop_chunks = (uint32_t)((ncells + chunk_items - 1) / chunk_items);
    auto enqueue_batch = [&](int slot) {
        const uint32_t seq = probe.arm(slot);
        const unsigned blocks = last_taken < 2048u ? std::min(prop_blocks, 256u) : prop_blocks;
        // large rounds propagate item-major (the same keys, fewer evaluations);
        // before the first frontier size is known, a sixteenth of the cells
        // stands for it (grid_item_major_min = 0 forces it everywhere)
        const uint64_t round_cells = last_taken == UINT32_MAX ? ncells / 16 : last_taken;
        const int64_t im_min = tuning().grid_item_major_min;
        const bool item_major = im_min == 0 || round_cells >= (uint64_t)im_min;
        for (int b = 0; b < batch; ++b, ++round) {
            SDFX_LAUNCH(k_frontier_compact<MODE>, dim3(compact_blocks), dim3(256), 0, s, keys, rv, tile_req, ntiles,
                        changed, tile_changed, items, region, ctr[cur], ctr[cur ^ 1], minw[cur], minw[cur ^ 1],
                        taken[cur ^ 1], taken[cur], released, thin_round, (uint32_t)(round + 1), (int)SIGN);
            for (uint32_t k = 0; k < prop_chunks; ++k) {
                uint32_t* probe_w = b == batch - 1 && k + 1 == prop_chunks ? probe.probe(slot) : nullptr;
                if (item_major)
                    SDFX_LAUNCH(k_frontier_propagate_items<MODE>, dim3(blocks), dim3(256), 0, s, items, region,
                                ctr[cur], rv, nreq, keys, dirs, changed, tile_changed, minw[cur ^ 1], minw[cur],
                                taken[cur ^ 1], totals + cur, probe_w, seq, k * chunk_items, kmax, d_flags,
                                evals ? evals + seed_slots : nullptr);
                else
                    SDFX_LAUNCH(k_frontier_propagate<MODE>, dim3(blocks), dim3(256), 0, s, items, region, ctr[cur],
                                rv, nreq, keys, dirs, changed, tile_changed, minw[cur ^ 1], minw[cur], taken[cur ^ 1],
                                totals + cur, probe_w, seq, k * chunk_items, kmax, d_flags,
                                evals ? evals + seed_slots : nullptr);
            }
            cur ^= 1;
        }
        SDFX_HIP(hipGetLastError());
    };
    enqueue_batch(0);
    for (int slot = 0;; slot ^= 1) {
        if (round > max_rounds) throw Error(SDFX_ERR_DEVICE, "grid propagation did not converge");
        enqueue_batch(slot ^ 1);
        const uint32_t taken_cells = probe.wait(slot, s);
        last_taken = taken_cells;
        if (taken_cells == 0) break;  // the batch after it only runs empty rounds
        batch = taken_cells < 512 ? 2 : (taken_cells < 8192 ? 4 : 8);
    }
    }  // frontier rounds

    // 3. sign
    uint32_t* bits[3] = {nullptr, nullptr, nullptr};
    if (SIGN == SDFX_SIGN_RAYCAST) {
        bits[0] = words + runs.end[4];  // cleared by k_grid_init
        bits[1] = bits[0] + words_total[0];
        bits[2] = bits[1] + words_total[1];
        RayBig* big = A.alloc<RayBig>(3 * nt);
        uint32_t* n_big = totals + 2 + nreq;  // cleared by k_grid_init
        SDFX_LAUNCH(k_ray_small<MODE>, dim3(blocks_for(3 * nt, 256)), dim3(256), 0, s, nt, rv, tri_req, big, n_big,
                    bits[0], bits[1], bits[2]);
        SDFX_LAUNCH(k_ray_big<MODE>, dim3(1024), dim3(256), 0, s, rv, tri_req, big, n_big, bits[0], bits[1], bits[2]);
        SDFX_LAUNCH(k_ray_suffix<MODE>, dim3(blocks_for(rows_total[0] + rows_total[1] + rows_total[2], 256)),
                    dim3(256), 0, s, bits[0], bits[1], bits[2], rows_total[0], rows_total[1], rows_total[2], rv,
                    nreq);
        SDFX_HIP(hipGetLastError());
    }
    SDFX_LAUNCH(k_grid_finish<MODE>, dim3(stride_blocks(ncells)), dim3(256), 0, s, keys, ncells, rv, tile_req,
                bits[0], bits[1], bits[2]);
    SDFX_HIP(hipGetLastError());
    if (evals) {
        profile_count("grid.seed_distance_evals", sum_block_counters(evals, seed_slots, s));
        profile_count("grid.propagate_distance_evals", sum_block_counters(evals + seed_slots, prop_slots, s));
    }
}

// Tile-padded cells one run of a batch may hold: a run's device state is about
// 25 bytes per cell (keys, flags, frontier items), so larger batches run as
// several runs, each with its own scratch returned before the next starts; a
// single request beyond it runs alone (up to the 2^32 cells the entry points
// admit).
constexpr uint64_t RUN_MAX_CELLS = 1ull << 28;

static uint64_t padded_cells(const GridParams& g) {
    const uint64_t nc = g.count[0] * g.count[1] * g.count[2];
    return (nc + COMPACT_TILE - 1) / COMPACT_TILE * COMPACT_TILE;
}

template <int SIGN>
static void grid_run_any(uint64_t nt, const std::vector<GridBatchItem>& jobs, DevFlags* d_flags, Arena& A) {
    if (jobs.size() > 1) grid_run<SIGN | 2>(nt, jobs, d_flags, A);
    else grid_run<SIGN>(nt, jobs, d_flags, A);
}

void grid_sdf_batch(const std::vector<GridBatchItem>& items, sdfx_sign_method sign, DevFlags* d_flags, Arena& A,
                    std::vector<std::pair<sdfx_status, std::string>>* run_errors) {
    if (run_errors) run_errors->assign(items.size(), {SDFX_OK, std::string()});
    // requests with cells and triangles, in order, cut into runs
    std::vector<uint64_t> cells(items.size(), 0);
    for (size_t i = 0; i < items.size(); ++i) {
        const GridBatchItem& it = items[i];
        const uint64_t ncells = it.gp.count[0] * it.gp.count[1] * it.gp.count[2];
        if (!ncells) continue;
        if (it.n_tris == 0) {
            // no triangle: every cell keeps f32::MAX (grid.rs:137-143), no ray hits
            SDFX_LAUNCH(k_fill_u32g, dim3(stride_blocks(ncells)), dim3(256), 0, A.stream, (uint32_t*)it.d_out, ncells,
                        fbits(FLT_MAX));
            continue;
        }
        cells[i] = padded_cells(it.gp);
    }
    SDFX_HIP(hipGetLastError());
    const std::vector<std::vector<size_t>> runs = plan_grid_runs(cells, RUN_MAX_CELLS);
    for (const std::vector<size_t>& run : runs) {
        // triangle numbering local to the run
        uint64_t base = UINT64_MAX, nt = 0;
        for (size_t i : run) base = std::min(base, items[i].tri_off);
        std::vector<GridBatchItem> jobs;
        for (size_t i : run) {
            GridBatchItem j = items[i];
            j.tri_off -= base;
            nt = std::max(nt, j.tri_off + j.n_tris);
            jobs.push_back(j);
        }
        try {
            Arena R(A.stream);  // the run's scratch, returned (after a sync) before the next run
            if (sign == SDFX_SIGN_NORMAL) grid_run_any<SDFX_SIGN_NORMAL>(nt, jobs, d_flags, R);
            else grid_run_any<SDFX_SIGN_RAYCAST>(nt, jobs, d_flags, R);
        } catch (const Error& e) {
            if (!run_errors) throw;
            for (size_t i : run) (*run_errors)[i] = {e.status, e.what()};
        }
    }
}

void grid_sdf(const TriRec* d_tris, uint64_t nt, const GridParams& gp, sdfx_sign_method sign, float* d_out,
              DevFlags* d_flags, Arena& A) {
    grid_sdf_batch({GridBatchItem{d_tris, 0, nt, gp, d_out, 0}}, sign, d_flags, A, nullptr);
}

void grid_sdf_exact(const TriRec* d_tris, uint64_t nt, const GridParams& gp, sdfx_sign_method sign, float* d_out,
                    DevFlags* d_flags, Arena& A) {
    const uint64_t ncells = gp.count[0] * gp.count[1] * gp.count[2];
    if (!ncells || !nt) {  // the same empty-grid and empty-mesh conventions
        grid_sdf(d_tris, nt, gp, sign, d_out, d_flags, A);
        return;
    }
    Arena R(A.stream);  // the tree, the queries and the values, returned after the run
    Lbvh bvh;
    lbvh_build(d_tris, nt, bvh, R);
    GridDev g{};
    for (int k = 0; k < 3; ++k) {
        g.first[k] = gp.first[k];
        g.size[k] = gp.size[k];
        g.n[k] = (uint32_t)gp.count[k];
    }
    float* q = R.alloc<float>(3 * ncells);
    SDFX_LAUNCH(k_cell_centres, dim3(stride_blocks(ncells)), dim3(256), 0, R.stream, g, ncells, q);
    float* values = R.alloc<float>(ncells);
    // Raycast: nearest by (d^2, topology index), signed distance to it (its
    // magnitude is the reference's point_triangle_distance to the nearest
    // triangle); Normal: the compare_distances fold over every triangle
    lbvh_query(bvh, q, ncells, sign == SDFX_SIGN_NORMAL ? QM_NORMAL_FOLD : QM_NEAREST_SIGNED, values, d_flags, R);
    GridBatchItem it{d_tris, 0, nt, gp, d_out, 0};
    it.exact = values;
    if (sign == SDFX_SIGN_NORMAL) grid_run<SDFX_SIGN_NORMAL>(nt, {it}, d_flags, R);
    else grid_run<SDFX_SIGN_RAYCAST>(nt, {it}, d_flags, R);
}

}  // namespace sdfx
