// sdfx_none.hip — AccelerationMethod::None (generate/generic/default.rs:22-85).
//
// One lane per query; the triangle loop index is wave-uniform, so each
// TriRec is fetched with scalar loads into SGPRs and shared by the 64 lanes
// (no LDS staging needed on CDNA: VALU operands read SGPRs directly).
//
// Raycast: min over triangles of the distance (the min of sqrt(d^2) equals
// sqrt of the min d^2, so sqrt runs once per query) and the parity of +X
// aligned-ray hits (default.rs:45-49, 76-83). The triangle range may be split
// across blocks (min via atomicMin on the d^2 bit pattern, which orders
// non-negative floats, and hit counts via atomicAdd); both reductions are exact.
// Normal: the compare_distances fold runs in triangle order in one lane per
// query (default.rs:55-74), so non-transitive near-ties resolve exactly as in
// the reference.
#include "sdfx_internal.hpp"

namespace sdfx {

constexpr int NONE_THREADS = 256;

template <bool CHUNKED>
__global__ void __launch_bounds__(NONE_THREADS) k_none_raycast(const TriRec* __restrict__ tris, uint64_t nt,
                                                                uint64_t chunk, const float* __restrict__ q,
                                                                uint64_t nq, float* __restrict__ out,
                                                                uint32_t* __restrict__ best_bits,
                                                                uint32_t* __restrict__ counts) {
    uint64_t i = (uint64_t)blockIdx.x * NONE_THREADS + threadIdx.x;
    bool active = i < nq;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (active) {
        px = q[3 * i];
        py = q[3 * i + 1];
        pz = q[3 * i + 2];
    }
    uint64_t t0 = CHUNKED ? (uint64_t)blockIdx.y * chunk : 0;
    uint64_t t1 = CHUNKED ? (t0 + chunk < nt ? t0 + chunk : nt) : nt;
    float best = INFINITY;
    uint32_t hits = 0;
    for (uint64_t t = t0; t < t1; ++t) {
        const TriRec tr = tris[t];
        float dx, dy, dz, cx, cy, cz;
        closest_diff_flat(px, py, pz, tr, dx, dy, dz, cx, cy, cz);
        float d2 = dot3(dx, dy, dz, dx, dy, dz);
        if (d2 < best) best = d2;  // NaN never replaces (f32::min ignores NaN)
        hits += ray_hit<0>(px, py, pz, tr, nullptr) ? 1u : 0u;
    }
    if (!active) return;
    if (CHUNKED) {
        if (best < INFINITY) atomicMin(&best_bits[i], fbits(best));
        if (hits) atomicAdd(&counts[i], hits);
    } else {
        float d = fminf(FLT_MAX, sqrtf(best));
        out[i] = (hits & 1u) ? -d : d;
    }
}

__global__ void __launch_bounds__(NONE_THREADS) k_none_raycast_finish(const uint32_t* __restrict__ best_bits,
                                                                       const uint32_t* __restrict__ counts,
                                                                       uint64_t nq, float* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * NONE_THREADS + threadIdx.x;
    if (i >= nq) return;
    float d = fminf(FLT_MAX, sqrtf(bitsf(best_bits[i])));
    out[i] = (counts[i] & 1u) ? -d : d;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(NONE_THREADS) k_none_normal(const TriRec* __restrict__ tris, uint64_t nt,
                                                               const float* __restrict__ q, uint64_t nq,
                                                               float* __restrict__ out, DevFlags* flags) {
    uint64_t i = (uint64_t)blockIdx.x * NONE_THREADS + threadIdx.x;
    bool active = i < nq;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (active) {
        px = q[3 * i];
        py = q[3 * i + 1];
        pz = q[3 * i + 2];
    }
    float acc = FLT_MAX;
    bool nan_seen = false;
    for (uint64_t t = 0; t < nt; ++t) {
        const TriRec tr = tris[t];
        float d = tri_signed_dist(px, py, pz, tr);
        int c = compare_distances(d, acc);  // default.rs:64: keep `distance` when Less
        nan_seen |= c == 2;
        if (c == -1) acc = d;
    }
    if (!active) return;
    if (nan_seen) atomicOr(&flags->nan_distance, 1u);
    out[i] = acc;
}

// ------------------------------------------------------------------ exhaustive
// forms of the accelerated modes. By the conservativeness of the reference's
// BVH/R-tree searches their results are functions of all triangles: the
// minimum d^2 (lowest topology index on ties) and the +X/+Y/+Z crossing
// parities. For small query x triangle counts these kernels replace the LBVH
// (no build, every CU busy through triangle-range splitting).

// Per query: min d^2 (as bits, with the topology index in the low word for the
// nearest-triangle modes) and a 3-bit crossing-parity mask.
// KIND: 1 = keep the nearest triangle's index, 2 = accumulate crossing parities
template <int KIND>
__global__ void __launch_bounds__(NONE_THREADS) k_exhaustive(const TriRec* __restrict__ tris, uint64_t nt,
                                                              uint64_t chunk, const float* __restrict__ q,
                                                              uint64_t nq, unsigned long long* __restrict__ best,
                                                              uint32_t* __restrict__ parity) {
    uint64_t i = (uint64_t)blockIdx.x * NONE_THREADS + threadIdx.x;
    bool active = i < nq;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (active) {
        px = q[3 * i];
        py = q[3 * i + 1];
        pz = q[3 * i + 2];
    }
    uint64_t t0 = (uint64_t)blockIdx.y * chunk;
    uint64_t t1 = t0 + chunk < nt ? t0 + chunk : nt;
    float bd = INFINITY;
    uint32_t bi = 0xFFFFFFFFu;
    uint32_t par = 0;
    constexpr bool WANT_INDEX = (KIND & 1) != 0;
    constexpr bool WANT_PARITY = (KIND & 2) != 0;
    for (uint64_t t = t0; t < t1; ++t) {
        const TriRec tr = tris[t];
        // The tests the LBVH traversal makes at its leaves: the exact distance
        // only where the padded box is not clearly farther than the best so far
        // (such a triangle cannot be the first of the minima), a ray only where
        // it meets the padded box (a hit point lies inside it).
        Box b;
        tri_bbox(tr.ax, tr.ay, tr.az, tr.bx, tr.by, tr.bz, tr.cx, tr.cy, tr.cz, b);
        if (box_dist2(px, py, pz, b.lo, b.hi) * PRUNE_SLACK <= bd) {
            float dx, dy, dz, cx, cy, cz;
            closest_diff_flat(px, py, pz, tr, dx, dy, dz, cx, cy, cz);
            float d2 = dot3(dx, dy, dz, dx, dy, dz);
            if (d2 < bd) {  // ascending t: the first (lowest index) of equal minima stays
                bd = d2;
                bi = (uint32_t)t;
            }
        }
        if (WANT_PARITY) {
            const uint32_t axes = ray_axes(b.lo, b.hi, px, py, pz);
            if ((axes & 1u) && ray_hit<0>(px, py, pz, tr, nullptr)) par ^= 1u;
            if ((axes & 2u) && ray_hit<1>(px, py, pz, tr, nullptr)) par ^= 2u;
            if ((axes & 4u) && ray_hit<2>(px, py, pz, tr, nullptr)) par ^= 4u;
        }
    }
    if (!active) return;
    if (bd < INFINITY) {
        unsigned long long key = ((unsigned long long)fbits(bd) << 32) | (WANT_INDEX ? bi : 0u);
        if (key < best[i]) atomicMin(&best[i], key);
    }
    if (WANT_PARITY && par) atomicXor(&parity[i], par);
}

template <int MODE>
__global__ void __launch_bounds__(NONE_THREADS) k_exhaustive_finish(const TriRec* __restrict__ tris,
                                                                     const float* __restrict__ q, uint64_t nq,
                                                                     const unsigned long long* __restrict__ best,
                                                                     const uint32_t* __restrict__ parity,
                                                                     float* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * NONE_THREADS + threadIdx.x;
    if (i >= nq) return;
    unsigned long long b = best[i];
    float d2 = bitsf((uint32_t)(b >> 32));
    if (MODE == QM_NEAREST_SIGNED) {
        uint32_t t = (uint32_t)b;
        out[i] = b != ~0ull ? tri_signed_dist(q[3 * i], q[3 * i + 1], q[3 * i + 2], tris[t]) : sqrtf(d2);
        return;
    }
    float d = sqrtf(b == ~0ull ? INFINITY : d2);
    if (MODE == QM_RAYCAST_VOTE) d = fminf(FLT_MAX, d);
    uint32_t p = parity[i];
    int insides = (int)(p & 1u) + (int)((p >> 1) & 1u) + (int)((p >> 2) & 1u);
    out[i] = insides > 1 ? -d : d;
}

__global__ void k_fill_u64n(unsigned long long* p, uint64_t n, unsigned long long v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void exhaustive_query(const TriRec* d_tris, uint64_t nt, const float* d_q, uint64_t nq, QueryMode mode,
                      float* d_out, DevFlags* d_flags, Arena& A) {
    if (!nq) return;
    hipStream_t s = A.stream;
    if (mode == QM_NORMAL_FOLD) {  // the fold is order-dependent: one lane folds all triangles in order
        none_query(d_tris, nt, d_q, nq, SDFX_SIGN_NORMAL, d_out, d_flags, A);
        return;
    }
    unsigned qblocks = blocks_for(nq, NONE_THREADS);
    const uint64_t target_blocks = 2048, min_chunk = 32;
    uint64_t chunks = 1;
    if (qblocks < target_blocks && nt > min_chunk) {
        chunks = (target_blocks + qblocks - 1) / qblocks;
        uint64_t max_chunks = (nt + min_chunk - 1) / min_chunk;
        if (chunks > max_chunks) chunks = max_chunks;
        if (chunks > 65535) chunks = 65535;
    }
    uint64_t chunk = nt ? (nt + chunks - 1) / chunks : 1;
    chunks = nt ? (nt + chunk - 1) / chunk : 1;
    unsigned long long* best = A.alloc<unsigned long long>(nq);
    uint32_t* parity = A.alloc<uint32_t>(nq);
    SDFX_LAUNCH(k_fill_u64n, dim3(blocks_for(nq, 256)), dim3(256), 0, s, best, nq, ~0ull);
    SDFX_LAUNCH(k_fill_u32, dim3(blocks_for(nq, 256)), dim3(256), 0, s, parity, nq, 0u);
    dim3 grid(qblocks, (unsigned)chunks);
    if (mode == QM_NEAREST_SIGNED)
        SDFX_LAUNCH(k_exhaustive<1>, grid, dim3(NONE_THREADS), 0, s, d_tris, nt, chunk, d_q, nq, best,
                    parity);
    else
        SDFX_LAUNCH(k_exhaustive<2>, grid, dim3(NONE_THREADS), 0, s, d_tris, nt, chunk, d_q, nq, best,
                    parity);
    switch (mode) {
        case QM_NEAREST_SIGNED:
            SDFX_LAUNCH(k_exhaustive_finish<QM_NEAREST_SIGNED>, dim3(qblocks), dim3(NONE_THREADS), 0, s, d_tris, d_q,
                        nq, best, parity, d_out);
            break;
        case QM_RAYCAST_VOTE:
            SDFX_LAUNCH(k_exhaustive_finish<QM_RAYCAST_VOTE>, dim3(qblocks), dim3(NONE_THREADS), 0, s, d_tris, d_q,
                        nq, best, parity, d_out);
            break;
        default:
            SDFX_LAUNCH(k_exhaustive_finish<QM_NEAREST_VOTE>, dim3(qblocks), dim3(NONE_THREADS), 0, s, d_tris, d_q,
                        nq, best, parity, d_out);
            break;
    }
    SDFX_HIP(hipGetLastError());
}

void none_query(const TriRec* d_tris, uint64_t nt, const float* d_q, uint64_t nq, sdfx_sign_method sign,
                float* d_out, DevFlags* d_flags, Arena& A) {
    if (!nq) return;
    hipStream_t s = A.stream;
    unsigned qblocks = blocks_for(nq, NONE_THREADS);
    if (sign == SDFX_SIGN_NORMAL) {
        SDFX_LAUNCH(k_none_normal, dim3(qblocks), dim3(NONE_THREADS), 0, s, d_tris, nt, d_q, nq, d_out,
                           d_flags);
        SDFX_HIP(hipGetLastError());
        return;
    }
    // Split the triangle range so that small query sets still fill the 256 CUs.
    const uint64_t target_blocks = 2048, min_chunk = 64;
    uint64_t chunks = 1;
    if (qblocks < target_blocks && nt > min_chunk) {
        chunks = (target_blocks + qblocks - 1) / qblocks;
        uint64_t max_chunks = (nt + min_chunk - 1) / min_chunk;
        if (chunks > max_chunks) chunks = max_chunks;
        if (chunks > 65535) chunks = 65535;
    }
    if (chunks <= 1) {
        SDFX_LAUNCH(k_none_raycast<false>, dim3(qblocks), dim3(NONE_THREADS), 0, s, d_tris, nt, nt, d_q, nq,
                           d_out, nullptr, nullptr);
        SDFX_HIP(hipGetLastError());
        return;
    }
    uint64_t chunk = (nt + chunks - 1) / chunks;
    chunks = (nt + chunk - 1) / chunk;
    uint32_t* best = A.alloc<uint32_t>(nq);
    uint32_t* counts = A.alloc<uint32_t>(nq);
    SDFX_LAUNCH(k_fill_u32, dim3(blocks_for(nq, 256)), dim3(256), 0, s, best, nq, 0x7f800000u);
    SDFX_LAUNCH(k_fill_u32, dim3(blocks_for(nq, 256)), dim3(256), 0, s, counts, nq, 0u);
    SDFX_LAUNCH(k_none_raycast<true>, dim3(qblocks, (unsigned)chunks), dim3(NONE_THREADS), 0, s, d_tris, nt,
                       chunk, d_q, nq, d_out, best, counts);
    SDFX_LAUNCH(k_none_raycast_finish, dim3(qblocks), dim3(NONE_THREADS), 0, s, best, counts, nq, d_out);
    SDFX_HIP(hipGetLastError());
}

}  // namespace sdfx
