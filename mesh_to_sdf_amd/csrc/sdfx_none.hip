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
        closest_diff(px, py, pz, tr, dx, dy, dz, cx, cy, cz);
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
