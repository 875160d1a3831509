// sdfx_mesh_check.hip — check_mesh_triangle_list (generate/mesh_check.rs:24-81)
// on the device, with the reference's verdict and message.
//
// The reference runs three O(n^2) scans in order and returns the first
// failure:
//  1. the first vertex equal to an earlier one, equality being that of
//     `format!("{:?}_{:?}_{:?}")` (:37-43): bitwise, except that every NaN
//     prints "NaN";
//  2. the first face equal to an earlier one as an ordered triple (:46-49);
//  3. the edge loop (:52-78): edges (f0,f1), (f1,f2), (f2,f0) per face in face
//     order; it pops the last edge, fails if a copy of it remains ("Duplicate
//     edge") or if its reverse does not ("not waterproof"), and removes every
//     copy of the reverse.
// Here each scan is a stable sort by key plus a pass over the sorted runs:
//  1, 2. the first duplicate is the smallest index among run members that are
//     not the first of their run (a stable sort keeps each run in index order);
//  3. a directed edge value and its reverse only ever interact with each other,
//     and the one whose last copy has the larger index is popped first. If it
//     has another copy, that pop fails as a duplicate; if its reverse is absent
//     (or it is a loop (a, a)), it fails as not waterproof; otherwise it
//     removes the reverse with all its copies and the pair is done. The loop
//     reports the failing pair whose popped copy has the largest index.
#include <algorithm>
#include <cstring>
#include <string>

#include "sdfx_internal.hpp"

namespace sdfx {

namespace {

__global__ void __launch_bounds__(256) k_vertex_cols(const float* __restrict__ v, uint64_t n, uint32_t* __restrict__ cx,
                                                     uint32_t* __restrict__ cy, uint32_t* __restrict__ cz) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* cols[3] = {cx, cy, cz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float x = v[3 * i + k];
        cols[k][i] = x != x ? 0x7fc00000u : fbits(x);  // every NaN prints "NaN"
    }
}

// Face columns (f0, f1, f2) or directed-edge columns (from, to) of edge 3f + k.
__global__ void __launch_bounds__(256) k_face_cols(const uint32_t* __restrict__ faces, uint64_t nt, int edges,
                                                   uint32_t* __restrict__ c0, uint32_t* __restrict__ c1,
                                                   uint32_t* __restrict__ c2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (edges) {
        if (i >= 3 * nt) return;
        const uint64_t f = i / 3;
        const uint32_t k = (uint32_t)(i - 3 * f);
        c0[i] = faces[3 * f + k];
        c1[i] = faces[3 * f + (k + 1) % 3];
    } else {
        if (i >= nt) return;
        c0[i] = faces[3 * i];
        c1[i] = faces[3 * i + 1];
        c2[i] = faces[3 * i + 2];
    }
}

__global__ void __launch_bounds__(256) k_iota(uint32_t* __restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(256) k_gather_cols(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm,
                                                     uint64_t n, uint32_t* __restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// Smallest original index among sorted positions i whose key equals the key at
// i - 1 (UINT64_MAX when there is none).
__global__ void __launch_bounds__(256) k_first_duplicate(const uint32_t* __restrict__ perm, uint64_t n,
                                                         const uint32_t* __restrict__ c0, const uint32_t* __restrict__ c1,
                                                         const uint32_t* __restrict__ c2,
                                                         unsigned long long* __restrict__ best) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cand = ~0ull;
    if (i > 0 && i < n) {
        const uint32_t a = perm[i], b = perm[i - 1];
        if (c0[a] == c0[b] && c1[a] == c1[b] && c2[a] == c2[b]) cand = a;
    }
    for (int off = 32; off > 0; off >>= 1) cand = min(cand, (unsigned long long)__shfl_xor(cand, off, 64));
    if ((threadIdx.x & 63u) == 0 && cand != ~0ull) atomicMin(best, cand);
}

// One lane per run of equal directed edges in the (from, to)-sorted list (sa,
// sb); perm gives each position's edge index. A failing pair folds
// (index of the popped copy << 3 | kind) into *worst with atomicMax.
__global__ void __launch_bounds__(256) k_edge_runs(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ sa,
                                                   const uint32_t* __restrict__ sb, uint64_t n,
                                                   unsigned long long* __restrict__ worst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cand = 0;
    if (i < n && !(i > 0 && sa[i] == sa[i - 1] && sb[i] == sb[i - 1])) {
        const uint32_t a = sa[i], b = sb[i];
        uint64_t t = i + 1;
        while (t < n && sa[t] == a && sb[t] == b) ++t;
        const uint64_t count = t - i;
        const uint32_t hi = perm[t - 1];  // the last copy: runs keep index order
        // the reverse run (b, a): lower bound in the sorted list
        uint64_t lo = 0, up = n;
        while (lo < up) {
            const uint64_t mid = (lo + up) >> 1;
            if (sa[mid] < b || (sa[mid] == b && sb[mid] < a)) lo = mid + 1;
            else up = mid;
        }
        const bool loop = a == b;
        const bool has_rev = !loop && lo < n && sa[lo] == b && sb[lo] == a;
        uint32_t rev_hi = 0;
        if (has_rev) {
            uint64_t r = lo;
            while (r + 1 < n && sa[r + 1] == b && sb[r + 1] == a) ++r;
            rev_hi = perm[r];
        }
        if (!has_rev || hi > rev_hi) {  // this value is popped first
            const uint32_t kind = count >= 2 ? 4u : (!has_rev ? 5u : 0u);
            if (kind) cand = ((unsigned long long)hi << 3) | kind;
        }
    }
    for (int off = 32; off > 0; off >>= 1) cand = max(cand, (unsigned long long)__shfl_xor(cand, off, 64));
    if ((threadIdx.x & 63u) == 0 && cand) atomicMax(worst, cand);
}

// Sorts positions 0..n-1 by the columns (most significant first), stably: one
// LSD radix pass per column from the least significant.
void sort_by_columns(uint32_t* const* cols, int ncols, uint64_t n, uint32_t* perm, Arena& A) {
    hipStream_t s = A.stream;
    uint32_t* keys = A.alloc<uint32_t>(n);
    SDFX_LAUNCH(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, s, perm, n);
    for (int c = ncols - 1; c >= 0; --c) {
        SDFX_LAUNCH(k_gather_cols, dim3(blocks_for(n, 256)), dim3(256), 0, s, cols[c], perm, n, keys);
        radix_sort_pairs(keys, perm, n, 32, A);
    }
    SDFX_HIP(hipGetLastError());
}

std::string face_text(const uint32_t* f) {
    return "[" + std::to_string(f[0]) + ", " + std::to_string(f[1]) + ", " + std::to_string(f[2]) + "]";
}

}  // namespace

MeshCheck check_mesh_triangle_list(const float* vertices, uint64_t nv, const uint32_t* faces, uint64_t nf,
                                   hipStream_t stream) {
    MeshCheck res;
    if (nf % 3 != 0) {  // mesh_check.rs:30-32
        res.kind = 1;
        res.message = "Faces must be a multiple of 3";
        return res;
    }
    const uint64_t nt = nf / 3;
    if (nv >= (1ull << 32) || 3 * nt >= (1ull << 32))
        throw Error(SDFX_ERR_INVALID_ARGUMENT, "mesh check supports fewer than 2^32 vertices and edges");
    Arena A(stream);
    unsigned long long* words = A.alloc<unsigned long long>(3);  // first dup vertex, first dup face, worst edge
    fill_u32(reinterpret_cast<uint32_t*>(words), 4, UINT32_MAX, stream);
    fill_u32(reinterpret_cast<uint32_t*>(words + 2), 2, 0u, stream);
    float* d_v = nullptr;
    uint32_t* d_f = nullptr;
    if (nv) {
        d_v = A.alloc<float>(3 * nv);
        SDFX_HIP(hipMemcpyAsync(d_v, vertices, 12 * nv, hipMemcpyHostToDevice, stream));
        uint32_t* cols[3] = {A.alloc<uint32_t>(nv), A.alloc<uint32_t>(nv), A.alloc<uint32_t>(nv)};
        uint32_t* perm = A.alloc<uint32_t>(nv);
        SDFX_LAUNCH(k_vertex_cols, dim3(blocks_for(nv, 256)), dim3(256), 0, stream, d_v, nv, cols[0], cols[1],
                    cols[2]);
        sort_by_columns(cols, 3, nv, perm, A);
        SDFX_LAUNCH(k_first_duplicate, dim3(blocks_for(nv, 256)), dim3(256), 0, stream, perm, nv, cols[0], cols[1],
                    cols[2], words);
    }
    if (nt) {
        d_f = A.alloc<uint32_t>(nf);
        SDFX_HIP(hipMemcpyAsync(d_f, faces, 4 * nf, hipMemcpyHostToDevice, stream));
        uint32_t* cols[3] = {A.alloc<uint32_t>(nt), A.alloc<uint32_t>(nt), A.alloc<uint32_t>(nt)};
        uint32_t* perm = A.alloc<uint32_t>(nt);
        SDFX_LAUNCH(k_face_cols, dim3(blocks_for(nt, 256)), dim3(256), 0, stream, d_f, nt, 0, cols[0], cols[1],
                    cols[2]);
        sort_by_columns(cols, 3, nt, perm, A);
        SDFX_LAUNCH(k_first_duplicate, dim3(blocks_for(nt, 256)), dim3(256), 0, stream, perm, nt, cols[0], cols[1],
                    cols[2], words + 1);
        const uint64_t ne = 3 * nt;
        uint32_t* ecols[2] = {A.alloc<uint32_t>(ne), A.alloc<uint32_t>(ne)};
        uint32_t* eperm = A.alloc<uint32_t>(ne);
        SDFX_LAUNCH(k_face_cols, dim3(blocks_for(ne, 256)), dim3(256), 0, stream, d_f, nt, 1, ecols[0], ecols[1],
                    (uint32_t*)nullptr);
        sort_by_columns(ecols, 2, ne, eperm, A);
        uint32_t* sa = A.alloc<uint32_t>(ne);
        uint32_t* sb = A.alloc<uint32_t>(ne);
        SDFX_LAUNCH(k_gather_cols, dim3(blocks_for(ne, 256)), dim3(256), 0, stream, ecols[0], eperm, ne, sa);
        SDFX_LAUNCH(k_gather_cols, dim3(blocks_for(ne, 256)), dim3(256), 0, stream, ecols[1], eperm, ne, sb);
        SDFX_LAUNCH(k_edge_runs, dim3(blocks_for(ne, 256)), dim3(256), 0, stream, eperm, sa, sb, ne, words + 2);
    }
    SDFX_HIP(hipGetLastError());
    unsigned long long h[3];
    SDFX_HIP(hipMemcpyAsync(h, words, sizeof h, hipMemcpyDeviceToHost, stream));
    SDFX_HIP(hipStreamSynchronize(stream));
    if (h[0] != ~0ull) {  // mesh_check.rs:37-43
        const float* p = vertices + 3 * h[0];
        res.kind = 2;
        res.where = h[0];
        res.message = "Duplicate vertex found at (" + rust_debug_f32(p[0]) + ", " + rust_debug_f32(p[1]) + ", " +
                      rust_debug_f32(p[2]) + ")";
    } else if (h[1] != ~0ull) {  // mesh_check.rs:46-49
        res.kind = 3;
        res.where = h[1];
        res.message = "Duplicate face found: " + face_text(faces + 3 * h[1]);
    } else if (h[2] != 0ull) {  // mesh_check.rs:62-77
        res.kind = (int)(h[2] & 7u);
        res.where = h[2] >> 3;
        const uint64_t f = res.where / 3, k = res.where % 3;
        const std::string a = std::to_string(faces[3 * f + k]), b = std::to_string(faces[3 * f + (k + 1) % 3]);
        if (res.kind == 4) res.message = "Duplicate edge " + a + " -> " + b + " - do you have an inverted normal?";
        else
            res.message = "Mesh is not waterproof - edge " + a + " -> " + b + " does not have a corresponding edge " +
                          b + " -> " + a;
    }
    return res;
}

}  // namespace sdfx
