// sdfx_api.cpp — the C ABI (include/sdfx.h). Conditions are checked in the
// order the reference would panic on them: lib.rs:374-376 (finite queries),
// lib.rs:202-208 (index / vertex counts), then vertex reads
// (index out of range), then the path-specific cases.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <sstream>
#include <vector>

#include "sdfx_internal.hpp"

using namespace sdfx;

struct sdfx_mesh {
    int device = 0;
    sdfx_acceleration_method method = SDFX_ACCEL_RTREE_BVH;
    sdfx_sign_method sign = SDFX_SIGN_RAYCAST;
    uint64_t n_tris = 0;
    // Error the reference raises on the first query with at least one point
    // (the None path decodes its topology per query, default.rs:41).
    sdfx_status deferred = SDFX_OK;
    std::string deferred_msg;
    DevBuf<TriRec> tris;  // topology order (None path)
    Lbvh bvh;             // Bvh / Rtree / RtreeBvh
};

namespace {

template <typename F>
sdfx_status guarded(F&& f) {
    try {
        clear_last_error();
        return f();
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return SDFX_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return SDFX_ERR_DEVICE;
    } catch (...) {
        set_last_error("unknown error");
        return SDFX_ERR_DEVICE;
    }
}

[[noreturn]] void fail(sdfx_status s, const std::string& msg) { throw Error(s, msg); }


std::string topology_error_message(sdfx_status st, uint64_t nv, uint64_t ni) {
    std::ostringstream o;
    if (st == SDFX_ERR_INDEX_COUNT) o << "TriangleList indices length (" << ni << ") must be divisible by 3";
    else if (st == SDFX_ERR_VERTEX_COUNT)
        o << "TriangleList vertices length (" << nv << ") must be divisible by 3 when no indices are provided";
    else o << "invalid topology";
    return o.str();
}

void check_enums(sdfx_topology topo, int method, int sign) {
    if (topo != SDFX_TRIANGLE_LIST && topo != SDFX_TRIANGLE_STRIP) fail(SDFX_ERR_INVALID_ARGUMENT, "unknown topology");
    if (method < SDFX_ACCEL_NONE || method > SDFX_ACCEL_RTREE_BVH)
        fail(SDFX_ERR_INVALID_ARGUMENT, "unknown acceleration method");
    if (sign != SDFX_SIGN_RAYCAST && sign != SDFX_SIGN_NORMAL) fail(SDFX_ERR_INVALID_ARGUMENT, "unknown sign method");
}

// lib.rs:374-376 on host memory, with the reference's message.
void host_check_finite(const float* q, uint64_t nq) {
    for (uint64_t i = 0; i < nq; ++i) {
        const float* p = q + 3 * i;
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]))
            fail(SDFX_ERR_NON_FINITE_QUERY, "Query point [" + rust_debug_f32(p[0]) + ", " + rust_debug_f32(p[1]) + ", " +
                                                rust_debug_f32(p[2]) + "] contains non-finite values");
    }
}

DevFlags* new_flags(Arena& A) {
    DevFlags* f = A.alloc<DevFlags>(1);
    fill_u32(reinterpret_cast<uint32_t*>(f), sizeof(DevFlags) / sizeof(uint32_t), 0u, A.stream);
    return f;
}

DevFlags read_flags(const DevFlags* d, hipStream_t s) {
    DevFlags h{};
    SDFX_HIP(hipMemcpyAsync(&h, d, sizeof(DevFlags), hipMemcpyDeviceToHost, s));
    SDFX_HIP(hipStreamSynchronize(s));
    return h;
}

[[noreturn]] void fail_non_finite() {
    fail(SDFX_ERR_NON_FINITE_QUERY, "Query point contains non-finite values");
}
[[noreturn]] void fail_index(uint64_t nv) {
    fail(SDFX_ERR_INDEX_OUT_OF_RANGE,
         "index out of bounds: the len is " + std::to_string(nv) + " but a triangle index is >= " + std::to_string(nv));
}
[[noreturn]] void fail_nan() { fail(SDFX_ERR_NAN_DISTANCE, "NaN distance"); }
void check_internal(const DevFlags& f) {
    if (f.internal) fail(SDFX_ERR_DEVICE, "device consistency check failed (traversal stack or work decode)");
}
[[noreturn]] void fail_empty_mesh() { fail(SDFX_ERR_EMPTY_MESH, "called `Option::unwrap()` on a `None` value"); }

bool use_exhaustive(uint64_t nq, uint64_t nt) {
    const int64_t lim = tuning().exhaustive_max_pairs;
    if (lim <= 0) return false;
    return (double)nq * (double)nt <= (double)lim;
}

QueryMode mode_for(sdfx_acceleration_method m, sdfx_sign_method s) {
    switch (m) {
        case SDFX_ACCEL_BVH: return s == SDFX_SIGN_NORMAL ? QM_NORMAL_FOLD : QM_RAYCAST_VOTE;
        case SDFX_ACCEL_RTREE: return QM_NEAREST_SIGNED;
        default: return QM_NEAREST_VOTE;
    }
}

// generate_sdf on device memory. `finite_checked`: the host already ran lib.rs:374-376.
sdfx_status generate_sdf_core(const float* d_v, uint64_t nv, const uint32_t* d_idx, uint64_t ni, sdfx_topology topo,
                              const float* d_q, uint64_t nq, sdfx_acceleration_method method, sdfx_sign_method sign,
                              float* d_out, uint64_t* out_count, hipStream_t s, bool finite_checked) {
    check_enums(topo, method, sign);
    if (out_count) *out_count = 0;
    DevBuf<TriRec> tris;
    Lbvh bvh;
    Arena A(s);  // declared last: synchronises before the buffers above are released
    DevFlags* flags = new_flags(A);
    if (!finite_checked) check_queries_finite(d_q, nq, flags, s);
    auto finite_or = [&](void) {
        if (!finite_checked && read_flags(flags, s).non_finite_query) fail_non_finite();
    };
    // default.rs:38-41: without queries the None path never decodes its topology
    if (method == SDFX_ACCEL_NONE && nq == 0) {
        finite_or();
        return SDFX_OK;
    }
    uint64_t nt = 0;
    sdfx_status ts = topology_triangle_count(nv, d_idx != nullptr, ni, topo, &nt);
    if (ts != SDFX_OK) {
        finite_or();
        fail(ts, topology_error_message(ts, nv, ni));
    }
    tris.reset(nt);
    gather_triangles(d_v, nv, d_idx, topo, nt, tris.get(), flags, s);
    if (method == SDFX_ACCEL_NONE) {
        none_query(tris.get(), nt, d_q, nq, sign, d_out, flags, A);
    } else if (nt > 0 && use_exhaustive(nq, nt)) {
        exhaustive_query(tris.get(), nt, d_q, nq, mode_for(method, sign), d_out, flags, A);
    } else if (nt > 0) {
        lbvh_build(tris.get(), nt, bvh, A);
        lbvh_query(bvh, d_q, nq, mode_for(method, sign), d_out, flags, A);
    }
    DevFlags f = read_flags(flags, s);
    if (f.non_finite_query) fail_non_finite();
    if (f.index_out_of_range) fail_index(nv);
    check_internal(f);
    if (method != SDFX_ACCEL_NONE && nt == 0) {
        // generic/rtree_bvh.rs:96-98 returns an empty Vec; Rtree unwraps a
        // missing nearest neighbour (generic/rtree.rs:148); the bvh crate has no
        // root to traverse. Without queries neither panics.
        if (method == SDFX_ACCEL_RTREE_BVH) return SDFX_EMPTY_RESULT;
        if (nq > 0) fail_empty_mesh();
        return SDFX_OK;
    }
    if (f.nan_distance) fail_nan();
    if (out_count) *out_count = nq;
    return SDFX_OK;
}

sdfx_status grid_core(const float* d_v, uint64_t nv, const uint32_t* d_idx, uint64_t ni, sdfx_topology topo,
                      const float* first, const float* size, const uint64_t* count, sdfx_sign_method sign,
                      float* d_out, hipStream_t s, bool exact = false) {
    check_enums(topo, SDFX_ACCEL_NONE, sign);
    if (!first || !size || !count) fail(SDFX_ERR_INVALID_ARGUMENT, "null grid description");
    uint64_t nt = 0;
    sdfx_status ts = topology_triangle_count(nv, d_idx != nullptr, ni, topo, &nt);
    if (ts != SDFX_OK) fail(ts, topology_error_message(ts, nv, ni));
    GridParams g;
    for (int k = 0; k < 3; ++k) {
        g.first[k] = first[k];
        g.size[k] = size[k];
        g.count[k] = count[k];
    }
    const bool zero_axis = count[0] == 0 || count[1] == 0 || count[2] == 0;
    if (!zero_axis) {
        uint64_t c = count[0];
        bool big = false;
        for (int k = 1; k < 3; ++k) {
            if (count[k] > (UINT64_MAX / c)) big = true;
            else c *= count[k];
        }
        if (big || c >= (1ull << 32) || count[0] >= (1ull << 31) || count[1] >= (1ull << 31) ||
            count[2] >= (1ull << 31))
            fail(SDFX_ERR_INVALID_ARGUMENT, "grids are limited to fewer than 2^32 cells");
    }
    DevBuf<TriRec> tris(nt);
    Arena A(s);
    DevFlags* flags = new_flags(A);
    gather_triangles(d_v, nv, d_idx, topo, nt, tris.get(), flags, s);
    if (!(zero_axis && nt > 0)) {
        if (exact) grid_sdf_exact(tris.get(), nt, g, sign, d_out, flags, A);
        else grid_sdf(tris.get(), nt, g, sign, d_out, flags, A);
    }
    DevFlags f = read_flags(flags, s);
    if (f.index_out_of_range) fail_index(nv);
    check_internal(f);
    // grid.rs:311-315: snap_point_to_grid's clamp(0, n - 1) panics when n == 0
    if (zero_axis && nt > 0) fail(SDFX_ERR_INVALID_GRID, "assertion failed: min <= max");
    if (f.nan_distance) fail_nan();
    return SDFX_OK;
}

// One request of a batched grid call after the host-side checks grid_core
// makes (enums, topology counts, size limits); a failed check sets `status`.
struct BatchPrep {
    sdfx_status status = SDFX_OK;
    std::string message;
    uint64_t nt = 0;
    bool zero_axis = false;
    GridParams g{};
};

BatchPrep prepare_grid_request(const sdfx_grid_request& r, sdfx_sign_method sign) {
    BatchPrep p;
    try {
        check_enums(r.topology, SDFX_ACCEL_NONE, sign);
        if (!r.vertices && r.vertex_count) fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: vertices");
        if (!r.out && r.cell_count[0] * r.cell_count[1] * r.cell_count[2] != 0)
            fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: out");
        sdfx_status ts = topology_triangle_count(r.vertex_count, r.indices != nullptr, r.index_count, r.topology, &p.nt);
        if (ts != SDFX_OK) fail(ts, topology_error_message(ts, r.vertex_count, r.index_count));
        for (int k = 0; k < 3; ++k) {
            p.g.first[k] = r.first_cell[k];
            p.g.size[k] = r.cell_size[k];
            p.g.count[k] = r.cell_count[k];
        }
        p.zero_axis = r.cell_count[0] == 0 || r.cell_count[1] == 0 || r.cell_count[2] == 0;
        if (!p.zero_axis) {
            uint64_t c = r.cell_count[0];
            bool big = false;
            for (int k = 1; k < 3; ++k) {
                if (r.cell_count[k] > (UINT64_MAX / c)) big = true;
                else c *= r.cell_count[k];
            }
            if (big || c >= (1ull << 32) || r.cell_count[0] >= (1ull << 31) || r.cell_count[1] >= (1ull << 31) ||
                r.cell_count[2] >= (1ull << 31))
                fail(SDFX_ERR_INVALID_ARGUMENT, "grids are limited to fewer than 2^32 cells");
        }
    } catch (const Error& e) {
        p.status = e.status;
        p.message = e.what();
    }
    return p;
}

// Runs the requests whose device pointers are in `reqs`; fills each status.
sdfx_status grid_batch_core(sdfx_grid_request* reqs, uint64_t n, sdfx_sign_method sign, hipStream_t s) {
    std::vector<BatchPrep> prep(n);
    uint64_t nt_total = 0;
    for (uint64_t i = 0; i < n; ++i) {
        prep[i] = prepare_grid_request(reqs[i], sign);
        if (prep[i].status == SDFX_OK) nt_total += prep[i].nt;
    }
    DevBuf<TriRec> tris(nt_total);
    std::vector<GridBatchItem> items;
    Arena A(s);
    DevFlags* flags = A.alloc<DevFlags>(n ? n : 1);
    fill_u32(reinterpret_cast<uint32_t*>(flags), (n ? n : 1) * (sizeof(DevFlags) / sizeof(uint32_t)), 0u, s);
    uint64_t off = 0;
    std::vector<GatherJob> gathers;
    std::vector<uint64_t> item_req;  // request of each item
    for (uint64_t i = 0; i < n; ++i) {
        if (prep[i].status != SDFX_OK) continue;
        const sdfx_grid_request& r = reqs[i];
        gathers.push_back(GatherJob{r.vertices, r.vertex_count, r.indices, r.topology, prep[i].nt, tris.get() + off,
                                    flags + i});
        if (!(prep[i].zero_axis && prep[i].nt > 0)) {
            items.push_back(GridBatchItem{tris.get() + off, off, prep[i].nt, prep[i].g, r.out, (uint32_t)i});
            item_req.push_back(i);
        }
        off += prep[i].nt;
    }
    gather_triangles_batch(gathers, A);
    // a run that fails (its scratch, its launches) fails only its own requests
    std::vector<std::pair<sdfx_status, std::string>> run_errors;
    grid_sdf_batch(items, sign, flags, A, &run_errors);
    for (size_t k = 0; k < items.size(); ++k)
        if (run_errors[k].first != SDFX_OK) {
            prep[item_req[k]].status = run_errors[k].first;
            prep[item_req[k]].message = run_errors[k].second;
        }
    std::vector<DevFlags> hf(n ? n : 1);
    SDFX_HIP(hipMemcpyAsync(hf.data(), flags, (n ? n : 1) * sizeof(DevFlags), hipMemcpyDeviceToHost, s));
    SDFX_HIP(hipStreamSynchronize(s));
    sdfx_status first = SDFX_OK;
    std::string first_msg;
    for (uint64_t i = 0; i < n; ++i) {
        sdfx_status st = prep[i].status;
        std::string msg = prep[i].message;
        if (st == SDFX_OK) {  // in grid_core's order
            const DevFlags& f = hf[i];
            if (f.index_out_of_range) {
                st = SDFX_ERR_INDEX_OUT_OF_RANGE;
                msg = "index out of bounds: the len is " + std::to_string(reqs[i].vertex_count) +
                      " but a triangle index is >= " + std::to_string(reqs[i].vertex_count);
            } else if (f.internal) {
                st = SDFX_ERR_DEVICE;
                msg = "device consistency check failed (traversal stack or work decode)";
            } else if (prep[i].zero_axis && prep[i].nt > 0) {
                st = SDFX_ERR_INVALID_GRID;
                msg = "assertion failed: min <= max";
            } else if (f.nan_distance) {
                st = SDFX_ERR_NAN_DISTANCE;
                msg = "NaN distance";
            }
        }
        reqs[i].status = st;
        if (st != SDFX_OK && first == SDFX_OK) {
            first = st;
            first_msg = "request " + std::to_string(i) + ": " + msg;
        }
    }
    if (first != SDFX_OK) set_last_error(first_msg);
    return first;
}

template <typename T>
void upload(DevBuf<T>& d, const T* h, uint64_t n, hipStream_t s) {
    d.reset(n);
    if (n) SDFX_HIP(hipMemcpyAsync(d.get(), h, n * sizeof(T), hipMemcpyHostToDevice, s));
}

void require(bool ok, const char* what) {
    if (!ok) fail(SDFX_ERR_INVALID_ARGUMENT, std::string("null pointer: ") + what);
}

}  // namespace

extern "C" {

const char* sdfx_last_error(void) { return last_error().c_str(); }

int sdfx_version(void) { return SDFX_VERSION; }

sdfx_status sdfx_generate_sdf(const float* vertices, uint64_t vertex_count, const uint32_t* indices,
                              uint64_t index_count, sdfx_topology topology, const float* query_points,
                              uint64_t query_count, sdfx_acceleration_method method, sdfx_sign_method sign,
                              float* out, uint64_t* out_count) {
    return guarded([&]() -> sdfx_status {
        require(vertices || vertex_count == 0, "vertices");
        require(query_points || query_count == 0, "query_points");
        require(out || query_count == 0, "out");
        if (out_count) *out_count = 0;
        host_check_finite(query_points, query_count);
        hipStream_t s = thread_stream();
        DevBuf<float> dv, dq, dout(query_count);
        DevBuf<uint32_t> di;
        upload(dv, vertices, 3 * vertex_count, s);
        if (indices) upload(di, indices, index_count, s);
        upload(dq, query_points, 3 * query_count, s);
        uint64_t n = 0;
        sdfx_status st = generate_sdf_core(dv.get(), vertex_count, indices ? di.get() : nullptr, index_count,
                                           topology, dq.get(), query_count, method, sign, dout.get(), &n, s, true);
        if (st == SDFX_OK && n) {
            SDFX_HIP(hipMemcpyAsync(out, dout.get(), n * sizeof(float), hipMemcpyDeviceToHost, s));
            SDFX_HIP(hipStreamSynchronize(s));
        }
        if (out_count) *out_count = n;
        return st;
    });
}

sdfx_status sdfx_generate_sdf_device(const float* d_vertices, uint64_t vertex_count, const uint32_t* d_indices,
                                     uint64_t index_count, sdfx_topology topology, const float* d_query_points,
                                     uint64_t query_count, sdfx_acceleration_method method,
                                     sdfx_sign_method sign, float* d_out, uint64_t* out_count, void* stream) {
    return guarded([&]() -> sdfx_status {
        require(d_vertices || vertex_count == 0, "vertices");
        require(d_query_points || query_count == 0, "query_points");
        require(d_out || query_count == 0, "out");
        return generate_sdf_core(d_vertices, vertex_count, d_indices, index_count, topology, d_query_points,
                                 query_count, method, sign, d_out, out_count, (hipStream_t)stream, false);
    });
}

sdfx_status sdfx_generate_grid_sdf(const float* vertices, uint64_t vertex_count, const uint32_t* indices,
                                   uint64_t index_count, sdfx_topology topology, const float first_cell[3],
                                   const float cell_size[3], const uint64_t cell_count[3], sdfx_sign_method sign,
                                   float* out) {
    return guarded([&]() -> sdfx_status {
        require(vertices || vertex_count == 0, "vertices");
        require(cell_count != nullptr, "cell_count");
        uint64_t ncells = cell_count[0] * cell_count[1] * cell_count[2];
        require(out || ncells == 0, "out");
        hipStream_t s = thread_stream();
        DevBuf<float> dv, dout(ncells);
        DevBuf<uint32_t> di;
        upload(dv, vertices, 3 * vertex_count, s);
        if (indices) upload(di, indices, index_count, s);
        sdfx_status st = grid_core(dv.get(), vertex_count, indices ? di.get() : nullptr, index_count, topology,
                                   first_cell, cell_size, cell_count, sign, dout.get(), s);
        if (st == SDFX_OK && ncells) {
            SDFX_HIP(hipMemcpyAsync(out, dout.get(), ncells * sizeof(float), hipMemcpyDeviceToHost, s));
            SDFX_HIP(hipStreamSynchronize(s));
        }
        return st;
    });
}

sdfx_status sdfx_generate_grid_sdf_device(const float* d_vertices, uint64_t vertex_count,
                                          const uint32_t* d_indices, uint64_t index_count, sdfx_topology topology,
                                          const float first_cell[3], const float cell_size[3],
                                          const uint64_t cell_count[3], sdfx_sign_method sign, float* d_out,
                                          void* stream) {
    return guarded([&]() -> sdfx_status {
        require(d_vertices || vertex_count == 0, "vertices");
        require(cell_count != nullptr, "cell_count");
        return grid_core(d_vertices, vertex_count, d_indices, index_count, topology, first_cell, cell_size,
                         cell_count, sign, d_out, (hipStream_t)stream);
    });
}

sdfx_status sdfx_generate_grid_sdf_exact(const float* vertices, uint64_t vertex_count, const uint32_t* indices,
                                         uint64_t index_count, sdfx_topology topology, const float first_cell[3],
                                         const float cell_size[3], const uint64_t cell_count[3],
                                         sdfx_sign_method sign, float* out) {
    return guarded([&]() -> sdfx_status {
        require(vertices || vertex_count == 0, "vertices");
        require(cell_count != nullptr, "cell_count");
        uint64_t ncells = cell_count[0] * cell_count[1] * cell_count[2];
        require(out || ncells == 0, "out");
        hipStream_t s = thread_stream();
        DevBuf<float> dv, dout(ncells);
        DevBuf<uint32_t> di;
        upload(dv, vertices, 3 * vertex_count, s);
        if (indices) upload(di, indices, index_count, s);
        sdfx_status st = grid_core(dv.get(), vertex_count, indices ? di.get() : nullptr, index_count, topology,
                                   first_cell, cell_size, cell_count, sign, dout.get(), s, true);
        if (st == SDFX_OK && ncells) {
            SDFX_HIP(hipMemcpyAsync(out, dout.get(), ncells * sizeof(float), hipMemcpyDeviceToHost, s));
            SDFX_HIP(hipStreamSynchronize(s));
        }
        return st;
    });
}

sdfx_status sdfx_generate_grid_sdf_exact_device(const float* d_vertices, uint64_t vertex_count,
                                                const uint32_t* d_indices, uint64_t index_count,
                                                sdfx_topology topology, const float first_cell[3],
                                                const float cell_size[3], const uint64_t cell_count[3],
                                                sdfx_sign_method sign, float* d_out, void* stream) {
    return guarded([&]() -> sdfx_status {
        require(d_vertices || vertex_count == 0, "vertices");
        require(cell_count != nullptr, "cell_count");
        return grid_core(d_vertices, vertex_count, d_indices, index_count, topology, first_cell, cell_size,
                         cell_count, sign, d_out, (hipStream_t)stream, true);
    });
}

sdfx_status sdfx_build_sdf_acceleration(const float* vertices, uint64_t vertex_count, const uint32_t* indices,
                                        uint64_t index_count, sdfx_topology topology,
                                        sdfx_acceleration_method method, sdfx_sign_method sign,
                                        sdfx_mesh** mesh_out) {
    return guarded([&]() -> sdfx_status {
        require(mesh_out != nullptr, "mesh_out");
        *mesh_out = nullptr;
        require(vertices || vertex_count == 0, "vertices");
        check_enums(topology, method, sign);
        hipStream_t s = thread_stream();
        std::unique_ptr<sdfx_mesh> m(new sdfx_mesh());
        m->device = current_device();
        m->method = method;
        m->sign = method == SDFX_ACCEL_RTREE ? SDFX_SIGN_NORMAL
                                             : (method == SDFX_ACCEL_RTREE_BVH ? SDFX_SIGN_RAYCAST : sign);
        uint64_t nt = 0;
        sdfx_status ts = topology_triangle_count(vertex_count, indices != nullptr, index_count, topology, &nt);
        if (ts != SDFX_OK) {
            // default.rs:87-97 stores the topology unchecked; the query decodes it
            if (method != SDFX_ACCEL_NONE) fail(ts, topology_error_message(ts, vertex_count, index_count));
            m->deferred = ts;
            m->deferred_msg = topology_error_message(ts, vertex_count, index_count);
            *mesh_out = m.release();
            return SDFX_OK;
        }
        DevBuf<float> dv;
        DevBuf<uint32_t> di;
        DevBuf<TriRec> tris(nt);
        {
            Arena A(s);
            upload(dv, vertices, 3 * vertex_count, s);
            if (indices) upload(di, indices, index_count, s);
            DevFlags* flags = new_flags(A);
            gather_triangles(dv.get(), vertex_count, indices ? di.get() : nullptr, topology, nt, tris.get(), flags,
                             s);
            DevFlags f = read_flags(flags, s);
            if (f.index_out_of_range) {
                if (method != SDFX_ACCEL_NONE) fail_index(vertex_count);
                m->deferred = SDFX_ERR_INDEX_OUT_OF_RANGE;
                m->deferred_msg = "index out of bounds: the len is " + std::to_string(vertex_count);
            }
            if (method == SDFX_ACCEL_RTREE_BVH && nt == 0) return SDFX_EMPTY_RESULT;  // rtree_bvh.rs:124-126
            m->n_tris = nt;
            if (method == SDFX_ACCEL_NONE) {
                m->tris = std::move(tris);
            } else if (nt > 0) {
                lbvh_build(tris.get(), nt, m->bvh, A);
            }
        }
        *mesh_out = m.release();
        return SDFX_OK;
    });
}

sdfx_status sdfx_query_sdf(const sdfx_mesh* mesh, const float* query_points, uint64_t query_count, float* out,
                           uint64_t* out_count) {
    return guarded([&]() -> sdfx_status {
        require(mesh != nullptr, "mesh");
        require(query_points || query_count == 0, "query_points");
        require(out || query_count == 0, "out");
        if (out_count) *out_count = 0;
        if (query_count == 0) return SDFX_OK;
        if (mesh->deferred != SDFX_OK) fail(mesh->deferred, mesh->deferred_msg);
        if (mesh->method != SDFX_ACCEL_NONE && mesh->n_tris == 0) fail_empty_mesh();
        SDFX_HIP(hipSetDevice(mesh->device));
        hipStream_t s = thread_stream();
        DevBuf<float> dq, dout(query_count);
        upload(dq, query_points, 3 * query_count, s);
        {
            Arena A(s);
            DevFlags* flags = new_flags(A);
            if (mesh->method == SDFX_ACCEL_NONE)
                none_query(mesh->tris.get(), mesh->n_tris, dq.get(), query_count, mesh->sign, dout.get(), flags, A);
            else
                lbvh_query(mesh->bvh, dq.get(), query_count, mode_for(mesh->method, mesh->sign), dout.get(), flags,
                           A);
            DevFlags f = read_flags(flags, s);
            check_internal(f);
            if (f.nan_distance) fail_nan();
        }
        SDFX_HIP(hipMemcpyAsync(out, dout.get(), query_count * sizeof(float), hipMemcpyDeviceToHost, s));
        SDFX_HIP(hipStreamSynchronize(s));
        if (out_count) *out_count = query_count;
        return SDFX_OK;
    });
}

// ---- sdfx_query_sdf_noalloc: sdfx_query_sdf on device memory without any
// allocation. Its temporaries (error flags, the Morton order of the queries
// and its radix-sort buffers, the deferred-search records) come from
// `scratch`, a device array each GPU holds from the library's load to the end
// of the process. In host code `scratch` names a host-side stand-in; its
// device address on the current device comes from hipGetSymbolAddress, once
// per device, and only that address is used.

namespace {

constexpr uint64_t NOALLOC_MAX_QUERIES = 262144;
constexpr size_t NOALLOC_SCRATCH_BYTES = 16u << 20;  // 48 B per query at most, plus small tables
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

struct NoallocDevice {
    std::mutex mu;  // held for a whole call: calls on one device run one after another
    void* addr = nullptr;
};
NoallocDevice g_noalloc[NOALLOC_MAX_DEVICES];

// The device address of `scratch` on the current device `dev`, looked up on
// the first call on that device.
void* noalloc_scratch(int dev) {
    NoallocDevice& D = g_noalloc[dev];
    if (!D.addr) {
        void* p = nullptr;
        SDFX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(SDFX_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        D.addr = p;
    }
    return D.addr;
}

}  // namespace

sdfx_status sdfx_query_sdf_noalloc(const sdfx_mesh* mesh, const float* d_query_points, uint64_t query_count,
                                   float* d_out, uint64_t* out_count, void* stream) {
    return guarded([&]() -> sdfx_status {
        require(mesh != nullptr, "mesh");
        require(d_query_points || query_count == 0, "query_points");
        require(d_out || query_count == 0, "out");
        if (out_count) *out_count = 0;
        if (query_count > NOALLOC_MAX_QUERIES)
            fail(SDFX_ERR_INVALID_ARGUMENT, "sdfx_query_sdf_noalloc takes at most " +
                                                std::to_string(NOALLOC_MAX_QUERIES) + " queries per call (got " +
                                                std::to_string(query_count) + ")");
        if (query_count == 0) return SDFX_OK;
        if (mesh->deferred != SDFX_OK) fail(mesh->deferred, mesh->deferred_msg);
        if (mesh->method != SDFX_ACCEL_NONE && mesh->n_tris == 0) fail_empty_mesh();
        if (mesh->device < 0 || mesh->device >= NOALLOC_MAX_DEVICES) fail(SDFX_ERR_INVALID_ARGUMENT, "device index");
        SDFX_HIP(hipSetDevice(mesh->device));
        hipStream_t s = static_cast<hipStream_t>(stream);
        std::lock_guard<std::mutex> lk(g_noalloc[mesh->device].mu);
        void* base = noalloc_scratch(mesh->device);
        Arena A(s, base, NOALLOC_SCRATCH_BYTES);
        DevFlags* flags = new_flags(A);
        // Under stream capture the work is only recorded: the device-side
        // checks cannot be read back before a replay has run, so they go
        // unreported (every kernel and fill here is capturable: no memset
        // nodes, no host synchronisation).
        hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
        SDFX_HIP(hipStreamIsCapturing(s, &capture));
        try {
            if (mesh->method == SDFX_ACCEL_NONE)
                none_query(mesh->tris.get(), mesh->n_tris, d_query_points, query_count, mesh->sign, d_out, flags, A);
            else
                lbvh_query(mesh->bvh, d_query_points, query_count, mode_for(mesh->method, mesh->sign), d_out, flags,
                           A);
        } catch (...) {
            // work already enqueued on the scratch ends before the next call may reuse it
            if (capture != hipStreamCaptureStatusActive) (void)hipStreamSynchronize(s);
            throw;
        }
        if (capture == hipStreamCaptureStatusActive) {
            if (out_count) *out_count = query_count;
            return SDFX_OK;
        }
        DevFlags f = read_flags(flags, s);
        check_internal(f);
        if (f.nan_distance) fail_nan();
        if (out_count) *out_count = query_count;
        return SDFX_OK;
    });
}

sdfx_status sdfx_query_sdf_batch_device(sdfx_point_request* requests, uint64_t request_count, void* stream) {
    return guarded([&]() -> sdfx_status {
        require(requests || request_count == 0, "requests");
        if (request_count == 0) return SDFX_OK;
        hipStream_t s = static_cast<hipStream_t>(stream);
        // sdfx_query_sdf's host-side checks, per request
        std::vector<std::string> msg(request_count);
        int device = -1;
        for (uint64_t i = 0; i < request_count; ++i) {
            sdfx_point_request& r = requests[i];
            r.status = SDFX_OK;
            try {
                require(r.mesh != nullptr, "mesh");
                require(r.query_points || r.query_count == 0, "query_points");
                require(r.out || r.query_count == 0, "out");
                if (r.query_count == 0) continue;
                if (r.mesh->deferred != SDFX_OK) fail(r.mesh->deferred, r.mesh->deferred_msg);
                if (r.mesh->method != SDFX_ACCEL_NONE && r.mesh->n_tris == 0) fail_empty_mesh();
                if (device < 0) device = r.mesh->device;
                if (r.mesh->device != device) fail(SDFX_ERR_INVALID_ARGUMENT, "the meshes of a batch live on one device");
            } catch (const Error& e) {
                r.status = e.status;
                msg[i] = e.what();
            }
        }
        if (device >= 0) {
            SDFX_HIP(hipSetDevice(device));
            Arena A(s);
            DevFlags* flags = A.alloc<DevFlags>(request_count);
            fill_u32(reinterpret_cast<uint32_t*>(flags), request_count * (sizeof(DevFlags) / sizeof(uint32_t)), 0u, s);
            std::vector<PointBatchItem> by_mode[4];
            for (uint64_t i = 0; i < request_count; ++i) {
                const sdfx_point_request& r = requests[i];
                if (r.status != SDFX_OK || r.query_count == 0) continue;
                if (r.mesh->method == SDFX_ACCEL_NONE)
                    none_query(r.mesh->tris.get(), r.mesh->n_tris, r.query_points, r.query_count, r.mesh->sign, r.out,
                               flags + i, A);
                else
                    by_mode[mode_for(r.mesh->method, r.mesh->sign)].push_back(
                        PointBatchItem{&r.mesh->bvh, r.query_points, r.query_count, r.out, (uint32_t)i});
            }
            for (int m = 0; m < 4; ++m)
                if (!by_mode[m].empty()) lbvh_query_batch(by_mode[m], (QueryMode)m, flags, A);
            std::vector<DevFlags> hf(request_count);
            SDFX_HIP(hipMemcpyAsync(hf.data(), flags, request_count * sizeof(DevFlags), hipMemcpyDeviceToHost, s));
            SDFX_HIP(hipStreamSynchronize(s));
            for (uint64_t i = 0; i < request_count; ++i) {
                if (requests[i].status != SDFX_OK || requests[i].query_count == 0) continue;
                if (hf[i].internal) {
                    requests[i].status = SDFX_ERR_DEVICE;
                    msg[i] = "device consistency check failed (traversal stack or work decode)";
                } else if (hf[i].nan_distance) {
                    requests[i].status = SDFX_ERR_NAN_DISTANCE;
                    msg[i] = "NaN distance";
                }
            }
        }
        for (uint64_t i = 0; i < request_count; ++i)
            if (requests[i].status != SDFX_OK) {
                set_last_error("request " + std::to_string(i) + ": " + msg[i]);
                return requests[i].status;
            }
        return SDFX_OK;
    });
}

sdfx_status sdfx_query_vector_to_closest_point(const sdfx_mesh* mesh, const float* query_points,
                                               uint64_t query_count, float* vectors_out, uint8_t* inside_out) {
    return guarded([&]() -> sdfx_status {
        require(mesh != nullptr, "mesh");
        require(query_points || query_count == 0, "query_points");
        require((vectors_out && inside_out) || query_count == 0, "outputs");
        if (mesh->method != SDFX_ACCEL_RTREE_BVH)
            fail(SDFX_ERR_INVALID_ARGUMENT, "query_vector_to_closest_point needs an RtreeBvh acceleration");
        if (query_count == 0) return SDFX_OK;
        SDFX_HIP(hipSetDevice(mesh->device));
        hipStream_t s = thread_stream();
        DevBuf<float> dq, dvec(3 * query_count);
        DevBuf<uint8_t> din(query_count);
        upload(dq, query_points, 3 * query_count, s);
        {
            Arena A(s);
            DevFlags* flags = new_flags(A);
            lbvh_closest_vectors(mesh->bvh, dq.get(), query_count, dvec.get(), din.get(), flags, A);
            check_internal(read_flags(flags, s));
        }
        SDFX_HIP(hipMemcpyAsync(vectors_out, dvec.get(), 3 * query_count * sizeof(float), hipMemcpyDeviceToHost, s));
        SDFX_HIP(hipMemcpyAsync(inside_out, din.get(), query_count, hipMemcpyDeviceToHost, s));
        SDFX_HIP(hipStreamSynchronize(s));
        return SDFX_OK;
    });
}

void sdfx_free_sdf_acceleration(sdfx_mesh* mesh) {
    if (!mesh) return;
    (void)hipSetDevice(mesh->device);
    delete mesh;
}

sdfx_status sdfx_check_mesh_triangle_list(const float* vertices, uint64_t vertex_count, const uint32_t* faces,
                                          uint64_t face_index_count, int* kind_out, uint64_t* where_out,
                                          char* message, uint64_t message_capacity) {
    return guarded([&]() -> sdfx_status {
        require(vertices || vertex_count == 0, "vertices");
        require(faces || face_index_count == 0, "faces");
        const MeshCheck r = check_mesh_triangle_list(vertices, vertex_count, faces, face_index_count,
                                                     thread_stream());
        if (kind_out) *kind_out = r.kind;
        if (where_out) *where_out = r.where;
        if (message && message_capacity) {
            const uint64_t n = std::min<uint64_t>(r.message.size(), message_capacity - 1);
            std::memcpy(message, r.message.data(), n);
            message[n] = '\0';
        }
        return r.kind ? SDFX_MESH_INVALID : SDFX_OK;
    });
}

sdfx_status sdfx_generate_grid_sdf_batch_device(sdfx_grid_request* requests, uint64_t request_count,
                                                sdfx_sign_method sign, void* stream) {
    return guarded([&]() -> sdfx_status {
        require(requests || request_count == 0, "requests");
        return grid_batch_core(requests, request_count, sign, (hipStream_t)stream);
    });
}

sdfx_status sdfx_generate_grid_sdf_batch(sdfx_grid_request* requests, uint64_t request_count,
                                         sdfx_sign_method sign) {
    return guarded([&]() -> sdfx_status {
        require(requests || request_count == 0, "requests");
        hipStream_t s = thread_stream();
        // device copies of every request's inputs and outputs
        std::vector<DevBuf<float>> dv(request_count), dout(request_count);
        std::vector<DevBuf<uint32_t>> di(request_count);
        std::vector<sdfx_grid_request> dreq(requests, requests + request_count);
        // the host-side checks first, so only valid requests' buffers are
        // allocated and uploaded (a failing request keeps its status)
        std::vector<uint64_t> ncells(request_count, 0);
        for (uint64_t i = 0; i < request_count; ++i) {
            const sdfx_grid_request& r = requests[i];
            const BatchPrep p = prepare_grid_request(r, sign);
            if (p.status != SDFX_OK) continue;  // grid_batch_core reports it
            if (!p.zero_axis) ncells[i] = r.cell_count[0] * r.cell_count[1] * r.cell_count[2];
            if (r.vertices) upload(dv[i], r.vertices, 3 * r.vertex_count, s);
            if (r.indices) upload(di[i], r.indices, r.index_count, s);
            dout[i].reset(ncells[i]);
            dreq[i].vertices = r.vertices ? dv[i].get() : nullptr;
            dreq[i].indices = r.indices ? di[i].get() : nullptr;
            dreq[i].out = r.out ? dout[i].get() : nullptr;
        }
        const sdfx_status st = grid_batch_core(dreq.data(), request_count, sign, s);
        for (uint64_t i = 0; i < request_count; ++i) {
            requests[i].status = dreq[i].status;
            if (dreq[i].status == SDFX_OK && ncells[i] && requests[i].out)
                SDFX_HIP(hipMemcpyAsync(requests[i].out, dout[i].get(), ncells[i] * sizeof(float),
                                        hipMemcpyDeviceToHost, s));
        }
        SDFX_HIP(hipStreamSynchronize(s));
        return st;
    });
}

}  // extern "C"
